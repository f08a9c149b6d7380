"""The .pac1 writer and reader (Classes_IO.cs:57-261): a direct-sound and a
ray-traced section written for a traced-like receiver bank read back
unchanged, and the byte layout matches the field order and .NET
BinaryWriter encoding of Write_Pac1 / Direct_Sound.Write_Data /
Receiver_Bank.Write_Data (file size from the layout; 7-bit length-prefixed
strings; no 1/3-octave flag below version 2.7)."""
import struct

import numpy as np

from pachydermacoustic_universal_amd import pac1
from pachydermacoustic_universal_amd import pachyderm as P
from pachydermacoustic_universal_amd import rooms


def bank_and_direct(seed=0, cut_ms=120.0):
    room = rooms.shoebox()
    sc = P.Polygon_Scene.from_room(room)
    src = P.GeodesicSource([120] * 8, room.source)
    rb = P.Receiver_Bank(room.receivers, src, sc, 1000, cut_ms, radius=room.radius)
    rng = np.random.default_rng(seed)
    rb.planes[:] = rng.normal(size=rb.planes.shape)
    ds = P.Direct_Sound(src, rb, sc, [0, 7])
    ds.Begin()
    return rb, ds, src


def test_round_trip(tmp_path):
    rb, ds, src = bank_and_direct()
    f = tmp_path / "t.pac1"
    pac1.write_pac1(f, ds, rb, date="12/03/2026 10:00:00 PM")
    d = pac1.read_pac1(f)
    assert d["version"] == pac1.VERSION and d["sample_rate"] == 1000 and d["cutoff_ms"] == 120.0
    np.testing.assert_array_equal(d["sources"][0], src.Origin)
    np.testing.assert_array_equal(d["receivers"], rb.Origins)
    np.testing.assert_allclose(d["rho_c"], rb.scene.rho * rb.scene.c)
    rt = d["ray_traced"][0]
    assert rt["sample_ct"] == rb.SampleCT
    np.testing.assert_array_equal(rt["planes"], rb.planes[:8])
    dd = d["direct"][0]
    assert dd["type"] == "Geodesic" and np.allclose(dd["SWL"], 120)
    for q in range(len(rb.Origins)):
        r = dd["receivers"][q]
        assert r["n"] == 100 and r["valid"] == ds.IsOccluded(q) and r["time"] == ds.Time(q)
        np.testing.assert_array_equal(r["energy"].T, ds.Io[q])
        # direction sums: the arrival direction (receiver - source) times the
        # energy, on the side of the sign of -direction (Classes_Direct.cs:828-835)
        dvec = rb.Origins[q] - src.Origin
        dvec /= np.linalg.norm(dvec)
        for a in range(3):
            got = r["pos"][:, 0, a] if -dvec[a] * ds.Io[q, 0, 0] > 0 else r["neg"][:, 0, a]
            np.testing.assert_allclose(got, (-dvec[a] * ds.Io[q, :, 0]).astype(np.float32), rtol=1e-6)


def test_byte_layout(tmp_path):
    rb, ds, src = bank_and_direct(cut_ms=5.0)
    f = tmp_path / "t.pac1"
    date = "12/03/2026 10:00:00 PM"
    pac1.write_pac1(f, ds, rb, date=date)
    b = f.read_bytes()
    R, n, S = 3, 100, rb.SampleCT
    head = (1 + len(date)) + (1 + 8) + 8 + 4 + 4 + 24 + 4 + R * 32
    direct = (1 + 25) + (1 + 8) + 8 * 8 + 8 + R * (4 + 1 + 8 + n * (8 * 8 + 8 * 3 * 2 * 4))
    traced = (1 + 15) + 4 + 4 + 4 + 8 + R * (8 + S * 8 * 8 * 8)
    assert len(b) == head + direct + traced + (1 + 3)
    assert b[0] == len(date) and b[1:1 + len(date)].decode() == date
    o = 1 + len(date)
    assert b[o] == 8 and b[o + 1:o + 9] == b"2.6.0.20"
    o += 9
    assert struct.unpack("<d", b[o:o + 8])[0] == 5.0 and struct.unpack("<i", b[o + 8:o + 12])[0] == 1000
    assert struct.unpack("<i", b[o + 12:o + 16])[0] == 1  # one source, no 1/3-octave flag before it
    assert b[head] == 25 and b[head + 1:head + 26] == b"Direct_Sound w sourcedata"
    assert b[head + direct] == 15 and b[head + direct + 1:head + direct + 16] == b"Ray-Traced_Data"
    assert b[-4:] == b"\x03End"


def test_long_strings_use_two_length_bytes(tmp_path):
    rb, ds, _ = bank_and_direct(cut_ms=2.0)
    f = tmp_path / "t.pac1"
    date = "x" * 200  # 200 = 0xC8 -> 0xC8 0x01 in 7-bit groups
    pac1.write_pac1(f, None, rb, date=date)
    b = f.read_bytes()
    assert b[0] == 0xC8 and b[1] == 0x01 and b[2:202] == date.encode()
    assert pac1.read_pac1(f)["date"] == date


class _IS:
    """ImageSourceData stand-in holding Specular_Paths (Paths[rec])."""

    def __init__(self, paths):
        self.Paths = paths


def image_source_data(rb):
    from oracle import pyoracle as O
    room = rooms.shoebox()
    med = O.standard_medium()
    raw = O.Scene(room, grid=(4, 4, 4)).image_sources(2, room.source, np.ones(8), room.receivers, med)
    paths = [[] for _ in range(len(rb.Origins))]
    for p in raw:
        paths[p["rec"]].append(P.Specular_Path(p["rec"], p["order"], p["seq"], p["poly"], p["pts"], p["length"],
                                               p["time"], p["energy"], med.rho_c))
    return _IS(paths), raw


def test_image_source_section_round_trip(tmp_path):
    rb, ds, src = bank_and_direct(cut_ms=5.0)
    isd, raw = image_source_data(rb)
    for special in (True, False):
        f = tmp_path / f"is{int(special)}.pac1"
        pac1.write_pac1(f, ds, rb, date="d", image_sources=isd, special_filter=special)
        d = pac1.read_pac1(f)
        got = d["image_sources"][0]
        assert [len(x) for x in got] == [len(x) for x in isd.Paths]
        for q, paths in enumerate(isd.Paths):
            for g, p in zip(got[q], paths):
                np.testing.assert_array_equal(g["pts"], p.Path[0])
                np.testing.assert_array_equal(g["energy"], p.PathEnergy)
                np.testing.assert_array_equal(g["prms"], p.prms)
                assert g["time"] == p.TravelTime and g["seq"] == p.Reflection_Sequence
                if special:
                    assert g["filter"].shape == (16384,) and np.all(g["filter"] == 1)
                else:
                    assert g["filter"] is None
        # the section's size from Write_Data's field list
        n_paths = sum(len(x) for x in isd.Paths)
        pts = sum(len(p.Path[0]) for x in isd.Paths for p in x)
        seq = sum(len(p.Reflection_Sequence) for x in isd.Paths for p in x)
        size = (1 + len("Image-Source_Data")) + 8 * len(isd.Paths) + n_paths * (2 + 4 + 64 + 1 + 64 + 8) \
            + 24 * pts + 4 * seq + (n_paths * (4 + 16 * 16384) if special else 0)
        f0 = tmp_path / "none.pac1"
        pac1.write_pac1(f0, ds, rb, date="d")
        assert f.stat().st_size - f0.stat().st_size == size
