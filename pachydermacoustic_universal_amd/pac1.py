"""Pachyderm's .pac1 result file (FileIO.Write_Pac1 / Read_Pac1,
Classes_IO.cs:57-261), for the direct sound and ray-traced data this package
computes, so results open in existing Pachyderm tooling.

Layout (.NET BinaryWriter: little-endian numbers, strings as a 7-bit-encoded
byte length followed by UTF-8; bool 1 byte, int 4, float 4, double 8):

    string date, string version ("2.6.0.20": the file layout of that
        version; the 1/3-octave flag is written only from 2.7 on)
    double cut-off time (ms), int sample rate
    int source count; per source double x, y, z
    int receiver count; per receiver double x, y, z, rho c
    per source:
      Direct_Sound.Write_Data (Classes_Direct.cs:61-104):
        string "Direct_Sound w sourcedata", string source type, double SWL[8],
        double delay (ms); per receiver: int window length n, bool validity,
        double arrival time (s); per window sample: double energy[8], then for
        each band and axis float positive, float negative direction sums
      ImageSourceData.Write_Data (Classes_ImageSource.cs:1443-1583), when given:
        string "Image-Source_Data"; per receiver: int receiver, int path count;
        per Specular_Path: short 0, int point count, points (double x, y, z),
        double energy[8], bool special filter (the reference writes true and
        16,384 complex ones: int length, then double real, imaginary per
        entry), double prms[8], double travel time, int sequence[points - 2]
        string "Ray-Traced_Data", int receiver type (0 stationary), int sample
        rate, int sample count, double cut-off time (ms); per receiver: double
        speed of sound; per sample and band: double energy, double pressure,
        double positive x, y, z, double negative x, y, z
    string "End"

Compound paths (edge diffraction, curved surfaces) are not produced by this
package and are neither written nor read.
"""
from __future__ import annotations

import datetime
import struct

import numpy as np

VERSION = "2.6.0.20"


class _W:
    def __init__(self, f):
        self.f = f

    def string(self, s):
        b = s.encode("utf-8")
        n = len(b)
        out = bytearray()
        while True:
            byte = n & 0x7F
            n >>= 7
            if n:
                out.append(byte | 0x80)
            else:
                out.append(byte)
                break
        self.f.write(bytes(out) + b)

    def i32(self, v):
        self.f.write(struct.pack("<i", int(v)))

    def f64(self, v):
        self.f.write(struct.pack("<d", float(v)))

    def f64s(self, a):
        self.f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())

    def f32(self, v):
        self.f.write(struct.pack("<f", float(v)))

    def i16(self, v):
        self.f.write(struct.pack("<h", int(v)))

    def boolean(self, v):
        self.f.write(b"\x01" if v else b"\x00")


class _R:
    def __init__(self, data):
        self.b = memoryview(data)
        self.o = 0

    def _take(self, n):
        if self.o + n > len(self.b):
            raise ValueError(".pac1 file ends early")
        v = self.b[self.o:self.o + n]
        self.o += n
        return v

    def string(self):
        n, shift = 0, 0
        while True:
            byte = self._take(1)[0]
            n |= (byte & 0x7F) << shift
            shift += 7
            if not byte & 0x80:
                break
        return bytes(self._take(n)).decode("utf-8")

    def i32(self):
        return struct.unpack("<i", self._take(4))[0]

    def f64(self):
        return struct.unpack("<d", self._take(8))[0]

    def f64s(self, n):
        return np.frombuffer(bytes(self._take(8 * n)), dtype="<f8").copy()

    def f32(self):
        return struct.unpack("<f", self._take(4))[0]

    def i16(self):
        return struct.unpack("<h", self._take(2))[0]

    def boolean(self):
        return self._take(1)[0] != 0


def _direction_sums(direct, rec):
    """Direct_Sound's Dir_Rec_Pos / Dir_Rec_Neg of a point source
    (Classes_Direct.cs:828-835): the arrival direction times the energy, split
    by sign, at sample 0 of the window."""
    n = direct.Io.shape[2]
    pos = np.zeros((8, n, 3), dtype=np.float32)
    neg = np.zeros((8, n, 3), dtype=np.float32)
    d = np.asarray(direct.Rec.Origins[rec], dtype=np.float64) - np.asarray(direct.Src.Origin, dtype=np.float64)
    d /= np.linalg.norm(d)
    for oct in range(8):
        V = d * direct.Io[rec, oct, 0]
        for a in range(3):
            if -V[a] > 0:
                pos[oct, 0, a] -= np.float32(V[a])
            else:
                neg[oct, 0, a] -= np.float32(V[a])
    return pos, neg


_FILTER_LEN = 16384


def _write_image_sources(w, isd, special_filter):
    """ImageSourceData.Write_Data (Classes_ImageSource.cs:1443-1583), specular paths."""
    w.string("Image-Source_Data")
    ones = np.zeros(2 * _FILTER_LEN)
    ones[0::2] = 1.0
    for q, paths in enumerate(isd.Paths):
        w.i32(q)
        w.i32(len(paths))
        for p in paths:
            w.i16(0)
            pts = p.Path[0]
            w.i32(len(pts))
            w.f64s(pts)
            w.f64s(p.PathEnergy)
            w.boolean(special_filter)
            if special_filter:
                w.i32(_FILTER_LEN)
                w.f64s(ones)
            w.f64s(p.prms)
            w.f64(p.TravelTime)
            for d in p.Reflection_Sequence:
                w.i32(d)


def write_pac1(filename, direct, receiver, version=VERSION, date=None, image_sources=None, special_filter=True):
    """Write one source's Direct_Sound (pachyderm.Direct_Sound, after Begin),
    ImageSourceData (pachyderm.ImageSourceData after Begin; optional) and
    ray-traced Receiver_Bank (pachyderm.Receiver_Bank after Combine, scaled).
    Direct and receiver may not both be None.  special_filter=False writes
    each path's materials filter as absent (a form the reference's reader
    accepts) instead of the 16,384 unit entries the reference writes."""
    if direct is None and receiver is None:
        raise ValueError("There is no simulated data to save.")
    rb = receiver if receiver is not None else direct.Rec
    src = direct.Src if direct is not None else None
    with open(filename, "wb") as f:
        w = _W(f)
        w.string(date or datetime.datetime.now().strftime("%m/%d/%Y %I:%M:%S %p"))
        w.string(version)
        w.f64(rb.CutOffTime)
        w.i32(rb.SampleRate)
        if float(version[:2]) >= 2.7:  # the writer's own test (Classes_IO.cs:81)
            w.boolean(False)
        origin = src.Origin if src is not None else np.zeros(3)
        w.i32(1)
        for v in origin:
            w.f64(v)
        R = len(rb.Origins)
        w.i32(R)
        rho_c = rb.scene.rho * rb.scene.c
        for q in range(R):
            for v in rb.Origins[q]:
                w.f64(v)
            w.f64(rho_c)
        if direct is not None:
            w.string("Direct_Sound w sourcedata")
            w.string(src.type_name)
            for o in range(8):
                w.f64(src.SWL(o))
            w.f64(direct.Delay_ms)
            for q in range(R):
                n = direct.Io.shape[2]
                w.i32(n)
                w.boolean(direct.IsOccluded(q))
                w.f64(direct.Time(q))
                pos, neg = _direction_sums(direct, q)
                for t in range(n):
                    w.f64s(direct.Io[q, :, t])
                    for oct in range(8):
                        for a in range(3):
                            w.f32(pos[oct, t, a])
                            w.f32(neg[oct, t, a])
        if image_sources is not None:
            _write_image_sources(w, image_sources, special_filter)
        if receiver is not None:
            w.string("Ray-Traced_Data")
            w.i32(0)  # Stationary
            w.i32(rb.SampleRate)
            w.i32(rb.SampleCT)
            w.f64(rb.CutOffTime)
            P = rb.planes
            for q in range(R):
                w.f64(rb.scene.c)
                # [sample][band][energy, pressure, +x, +y, +z, -x, -y, -z]
                blk = np.empty((rb.SampleCT, 8, 8))
                blk[:, :, 0] = P[0, q].T
                blk[:, :, 1] = P[1, q].T
                for a in range(3):
                    blk[:, :, 2 + a] = P[2 + a, q].T
                    blk[:, :, 5 + a] = P[5 + a, q].T
                w.f64s(blk)
        w.string("End")


def read_pac1(filename):
    """Read a file written by write_pac1 (or by Pachyderm 2.x for these
    sections): returns a dict with the header, 'direct' (per source: type,
    SWL, delay, per receiver n / validity / time / energy [n, 8] / pos, neg
    [8, n, 3]) and 'ray_traced' (per source: sample rate, cut-off, speed of
    sound [R], planes [8, R, 8, SampleCT]: energy, pressure, +x, +y, +z, -x,
    -y, -z)."""
    with open(filename, "rb") as f:
        r = _R(f.read())
    out = {"date": r.string(), "version": r.string()}
    out["cutoff_ms"] = r.f64()
    out["sample_rate"] = r.i32()
    v3 = float(out["version"][:3])
    out["third_octave"] = r.boolean() if v3 >= 2.7 else False
    n_src = r.i32() if v3 >= 1.1 else 1
    out["sources"] = np.array([[r.f64(), r.f64(), r.f64()] for _ in range(n_src)])
    R = r.i32()
    recs, rho_c = [], []
    for _ in range(R):
        recs.append([r.f64(), r.f64(), r.f64()])
        rho_c.append(r.f64() if v3 >= 2.0 else 400.0)
    out["receivers"] = np.array(recs)
    out["rho_c"] = np.array(rho_c)
    out["direct"], out["ray_traced"], out["image_sources"] = [], [], []
    while True:
        tag = r.string()
        if tag in ("End", ""):
            return out
        if tag in ("Direct_Sound", "Direct_Sound w sourcedata"):
            d = {"type": r.string() if tag.endswith("sourcedata") else ""}
            d["SWL"] = np.array([r.f64() for _ in range(8)]) if tag.endswith("sourcedata") else None
            d["delay_ms"] = r.f64() if tag.endswith("sourcedata") else 0.0
            d["receivers"] = []
            for _ in range(R):
                n = r.i32()
                rec = {"n": n, "valid": r.boolean(), "time": r.f64(), "energy": np.zeros((n, 8)),
                       "pos": np.zeros((8, n, 3), dtype=np.float32), "neg": np.zeros((8, n, 3), dtype=np.float32)}
                for t in range(n):
                    rec["energy"][t] = r.f64s(8)
                    for oct in range(8):
                        for a in range(3):
                            rec["pos"][oct, t, a] = r.f32()
                            rec["neg"][oct, t, a] = r.f32()
                d["receivers"].append(rec)
            out["direct"].append(d)
        elif tag == "Image-Source_Data":
            per_rec = []
            for _ in range(R):
                r.i32()
                paths = []
                for _ in range(r.i32()):
                    kind = r.i16()
                    if kind != 0:
                        raise ValueError("compound image-source paths are not supported")
                    n = r.i32()
                    p = {"pts": r.f64s(3 * n).reshape(n, 3), "energy": r.f64s(8)}
                    p["filter"] = None
                    if r.boolean():
                        m = r.i32()
                        f = r.f64s(2 * m)
                        p["filter"] = f[0::2] + 1j * f[1::2]
                    p["prms"] = r.f64s(8)
                    p["time"] = r.f64()
                    p["seq"] = [r.i32() for _ in range(n - 2)]
                    paths.append(p)
                per_rec.append(paths)
            out["image_sources"].append(per_rec)
        elif tag == "Ray-Traced_Data":
            d = {"type": r.i32(), "sample_rate": r.i32()}
            n = r.i32()
            d["sample_ct"] = n
            d["cutoff_ms"] = r.f64()
            d["c"] = np.zeros(R)
            planes = np.zeros((8, R, 8, n))
            for q in range(R):
                d["c"][q] = r.f64()
                blk = r.f64s(n * 64).reshape(n, 8, 8)
                planes[:, q] = blk.transpose(2, 1, 0)
            d["planes"] = planes
            out["ray_traced"].append(d)
        else:
            raise ValueError(f"unsupported .pac1 section {tag!r}")
