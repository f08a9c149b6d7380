"""Denoised inference at scale on the procedural concert hall (50,120
triangles, 64 receivers over the audience plane; never seen in training).

Every trace uses IS_Order 0 and the direct sound of Direct_Sound is added to
each energy-time curve as IR_Construction.ETCurve does.  A plain 4e5-ray trace
is the reference here (its own mean errors against a 1e7-ray trace are about
0.2 % T30, 1 % EDT, 0.15 dB C80); plain and denoised 2e4- and 5e4-ray traces
are judged against it per (receiver, band): T30 and EDT relative, C80 in dB
measured from the direct sound's arrival (AcousticalMath.Clarity with its
startTime there).  The shipped network must lower the mean error of every
parameter and raise the share of pairs inside the just-noticeable
differences (5 %, 5 %, 1 dB) for EDT and C80 at both ray counts; the shares
and wall times are printed."""
import os
import time

import numpy as np
import pytest

from pachydermacoustic_universal_amd import _abi as A
from pachydermacoustic_universal_amd import denoise as DN
from pachydermacoustic_universal_amd import rooms

pytestmark = pytest.mark.gpu


def test_denoiser_gain_on_the_concert_hall():
    assert os.path.exists(DN.DEFAULT_WEIGHTS), "trained weights missing"
    room = rooms.concert_hall()
    at, c, rho = A.uniform_medium()
    dE, dbin = DN.direct_part(room)
    etc = lambda E: DN.with_direct(E, dE, dbin)  # noqa: E731
    desc = DN.descriptors(room)
    dn = DN.Denoiser.load(max_batch=64)
    tr = A.Tracer(0)
    walls = {}
    try:
        tr.upload_scene(room.verts, room.surf, room.alpha, room.scat, room.trans, (0, 0, 0))
        tr.set_source(room.source, np.ones(8), A.DIR_GEODESIC)
        tr.set_receivers(room.receivers, room.radius, at, c, c * rho, 1000, 3000)

        def trace(rays, seed):
            print(f"[case] concert hall {rays} rays", flush=True)
            t0 = time.perf_counter()
            tr.trace(rays=rays, cutoff_m=3.0 * c, seed=seed, is_order=0)
            walls[rays] = time.perf_counter() - t0
            return tr.histogram(0, scaled=True)

        P_ref = DN.decay_params(etc(trace(400_000, 1)), start_bin=dbin)
        rep = {}
        for rays in (20_000, 50_000):
            E = trace(rays, 1000 + rays)
            t0 = time.perf_counter()
            E_dn = dn.denoise(E, desc, rays=rays)
            t_dn = time.perf_counter() - t0
            rep[rays] = {k: DN.summarize(DN.errors(DN.decay_params(etc(X), start_bin=dbin), P_ref))
                         for k, X in (("plain", E), ("denoised", E_dn))}
            rep[rays]["wall_s"] = (walls[rays], walls[rays] + t_dn)
    finally:
        tr.close()
        dn.close()
    for rays, r in rep.items():
        print(f"[concert] {rays} rays, wall plain / denoised {r['wall_s'][0]:.2f} / {r['wall_s'][1]:.2f} s "
              f"(reference 4e5: {walls[400_000]:.2f} s)", flush=True)
        for k in ("t30", "edt", "c80"):
            p, d = r["plain"][k], r["denoised"][k]
            print(f"[concert]   {k}: mean {p['mean']:.4f} -> {d['mean']:.4f}, within {p['within']:.3f} -> "
                  f"{d['within']:.3f}", flush=True)
        for k in ("t30", "edt", "c80"):
            assert r["denoised"][k]["mean"] < r["plain"][k]["mean"], (rays, k)
        for k in ("edt", "c80"):
            assert r["denoised"][k]["within"] >= r["plain"][k]["within"], (rays, k)
