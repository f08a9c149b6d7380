"""What the trained denoiser gains over the plain trace it starts from.

Two hall variants the training never saw (16 receivers each) are traced
plainly at 1e4 source rays and, as the reference, at 3e5 rays, with
IS_Order 0 as the network was trained: the direct sound comes from
Direct_Sound and is added to every energy-time curve (ETCurve).  The shipped
network (pachydermacoustic_universal_amd/weights/denoiser.npz) denoises the
1e4-ray histograms; its mean EDT error and mean C80 error against the
reference must be lower than the plain 1e4-ray trace's, and its T30 error no
worse than the plain trace's by more than 0.5 percentage points.
"""
import os

import numpy as np
import pytest

from pachydermacoustic_universal_amd import _abi as A
from pachydermacoustic_universal_amd import denoise as DN
from pachydermacoustic_universal_amd import rooms

pytestmark = pytest.mark.gpu


def test_denoiser_beats_the_plain_trace_it_starts_from():
    assert os.path.exists(DN.DEFAULT_WEIGHTS), "trained weights missing"
    dn = DN.Denoiser.load(max_batch=32)
    tr = A.Tracer(0)
    err_plain, err_dn = [], []
    try:
        for seed in (200_001, 200_002):
            room = rooms.hall_variant(seed)
            print(f"[case] denoiser gain {room.name}", flush=True)
            dE, dbin = DN.direct_part(room)
            etc = lambda E: DN.with_direct(E, dE, dbin)  # noqa: E731
            A.trace_room(tr, room, rays=300_000, cutoff_s=3.0, seed=1, is_order=0)
            E_ref = tr.histogram(0, scaled=True)
            A.trace_room(tr, room, rays=10_000, cutoff_s=3.0, seed=2, is_order=0)
            E_lo = tr.histogram(0, scaled=True)
            E_dn = dn.denoise(E_lo, DN.descriptors(room))
            P_ref = DN.decay_params(etc(E_ref))
            err_plain.append(DN.errors(DN.decay_params(etc(E_lo)), P_ref))
            err_dn.append(DN.errors(DN.decay_params(etc(E_dn)), P_ref))
    finally:
        tr.close()
        dn.close()
    mean = lambda errs, k: float(np.nanmean(np.concatenate([e[k].ravel() for e in errs])))
    rep = {k: (mean(err_plain, k), mean(err_dn, k)) for k in ("t30", "edt", "c80")}
    print("[gain] plain vs denoised mean errors:", rep, flush=True)
    assert rep["edt"][1] < rep["edt"][0]
    assert rep["c80"][1] < rep["c80"][0]
    assert rep["t30"][1] < rep["t30"][0] + 0.005
