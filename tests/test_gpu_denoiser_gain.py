"""What the trained denoiser gains over the plain trace it starts from.

A concert-hall variant and a hall variant the training never saw (64 and 16
receivers) are traced plainly at 2e4 source rays and, as the reference, at
3e5 rays, with IS_Order 0 as the network was trained: the direct sound comes
from Direct_Sound and is added to every energy-time curve (ETCurve), and C80
is measured from its arrival.  The shipped network
(pachydermacoustic_universal_amd/weights/denoiser.npz) denoises the 2e4-ray
histograms; its mean T30, EDT and C80 errors against the reference must be
lower than the plain 2e4-ray trace's.
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
    dn = DN.Denoiser.load(max_batch=64)
    tr = A.Tracer(0)
    err_plain, err_dn = [], []
    try:
        for room in (rooms.concert_variant(200_001), rooms.hall_variant(200_002)):
            print(f"[case] denoiser gain {room.name}", flush=True)
            dE, dbin = DN.direct_part(room)
            etc = lambda E: DN.with_direct(E, dE, dbin)  # noqa: E731
            A.trace_room(tr, room, rays=300_000, cutoff_s=3.0, seed=1, is_order=0)
            E_ref = tr.histogram(0, scaled=True)
            A.trace_room(tr, room, rays=20_000, cutoff_s=3.0, seed=2, is_order=0)
            E_lo = tr.histogram(0, scaled=True)
            E_dn = dn.denoise(E_lo, DN.descriptors(room), rays=20_000)
            P_ref = DN.decay_params(etc(E_ref), start_bin=dbin)
            err_plain.append(DN.errors(DN.decay_params(etc(E_lo), start_bin=dbin), P_ref))
            err_dn.append(DN.errors(DN.decay_params(etc(E_dn), start_bin=dbin), P_ref))
    finally:
        tr.close()
        dn.close()
    mean = lambda errs, k: float(np.nanmean(np.concatenate([e[k].ravel() for e in errs])))  # noqa: E731
    rep = {k: (mean(err_plain, k), mean(err_dn, k)) for k in ("t30", "edt", "c80")}
    print("[gain] plain vs denoised mean errors:", rep, flush=True)
    for k in ("t30", "edt", "c80"):
        assert rep[k][1] < rep[k][0], k
