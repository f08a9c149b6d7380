"""Convergence mode (rayCount < 1, TaskAll:111-176, :288-329) on the GPU.

Every source ray keeps its global index as its Philox stream, so a
convergence-mode trace that stopped after N = 80 k rays must equal a plain
trace of N rays bin for bin (FP64), and replaying the check on the plain
traces of the last revolutions (N - 80 j rays) must first pass at N.
"""
import numpy as np
import pytest

from pachydermacoustic_universal_amd import _abi as A
from pachydermacoustic_universal_amd import rooms

pytestmark = pytest.mark.gpu


def room_with_obstacle():
    """Shoebox 20 x 12 x 6 m with a 2 x 4 x 5 m block between the source and
    receiver 1, so receiver 1 is occluded and receivers 0 (the farthest) and 2
    see the source."""
    box = rooms.shoebox(20, 12, 6, alpha=0.15, scat=0.3, source=(3, 6, 1.5),
                        receivers=[(16, 10.5, 1.5), (14, 6, 1.5), (9, 2, 1.5)], radius=0.5)
    block = rooms._box_tris((10.0, 4.0, 0.0), (12.0, 8.0, 5.0), bottom=False)
    tris = np.concatenate([box.verts, block])
    surf = np.concatenate([box.surf, 6 + np.arange(block.shape[0], dtype=np.int32) // 2])
    S = 6 + block.shape[0] // 2
    return rooms.Room(name="box_with_block", verts=tris, surf=surf, alpha=np.full((S, 8), 0.15),
                      scat=np.full((S, 8), 0.3), trans=np.zeros((S, 8)), source=box.source,
                      receivers=box.receivers, radius=0.5)


@pytest.mark.parametrize("kind,ray_count", [(A.CONV_MINIMUM, -1), (A.CONV_SPECMIN, 0)])
def test_convergence_mode_equals_plain_trace_and_stops_where_the_check_passes(kind, ray_count):
    room = room_with_obstacle()
    at, c, rho = A.uniform_medium()
    cut = 0.6
    tr = A.Tracer(0)
    try:
        print(f"[case] convergence kind={kind}", flush=True)
        st = A.trace_room(tr, room, rays=ray_count, cutoff_s=cut, seed=21, precision=A.PREC_FP64)
        assert st["converged"] == 1
        assert st["conv_receiver_visible"] == 0 and st["conv_receiver_occluded"] == 1
        N = st["rays_done"]
        assert N == 80 * st["revolutions"] and N >= 160
        conv = tr.planes(scaled=False)
        A.trace_room(tr, room, rays=N, cutoff_s=cut, seed=21, precision=A.PREC_FP64)
        plain = tr.planes(scaled=False)
        scale = np.abs(plain).max()
        assert np.all(np.abs(conv - plain) <= 1e-9 * np.abs(plain) + 1e-15 * scale)
        h = st["h_oct"]
        # replay both receivers' checks on plain traces of 80 k rays, k = 1 .. revolutions
        chks = {}
        for r in (0, 1):
            d = np.linalg.norm(room.receivers[r] - room.source) / c
            chks[r] = A.ConvergenceCheck(kind, conv.shape[-1], 1000, direct_s=d)
        passed = []
        for k in range(1, st["revolutions"] + 1):
            A.trace_room(tr, room, rays=80 * k, cutoff_s=cut, seed=21, precision=A.PREC_FP64)
            P = tr.planes(scaled=False)
            ok = [chks[r].step(P[0, r, h], P[8, r, h], P[9, r, h], P[10, r, h]) for r in (0, 1)]
            passed.append(all(ok))
        assert passed[-1] and not any(passed[:-1])
    finally:
        tr.close()
