"""Denoised inference at scale: the procedural concert hall (~50,000
triangles, 64 receivers) traced plainly at 1e7 rays (the reference), and at a
ladder of smaller ray counts with and without the trained denoiser.

    python -m pachydermacoustic_universal_amd.denoise_eval [--ref-rays 10000000] [--rays 20000]

By default every trace is made with IS_Order 0 and the deterministic direct
sound (Direct_Sound) is added to each energy-time curve as
IR_Construction.ETCurve does, before the decay parameters are compared.
Reports per (receiver, band) T30 / EDT relative errors and C80 absolute
errors (dB) against the reference, the share inside the just-noticeable
differences (5 %, 5 %, 1 dB), and wall times: the plain reference trace, the
low-ray trace, the denoiser (normalisation, network, inverse), so the
denoised path's time can be set against the plain trace's.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import _abi as A
from . import denoise as DN
from . import rooms as RM


def log(msg):
    print(f"[denoise-eval] {msg}", flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref-rays", type=int, default=10_000_000)
    ap.add_argument("--rays", type=int, default=20_000)
    ap.add_argument("--ladder", default="100000,400000",
                    help="further plain ray counts, for the plain trace's error against its time")
    ap.add_argument("--cutoff", type=float, default=3.0)
    ap.add_argument("--is-order", type=int, default=0,
                    help="IS_Order of every trace; with 0 the direct sound comes from Direct_Sound (ETCurve)")
    ap.add_argument("--weights", default=DN.DEFAULT_WEIGHTS)
    ap.add_argument("--ref-cache", default="", help="npz with a reference histogram to reuse (and to write)")
    ap.add_argument("--autotune", type=int, default=-1,
                    help="sample source rays of pach_autotune (-1: sized by the tuner, 0: no tuning)")
    ap.add_argument("--report", default="")
    args = ap.parse_args(argv)

    room = RM.concert_hall()
    at, c, rho = A.uniform_medium()
    tr = A.Tracer(0)
    sct = int(round(args.cutoff * 1000))
    tr.upload_scene(room.verts, room.surf, room.alpha, room.scat, room.trans, (0, 0, 0))
    tr.set_source(room.source, np.ones(8), A.DIR_GEODESIC)
    tr.set_receivers(room.receivers, room.radius, at, c, c * rho, 1000, sct)
    log(f"concert hall: {room.n_tris} triangles, {room.n_surfaces} polygons, {len(room.receivers)} receivers")
    tuning = None
    if args.autotune:
        t0 = time.perf_counter()
        tuning = tr.autotune(rays=args.ref_rays, cutoff_m=args.cutoff * c, sample_rays=args.autotune, seed=99)
        log(f"autotune ({time.perf_counter() - t0:.1f} s): {tuning}")
    rep = {"room": {"triangles": room.n_tris, "polygons": room.n_surfaces, "receivers": len(room.receivers),
                    "volume_m3": room.volume}, "cutoff_s": args.cutoff, "is_order": args.is_order, "tuning": tuning}
    dE, dbin = DN.direct_part(room)
    if args.is_order >= 0:
        rep["direct_sound"] = {"audible": int((dE.sum(1) > 0).sum()), "of": len(room.receivers)}

    def etc(E):
        # energy-time curves: with IS_Order >= 0 the direct sound is added (ETCurve)
        return DN.with_direct(E, dE, dbin) if args.is_order >= 0 else E

    def plain(rays, seed):
        wall = DN.trace_with_progress(tr, log, rays=rays, cutoff_m=args.cutoff * c, seed=seed,
                                      is_order=args.is_order)
        st = tr.stats()
        return tr.histogram(0, scaled=True), wall, st

    E_ref = None
    if args.ref_cache:
        try:
            z = np.load(args.ref_cache, allow_pickle=False)
            if int(z["rays"]) == args.ref_rays and int(z["is_order"]) == args.is_order:
                E_ref, rep["ref"] = z["E"], json.loads(str(z["info"]))
                log(f"reference from {args.ref_cache}")
        except (OSError, KeyError):
            pass
    if E_ref is None:
        log(f"reference: {args.ref_rays} rays")
        E_ref, wall, st = plain(args.ref_rays, seed=1)
        rep["ref"] = {"rays": args.ref_rays, "wall_s": wall, "bounces": st["bounces"],
                      "bounces_per_s": st["bounces"] / wall}
        if args.ref_cache:
            np.savez_compressed(args.ref_cache, E=E_ref.astype(np.float32), rays=args.ref_rays,
                                is_order=args.is_order, info=json.dumps(rep["ref"]))
    log(f"reference: {rep['ref']}")
    P_ref = DN.decay_params(etc(E_ref))

    dn = DN.Denoiser.load(args.weights, max_batch=64)
    desc = DN.descriptors(room)
    rep["plain"], rep["denoised"] = {}, {}
    for rays in [args.rays] + [int(r) for r in args.ladder.split(",") if r]:
        E, wall, st = plain(rays, seed=1000 + rays)
        rep["plain"][rays] = {"wall_s": wall, "errors": DN.summarize(DN.errors(DN.decay_params(etc(E)), P_ref))}
        log(f"plain {rays}: {rep['plain'][rays]}")
        if rays == args.rays:
            t0 = time.perf_counter()
            E_dn = dn.denoise(E, desc)
            t_dn = time.perf_counter() - t0
            rep["denoised"] = {"rays": rays, "trace_wall_s": wall, "denoise_wall_s": t_dn,
                               "total_wall_s": wall + t_dn,
                               "time_fraction_of_ref": (wall + t_dn) / rep["ref"]["wall_s"],
                               "errors": DN.summarize(DN.errors(DN.decay_params(etc(E_dn)), P_ref))}
            log(f"denoised {rays}: {rep['denoised']}")
    dn.close()
    tr.close()
    out = json.dumps(rep, indent=1, default=str)
    print(out, flush=True)
    if args.report:
        with open(args.report, "w") as f:
            f.write(out)


if __name__ == "__main__":
    main()
