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
denoised path's time can be set against the plain trace's.  A network that
takes the ray count (13 descriptors) denoises every ladder count; one that
does not denoises --rays only.

Time to JND accuracy: for each share q in --shares, the wall time of the
cheapest ladder trace (plain, or plain + network) whose T30, EDT and C80 each
have at least a share q of the (receiver, band) pairs inside 5 % / 5 % / 1 dB
of the reference; its ratio to the reference trace's wall time is what the
denoised path must bring under 1/20.
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
    ap.add_argument("--ladder", default="50000,100000,200000,400000",
                    help="further plain ray counts, for the plain trace's error against its time")
    ap.add_argument("--shares", default="0.9,0.95,1.0",
                    help="shares of (receiver, band) pairs inside the JND that define JND accuracy")
    ap.add_argument("--cutoff", type=float, default=3.0)
    ap.add_argument("--is-order", type=int, default=0,
                    help="IS_Order of every trace; with 0 the direct sound comes from Direct_Sound (ETCurve)")
    ap.add_argument("--weights", default=DN.DEFAULT_WEIGHTS)
    ap.add_argument("--ref-cache", default="", help="npz with a reference histogram to reuse (and to write)")
    ap.add_argument("--autotune", type=int, default=-1,
                    help="sample source rays of pach_autotune (-1: sized by the tuner, 0: no tuning)")
    ap.add_argument("--report", default="")
    ap.add_argument("--dump", default="", help="npz of the plain histograms of every ladder count")
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
    # C80 from the direct sound's arrival (Clarity's startTime), the same bin for every trace
    sb = dbin if args.is_order >= 0 else None
    P_ref = DN.decay_params(etc(E_ref), start_bin=sb)

    dn = DN.Denoiser.load(args.weights, max_batch=64)
    desc = DN.descriptors(room)
    takes_rays = dn.cfg["n_desc"] == desc.shape[1] + 1
    rep["weights_meta"] = json.loads(str(np.load(args.weights, allow_pickle=False)["meta"]))
    rep["plain"], rep["denoised"] = {}, {}
    dump = {}
    for rays in sorted(set([args.rays] + [int(float(r)) for r in args.ladder.split(",") if r])):
        E, wall, st = plain(rays, seed=1000 + rays)
        if args.dump:
            dump[f"E{rays}"] = E.astype(np.float32)
        rep["plain"][rays] = {"wall_s": wall, "errors": DN.summarize(DN.errors(DN.decay_params(etc(E), start_bin=sb), P_ref))}
        log(f"plain {rays}: {rep['plain'][rays]}")
        if rays == args.rays or takes_rays:
            t0 = time.perf_counter()
            E_dn = dn.denoise(E, desc, rays=rays)
            t_dn = time.perf_counter() - t0
            rep["denoised"][rays] = {"trace_wall_s": wall, "denoise_wall_s": t_dn, "wall_s": wall + t_dn,
                                     "time_fraction_of_ref": (wall + t_dn) / rep["ref"]["wall_s"],
                                     "errors": DN.summarize(DN.errors(DN.decay_params(etc(E_dn), start_bin=sb), P_ref))}
            log(f"denoised {rays}: {rep['denoised'][rays]}")
    dn.close()
    tr.close()
    if args.dump:
        np.savez_compressed(args.dump, E_ref=E_ref.astype(np.float32), dE=dE, dbin=dbin, desc=desc, **dump)
    rep["time_to_jnd"] = {}
    for q in [float(v) for v in args.shares.split(",") if v]:
        row = {}
        for kind in ("plain", "denoised"):
            ok = [(r, v["wall_s"]) for r, v in sorted(rep[kind].items())
                  if min(v["errors"][k]["within"] for k in ("t30", "edt", "c80")) >= q]
            row[kind] = ({"rays": ok[0][0], "wall_s": ok[0][1], "fraction_of_ref": ok[0][1] / rep["ref"]["wall_s"]}
                         if ok else None)
        rep["time_to_jnd"][q] = row
        log(f"time to JND accuracy at share {q}: {row}")
    out = json.dumps(rep, indent=1, default=str)
    print(out, flush=True)
    if args.report:
        with open(args.report, "w") as f:
            f.write(out)


if __name__ == "__main__":
    main()
