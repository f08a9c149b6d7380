"""Trains the histogram denoiser on low/high ray-count pairs the tracer makes.

    python -m pachydermacoustic_universal_amd.denoise_train [--rooms 96] [--steps 3000] ...

Training rooms are seeded hall variants (rooms.hall_variant: the polyhedral
hall generator with random proportions, ~2,000 triangles, 16 receivers);
validation rooms use seeds the training never sees.  Every room is traced in
packed batches of 16 rooms (one launch sequence per batch) at lo_rays
(`repeats` independent seeds, each a separate input) and once at hi_rays
(the target), 3 s cut-off, 1 ms bins, FP32 traversal, IS_Order 0 (the
direct sound is left to Direct_Sound and added back by the ETCurve wherever
decay parameters are compared).  The network trains
with AdamW (batch 256, linear warm-up, cosine decay, per-bin loss weights
1 + 20 exp(-t / 0.15 s)) through libpach.so's BF16 MFMA kernels, once per
learning rate in --lrs, each run keeping the parameters of its lowest
validation loss.  The run whose denoised validation traces have the smallest
T30 / EDT / C80 errors (in units of 5 %, 5 %, 1 dB) against the high-ray
traces is saved, with a JSON report of both runs and of the plain low-ray
traces' errors.
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

from . import _abi as A
from . import denoise as DN
from . import rooms as RM


def log(msg):
    print(f"[denoise-train] {msg}", flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--rooms", type=int, default=160)
    ap.add_argument("--val-rooms", type=int, default=16)
    ap.add_argument("--lo-rays", type=int, default=10000)
    ap.add_argument("--hi-rays", type=int, default=40000)
    ap.add_argument("--repeats", type=int, default=2)
    ap.add_argument("--cutoff", type=float, default=3.0)
    ap.add_argument("--is-order", type=int, default=0,
                    help="IS_Order of the traces: 0 leaves the direct sound to Direct_Sound, added by ETCurve")
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lrs", default="3e-4,1e-3", help="learning rates tried; the best on validation is kept")
    ap.add_argument("--weight-decay", type=float, default=0.0)
    ap.add_argument("--early-gain", type=float, default=20.0, help="loss weight 1 + gain exp(-t / tau)")
    ap.add_argument("--early-tau", type=float, default=0.15)
    ap.add_argument("--eval-every", type=int, default=250)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=DN.DEFAULT_WEIGHTS)
    ap.add_argument("--report", default="")
    args = ap.parse_args(argv)

    t_all = time.perf_counter()
    tr = A.Tracer(0)
    train_rooms = [RM.hall_variant(s) for s in range(args.rooms)]
    val_rooms = [RM.hall_variant(100_000 + s) for s in range(args.val_rooms)]
    log(f"{len(train_rooms)} training rooms, {len(val_rooms)} validation rooms")
    t0 = time.perf_counter()
    data = DN.make_pairs(tr, train_rooms, args.lo_rays, args.hi_rays, args.cutoff, args.repeats, seed=11,
                         log=log, is_order=args.is_order)
    val = DN.make_pairs(tr, val_rooms, args.lo_rays, args.hi_rays, args.cutoff, 1, seed=977, log=log,
                        is_order=args.is_order)
    t_gen = time.perf_counter() - t0
    log(f"pairs: {data['x'].shape[0]} training, {val['x'].shape[0]} validation samples in {t_gen:.0f} s")
    tr.close()

    dn = DN.Denoiser(max_batch=args.batch, weight_decay=args.weight_decay)
    log(f"network: {dn.n_params} parameters")
    w = DN.early_weights(dn.cfg["bins"], gain=args.early_gain, tau_s=args.early_tau)
    # decay parameters of energy-time curves: the traces plus, with IS_Order 0,
    # the deterministic direct sound (ETCurve)
    def etc(E):
        return DN.with_direct(E, val["dE"], val["dbin"]) if args.is_order >= 0 else E

    P_hi = DN.decay_params(etc(DN.denormalize(val["t"], val["S"])))
    plain = DN.summarize(DN.errors(DN.decay_params(etc(DN.denormalize(val["x"], val["S"]))), P_hi))

    def score(s):
        # mean errors in units of their just-noticeable differences
        return s["t30"]["mean"] / 0.05 + s["edt"]["mean"] / 0.05 + s["c80"]["mean"] / 1.0

    variants, best = [], None
    t0 = time.perf_counter()
    for lr in [float(v) for v in args.lrs.split(",") if v]:
        dn.set_params(DN.init_params(dn.cfg, seed=args.seed))
        hist, vhist = DN.train(dn, data, args.steps, batch=args.batch, lr=lr, seed=args.seed, log=log, val=val,
                               eval_every=args.eval_every, weights=w)
        out = dn.forward(val["x"], val["desc"])
        den = DN.summarize(DN.errors(DN.decay_params(etc(DN.denormalize(out, val["S"]))), P_hi))
        v = {"lr": lr, "score": score(den), "loss_first": float(hist[0]), "loss_last100": float(np.mean(hist[-100:])),
             "val_history": vhist, "val_network": float(np.mean(w * (out - val["t"]) ** 2)), "denoised": den}
        variants.append(v)
        log(f"lr {lr}: score {v['score']:.3f} (plain {score(plain):.3f}) {json.dumps(den)}")
        if best is None or v["score"] < best[0]["score"]:
            best = (v, dn.get_params())
    t_train = time.perf_counter() - t0
    dn.set_params(best[1])
    rep = {
        "rooms": args.rooms, "val_rooms": args.val_rooms, "lo_rays": args.lo_rays, "hi_rays": args.hi_rays,
        "repeats": args.repeats, "cutoff_s": args.cutoff, "is_order": args.is_order, "steps": args.steps,
        "batch": args.batch,
        "weight_decay": args.weight_decay, "early_gain": args.early_gain, "early_tau": args.early_tau,
        "params": dn.n_params, "train_samples": int(data["x"].shape[0]), "val_samples": int(val["x"].shape[0]),
        "seconds": {"pairs": t_gen, "train": t_train, "total": time.perf_counter() - t_all},
        "ms_per_step": 1e3 * t_train / max(args.steps * len(variants), 1),
        "val_identity_loss": float(np.mean(w * (val["x"] - val["t"]) ** 2)),
        "variants": variants, "chosen_lr": best[0]["lr"], "plain_score": score(plain),
        "val_plain_vs_hi": plain, "val_denoised_vs_hi": best[0]["denoised"],
    }
    log(json.dumps(rep))
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    dn.save(args.out, meta={k: rep[k] for k in ("rooms", "lo_rays", "hi_rays", "repeats", "steps", "chosen_lr",
                                                "early_gain", "early_tau", "is_order")})
    if args.report:
        with open(args.report, "w") as f:
            json.dump(rep, f, indent=1)
    dn.close()
    log(f"saved {args.out}")


if __name__ == "__main__":
    main()
