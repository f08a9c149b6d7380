"""Trains the histogram denoiser on low/high ray-count pairs the tracer makes.

    python -m pachydermacoustic_universal_amd.denoise_train [--rooms 96] [--steps 3000] ...

Training rooms are seeded concert-hall variants (rooms.concert_variant:
shell, seat blocks and balconies of seeded proportions, ~8,000-20,000
triangles, 64 receivers over the audience floor) and hall variants
(rooms.hall_variant: the polyhedral hall generator with random proportions,
~2,000 triangles, 16 receivers); validation rooms use seeds the training
never sees, and rooms.concert_hall is never among them.  Every room is traced
in packed batches (one launch sequence per batch) at each input ray count of
--lo (a count, optionally "xK" for K independent seeds; every trace is a
separate input) and once at hi_rays (the target), 3 s cut-off, 1 ms bins,
FP32 traversal, IS_Order 0 (the direct sound is left to Direct_Sound and
added back by the ETCurve wherever decay parameters are compared).  With
--ray-desc the input's ray count is a 13th descriptor, so one network serves
several ray counts.  The network trains
with AdamW (batch 256, linear warm-up, cosine decay, per-bin loss weights
1 + gain exp(-t / tau), by default 1 + 20 exp(-t / 0.15 s)) through
libpach.so's BF16 MFMA kernels, once per variant in --lrs (learning rate and
optionally the loss weights), each run keeping the parameters of its lowest
validation loss.  The run whose denoised validation traces have the smallest
T30 / EDT / C80 errors (in units of 5 %, 5 %, 1 dB; C80 from the direct
sound's arrival) against the high-ray traces is saved, with a JSON report of
every run and of the plain low-ray traces' errors.
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
    ap.add_argument("--rooms", type=int, default=32, help="hall variants")
    ap.add_argument("--concert-rooms", type=int, default=64, help="concert-hall variants")
    ap.add_argument("--val-rooms", type=int, default=8, help="validation rooms of each kind")
    ap.add_argument("--lo", default="20000x2,50000", help="input ray counts, each optionally xK seeds")
    ap.add_argument("--hi-rays", type=int, default=200000)
    ap.add_argument("--ray-desc", type=int, default=1, help="1: the input's ray count is a 13th descriptor")
    ap.add_argument("--cutoff", type=float, default=3.0)
    ap.add_argument("--is-order", type=int, default=0,
                    help="IS_Order of the traces: 0 leaves the direct sound to Direct_Sound, added by ETCurve")
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lrs", default="3e-4,1e-3",
                    help="variants tried, each lr[:gain[:tau]] (loss-weight gain and tau default to --early-gain "
                         "and --early-tau); the best on validation is kept")
    ap.add_argument("--weight-decay", type=float, default=0.0)
    ap.add_argument("--early-gain", type=float, default=20.0, help="loss weight 1 + gain exp(-t / tau)")
    ap.add_argument("--early-tau", type=float, default=0.15)
    ap.add_argument("--eval-every", type=int, default=250)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=DN.DEFAULT_WEIGHTS)
    ap.add_argument("--report", default="")
    args = ap.parse_args(argv)

    t_all = time.perf_counter()
    lo = []
    for item in args.lo.split(","):
        n, _, k = item.partition("x")
        lo += [int(float(n))] * (int(k) if k else 1)
    rd = bool(args.ray_desc)
    tr = A.Tracer(0)
    ch = [RM.concert_variant(s) for s in range(args.concert_rooms)]
    hv = [RM.hall_variant(s) for s in range(args.rooms)]
    vch = [RM.concert_variant(100_000 + s) for s in range(args.val_rooms)]
    vhv = [RM.hall_variant(100_000 + s) for s in range(args.val_rooms)]
    log(f"{len(ch)} concert-hall and {len(hv)} hall variants, {len(vch) + len(vhv)} validation rooms; "
        f"inputs {lo}, targets {args.hi_rays} rays")
    t0 = time.perf_counter()
    sets, vsets = [], []
    kw = dict(cutoff_s=args.cutoff, log=log, is_order=args.is_order, ray_desc=rd)
    if ch:
        sets.append(DN.make_pairs(tr, ch, lo, args.hi_rays, seed=11, pack=8, **kw))
    if hv:
        sets.append(DN.make_pairs(tr, hv, lo, args.hi_rays, seed=13, pack=16, room_ids=range(10_000, 10_000 + len(hv)),
                                  **kw))
    if vch:
        vsets.append(DN.make_pairs(tr, vch, lo[0], args.hi_rays, seed=977, pack=8, **kw))
    if vhv:
        vsets.append(DN.make_pairs(tr, vhv, lo[0], args.hi_rays, seed=979, pack=16,
                                   room_ids=range(10_000, 10_000 + len(vhv)), **kw))
    data, val = DN.merge_pairs(*sets), DN.merge_pairs(*vsets)
    t_gen = time.perf_counter() - t0
    log(f"pairs: {data['x'].shape[0]} training, {val['x'].shape[0]} validation samples in {t_gen:.0f} s")
    tr.close()

    dn = DN.Denoiser(max_batch=args.batch, weight_decay=args.weight_decay, n_desc=13 if rd else 12)
    log(f"network: {dn.n_params} parameters")
    # decay parameters of energy-time curves: the traces plus, with IS_Order 0,
    # the deterministic direct sound (ETCurve)
    def etc(E):
        return DN.with_direct(E, val["dE"], val["dbin"]) if args.is_order >= 0 else E

    # C80 from the direct sound's arrival (Clarity's startTime)
    sb = val["dbin"] if args.is_order >= 0 else None
    P_hi = DN.decay_params(etc(DN.denormalize(val["t"], val["S"])), start_bin=sb)
    plain = DN.summarize(DN.errors(DN.decay_params(etc(DN.denormalize(val["x"], val["S"])), start_bin=sb), P_hi))

    def score(s):
        # mean errors in units of their just-noticeable differences
        return s["t30"]["mean"] / 0.05 + s["edt"]["mean"] / 0.05 + s["c80"]["mean"] / 1.0

    variants, best = [], None
    t0 = time.perf_counter()
    for spec in [v for v in args.lrs.split(",") if v]:
        f = [float(u) for u in spec.split(":")]
        lr, gain, tau = f[0], f[1] if len(f) > 1 else args.early_gain, f[2] if len(f) > 2 else args.early_tau
        w = DN.early_weights(dn.cfg["bins"], gain=gain, tau_s=tau)
        dn.set_params(DN.init_params(dn.cfg, seed=args.seed))
        hist, vhist = DN.train(dn, data, args.steps, batch=args.batch, lr=lr, seed=args.seed, log=log, val=val,
                               eval_every=args.eval_every, weights=w)
        out = dn.forward(val["x"], val["desc"])
        den = DN.summarize(DN.errors(DN.decay_params(etc(DN.denormalize(out, val["S"])), start_bin=sb), P_hi))
        v = {"lr": lr, "early_gain": gain, "early_tau": tau, "score": score(den), "loss_first": float(hist[0]), "loss_last100": float(np.mean(hist[-100:])),
             "val_history": vhist, "val_network": float(np.mean(w * (out - val["t"]) ** 2)), "denoised": den}
        variants.append(v)
        log(f"lr {lr}, loss weights 1 + {gain} exp(-t / {tau}): score {v['score']:.3f} (plain {score(plain):.3f}) {json.dumps(den)}")
        if best is None or v["score"] < best[0]["score"]:
            best = (v, dn.get_params(), w)
    t_train = time.perf_counter() - t0
    dn.set_params(best[1])
    w = best[2]
    rep = {
        "rooms": args.rooms, "concert_rooms": args.concert_rooms, "val_rooms": args.val_rooms, "lo_rays": lo,
        "hi_rays": args.hi_rays, "ray_desc": rd, "cutoff_s": args.cutoff, "is_order": args.is_order,
        "steps": args.steps,
        "batch": args.batch,
        "weight_decay": args.weight_decay, "early_gain": best[0]["early_gain"], "early_tau": best[0]["early_tau"],
        "params": dn.n_params, "train_samples": int(data["x"].shape[0]), "val_samples": int(val["x"].shape[0]),
        "seconds": {"pairs": t_gen, "train": t_train, "total": time.perf_counter() - t_all},
        "ms_per_step": 1e3 * t_train / max(args.steps * len(variants), 1),
        "val_identity_loss": float(np.mean(w * (val["x"] - val["t"]) ** 2)),
        "variants": variants, "chosen_lr": best[0]["lr"], "plain_score": score(plain),
        "val_plain_vs_hi": plain, "val_denoised_vs_hi": best[0]["denoised"],
    }
    log(json.dumps(rep))
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    dn.save(args.out, meta={k: rep[k] for k in ("rooms", "concert_rooms", "lo_rays", "hi_rays", "ray_desc",
                                                "steps", "chosen_lr", "early_gain", "early_tau", "is_order")})
    if args.report:
        with open(args.report, "w") as f:
            json.dump(rep, f, indent=1)
    dn.close()
    log(f"saved {args.out}")


if __name__ == "__main__":
    main()
