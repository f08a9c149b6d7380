"""Throughput of the denoiser's training step on synthetic data.

    python -m pachydermacoustic_universal_amd.denoise_bench [--steps 50] [--batch 256]

Prints one JSON line: ms per AdamW step (forward + backward + update, batch
256, 8 x 3,000-bin inputs, 1,978,048 parameters), the GEMM FLOPs of a step
(2 M N K summed over the 41 products) and the resulting TFLOP/s against the
MI355X dense BF16 MFMA peak (2.5 PFLOP/s).
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import denoise as DN

BF16_PEAK_TFLOPS = 2500.0


def step_flops(cfg, batch):
    """GEMM FLOPs of one training step (forward 14 products, backward 27)."""
    P = cfg["bins"] // cfg["patch"]
    F = 8 * cfg["patch"]
    KE = (F + cfg["n_desc"] + 31) // 32 * 32
    D = cfg["width"]
    M = batch * P
    fwd = [(M, D, KE)] + [(M, D, 3 * D), (M, 2 * D, D), (M, D, 2 * D)] * cfg["blocks"] + [(M, F, D)]
    # each forward product has a data-gradient and a weight-gradient product of
    # the same size, except the first (no gradient for the input)
    bwd = [(M, D, KE)] + [(M, D, 3 * D), (M, 2 * D, D), (M, D, 2 * D)] * 2 * cfg["blocks"] + [(M, F, D)] * 2
    return sum(2 * m * n * k for m, n, k in fwd + bwd)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=256)
    args = ap.parse_args(argv)
    dn = DN.Denoiser(max_batch=args.batch)
    cfg = dn.cfg
    rng = np.random.default_rng(0)
    n = args.batch * 2
    x = rng.uniform(-2, 2, (n, 8, cfg["bins"])).astype(np.float32)
    t = (x + rng.normal(0, 0.05, x.shape)).astype(np.float32)
    desc = rng.normal(0, 1, (n, cfg["n_desc"])).astype(np.float32)
    dn.set_params(DN.init_params(cfg, seed=1))
    dn.set_dataset(x, desc, t)
    for k in range(args.warmup):
        dn.train_step(rng.choice(n, args.batch, replace=False), 1e-4)
    t0 = time.perf_counter()
    for k in range(args.steps):
        dn.train_step(rng.choice(n, args.batch, replace=False), 1e-4)
    dn.get_params()  # waits for the last step
    dt = (time.perf_counter() - t0) / args.steps
    fl = step_flops(cfg, args.batch)
    out = {"metric": "denoiser train step", "ms_per_step": dt * 1e3, "batch": args.batch, "params": dn.n_params,
           "gemm_flops_per_step": fl, "tflops": fl / dt / 1e12, "peak_tflops_bf16": BF16_PEAK_TFLOPS,
           "frac": fl / dt / 1e12 / BF16_PEAK_TFLOPS}
    print(json.dumps(out), flush=True)
    dn.close()


if __name__ == "__main__":
    main()
