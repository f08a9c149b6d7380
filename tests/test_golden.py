"""Golden fixtures (tests/golden/*.json, written by make_golden.py from the
oracle): the oracle must keep reproducing them; the GPU FP64 path must match
them bin by bin within 1e-5 relative."""
import glob
import json
import os

import numpy as np
import pytest

import pyoracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
import sys  # noqa: E402

sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as G  # noqa: E402

FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.json")))


def load(path):
    with open(path) as f:
        return json.load(f)


def dense(fx):
    E = np.zeros(fx["shape"])
    for r, b, s, v in fx["bins"]:
        E[r, b, s] = v
    return E


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p))
def test_oracle_reproduces_fixture(path):
    fx = load(path)
    pl, cnt = G.run(fx)
    for k, v in fx["counters"].items():
        assert cnt[k] == v, k
    np.testing.assert_allclose(pl[0], dense(fx), rtol=1e-12, atol=0)


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p))
def test_gpu_fp64_matches_fixture(path):
    from pachydermacoustic_universal_amd import _abi as A

    fx = load(path)
    print(f"[case] fixture {fx['name']}", flush=True)
    with A.Tracer(0) as tr:
        st = A.trace_room(tr, G.room_of(fx["room"]), rays=fx["rays"], cutoff_s=fx["cutoff_s"],
                          seed=fx["seed"], mode=fx["mode"], precision=A.PREC_FP64,
                          dir_mode=fx["dir_mode"], dirs=G.dirs_of(fx))
        E = tr.histogram(0, scaled=False)
    c = fx["counters"]
    assert st["bounces"] == c["bounces_broad"] + c["bounces_late"] + c["bounces_verylate"]
    assert st["deposits"] == c["deposits"]
    ref = dense(fx)
    bad = np.abs(E - ref) > 1e-5 * np.abs(ref) + 1e-12 * np.abs(ref).max()
    assert not bad.any(), bad.sum()
