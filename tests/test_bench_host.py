"""bench.py's host-side contract, on the CPU: more than one GPU is refused
with a message, the measured-traffic figure scales the profiled bytes per
octave bounce by a run's bounces per launch (per precision, from that
precision's committed profile), and the VALU-issue figure that names the
binding resource is given only for the profiled workload."""
import importlib.util
import json
import os
import subprocess
import sys
import types

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def bench_module():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def default_args(**kw):
    a = dict(workload="hall", cutoff=3.0, grid=None, precision="fp32", tune=[], autotune=0)
    a.update(kw)
    return types.SimpleNamespace(**a)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)


PROFILED = [p for p in ("fp32", "mixed") if os.path.exists(
    os.path.join(ROOT, "profiles", f"k_octave_{p}_hall_traffic.json"))]


@pytest.mark.parametrize("precision", PROFILED)
def test_traffic_and_valu_issue_follow_the_profile(precision):
    b = bench_module()
    prof = json.load(open(b.traffic_file(precision)))
    assert prof["precision"] == precision
    tot = dict(oct=4.0e10, n_oct=20, ms_oct=9000.0)
    args = default_args(precision=precision)
    traffic, note = b.measured_traffic(args, tot, prof["rays"])
    assert traffic == pytest.approx(prof["hbm_bytes_per_bounce"] * 4.0e10 / 20)
    assert "this workload" in note
    v = b.valu_issue(args, tot, prof["rays"])
    rate = 4.0e10 / 9.0
    assert v["resource"] == "valu_issue"
    assert v["frac"] == pytest.approx(prof["valu_instr_per_oct_bounce"] * 4 * rate / (1024 * 2.4e9))
    assert v["lane_frac"] == pytest.approx(v["frac"] * prof["valu_lane_utilisation"])
    assert v["instr_per_bounce"] == prof["valu_instr_per_oct_bounce"]


@pytest.mark.parametrize("change", [dict(grid=48), dict(tune=["walk_steps=8"]), dict(cutoff=2.0)])
@pytest.mark.parametrize("precision", PROFILED)
def test_valu_issue_only_for_the_profiled_workload(change, precision):
    b = bench_module()
    prof = json.load(open(b.traffic_file(precision)))
    tot = dict(oct=4.0e10, n_oct=20, ms_oct=9000.0)
    args = default_args(precision=precision, **change)
    assert b.valu_issue(args, tot, prof["rays"]) is None
    # the traffic figure is still given, labelled as measured on the profiled hall
    traffic, note = b.measured_traffic(args, tot, prof["rays"])
    assert traffic > 0 and "1e6-ray hall" in note


def test_no_figures_without_a_profile_of_that_precision_or_workload():
    b = bench_module()
    tot = dict(oct=4.0e10, n_oct=20, ms_oct=9000.0)
    assert b.measured_traffic(default_args(precision="fp64"), tot, 1_000_000) == (None, None)
    assert b.valu_issue(default_args(precision="fp64"), tot, 1_000_000) is None
    assert b.measured_traffic(default_args(workload="map"), tot, 1_000_000) == (None, None)
