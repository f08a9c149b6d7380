"""Build libpach.so (gfx950) and the CPU oracle.

    python -m pachydermacoustic_universal_amd.build [--oracle-only | --lib-only]

The HIP library is compiled with hipcc for gfx950 only.  Every HIP object is
compiled with packed-FP32 instructions disabled; the FP64 kernel
instantiation is additionally compiled with -ffp-contract=off.  Objects go to
pachydermacoustic_universal_amd/_build/, the library next to this file.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libpach.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = "gfx950"
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
          f"-I{os.path.join(ROOT, 'include')}"]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics"]

UNITS = [
    # (source, extra flags, is_hip)
    ("runtime.cpp", [], False),
    ("grid.cpp", [], False),
    ("direct.cpp", [], False),
    ("trace_f32.hip", [], True),
    ("trace_f64.hip", ["-ffp-contract=off"], True),
    ("trace_mix.hip", ["-ffp-contract=off"], True),
    ("image_source.hip", ["-ffp-contract=off"], True),
    ("filters.hip", [], True),
    ("denoise.hip", [], True),
]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{p.stdout}\n{p.stderr}")
    return p.stderr


def build_lib(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for src, extra, is_hip in UNITS:
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, "-c", os.path.join(CSRC, src), "-o", obj] + COMMON + extra
        if is_hip:
            cmd += HIP_FLAGS
        jobs.append((cmd, obj))
    with cf.ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        outs = list(ex.map(lambda j: _run(j[0]), jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o, file=sys.stderr)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + [o for _, o in jobs])
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    _run(["make", "-s", "-C", ORACLE_DIR])
    return os.path.join(ORACLE_DIR, "_build", "liboracle.so")


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if "--lib-only" not in argv:
        print(build_oracle())
    if "--oracle-only" not in argv:
        print(build_lib(verbose="-v" in argv))


if __name__ == "__main__":
    main()
