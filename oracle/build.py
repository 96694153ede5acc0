"""Builds the CPU oracle libraries (one per application) with g++ and OpenMP.

TEST INFRASTRUCTURE: the oracle is the FP64 CPU restatement of the reference's
algorithm that the HIP path is checked against; only the tests, the smoke check and the
benchmark's CPU baseline load it.  Outputs go to ``oracle/build/liboracle_<app>.so``.

The reference itself (PRISMS-PF on deal.II >= 9.6 with p4est and MPI) cannot be built
here, so there is no ``oracle/_ref`` build: the oracle is pinned by the checks in
``tests/test_oracle.py`` instead.
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
BUILD = HERE / "build"

APPS = [
    "allen_cahn_2d_q1",
    "allen_cahn_implicit_2d_q1",
    "allen_cahn_3d_q1",
    "cahn_hilliard_3d_q2",
    "cahn_hilliard_implicit_2d_q2",
    "mechanics_3d_q2",
    "poisson_2d_q2",
    "poisson_3d_q2",
    "precipitate_3d_q1",
    "switching_2d_q2",
    "testbed_2d_q2",
    "testbed_3d_q1",
]

CXX = os.environ.get("CXX", "g++")
FLAGS = ["-O3", "-march=x86-64-v3", "-std=c++20", "-fopenmp", "-fPIC", "-shared"]


def library_path(app: str) -> Path:
    return BUILD / f"liboracle_{app}.so"


def compile_app(app: str) -> Path:
    BUILD.mkdir(exist_ok=True)
    out = library_path(app)
    tmp = out.with_suffix(".so.tmp")
    cmd = [
        CXX, *FLAGS,
        f"-I{HERE / 'include'}", f"-I{ROOT / 'phasefield_amd' / 'include'}",
        str(HERE / "apps" / f"{app}.cc"), "-o", str(tmp),
    ]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build of {app} failed:\n{res.stderr[-4000:]}")
    os.replace(tmp, out)
    return out


# CPU baseline of the Cahn-Hilliard increment (baseline/ch_baseline.cc): one build per
# vector ISA; oracle/baseline.py loads the one the host supports.
BASELINE_VARIANTS = {"avx512": "-march=x86-64-v4", "avx2": "-march=x86-64-v3"}


def baseline_path(variant: str) -> Path:
    return BUILD / f"libch_baseline_{variant}.so"


def compile_baseline(variant: str) -> Path:
    BUILD.mkdir(exist_ok=True)
    out = baseline_path(variant)
    tmp = out.with_suffix(".so.tmp")
    cmd = [CXX, "-O3", BASELINE_VARIANTS[variant], "-std=c++20", "-fopenmp", "-fPIC", "-shared",
           str(HERE / "baseline" / "ch_baseline.cc"), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"baseline build ({variant}) failed:\n{res.stderr[-4000:]}")
    os.replace(tmp, out)
    return out


def build(apps: list[str] | None = None) -> list[Path]:
    apps = apps or APPS
    with ThreadPoolExecutor(max_workers=min(8, len(apps))) as pool:
        out = list(pool.map(compile_app, apps))
    out += [compile_baseline(v) for v in BASELINE_VARIANTS]
    return out


if __name__ == "__main__":
    for p in build(sys.argv[1:] or None):
        print("built", p)
