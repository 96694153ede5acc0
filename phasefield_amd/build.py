"""Builds the per-application HIP libraries (gfx950) of the matrix-free operator path.

Each application library is one translation unit: the application's equations and
solve-block declarations compiled together with the kernels and the C-ABI of
``include/ppf.h``.  Libraries are written in-tree to ``phasefield_amd/lib`` so they
travel with the repository snapshot.

Usage: ``python -m phasefield_amd.build [app ...]``.
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB_DIR = PKG / "lib"

# application name -> source file (relative to the package)
APPS = {
    "allen_cahn_2d_q1": "apps/allen_cahn/app_2d.hip",
    "allen_cahn_implicit_2d_q1": "apps/allen_cahn_implicit/app_2d.hip",
    "allen_cahn_3d_q1": "apps/allen_cahn/app_3d.hip",
    "cahn_hilliard_3d_q2": "apps/cahn_hilliard/app.hip",
    "cahn_hilliard_implicit_2d_q2": "apps/cahn_hilliard_implicit/app_2d.hip",
    "mechanics_3d_q2": "apps/mechanics/app_3d.hip",
    "poisson_2d_q2": "apps/poisson/app_2d.hip",
    "poisson_3d_q2": "apps/poisson/app_3d.hip",
    "precipitate_3d_q1": "apps/precipitate/app.hip",
    "switching_2d_q2": "apps/switching/app_2d.hip",
    "testbed_2d_q2": "apps/testbed/app_2d_q2.hip",
    "testbed_3d_q1": "apps/testbed/app_3d_q1.hip",
}

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -fno-access-control: application hooks are commonly declared private (as in the
#   reference's applications); the kernels call them by qualified name so that they are
#   inlined without any virtual dispatch on the device.
# -packed-fp32-ops: packed FP32 instructions are kept out of the generated code.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-shared",
    "-Xclang",
    "-fno-access-control",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def library_path(app: str) -> Path:
    return LIB_DIR / f"libppf_{app}.so"


def include_flags() -> list[str]:
    return [
        f"-I{PKG / 'include'}",
        f"-I{ROOT / 'include'}",
        f"-I{PKG / 'csrc'}",
    ]


def compile_app(app: str, verbose: bool = False) -> Path:
    src = PKG / APPS[app]
    out = library_path(app)
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    tmp = out.with_suffix(".so.tmp")
    cmd = [HIPCC, *HIP_FLAGS, *include_flags(), str(src), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"building {app} failed:\n{res.stderr[-6000:]}")
    os.replace(tmp, out)
    return out


def build(apps: list[str] | None = None, jobs: int | None = None, verbose: bool = False) -> list[Path]:
    """Compiles the given applications (all by default) in parallel."""
    apps = list(apps or APPS)
    unknown = [a for a in apps if a not in APPS]
    if unknown:
        raise ValueError(f"unknown applications: {unknown}")
    jobs = jobs or min(len(apps), max(1, (os.cpu_count() or 2) - 1), 8)
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        return list(pool.map(lambda a: compile_app(a, verbose), apps))


if __name__ == "__main__":
    names = sys.argv[1:] or None
    for path in build(names, verbose=True):
        print("built", path)
