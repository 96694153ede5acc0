"""ctypes binding of the CPU baseline of the Cahn-Hilliard increment
(oracle/baseline/ch_baseline.cc).  TEST AND BENCHMARK INFRASTRUCTURE: the benchmark's
`cpu_baseline` leg and the tests load it; the product never does."""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int
from pathlib import Path

import numpy as np

from oracle.build import baseline_path


def host_isa() -> str:
    """'avx512' if the host CPU has AVX-512F, else 'avx2'."""
    try:
        flags = Path("/proc/cpuinfo").read_text()
    except OSError:
        return "avx2"
    return "avx512" if " avx512f" in flags else "avx2"


_lib = {}


def library(variant: str | None = None):
    variant = variant or host_isa()
    if variant not in _lib:
        path = baseline_path(variant)
        if not path.exists():
            raise FileNotFoundError(f"{path} missing: run `python -m oracle.build`")
        lib = ctypes.CDLL(str(path))
        dp = POINTER(c_double)
        lib.ch_baseline_run.argtypes = [c_int, c_double, c_double, c_double, c_double, c_int,
                                        c_int, dp, dp, dp, dp]
        lib.ch_baseline_isa_avx512.argtypes = []
        _lib[variant] = lib
    return _lib[variant]


def run(c0: np.ndarray, n_cells: int, h: float, dt: float, mob: float, kappa: float,
        n_increments: int, threads: int = 0, variant: str | None = None):
    """(c_n, mu_n, seconds): n_increments Cahn-Hilliard increments from c0 (stored
    periodic lattice, (2 n_cells)^3 values, x fastest); seconds excludes increment 0."""
    lib = library(variant)
    c0 = np.ascontiguousarray(c0, dtype=np.float64)
    c, mu, sec = np.empty_like(c0), np.empty_like(c0), c_double()
    p = lambda a: a.ctypes.data_as(POINTER(c_double))
    st = lib.ch_baseline_run(n_cells, h, dt, mob, kappa, n_increments, threads, p(c0), p(c), p(mu),
                             ctypes.byref(sec))
    if st != 0:
        raise ValueError("ch_baseline_run: the cell count must be a positive multiple of 8")
    return c, mu, sec.value


def avx512(variant: str | None = None) -> bool:
    return bool(library(variant).ch_baseline_isa_avx512())
