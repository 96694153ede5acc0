"""Committed golden fixtures (tests/golden, written by tests/golden/make_golden.py).

CPU: the oracle reproduces every fixture (a regression pin of the checked oracle).
GPU: the HIP path through the C-ABI reproduces the same fixtures, within 1e-10 relative
for explicit runs and adjoint quantities and 1e-8 for linear solves (the solve stops
at an absolute residual of 1e-10)."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

HERE = Path(__file__).resolve().parent / "golden"
META = json.loads((HERE / "fixtures.json").read_text())
RUNS = [k for k in META if not k.endswith("_adjoint")]


def rel(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def fixture(name):
    return np.load(HERE / f"{name}.npz", allow_pickle=False)


@pytest.mark.parametrize("name", RUNS)
def test_oracle_reproduces_fixture(oracle_libs, name):
    from oracle.binding import Oracle

    m, ref = META[name], fixture(name)
    o = Oracle(m["app"], m["prm"], threads=2)
    assert o.run(m["increments"]) == int(ref["status"])
    for f in range(o.nfields):
        assert rel(o.get_field(f, -1), ref[f"field_{f}"]) < 1e-12, f"field {f}"


def test_oracle_reproduces_adjoint_fixture(oracle_libs):
    from oracle.binding import Oracle

    m, ref = META["allen_cahn_2d_q1_adjoint"], fixture("allen_cahn_2d_q1_adjoint")
    o = Oracle(m["app"], m["prm"], threads=2)
    final = o.adjoint_forward(ref["phi0"], m["L"], m["kappa"], m["steps"])
    assert rel(final, ref["final"]) < 1e-13
    dL, dk, g = o.adjoint_vjp(ref["phi0"], ref["lam"], m["L"], m["kappa"], m["steps"])
    assert abs(dL - float(ref["dL"])) < 1e-12 * abs(float(ref["dL"]))
    assert abs(dk - float(ref["dkappa"])) < 1e-12 * abs(float(ref["dkappa"]))
    assert rel(g, ref["dphi0"]) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", RUNS)
def test_gpu_reproduces_fixture(device, name):
    from phasefield_amd import Context

    m, ref = META[name], fixture(name)
    ctx = Context(m["app"], m["prm"])
    assert ctx.run(m["increments"]) == int(ref["status"])
    tol = 1e-8 if name.startswith("poisson") else 1e-10
    for f in range(ctx.info.n_fields):
        assert rel(ctx.get_field(f, -1), ref[f"field_{f}"]) < tol, f"field {f}"


@pytest.mark.gpu
def test_gpu_reproduces_adjoint_fixture(device):
    import torch

    from phasefield_amd import Context
    from phasefield_amd.capi import FP64
    from tests.helpers import to_full, to_stored

    m, ref = META["allen_cahn_2d_q1_adjoint"], fixture("allen_cahn_2d_q1_adjoint")
    ctx = Context(m["app"], m["prm"])
    info = ctx.info
    d_phi0 = torch.tensor(to_stored(ref["phi0"], info), device="cuda")
    d_out = torch.zeros_like(d_phi0)
    ctx.adjoint_forward(d_phi0.data_ptr(), m["L"], m["kappa"], m["steps"], FP64, d_out.data_ptr())
    assert rel(to_full(d_out.cpu().numpy(), info), ref["final"]) < 1e-12
    d_lam = torch.tensor(to_stored(ref["lam"], info), device="cuda")
    d_g = torch.zeros_like(d_phi0)
    dL, dk = ctx.adjoint_vjp(d_phi0.data_ptr(), d_lam.data_ptr(), m["L"], m["kappa"],
                             m["steps"], 4, FP64, d_g.data_ptr())
    assert abs(dL - float(ref["dL"])) < 1e-10 * abs(float(ref["dL"]))
    assert abs(dk - float(ref["dkappa"])) < 1e-10 * abs(float(ref["dkappa"]))
    assert rel(d_g.cpu().numpy(), to_stored(ref["dphi0"], info)) < 1e-10
