"""Adjoint gradients of the explicit Allen-Cahn step through torch.autograd.

Loss J = 1/2 sum_i m_i (phi_N,i - target_i)^2 (lumped-mass L2 misfit at the final
step); gradients with respect to the mobility L, the gradient coefficient kappa and the
initial field phi0.  Checked against central finite differences of the GPU forward
(step 1e-6, relative error <= 1e-6) and against the FP64 CPU oracle's adjoint (<= 1e-10);
the FP32-forward variant is compared with the FP64 gradient.
"""

from __future__ import annotations

import numpy as np
import pytest

from phasefield_amd import Context, prm
from phasefield_amd.capi import AGE_LATEST, FP32, FP64
from tests.helpers import to_full, to_stored

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def setup(text, n, every, precision, target_L=1.2, target_kappa=1.3, seed=0):
    from tests.adjoint_autograd import allen_cahn_steps, lumped_mass

    ctx = Context("allen_cahn_3d_q1", text)
    ctx.run(0)  # the application's seeded initial condition
    info = ctx.info
    phi0 = torch.tensor(to_stored(ctx.get_field(0, AGE_LATEST).ravel(), info), device="cuda")
    m = lumped_mass(ctx)
    one = lambda v: torch.tensor(v, dtype=torch.float64, device="cuda")
    with torch.no_grad():
        target = allen_cahn_steps(phi0, one(target_L), one(target_kappa), ctx, n, every, FP64)

    def loss(p0, L, kappa, prec=precision):
        phiN = allen_cahn_steps(p0, L, kappa, ctx, n, every, prec)
        return 0.5 * torch.sum(m * (phiN - target) ** 2)

    return ctx, phi0, loss, one, m, target


def grads(loss, phi0, one, L=1.0, kappa=1.5, prec=FP64):
    p0 = phi0.clone().requires_grad_(True)
    Lt, kt = one(L).requires_grad_(True), one(kappa).requires_grad_(True)
    J = loss(p0, Lt, kt, prec)
    J.backward()
    return float(J.detach()), p0.grad, float(Lt.grad), float(kt.grad)


def fd_checks(loss, phi0, one, dphi, dL, dk, eps=1e-6, L=1.0, kappa=1.5, seed=3):
    with torch.no_grad():
        fL = (float(loss(phi0, one(L + eps), one(kappa))) -
              float(loss(phi0, one(L - eps), one(kappa)))) / (2 * eps)
        fk = (float(loss(phi0, one(L), one(kappa + eps))) -
              float(loss(phi0, one(L), one(kappa - eps)))) / (2 * eps)
        g = torch.Generator(device="cuda").manual_seed(seed)
        v = torch.randn(phi0.shape, dtype=torch.float64, device="cuda", generator=g)
        fv = (float(loss(phi0 + eps * v, one(L), one(kappa))) -
              float(loss(phi0 - eps * v, one(L), one(kappa)))) / (2 * eps)
        gv = float(torch.dot(dphi, v))
    return abs(dL - fL) / abs(fL), abs(dk - fk) / abs(fk), abs(gv - fv) / abs(fv)


def test_autograd_gradients_small(device):
    from oracle.binding import Oracle

    n, every = 20, 6
    text = prm.allen_cahn(dim=3, size=16.0, refinement=4, steps=n, dt=0.05, kappa=1.5,
                          n_circles=3)
    ctx, phi0, loss, one, m, target = setup(text, n, every, FP64)
    J, dphi, dL, dk = grads(loss, phi0, one)
    eL, ek, ev = fd_checks(loss, phi0, one, dphi, dL, dk)
    print(f"small: J={J:.6e} dJ/dL={dL:.6e} dJ/dkappa={dk:.6e} FD rel errors {eL:.2e} {ek:.2e} {ev:.2e}")
    assert max(eL, ek, ev) <= 1e-6
    # FP64 CPU reference gradient
    info = ctx.info
    orc = Oracle("allen_cahn_3d_q1", text)
    p0 = to_full(phi0.cpu().numpy(), info)
    with torch.no_grad():
        from tests.adjoint_autograd import allen_cahn_steps
        phiN = allen_cahn_steps(phi0, one(1.0), one(1.5), ctx, n, every, FP64)
    lam = (m * (phiN - target)).cpu().numpy()
    lam_full = to_full(lam, info)
    mfull = orc.lumped_mass()
    lam_full[mfull == 0] = 0.0  # image entries: the oracle folds them onto masters
    rdl, rdk, rg = orc.adjoint_vjp(p0, lam_full, 1.0, 1.5, n)
    assert abs(dL - rdl) <= 1e-10 * abs(rdl)
    assert abs(dk - rdk) <= 1e-10 * abs(rdk)
    assert np.abs(dphi.cpu().numpy() - to_stored(rg, info)).max() <= 1e-10 * np.abs(rg).max()


def test_autograd_gradients_full_config(device):
    """AC 3D adjoint: Q1, 128^3 cells on [0,128]^3, 500 steps of dt = 0.01, target from
    L = 1.2, kappa = 1.3, checkpoint every 50 steps, FD step 1e-6; FP64 and
    FP32-forward variants."""
    n, every = 500, 50
    text = prm.config_prm("allen_cahn_adjoint_3d")
    ctx, phi0, loss, one, m, target = setup(text, n, every, FP64)
    J, dphi, dL, dk = grads(loss, phi0, one)
    eL, ek, ev = fd_checks(loss, phi0, one, dphi, dL, dk)
    print(f"full: J={J:.6e} dJ/dL={dL:.6e} dJ/dkappa={dk:.6e} FD rel errors {eL:.2e} {ek:.2e} {ev:.2e}")
    assert max(eL, ek, ev) <= 1e-6
    J32, dphi32, dL32, dk32 = grads(loss, phi0, one, prec=FP32)
    e32 = (abs(dL32 - dL) / abs(dL), abs(dk32 - dk) / abs(dk),
           float((dphi32 - dphi).abs().max() / dphi.abs().max()))
    print(f"FP32 forward vs FP64: J rel {abs(J32 - J) / J:.2e}, gradient rel errors "
          f"{e32[0]:.2e} {e32[1]:.2e} {e32[2]:.2e}")
    assert max(e32) <= 1e-3
