"""Pins the CPU oracle (the FP64 restatement the HIP path is checked against).

No reference test covers the matrix-free operator, the explicit update, CG or
multigrid, and the reference cannot be built here (it needs deal.II, p4est and MPI).
Each oracle function is therefore checked against something that fixes it
independently:

* brute-force global sparse assembly with explicit GLL quadrature sums
  (tests/bruteforce.py) on tiny meshes, Q1/Q2, 2-D/3-D, periodic and Dirichlet;
* closed forms: the Q1 lumped operator is the FD Laplacian, Q2 lumped masses are
  h^3 {1,2,4,8}/27, the 1-D GLL derivative matrix;
* exactness on polynomials of the discrete space (-Laplacian of tensor quadratics,
  prolongation of coarse quadratics);
* invariants: symmetry, positivity, constant null space, CH mass conservation, AC
  energy decay, restriction = prolongation^T;
* scipy's CG iteration count on the assembled matrix (same stopping rule, +-1);
* central finite differences for the adjoint gradients.
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from phasefield_amd import prm
from tests import bruteforce as bf

RHS, LHS, LHS_PLAIN = 0, 1, 2
NONE, INVM, INVM_SQRT = 0, 1, 2


def oracle(app, text):
    from oracle.binding import Oracle

    return Oracle(app, text, threads=2)


def rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


# --------------------------------------------------------------------------- basis
def test_gll_tables_closed_form():
    xs, D, W = bf.lagrange_tables(2)
    assert np.allclose(xs, [0.0, 0.5, 1.0], atol=1e-15)
    assert np.allclose(D, [[-3, 4, -1], [-1, 0, 1], [1, -4, 3]], atol=1e-13)
    assert np.allclose(W, [1 / 6, 2 / 3, 1 / 6], atol=1e-15)
    xs, D, W = bf.lagrange_tables(1)
    assert np.allclose(D, [[-1, 1], [-1, 1]]) and np.allclose(W, [0.5, 0.5])


def test_device_gll_header_matches_bruteforce(tmp_path):
    """The tables compiled into the kernels (gll_basis.h) equal the derived ones."""
    import subprocess
    from pathlib import Path

    root = Path(__file__).resolve().parent.parent
    src = tmp_path / "gll.cc"
    src.write_text(
        "#include <prismspf/core/gll_basis.h>\n#include <cstdio>\n"
        "template<int p> void dump(){using B=prismspf::GLLBasis<p>;"
        "for(int i=0;i<=p;++i){std::printf(\"%.17g\\n\",B::w[i]);}"
        "for(int i=0;i<=p;++i)for(int j=0;j<=p;++j)std::printf(\"%.17g\\n\",B::D[i][j]);"
        "for(int i=0;i<=p;++i)for(int j=0;j<=p;++j)std::printf(\"%.17g\\n\",B::K[i][j]);}\n"
        "int main(){dump<1>();dump<2>();}\n")
    exe = tmp_path / "gll"
    subprocess.run(["g++", "-std=c++20", f"-I{root / 'phasefield_amd' / 'include'}",
                    str(src), "-o", str(exe)], check=True)
    vals = [float(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                             check=True).stdout.split()]
    pos = 0
    for p in (1, 2):
        n = p + 1
        xs, D, W = bf.lagrange_tables(p)
        K = np.einsum("q,qj,qk->jk", W, D, D)
        w_h = np.array(vals[pos:pos + n]); pos += n
        D_h = np.array(vals[pos:pos + n * n]).reshape(n, n); pos += n * n
        K_h = np.array(vals[pos:pos + n * n]).reshape(n, n); pos += n * n
        assert np.allclose(w_h, W, atol=1e-15)
        assert np.allclose(D_h, D, atol=1e-13)
        assert np.allclose(K_h, K, atol=1e-13)


# ----------------------------------------------------------------- closed forms
def test_q2_lumped_mass_closed_form(oracle_libs):
    size, ref = 8.0, 1
    o = oracle("cahn_hilliard_3d_q2", prm.cahn_hilliard(size=size, refinement=ref, steps=1))
    h = size / 2 ** ref
    n = 2 * 2 ** ref + 1
    m = o.lumped_mass().reshape(n, n, n)
    I = np.indices((n, n, n))
    odd = (I % 2 == 1).sum(axis=0)
    masters = (I < n - 1).all(axis=0)
    expected = h ** 3 / 27.0 * 2.0 ** odd
    assert np.allclose(m[masters], expected[masters], rtol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_q1_explicit_update_is_fd_laplacian(oracle_libs, dim):
    """Q1 + GLL lumping: M^-1 K is the (2d+1)-point FD Laplacian / h^2 (periodic)."""
    size, ref, dt, L, kappa = 10.0, 2, 0.01, 1.0, 1.5
    app = f"allen_cahn_{dim}d_q1"
    o = oracle(app, prm.allen_cahn(dim=dim, size=size, refinement=ref, dt=dt, steps=1,
                                   L=L, kappa=kappa))
    cells = 2 ** ref
    h = size / cells
    rng = np.random.default_rng(1)
    u = rng.uniform(-1, 1, size=(cells,) * dim)
    full = np.pad(u, [(0, 1)] * dim, mode="wrap")
    o.set_field(0, full.ravel(), age=1)
    got = o.apply_operator(1, RHS, INVM, None).reshape((cells + 1,) * dim)
    lap = sum(np.roll(u, 1, axis=a) - 2 * u + np.roll(u, -1, axis=a) for a in range(dim)) / h ** 2
    expected = u - dt * L * (u ** 3 - u) + dt * L * kappa * lap
    got = got[(slice(0, cells),) * dim]
    assert rel(got, expected) < 1e-13


# ------------------------------------------------------- brute-force assembly
def poisson_mesh(dim, ref):
    return bf.Mesh(dim, 2, [2 ** ref] * dim, [1.0] * dim, [False] * dim,
                   dirichlet_faces=range(2 * dim))


@pytest.mark.parametrize("dim,ref", [(2, 2), (3, 1)])
def test_poisson_operator_vs_bruteforce(oracle_libs, dim, ref):
    o = oracle(f"poisson_{dim}d_q2", prm.poisson(dim=dim, refinement=ref, mg_depth=1))
    mesh = poisson_mesh(dim, ref)
    K, mass = bf.assemble(mesh)
    rng = np.random.default_rng(7 + dim)
    x = rng.standard_normal(mesh.ndofs)
    f = rng.standard_normal(mesh.ndofs)
    o.set_field(1, f)
    # lhs with constrained reads (Dirichlet src -> 0) and plain reads
    exp_c = bf.scatter_full(mesh, K @ bf.read_unique(mesh, x, True))
    exp_p = bf.scatter_full(mesh, K @ bf.read_unique(mesh, x, False))
    assert rel(o.apply_operator(1, LHS, NONE, x), exp_c) < 1e-12
    assert rel(o.apply_operator(1, LHS_PLAIN, NONE, x), exp_p) < 1e-12
    # scalings: invm and invm^(1/2) of the lumped mass (generic = periodic-only constraints)
    mfull = bf.scatter_full(mesh, mass, dirichlet_zero=False)
    with np.errstate(divide="ignore", invalid="ignore"):
        exp_i = np.where(mfull > 0, exp_c / mfull, 0.0)
        exp_s = np.where(mfull > 0, exp_c / np.sqrt(mfull), 0.0)
    assert rel(o.apply_operator(1, LHS, INVM, x), exp_i) < 1e-12
    assert rel(o.apply_operator(1, LHS, INVM_SQRT, x), exp_s) < 1e-12
    # rhs: b = M f, Dirichlet rows dropped
    assert rel(o.apply_operator(1, RHS, NONE, None), bf.scatter_full(mesh, mass * f)) < 1e-12
    # inverse diagonal (Dirichlet rows -> 1)
    diag = bf.scatter_full(mesh, K.diagonal(), dirichlet_zero=False)
    diag[mesh.dirichlet] = 1.0
    assert rel(o.inverse_diagonal(1, NONE), 1.0 / diag) < 1e-12


def make_testbed_mesh(dim, degree, ref, size=4.0):
    return bf.Mesh(dim, degree, [2 ** ref] * dim, [size] * dim, [False] * dim,
                   dirichlet_faces=(0, 3))


@pytest.mark.parametrize("app,dim,degree", [("testbed_2d_q2", 2, 2), ("testbed_3d_q1", 3, 1)])
def test_generic_hooks_vs_bruteforce(oracle_libs, app, dim, degree):
    """Hooks without the split structure: nonlinear diffusion (explicit), variable
    coefficient Helmholtz (linear lhs) and |grad c| (postprocess)."""
    ref, dt, D0, alpha = 2, 1e-3, 0.7, 0.5
    o = oracle(app, prm.testbed(dim=dim, refinement=ref, dt=dt))
    mesh = make_testbed_mesh(dim, degree, ref)
    rng = np.random.default_rng(11)
    c = bf.distribute(mesh, rng.uniform(-1, 1, mesh.ndofs))
    v = rng.standard_normal(mesh.ndofs)
    _, mass = bf.assemble(mesh)
    mfull = bf.scatter_full(mesh, mass, dirichlet_zero=False)

    # block 0: value c_old, flux -dt D0 (1 + c^2) grad c
    o.set_field(0, c, age=1)
    b = bf.explicit_rhs(mesh, c, lambda cv, g: cv,
                        lambda cv, g: -dt * D0 * (1 + cv ** 2)[:, None] * g)
    exp = bf.scatter_full(mesh, b) / np.where(mfull > 0, mfull, 1.0)
    assert rel(o.apply_operator(0, RHS, INVM, None), exp) < 1e-12

    # block 1 lhs: alpha v + (1 + c^2) grad v, with constrained reads of v
    o.set_field(0, c, age=0)
    A, _ = bf.assemble(mesh, coeff=1 + c ** 2, mass_coeff=alpha)
    exp = bf.scatter_full(mesh, A @ bf.read_unique(mesh, v, True))
    assert rel(o.apply_operator(1, LHS, NONE, v), exp) < 1e-12

    # block 2: |grad c| (cell-wise gradients averaged by the lumped projection)
    b = bf.explicit_rhs(mesh, c, lambda cv, g: np.linalg.norm(g, axis=1),
                        lambda cv, g: np.zeros_like(g))
    out = bf.scatter_full(mesh, b, dirichlet_zero=False) / np.where(mfull > 0, mfull, 1.0)
    assert rel(o.apply_operator(2, RHS, INVM, None), out) < 1e-12


def test_cahn_hilliard_blocks_vs_bruteforce(oracle_libs):
    size, ref, dt, M, kappa = 8.0, 1, 1e-3, 1.0, 1.0
    o = oracle("cahn_hilliard_3d_q2", prm.cahn_hilliard(size=size, refinement=ref, dt=dt,
                                                        steps=1, M=M, kappa=kappa))
    mesh = bf.Mesh(3, 2, [2 ** ref] * 3, [size] * 3, [True] * 3)
    K, mass = bf.assemble(mesh)
    rng = np.random.default_rng(3)
    c = bf.distribute(mesh, rng.uniform(0.2, 0.8, mesh.ndofs))
    mu = bf.distribute(mesh, rng.standard_normal(mesh.ndofs))
    o.set_field(0, c, age=1)
    o.set_field(1, mu, age=1)
    cu, muu = bf.read_unique(mesh, c, False), bf.read_unique(mesh, mu, False)
    exp = bf.scatter_full(mesh, cu - dt * M * (K @ muu) / mass)
    assert rel(o.apply_operator(0, RHS, INVM, None), exp) < 1e-12
    o.set_field(0, c, age=0)
    fp = 2 * cu * (1 - cu) * (1 - 2 * cu)
    exp = bf.scatter_full(mesh, fp + kappa * (K @ cu) / mass)
    assert rel(o.apply_operator(1, RHS, INVM, None), exp) < 1e-12


# ------------------------------------------------------------- invariants
def test_poisson_operator_symmetric_positive(oracle_libs):
    o = oracle("poisson_3d_q2", prm.poisson(dim=3, refinement=2, mg_depth=1))
    rng = np.random.default_rng(5)
    u, v = rng.standard_normal((2, o.ndofs))
    Au, Av = o.apply_operator(1, LHS, NONE, u), o.apply_operator(1, LHS, NONE, v)
    mesh = poisson_mesh(3, 2)
    uf, vf = u.copy(), v.copy()
    uf[mesh.dirichlet] = 0
    vf[mesh.dirichlet] = 0
    assert abs(uf @ Av - vf @ Au) < 1e-12 * abs(uf @ Au)
    assert uf @ Au > 0


@pytest.mark.parametrize("dim", [2, 3])
def test_q2_tensor_quadratics_are_exact(oracle_libs, dim):
    """M^-1 K of a tensor-product quadratic equals -Laplacian at interior nodes."""
    ref = 2
    o = oracle(f"poisson_{dim}d_q2", prm.poisson(dim=dim, refinement=ref, mg_depth=1))
    mesh = poisson_mesh(dim, ref)
    X = mesh.node_x()
    if dim == 2:
        x, y = X[:, 0], X[:, 1]
        u, lap = x ** 2 * y ** 2, 2 * y ** 2 + 2 * x ** 2
    else:
        x, y, z = X[:, 0], X[:, 1], X[:, 2]
        u, lap = x ** 2 * y * z + z ** 2, 2 * y * z + 2
    got = o.apply_operator(1, LHS_PLAIN, INVM, u)
    interior = ~mesh.dirichlet
    assert np.abs(got[interior] + lap[interior]).max() < 1e-11 * np.abs(lap).max()


def test_constant_null_space_periodic(oracle_libs):
    """Periodic Q2: K 1 = 0, so the mu block of a constant c is f'(c) exactly."""
    o = oracle("cahn_hilliard_3d_q2", prm.cahn_hilliard(size=8.0, refinement=1, steps=1))
    c0 = 0.3
    o.set_field(0, np.full(o.ndofs, c0), age=0)
    got = o.apply_operator(1, RHS, INVM, None)
    n = 5
    got = got.reshape(n, n, n)[:-1, :-1, :-1]
    assert np.allclose(got, 2 * c0 * (1 - c0) * (1 - 2 * c0), rtol=0, atol=1e-14)


def test_cahn_hilliard_conserves_mass(oracle_libs):
    o = oracle("cahn_hilliard_3d_q2", prm.cahn_hilliard(size=8.0, refinement=2, steps=30,
                                                        dt=5e-3, outputs=0))
    m = o.lumped_mass()
    o.solve_increment()
    c0 = o.get_field(0, -1)
    for _ in range(30):
        o.solve_increment()
    c1 = o.get_field(0, -1)
    assert not np.allclose(c0, c1)
    assert abs(m @ c1 - m @ c0) < 1e-12 * abs(m @ c0)


def test_allen_cahn_energy_decreases(oracle_libs):
    size, ref, kappa = 20.0, 4, 1.5
    o = oracle("allen_cahn_2d_q1", prm.allen_cahn(dim=2, size=size, refinement=ref, dt=0.02,
                                                  steps=40, kappa=kappa, outputs=0))
    mesh = bf.Mesh(2, 1, [2 ** ref] * 2, [size] * 2, [True, True])
    K, mass = bf.assemble(mesh)

    def energy(phi):
        u = bf.read_unique(mesh, phi, False)
        return mass @ (0.25 * u ** 4 - 0.5 * u ** 2) + 0.5 * kappa * u @ (K @ u)

    o.solve_increment()
    e = [energy(o.get_field(0, -1))]
    for _ in range(40):
        o.solve_increment()
        e.append(energy(o.get_field(0, -1)))
    assert all(b < a for a, b in zip(e, e[1:]))


# --------------------------------------------------------------- CG and GMG
def test_cg_iterations_match_scipy(oracle_libs):
    dim, ref, tol = 2, 3, 1e-10
    text = prm.poisson(dim=dim, refinement=ref, mg_depth=1, preconditioner="None",
                       tolerance=tol)
    o = oracle("poisson_2d_q2", text)
    mesh = poisson_mesh(dim, ref)
    K, mass = bf.assemble(mesh)
    rng = np.random.default_rng(9)
    f = rng.standard_normal(mesh.ndofs)
    o.set_field(1, f)
    o.set_time(1, 0.0)
    st = o.linear_solve(1)
    free = ~mesh.dirichlet[mesh.unique]
    A = K[free][:, free]
    b = (mass * f[mesh.unique])[free]
    its = [0]
    x, info = spla.cg(A, b, x0=np.zeros_like(b), rtol=0.0, atol=tol, maxiter=1000,
                      callback=lambda xk: its.__setitem__(0, its[0] + 1))
    assert info == 0 and st["converged"]
    assert abs(st["iterations"] - its[0]) <= 1
    u = o.get_field(0)[mesh.unique][free]
    assert rel(u, x) < 1e-8


def test_mg_prolongation_reproduces_coarse_quadratics(oracle_libs):
    dim, ref = 2, 3
    o = oracle("poisson_2d_q2", prm.poisson(dim=dim, refinement=ref, mg_depth=3))
    fine, coarse = poisson_mesh(dim, ref), poisson_mesh(dim, ref - 1)

    def q(mesh):
        X = mesh.node_x()
        return X[:, 0] * (1 - X[:, 0]) * X[:, 1] * (1 - X[:, 1])

    got = o.mg_prolongate(1, 2, q(coarse))
    assert np.abs(got - q(fine)).max() < 1e-14


def test_mg_restriction_is_prolongation_transpose(oracle_libs):
    o = oracle("poisson_3d_q2", prm.poisson(dim=3, refinement=2, mg_depth=2))
    fine, coarse = poisson_mesh(3, 2), poisson_mesh(3, 1)
    rng = np.random.default_rng(2)
    cvec = rng.standard_normal(coarse.ndofs)
    cvec[coarse.dirichlet] = 0
    fvec = rng.standard_normal(fine.ndofs)
    fvec[fine.dirichlet] = 0
    lhs = fvec @ o.mg_prolongate(1, 1, cvec)
    rhs = cvec @ o.mg_restrict(1, 1, fvec)
    assert abs(lhs - rhs) < 1e-13 * abs(lhs)


def test_gmg_preconditioned_cg_converges_fast(oracle_libs):
    text = prm.poisson(dim=2, refinement=5, mg_depth=4, preconditioner="GMG", tolerance=1e-10)
    o = oracle("poisson_2d_q2", text)
    rng = np.random.default_rng(4)
    o.set_field(1, rng.standard_normal(o.ndofs))
    o.set_time(1, 0.0)
    st = o.linear_solve(1)
    assert st["converged"] and st["iterations"] <= 12


# -------------------------------------------------------------------- adjoint
def test_adjoint_gradients_match_finite_differences(oracle_libs):
    size, ref, n, dt = 16.0, 3, 12, 0.05
    o = oracle("allen_cahn_2d_q1", prm.allen_cahn(dim=2, size=size, refinement=ref, dt=dt,
                                                  steps=n, outputs=0))
    mesh = bf.Mesh(2, 1, [2 ** ref] * 2, [size] * 2, [True, True])
    m = o.lumped_mass()
    X = mesh.node_x()
    phi0 = bf.distribute(mesh, 0.6 * np.sin(2 * np.pi * X[:, 0] / size)
                         * np.cos(2 * np.pi * X[:, 1] / size) + 0.1)
    target = o.adjoint_forward(phi0, 1.2, 1.3, n)

    def J(L, kappa, p0):
        r = o.adjoint_forward(p0, L, kappa, n) - target
        return 0.5 * float(m @ (r * r))

    L, kappa = 1.0, 1.5
    lam = m * (o.adjoint_forward(phi0, L, kappa, n) - target)
    dL, dk, dphi = o.adjoint_vjp(phi0, lam, L, kappa, n)
    eps = 1e-6
    fd_L = (J(L + eps, kappa, phi0) - J(L - eps, kappa, phi0)) / (2 * eps)
    fd_k = (J(L, kappa + eps, phi0) - J(L, kappa - eps, phi0)) / (2 * eps)
    rng = np.random.default_rng(8)
    v = bf.distribute(mesh, rng.standard_normal(mesh.ndofs))
    fd_v = (J(L, kappa, phi0 + eps * v) - J(L, kappa, phi0 - eps * v)) / (2 * eps)
    assert abs(dL - fd_L) <= 1e-6 * abs(fd_L)
    assert abs(dk - fd_k) <= 1e-6 * abs(fd_k)
    assert abs(dphi @ v - fd_v) <= 1e-6 * abs(fd_v)


@pytest.mark.parametrize("periodic", [True, False], ids=["periodic", "natural"])
def test_cahn_hilliard_adjoint_matches_finite_differences(oracle_libs, periodic):
    """The oracle's Cahn-Hilliard adjoint (c' = c - dt M L mu, mu = f'(c) + kappa L c):
    dJ/dM, dJ/dkappa and a directional derivative of dJ/dc0 against central finite
    differences of its forward map, J = 1/2 sum m (c_N - target)^2."""
    size, ref, n, dt = 8.0, 2, 10, 1e-3
    text = prm.cahn_hilliard(size=size, refinement=ref, dt=dt, steps=n, outputs=0)
    if not periodic:
        text = text.replace("periodic = true", "periodic = false")
    o = oracle("cahn_hilliard_3d_q2", text)
    mesh = bf.Mesh(3, 2, [2 ** ref] * 3, [size] * 3, [periodic] * 3)
    m = o.lumped_mass()
    X = mesh.node_x()
    c0 = bf.distribute(mesh, 0.5 + 0.3 * np.sin(2 * np.pi * X[:, 0] / size)
                       * np.cos(2 * np.pi * X[:, 1] / size) * np.cos(2 * np.pi * X[:, 2] / size))
    target = o.adjoint_forward(c0, 1.2, 1.3, n)

    def J(mob, kappa, p0):
        r = o.adjoint_forward(p0, mob, kappa, n) - target
        return 0.5 * float(m @ (r * r))

    mob, kappa = 1.0, 1.5
    lam = m * (o.adjoint_forward(c0, mob, kappa, n) - target)
    dmob, dk, dc = o.adjoint_vjp(c0, lam, mob, kappa, n)
    eps = 1e-6
    fd_m = (J(mob + eps, kappa, c0) - J(mob - eps, kappa, c0)) / (2 * eps)
    fd_k = (J(mob, kappa + eps, c0) - J(mob, kappa - eps, c0)) / (2 * eps)
    rng = np.random.default_rng(8)
    v = bf.distribute(mesh, rng.standard_normal(mesh.ndofs))
    fd_v = (J(mob, kappa, c0 + eps * v) - J(mob, kappa, c0 - eps * v)) / (2 * eps)
    assert abs(dmob - fd_m) <= 1e-6 * abs(fd_m), (dmob, fd_m)
    assert abs(dk - fd_k) <= 1e-6 * abs(fd_k), (dk, fd_k)
    assert abs(dc @ v - fd_v) <= 1e-6 * abs(fd_v), (dc @ v, fd_v)


def test_cahn_hilliard_adjoint_forward_is_the_explicit_blocks(oracle_libs):
    """The adjoint's forward map is the application's c and mu blocks: n increments of
    the explicit run from the same c0 give the same c (1e-13), and the lumped mass
    sum m c is conserved (periodic)."""
    n = 6
    text = prm.cahn_hilliard(size=8.0, refinement=2, dt=1e-3, steps=n, outputs=0)
    o = oracle("cahn_hilliard_3d_q2", text)
    o.run(0)  # increment 0: the application's initial condition and mu from it
    c0 = o.get_field(0, -1)
    o.run(n)
    cn = o.get_field(0, -1)
    fwd = o.adjoint_forward(c0, 1.0, 1.0, n)  # the application's M = kappa = 1
    assert rel(fwd, cn) <= 1e-13
    m = o.lumped_mass()
    assert abs(m @ fwd - m @ c0) <= 1e-13 * abs(m @ c0)


@pytest.mark.parametrize("periodic", [False, True], ids=["natural", "periodic"])
def test_newton_matches_dense_newton(oracle_libs, periodic):
    """One implicit Allen-Cahn increment: the oracle's Newton loop against a dense Newton
    on the brute-force assembled residual
        R(n) = M (n_old - n - dt m f'(n)) - dt kappa m K n,  f'(n) = 4 n (n - 1)(n - 1/2),
    with exact linear solves and the same stopping rule (RMSEPerField, one field:
    ||M^-1/2 R||_2 <= tol sqrt(V), checked before each update, start from n_old).
    Same iteration count, same root."""
    dt, m, kappa, tol = 0.5, 1.0, 2.0, 1e-10  # 1 + dt m f'' > 0: SPD Jacobian for CG
    text = prm.allen_cahn_implicit(refinement=3, steps=2, dt=dt, m_well=m, kappa=kappa,
                                   linear_tolerance=1e-14, newton_tolerance=tol,
                                   periodic=periodic)
    o = oracle("allen_cahn_implicit_2d_q1", text)
    o.solve_increment()  # increment 0: initial condition
    n0 = o.get_field(0, -1)
    o.solve_increment()
    n1 = o.get_field(0, -1)
    ns = o.last_newton_stats(1)

    mesh = bf.Mesh(2, 1, [8, 8], [100.0, 100.0], [periodic, periodic])
    K, mass = bf.assemble(mesh)
    Kd = K.toarray()
    u_old = n0[mesh.unique]

    def residual(u):
        return mass * (u_old - u - dt * m * 4.0 * u * (u - 1.0) * (u - 0.5)) - dt * kappa * m * (Kd @ u)

    def jacobian(u):  # dR/du
        return -np.diag(mass * (1.0 + dt * m * (12.0 * u * (u - 1.0) + 2.0))) - dt * kappa * m * Kd

    u, its = u_old.copy(), 0
    while np.linalg.norm(residual(u) / np.sqrt(mass)) > tol * 100.0 and its < 100:
        u = u - np.linalg.solve(jacobian(u), residual(u))
        its += 1
    assert its >= 3  # a genuinely nonlinear step
    assert ns["converged"] and ns["iterations"] == its
    assert abs(ns["tolerance"] - tol * 100.0) <= 1e-15 * tol * 100.0
    assert rel(n1[mesh.unique], u) < 1e-9


def test_newton_bicgstab_matches_dense_newton_and_scipy(oracle_libs):
    """Implicit Cahn-Hilliard increment (two fields, non-symmetric Jacobian, Q2): the
    oracle's Newton loop with its BiCGStab restatement against

    * a dense Newton on the brute-force assembled residual
        R_c  = M (c_old - c) - dt Mc K mu,
        R_mu = M (mu - W c (c - 1)(c - 1/2)) - Kc K c,
      with exact linear solves and the same stopping rule (RMSEPerField, two fields:
      ||M^-1/2 R||_2 <= tol sqrt(V) sqrt(2)): same iteration count, same root;
    * scipy's BiCGStab on the same scaled Newton systems M^-1/2 J du = M^-1/2 R, each
      started from the previous update as the Newton solver does (recurrence residual,
      i.e. `exact residual = false`, no preconditioner): total BiCGStab steps within one
      per Newton update plus 5 %.
    """
    dt, Mc, Kc, W, tol, lin_tol = 0.05, 1.0, 1.5, 16.0, 1e-10, 1e-12
    size, ref = 16.0, 2
    text = prm.cahn_hilliard_implicit(size=size, refinement=ref, steps=2, dt=dt, McV=Mc, KcV=Kc,
                                      WcV=W, ic_type=0, icamplitude=0.2, linear_tolerance=lin_tol,
                                      newton_tolerance=tol, exact_residual=False)
    o = oracle("cahn_hilliard_implicit_2d_q2", text)
    o.solve_increment()
    c0 = o.get_field(0, -1)
    o.solve_increment()
    c1, mu1 = o.get_field(0, -1), o.get_field(1, 0)
    ns = o.last_newton_stats(0)

    mesh = bf.Mesh(2, 2, [4, 4], [size, size], [False, False])
    K, mass = bf.assemble(mesh)
    Kd = K.toarray()
    nu = mass.size
    cold = c0[mesh.unique]
    Dinv = np.concatenate([1.0 / np.sqrt(mass)] * 2)

    def residual(x):
        c, mu = x[:nu], x[nu:]
        return np.concatenate([mass * (cold - c) - dt * Mc * (Kd @ mu),
                               mass * (mu - W * c * (c - 1.0) * (c - 0.5)) - Kc * (Kd @ c)])

    def lhs(x):  # the Jacobian operator the lhs hooks assemble (= -dR/dx)
        c = x[:nu]
        fp = W * (0.5 + 3.0 * c * (c - 1.0))
        return np.block([[np.diag(mass), dt * Mc * Kd],
                         [np.diag(mass * fp) + Kc * Kd, -np.diag(mass)]])

    norm = tol * size * np.sqrt(2.0)
    x = np.concatenate([cold, np.zeros(nu)])  # mu starts from its initial value 0
    du = np.zeros(2 * nu)
    its, scipy_steps = 0, 0
    while np.linalg.norm(Dinv * residual(x)) > norm and its < 50:
        A, b = Dinv[:, None] * lhs(x), Dinv * residual(x)
        count = [0]
        du, info = spla.bicgstab(A, b, x0=du, rtol=0.0, atol=lin_tol * size * np.sqrt(2.0),
                                 maxiter=10000, callback=lambda xk: count.__setitem__(0, count[0] + 1))
        assert info == 0
        scipy_steps += count[0]
        x = x + np.linalg.solve(lhs(x), residual(x))  # exact Newton update
        its += 1
    assert its >= 2
    assert ns["converged"] and ns["iterations"] == its
    assert rel(c1[mesh.unique], x[:nu]) < 1e-9
    assert rel(mu1[mesh.unique], x[nu:]) < 1e-9
    # scipy counts a step that ends at the half-update test differently (one per solve),
    # and a BiCGStab step count at a tight tolerance moves by a few with the last bits of
    # its dot products: 5 % slack on top.
    assert abs(ns["linear_iterations"] - scipy_steps) <= its + scipy_steps // 20, (ns, scipy_steps)


@pytest.mark.parametrize("time_dependent", [False, True], ids=["fixed", "time_dependent"])
def test_time_dependent_dirichlet_values(oracle_libs, time_dependent):
    """Dirichlet values of the test-bed application, g = 0.25 (boundary id + 1) + 0.1 x
    + 2 t, on the explicit field c and the linear field v after k increments: g at the
    time of increment k when the conditions are declared time dependent (re-evaluated at
    the start of every increment, reference problem.cc:428-430), g at t = 0 otherwise
    (evaluated once).  At a corner the lower boundary id's value holds."""
    dt, k = 0.01, 4
    o = oracle("testbed_2d_q2", prm.testbed(dim=2, refinement=3, steps=k, dt=dt,
                                             dirichlet_faces=(0, 3), time_dependent=time_dependent))
    assert o.run(k) == 0
    mesh = make_testbed_mesh(2, 2, 3)
    x = mesh.node_x()
    t = k * dt if time_dependent else 0.0
    lower_x = mesh.coords()[:, 0] == 0                # boundary id 0
    upper_y = mesh.coords()[:, 1] == mesh.N[1] - 1    # boundary id 3
    g = np.where(lower_x, 0.25, 1.0) + 0.1 * x[:, 0] + 2.0 * t
    mask = lower_x | upper_y
    for f in (0, 1):
        u = o.get_field(f, -1)
        assert np.abs(u[mask] - g[mask]).max() <= 1e-13, f


def test_newton_gmres_matches_dense_newton_and_scipy(oracle_libs):
    """The oracle's GMRES restatement (left preconditioning, basis 30, no preconditioner)
    inside the implicit Cahn-Hilliard Newton loop: same root and Newton count as a dense
    Newton on the brute-force residual, and scipy's GMRES (restart 30) on the same scaled
    systems, each started from the previous update, needs the same number of steps
    within one per update plus 5 %."""
    dt, Mc, Kc, W, tol, lin_tol = 0.05, 1.0, 1.5, 16.0, 1e-10, 1e-12
    size = 16.0
    text = prm.cahn_hilliard_implicit(size=size, refinement=2, steps=2, dt=dt, McV=Mc, KcV=Kc,
                                      WcV=W, icamplitude=0.2, solver="gmres",
                                      linear_tolerance=lin_tol, newton_tolerance=tol)
    o = oracle("cahn_hilliard_implicit_2d_q2", text)
    o.solve_increment()
    c0 = o.get_field(0, -1)
    o.solve_increment()
    c1, mu1 = o.get_field(0, -1), o.get_field(1, 0)
    ns = o.last_newton_stats(0)
    mesh = bf.Mesh(2, 2, [4, 4], [size, size], [False, False])
    K, mass = bf.assemble(mesh)
    Kd, nu = K.toarray(), mass.size
    cold = c0[mesh.unique]
    Dinv = np.concatenate([1.0 / np.sqrt(mass)] * 2)

    def residual(x):
        c, mu = x[:nu], x[nu:]
        return np.concatenate([mass * (cold - c) - dt * Mc * (Kd @ mu),
                               mass * (mu - W * c * (c - 1.0) * (c - 0.5)) - Kc * (Kd @ c)])

    def lhs(x):
        c = x[:nu]
        return np.block([[np.diag(mass), dt * Mc * Kd],
                         [np.diag(mass * W * (0.5 + 3.0 * c * (c - 1.0))) + Kc * Kd, -np.diag(mass)]])

    x = np.concatenate([cold, np.zeros(nu)])
    du = np.zeros(2 * nu)
    its, steps = 0, 0
    while np.linalg.norm(Dinv * residual(x)) > tol * size * np.sqrt(2.0) and its < 50:
        count = [0]
        du, info = spla.gmres(Dinv[:, None] * lhs(x), Dinv * residual(x), x0=du, rtol=0.0,
                              atol=lin_tol * size * np.sqrt(2.0), restart=30, maxiter=1000,
                              callback=lambda r: count.__setitem__(0, count[0] + 1),
                              callback_type="pr_norm")
        assert info == 0
        steps += count[0]
        x = x + np.linalg.solve(lhs(x), residual(x))
        its += 1
    assert its >= 2 and ns["converged"] and ns["iterations"] == its
    assert rel(c1[mesh.unique], x[:nu]) < 1e-9 and rel(mu1[mesh.unique], x[nu:]) < 1e-9
    assert abs(ns["linear_iterations"] - steps) <= its + steps // 20, (ns, steps)


def test_newton_richardson_matches_dense_newton_and_numpy(oracle_libs):
    """The oracle's Richardson restatement (omega 0.065, no preconditioner) inside the
    implicit Allen-Cahn Newton loop: same root and Newton count as a dense Newton, and
    the same number of Richardson steps as x += omega (b - A x) on the assembled scaled
    systems (each started from the previous update), within one per update."""
    dt, m, kappa, tol, lin_tol, omega = 0.5, 1.0, 2.0, 1e-10, 1e-12, 0.065
    text = prm.allen_cahn_implicit(refinement=3, steps=2, dt=dt, m_well=m, kappa=kappa,
                                   linear_tolerance=lin_tol, newton_tolerance=tol,
                                   solver="richardson",
                                   linear_extra=["  subsection Richardson",
                                                 f"    set omega = {omega}", "  end"])
    o = oracle("allen_cahn_implicit_2d_q1", text)
    o.solve_increment()
    n0 = o.get_field(0, -1)
    o.solve_increment()
    n1 = o.get_field(0, -1)
    ns = o.last_newton_stats(1)
    mesh = bf.Mesh(2, 1, [8, 8], [100.0, 100.0], [False, False])
    K, mass = bf.assemble(mesh)
    Kd = K.toarray()
    u_old = n0[mesh.unique]
    D = 1.0 / np.sqrt(mass)

    def residual(u):
        return mass * (u_old - u - dt * m * 4.0 * u * (u - 1.0) * (u - 0.5)) - dt * kappa * m * (Kd @ u)

    def lhs(u):
        return np.diag(mass * (1.0 + dt * m * (12.0 * u * (u - 1.0) + 2.0))) + dt * kappa * m * Kd

    u, du, its, steps = u_old.copy(), np.zeros_like(u_old), 0, 0
    while np.linalg.norm(D * residual(u)) > tol * 100.0 and its < 50:
        A, b = D[:, None] * lhs(u), D * residual(u)
        k = 0
        while np.linalg.norm(b - A @ du) > lin_tol * 100.0 and k < 1000:
            du = du + omega * (b - A @ du)
            k += 1
        steps += k
        u = u + np.linalg.solve(lhs(u), residual(u))
        its += 1
    assert its >= 3 and ns["converged"] and ns["iterations"] == its
    assert rel(n1[mesh.unique], u) < 1e-9
    assert abs(ns["linear_iterations"] - steps) <= its, (ns, steps)


def test_fgmres_equals_right_preconditioned_gmres(oracle_libs):
    """With a fixed preconditioner, flexible GMRES (right preconditioning keeping the
    preconditioned basis) produces the iterates of right-preconditioned GMRES: same
    Newton and Krylov step counts and the same implicit Cahn-Hilliard state."""
    right = ["  subsection GMRES", "    set right preconditioning = true", "  end"]
    out = {}
    for solver, extra in (("fgmres", ()), ("gmres", right)):
        text = prm.cahn_hilliard_implicit(refinement=4, steps=2, dt=1e-3, solver=solver,
                                          preconditioner="Chebyshev", linear_tolerance=1e-12,
                                          newton_tolerance=1e-10, linear_extra=extra)
        o = oracle("cahn_hilliard_implicit_2d_q2", text)
        o.solve_increment()
        o.solve_increment()
        out[solver] = (o.last_newton_stats(0), o.get_field(0, -1), o.get_field(1, 0))
    a, b = out["fgmres"], out["gmres"]
    assert a[0]["iterations"] == b[0]["iterations"]
    assert a[0]["linear_iterations"] == b[0]["linear_iterations"]
    assert rel(a[1], b[1]) < 1e-12 and rel(a[2], b[2]) < 1e-10


def test_newton_minres_matches_dense_newton_and_scipy(oracle_libs):
    """The oracle's MINRES restatement (no preconditioner, absolute residual: the
    unscaled symmetric Jacobian) inside the implicit Allen-Cahn Newton loop: same root and
    Newton count as a dense Newton, and per update the first step at which scipy's MINRES
    iterate (same system, started from the previous update) has |b - A x| <= tol, within
    one per update."""
    dt, m, kappa, tol, lin_tol = 0.5, 1.0, 2.0, 1e-9, 1e-11
    text = prm.allen_cahn_implicit(refinement=3, steps=2, dt=dt, m_well=m, kappa=kappa,
                                   linear_tolerance=lin_tol, newton_tolerance=tol,
                                   solver="minres", tolerance_type="AbsoluteResidual")
    o = oracle("allen_cahn_implicit_2d_q1", text)
    o.solve_increment()
    n0 = o.get_field(0, -1)
    o.solve_increment()
    n1 = o.get_field(0, -1)
    ns = o.last_newton_stats(1)
    mesh = bf.Mesh(2, 1, [8, 8], [100.0, 100.0], [False, False])
    K, mass = bf.assemble(mesh)
    Kd = K.toarray()
    u_old = n0[mesh.unique]

    def residual(u):
        return mass * (u_old - u - dt * m * 4.0 * u * (u - 1.0) * (u - 0.5)) - dt * kappa * m * (Kd @ u)

    def lhs(u):
        return np.diag(mass * (1.0 + dt * m * (12.0 * u * (u - 1.0) + 2.0))) + dt * kappa * m * Kd

    class Reached(Exception):
        pass

    u, du, its, steps = u_old.copy(), np.zeros_like(u_old), 0, 0
    while np.linalg.norm(residual(u)) > tol and its < 50:
        A, b = lhs(u), residual(u)
        k, x_last = 0, du
        if np.linalg.norm(b - A @ du) > lin_tol:
            count = [0]

            def cb(xk):
                count[0] += 1
                if np.linalg.norm(b - A @ xk) <= lin_tol:
                    x_last_holder[0] = xk.copy()
                    raise Reached

            x_last_holder = [None]
            with pytest.raises(Reached):
                spla.minres(A, b, x0=du, rtol=1e-30, maxiter=1000, callback=cb)
            k, x_last = count[0], x_last_holder[0]
        du = x_last
        steps += k
        u = u + np.linalg.solve(lhs(u), residual(u))
        its += 1
    assert its >= 3 and ns["converged"] and ns["iterations"] == its
    assert rel(n1[mesh.unique], u) < 1e-9
    assert abs(ns["linear_iterations"] - steps) <= its, (ns, steps)


def test_mechanics_matches_assembled_elasticity(oracle_libs):
    """The oracle's linear-elasticity solve (one vector field stored as three component
    fields, symmetric-gradient stress with isotropic constants) equals the solution of
    the brute-force assembled system K[(i,a),(j,b)] = sum_q w_q (lambda G_ia G_jb +
    mu delta_ab G_i.G_j + mu G_ib G_ja) with the same Dirichlet values (u_x = -1 on
    x = 0, zero on the other Dirichlet faces), to 1e-8."""
    E, nu, size = 2.0, 0.3, 100.0
    text = prm.mechanics(size=size, refinement=1, tolerance=1e-13,
                         tolerance_type="AbsoluteResidual")
    o = oracle("mechanics_3d_q2", text)
    o.solve_increment()
    faces = (0, 1, 3, 5)
    mesh = bf.Mesh(3, 2, [2, 2, 2], [size] * 3, [False] * 3, dirichlet_faces=faces)
    G, wq = bf.element_gradients(mesh)
    lam = nu * E / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    n = mesh.ndofs
    K = np.zeros((3 * n, 3 * n))
    for cell in mesh.cells_iter():
        dofs, _ = mesh.cell_dofs(cell)
        for a in range(3):
            for b in range(3):
                Ke = (lam * np.einsum("q,qi,qj->ij", wq, G[:, :, a], G[:, :, b])
                      + mu * np.einsum("q,qi,qj->ij", wq, G[:, :, b], G[:, :, a]))
                if a == b:
                    Ke += mu * np.einsum("q,qid,qjd->ij", wq, G, G)
                K[np.ix_(a * n + dofs, b * n + dofs)] += Ke
    x = mesh.coords()
    fixed = np.concatenate([mesh.dirichlet] * 3)
    g = np.zeros(3 * n)
    g[:n][mesh.dirichlet & (x[:, 0] == 0)] = -1.0
    u = g.copy()
    free = ~fixed
    u[free] = np.linalg.solve(K[np.ix_(free, free)], -K[np.ix_(free, fixed)] @ g[fixed])
    got = np.concatenate([o.get_field(c, 0) for c in range(3)])
    assert np.abs(u[:n]).max() > 0.1 and np.abs(u[n:]).max() > 1e-3
    assert np.abs(got - u).max() <= 1e-8 * np.abs(u).max(), np.abs(got - u).max()


@pytest.mark.parametrize("ttype", ["RMSEPerField", "IntegratedPerField", "RMSETotal",
                                   "AbsoluteResidual"])
def test_tolerance_normalization_counts_declared_fields(oracle_libs, ttype):
    """normalization_value (reference linear_solver.h:414-427): sqrt(volume) for the RMSE
    types times sqrt(number of the block's declared fields) for the PerField types.  The
    elasticity block holds ONE declared field, the vector u (stored as three component
    fields), so the factor is sqrt(1), not sqrt(3).  The absolute tolerance the solver
    reports is tolerance value x that normalization."""
    size, tol = 100.0, 1e-6
    o = oracle("mechanics_3d_q2", prm.mechanics(size=size, refinement=1, tolerance=tol,
                                                tolerance_type=ttype))
    o.run(0)
    got = o.linear_solve(1)["tolerance"]
    vol = size ** 3
    want = tol * {"RMSEPerField": np.sqrt(vol), "IntegratedPerField": 1.0,
                  "RMSETotal": np.sqrt(vol), "AbsoluteResidual": 1.0}[ttype]
    assert got == pytest.approx(want, rel=1e-14), (ttype, got, want)


def test_tolerance_normalization_two_scalar_fields(oracle_libs):
    """A Newton block of two scalar fields (implicit Cahn-Hilliard: c and mu) counts two
    declared fields: tolerance = tolerance value x sqrt(volume) x sqrt(2) (RMSEPerField)."""
    size = 100.0
    text = prm.cahn_hilliard_implicit(size=size, refinement=2, steps=1, outputs=0,
                                      newton_tolerance=1e-6)
    o = oracle("cahn_hilliard_implicit_2d_q2", text)
    o.run(1)
    got = o.last_newton_stats(0)["tolerance"]
    assert got == pytest.approx(1e-6 * np.sqrt(size ** 2) * np.sqrt(2.0), rel=1e-14)
