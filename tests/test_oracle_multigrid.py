"""Pins of the oracle's multigrid smoother, its eigenvalue estimate, the V-cycle and the
preconditioned GMRES / FGMRES iterates, each against the mathematics on matrices
assembled independently by tests/bruteforce.py (Poisson, Q2, zero Dirichlet, tolerance
type AbsoluteResidual so the level operators are the unscaled stiffness matrices; one
case with the M^-1/2 scaling of the RMSE types).  On the free nodes, A is the assembled
matrix and D = diag(A).

* Chebyshev smoother (reference linear_solver.h:126-142, deal.II PreconditionChebyshev;
  Saad, Iterative Methods, Alg. 12.1): from a zero start, k steps give
  x = (I - p_k(D^-1 A)) A^-1 b with p_k(l) = T_k((theta - l)/delta) / T_k(theta/delta),
  theta, delta the centre and half-width of [lmin, lmax].  Checked on eigenvectors of
  D^-1 A (closed-form polynomial value) and as a matrix function on a random vector.
* Eigenvalue estimate: CG with the diagonal preconditioner is the Lanczos process on
  B = D^-1/2 A D^-1/2 started from D^-1/2 r0; the estimate is 1.2 x the largest Ritz
  value of the k-step Lanczos matrix, lmin = lmax / smoothing range.  Checked against
  an independent numpy Lanczos with full reorthogonalization.
* V-cycle (reference linear_solver.h:148-157, mg::Matrix / Multigrid with
  MGCoarseGridApplySmoother or a coarse CG): pre-smooth from zero, restrict the
  residual with R = P^T, recurse, prolongate-add, post-smooth.  Checked against a numpy
  V-cycle on the assembled level matrices, P from Q2 nodal interpolation of the coarse
  space (built here from Lagrange polynomials on the GLL points), each smoother as the
  matrix function above with the oracle's bounds, and an exact coarse solve.
* GMRES / FGMRES (reference linear_solver.h:461-490, deal.II SolverGMRES /
  SolverFGMRES): the k-th iterate is the unique minimizer of ||M (b - A x)|| over
  x0 + K_k(MA, M r0) (left) or of ||b - A x|| over x0 + M K_k(AM, r0) (right, and FGMRES
  with a fixed preconditioner), restarted every `max basis size` steps.  M is the
  oracle's preconditioner probed column by column (Chebyshev, or a V-cycle with the
  smoother as coarse solver; both linear).
"""

from __future__ import annotations

import numpy as np
import pytest
from numpy.polynomial import chebyshev as npcheb

from phasefield_amd import prm
from tests import bruteforce as bf

BLOCK = 1  # the Poisson application's linear block (u)


def oracle(app, text):
    from oracle.binding import Oracle

    return Oracle(app, text, threads=2)


def poisson_mesh(dim, ref):
    return bf.Mesh(dim, 2, [2 ** ref] * dim, [1.0] * dim, [False] * dim,
                   dirichlet_faces=range(2 * dim))


def free_system(mesh, scaled=False):
    """A on the free nodes (optionally M^-1/2 K M^-1/2) and the free mask (full lattice)."""
    K, mass = bf.assemble(mesh)
    free = ~mesh.dirichlet
    A = K.toarray()[np.ix_(free, free)]
    if scaled:
        # rows scaled by M^-1/2 (the operator output is scaled, reference
        # mf_operator.cc:25-31; the source is not)
        A = (1.0 / np.sqrt(mass[free]))[:, None] * A
    return A, free


def full(mesh, free, v):
    out = np.zeros(mesh.ndofs)
    out[free] = v
    return out


def cheb_t(k, t):
    return npcheb.chebval(t, [0.0] * k + [1.0])


def smoother_matrix(A, lmax, lmin, k):
    """x = S b: k Chebyshev steps from zero, as a matrix function of D^-1 A
    (S = q(D^-1 A) D^-1 with q(l) = (1 - p_k(l)) / l; D^-1 A is diagonalizable with a real
    spectrum: it is similar to a symmetric matrix for the matrices used here)."""
    dinv = 1.0 / np.diag(A)
    lam, V = np.linalg.eig(dinv[:, None] * A)
    lam, V = lam.real, V.real
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    p = cheb_t(k, (theta - lam) / delta) / cheb_t(k, theta / delta)
    q = (1.0 - p) / lam
    return ((V * q) @ np.linalg.inv(V)) * dinv[None, :]


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


# ------------------------------------------------------------------ smoother
@pytest.mark.parametrize("level", [0, 1, 2])
def test_chebyshev_smoother_is_the_chebyshev_polynomial(oracle_libs, level):
    ref, depth, degree = 3, 3, 4
    o = oracle("poisson_2d_q2", prm.poisson(dim=2, refinement=ref, mg_depth=depth,
                                            smoother_degree=degree, coarse_solver="smoother"))
    mesh = poisson_mesh(2, ref - depth + 1 + level)
    A, free = free_system(mesh)
    # The level's inverse diagonal is the assembled one (constrained rows: one).
    inv = o.mg_inv_diag(BLOCK, level)
    assert rel(inv[free], 1.0 / np.diag(A)) < 1e-12
    assert np.all(inv[~free] == 1.0)
    D = np.diag(A)
    B = A / np.sqrt(np.outer(D, D))
    mu, W = np.linalg.eigh(B)
    _, lmax, lmin = o.mg_smooth(BLOCK, level, np.zeros(mesh.ndofs))
    assert lmin == pytest.approx(lmax / 15.0, rel=1e-15)
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    # Eigenvectors of D^-1 A: the smoother scales them by 1 - p_k(mu).
    for j in (0, len(mu) // 3, len(mu) // 2, len(mu) - 1):
        v = W[:, j] / np.sqrt(D)
        x, _, _ = o.mg_smooth(BLOCK, level, full(mesh, free, A @ v))
        pk = cheb_t(degree, (theta - mu[j]) / delta) / cheb_t(degree, theta / delta)
        assert np.abs(x[free] - (1.0 - pk) * v).max() <= 1e-11 * np.abs(v).max(), (j, mu[j])
        assert np.all(x[~free] == 0.0)
    # A random right-hand side: the matrix function.
    rng = np.random.default_rng(level)
    b = rng.standard_normal(free.sum())
    x, _, _ = o.mg_smooth(BLOCK, level, full(mesh, free, b))
    assert rel(x[free], smoother_matrix(A, lmax, lmin, degree) @ b) < 1e-11
    # A smoother of another degree would not pass: degree 3 and 5 differ from 4.
    for other in (degree - 1, degree + 1):
        assert rel(smoother_matrix(A, lmax, lmin, other) @ b, x[free]) > 1e-3


# ---------------------------------------------------------- eigenvalue estimate
def lanczos_ritz_max(B, v, k):
    n = len(v)
    Q = np.zeros((n, k + 1))
    alpha, beta = np.zeros(k), np.zeros(k)
    Q[:, 0] = v / np.linalg.norm(v)
    for j in range(k):
        w = B @ Q[:, j]
        alpha[j] = Q[:, j] @ w
        for _ in range(2):  # full reorthogonalization, twice
            w -= Q[:, :j + 1] @ (Q[:, :j + 1].T @ w)
        beta[j] = np.linalg.norm(w)
        Q[:, j + 1] = w / beta[j]
    T = np.diag(alpha) + np.diag(beta[:-1], 1) + np.diag(beta[:-1], -1)
    return np.linalg.eigvalsh(T).max()


@pytest.mark.parametrize("level", [1, 2])
def test_eigenvalue_estimate_is_the_lanczos_ritz_value(oracle_libs, level):
    ref, depth, its = 3, 3, 10
    o = oracle("poisson_2d_q2", prm.poisson(dim=2, refinement=ref, mg_depth=depth))
    mesh = poisson_mesh(2, ref - depth + 1 + level)
    A, free = free_system(mesh)
    rng = np.random.default_rng(7 + level)
    x0 = rng.standard_normal(mesh.ndofs)
    lmax, lmin = o.mg_estimate(BLOCK, level, x0)
    ds = 1.0 / np.sqrt(np.diag(A))
    B = ds[:, None] * A * ds[None, :]
    r0 = A @ x0[free]  # residual of A x = 0 from x0 (constrained entries read as zero)
    ritz = lanczos_ritz_max(B, ds * r0, its)
    assert lmax == pytest.approx(1.2 * ritz, rel=1e-10)
    assert lmin == pytest.approx(lmax / 15.0, rel=1e-15)
    # A Ritz value lies inside the spectrum, close to its top after 10 steps.
    top = np.linalg.eigvalsh(B).max()
    assert 0.9 * top <= ritz <= top * (1 + 1e-12)
    # Without the 1.2 safety factor, or with one step fewer, the estimate differs.
    assert abs(lmax - ritz) > 0.1 * ritz
    assert abs(1.2 * lanczos_ritz_max(B, ds * r0, its - 1) - lmax) > 1e-8 * lmax


# ------------------------------------------------------------------- V-cycle
def lagrange_1d(xs, t):
    """Values of the Lagrange polynomials on nodes xs at points t: [len(t), len(xs)]."""
    L = np.ones((len(t), len(xs)))
    for b in range(len(xs)):
        for m in range(len(xs)):
            if m != b:
                L[:, b] *= (t - xs[m]) / (xs[b] - xs[m])
    return L


def prolongation_1d(nc_cells, degree):
    """Nodal interpolation of the coarse Q_degree space (nc_cells cells on [0,1]) at the
    nodes of the mesh with twice as many cells: [fine nodes, coarse nodes]."""
    xs, _, _ = bf.lagrange_tables(degree)
    H = 1.0 / nc_cells
    nf_nodes = degree * 2 * nc_cells + 1
    P = np.zeros((nf_nodes, degree * nc_cells + 1))
    xf = np.arange(nf_nodes) * (H / 2) / degree
    for i, x in enumerate(xf):
        c = min(int(x / H), nc_cells - 1)
        vals = lagrange_1d(xs * H + c * H, np.array([x]))[0]
        P[i, degree * c:degree * c + degree + 1] = vals
    return P


@pytest.mark.parametrize("scaled", [False, True])
def test_v_cycle_matches_assembled_levels(oracle_libs, scaled):
    dim, ref, depth, degree = 2, 3, 3, 4
    ttype = "RMSEPerField" if scaled else "AbsoluteResidual"
    o = oracle("poisson_2d_q2", prm.poisson(dim=dim, refinement=ref, mg_depth=depth,
                                            smoother_degree=degree, coarse_solver="cg",
                                            tolerance_type=ttype))
    meshes = [poisson_mesh(dim, ref - depth + 1 + l) for l in range(depth)]
    systems = [free_system(m, scaled) for m in meshes]
    bounds = [o.mg_smooth(BLOCK, l, np.zeros(m.ndofs))[1:] for l, m in enumerate(meshes)]
    # The level operators are the assembled level matrices.
    rng = np.random.default_rng(11)
    for l, m in enumerate(meshes):
        A, free = systems[l]
        v = rng.standard_normal(free.sum())
        assert rel(o.mg_level_vmult(BLOCK, l, full(m, free, v))[free], A @ v) < 1e-12
    transfers = [None]
    for l in range(1, depth):
        p1 = prolongation_1d(2 ** (ref - depth + l), 2)
        P = np.kron(p1, p1)  # x fastest: index i0 + N0 i1
        transfers.append(P[np.ix_(systems[l][1], systems[l - 1][1])])
        # the oracle's prolongation is this interpolation
        vc = rng.standard_normal(systems[l - 1][1].sum())
        got = o.mg_prolongate(BLOCK, l, full(meshes[l - 1], systems[l - 1][1], vc))
        assert rel(got[systems[l][1]], transfers[l] @ vc) < 1e-13

    def vcycle(l, d):
        A, _ = systems[l]
        if l == 0:
            return np.linalg.solve(A, d)
        S = smoother_matrix(A, *bounds[l], degree)
        u = S @ d
        u = u + transfers[l] @ vcycle(l - 1, transfers[l].T @ (d - A @ u))
        return u + S @ (d - A @ u)

    d = rng.standard_normal(systems[-1][1].sum())
    u = o.mg_vcycle(BLOCK, full(meshes[-1], systems[-1][1], d))
    want = vcycle(depth - 1, d)
    assert rel(u[systems[-1][1]], want) < 1e-9
    # Without post-smoothing the result is far off.
    A, _ = systems[-1]
    S = smoother_matrix(A, *bounds[-1], degree)
    pre = S @ d
    pre = pre + transfers[-1] @ vcycle(depth - 2, transfers[-1].T @ (d - A @ pre))
    assert rel(pre, want) > 1e-3


# -------------------------------------------------------- preconditioned GMRES
def with_solver(text, solver, right, basis):
    lines = [f"  set solver type = {solver}", "  subsection GMRES",
             f"    set max basis size = {basis}",
             f"    set right preconditioning = {'true' if right else 'false'}", "  end"]
    return text.replace("  subsection Chebyshev", "\n".join(lines) + "\n  subsection Chebyshev", 1)


def arnoldi_basis(op, v, k):
    Q = np.zeros((len(v), k))
    Q[:, 0] = v / np.linalg.norm(v)
    for j in range(1, k):
        w = op(Q[:, j - 1])
        for _ in range(2):
            w -= Q[:, :j] @ (Q[:, :j].T @ w)
        Q[:, j] = w / np.linalg.norm(w)
    return Q


def gmres_iterate(A, M, x0, b, steps, basis, right):
    x = x0.copy()
    done = 0
    while done < steps:
        k = min(basis, steps - done)
        r = b - A @ x
        if right:
            Q = arnoldi_basis(lambda v: A @ (M @ v), r, k)
            y, *_ = np.linalg.lstsq(A @ M @ Q, r, rcond=None)
            x = x + M @ (Q @ y)
        else:
            z = M @ r
            Q = arnoldi_basis(lambda v: M @ (A @ v), z, k)
            y, *_ = np.linalg.lstsq(M @ A @ Q, z, rcond=None)
            x = x + Q @ y
        done += k
    return x


@pytest.mark.parametrize("precond", ["Chebyshev", "GMG"])
@pytest.mark.parametrize("solver,right,basis,steps", [
    ("gmres", False, 30, 6), ("gmres", True, 30, 6), ("gmres", False, 4, 7),
    ("gmres", True, 3, 7), ("fgmres", False, 30, 6), ("fgmres", False, 4, 7)])
def test_preconditioned_gmres_iterates_are_residual_minimizers(oracle_libs, precond, solver,
                                                               right, basis, steps):
    ref = 2
    if precond == "GMG":
        # the V-cycle preconditioner converges within a few steps: fewer steps (and a
        # restart after 2 where the case restarts) keep the iterates informative
        steps, basis = 3, (2 if basis < steps else 30)
    text = prm.poisson(dim=2, refinement=ref, mg_depth=2, preconditioner=precond,
                       smoother_degree=3, coarse_solver="smoother")
    o = oracle("poisson_2d_q2", with_solver(text, solver, right, basis))
    mesh = poisson_mesh(2, ref)
    A, free = free_system(mesh)
    n = int(free.sum())
    # The preconditioner, column by column; it is linear.
    M = np.stack([o.precondition(BLOCK, full(mesh, free, np.eye(n)[j]))[free]
                  for j in range(n)], axis=1)
    rng = np.random.default_rng(3)
    a, c = rng.standard_normal(n), rng.standard_normal(n)
    assert rel(o.precondition(BLOCK, full(mesh, free, a + 2 * c))[free], M @ (a + 2 * c)) < 1e-12
    x0, b = rng.standard_normal(n), rng.standard_normal(n)
    got = o.krylov_steps(BLOCK, full(mesh, free, x0), full(mesh, free, b), steps)
    want = gmres_iterate(A, M, x0, b, steps, basis, right or solver == "fgmres")
    assert rel(got[free], want) < 1e-9, rel(got[free], want)
    # The other preconditioning side gives another iterate.
    other = gmres_iterate(A, M, x0, b, steps, basis, not (right or solver == "fgmres"))
    assert rel(other, want) > 1e-6
