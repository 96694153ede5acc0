"""GMRES, Richardson and MINRES inside Newton blocks on the GPU against the oracle (deal.II's
SolverGMRES, SolverRichardson and SolverMinRes as the reference selects them by `solver type`,
include/prismspf/solvers/linear_solver.h:461-490, restated on both sides).

Per increment: the same number of Newton updates, Krylov steps within 15 % (their count
at a tight tolerance moves with the last bits of the dot products, which the two sides
sum in different orders), and every field within 1e-9 relative.  Cases: GMRES with left
preconditioning (none, Chebyshev) and right preconditioning (GMG) on the two-field
implicit Cahn-Hilliard application, a GMRES basis small enough to restart, flexible
GMRES (SolverFGMRES) with GMG,
Richardson (unpreconditioned with a damping factor, and Chebyshev-preconditioned) and
MINRES (SolverMinRes; unpreconditioned and Chebyshev-preconditioned, absolute-residual
tolerance so that the Newton system is the unscaled symmetric Jacobian) on the implicit
Allen-Cahn application.
"""

from __future__ import annotations

import pytest

from phasefield_amd import Context, prm
from phasefield_amd.capi import AGE_LATEST
from oracle.binding import Oracle
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

CH = "cahn_hilliard_implicit_2d_q2"
AC = "allen_cahn_implicit_2d_q1"
RIGHT = ["  subsection GMRES", "    set right preconditioning = true", "  end"]

CASES = {
    "gmres_left_none": (CH, 0, prm.cahn_hilliard_implicit(
        refinement=5, steps=3, dt=1e-3, solver="gmres", linear_tolerance=1e-12,
        newton_tolerance=1e-10)),
    "gmres_left_chebyshev_periodic": (CH, 0, prm.cahn_hilliard_implicit(
        refinement=5, steps=3, dt=1e-3, solver="gmres", linear_tolerance=1e-12,
        newton_tolerance=1e-10, preconditioner="Chebyshev", periodic=True)),
    "gmres_right_gmg": (CH, 0, prm.cahn_hilliard_implicit(
        refinement=5, steps=3, dt=1e-3, solver="gmres", linear_tolerance=1e-12,
        newton_tolerance=1e-10, preconditioner="GMG", mg_depth=3, linear_extra=RIGHT)),
    "gmres_restarts": (CH, 0, prm.cahn_hilliard_implicit(
        refinement=5, steps=2, dt=1e-3, solver="gmres", linear_tolerance=1e-12,
        newton_tolerance=1e-10,
        linear_extra=["  subsection GMRES", "    set max basis size = 8", "  end"])),
    "fgmres_gmg": (CH, 0, prm.cahn_hilliard_implicit(
        refinement=5, steps=3, dt=1e-3, solver="fgmres", linear_tolerance=1e-12,
        newton_tolerance=1e-10, preconditioner="GMG", mg_depth=3)),
    "richardson_damped": (AC, 1, prm.allen_cahn_implicit(
        refinement=3, steps=3, dt=0.5, solver="richardson", linear_tolerance=1e-12,
        newton_tolerance=1e-10,
        linear_extra=["  subsection Richardson", "    set omega = 0.065", "  end"])),
    "richardson_chebyshev": (AC, 1, prm.allen_cahn_implicit(
        refinement=5, steps=3, dt=0.5, solver="richardson", linear_tolerance=1e-12,
        newton_tolerance=1e-10, preconditioner="Chebyshev")),
    "minres_none": (AC, 1, prm.allen_cahn_implicit(
        refinement=5, steps=3, dt=0.5, solver="minres", linear_tolerance=1e-11,
        newton_tolerance=1e-9, tolerance_type="AbsoluteResidual")),
    "minres_chebyshev": (AC, 1, prm.allen_cahn_implicit(
        refinement=5, steps=3, dt=0.5, solver="minres", linear_tolerance=1e-11,
        newton_tolerance=1e-9, tolerance_type="AbsoluteResidual", preconditioner="Chebyshev")),
}


@pytest.mark.parametrize("name", list(CASES))
def test_krylov_in_newton_matches_oracle(device, name):
    app, block, text = CASES[name]
    ctx, orc = Context(app, text), Oracle(app, text)
    assert ctx.solve_increment() == 0 and orc.solve_increment() == 0  # initial condition
    for inc in range(1, ctx.info.n_increments + 1):
        assert ctx.solve_increment() == 0
        assert orc.solve_increment() == 0
        g, o = ctx.last_newton_stats(block), orc.last_newton_stats(block)
        print(f"{name} increment {inc}: newton {g.iterations}/{o['iterations']}, "
              f"linear {g.linear_iterations}/{o['linear_iterations']}", flush=True)
        assert g.iterations == o["iterations"] and g.converged and o["converged"]
        slack = max(g.iterations, (15 * o["linear_iterations"]) // 100)
        assert abs(g.linear_iterations - o["linear_iterations"]) <= slack
        for f in range(ctx.info.n_fields):
            err = rel_err(ctx.get_field(f, AGE_LATEST).ravel(), orc.get_field(f, AGE_LATEST))
            assert err <= 1e-9, (name, inc, f, err)
