"""Newton blocks on the GPU against the oracle's Newton loop (reference
include/prismspf/solvers/newton_solver.h:73-156), implicit Allen-Cahn application.

Per increment: the same number of Newton updates, CG steps within one per update (a
CG residual can land on either side of its tolerance by rounding), and the new state,
the post-processed |grad n| and free-energy density within 1e-9 relative.  Cases: no
preconditioner, Chebyshev and GMG; natural and periodic boundaries; a lattice whose
operators run on the tiled stencil kernel (129 x 129 nodes) and small ones; a Newton
loop stopped by its iteration limit.  The time step keeps the Jacobian
M (1 + dt m f'') + dt kappa m K positive definite (dt m min f'' = -dt > -1), as CG
requires.
"""

from __future__ import annotations

import numpy as np
import pytest

from phasefield_amd import Context, prm
from phasefield_amd.capi import AGE_LATEST
from oracle.binding import Oracle
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

APP = "allen_cahn_implicit_2d_q1"

CASES = {
    "none_natural": prm.allen_cahn_implicit(refinement=5, steps=3, dt=0.5, linear_tolerance=1e-12,
                                            newton_tolerance=1e-10),
    "chebyshev_periodic": prm.allen_cahn_implicit(refinement=5, steps=3, dt=0.5,
                                                  linear_tolerance=1e-12, newton_tolerance=1e-10,
                                                  preconditioner="Chebyshev", periodic=True),
    "gmg_natural_129": prm.allen_cahn_implicit(refinement=7, steps=3, dt=0.5,
                                               linear_tolerance=1e-12, newton_tolerance=1e-10,
                                               preconditioner="GMG", mg_depth=4),
    "gmg_periodic_rect": prm.allen_cahn_implicit(refinement=4, steps=3, dt=0.25,
                                                 linear_tolerance=1e-12, newton_tolerance=1e-10,
                                                 preconditioner="GMG", mg_depth=3, periodic=True,
                                                 subdivisions=[3, 2]),
    # Stopped by the iteration limit: the state after two updates, not converged.
    "iteration_limit": prm.allen_cahn_implicit(refinement=5, steps=2, dt=0.5,
                                               linear_tolerance=1e-12, newton_tolerance=1e-14,
                                               newton_max_iterations=2),
}


@pytest.mark.parametrize("name", list(CASES))
def test_newton_matches_oracle(device, name):
    text = CASES[name]
    ctx, orc = Context(APP, text), Oracle(APP, text)
    n_inc = ctx.info.n_increments
    assert ctx.solve_increment() == 0 and orc.solve_increment() == 0  # initial condition
    for inc in range(1, n_inc + 1):
        assert ctx.solve_increment() == 0
        assert orc.solve_increment() == 0
        g, o = ctx.last_newton_stats(1), orc.last_newton_stats(1)
        print(f"{name} increment {inc}: newton {g.iterations}/{o['iterations']}, "
              f"cg {g.linear_iterations}/{o['linear_iterations']}, "
              f"residual {g.residual:.3e}/{o['residual']:.3e}", flush=True)
        assert g.iterations == o["iterations"]
        assert abs(g.linear_iterations - o["linear_iterations"]) <= g.iterations
        assert bool(g.converged) == o["converged"]
        assert g.tolerance == pytest.approx(o["tolerance"], rel=1e-14)
        if name == "iteration_limit":
            assert g.iterations == 2 and not g.converged
        else:
            assert g.converged and g.residual <= g.tolerance
        for f in range(3):
            err = rel_err(ctx.get_field(f, AGE_LATEST).ravel(), orc.get_field(f, AGE_LATEST))
            assert err <= 1e-9, (name, inc, f, err)
