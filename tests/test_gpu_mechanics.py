"""Vector fields on the GPU: the linear elasticity application (the reference's
applications/mechanics/boundary_value_problem; one vector field u stored as the component
fields u_x, u_y, u_z; lhs stress C eps(u) from get_symmetric_gradient) against the
oracle: CG iterations within one, displacements within 1e-8 relative.  Cases: the
reference's boundary conditions and RMSE tolerance without a preconditioner, with
Chebyshev and with GMG (two levels), and an absolute tolerance on a finer mesh.  The oracle's solve is pinned
against a brute-force assembled elasticity system in tests/test_oracle.py.
"""

from __future__ import annotations

import numpy as np
import pytest

from phasefield_amd import Context, prm
from phasefield_amd.capi import AGE_LATEST
from oracle.binding import Oracle
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

APP = "mechanics_3d_q2"
CASES = {
    "rmse_none": dict(refinement=2, tolerance=1e-10),
    "rmse_chebyshev": dict(refinement=2, tolerance=1e-10, preconditioner="Chebyshev"),
    "absolute_none": dict(refinement=3, tolerance=1e-9, tolerance_type="AbsoluteResidual"),
    "rmse_gmg": dict(refinement=3, tolerance=1e-10, preconditioner="GMG", mg_depth=2),
}


@pytest.mark.parametrize("name", list(CASES))
def test_mechanics_matches_oracle(device, name):
    text = prm.mechanics(**CASES[name])
    ctx, orc = Context(APP, text), Oracle(APP, text)
    assert ctx.field_names() == ["u_x", "u_y", "u_z"]
    ctx.run(0)
    orc.run(0)
    st = ctx.last_solve_stats(1)
    ost = Oracle(APP, text).linear_solve(1)
    print(name, "gpu its", st.iterations, "oracle its", ost["iterations"], flush=True)
    assert st.converged and ost["converged"]
    assert abs(st.iterations - ost["iterations"]) <= 1
    # same normalization of the tolerance (one declared field, reference
    # linear_solver.h:414-427)
    assert st.tolerance == pytest.approx(ost["tolerance"], rel=1e-14)
    for c in range(3):
        g, o = ctx.get_field(c, AGE_LATEST).ravel(), orc.get_field(c, AGE_LATEST)
        assert np.abs(o).max() > 0.0 or c > 0
        assert rel_err(g, o) <= 1e-8, (name, c, rel_err(g, o))
