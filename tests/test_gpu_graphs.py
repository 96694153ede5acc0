"""Multigrid V-cycles replayed from HIP graphs (option mg_graphs) against the same
V-cycles launched one by one: the same kernels with the same arguments run in both, so
solutions and iteration counts agree bit for bit.  Every case solves several times in a
row, with eager work of the program in between (operator applications, host reads and
synchronizations), so each replay is checked against the eager result: Poisson 3-D and
2-D with FP64 and FP32 levels (the graph key includes the precision), and the implicit
Allen-Cahn time loop, whose launches alternate between two halt flags from one
increment to the next (two graphs)."""

from __future__ import annotations

import numpy as np
import pytest

from phasefield_amd import Context, prm
from phasefield_amd.capi import AGE_LATEST, FP64, OP_LHS, SCALE_NONE

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _eager_work(ctx, rng):
    """An operator application and a host read between two solves."""
    n = ctx.info.n_nodes
    src = torch.tensor(rng.uniform(-1, 1, n), device="cuda")
    dst = torch.zeros_like(src)
    ctx.apply_operator(1, OP_LHS, SCALE_NONE, FP64, src.data_ptr(), dst.data_ptr())
    ctx.synchronize()
    return float(dst.abs().max())


CASES = [
    ("poisson_3d_q2", dict(dim=3, refinement=4, mg_depth=3), "double"),
    ("poisson_3d_q2", dict(dim=3, refinement=4, mg_depth=3), "single"),
    ("poisson_2d_q2", dict(dim=2, refinement=6, mg_depth=4), "double"),
    ("poisson_2d_q2", dict(dim=2, refinement=6, mg_depth=4), "single"),
]


@pytest.mark.parametrize("app,kw,precision", CASES,
                         ids=[f"{c[0]}-{c[2]}" for c in CASES])
def test_graph_vcycles_match_eager(device, app, kw, precision):
    text = prm.poisson(tolerance=1e-10, coarse_solver="cg", **kw)
    results = {}
    for graphs in ("0", "1"):
        ctx = Context(app, text)
        ctx.set_option("mg_graphs", graphs)
        ctx.set_option("mg_level_precision", precision)
        ctx.set_option("linear_initial_guess", "zero")
        ctx.solve_block(0)
        rng = np.random.default_rng(3)
        info = ctx.info
        out = []
        for rep in range(4):
            # a new right-hand side field each time, then the GMG-CG solve
            ctx.set_field(1, rng.normal(size=info.n_dofs), 0)
            ctx.solve_block(1)
            st = ctx.last_solve_stats(1)
            assert st.converged
            out.append((st.iterations, ctx.get_field(0, AGE_LATEST).copy(),
                        _eager_work(ctx, rng)))
        results[graphs] = out
        ctx.close()
    for (ia, ua, ea), (ib, ub, eb) in zip(results["0"], results["1"]):
        print(app, precision, ia, ib, flush=True)
        assert ia == ib
        assert np.array_equal(ua, ub)
        assert ea == eb


def test_graph_vcycles_implicit_time_loop(device):
    """Implicit Allen-Cahn with a GMG preconditioner: each increment's launches carry
    the halt flag of its parity, so replays alternate between two captured graphs."""
    fields = {}
    for graphs in ("0", "1"):
        text = prm.allen_cahn_implicit(refinement=6, steps=8, outputs=0, preconditioner="GMG",
                                       mg_depth=3)
        ctx = Context("allen_cahn_implicit_2d_q1", text)
        ctx.set_option("mg_graphs", graphs)
        states = []
        for _ in range(8):
            assert ctx.solve_increment() == 0
            states.append(ctx.get_field(0, AGE_LATEST).copy())
        fields[graphs] = states
        ctx.close()
    for a, b in zip(fields["0"], fields["1"]):
        assert np.array_equal(a, b)
