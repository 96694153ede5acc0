"""Routing of equation hooks between the split (stencil) path and the generic cell path
is safe for hooks whose flux coefficients branch (application switching_2d_q2):

* block 1, coefficient switching at c > 1: the structure probe (random states with
  values up to +-50) sees a non-constant coefficient and routes it to the generic path;
* block 2, coefficient switching at increment 3: the probe sees the dependence on the
  increment and re-derives the plan whenever the increment changes;
* block 3, coefficient switching at c > 60 (beyond the probe's values): the comparison
  with the generic path at the plan's first use on the real fields (c reaches 70)
  drops the split plan and says so in the output log.

Every field matches the oracle (which evaluates every hook at every quadrature point)
after 6 increments, across the switch at increment 3.
"""

from __future__ import annotations

import pytest

from phasefield_amd import Context, PPFError, prm
from phasefield_amd.capi import AGE_LATEST
from oracle.binding import Oracle
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

APP = "switching_2d_q2"


def test_branching_hooks_match_oracle(device):
    text = prm.switching(steps=6)
    ctx, orc = Context(APP, text), Oracle(APP, text)
    assert ctx.run(6) == 0 and orc.run(6) == 0
    for f in range(4):
        err = rel_err(ctx.get_field(f, AGE_LATEST).ravel(), orc.get_field(f, AGE_LATEST))
        print(f"field {f}: relative difference from the oracle {err:.2e}", flush=True)
        assert err <= 1e-10, (f, err)
    log = ctx.output_log()
    print(log, flush=True)
    assert "solve block 3 (rhs): the split operator differs from the generic path" in log
    assert "solve block 1 " not in log and "solve block 0 " not in log


def test_forced_split_path_refuses_value_branch(device):
    ctx = Context(APP, prm.switching(steps=2))
    ctx.set_option("operator_path", "split")
    with pytest.raises(PPFError) as e:
        ctx.run(2)
    assert "does not qualify" in str(e.value)
