"""Consecutive explicit blocks in one launch (pair_kernel, option explicit_pairs).

The Cahn-Hilliard c block and the mu block that reads the new c through its stencil, and
the precipitate model's c and mu blocks (mu also reads eta, written by the block before
them), run as one z-march after the first increment.  Every case runs the same
increments with the pair on and off and compares every field bit for bit: the fused
kernel evaluates each node with the taps, hook and summation order of the two
launches.  (The two launches run on the tiled stencil kernel at every size:
stencil_min_nodes = 0.)  The cases cover periodic and mirror (natural) boundaries, lattices that are
not multiples of the tile (partial tiles, and remainder columns processed node by node),
both tile heights and short z chunks (several chunks, odd lengths); the log names the
fused blocks.
"""

from __future__ import annotations

import numpy as np
import pytest

from phasefield_amd import Context, prm
from phasefield_amd.capi import AGE_LATEST

pytestmark = pytest.mark.gpu


def _mirror(text):
    return text.replace("periodic = true", "periodic = false")


CH = "cahn_hilliard_3d_q2"
PR = "precipitate_3d_q1"
CASES = [
    ("ch-8", CH, prm.cahn_hilliard(size=8.0, refinement=2, steps=8, dt=1e-3, outputs=4), {}),
    ("ch-48x32x16", CH, prm.cahn_hilliard(size=[24.0, 16.0, 8.0], refinement=3, subdivisions=[3, 2, 1],
                                          steps=8, dt=1e-3, outputs=0), {}),
    ("ch-128x32x32-rows2-zc3", CH, prm.cahn_hilliard(size=[64.0, 16.0, 16.0], refinement=4,
                                                     subdivisions=[4, 1, 1], steps=6, dt=1e-3,
                                                     outputs=0), {"pair_rows": "2", "pair_z_chunk": "3"}),
    ("ch-129x33x33-mirror", CH, _mirror(prm.cahn_hilliard(size=[64.0, 16.0, 16.0], refinement=4,
                                                          subdivisions=[4, 1, 1], steps=6, dt=1e-3,
                                                          outputs=0)), {}),
    ("ch-49x33x17-mirror-zc5", CH, _mirror(prm.cahn_hilliard(size=[24.0, 16.0, 8.0], refinement=3,
                                                             subdivisions=[3, 2, 1], steps=6, dt=1e-3,
                                                             outputs=0)), {"pair_z_chunk": "5"}),
    ("ch-130x8x8-remainder", CH, prm.cahn_hilliard(size=[65.0, 4.0, 4.0], refinement=0,
                                                   subdivisions=[65, 4, 4], steps=6, dt=1e-3,
                                                   outputs=0), {}),
    ("precipitate-8", PR, prm.precipitate(size=16.0, refinement=3, subdivisions=(1, 1, 1), steps=6,
                                          outputs=3), {}),
    ("precipitate-128x16x8", PR, prm.precipitate(size=[128.0, 16.0, 8.0], refinement=3,
                                                 subdivisions=(16, 2, 1), steps=6, outputs=0),
     {"pair_rows": "2"}),
    ("precipitate-129x9x9-remainder", PR, prm.precipitate(size=[129.0, 9.0, 9.0], refinement=0,
                                                          subdivisions=(129, 9, 9), steps=6,
                                                          outputs=0), {}),
]


def _run(app, text, steps, opts, pairs):
    ctx = Context(app, text)
    ctx.set_option("explicit_pairs", "1" if pairs else "0")
    # the two launches on the tiled stencil kernel at every lattice size (small lattices
    # otherwise take the node-by-node kernel, whose sums are ordered differently)
    ctx.set_option("stencil_min_nodes", "0")
    for k, v in opts.items():
        ctx.set_option(k, v)
    assert ctx.run(steps) == 0
    out = [ctx.get_field(f, AGE_LATEST).copy() for f in range(ctx.info.n_fields)]
    log = ctx.output_log()
    ctx.close()
    return out, log


@pytest.mark.parametrize("name,app,text,opts", CASES, ids=[c[0] for c in CASES])
def test_pair_matches_two_launches(name, app, text, opts):
    steps = 5
    fused, log = _run(app, text, steps, opts, True)
    plain, log0 = _run(app, text, steps, opts, False)
    print(name, flush=True)
    print(log, flush=True)
    assert "one launch" in log
    assert "one launch" not in log0
    for f, (a, b) in enumerate(zip(fused, plain)):
        d = float(np.abs(a - b).max())
        print(f"{name} field {f}: max difference {d:.3e}", flush=True)
        assert np.array_equal(a, b), (name, f, d)
