"""The increment loop of ppf_run enqueues explicit increments up to pipeline_depth (1, 2
by default, or 3) ahead of the NaN probe.

Every case diverges (explicit steps far above the stability limit), so the run stops
with status 2 part-way.  The pipelined ppf_run must leave exactly what a loop of
ppf_solve_increment leaves: the same status, increment counter, time, output log and
every field at every age, bit for bit (same kernels on the same device).  The oracle
stops within one increment of the same place: near overflow, the GPU's fused
multiply-adds and the oracle's separate roundings can put the first inf - inf one
increment apart.
"""

from __future__ import annotations

import numpy as np
import pytest

from phasefield_amd import Context, prm
from oracle.binding import Oracle

pytestmark = pytest.mark.gpu


# (app, prm text): NaN inside a run of speculated increments, NaN one increment before an
# output increment (which is solved one at a time), NaN on an output increment, and the
# stencil path of the Q2 Cahn-Hilliard blocks.
DIVERGING = [
    ("allen_cahn_2d_q1", prm.allen_cahn(dim=2, size=16.0, refinement=4, steps=40, dt=0.5, outputs=4)),
    ("allen_cahn_2d_q1", prm.allen_cahn(dim=2, size=16.0, refinement=4, steps=35, dt=0.5, outputs=5)),
    ("allen_cahn_2d_q1", prm.allen_cahn(dim=2, size=16.0, refinement=4, steps=36, dt=0.5, outputs=6)),
    ("cahn_hilliard_3d_q2", prm.cahn_hilliard(size=8.0, refinement=2, steps=40, dt=0.1, outputs=4)),
    ("cahn_hilliard_3d_q2", prm.cahn_hilliard(size=8.0, refinement=2, steps=40, dt=0.05, outputs=40)),
]


def state(ctx: Context):
    """Every field at every age, as raw bytes (NaN payloads included)."""
    out = []
    for f in range(ctx.info.n_fields):
        age = 0
        while True:
            try:
                out.append(ctx.get_field(f, age).tobytes())
            except Exception:
                break
            age += 1
    return out


# FP32 operator arithmetic on the FP64 state (explicit_precision): the stencil kernel on
# the FP64 arrays (single), and FP32 copies converted into the state (single_copies),
# whose conversions also skip once the halt flag is set.
DIVERGING_DEPTH = [
    (DIVERGING[0][0], DIVERGING[0][1], {"pipeline_depth": "1"}),
    (DIVERGING[0][0], DIVERGING[0][1], {"pipeline_depth": "3"}),
    (DIVERGING[1][0], DIVERGING[1][1], {"pipeline_depth": "3"}),
    (DIVERGING[3][0], DIVERGING[3][1], {"pipeline_depth": "1"}),
    (DIVERGING[3][0], DIVERGING[3][1], {"pipeline_depth": "3"}),
]
DIVERGING_FP32 = [
    ("cahn_hilliard_3d_q2", DIVERGING[3][1], {"explicit_precision": "single", "stencil_min_nodes": "0"}),
    ("allen_cahn_2d_q1", DIVERGING[0][1], {"explicit_precision": "single", "stencil_min_nodes": "0"}),
    ("allen_cahn_2d_q1", DIVERGING[0][1], {"explicit_precision": "single_copies"}),
]


def _context(app, text, opts):
    ctx = Context(app, text)
    for k, v in opts.items():
        ctx.set_option(k, v)
    return ctx


@pytest.mark.parametrize("app,text,opts",
                         [c + ({},) for c in DIVERGING] + DIVERGING_DEPTH + DIVERGING_FP32,
                         ids=[f"{c[0]}-{i}" for i, c in enumerate(DIVERGING)] +
                             [f"{c[0]}-depth{c[2]['pipeline_depth']}-{i}" for i, c in enumerate(DIVERGING_DEPTH)] +
                             [f"{c[0]}-{c[2]['explicit_precision']}" for c in DIVERGING_FP32])
def test_pipelined_run_stops_like_single_increments(app, text, opts):
    n = 40
    piped = _context(app, text, opts)
    status_p = piped.run(n)
    single = _context(app, text, opts)
    status_s = 0
    while status_s == 0 and single.get_info().increment <= n:
        status_s = single.solve_increment()
    ip, is_ = piped.get_info(), single.get_info()
    print(app, "status", status_p, status_s, "increment", ip.increment, is_.increment, flush=True)
    assert status_p == status_s == 2
    assert (ip.increment, ip.time) == (is_.increment, is_.time)
    assert piped.output_log() == single.output_log()
    assert state(piped) == state(single)
    if "explicit_precision" in opts:
        return  # FP32 rounding moves the divergence away from the FP64 oracle's

    orc = Oracle(app, text)
    assert orc.run(n) == 2
    last = lambda log: int([l for l in log.splitlines() if l.startswith("Iteration")][-1].split()[1])
    print("NaN at increment", last(piped.output_log()), "oracle", last(orc.output_log()))
    assert abs(last(piped.output_log()) - last(orc.output_log())) <= 1


def test_pipelined_run_continues_across_calls():
    """Two ppf_run calls (the second resumes at the first's counter) equal one call."""
    text = prm.cahn_hilliard(size=8.0, refinement=2, steps=12, dt=1e-3, outputs=3)
    a, b = Context("cahn_hilliard_3d_q2", text), Context("cahn_hilliard_3d_q2", text)
    assert a.run(12) == 0
    assert b.run(5) == 0
    assert b.get_info().increment == 6
    assert b.run(12) == 0
    ia, ib = a.get_info(), b.get_info()
    assert (ia.increment, ia.time) == (ib.increment, ib.time)
    assert a.output_log() == b.output_log()
    assert state(a) == state(b)
    assert np.isfinite(a.get_field(0, 1)).all()
