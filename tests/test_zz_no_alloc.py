"""ppf_apply_operator_noalloc (the operator application that allocates no device memory,
for callers that cannot allocate) against ppf_apply_operator on the same inputs: the
same kernels run, so results agree bit for bit.  Each case makes two calls in a row, the
second on a smaller lattice (less scratch), so nothing left in the shared device array
from the first call may leak into the second.  Covered: the tiled stencil path
(Cahn-Hilliard), the generic path's one-pass cell-block kernel (linear elasticity, FP64;
no scratch) and its two-pass kernels whose per-cell results live in the scratch (linear
elasticity, FP64; test bed, FP32 with FP32 copies of the non-src fields), an operator
whose scratch need exceeds the array (a clear error), and the free device memory before
and after a first call on a fresh context (unchanged)."""

from __future__ import annotations

import numpy as np
import pytest

from phasefield_amd import Context, PPFError, prm
from phasefield_amd.capi import FP32, FP64, OP_LHS, OP_RHS, SCALE_INVM, SCALE_NONE

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SCRATCH_BYTES = 64 << 20


def seeded(ctx, rng):
    info = ctx.info
    for f in range(info.n_fields):
        try:
            ctx.set_field(f, rng.uniform(0.1, 0.9, info.n_dofs), 0)
        except Exception:
            pass


def both(ctx, block, which, scaling, prec, nb, rng):
    """(regular result, allocation-free result) of one application on random input."""
    dt = torch.float64 if prec == FP64 else torch.float32
    n = nb * ctx.info.n_nodes
    src = torch.tensor(rng.uniform(-1, 1, n), dtype=dt, device="cuda")
    a = torch.full((n,), float("nan"), dtype=dt, device="cuda")
    b = torch.full((n,), float("nan"), dtype=dt, device="cuda")
    ptr = src.data_ptr() if which != OP_RHS else 0
    ctx.apply_operator(block, which, scaling, prec, ptr, a.data_ptr())
    ctx.apply_operator_noalloc(block, which, scaling, prec, ptr, b.data_ptr())
    ctx.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


CASES = [
    # (app, [prm text of the first, larger lattice; of the second, smaller one],
    #  block, which, scaling, precision, fields in the block, operator path,
    #  generic-path kernel)
    ("cahn_hilliard_3d_q2", [prm.cahn_hilliard(size=16.0, refinement=4),
                             prm.cahn_hilliard(size=8.0, refinement=3)],
     0, OP_RHS, SCALE_INVM, FP64, 1, "auto", "cell_block"),
    ("mechanics_3d_q2", [prm.mechanics(refinement=3), prm.mechanics(refinement=2)],
     1, OP_LHS, SCALE_NONE, FP64, 3, "auto", "cell_block"),
    ("mechanics_3d_q2", [prm.mechanics(refinement=3), prm.mechanics(refinement=2)],
     1, OP_LHS, SCALE_NONE, FP64, 3, "auto", "two_pass"),
    ("testbed_2d_q2", [prm.testbed(dim=2, refinement=5), prm.testbed(dim=2, refinement=3)],
     1, OP_LHS, SCALE_NONE, FP32, 1, "generic", "two_pass"),
]


@pytest.mark.parametrize("app,texts,block,which,scaling,prec,nb,path,kernel", CASES,
                         ids=[f"{c[0]}-{c[-1]}" for c in CASES])
def test_noalloc_matches_operator_two_calls(device, app, texts, block, which, scaling, prec,
                                            nb, path, kernel):
    rng = np.random.default_rng(5)
    for text in texts:  # the second call runs on a smaller lattice
        ctx = Context(app, text)
        ctx.set_option("operator_path", path)
        ctx.set_option("generic_kernel", kernel)
        ctx.set_option("stencil_min_nodes", 0)  # split hooks: the tiled kernel
        seeded(ctx, rng)
        print(app, ctx.info.n_nodes, flush=True)
        a, b = both(ctx, block, which, scaling, prec, nb, rng)
        assert np.isfinite(a).all()
        assert np.array_equal(a, b), float(np.abs(a - b).max())
        ctx.close()


def test_noalloc_rejects_oversized_scratch(device):
    # 64^3 Q2 elasticity cells on the two-pass generic kernels: 3 x 27 x 8 B per cell =
    # 170 MB of scratch.
    ctx = Context("mechanics_3d_q2", prm.mechanics(refinement=6))
    ctx.set_option("generic_kernel", "two_pass")
    n = 3 * ctx.info.n_nodes
    src = torch.zeros(n, dtype=torch.float64, device="cuda")
    dst = torch.zeros_like(src)
    with pytest.raises(PPFError) as e:
        ctx.apply_operator_noalloc(1, OP_LHS, SCALE_NONE, FP64, src.data_ptr(), dst.data_ptr())
    assert e.value.status == -1 and "device scratch" in str(e.value)
    # the regular entry point still works on the same context
    ctx.apply_operator(1, OP_LHS, SCALE_NONE, FP64, src.data_ptr(), dst.data_ptr())
    ctx.close()


def test_noalloc_first_call_allocates_nothing(device):
    text = prm.mechanics(refinement=4)  # two-pass kernels: 4096 cells, 2.7 MB of scratch
    warm = Context("mechanics_3d_q2", prm.mechanics(refinement=2))
    warm.set_option("generic_kernel", "two_pass")
    rng = np.random.default_rng(9)
    both(warm, 1, OP_LHS, SCALE_NONE, FP64, 3, rng)  # kernels loaded
    warm.close()
    ctx = Context("mechanics_3d_q2", text)
    ctx.set_option("generic_kernel", "two_pass")
    n = 3 * ctx.info.n_nodes
    src = torch.tensor(rng.uniform(-1, 1, n), device="cuda")
    dst = torch.zeros_like(src)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    ctx.apply_operator_noalloc(1, OP_LHS, SCALE_NONE, FP64, src.data_ptr(), dst.data_ptr())
    ctx.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert free1 == free0, (free0, free1)
    # the regular entry point's first call allocates its scratch (the check can see it)
    ctx.apply_operator(1, OP_LHS, SCALE_NONE, FP64, src.data_ptr(), dst.data_ptr())
    ctx.synchronize()
    assert torch.cuda.mem_get_info()[0] < free1
    ctx.close()


def test_noalloc_cell_block_needs_no_scratch(device):
    # The one-pass generic kernel keeps its per-cell results in shared memory: the
    # 64^3 elasticity operator that overflows the array on the two-pass kernels runs.
    rng = np.random.default_rng(11)
    ctx = Context("mechanics_3d_q2", prm.mechanics(refinement=6))
    a, b = both(ctx, 1, OP_LHS, SCALE_NONE, FP64, 3, rng)
    assert np.isfinite(a).all()
    assert np.array_equal(a, b), float(np.abs(a - b).max())
    ctx.close()
