"""Parity of the HIP path (through the C-ABI) with the CPU oracle on identical inputs.

Bars: operator application within 1e-12 relative in FP64 (1e-5 in FP32); explicit runs
within 1e-10 after several increments; CG iteration counts within +-1 of the oracle with
the same stopping rule; adjoint quantities within 1e-10 of the oracle's FP64 adjoint.
"""

from __future__ import annotations

import numpy as np
import pytest

from phasefield_amd import Context, PPFError, prm
from phasefield_amd.capi import (AGE_LATEST, FP32, FP64, OP_LHS, OP_LHS_PLAIN, OP_RHS,
                                 SCALE_INVM, SCALE_INVM_SQRT, SCALE_NONE)
from oracle.binding import Oracle
from tests.helpers import (block_to_full, block_to_stored, rel_err, to_full, to_stored)

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a, dtype=None):
    return torch.tensor(np.asarray(a), dtype=dtype or torch.float64, device="cuda")


def seed_fields(ctx: Context, orc: Oracle, rng, lo=0.0, hi=1.0):
    """Same random state (every field, every age) on both sides; images = masters."""
    info = ctx.info
    for f in range(info.n_fields):
        age = 0
        while True:
            vals = to_full(rng.uniform(lo, hi, info.n_nodes), info)
            try:
                ctx.set_field(f, vals, age)
            except Exception:
                break
            orc.set_field(f, vals, age)
            age += 1


# (app, prm text, [(block_id, which, n_block)])
OPERATOR_CASES = [
    ("allen_cahn_2d_q1", prm.allen_cahn(dim=2, size=6.0, refinement=2, subdivisions=[2, 3]),
     [(1, OP_RHS, 1), (2, OP_RHS, 1)]),
    ("allen_cahn_3d_q1", prm.allen_cahn(dim=3, size=4.0, refinement=2), [(1, OP_RHS, 1)]),
    ("cahn_hilliard_3d_q2", prm.cahn_hilliard(size=4.0, refinement=2),
     [(0, OP_RHS, 1), (1, OP_RHS, 1), (2, OP_RHS, 1)]),
    ("precipitate_3d_q1", prm.precipitate(size=8.0, refinement=2, subdivisions=(1, 1, 1)),
     [(0, OP_RHS, 1), (1, OP_RHS, 1), (2, OP_RHS, 1)]),
    ("poisson_2d_q2", prm.poisson(dim=2, refinement=3, mg_depth=2),
     [(1, OP_RHS, 1), (1, OP_LHS, 1), (1, OP_LHS_PLAIN, 1)]),
    ("poisson_3d_q2", prm.poisson(dim=3, refinement=2, mg_depth=2),
     [(1, OP_RHS, 1), (1, OP_LHS, 1), (1, OP_LHS_PLAIN, 1)]),
    ("testbed_2d_q2", prm.testbed(dim=2, refinement=3),
     [(0, OP_RHS, 1), (1, OP_RHS, 1), (1, OP_LHS, 1), (1, OP_LHS_PLAIN, 1), (2, OP_RHS, 1)]),
    ("testbed_3d_q1", prm.testbed(dim=3, refinement=2, dirichlet_faces=(1, 4)),
     [(0, OP_RHS, 1), (1, OP_LHS, 1), (1, OP_LHS_PLAIN, 1), (2, OP_RHS, 1)]),
]


@pytest.mark.parametrize("app,text,ops", OPERATOR_CASES, ids=[c[0] for c in OPERATOR_CASES])
@pytest.mark.parametrize("path", ["auto", "auto_node", "generic", "generic_two_pass",
                                  "split_node"])
def test_operator_matches_oracle(app, text, ops, path):
    """auto: tiled stencil kernel for split hooks (at any lattice size), generic cell
    path otherwise; auto_node: the default routing, in which split hooks on lattices
    below 100000 nodes run node by node;
    generic: cell path for every hook (one-pass cell-block kernel); generic_two_pass: the
    cell path's two-pass kernels (per-cell results through a global array, then a node
    gather); split_node: one-thread-per-node split kernel (hooks without the split
    structure are rejected with an error)."""
    ctx, orc = Context(app, text), Oracle(app, text)
    info = ctx.info
    rng = np.random.default_rng(7)
    seed_fields(ctx, orc, rng)
    ctx.set_time(3, 0.37)
    orc.set_time(3, 0.37)
    if path == "generic_two_pass":
        ctx.set_option("generic_kernel", "two_pass")
        path = "generic"
    if path == "auto_node":
        # split hooks on these small lattices run node by node by default
        path = "auto"
    else:
        ctx.set_option("stencil_min_nodes", 0)  # auto: the tiled kernel at any size
    ctx.set_option("operator_path", path)
    for block_id, which, nb in ops:
        for scaling in (SCALE_NONE, SCALE_INVM, SCALE_INVM_SQRT):
            src_stored = rng.uniform(-1, 1, nb * info.n_nodes)
            src_full = block_to_full(src_stored, info, nb)
            ref = orc.apply_operator(block_id, which, scaling, src_full, nb)
            d_src = dev(src_stored)
            d_dst = torch.zeros_like(d_src)
            if path == "split_node":
                try:
                    ctx.apply_operator(block_id, which, scaling, FP64, d_src.data_ptr(),
                                       d_dst.data_ptr())
                except PPFError as e:
                    assert "does not qualify" in str(e)
                    continue
            else:
                ctx.apply_operator(block_id, which, scaling, FP64, d_src.data_ptr(), d_dst.data_ptr())
            got = d_dst.cpu().numpy()
            err = rel_err(got, block_to_stored(ref, info, nb))
            print(app, path, block_id, which, scaling, err, flush=True)
            assert err <= 1e-12, (block_id, which, scaling, err)


# Lattices for the one-pass generic kernels: (2-D) blocks of 16 x 16 evaluated cells,
# 15 x 15 owned; (3-D) columns of 9 x 9 cells, 8 x 8 owned, marching over chunks of z
# layers in slabs of 3 (chunk 0: automatic; otherwise the given number of layers).
CELL_BLOCK_CASES = [
    # 128 x 16 x 16 Q2 cells, Dirichlet on every face: 32 columns, short chunks.
    ("poisson_3d_q2", prm.poisson(dim=3, refinement=4, mg_depth=1, subdivisions=[8, 1, 1]),
     [(1, OP_LHS, 1), (1, OP_LHS_PLAIN, 1), (1, OP_RHS, 1)], 0),
    # The same lattice turned: 16 x 16 x 128 cells, z chunks of 7 layers (3 slabs each,
    # the last one past the chunk), kept sums carried between slabs.
    ("poisson_3d_q2", prm.poisson(dim=3, refinement=4, mg_depth=1, subdivisions=[1, 1, 8]),
     [(1, OP_LHS, 1), (1, OP_LHS_PLAIN, 1)], 7),
    # Periodic, 2 cells per direction: fewer cells than a column, wrapped halo.
    ("cahn_hilliard_3d_q2", prm.cahn_hilliard(size=4.0, refinement=1),
     [(0, OP_RHS, 1), (1, OP_RHS, 1), (2, OP_RHS, 1)], 0),
    # Periodic, 16^3 Q2 cells in one chunk per column (6 slabs), wrapped first layer.
    ("cahn_hilliard_3d_q2", prm.cahn_hilliard(size=32.0, refinement=4),
     [(0, OP_RHS, 1), (1, OP_RHS, 1)], 16),
    # Periodic, 160 x 96 Q1 cells: ragged blocks along both axes.
    ("allen_cahn_2d_q1", prm.allen_cahn(dim=2, size=30.0, refinement=5, subdivisions=[5, 3]),
     [(1, OP_RHS, 1), (2, OP_RHS, 1)], 0),
    # Non-separable hooks, mixed Dirichlet/natural faces, 16^3 Q1 cells; chunks of 5.
    ("testbed_3d_q1", prm.testbed(dim=3, refinement=4, dirichlet_faces=(1, 4)),
     [(0, OP_RHS, 1), (1, OP_LHS, 1), (1, OP_LHS_PLAIN, 1), (2, OP_RHS, 1)], 5),
    ("testbed_3d_q1", prm.testbed(dim=3, refinement=4, dirichlet_faces=(0, 5)),
     [(0, OP_RHS, 1), (1, OP_LHS, 1), (2, OP_RHS, 1)], 0),
    ("testbed_2d_q2", prm.testbed(dim=2, refinement=5),
     [(0, OP_RHS, 1), (1, OP_RHS, 1), (1, OP_LHS, 1), (2, OP_RHS, 1)], 0),
    # Three coupled component fields (linear elasticity), 8^3 and 16^3 Q2 cells.
    ("mechanics_3d_q2", prm.mechanics(refinement=3), [(1, OP_LHS, 3), (1, OP_RHS, 3)], 0),
    ("mechanics_3d_q2", prm.mechanics(refinement=4), [(1, OP_LHS, 3)], 16),
]


@pytest.mark.parametrize("app,text,ops,chunk", CELL_BLOCK_CASES,
                         ids=[f"{c[0]}-{i}" for i, c in enumerate(CELL_BLOCK_CASES)])
def test_cell_block_kernel_matches_oracle_and_two_pass(app, text, ops, chunk):
    """The one-pass generic kernel against the oracle (1e-12) and against the two-pass
    kernels, which sum the same per-cell contributions in the same order (1e-14)."""
    ctx, orc = Context(app, text), Oracle(app, text)
    info = ctx.info
    rng = np.random.default_rng(17)
    seed_fields(ctx, orc, rng)
    ctx.set_time(2, 0.21)
    orc.set_time(2, 0.21)
    ctx.set_option("operator_path", "generic")
    ctx.set_option("generic_column_chunk", chunk)
    for block_id, which, nb in ops:
        for scaling in (SCALE_NONE, SCALE_INVM):
            src_stored = rng.uniform(-1, 1, nb * info.n_nodes)
            ref = block_to_stored(orc.apply_operator(block_id, which, scaling,
                                                     block_to_full(src_stored, info, nb), nb),
                                  info, nb)
            d_src = dev(src_stored)
            got = {}
            for kernel in ("cell_block", "two_pass"):
                ctx.set_option("generic_kernel", kernel)
                d_dst = torch.full_like(d_src, float("nan"))
                ptr = d_src.data_ptr() if which != OP_RHS else 0
                ctx.apply_operator(block_id, which, scaling, FP64, ptr, d_dst.data_ptr())
                got[kernel] = d_dst.cpu().numpy()
            err = rel_err(got["cell_block"], ref)
            diff = rel_err(got["cell_block"], got["two_pass"])
            print(app, block_id, which, scaling, err, diff,
                  np.array_equal(got["cell_block"], got["two_pass"]), flush=True)
            assert err <= 1e-12, (block_id, which, scaling, err)
            assert diff <= 1e-14, (block_id, which, scaling, diff)


# Tile shapes of the stencil kernel: (waves per workgroup, rows per wave, z-planes per
# workgroup; 0 = automatic; Q2 rounds odd chunks up to even).
TILINGS = [(4, 2, 0), (4, 4, 0), (4, 2, 5), (4, 4, 7), (4, 2, 64), (4, 4, 1),
           (8, 1, 0), (8, 2, 0), (8, 1, 3), (8, 2, 6)]
TILING_CASES = [
    # 257 x 33 x 33 nodes, Dirichlet on every face: tiles in the interior run the
    # variant without face tests, the others the masked variant.
    ("poisson_3d_q2", prm.poisson(dim=3, refinement=4, mg_depth=1, subdivisions=[8, 1, 1]),
     [(1, OP_LHS, 1), (1, OP_LHS_PLAIN, 1), (1, OP_RHS, 1)]),
    # Periodic lattice of 4 nodes per direction (2 cells): ring wrap-around within a tile.
    ("cahn_hilliard_3d_q2", prm.cahn_hilliard(size=4.0, refinement=1),
     [(0, OP_RHS, 1), (1, OP_RHS, 1)]),
    ("allen_cahn_2d_q1", prm.allen_cahn(dim=2, size=30.0, refinement=5, subdivisions=[5, 3]),
     [(1, OP_RHS, 1)]),
    # Blocks with two staged operands (4 x 2 and 8 x 1 tiles only).
    ("precipitate_3d_q1", prm.precipitate(size=8.0, refinement=2, subdivisions=(1, 1, 1)),
     [(0, OP_RHS, 1), (1, OP_RHS, 1), (2, OP_RHS, 1)]),
]


@pytest.mark.parametrize("app,text,ops", TILING_CASES, ids=[c[0] for c in TILING_CASES])
def test_stencil_tilings_match_oracle(app, text, ops):
    ctx, orc = Context(app, text), Oracle(app, text)
    info = ctx.info
    rng = np.random.default_rng(21)
    seed_fields(ctx, orc, rng)
    ctx.set_option("stencil_min_nodes", 0)
    for i, (waves, rows, zc) in enumerate(TILINGS):
        ctx.set_option("stencil_waves", waves)
        ctx.set_option("stencil_rows", rows)
        ctx.set_option("stencil_z_chunk", zc)
        # Both cache policies of the output stores, alternating over the tilings.
        ctx.set_option("stencil_store_policy", ("temporal", "nontemporal")[i % 2])
        for block_id, which, nb in ops:
            for scaling in (SCALE_NONE, SCALE_INVM, SCALE_INVM_SQRT):
                src_stored = rng.uniform(-1, 1, nb * info.n_nodes)
                ref = orc.apply_operator(block_id, which, scaling,
                                         block_to_full(src_stored, info, nb), nb)
                d_src = dev(src_stored)
                d_dst = torch.zeros_like(d_src)
                ctx.apply_operator(block_id, which, scaling, FP64, d_src.data_ptr(),
                                   d_dst.data_ptr())
                err = rel_err(d_dst.cpu().numpy(), block_to_stored(ref, info, nb))
                assert err <= 1e-12, (waves, rows, zc, i % 2, block_id, which, scaling, err)


@pytest.mark.parametrize("app,text,ops", OPERATOR_CASES[2:6], ids=[c[0] for c in OPERATOR_CASES[2:6]])
def test_operator_fp32_matches_oracle(app, text, ops):
    ctx, orc = Context(app, text), Oracle(app, text)
    info = ctx.info
    rng = np.random.default_rng(11)
    seed_fields(ctx, orc, rng)
    for block_id, which, nb in ops:
        src_stored = rng.uniform(-1, 1, nb * info.n_nodes)
        ref = orc.apply_operator(block_id, which, SCALE_NONE, block_to_full(src_stored, info, nb), nb)
        d_src = dev(src_stored, torch.float32)
        d_dst = torch.zeros_like(d_src)
        ctx.apply_operator(block_id, which, SCALE_NONE, FP32, d_src.data_ptr(), d_dst.data_ptr())
        err = rel_err(d_dst.double().cpu().numpy(), block_to_stored(ref, info, nb))
        assert err <= 1e-5, (block_id, which, err)


DIAGONAL_CASES = [
    ("poisson_2d_q2", prm.poisson(dim=2, refinement=3, mg_depth=2), 1),
    ("poisson_3d_q2", prm.poisson(dim=3, refinement=2, mg_depth=2), 1),
    ("testbed_2d_q2", prm.testbed(dim=2, refinement=3), 1),
    ("testbed_3d_q1", prm.testbed(dim=3, refinement=2), 1),
]


@pytest.mark.parametrize("app,text,block_id", DIAGONAL_CASES, ids=[c[0] for c in DIAGONAL_CASES])
@pytest.mark.parametrize("scaling", [SCALE_NONE, SCALE_INVM_SQRT])
def test_inverse_diagonal_matches_oracle(app, text, block_id, scaling):
    ctx, orc = Context(app, text), Oracle(app, text)
    info = ctx.info
    seed_fields(ctx, orc, np.random.default_rng(3))
    ref = orc.inverse_diagonal(block_id, scaling)
    d = torch.zeros(info.n_nodes, dtype=torch.float64, device="cuda")
    ctx.inverse_diagonal(block_id, scaling, FP64, d.data_ptr())
    assert rel_err(d.cpu().numpy(), to_stored(ref, info)) <= 1e-12


RUN_CASES = [
    ("allen_cahn_2d_q1", prm.allen_cahn(dim=2, size=16.0, refinement=4, steps=20, dt=0.01, outputs=4), 20),
    ("cahn_hilliard_3d_q2", prm.cahn_hilliard(size=8.0, refinement=2, steps=10, dt=1e-3, outputs=2), 10),
    ("precipitate_3d_q1", prm.precipitate(size=16.0, refinement=3, subdivisions=(1, 1, 1), steps=10, outputs=2), 10),
    ("testbed_2d_q2", prm.testbed(dim=2, refinement=3, steps=5), 5),
    ("testbed_3d_q1", prm.testbed(dim=3, refinement=2, steps=3), 3),
    # Time-dependent Dirichlet values (explicit c and linear v re-constrained every
    # increment, reference constraint_manager.cc:380-423).
    ("testbed_2d_q2", prm.testbed(dim=2, refinement=3, steps=5, dt=0.01, time_dependent=True), 5),
    ("testbed_3d_q1", prm.testbed(dim=3, refinement=2, steps=3, dt=0.01, dirichlet_faces=(1, 4),
                                  time_dependent=True), 3),
]
RUN_IDS = [c[0] + ("_time_dependent_bc" if "time dependent = true" in c[1] else "")
           for c in RUN_CASES]


@pytest.mark.parametrize("app,text,steps", RUN_CASES, ids=RUN_IDS)
def test_run_matches_oracle(app, text, steps):
    """Whole increment loop (ICs, blocks in order, postprocess, rotation)."""
    ctx, orc = Context(app, text), Oracle(app, text)
    assert ctx.run(steps) == 0
    assert orc.run(steps) == 0
    for f in range(ctx.info.n_fields):
        got, ref = ctx.get_field(f, AGE_LATEST).ravel(), orc.get_field(f, AGE_LATEST)
        err = rel_err(got, ref)
        print(app, "field", f, err, flush=True)
        assert err <= 1e-10, (f, err)
    # Output diagnostics: l2-norm over deal.II's node set and integral.
    for f in range(ctx.info.n_fields):
        a, b = ctx.field_norms(f, AGE_LATEST), orc.field_norms(f, AGE_LATEST)
        assert abs(a[0] - b[0]) <= 1e-10 * max(1.0, abs(b[0]))
        assert abs(a[1] - b[1]) <= 1e-10 * max(1.0, abs(b[1]))


FP32_EXPLICIT_CASES = [
    ("precipitate_3d_q1", prm.precipitate(size=16.0, refinement=3, subdivisions=(1, 1, 1), steps=20, outputs=1), 20),
    ("cahn_hilliard_3d_q2", prm.cahn_hilliard(size=8.0, refinement=2, steps=20, dt=1e-3, outputs=1), 20),
]


@pytest.mark.parametrize("app,text,steps", FP32_EXPLICIT_CASES, ids=[c[0] for c in FP32_EXPLICIT_CASES])
def test_fp32_operator_fp64_state_error(app, text, steps):
    """explicit_precision = single: FP32 operator arithmetic, FP64 state.  Error against
    the FP64 oracle after the run (max over fields of max |error| / max |field|): at most
    1e-4 (FP32 rounding of the inputs of every increment; the CH mu block's Laplacian of
    c ~ 0.5 cancels to a few 1e-5, the precipitate model stays near 1e-6), and above the
    FP64 path's error (the FP32 path really ran)."""
    orc = Oracle(app, text)
    assert orc.run(steps) == 0
    errs = {}
    for precision in ("double", "single"):
        ctx = Context(app, text)
        ctx.set_option("explicit_precision", precision)
        assert ctx.run(steps) == 0
        errs[precision] = max(rel_err(ctx.get_field(f, AGE_LATEST).ravel(), orc.get_field(f, AGE_LATEST))
                              for f in range(ctx.info.n_fields))
    print(app, "error vs FP64 oracle:", errs, flush=True)
    assert errs["double"] <= 1e-10
    assert errs["single"] <= 1e-4
    assert errs["single"] > errs["double"]


@pytest.mark.parametrize("app,text,steps", FP32_EXPLICIT_CASES, ids=[c[0] for c in FP32_EXPLICIT_CASES])
def test_fp32_stencil_on_fp64_state_matches_fp32_copies(app, text, steps):
    """explicit_precision = single runs the stencil kernel with FP32 arithmetic on the
    FP64 state (loads rounded to FP32 in registers, FP64 stores) after a first increment
    through FP32 copies; single_copies rounds the inputs into FP32 copies every
    increment.  Both round every input once and evaluate the same FP32 arithmetic, so
    the states agree to a few FP32 roundings (bound 1e-6 of the field's size; the FP32
    path's own error is ~1e-6 .. 1e-5), the log names the block that took the kernel, and
    the result keeps the single-precision error bound against the FP64 oracle."""
    orc = Oracle(app, text)
    assert orc.run(steps) == 0
    res, logs = {}, {}
    for mode in ("single", "single_copies"):
        ctx = Context(app, text)
        ctx.set_option("explicit_precision", mode)
        ctx.set_option("stencil_min_nodes", "0")
        assert ctx.run(steps) == 0
        res[mode] = [ctx.get_field(f, AGE_LATEST).ravel().copy() for f in range(ctx.info.n_fields)]
        logs[mode] = ctx.output_log()
        ctx.close()
    print(logs["single"], flush=True)
    assert "FP32 stencil arithmetic on the FP64 state" in logs["single"]
    assert "FP32 stencil arithmetic" not in logs["single_copies"]
    for f, (a, b) in enumerate(zip(res["single"], res["single_copies"])):
        d = float(np.abs(a - b).max())
        scale = float(np.abs(b).max())
        print(app, "field", f, "single vs single_copies max diff", d, "exact", d == 0.0,
              flush=True)
        assert d <= 1e-6 * scale
        assert rel_err(a, orc.get_field(f, AGE_LATEST)) <= 1e-4


LINEAR_CASES = [
    ("poisson_2d_q2", dict(dim=2, refinement=4, preconditioner="None", mg_depth=1)),
    ("poisson_2d_q2", dict(dim=2, refinement=4, preconditioner="Chebyshev", mg_depth=1)),
    ("poisson_2d_q2", dict(dim=2, refinement=4, preconditioner="GMG", mg_depth=3, coarse_solver="smoother")),
    ("poisson_2d_q2", dict(dim=2, refinement=4, preconditioner="GMG", mg_depth=4, coarse_solver="cg")),
    ("poisson_3d_q2", dict(dim=3, refinement=3, preconditioner="GMG", mg_depth=3, coarse_solver="cg")),
    ("poisson_2d_q2", dict(dim=2, refinement=4, preconditioner="None", tolerance_type="RMSEPerField", size=2.0)),
]


@pytest.mark.parametrize("app,kw", LINEAR_CASES)
def test_linear_solve_matches_oracle(app, kw):
    text = prm.poisson(tolerance=1e-9, **kw)
    ctx, orc = Context(app, text), Oracle(app, text)
    ctx.run(0)
    orc.run(0)
    st = ctx.last_solve_stats(1)
    orc2 = Oracle(app, text)
    orc2.solve_block(0)
    ost = orc2.linear_solve(1)
    print(app, kw, "gpu its", st.iterations, "oracle its", ost["iterations"], flush=True)
    assert st.converged and ost["converged"]
    assert abs(st.iterations - ost["iterations"]) <= 1
    u_gpu, u_ref = ctx.get_field(0, AGE_LATEST).ravel(), orc.get_field(0, AGE_LATEST)
    assert rel_err(u_gpu, u_ref) <= 1e-6


def test_testbed_gmg_variable_coefficient(device):
    """Generic operator, generic diagonal, non-src dependency injected to the levels."""
    text = prm.testbed(dim=2, refinement=4, preconditioner="GMG", mg_depth=3, steps=0)
    ctx, orc = Context("testbed_2d_q2", text), Oracle("testbed_2d_q2", text)
    ctx.run(0)
    orc.run(0)
    st = ctx.last_solve_stats(1)
    orc2 = Oracle("testbed_2d_q2", text)
    orc2.solve_block(0)
    ost = orc2.linear_solve(1)
    assert st.converged and abs(st.iterations - ost["iterations"]) <= 1
    assert rel_err(ctx.get_field(1, AGE_LATEST).ravel(), orc.get_field(1, AGE_LATEST)) <= 1e-7


def test_mixed_precision_gmg_reaches_fp64_tolerance(device):
    text = prm.poisson(dim=3, refinement=3, mg_depth=3, tolerance=1e-9, level_precision="single")
    ctx = Context("poisson_3d_q2", text)
    ctx.run(0)
    st = ctx.last_solve_stats(1)
    orc = Oracle("poisson_3d_q2", prm.poisson(dim=3, refinement=3, mg_depth=3, tolerance=1e-9))
    orc.run(0)
    assert st.converged and st.final_residual <= st.tolerance
    assert rel_err(ctx.get_field(0, AGE_LATEST).ravel(), orc.get_field(0, AGE_LATEST)) <= 1e-6


def test_linear_initial_guess_option(device):
    """A repeated solve starts from the previous solution (the reference's warm start,
    0 iterations); with linear_initial_guess=zero it repeats the first solve exactly."""
    text = prm.poisson(dim=3, refinement=3, mg_depth=3, tolerance=1e-9)
    ctx = Context("poisson_3d_q2", text)
    ctx.run(0)
    first = ctx.last_solve_stats(1)
    u1 = ctx.get_field(0, 0).copy()
    warm = ctx.linear_solve(1)
    assert first.iterations > 0 and warm.iterations == 0
    ctx.set_option("linear_initial_guess", "zero")
    cold = ctx.linear_solve(1)
    assert cold.iterations == first.iterations
    assert np.array_equal(ctx.get_field(0, 0), u1)
    with pytest.raises(PPFError):
        ctx.set_option("linear_initial_guess", "random")


def _adj_text(dim, periodic, subdivisions=None):
    text = prm.allen_cahn(dim=dim, size=8.0, refinement=3, steps=0, dt=0.01, kappa=1.5,
                          subdivisions=subdivisions)
    return text if periodic else text.replace("periodic = true", "periodic = false")


# Periodic and mirror-ghost (natural boundary) lattices, cubic and with unequal node
# counts per axis (catches a transposed plane index), in 3D and 2D.
ADJ_CASES = [
    ("allen_cahn_3d_q1", _adj_text(3, True)),
    ("allen_cahn_3d_q1", _adj_text(3, False, [2, 1, 3])),
    ("allen_cahn_2d_q1", _adj_text(2, True, [3, 1])),
    ("allen_cahn_2d_q1", _adj_text(2, False, [1, 2])),
]


@pytest.mark.parametrize("app,text", ADJ_CASES,
                         ids=["3d_periodic", "3d_natural_2x1x3", "2d_periodic_3x1", "2d_natural_1x2"])
def test_adjoint_matches_oracle(device, app, text, precision=FP64):
    ctx, orc = Context(app, text), Oracle(app, text)
    info = ctx.info
    rng = np.random.default_rng(5)
    phi0 = rng.uniform(-1, 1, info.n_nodes)
    lam = rng.uniform(-1, 1, info.n_nodes)
    n, L, kappa = 12, 1.1, 1.3
    d_phi0, d_out = dev(phi0), torch.zeros(info.n_nodes, dtype=torch.float64, device="cuda")
    ctx.adjoint_forward(d_phi0.data_ptr(), L, kappa, n, precision, d_out.data_ptr())
    ref = orc.adjoint_forward(to_full(phi0, info), L, kappa, n)
    assert rel_err(d_out.cpu().numpy(), to_stored(ref, info)) <= 1e-12
    for every in (1, 5, 12, 50):
        d_lam, d_g = dev(lam), torch.zeros(info.n_nodes, dtype=torch.float64, device="cuda")
        dl, dk = ctx.adjoint_vjp(d_phi0.data_ptr(), d_lam.data_ptr(), L, kappa, n, every,
                                 precision, d_g.data_ptr())
        lam_full = np.zeros(info.n_dofs)
        lam_full[:] = to_full(lam, info)
        # The oracle folds image entries onto masters: pass zeros at images.
        lam_masters = to_stored(lam_full, info)
        rdl, rdk, rg = orc.adjoint_vjp(to_full(phi0, info), _zero_images(lam_masters, info), L, kappa, n)
        assert abs(dl - rdl) <= 1e-10 * abs(rdl)
        assert abs(dk - rdk) <= 1e-10 * abs(rdk)
        assert rel_err(d_g.cpu().numpy(), to_stored(rg, info)) <= 1e-10


def _ch_adj_text(periodic, subdivisions=None, dt=1e-3):
    # dt below the explicit stability limit ~ h^4 / (M kappa) of the finest axis
    text = prm.cahn_hilliard(size=8.0, refinement=2, steps=0, dt=dt, subdivisions=subdivisions)
    return text if periodic else text.replace("periodic = true", "periodic = false")


CH_ADJ_CASES = [_ch_adj_text(True), _ch_adj_text(False, [2, 1, 3], dt=2e-5)]


@pytest.mark.parametrize("text", CH_ADJ_CASES, ids=["3d_periodic", "3d_natural_2x1x3"])
def test_cahn_hilliard_adjoint_matches_oracle(device, text):
    """Adjoint of the explicit Cahn-Hilliard step (c, mu blocks) against the oracle's:
    forward within 1e-12, dJ/dM, dJ/dkappa and dJ/dc0 within 1e-10, for several
    checkpoint spacings."""
    app = "cahn_hilliard_3d_q2"
    ctx, orc = Context(app, text), Oracle(app, text)
    info = ctx.info
    rng = np.random.default_rng(6)
    c0 = 0.5 + rng.uniform(-0.2, 0.2, info.n_nodes)
    lam = rng.uniform(-1, 1, info.n_nodes)
    n, mob, kappa = 8, 1.1, 1.3
    d_c0, d_out = dev(c0), torch.zeros(info.n_nodes, dtype=torch.float64, device="cuda")
    ctx.adjoint_forward(d_c0.data_ptr(), mob, kappa, n, FP64, d_out.data_ptr())
    ref = orc.adjoint_forward(to_full(c0, info), mob, kappa, n)
    assert rel_err(d_out.cpu().numpy(), to_stored(ref, info)) <= 1e-12
    rdm, rdk, rg = orc.adjoint_vjp(to_full(c0, info), _zero_images(lam, info), mob, kappa, n)
    for every in (1, 3, 8, 50):
        d_lam, d_g = dev(lam), torch.zeros(info.n_nodes, dtype=torch.float64, device="cuda")
        dm, dk = ctx.adjoint_vjp(d_c0.data_ptr(), d_lam.data_ptr(), mob, kappa, n, every, FP64,
                                 d_g.data_ptr())
        print("ch adjoint", every, dm, rdm, dk, rdk, flush=True)
        assert abs(dm - rdm) <= 1e-10 * abs(rdm)
        assert abs(dk - rdk) <= 1e-10 * abs(rdk)
        assert rel_err(d_g.cpu().numpy(), to_stored(rg, info)) <= 1e-10


def test_cahn_hilliard_explicit_matches_adjoint_forward(device):
    """The Cahn-Hilliard adjoint's forward map equals the application's c and mu blocks."""
    n = 4
    text = prm.cahn_hilliard(size=8.0, refinement=3, steps=n, dt=1e-3)
    ctx = Context("cahn_hilliard_3d_q2", text)
    ctx.run(0)
    c0 = to_stored(ctx.get_field(0, AGE_LATEST).ravel(), ctx.info)
    ctx.run(n)
    cn = to_stored(ctx.get_field(0, AGE_LATEST).ravel(), ctx.info)
    d0, dn = dev(c0), torch.zeros(ctx.info.n_nodes, dtype=torch.float64, device="cuda")
    ctx.adjoint_forward(d0.data_ptr(), 1.0, 1.0, n, FP64, dn.data_ptr())
    assert rel_err(dn.cpu().numpy(), cn) <= 1e-12


def _zero_images(stored, info):
    full = to_full(stored, info).reshape(tuple(reversed([info.degree * info.cells[d] + 1 for d in range(info.dim)])))
    for ax in range(info.dim):
        d = info.dim - 1 - ax
        if info.periodic[d]:
            idx = [slice(None)] * info.dim
            idx[ax] = -1
            full[tuple(idx)] = 0.0
    return full.ravel()


def test_explicit_matches_adjoint_forward(device):
    """The adjoint module's forward step equals the application's explicit block."""
    text = prm.allen_cahn(dim=3, size=8.0, refinement=3, steps=4, dt=0.01, L=1.0, kappa=1.5)
    ctx = Context("allen_cahn_3d_q1", text)
    ctx.run(0)
    phi0 = to_stored(ctx.get_field(0, AGE_LATEST).ravel(), ctx.info)
    ctx.run(4)
    phi4 = to_stored(ctx.get_field(0, AGE_LATEST).ravel(), ctx.info)
    d0, d4 = dev(phi0), torch.zeros(ctx.info.n_nodes, dtype=torch.float64, device="cuda")
    ctx.adjoint_forward(d0.data_ptr(), 1.0, 1.5, 4, FP64, d4.data_ptr())
    assert rel_err(d4.cpu().numpy(), phi4) <= 1e-12


def test_solution_files_at_output_increments(device, tmp_path):
    """Option write_output: the reference's solution files (`<directory>/<file
    name>_<increment>.vtu` at every output increment, problem.cc:500-515) hold every
    field's state of that increment."""
    from tests.test_vtu_writer import read_vtu

    text = prm.allen_cahn(dim=2, size=16.0, refinement=4, steps=4, dt=0.01, outputs=2)
    text = text.replace("subsection output\n", f"subsection output\n  set directory = {tmp_path}\n")
    ctx = Context("allen_cahn_2d_q1", text)
    ctx.set_option("write_output", "1")
    assert ctx.run(4) == 0
    names = sorted(p.name for p in tmp_path.glob("*.vtu"))
    assert names == ["solution_0.vtu", "solution_2.vtu", "solution_4.vtu"]
    arrs, n_points, _ = read_vtu(tmp_path / "solution_4.vtu")
    assert n_points == ctx.info.n_dofs
    for f, name in enumerate(("n", "f_tot")):
        np.testing.assert_array_equal(arrs[name], ctx.get_field(f, AGE_LATEST).ravel())
    assert arrs["CYCLE"][0] == 4
