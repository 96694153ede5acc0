#!/usr/bin/env python
"""Benchmark of the MI355X matrix-free operator path.

Workload (one GPU): the Cahn-Hilliard 3D explicit configuration -- Q2 elements,
64^3 cells on [0,64]^3, periodic, fields c and mu, f = c^2(1-c)^2, M = kappa = 1,
dt = 1e-4.  A step is one increment: the c block (reads c_old, mu_old; writes c) and
the mu block (reads c; writes mu), each one operator application over
N_d = 129^3 = 2,146,689 DoFs.

Printed JSON line:
  value        operator DoFs/s = 2 * N_d * steps / wall time of the timed steps
               (inputs resident in HBM, increment loop driven through the C-ABI);
  roofline     dominant kernel (the c-block operator): algorithmic bytes per launch
               (3 arrays x N_d x 8 B) / average launch time from HIP events on the
               library's stream, against MI355X's 8 TB/s;
  cpu_baseline the CPU oracle (FP64 restatement of the reference algorithm, OpenMP) on
               the host cores, timed on a bounded number of increments of the same
               configuration;
  extra        Poisson 3D Q2 128^3: FP64 and FP32 operator application rooflines and
               conjugate-gradient + geometric-multigrid time to 1e-8 and 1e-10 relative
               residual, all-FP64 versus FP32 levels with FP64 outer iteration; Allen-Cahn
               3D adjoint: forward and vector-Jacobian-product times (FP64 and
               FP32-forward) with their algorithmic GB/s (16 and 40 B per DoF and step);
               implicit Allen-Cahn 2D (Newton blocks): time, Newton updates and CG steps
               per increment, without and with GMG; implicit Cahn-Hilliard 2D (Newton
               with BiCGStab) at the reference application's parameters; coupled
               precipitate 3D: increment time in FP64 and FP32-operator/FP64-state, and
               the difference between the two.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--cpu-seconds", type=float, default=15.0,
                   help="target duration of the CPU-baseline sample (0 disables it)")
    p.add_argument("--no-extra", action="store_true", help="skip every extra measurement")
    p.add_argument("--extras", default="poisson,adjoint,newton,precipitate,mechanics",
                   help="comma-separated extra measurements to run")
    p.add_argument("--kernel-reps", type=int, default=50)
    p.add_argument("--copy-calibration", action="store_true",
                   help="also run a plain device copy of 16,974,593 doubles (counter "
                        "calibration for rocprofv3 FETCH_SIZE / WRITE_SIZE runs)")
    return p.parse_args()


def cpu_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def time_kernel(ctx, torch, block_id, which, scaling, d_src, d_dst, reps, flush=None,
                precision=None):
    """Average duration of one operator launch, from HIP events on the library's stream.

    flush: optional device tensor read between launches (its 1 GiB evicts L2 and the
    Infinity Cache with clean lines, so no write-back lands in the timed launch); the
    launches are then timed one by one."""
    from phasefield_amd.capi import FP64

    prec = FP64 if precision is None else precision
    stream = torch.cuda.ExternalStream(ctx.stream())
    ctx.set_option("synchronize", "0")
    src = d_src.data_ptr() if d_src is not None else 0
    for _ in range(3):
        ctx.apply_operator(block_id, which, scaling, prec, src, d_dst.data_ptr())
    ctx.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    if flush is None:
        start.record(stream)
        for _ in range(reps):
            ctx.apply_operator(block_id, which, scaling, prec, src, d_dst.data_ptr())
        stop.record(stream)
        stop.synchronize()
        total = start.elapsed_time(stop)
    else:
        total = 0.0
        for r in range(reps):
            with torch.cuda.stream(stream):
                flush.sum()
            start.record(stream)
            ctx.apply_operator(block_id, which, scaling, prec, src, d_dst.data_ptr())
            stop.record(stream)
            stop.synchronize()
            total += start.elapsed_time(stop)
    ctx.set_option("synchronize", "1")
    return total / reps * 1e-3  # seconds


def profiled(name, key="traffic_bytes_per_launch"):
    """A per-launch figure from a committed rocprofv3 counter summary: HBM bytes
    (FETCH_SIZE and WRITE_SIZE scaled by the copy-kernel calibration) or FP64 flops (VALU
    FMA/ADD/MUL instruction counts); None if the summary or the figure is absent."""
    path = Path(__file__).resolve().parent / "profiles" / name
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(steps_target_s: float):
    """The CPU baseline on the host cores: oracle/baseline/ch_baseline.cc, the reference's
    matrix-free cell loop for this workload (8-lane SIMD cell batches, sum-factorized
    evaluate/integrate, scatter-add, OpenMP over coloured z-slabs), checked against the
    oracle by tests/test_cpu_baseline.py.  Timed on a bounded number of increments of the
    full 64^3 Q2 problem."""
    from oracle import baseline

    cores = cpu_cores()
    n_cells, n_d = 64, 129 ** 3
    rng = np.random.default_rng(2025)
    c0 = 0.5 + rng.uniform(-0.01, 0.01, (2 * n_cells) ** 3)
    _, _, t1 = baseline.run(c0, n_cells, 1.0, 1e-4, 1.0, 1.0, 2, threads=cores)
    n = max(2, int(steps_target_s / max(t1 / 2, 1e-6)))
    _, _, sec = baseline.run(c0, n_cells, 1.0, 1e-4, 1.0, 1.0, n, threads=cores)
    _, _, sec1 = baseline.run(c0, n_cells, 1.0, 1e-4, 1.0, 1.0, 2, threads=1)
    isa = "AVX-512" if baseline.avx512() else "AVX2"
    value = 2 * n_d * n / sec
    per_core_1 = 2 * n_d * 2 / sec1
    return {
        "value": value,
        "unit": "DoF/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{n} increments (c and mu blocks) of the full 64^3 Q2 problem in {sec:.1f} s "
                  f"on {cores} OpenMP threads, {isa} 8-lane cell batches "
                  f"(oracle/baseline/ch_baseline.cc); {value / cores:.3g} DoF/s per core at "
                  f"{cores} threads, {per_core_1:.3g} DoF/s on one core (the reference's own "
                  f"transcript: 1.05e8 DoF-updates/s on one AVX-512 core for 2-D Q2 "
                  f"Allen-Cahn, tests/regression_tests/prismspf_3_6/gold_output.txt:69)",
    }


def poisson_extra(torch):
    """Poisson 128^3 Q2: FP64 operator roofline and GMG-CG time-to-solution."""
    from phasefield_amd import Context, prm
    from phasefield_amd.capi import OP_LHS, OP_RHS, SCALE_NONE

    out = {}
    text = prm.poisson(dim=3, refinement=7, mg_depth=7, tolerance=1e-8, coarse_solver="cg",
                       smoother_degree=4)
    ctx = Context("poisson_3d_q2", text)
    info = ctx.info
    n_d = info.n_dofs
    ctx.solve_block(0)  # the right-hand side field f
    # ||b|| for the relative criterion: b = M f with Dirichlet rows dropped.
    b = torch.zeros(info.n_nodes, dtype=torch.float64, device="cuda")
    ctx.apply_operator(1, OP_RHS, SCALE_NONE, 0, 0, b.data_ptr())
    bnorm = float(torch.linalg.vector_norm(b))
    x = torch.randn(info.n_nodes, dtype=torch.float64, device="cuda")
    y = torch.zeros_like(x)
    t_apply = time_kernel(ctx, torch, 1, OP_LHS, SCALE_NONE, x, y, 20)
    bytes_apply = 2 * 8 * n_d
    out["fp64_apply"] = {
        "workload": "Poisson Q2 128^3 lhs application (grad.grad), FP64, zero Dirichlet",
        "kernel": "stencil_operator_kernel (Poisson lhs block, Q2, FP64)",
        "dofs": n_d,
        "seconds": t_apply,
        "dofs_per_s": n_d / t_apply,
        "bytes_per_launch": bytes_apply,
        "achieved_gbs": bytes_apply / t_apply / 1e9,
        "frac": bytes_apply / t_apply / 1e9 / HBM_PEAK_GBS,
        "traffic": profiled("stencil_poisson_q2_128_fp64_pmc.md"),
    }
    pflops = profiled("stencil_poisson_q2_128_fp64_pmc.md", "fp64_flops_per_launch")
    if pflops is not None:
        out["fp64_apply"]["fp64_tflops"] = pflops / t_apply / 1e12
    del x, y
    # The same application in FP32 (the multigrid levels' arithmetic).
    from phasefield_amd.capi import FP32
    xf = torch.randn(info.n_nodes, dtype=torch.float32, device="cuda")
    yf = torch.zeros_like(xf)
    t32 = time_kernel(ctx, torch, 1, OP_LHS, SCALE_NONE, xf, yf, 20, precision=FP32)
    out["fp32_apply"] = {
        "workload": "Poisson Q2 128^3 lhs application (grad.grad), FP32, zero Dirichlet",
        "seconds": t32,
        "dofs_per_s": n_d / t32,
        "bytes_per_launch": 2 * 4 * n_d,
        "achieved_gbs": 2 * 4 * n_d / t32 / 1e9,
        "frac": 2 * 4 * n_d / t32 / 1e9 / HBM_PEAK_GBS,
    }
    del xf, yf
    # Per level precision: one context; the first solve includes the one-time
    # allocations; then each tolerance is solved from a zero initial guess and timed
    # (every solve re-evaluates the level diagonals and eigenvalue estimates, as the
    # reference does).
    for rel_tol in (1e-8, 1e-10):
        out[f"gmg_cg_to_{rel_tol:g}"] = {}
    for label, precision in (("fp64_levels", "double"), ("fp32_levels", "single")):
        c = Context("poisson_3d_q2", text)
        c.set_option("mg_level_precision", precision)
        c.solve_block(0)
        c.set_time(1, 0.0)
        c.set_option("linear_tolerance", 1e-8 * bnorm)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        c.linear_solve(1)
        torch.cuda.synchronize()
        first = time.perf_counter() - t0
        c.set_option("linear_initial_guess", "zero")
        for rel_tol in (1e-8, 1e-10):
            c.set_option("linear_tolerance", rel_tol * bnorm)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st = c.linear_solve(1)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            out[f"gmg_cg_to_{rel_tol:g}"][label] = {
                "seconds": dt,
                "first_solve_seconds_with_setup": first,
                "iterations": st.iterations,
                "relative_residual": st.final_residual / bnorm,
                "converged": bool(st.converged),
            }
        del c
    for rel_tol in (1e-8, 1e-10):
        r = out[f"gmg_cg_to_{rel_tol:g}"]
        r["fp32_levels_speedup"] = r["fp64_levels"]["seconds"] / r["fp32_levels"]["seconds"]
    del ctx
    return out


def adjoint_extra(torch):
    """Allen-Cahn 3D adjoint configuration: Q1, 128^3 cells, 500 steps, checkpoints every
    50 steps; forward run and vector-Jacobian product (recompute + adjoint sweep)."""
    from phasefield_amd import Context, prm
    from phasefield_amd.capi import AGE_LATEST, FP32, FP64

    ctx = Context("allen_cahn_3d_q1", prm.config_prm("allen_cahn_adjoint_3d"))
    ctx.run(0)  # seeded initial condition
    info = ctx.info
    n_d, n, every = info.n_dofs, 500, 50
    full = ctx.get_field(0, AGE_LATEST).ravel()
    dims = tuple(reversed([info.degree * info.cells[d] + 1 for d in range(info.dim)]))
    stored = np.ascontiguousarray(full.reshape(dims)[: info.nodes[2], : info.nodes[1], : info.nodes[0]]).ravel()
    phi0 = torch.tensor(stored, dtype=torch.float64, device="cuda")
    out = torch.empty_like(phi0)
    lam = torch.randn_like(phi0)
    dphi = torch.empty_like(phi0)
    res = {"workload": "Allen-Cahn 3D adjoint, Q1, 128^3 cells, 500 steps of dt=0.01, "
                       "checkpoints every 50 steps"}
    for label, prec in (("fp64", FP64), ("fp32_forward", FP32)):
        ctx.adjoint_forward(phi0.data_ptr(), 1.0, 1.5, 5, prec, out.data_ptr())  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.adjoint_forward(phi0.data_ptr(), 1.0, 1.5, n, prec, out.data_ptr())
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        t0 = time.perf_counter()
        ctx.adjoint_vjp(phi0.data_ptr(), lam.data_ptr(), 1.0, 1.5, n, every, prec, dphi.data_ptr())
        torch.cuda.synchronize()
        t_vjp = time.perf_counter() - t0
        res[label] = {
            "forward_seconds": t_fwd,
            "forward_us_per_step": t_fwd / n * 1e6,
            "forward_gbs": 16 * n_d * n / t_fwd / 1e9,
            "vjp_seconds": t_vjp,
            "vjp_us_per_step": t_vjp / n * 1e6,
            "vjp_gbs": 40 * n_d * n / t_vjp / 1e9,
        }
    del ctx
    return res


def precipitate_extra(torch):
    """Coupled Allen-Cahn/Cahn-Hilliard precipitate configuration (Q1, 192^3 cells on
    [0,192]^3, periodic, fields c, mu, eta): increment time in FP64 and with FP32 operator
    arithmetic on the FP64 state, and the largest difference between the two states after
    the same increments."""
    from phasefield_amd import Context, prm
    from phasefield_amd.capi import AGE_LATEST

    n_warm, n_inc = 5, 40
    res = {"workload": "Coupled Allen-Cahn/Cahn-Hilliard precipitate 3D, Q1, 192^3 cells on "
                       f"[0,192]^3, fields c, mu, eta, dt=5e-3, {n_inc} increments after "
                       f"{n_warm} warm-up increments",
           "bytes_per_dof_per_increment": 72}
    states = {}
    for label, precision in (("fp64", "double"), ("fp32_operator_fp64_state", "single")):
        ctx = Context("precipitate_3d_q1",
                      prm.config_prm("precipitate_3d", steps=n_warm + n_inc + 1, outputs=0))
        ctx.set_option("explicit_precision", precision)
        if ctx.run(n_warm) != 0:
            raise RuntimeError("precipitate warm-up stopped")
        ctx.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if ctx.run(n_warm + n_inc) != 0:
            raise RuntimeError("precipitate increments stopped")
        ctx.synchronize()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        n_d = ctx.info.n_dofs
        res[label] = {
            "us_per_increment": dt / n_inc * 1e6,
            "dofs_per_s": 3 * n_d * n_inc / dt,
            "algorithmic_gbs": 72 * n_d * n_inc / dt / 1e9,
        }
        states[label] = [ctx.get_field(f, AGE_LATEST) for f in range(3)]
        del ctx
    res["fp32_operator_max_abs_difference"] = max(
        float(np.abs(a - b).max()) for a, b in zip(states["fp64"], states["fp32_operator_fp64_state"]))
    return res


def mechanics_extra(torch):
    """Linear elasticity (one vector field, three components) at the reference
    application's parameters (applications/mechanics/boundary_value_problem/
    parameters.prm: 64^3 Q2 cells on [0,100]^3, RMSEPerField 1e-10, CG without a
    preconditioner), and with a Chebyshev preconditioner."""
    from phasefield_amd import Context, prm
    from phasefield_amd.capi import OP_LHS, SCALE_NONE

    res = {"workload": "Linear elasticity 3D, Q2, 64^3 cells on [0,100]^3, fields u_x, u_y, "
                       "u_z, E=2, nu=0.3, CG to RMSEPerField 1e-10"}
    for label, pc in (("no_preconditioner", None), ("chebyshev", "Chebyshev")):
        ctx = Context("mechanics_3d_q2", prm.mechanics(preconditioner=pc))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.run(0)
        ctx.synchronize()
        secs = time.perf_counter() - t0
        st = ctx.last_solve_stats(1)
        r = {"seconds_with_setup": secs, "iterations": st.iterations,
             "converged": bool(st.converged), "ms_per_iteration": secs / max(1, st.iterations) * 1e3}
        if pc is None:
            n = 3 * ctx.info.n_nodes
            src = torch.randn(n, dtype=torch.float64, device="cuda")
            dst = torch.zeros_like(src)
            t = time_kernel(ctx, torch, 1, OP_LHS, SCALE_NONE, src, dst, 20)
            r["lhs_apply_us"] = t * 1e6
            r["lhs_apply_dofs_per_s"] = n / t
            del src, dst
        res[label] = r
        del ctx
    return res


def newton_extra(torch):
    """Implicit Allen-Cahn (Newton + CG per increment) at the reference application's
    parameters (applications/allen_cahn/implicit/parameters.prm: 256^2 cells on
    [0,100]^2, dt = 0.01, RMSEPerField tolerances 1e-6, no preconditioner), uniform mesh."""
    from phasefield_amd import Context, prm

    n_inc = 20
    res = {"workload": "Allen-Cahn 2D implicit (Newton, CG), Q1, 256^2 cells on [0,100]^2, "
                       f"dt=0.01, {n_inc} increments after 5 warm-up increments"}
    for label, pc, depth in (("no_preconditioner", None, None), ("gmg", "GMG", 5)):
        text = prm.allen_cahn_implicit(refinement=8, steps=n_inc + 6, outputs=0,
                                       preconditioner=pc, mg_depth=depth)
        ctx = Context("allen_cahn_implicit_2d_q1", text)
        for _ in range(6):  # initial condition + 5 warm-up increments
            ctx.solve_increment()
        newton = cg = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n_inc):
            if ctx.solve_increment() != 0:
                raise RuntimeError("implicit Allen-Cahn increment stopped")
            st = ctx.last_newton_stats(1)
            newton += st.iterations
            cg += st.linear_iterations
        ctx.synchronize()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        res[label] = {
            "us_per_increment": dt / n_inc * 1e6,
            "newton_iterations_per_increment": newton / n_inc,
            "cg_iterations_per_increment": cg / n_inc,
            "us_per_cg_iteration": dt / max(cg, 1) * 1e6,
        }
        del ctx
    # Implicit Cahn-Hilliard at the reference application's parameters
    # (applications/cahn_hilliard/implicit/parameters.prm: 128^2 Q2 cells on [0,100]^2,
    # dt = 5e-4, BiCGStab with exact residual, tolerances 1e-6).
    ctx = Context("cahn_hilliard_implicit_2d_q2",
                  prm.cahn_hilliard_implicit(refinement=7, steps=n_inc + 6, outputs=0))
    for _ in range(6):
        ctx.solve_increment()
    newton = its = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n_inc):
        if ctx.solve_increment() != 0:
            raise RuntimeError("implicit Cahn-Hilliard increment stopped")
        st = ctx.last_newton_stats(0)
        newton += st.iterations
        its += st.linear_iterations
    ctx.synchronize()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    res["cahn_hilliard_implicit"] = {
        "workload": "Cahn-Hilliard 2D implicit (Newton, BiCGStab), Q2, 128^2 cells on [0,100]^2, "
                    f"fields c and mu, dt=5e-4, {n_inc} increments after 5 warm-up increments",
        "us_per_increment": dt / n_inc * 1e6,
        "newton_iterations_per_increment": newton / n_inc,
        "bicgstab_iterations_per_increment": its / n_inc,
        "us_per_bicgstab_iteration": dt / max(its, 1) * 1e6,
    }
    del ctx
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box in this setup has a single MI355X; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from phasefield_amd import Context, prm
    from phasefield_amd.capi import OP_RHS, SCALE_INVM

    total = args.warmup + args.steps
    text = prm.cahn_hilliard(steps=total + 1, outputs=0)
    ctx = Context("cahn_hilliard_3d_q2", text)
    info = ctx.info
    n_d = info.n_dofs

    # Increment 0 (initial condition of c, mu from it) and the warmup increments.
    if ctx.run(args.warmup) != 0:
        raise RuntimeError("warmup increments stopped early")
    torch.cuda.synchronize()
    ctx.synchronize()
    t0 = time.perf_counter()
    # Increments warmup+1 .. warmup+steps through ppf_run (Problem::solve's loop, NaN
    # probe after every increment).
    status = ctx.run(args.warmup + args.steps)
    if status != 0 or ctx.get_info().increment != args.warmup + args.steps + 1:
        raise RuntimeError(f"increment loop stopped with status {status}")
    ctx.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = 2 * n_d * args.steps / elapsed

    # Dominant kernel: the c-block operator (reads c_old and mu_old, writes c).
    dst = torch.zeros(info.n_nodes, dtype=torch.float64, device="cuda")
    t_kernel = time_kernel(ctx, torch, 0, OP_RHS, SCALE_INVM, None, dst, args.kernel_reps)
    # Same kernel with L2 and Infinity Cache flushed before every launch (the 51.5 MB
    # working set of this workload otherwise stays in the 256 MB Infinity Cache).
    flush = torch.zeros(128 << 20, dtype=torch.float64, device="cuda")  # 1 GiB
    t_flushed = time_kernel(ctx, torch, 0, OP_RHS, SCALE_INVM, None, dst, 20, flush=flush)
    del flush
    alg_bytes = 3 * 8 * n_d
    achieved = alg_bytes / t_kernel / 1e9
    # Reference point at the same size: an elementwise add of two arrays into a third
    # (3 arrays of the stored lattice, the c block's traffic without the stencil).
    ea = torch.randn(info.n_nodes, dtype=torch.float64, device="cuda")
    eb, ec = torch.randn_like(ea), torch.empty_like(ea)
    for _ in range(5):
        torch.add(ea, eb, out=ec)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.kernel_reps):
        torch.add(ea, eb, out=ec)
    e1.record()
    e1.synchronize()
    t_add = e0.elapsed_time(e1) / args.kernel_reps * 1e-3
    add_gbs = 3 * 8 * info.n_nodes / t_add / 1e9
    del ea, eb, ec
    roofline = {
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": profiled("stencil_cahn_hilliard_q2_64_fp64_pmc.md"),
        "kernel": "stencil_operator_kernel (Cahn-Hilliard c block, Q2, FP64)",
        "launch_seconds": t_kernel,
        "bytes_per_launch": alg_bytes,
        "achieved_flushed": alg_bytes / t_flushed / 1e9,
        "frac_flushed": alg_bytes / t_flushed / 1e9 / HBM_PEAK_GBS,
        "launch_seconds_flushed": t_flushed,
        "elementwise_add_gbs": add_gbs,
        "frac_of_elementwise_add": achieved / add_gbs,
    }
    flops = profiled("stencil_cahn_hilliard_q2_64_fp64_pmc.md", "fp64_flops_per_launch")
    if flops is not None:
        roofline["fp64_flops_per_launch"] = flops
        roofline["fp64_tflops"] = flops / t_kernel / 1e12
        roofline["fp64_frac_of_peak"] = flops / t_kernel / 1e12 / FP64_PEAK_TFLOPS
    del ctx, dst
    torch.cuda.empty_cache()

    if args.copy_calibration:
        src = torch.randn(16974593, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    extra = None
    if not args.no_extra:
        wanted = set(args.extras.split(","))
        extra = poisson_extra(torch) if "poisson" in wanted else {}
        torch.cuda.empty_cache()
        for name, fn in (("adjoint", adjoint_extra), ("newton", newton_extra),
                         ("precipitate", precipitate_extra), ("mechanics", mechanics_extra)):
            if name in wanted:
                extra[name] = fn(torch)
                torch.cuda.empty_cache()
    cpu = cpu_baseline(args.cpu_seconds) if args.cpu_seconds > 0 else None

    line = {
        "metric": "operator DoFs/s & HBM roofline %; mixed vs FP64 time-to-solution at 1e-8",
        "value": value,
        "unit": "DoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded initial condition c = 0.5 + U[-0.01, 0.01])",
        "config": {
            "workload": "Cahn-Hilliard 3D explicit, Q2, 64^3 cells on [0,64]^3, periodic, "
                        "fields c and mu, dt=1e-4",
            "dofs_per_field": n_d,
            "operator_applications_per_step": 2,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
