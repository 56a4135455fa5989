"""Benchmark of the GPU factor-solve path on one MI355X.

Workload (the default line): the PTDF block on the 10,000-bus synthetic grid. B' (buses 1..N-1,
bus 0 the slack) is factored by AMD-ordered Cholesky on the host, uploaded once, and its 13,500
pair right-hand sides (+1/-1 at each branch's end buses) are staged in HBM; one step solves the
whole block forward and backward in FP64 on the device, the right-hand sides built in-kernel.
The metric is RHS x nnz(L) per second (nnz(L) counting the diagonal).

A second object, "precision", reports the FP32-solve + FP64-refinement path against the plain
FP64 path on the 25,000-bus AC Newton Jacobian (COLAMD LU, 512 N(0,1) right-hand sides, refined
to a relative residual of 1e-12).

The CPU baseline is the oracle (the C++ restatement of ybus_sparse's par_solve_many: one
right-hand side per thread) timed on this host on the same factor. The oracle also supplies the
factor itself: in the reference path the factor comes from ybus_sparse on the host, and that is
what the oracle restates; the timed GPU region never calls it.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-precision]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RHS·nnz(L)/s & HBM GB/s, FP32+refine vs FP64 at 1e-12 resid, 1×MI355X"
HBM_PEAK_GBS = 8000.0
SEED = 20261203


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


# ---- CPU-baseline leg: the reference's host side (ybus_sparse), restated by the oracle ----



def cpu_threads() -> int:
    """The host threads the CPU baseline runs on: every core this process may use (its CPU
    affinity), capped by OMP_NUM_THREADS where the environment sets it (the share of the host's
    cores given to this GPU's jobs)."""
    return min(cpu_affinity(), thread_cap() or cpu_affinity())


def cpu_affinity() -> int:
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except (AttributeError, OSError):
        return max(1, os.cpu_count() or 1)


def thread_cap() -> int:
    try:
        return max(0, int(os.environ.get("OMP_NUM_THREADS", "0")))
    except ValueError:
        return 0


CPU_SECONDS = 10.0


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_pair_baseline(chol, plus, minus, sample: int, threads: int = 0, seconds: float = CPU_SECONDS):
    """par_solve_many on `sample` of the pair right-hand sides (dense vectors, as the reference's
    CPU path takes them), repeated for `seconds` on `threads` of this host's cores (all it may
    use by default); returns (RHS per second, threads, the block)."""
    n = chol.n
    threads = threads or cpu_threads()
    dense = np.zeros((sample, n))
    for k in range(sample):
        if plus[k] >= 0:
            dense[k, plus[k]] = 1.0
        if minus[k] >= 0:
            dense[k, minus[k]] = -1.0
    chol.solve_many(dense[: min(64, sample)], threads)  # warm-up block
    t0 = time.perf_counter()
    reps = 0
    while True:
        out = chol.solve_many(dense, threads)
        reps += 1
        if time.perf_counter() - t0 >= seconds:
            break
    elapsed = time.perf_counter() - t0
    return reps * sample / elapsed, threads, out


def cpu_dense_baseline(factor, rhs):
    threads = cpu_threads()
    factor.solve_many(rhs[: min(32, len(rhs))], threads)
    t0 = time.perf_counter()
    reps = 0
    while True:
        out = factor.solve_many(rhs, threads)
        reps += 1
        if time.perf_counter() - t0 >= 1.0:
            break
    return reps * rhs.shape[0] / (time.perf_counter() - t0), threads, out


def profiled_traffic(kernel: str):
    """HBM bytes per launch of `kernel` on the PTDF block, from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE passes, each scaled by what the counter reads for a plain
    1 GiB device copy); None when there is no summary for this kernel."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_ptdf_10k_hbm_traffic.md")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1]), os.path.relpath(path, ROOT)
    return None, None


# ---- GPU legs ----

def pair_block(grid):
    """One pair per branch: +1 at its from bus, -1 at its to bus, the slack (bus 0) dropped."""
    from oracle import configs

    return configs.pair_lists(grid)


def run_ptdf(gpu, args):
    import ybus_amd as Y

    from oracle import configs

    t0 = time.perf_counter()
    grid, bprime, chol, plus, minus = configs.ptdf()
    factor_s = time.perf_counter() - t0
    rhs_count = len(plus)
    nnz_l = int(chol.l.nnz)
    log(f"[ptdf] n={chol.n} nnz(L)={nnz_l} RHS={rhs_count}; host grid and factor {factor_s * 1e3:.1f} ms")

    factor = Y.GpuFactor.from_cholesky(gpu, chol)
    factor.set_block_width(rhs_count)
    pairs = [Y.Pair(None if p < 0 else int(p), None if m < 0 else int(m)) for p, m in zip(plus, minus)]
    tuning = factor.autotune(rhs_count, pairs)  # the library's autotuner: path, threads, tile width
    log(f"[ptdf] autotuned: {tuning}")
    # the kernel a pair block takes: the column path's, or on the wide path the one that sweeps
    # each tile's reach forward (unless YBUS_HIP_REACH=0 turns that off)
    if factor.path(rhs_count, pairs=True) == "columns":
        kernel = "solve_lds_kernel"
    else:
        kernel = "solve64_kernel" if os.environ.get("YBUS_HIP_REACH") == "0" else "solve64_reach_kernel"
    factor.stage_pairs(pairs)
    for _ in range(args.warmup):
        factor.solve_staged(Y.FP64).free()
    factor.reset_kernel_stats()
    # timed region: K device solves of the staged block (inputs already in HBM; each call
    # returns with its work finished, and each frees the previous call's block first)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        block = factor.solve_staged(Y.FP64)
    elapsed = time.perf_counter() - t0
    stats = factor.kernel_stats()
    log(f"[ptdf] {args.steps} steps in {elapsed * 1e3:.2f} ms; kernel {stats.milliseconds / stats.launches:.3f} ms "
        f"avg over {stats.launches} launches")
    solutions = block.solutions()
    block.free()
    # end to end, as the reference's callers use it: the pair list from the host in, every
    # solution back in host memory (1.08 GB), the median of five
    walls = []
    for _ in range(5):
        t1 = time.perf_counter()
        with factor.pair_solves(pairs) as b:
            b.solutions()
        walls.append(time.perf_counter() - t1)
    end_to_end = sorted(walls)[len(walls) // 2]
    log(f"[ptdf] end to end (pairs in, solutions out in host memory) {end_to_end * 1e3:.1f} ms")
    return dict(grid=grid, bprime=bprime, chol=chol, plus=plus, minus=minus, nnz_l=nnz_l, rhs=rhs_count, kernel=kernel,
                tuning=tuning, end_to_end=end_to_end,
                elapsed=elapsed, stats=stats, solutions=solutions, factor=factor)


def run_precision(gpu, args):
    import oracle as O
    import ybus_amd as Y

    from oracle import configs

    t0 = time.perf_counter()
    jac, lu, rhs = configs.jacobian()
    factor_s = time.perf_counter() - t0
    nnz_lu = int(lu.l.nnz + lu.u.nnz)
    log(f"[precision] n={jac.rows} nnz(L+U)={nnz_lu}; host Jacobian and COLAMD LU {factor_s * 1e3:.1f} ms")
    factor = Y.GpuFactor.from_lu(gpu, lu)
    factor.set_matrix(jac)
    factor.set_block_width(512)
    tuning = factor.autotune(512)
    log(f"[precision] autotuned: {tuning}")
    factor.stage(rhs)
    out = {}
    for name, mode in [("fp64", Y.FP64), ("fp32_refined", Y.FP32_REFINED)]:
        for _ in range(max(1, args.warmup)):
            factor.solve_staged(mode, 1e-12, 10).free()
        factor.reset_kernel_stats()
        steps = max(3, args.steps // 4)
        t0 = time.perf_counter()
        for _ in range(steps):
            block = factor.solve_staged(mode, 1e-12, 10)
            if _ + 1 < steps:
                block.free()
        elapsed = (time.perf_counter() - t0) / steps
        stats = factor.kernel_stats()
        x = block.solutions()
        corr, ratio = block.refinement()
        block.free()
        if mode == Y.FP64:
            ratios = np.array([jac.relative_residual(x[k], rhs[k]) for k in range(0, 512, 8)])
        else:
            ratios = ratio
        out[name] = dict(seconds=elapsed, value=512 * nnz_lu / elapsed,
                         kernel_ms=stats.milliseconds / stats.launches,
                         gbs=stats.bytes / stats.launches / (stats.milliseconds / stats.launches) / 1e6,
                         max_relative_residual=float(np.max(ratios)),
                         max_steps=int(corr.max()) if mode != Y.FP64 else 0,
                         mean_steps=float(corr.mean()) if mode != Y.FP64 else 0.0)
        log(f"[precision] {name}: {elapsed * 1e3:.3f} ms/solve, {out[name]['value']:.3e} RHS*nnz/s, "
            f"residual <= {out[name]['max_relative_residual']:.2e}")
        if mode == Y.FP64:
            x64 = x
        else:
            x32 = x
    # FP32-refined with the floor stop rule: the correction that cannot halve the ratio saved
    factor.set_refine_stop(Y.STOP_AT_FLOOR)
    for _ in range(max(1, args.warmup)):
        factor.solve_staged(Y.FP32_REFINED, 1e-12, 10).free()
    factor.reset_kernel_stats()
    steps = max(3, args.steps // 4)
    t0 = time.perf_counter()
    for i in range(steps):
        block = factor.solve_staged(Y.FP32_REFINED, 1e-12, 10)
        if i + 1 < steps:
            block.free()
    elapsed = (time.perf_counter() - t0) / steps
    fstats = factor.kernel_stats()
    x32f = block.solutions()
    corr_f, ratio_f = block.refinement()
    block.free()
    factor.set_refine_stop(Y.STOP_NORMWISE)
    out["fp32_refined_floor_stop"] = dict(seconds=elapsed, value=512 * nnz_lu / elapsed,
                                          kernel_ms=fstats.milliseconds / fstats.launches,
                                          max_relative_residual=float(np.max(ratio_f)),
                                          mean_steps=float(corr_f.mean()))
    log(f"[precision] fp32_refined, floor stop: {elapsed * 1e3:.3f} ms/solve, residual <= {np.max(ratio_f):.2e}")
    # the FP32 sweeps alone: no correction (max_steps = 0), so one FP32 solve and one residual
    for _ in range(max(1, args.warmup)):
        factor.solve_staged(Y.FP32_REFINED, 1e-12, 0).free()
    factor.reset_kernel_stats()
    for _ in range(3):
        factor.solve_staged(Y.FP32_REFINED, 1e-12, 0).free()
    st = factor.kernel_stats()
    out["fp32_solve_and_residual_kernel_ms"] = st.milliseconds / st.launches
    log(f"[precision] one FP32 solve + FP64 residual: {out['fp32_solve_and_residual_kernel_ms']:.3f} ms")
    # the contraction per correction: each column's relative residual after k corrections (at most
    # k; a column that has stopped keeps its best iterate), the median over the block
    contraction = []
    for k in range(0, 6):
        with factor.solve_staged(Y.FP32_REFINED, 1e-12, k) as block:
            contraction.append(float(np.median(block.refinement()[1])))
    out["median_residual_after_k_corrections"] = contraction
    # against the CPU's FP64 solutions (the oracle), every column
    cpu = lu.solve_many(rhs, cpu_threads())
    scale = np.max(np.abs(cpu), axis=1)
    out["fp64_vs_cpu_max_rel_diff"] = float(np.max(np.max(np.abs(x64 - cpu), axis=1) / scale))
    out["fp64_bit_exact_columns"] = int(sum(a.tobytes() == b.tobytes() for a, b in zip(x64, cpu)))
    out["fp32_refined_vs_cpu_fp64_max_rel_diff"] = float(np.max(np.max(np.abs(x32 - cpu), axis=1) / scale))
    out["fp32_refined_floor_stop"]["vs_cpu_fp64_max_rel_diff"] = float(
        np.max(np.max(np.abs(x32f - cpu), axis=1) / scale))
    out["speedup_fp32_refined_floor_stop_over_fp64"] = out["fp32_refined_floor_stop"]["value"] / out["fp64"]["value"]
    # conditioning: the oracle's 1-norm estimate (accuracy.rs:207-279) and the infinity-norm one
    # that bounds ||x32 - x64|| by ||A^-1||inf times the residuals
    inv1 = configs.inverse_norm_1_estimate(lu.n, lu.solve, lu.solve_transpose)
    invinf = configs.lu_inverse_norm_inf_estimate(lu)
    out["condition_estimate_1"] = configs.norm_1(jac) * inv1
    out["condition_estimate_inf"] = configs.norm_inf(jac) * invinf
    # what the 1.7x target would need: (k + 1) FP32 solves with their residuals in the FP64
    # solve's time / 1.7, k the mean corrections
    k = out["fp32_refined"]["mean_steps"]
    out["fp32_solve_ms_needed_for_1_7x"] = out["fp64"]["kernel_ms"] / 1.7 / (k + 1.0)
    out["speedup_fp32_refined_over_fp64"] = out["fp32_refined"]["value"] / out["fp64"]["value"]
    out["workload"] = "25,000-bus AC Newton Jacobian (n=%d), COLAMD LU, 512 N(0,1) RHS" % jac.rows
    out["tuning"] = tuning
    out["nnz_factor"] = nnz_lu
    return out


def copy_calibration(mib: int):
    """A plain device-to-device copy of `mib` MiB, five times: the known-size kernel that scales
    the HBM counters (FETCH_SIZE, WRITE_SIZE) read for the solve kernels."""
    import torch

    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"[copy] {mib} MiB copied 5 times")


def gather_calibration(mib: int, row_bytes: int = 128):
    """A gather of known size, five times: a table of `mib` MiB as rows of `row_bytes` bytes, every
    row read once in a random order (torch.index_select over a seeded permutation) and written
    in that order. It scales FETCH_SIZE for a kernel whose reads are scattered whole lines, as
    the wide sweeps' dependency gathers are (TW 16 doubles: one 128-byte line per row and tile).
    Reads: the table plus the 8-byte indices; writes: the table's bytes."""
    import torch

    width = row_bytes // 8
    rows = mib * (1 << 20) // row_bytes
    table = torch.ones((rows, width), dtype=torch.float64, device="cuda")
    order = torch.randperm(rows, generator=torch.Generator().manual_seed(SEED)).to("cuda")
    out = torch.empty_like(table)
    for _ in range(5):
        torch.index_select(table, 0, order, out=out)
    torch.cuda.synchronize()
    log(f"[gather] {rows} rows of {row_bytes} B gathered 5 times: reads {rows * row_bytes + rows * 8} B "
        f"(table and indices), writes {rows * row_bytes} B")


def run_adjoint(gpu, args):
    """The adjoint on the 70,000-bus grid: 64 zero-sum N(0,1) injection scenarios, branch limits
    U[0.5, 2.0] pu, loss = sum relu(|f| - limit)^2 with f = Bf theta, theta = B'^-1 p. The whole
    step runs in device memory: the solve through ybus_amd.autograd.solve, the flows and the loss
    as PyTorch tensor ops (plumbing), and the backward pass through the device adjoint
    (b_bar = B'^-T x_bar, A_bar = -(sum b_bar theta^T) on B''s pattern)."""
    import torch

    import oracle as O
    import ybus_amd as Y
    from ybus_amd.autograd import solve

    from oracle import configs

    case = configs.adjoint()
    bprime, bf, chol, p, limit = case.bprime, case.bf, case.chol, case.p, case.limit
    factor = Y.GpuFactor.from_cholesky(gpu, chol)
    factor.set_matrix(bprime)
    factor.set_block_width(64)
    tuning = factor.autotune(64)  # the 64-scenario block's kernel configuration
    log(f"[adjoint] autotuned: {tuning}")
    # Bf has at most two entries per row: f = v0 * theta[c0] + v1 * theta[c1]
    counts = np.diff(bf.ptr)
    c0 = np.where(counts > 0, bf.idx[np.minimum(bf.ptr[:-1], len(bf.idx) - 1)], 0)
    v0 = np.where(counts > 0, bf.data[np.minimum(bf.ptr[:-1], len(bf.idx) - 1)], 0.0)
    c1 = np.where(counts > 1, bf.idx[np.minimum(bf.ptr[:-1] + 1, len(bf.idx) - 1)], 0)
    v1 = np.where(counts > 1, bf.data[np.minimum(bf.ptr[:-1] + 1, len(bf.idx) - 1)], 0.0)
    dev = torch.device("cuda")
    t_c0, t_c1 = torch.tensor(c0, device=dev), torch.tensor(c1, device=dev)
    t_v0, t_v1 = torch.tensor(v0, device=dev), torch.tensor(v1, device=dev)
    t_lim = torch.tensor(limit, device=dev)
    t_p = torch.tensor(p, device=dev)
    a_values = torch.tensor(bprime.data, device=dev)
    captured = {}

    def step(mode=0):
        a = a_values.detach().requires_grad_(True)
        b = t_p.detach().requires_grad_(True)
        theta = solve(factor, a, b, mode, 1e-12, 10)
        theta.register_hook(lambda g: captured.__setitem__("x_bar", g))
        f = t_v0 * theta[:, t_c0] + t_v1 * theta[:, t_c1]
        over = torch.relu(f.abs() - t_lim)
        loss = (over * over).sum()
        loss.backward()
        return loss, theta, a.grad, b.grad

    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    factor.reset_kernel_stats()
    reps = max(3, args.steps // 4)
    t0 = time.perf_counter()
    for _ in range(reps):
        loss, theta, a_grad, b_grad = step()
    torch.cuda.synchronize()
    seconds = (time.perf_counter() - t0) / reps
    stats = factor.kernel_stats()
    x_bar = captured["x_bar"].cpu().numpy()
    theta_h = theta.detach().cpu().numpy()
    want_b, want_a = O.adjoint(chol, bprime, x_bar, theta_h)
    # the same step with FP32 sweeps refined in FP64, forward and backward
    from ybus_amd.autograd import FP32_REFINED

    step(FP32_REFINED)
    torch.cuda.synchronize()
    factor.reset_kernel_stats()
    t0 = time.perf_counter()
    for _ in range(reps):
        _, theta32, a_grad32, b_grad32 = step(FP32_REFINED)
    torch.cuda.synchronize()
    seconds32 = (time.perf_counter() - t0) / reps
    stats32 = factor.kernel_stats()
    # and with the floor stop rule (the correction that cannot halve the ratio saved)
    factor.set_refine_stop(Y.STOP_AT_FLOOR)
    step(FP32_REFINED)
    torch.cuda.synchronize()
    factor.reset_kernel_stats()
    t0 = time.perf_counter()
    for _ in range(reps):
        _, theta32f, a_grad32f, b_grad32f = step(FP32_REFINED)
    torch.cuda.synchronize()
    seconds32f = (time.perf_counter() - t0) / reps
    stats32f = factor.kernel_stats()
    factor.set_refine_stop(Y.STOP_NORMWISE)

    def rel(got, want):
        return float((torch.max(torch.abs(got - want)) / torch.max(torch.abs(want))).detach())
    out = dict(workload="70,000-bus grid (n=%d), B' AMD Cholesky (nnz(L)=%d, nnz(B')=%d), 64 scenarios, "
                        "relu^2 flow-limit loss" % (chol.n, int(chol.l.nnz), bprime.nnz),
               loss=float(loss.detach()), ms_per_step=seconds * 1e3,
               library_kernel_ms_per_step=stats.milliseconds / reps,
               library_kernel_launches_per_step=stats.launches // reps, tuning=tuning,
               theta_bit_exact=bool(theta_h.tobytes() == chol.solve_many(p).tobytes()),
               b_bar_bit_exact=bool(b_grad.cpu().numpy().tobytes() == want_b.tobytes()),
               a_bar_bit_exact=bool(a_grad.cpu().numpy().tobytes() == want_a.tobytes()),
               fp32_refined=dict(ms_per_step=seconds32 * 1e3, library_kernel_ms_per_step=stats32.milliseconds / reps,
                                 theta_rel_err_vs_fp64=rel(theta32, theta), b_bar_rel_err_vs_fp64=rel(b_grad32, b_grad),
                                 a_bar_rel_err_vs_fp64=rel(a_grad32, a_grad)),
               fp32_refined_floor_stop=dict(ms_per_step=seconds32f * 1e3,
                                            library_kernel_ms_per_step=stats32f.milliseconds / reps,
                                            theta_rel_err_vs_fp64=rel(theta32f, theta),
                                            b_bar_rel_err_vs_fp64=rel(b_grad32f, b_grad),
                                            a_bar_rel_err_vs_fp64=rel(a_grad32f, a_grad)),
               note="one step = forward solve, flows and loss, backward (the library's transposed solve and "
                    "SDDMM), all in device memory; library kernels = forward solve + transposed solve + SDDMM")
    log(f"[adjoint] {out['ms_per_step']:.2f} ms per device step (library kernels "
        f"{out['library_kernel_ms_per_step']:.3f} ms in {out['library_kernel_launches_per_step']} launches); "
        f"bit-exact theta {out['theta_bit_exact']}, b_bar {out['b_bar_bit_exact']}, a_bar {out['a_bar_bit_exact']}; "
        f"FP32-refined step {out['fp32_refined']['ms_per_step']:.2f} ms, gradient error vs FP64 "
        f"{max(out['fp32_refined']['b_bar_rel_err_vs_fp64'], out['fp32_refined']['a_bar_rel_err_vs_fp64']):.2e}")
    return out


def run_batch(gpu, args):
    """256 synthetic grids with log-uniform sizes in [500, 20,000] buses, each with one random
    non-bridge branch out of service, factored by AMD Cholesky on the host; grid i gets r_i ~
    U{1..64} zero-sum N(0,1) injection scenarios. One step solves every grid's block: once as a
    batch (one persistent launch per LDS class, replayed from a HIP graph) and once factor by
    factor (one staged launch each), for comparison."""
    import oracle as O
    import ybus_amd as Y

    from oracle import configs

    t0 = time.perf_counter()
    case = configs.batch()
    grids, counts, cpu, rhs, mats = len(case.factors), case.counts, case.factors, case.rhs, case.matrices
    log(f"[batch] {grids} grids, sum n = {int(sum(c.n for c in cpu))}, sum r = {int(counts.sum())}; host "
        f"{time.perf_counter() - t0:.1f} s")
    factors = [Y.GpuFactor.from_cholesky(gpu, c) for c in cpu]
    for f, m in zip(factors, mats):
        f.set_matrix(m)  # for the FP32-refined batch's residuals
    batch = Y.GpuBatch.new(gpu)
    for f in factors:
        batch.add(f)
    batch.stage(rhs)
    for _ in range(max(1, args.warmup)):
        batch.solve_staged()
    batch.reset_kernel_stats()
    reps = max(3, args.steps // 2)
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.solve_staged()
    batch_s = (time.perf_counter() - t0) / reps
    stats = batch.kernel_stats()
    solved = batch.solutions()
    # the same batch as FP32 sweeps refined in FP64 (to 1e-12 or the FP64 floor), one graph replay
    batch.solve_staged(Y.FP32_REFINED, 1e-12, 10)
    batch.reset_kernel_stats()
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.solve_staged(Y.FP32_REFINED, 1e-12, 10)
    refined_s = (time.perf_counter() - t0) / reps
    refined_stats = batch.kernel_stats()
    refined = batch.solutions()
    steps_ratio = batch.refinement()
    max_ratio = max(float(r.max()) for _, r in steps_ratio if len(r))
    mean_steps = float(np.mean(np.concatenate([s for s, _ in steps_ratio])))
    refined_err = max(float(np.max(np.abs(x - y)) / np.max(np.abs(y))) for x, y in zip(refined, solved) if y.size)
    # and with the floor stop rule
    batch.set_refine_stop(Y.STOP_AT_FLOOR)
    batch.solve_staged(Y.FP32_REFINED, 1e-12, 10)
    batch.reset_kernel_stats()
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.solve_staged(Y.FP32_REFINED, 1e-12, 10)
    floor_s = (time.perf_counter() - t0) / reps
    floor_stats = batch.kernel_stats()
    floor_x = batch.solutions()
    floor_ratio = max(float(r.max()) for _, r in batch.refinement() if len(r))
    batch.set_refine_stop(Y.STOP_NORMWISE)
    # one launch per factor, the same blocks
    for f, b in zip(factors, rhs):
        f.stage(b)
        f.solve_staged(Y.FP64).free()
    for f in factors:
        f.reset_kernel_stats()
    t0 = time.perf_counter()
    for _ in range(reps):
        for f in factors:
            f.solve_staged(Y.FP64).free()
    single_s = (time.perf_counter() - t0) / reps
    single_kernel_ms = sum(f.kernel_stats().milliseconds for f in factors) / reps
    work = float(sum(int(r) * int(c.l.nnz) for r, c in zip(counts, cpu)))
    # every factor's block against the CPU's solves (the oracle), and the refined ones' distance
    cpu_x = [c.solve_many(b, cpu_threads()) for c, b in zip(cpu, rhs)]
    exact = sum(x.tobytes() == w.tobytes() for x, w in zip(solved, cpu_x))
    refined_vs_cpu = max(float(np.max(np.abs(x - w)) / np.max(np.abs(w))) for x, w in zip(refined, cpu_x) if w.size)
    floor_vs_cpu = max(float(np.max(np.abs(x - w)) / np.max(np.abs(w))) for x, w in zip(floor_x, cpu_x) if w.size)
    out = dict(workload="256 grids, n_i log-uniform in [499, 19,999] (sum n = %d), r_i in 1..64 (sum r = %d), "
                        "AMD Cholesky, one non-bridge branch out per grid" % (int(sum(c.n for c in cpu)),
                                                                              int(counts.sum())),
               value=work / batch_s, unit="RHS·nnz(L)/s", batch_ms=batch_s * 1e3,
               batch_kernel_ms=stats.milliseconds / stats.launches, one_launch_per_factor_ms=single_s * 1e3,
               one_launch_per_factor_kernel_ms=single_kernel_ms,
               speedup_over_one_launch_per_factor=single_s / batch_s,
               fp32_refined=dict(batch_ms=refined_s * 1e3, batch_kernel_ms=refined_stats.milliseconds / reps,
                                 max_relative_residual=max_ratio, mean_corrections=mean_steps,
                                 max_rel_diff_vs_fp64=refined_err, max_rel_diff_vs_cpu_fp64=refined_vs_cpu),
               fp32_refined_floor_stop=dict(batch_ms=floor_s * 1e3, batch_kernel_ms=floor_stats.milliseconds / reps,
                                            max_relative_residual=floor_ratio, max_rel_diff_vs_cpu_fp64=floor_vs_cpu),
               bit_exact_factors=f"{exact} of {grids}")
    log(f"[batch] batch {out['batch_ms']:.3f} ms (kernel {out['batch_kernel_ms']:.3f}), one launch per factor "
        f"{out['one_launch_per_factor_ms']:.3f} ms; bit-exact {out['bit_exact_factors']}; FP32-refined "
        f"{out['fp32_refined']['batch_ms']:.3f} ms, residual <= {max_ratio:.2e}, {mean_steps:.2f} corrections")
    return out


def host_screening(chol, bf, theta, plus, minus, inverse_weight, weights, rows, threads):
    """The CPU side of N-1 screening for the outages `rows`: each outage's pair solve
    (par_solve_many over dense pair vectors), its rank-one update of theta (pairs.rs:209-253 as
    gpu_pairs_test.rs:48-62 states it, alpha = 0), the flows Bf theta' with the outaged branch's
    zeroed, and the most loaded branch: (rows of the maxima, their loadings)."""
    n = chol.n
    dense = np.zeros((len(rows), n))
    for j, k in enumerate(rows):
        if plus[k] >= 0:
            dense[j, plus[k]] = 1.0
        if minus[k] >= 0:
            dense[j, minus[k]] = -1.0
    y = chol.solve_many(dense, threads)
    p, m = plus[rows], minus[rows]

    def at(v, idx):
        return np.where(idx >= 0, v[np.arange(len(idx)), np.maximum(idx, 0)], 0.0)

    x = theta[None, :] + 0.0 * y
    w = -(at(y, p) - at(y, m)) + inverse_weight[rows]
    z = (at(x, p) - at(x, m)) / w
    updated = x + z[:, None] * y
    flows = (bf_dense_product(bf, updated))
    flows[np.arange(len(rows)), rows] = 0.0
    loading = np.abs(flows) * weights[None, :]
    return np.argmax(loading, axis=1), np.max(loading, axis=1)


def bf_dense_product(bf, x):
    """Bf x for each row of x (Bf in CSR, at most two entries a row)."""
    out = np.zeros((x.shape[0], bf.rows))
    ptr, idx, val = np.asarray(bf.ptr), np.asarray(bf.idx), np.asarray(bf.data)
    for slot in range(int(np.max(np.diff(ptr)))):
        has = np.diff(ptr) > slot
        at = ptr[:-1][has] + slot
        out[:, has] += val[at][None, :] * x[:, idx[at]]
    return out


def run_screening(gpu, args):
    """N-1 screening on the 25,000-bus grid (the reference's outage_loadings, dc/gpu.rs:334-430):
    base angles theta = B'^-1 p for one zero-sum N(0,1) injection vector, then for every
    non-bridge branch k taken out alone, theta_k = theta plus the rank-one update of B' losing
    b_k (pair (from, to), alpha = 0, inverse weight 1/b_k), flows Bf theta_k with branch k's
    zeroed, and the most loaded branch under limits U[0.5, 2.0] pu. One step screens every
    outage: pair_updates, product_maxima and maxima on the device, one block."""
    import oracle as O
    import ybus_amd as Y

    buses, seed = 25000, SEED + 5
    t0 = time.perf_counter()
    grid = O.make_grid(buses, seed)
    bprime, bf, _ = grid.dc()
    chol = O.Cholesky(bprime)
    p = O.injections(seed, 4, buses, 1)[0]
    theta = chol.solve(p)
    limit = 0.5 + 1.5 * O.uniforms(seed, 5, bf.rows)
    weights = 1.0 / limit
    plus, minus = pair_block(grid)
    b = 1.0 / np.asarray(grid.x, dtype=np.float64)
    inverse_weight = 1.0 / b
    outages = np.flatnonzero(~grid.bridges())
    o_plus, o_minus = plus[outages], minus[outages]
    o_alpha, o_weight = np.zeros(len(outages)), inverse_weight[outages]
    zero = outages.astype(np.int64)
    log(f"[screening] n = {chol.n}, {len(outages)} outages of {grid.branches} branches; host "
        f"{time.perf_counter() - t0:.1f} s")
    factor = Y.GpuFactor.from_cholesky(gpu, chol)
    factor.set_block_width(len(outages))
    tuning = factor.autotune(len(outages), [Y.Pair(None if a < 0 else int(a), None if b < 0 else int(b))
                                            for a, b in zip(o_plus, o_minus)])
    log(f"[screening] autotuned: {tuning}")
    op = factor.operator(bf)
    op.set_offset(np.zeros(bf.rows))
    op.set_weights(weights)

    def step():
        block = factor.pair_updates_arrays(theta, o_plus, o_minus, o_alpha, o_weight)
        loadings = block.product_maxima_arrays(op, zero)
        angles = block.maxima_arrays()
        singular = block.singular()
        block.free()
        return loadings, angles, singular

    for _ in range(max(1, args.warmup)):
        step()
    reps = max(3, args.steps // 4)
    t0 = time.perf_counter()
    for _ in range(reps):
        loadings, angles, singular = step()
    step_s = (time.perf_counter() - t0) / reps
    # parity: sampled outages against the CPU composition (loadings within 1e-12; the same branch
    # unless two loadings tie to within rounding)
    sample = outages[:: max(1, len(outages) // 64)]
    positions = {int(k): j for j, k in enumerate(outages)}
    rows, values = host_screening(chol, bf, theta, plus, minus, inverse_weight, weights, sample, cpu_threads())
    worst, same_branch = 0.0, 0
    for k, r, v in zip(sample, rows, values):
        j = positions[int(k)]
        worst = max(worst, abs(loadings[1][j] - v) / max(v, 1e-300))
        same_branch += int(loadings[0][j] == int(r))
    cpu_baseline = None
    if not args.no_cpu:
        cpu_sample = outages[:2000]
        threads = cpu_threads()
        host_screening(chol, bf, theta, plus, minus, inverse_weight, weights, cpu_sample[:64], threads)
        t0, done = time.perf_counter(), 0
        while time.perf_counter() - t0 < CPU_SECONDS:
            host_screening(chol, bf, theta, plus, minus, inverse_weight, weights, cpu_sample, threads)
            done += len(cpu_sample)
        cpu_rate = done / (time.perf_counter() - t0)
        cpu_baseline = dict(value=cpu_rate, unit="outages/s", cores=threads, kind="port",
                            sample="the oracle's par_solve_many over the first 2,000 outages' pair vectors, then "
                                   "the update, flows and maxima in numpy, repeated for >= %g s" % CPU_SECONDS)
    n2 = run_double_outages(factor, op, chol, bf, theta, plus, minus, inverse_weight, weights, outages, args)
    # the caller's contract over every branch, bridges included (dc/gpu.rs:386-430): an outage
    # whose update is singular or whose angles or loadings fail the suspect rule goes to the
    # caller's own solve; here that solve only counts them, untimed
    from ybus_amd import screening

    suspects = []
    t0 = time.perf_counter()
    screening.outage_loadings(factor, op, theta, plus, minus, np.zeros(grid.branches), inverse_weight,
                              np.arange(grid.branches), suspects.append)
    every_s = time.perf_counter() - t0
    bridge_rows = set(np.flatnonzero(grid.bridges()).tolist())
    contract = dict(branches=int(grid.branches), suspect_outages=len(suspects),
                    suspects_are_the_bridges=set(suspects) == bridge_rows, screen_ms=every_s * 1e3)
    log(f"[screening] every branch: {len(suspects)} suspect outages of {grid.branches} "
        f"(the bridges: {contract['suspects_are_the_bridges']}), {every_s * 1e3:.1f} ms")
    out = dict(workload="N-1 screening: 25,000-bus grid, B' AMD Cholesky, every non-bridge branch out alone "
                        "(%d outages), rank-one update, Bf flows, most loaded branch" % len(outages),
               n=int(chol.n), nnz_l=int(chol.l.nnz), outages=int(len(outages)), tuning=tuning,
               value=len(outages) / step_s,
               unit="outages/s", step_ms=step_s * 1e3, singular=int(sum(singular)),
               parity=dict(sampled_outages=int(len(sample)), max_rel_diff_loading=worst,
                           same_branch=f"{same_branch} of {len(sample)}"),
               cpu_baseline=cpu_baseline, double_outages=n2, suspect_contract=contract)
    log(f"[screening] {step_s * 1e3:.2f} ms per step ({out['value']:.3g} outages/s); loadings within {worst:.1e}, "
        f"same branch {same_branch} of {len(sample)}; cpu {cpu_baseline['value'] if cpu_baseline else 0:.3g} "
        f"outages/s")
    return out


def run_flows(gpu, args, ptdf):
    """Whole PTDF results (the reference's DcGpuSensitivity::ptdf_transfers, dc/gpu.rs:531-611):
    the 13,500 branch transfers of the 10,000-bus grid solved as one pair block, turned into every
    branch's flow by the Bf operator on the device (transposed there into one vector per
    transfer), and copied to the host: 13,500 x 13,500 flows."""
    import oracle as O
    import ybus_amd as Y

    factor, chol = ptdf["factor"], ptdf["chol"]
    _, bf, _ = ptdf["grid"].dc()
    plus, minus = ptdf["plus"], ptdf["minus"]
    pairs = [Y.Pair(None if p < 0 else int(p), None if m < 0 else int(m)) for p, m in zip(plus, minus)]
    op = factor.operator(bf)
    op.set_offset(np.zeros(bf.rows))
    zero = [None] * len(pairs)
    block = factor.pair_solves(pairs)
    flows = block.product(op, zero)
    block.free()
    reps = 3
    solve_s = product_s = 0.0
    for _ in range(reps):
        t0 = time.perf_counter()
        block = factor.pair_solves(pairs)
        t1 = time.perf_counter()
        flows = block.product(op, zero)
        t2 = time.perf_counter()
        block.free()
        solve_s += t1 - t0
        product_s += t2 - t1
    solve_s /= reps
    product_s /= reps
    # parity: sampled transfers' flows against Bf times the CPU's solutions (bit for bit: both sum
    # each branch's two terms from 0 in index order)
    sample = np.arange(0, len(pairs), max(1, len(pairs) // 32))
    exact = 0
    for k in sample:
        b = np.zeros(chol.n)
        if plus[k] >= 0:
            b[plus[k]] = 1.0
        if minus[k] >= 0:
            b[minus[k]] = -1.0
        want = bf.mul_vec(chol.solve(b)) + 0.0
        exact += int(flows[k].tobytes() == want.tobytes())
    entries = len(pairs) * bf.rows
    out = dict(workload="PTDF transfers' flows: 10,000-bus grid, 13,500 branch transfers, Bf (13,500 branches) "
                        "applied on the device, every flow copied to the host",
               transfers=len(pairs), branches=int(bf.rows), solve_ms=solve_s * 1e3, product_and_copy_ms=product_s * 1e3,
               flows_per_second=entries / (solve_s + product_s), bytes_to_host=entries * 8,
               bit_exact_sampled_transfers=f"{exact} of {len(sample)}")
    log(f"[flows] pair solves {solve_s * 1e3:.2f} ms, product + copy of {entries * 8 / 1e9:.2f} GB "
        f"{product_s * 1e3:.1f} ms; bit-exact {exact} of {len(sample)}")
    return out


def run_double_outages(factor, op, chol, bf, theta, plus, minus, inverse_weight, weights, outages, args):
    """Pairs of branches out together on the screening grid (dc/outage.rs:126-241 with two
    branches): 8,192 random pairs of the non-bridge branches, each a rank-2 update, its flows and
    most loaded branch; the block holds the 16,384 branch solves. Checked against the CPU's
    composition on a sample (to rounding: the 2 x 2 systems are solved on the host)."""
    rng = np.random.default_rng(SEED)
    count = 8192
    first = rng.choice(outages, size=count)
    second = rng.choice(outages, size=count)
    second = np.where(second == first, outages[(np.searchsorted(outages, first) + 1) % len(outages)], second)
    ks = np.stack([first, second], axis=1).reshape(-1)
    set_ptr = np.arange(0, 2 * count + 1, 2, dtype=np.uint64)
    zero = first.astype(np.int64)  # the first branch's flow is zeroed (its row of the product)
    o_plus, o_minus = plus[ks], minus[ks]
    o_alpha, o_weight = np.zeros(len(ks)), inverse_weight[ks]

    def step():
        block = factor.outage_updates_arrays(theta, set_ptr, o_plus, o_minus, o_alpha, o_weight)
        loads = block.product_maxima_arrays(op, zero)
        singular = block.singular()
        block.free()
        return loads, singular

    step()
    reps = max(3, args.steps // 4)
    t0 = time.perf_counter()
    for _ in range(reps):
        loads, singular = step()
    step_s = (time.perf_counter() - t0) / reps
    # the CPU composition for a sample of sets
    worst = 0.0
    for s in range(0, count, count // 16):
        k1, k2 = int(first[s]), int(second[s])
        ys = []
        for k in (k1, k2):
            b = np.zeros(chol.n)
            if plus[k] >= 0:
                b[plus[k]] = 1.0
            if minus[k] >= 0:
                b[minus[k]] = -1.0
            ys.append(chol.solve(b))

        def cdot(k, v):
            return (v[plus[k]] if plus[k] >= 0 else 0.0) - (v[minus[k]] if minus[k] >= 0 else 0.0)
        w = np.array([[-cdot(ki, yj) + (inverse_weight[ki] if i == j else 0.0) for j, yj in enumerate(ys)]
                      for i, ki in enumerate((k1, k2))])
        if singular[s] or abs(np.linalg.det(w)) < 1e-300:
            continue
        z = np.linalg.solve(w, np.array([cdot(k1, theta), cdot(k2, theta)]))
        updated = theta + z[0] * ys[0] + z[1] * ys[1]
        flows = bf_dense_product(bf, updated[None, :])[0]
        flows[k1] = 0.0
        v = float(np.max(np.abs(flows) * weights))
        worst = max(worst, abs(loads[1][s] - v) / max(v, 1e-300))
    out = dict(workload="N-2: 8,192 random pairs of the 25,000-bus grid's non-bridge branches, each out together "
                        "(rank-2 update), flows and most loaded branch",
               sets=count, value=count / step_s, unit="outage pairs/s", step_ms=step_s * 1e3,
               singular=int(sum(singular)), max_rel_diff_loading_vs_cpu=worst)
    log(f"[screening] N-2: {step_s * 1e3:.2f} ms for {count} pairs ({out['value']:.3g} pairs/s), "
        f"{out['singular']} singular, loadings within {worst:.1e} of the CPU")
    return out


def run_complex(gpu, args):
    """A complex factor: the 10,000-bus grid's B' made Hermitian (each branch entry given a phase,
    so A stays positive definite), factored as P A P^T = L L^H by the oracle's restated complex
    Cholesky, and 1,024 complex right-hand sides solved per step in FP64 (host vectors in and
    out, as solve_many takes them); the kernel time comes from HIP events around each launch."""
    import oracle as O
    from oracle import complex_solve as Z
    import ybus_amd as Y

    t0 = time.perf_counter()
    grid = O.make_grid(10000, SEED)
    bprime, _, _ = grid.dc()
    re, im = Z.hermitian_from(bprime, 1)
    chol = Z.ComplexCholesky(bprime, re, im, O.Cholesky(bprime))
    count = 1024
    b = O.normals(SEED, 9, count * chol.n).reshape(count, chol.n) + 1j * O.normals(SEED, 10, count * chol.n).reshape(
        count, chol.n)
    log(f"[complex] n = {chol.n}, nnz(L) = {len(chol.l_idx)}; host {time.perf_counter() - t0:.1f} s")
    factor = Y.GpuFactor.from_cholesky_complex(gpu, chol)
    factor.set_block_width(count)
    for _ in range(max(1, args.warmup)):
        x = factor.solve_many(b)
    factor.reset_kernel_stats()
    reps = max(3, args.steps // 2)
    t0 = time.perf_counter()
    for _ in range(reps):
        x = factor.solve_many(b)
    step_s = (time.perf_counter() - t0) / reps
    stats = factor.kernel_stats()
    kernel_ms = stats.milliseconds / stats.launches
    achieved = stats.bytes / stats.launches / (kernel_ms / 1e3) / 1e9
    sample = np.arange(0, count, 128)
    exact = x[sample].tobytes() == chol.solve_many(b[sample]).tobytes()
    work = float(count) * len(chol.l_idx)
    out = dict(workload="10,000-bus grid's B' made Hermitian, complex AMD Cholesky (L L^H), 1,024 complex RHS, "
                        "FP64 complex forward+backward", n=int(chol.n), nnz_l=int(len(chol.l_idx)), rhs=count,
               value=work / (kernel_ms / 1e3), unit="RHS·nnz(L)/s (kernel time)",
               with_host_copies=work / step_s, step_ms=step_s * 1e3, kernel="solve_complex_kernel",
               kernel_ms=kernel_ms, bytes_per_launch=stats.bytes / stats.launches, achieved_gbs=achieved,
               hbm_frac=achieved / HBM_PEAK_GBS,
               bit_exact_sampled_columns=f"{len(sample)} of {len(sample)}" if exact else "MISMATCH")
    # the same block as FP32 sweeps refined in FP64 against A (the complex matrix attached), and
    # its distance from the CPU's FP64 solutions (the oracle), on the sampled columns
    import scipy.sparse as sp

    factor.set_matrix(sp.csr_matrix((re + 1j * im, np.asarray(bprime.idx), np.asarray(bprime.ptr)),
                                    shape=bprime.shape))
    factor.solve_many_refined(b, 1e-12, 10)
    factor.reset_kernel_stats()
    t0 = time.perf_counter()
    for _ in range(reps):
        xr, steps, ratio = factor.solve_many_refined(b, 1e-12, 10)
    refined_s = (time.perf_counter() - t0) / reps
    rstats = factor.kernel_stats()
    cpu = chol.solve_many(b[sample])
    scale = np.max(np.abs(cpu), axis=1)
    out["fp32_refined"] = dict(kernel_ms=rstats.milliseconds / rstats.launches, step_ms=refined_s * 1e3,
                               max_relative_residual=float(ratio.max()), mean_corrections=float(steps.mean()),
                               max_rel_diff_vs_cpu_fp64=float(np.max(np.max(np.abs(xr[sample] - cpu), axis=1) / scale)),
                               kernel="refine_complex_kernel")
    log(f"[complex] kernel {kernel_ms:.3f} ms ({achieved:.0f} GB/s), step with copies {step_s * 1e3:.3f} ms; "
        f"bit-exact {out['bit_exact_sampled_columns']}; FP32-refined kernel "
        f"{out['fp32_refined']['kernel_ms']:.3f} ms, residual <= {ratio.max():.2e}, {steps.mean():.2f} corrections")
    return out


def run_refactor(gpu, args):
    """A Newton iteration's refactorization: the 25,000-bus Jacobian (n = 44,998, COLAMD LU)
    factored once on the host and uploaded, then refactored on the device with the factor's own
    pivots for the Jacobian at a nearby operating point (every |V| and angle moved by up to 1e-3,
    as late Newton iterates are), alternating with the first Jacobian so every call makes new
    values. Each call uploads the matrix values (host to device), runs the refactor kernel and
    rewrites the solve sweeps' values. Beside it: the CPU refactor (the oracle's restatement of
    Lu::refactor, one thread) and what a caller without a device refactor does each iteration
    (a new upload of the whole factor)."""
    import oracle as O
    import ybus_amd as Y

    seed = SEED + 2
    grid = O.make_grid(25000, seed)
    a = grid.jacobian(seed)[0]
    b = grid.jacobian(seed, 0.001, 7)[0]
    lu = O.Lu(a, O.COLAMD)
    t0 = time.perf_counter()
    factor = Y.GpuFactor.from_lu(gpu, lu)
    upload_s = time.perf_counter() - t0
    work = lu.refactor_work()
    for k in range(max(2, args.warmup)):
        assert factor.refactor(b if k % 2 == 0 else a) == "reused"
    before = factor.refactor_stats()
    reps = max(4, args.steps // 2) // 2 * 2
    t0 = time.perf_counter()
    for k in range(reps):
        outcome = factor.refactor(b if k % 2 == 0 else a)
        assert outcome == "reused", outcome
    call_s = (time.perf_counter() - t0) / reps
    after = factor.refactor_stats()
    calls = after["calls"] - before["calls"]
    kernel_ms = (after["refactor_ms"] - before["refactor_ms"]) / calls
    refresh_ms = (after["refresh_ms"] - before["refresh_ms"]) / calls
    # parity: one more refactor with b on both sides, every value
    assert factor.refactor(b) == "reused"
    assert lu.refactor(b) == "reused"
    l_val, u_val = factor.lu_values()
    exact = l_val.tobytes() == lu.l.data.tobytes() and u_val.tobytes() == lu.u.data.tobytes()
    rhs = O.normals(seed, 11, lu.n)
    x = factor.solve_many(rhs[None, :])
    solve_exact = x.tobytes() == lu.solve_many(rhs[None, :]).tobytes()
    # a value update from a host refactor (the same values), and the algorithmic bytes of a refactor:
    # each multiply-subtract reads an L value and a target slot (12 B); each matrix entry its value
    # and slot (12 B); each new L and U value is written once (8 B)
    t0 = time.perf_counter()
    factor.update_lu_values(lu.l.data, lu.u.data)
    update_s = time.perf_counter() - t0
    nnz_lu = int(lu.l.nnz + lu.u.nnz)
    bytes_per = 12 * work + 12 * a.nnz + 8 * nnz_lu
    achieved = bytes_per / (kernel_ms / 1e3) / 1e9
    out = dict(workload="Newton refactor: 25,000-bus AC Jacobian (COLAMD LU, pivots reused), the operating point "
                        "moved by up to 1e-3", n=int(lu.n), nnz_a=int(a.nnz), nnz_lu=nnz_lu,
               multiply_subtracts=int(work), kernel="refactor_kernel", kernel_ms=kernel_ms, refresh_ms=refresh_ms,
               call_ms=call_s * 1e3, value=work / (kernel_ms / 1e3), unit="multiply-subtracts/s (kernel time)",
               bytes_per_launch=bytes_per, achieved_gbs=achieved, hbm_frac=achieved / HBM_PEAK_GBS,
               upload_factor_ms=upload_s * 1e3, update_values_ms=update_s * 1e3,
               bit_exact_values="all" if exact else "MISMATCH", bit_exact_solve=bool(solve_exact))
    if not args.no_cpu:
        cpu = O.Lu(a, O.COLAMD)
        cpu.refactor(b, read=False)
        count, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 2.0 or count < 4:
            cpu.refactor(b if count % 2 == 0 else a, read=False)
            count += 1
        cpu_ms = (time.perf_counter() - t0) / count * 1e3
        out["cpu_baseline"] = {"value": work / (cpu_ms / 1e3), "unit": "multiply-subtracts/s", "cores": 1,
                               "kind": "port", "ms": cpu_ms,
                               "sample": f"the oracle's Lu::refactor on the same two Jacobians, {count} calls"}
    log(f"[refactor] n={lu.n} work={work}: kernel {kernel_ms:.3f} ms + sweeps {refresh_ms:.3f} ms, call "
        f"{call_s * 1e3:.3f} ms (upload of the whole factor {upload_s * 1e3:.1f} ms); values "
        f"{out['bit_exact_values']}" + (f"; CPU {out['cpu_baseline']['ms']:.3f} ms" if "cpu_baseline" in out else ""))
    return out


def copy_peak_gbs(values: int = 1 << 28, repeats: int = 10) -> float:
    """The HBM bandwidth a plain device-to-device copy reaches on this GPU (2 GiB of FP64 read
    and written per copy, timed with events on PyTorch's stream), the practical ceiling beside the
    published 8 TB/s."""
    import torch

    src = torch.ones(values, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(repeats):
        dst.copy_(src)
    end.record()
    end.synchronize()
    ms = start.elapsed_time(end) / repeats
    del src, dst
    torch.cuda.empty_cache()
    return 2 * values * 8 / (ms / 1e3) / 1e9


def random_line_peaks_gbs(mib: int = 1024, row_bytes: int = 128, repeats: int = 5):
    """What this GPU moves for the access pattern of the wide sweeps, whole 128-byte lines at
    random rows of a large block, beside the plain copy: (gather, scatter) in GB/s of bytes read
    plus written. The gather reads `mib` MiB of rows in a seeded random order and writes them in
    order (torch.index_select); the scatter reads them in order and writes them to random rows
    (Tensor.index_copy_). Timed with events on PyTorch's stream."""
    import torch

    width = row_bytes // 8
    rows = mib * (1 << 20) // row_bytes
    table = torch.ones((rows, width), dtype=torch.float64, device="cuda")
    out = torch.empty_like(table)
    order = torch.randperm(rows, generator=torch.Generator().manual_seed(SEED)).to("cuda")
    rates = []
    for op in (lambda: torch.index_select(table, 0, order, out=out), lambda: out.index_copy_(0, order, table)):
        for _ in range(2):
            op()
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(repeats):
            op()
        end.record()
        end.synchronize()
        ms = start.elapsed_time(end) / repeats
        rates.append(2 * rows * row_bytes / (ms / 1e3) / 1e9)
    del table, out, order
    torch.cuda.empty_cache()
    return rates[0], rates[1]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-complex", action="store_true")
    ap.add_argument("--no-screening", action="store_true")
    ap.add_argument("--no-flows", action="store_true")
    ap.add_argument("--no-refactor", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB on the device (to scale HBM counters) and exit")
    ap.add_argument("--gather-calibration", type=int, default=0, metavar="MIB",
                    help="only gather MIB MiB as random whole rows on the device (to scale HBM counters) and exit")
    ap.add_argument("--gather-row-bytes", type=int, default=128)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the factor-solve path is measured on one GPU's share "
              "of the work (right-hand-side blocks shard across GPUs with no exchange); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gather_calibration:
        gather_calibration(args.gather_calibration, args.gather_row_bytes)
        return

    # PyTorch ships its own HIP runtime: it must be loaded before libybus_hip.so, so that both use
    # that one runtime (loaded the other way round, PyTorch finds no GPU)
    import torch  # noqa: F401

    import ybus_amd as Y

    gpu = Y.Gpu.new()
    if gpu is None:
        raise RuntimeError("no HIP device")
    log(f"[bench] device {gpu.name()}")
    ptdf = run_ptdf(gpu, args)
    nnz_l, rhs_count = ptdf["nnz_l"], ptdf["rhs"]
    ms_per_step = ptdf["elapsed"] / args.steps * 1e3
    value = rhs_count * nnz_l / (ms_per_step / 1e3)
    stats = ptdf["stats"]
    kernel_ms = stats.milliseconds / stats.launches
    bytes_per_launch = stats.bytes / stats.launches
    achieved = bytes_per_launch / (kernel_ms / 1e3) / 1e9
    traffic, traffic_source = profiled_traffic(ptdf["kernel"])
    copy_gbs = copy_peak_gbs()
    log(f"[bench] device copy of 2 GiB: {copy_gbs:.0f} GB/s (read + written)")
    gather_gbs, scatter_gbs = random_line_peaks_gbs()
    log(f"[bench] 128-byte lines at random rows of 1 GiB: gather {gather_gbs:.0f} GB/s, scatter {scatter_gbs:.0f} GB/s")

    # parity: every column against the CPU's solve of its dense vector (bit for bit)
    from oracle import configs

    chol = ptdf["chol"]
    bit_exact, max_rel = 0, 0.0
    for start in range(0, rhs_count, 1500):
        stop = min(rhs_count, start + 1500)
        cpu_cols = chol.solve_many(configs.dense_pairs(ptdf["plus"], ptdf["minus"], chol.n, start, stop), cpu_threads())
        got = ptdf["solutions"][start:stop]
        bit_exact += int(np.sum([a.tobytes() == b.tobytes() for a, b in zip(got, cpu_cols)]))
        max_rel = max(max_rel, float(np.max(np.abs(got - cpu_cols)) / max(np.max(np.abs(cpu_cols)), 1e-300)))

    cpu_baseline = None
    if not args.no_cpu:
        rate, threads, _ = cpu_pair_baseline(chol, ptdf["plus"], ptdf["minus"], min(2000, rhs_count))
        serial, _, _ = cpu_pair_baseline(chol, ptdf["plus"], ptdf["minus"], min(250, rhs_count), 1, 2.0)
        cpu_baseline = {"value": rate * nnz_l, "unit": "RHS·nnz(L)/s", "cores": threads, "kind": "port",
                        "cores_available": cpu_affinity(), "os_cpu_count": os.cpu_count(),
                        "omp_num_threads": thread_cap() or None,
                        "sample": "par_solve_many (oracle, one RHS per thread) on the first 2,000 of the 13,500 "
                                  "pair RHS as dense vectors, repeated for >= %g s; serial: solve_many on the "
                                  "first 250, repeated for >= 2 s" % CPU_SECONDS,
                        "serial_value": serial * nnz_l, "model": cpu_model(),
                        "ms_per_block": rhs_count / rate * 1e3, "serial_ms_per_block": rhs_count / serial * 1e3}

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "RHS·nnz(L)/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "PTDF block: 10,000-bus synthetic grid, B' AMD Cholesky, 13,500 pair RHS built "
                               "on the device, FP64 forward+backward", "buses": 10000, "n": int(chol.n),
                   "nnz_l": nnz_l, "rhs": rhs_count, "parallelism": "single GPU", "tuning": ptdf["tuning"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_source,
                     "kernel": ptdf["kernel"], "kernel_ms": kernel_ms, "bytes_per_launch": bytes_per_launch,
                     "copy_peak": copy_gbs, "frac_of_copy_peak": achieved / copy_gbs,
                     "random_line_gather_peak": gather_gbs, "random_line_scatter_peak": scatter_gbs,
                     "counted_traffic_gbs": (traffic / (kernel_ms / 1e3) / 1e9) if traffic else None},
        "cpu_baseline": cpu_baseline,
        "end_to_end_ms": ptdf["end_to_end"] * 1e3,
        "parity": {"columns": rhs_count, "bit_exact_columns": bit_exact, "max_rel_diff_vs_cpu": max_rel},
    }
    if not args.no_precision:
        line["precision"] = run_precision(gpu, args)
    if not args.no_adjoint:
        line["adjoint"] = run_adjoint(gpu, args)
    if not args.no_batch:
        line["batch"] = run_batch(gpu, args)
    if not args.no_flows:
        line["flows"] = run_flows(gpu, args, ptdf)
    if not args.no_screening:
        line["screening"] = run_screening(gpu, args)
    if not args.no_complex:
        line["complex"] = run_complex(gpu, args)
    if not args.no_refactor:
        line["refactor"] = run_refactor(gpu, args)
    # where each part of the metric's name is in this line
    parts = {
        "RHS·nnz(L)/s": "value: the FP64 PTDF block (13,500 pair RHS on the 10,000-bus grid's Cholesky factor)",
        "HBM GB/s": "roofline.achieved (algorithmic bytes / kernel time), roofline.counted_traffic_gbs",
        "1e-12 resid": "parity (bit-exact against the CPU); precision.*.max_relative_residual (the FP64 floor "
                       "of the 25,000-bus Jacobian is above 1e-12 for any FP64 solution)",
    }
    if "precision" in line:
        p = line["precision"]
        parts["FP32+refine vs FP64"] = {
            "fp64_rhs_nnz_per_s": p["fp64"]["value"],
            "fp32_refined_rhs_nnz_per_s": p["fp32_refined"]["value"],
            "ratio": p["speedup_fp32_refined_over_fp64"],
            "ratio_floor_stop": p.get("speedup_fp32_refined_floor_stop_over_fp64"),
        }
    line["metric_parts"] = parts
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
