"""Benchmark: time to a 1e-6 relative residual on one MI355X.

Workload (default): the 128x128x64 synthetic reservoir grid -- 1,048,576
cells, 10% inactive, 1% non-neighbour connections, 3x3 blocks
(943,550 block rows, 6.0M blocks, 2.83M unknowns), b ~ U[-1,1], x0 = 0.

One "step" is one complete solve to rel tol 1e-6 with b and x resident in
HBM.  Both forms are timed: the plain FP64 ILU0-BiCGSTAB and the precision
form (FP32 matrix/factor values and inner vectors, FP64 sums, scalars,
diagonal inverses and refinement).  ``value`` is the precision form's time
per solve; the FP64 time, iteration counts, the error of each against a
1e-12 FP64 solve, the roofline of the dominant kernel (the precision form's
FP32 SpMV, timed inside a solve) and the CPU oracle baseline are in the same
JSON line.

Beside the headline the line carries, measured in the same run:
  trisolve   back-to-back preconditioner applications (all L and U sweep
             launches), FP64 and FP32, against their algorithmic bytes;
  level_scheduling  the same system ordered by level sets instead of colours:
             FP64 and precision-form solves, one preconditioner application
             in each form and its launch count, beside the coloured figures;
  adjoint    the 64x64x32 grid's adjoint backward (A^T lambda = w to 1e-10,
             then the gradient -lambda x^T on the pattern), FP64 and in the
             precision form (with lambda's difference from the FP64 one);
  grid256x256x64  the 4x4-block grid: FP64 forward, precision-form forward and
             adjoint backward (FP64 and precision form), each with its time;
  cpu_baseline    the CPU oracle on the host cores: the 128x128x64 solve (all
             threads, median of 3; one thread, one full solve) and the
             32x32x16 solve of the CPU-runnable case (median of 5, all threads
             and one thread, with the colouring's SHA-256).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="grid128x128x64")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--inner-tol", type=float, default=1e-4)
    ap.add_argument("--mode", choices=["full", "kernels", "level"], default="full",
                    help="kernels: only repeated SpMV launches and preconditioner applications "
                         "(FP64 and FP32 values) and a reference copy of known size, for rocprofv3 "
                         "counter runs; level: only the level-scheduling block (its JSON), for a "
                         "kernel trace of the level-set launches")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the level-scheduling, adjoint (64x64x32) and 4x4 (256x256x64) blocks")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this solver runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding over several GPUs is not built)", flush=True)
        return 2
    import torch

    import fpga_amd as fa
    from fpga_amd.configs import WORKLOADS

    wl = WORKLOADS[args.workload]
    spec, tol, maxit = wl["spec"], 1e-6, wl["maxit"]
    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")

    t0 = time.perf_counter()
    s = fa.generate(spec)
    t_gen = time.perf_counter() - t0
    log(f"generated {args.workload}: n={s.n} nnzb={s.nnzb} N={s.N} in {t_gen:.2f}s")

    g = fa.BicgstabSolver(s.bs, s.n, s.nnzb)
    g.upload_system(s)
    t0 = time.perf_counter()
    g.analyse()
    t_an = time.perf_counter() - t0
    t0 = time.perf_counter()
    g.factor()
    torch.cuda.synchronize()
    t_fac = time.perf_counter() - t0
    t0 = time.perf_counter()
    g.prepare_mixed()     # the precision form's FP32 copies (paid again after every factorisation)
    t_dem = time.perf_counter() - t0
    log(f"analyse {t_an:.3f}s factor {t_fac:.3f}s fp32 copies {t_dem:.4f}s colours {g.num_colours}")

    b = torch.from_numpy(s.b).to(dev)
    x = torch.zeros(s.N, dtype=torch.float64, device=dev)

    if args.mode == "kernels":
        return kernels_only(args, g, s, torch, dev, fa)

    # reference solution for the error of both forms (untimed)
    g.solve_dev(b.data_ptr(), x.data_ptr(), 1e-12, 5000)
    x_ref = x.clone()
    nref = torch.linalg.vector_norm(x_ref).item()
    if args.mode == "level":
        g.close()
        print(json.dumps(level_block(fa, torch, dev, args.workload, s, b, x_ref, tol, maxit, args.steps, None)),
              flush=True)
        return 0

    def run_fp64():
        x.zero_()
        torch.cuda.synchronize()
        return g.solve_dev(b.data_ptr(), x.data_ptr(), tol, maxit)

    def run_mixed():
        x.zero_()
        torch.cuda.synchronize()
        return g.solve_mixed_dev(b.data_ptr(), x.data_ptr(), tol, 20, args.inner_tol, maxit, True)

    for _ in range(args.warmup):
        run_fp64()
        run_mixed()

    # FP64 form
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r64 = run_fp64()
    torch.cuda.synchronize()
    t_fp64 = (time.perf_counter() - t0) / args.steps
    err64 = (torch.linalg.vector_norm(x - x_ref).item() / nref)

    # one more FP64 solve, untimed, with every launch bracketed by HIP events
    # on the solver's stream (per-kernel-class averages)
    g.set_profiling(True)
    run_fp64()
    spmv64_in = in_solve_avg(g.kernel_times(fa.KIND_SPMV))
    n_low, ms_low = g.kernel_stats(fa.KIND_LOWER)
    n_up, ms_up = g.kernel_stats(fa.KIND_UPPER)
    n_vec, ms_vec = g.kernel_stats(fa.KIND_VEC)
    g.set_profiling(False)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rmx = run_mixed()
    torch.cuda.synchronize()
    t_mixed = (time.perf_counter() - t0) / args.steps
    errmx = (torch.linalg.vector_norm(x - x_ref).item() / nref)

    # one more precision-form solve, untimed, with every SpMV launch (the
    # streaming kernel alone) bracketed by HIP events on the solver's stream
    # (each event pair adds a few microseconds of its own to one launch)
    g.set_profiling(True)
    run_mixed()
    spmv32_in = in_solve_avg(g.kernel_times(fa.KIND_SPMV))
    g.set_profiling(False)

    spmv_bytes = g.algorithmic_bytes(fa.KIND_SPMV)
    spmv32_bytes = g.algorithmic_bytes(fa.KIND_SPMV, True)
    # roofline of the dominant kernel of the timed solve, the precision
    # form's FP32 SpMV: 20 back-to-back launches between two HIP events on
    # the solver's stream (per launch the same time as inside a solve in the
    # kernel trace; the FP64 one gains a few percent back to back, part of
    # its values staying in the 256 MB last-level cache between launches)
    spmv_ms = g.time_spmv(20, False)
    spmv32_ms = g.time_spmv(20, True)
    # the roofline kernel is the SpMV instance with the fused v.rt sums: its
    # algorithmic bytes are the SpMV's (values, columns, row pointers, x once,
    # v once) plus the rt operand it reads for the dot
    roof_bytes = spmv32_bytes + s.N * 4
    spmv32_gbs = roof_bytes / (spmv32_ms * 1e-3) / 1e9
    read_ceiling, copy_ceiling = g.measure_bandwidth(1 << 30, 10)
    traffic = None
    tfile = ROOT / "profiles" / f"traffic_spmv_{args.workload}.json"
    if tfile.exists():
        t = json.loads(tfile.read_text())
        if "fp32" in t:
            traffic = t["fp32"]["read_bytes"] + t["fp32"]["write_bytes"]
    low_bytes, up_bytes = g.algorithmic_bytes(fa.KIND_LOWER), g.algorithmic_bytes(fa.KIND_UPPER)
    iters64 = r64.iterations
    # algorithmic bytes per FP64 iteration: 2 SpMV + 2 (L + U) + 22 N vector passes
    it_bytes = 2 * spmv_bytes + 2 * (low_bytes + up_bytes) + 22 * s.N * 8
    solve_gbs = it_bytes * iters64 / t_fp64 / 1e9
    mixed_it_bytes = (2 * g.algorithmic_bytes(fa.KIND_SPMV, True)
                      + 2 * (g.algorithmic_bytes(fa.KIND_LOWER, True) + g.algorithmic_bytes(fa.KIND_UPPER, True))
                      + 22 * s.N * 4)
    # one preconditioner application (L then U sweep) back to back, both forms
    trisolve = {}
    for fp32 in (False, True):
        ms = g.time_precond(20, fp32)
        by = g.algorithmic_bytes(fa.KIND_LOWER, fp32) + g.algorithmic_bytes(fa.KIND_UPPER, fp32)
        trisolve["fp32" if fp32 else "fp64"] = {"avg_ms": ms, "algorithmic_bytes": by,
                                                "GBs": by / (ms * 1e-3) / 1e9,
                                                "frac": by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    ncolours = g.num_colours
    g.close()

    extra_adj = extra_big = extra_level = None
    if not args.no_extra:
        extra_level = level_block(fa, torch, dev, args.workload, s, b, x_ref, tol, maxit, min(args.steps, 3),
                                  {"fp64": (t_fp64, iters64, trisolve["fp64"]["avg_ms"]),
                                   "mixed": (t_mixed, rmx.iterations, trisolve["fp32"]["avg_ms"])})
        extra_adj = adjoint_block(fa, torch, dev, min(args.steps, 5))
        extra_big = big_block(fa, torch, dev, min(args.steps, 3))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(fa, s, args, tol, maxit)
    out = {
        "metric": "time-to-1e-6 rel-res (s), mixed vs FP64 on GPU; achieved HBM GB/s",
        "value": t_mixed,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mixed * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 values/vectors, f64 sums+refinement",
        "data": "synthetic",
        "config": {"workload": args.workload, "grid": [spec.nx, spec.ny, spec.nz], "block": s.bs,
                   "block_rows": s.n, "blocks": s.nnzb, "unknowns": s.N, "colours": ncolours,
                   "tol": tol, "inner_tol": args.inner_tol},
        "fp64": {"time_s": t_fp64, "iterations": iters64, "rel_residual": r64.rel_residual,
                 "x_rel_err_vs_1e-12": err64, "iters_per_s": iters64 / t_fp64,
                 "solve_GBs_algorithmic": solve_gbs},
        "mixed": {"time_s": t_mixed, "outer": rmx.outer_iterations, "inner_iterations": rmx.iterations,
                  "rel_residual": rmx.rel_residual, "x_rel_err_vs_1e-12": errmx,
                  "speedup_vs_fp64": t_fp64 / t_mixed,
                  "inner_GBs_algorithmic_approx": mixed_it_bytes * rmx.iterations / t_mixed / 1e9},
        "kernels_ms_avg": {"spmv_back_to_back": spmv_ms, "spmv_fp32_back_to_back": spmv32_ms,
                           "spmv_in_solve_event_pairs": spmv64_in, "spmv_fp32_in_solve_event_pairs": spmv32_in,
                           "fp64_in_solve_event_pairs": {"lower_sweep": ms_low / max(n_low, 1),
                                                         "upper_sweep": ms_up / max(n_up, 1),
                                                         "vector_updates_per_half_iter": ms_vec / max(n_vec, 1)},
                           "algorithmic_bytes": {"spmv": spmv_bytes, "spmv_fp32": spmv32_bytes}},
        "roofline": {"bound": "hbm", "achieved": spmv32_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": spmv32_gbs / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"k_stream<{s.bs}, SM_SPMV, SR_W0, float, float> (the precision form's SpMV "
                               "with the fused v.rt sums)",
                     "avg_launch_ms": spmv32_ms, "algorithmic_bytes_per_launch": roof_bytes,
                     "algorithmic_bytes_spmv_only": spmv32_bytes,
                     "measured_read_ceiling_GBs": read_ceiling, "measured_copy_ceiling_GBs": copy_ceiling,
                     "frac_of_measured_read_ceiling": spmv32_gbs / read_ceiling},
        "trisolve": trisolve,
        "setup_s": {"generate": t_gen, "analyse": t_an, "factor": t_fac, "fp32_copies": t_dem},
        "level_scheduling": extra_level,
        "adjoint": extra_adj,
        "grid256x256x64": extra_big,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def iteration_bytes(g, fa, s, fp32):
    """Algorithmic bytes of one BiCGSTAB iteration: 2 SpMV + 2 (L + U) + 22 N
    vector element passes (the same for the transposed operators)."""
    sx = 4 if fp32 else 8
    return (2 * g.algorithmic_bytes(fa.KIND_SPMV, fp32)
            + 2 * (g.algorithmic_bytes(fa.KIND_LOWER, fp32) + g.algorithmic_bytes(fa.KIND_UPPER, fp32))
            + 22 * s.N * sx)


def grad_bytes(s):
    """The gradient pass: one FP64 value written per matrix entry, the block's
    row and column indices and source slot read, lambda and x read."""
    return s.nnzb * s.bs * s.bs * 8 + s.nnzb * 12 + 2 * s.N * 8


def setup(fa, spec, torch):
    t0 = time.perf_counter()
    s = fa.generate(spec)
    t_gen = time.perf_counter() - t0
    g = fa.BicgstabSolver(s.bs, s.n, s.nnzb)
    g.upload_system(s)
    t0 = time.perf_counter()
    g.analyse()
    t_an = time.perf_counter() - t0
    t0 = time.perf_counter()
    g.factor()
    t_fac = time.perf_counter() - t0
    return s, g, {"generate": t_gen, "analyse": t_an, "factor": t_fac}


def level_block(fa, torch, dev, name, s, b, x_ref, tol, maxit, steps, coloured):
    """The same system and right-hand side with level scheduling instead of
    the colouring (the host API's level_scheduling flag,
    fpga_functions_bicgstab.cpp:84): one level set per sweep launch, so the
    preconditioner costs one launch per level and sweep, against fewer
    iterations from the natural-order ILU0.  FP64 and precision-form solves
    to the same tolerance, their errors against the same 1e-12 solution, and
    one preconditioner application back to back in each form."""
    g = fa.BicgstabSolver(s.bs, s.n, s.nnzb)
    g.upload_system(s)
    t0 = time.perf_counter()
    g.analyse(level_scheduling=True)
    t_an = time.perf_counter() - t0
    g.factor()
    g.prepare_mixed()
    sizes = np.diff(g.colouring()[2])
    x = torch.zeros_like(x_ref)
    nref = torch.linalg.vector_norm(x_ref).item()

    def timed(fn):
        x.zero_()
        fn()
        times, r = [], None
        for _ in range(steps):
            x.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = fn()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        return float(np.median(times)), r, torch.linalg.vector_norm(x - x_ref).item() / nref

    t64, r64, e64 = timed(lambda: g.solve_dev(b.data_ptr(), x.data_ptr(), tol, maxit))
    tmx, rmx, emx = timed(lambda: g.solve_mixed_dev(b.data_ptr(), x.data_ptr(), tol, 20, 1e-4, maxit, True))
    pre64, pre32 = g.time_precond(5, False), g.time_precond(5, True)
    launches = g.precond_launches()
    g.close()
    out = {"workload": f"{name}, same system, level sets of the lower-triangular graph",
           "levels": int(sizes.size), "rows_per_level": {"max": int(sizes.max()), "median": int(np.median(sizes)),
                                                         "min": int(sizes.min())},
           "analyse_s": t_an, "precond_launches_per_application": launches, "steps": steps,
           "fp64": {"time_s": t64, "iterations": r64.iterations, "converged": r64.converged,
                    "x_rel_err_vs_1e-12": e64, "precond_application_ms": pre64},
           "mixed": {"time_s": tmx, "outer": rmx.outer_iterations, "inner_iterations": rmx.iterations,
                     "converged": rmx.converged, "x_rel_err_vs_1e-12": emx, "precond_application_ms": pre32},
           "coloured": None if coloured is None else
                       {k: {"time_s": v[0], "iterations": v[1], "precond_application_ms": v[2]}
                        for k, v in coloured.items()}}
    log(f"level scheduling {name}: {sizes.size} levels, fp64 {t64 * 1e3:.1f} ms ({r64.iterations} it), "
        f"mixed {tmx * 1e3:.1f} ms ({rmx.iterations} inner it), precond {pre64:.3f} / {pre32:.3f} ms, "
        f"{launches} launches")
    return out


def adjoint_block(fa, torch, dev, steps):
    """The 64x64x32 grid's backward pass: A^T lambda = w (w ~ N(0,1)) to
    1e-10 with the transposed colour-ordered ILU0, then dL/dA = -lambda x^T."""
    from fpga_amd.configs import GRID_64x64x32
    s, g, st = setup(fa, GRID_64x64x32, torch)
    b = torch.from_numpy(s.b).to(dev)
    x = torch.zeros(s.N, dtype=torch.float64, device=dev)
    g.solve_dev(b.data_ptr(), x.data_ptr(), 1e-10, 1000)
    w = torch.from_numpy(np.random.default_rng(0).standard_normal(s.N)).to(dev)
    lam = torch.zeros_like(w)
    grad = torch.empty((s.nnzb, s.bs, s.bs), dtype=torch.float64, device=dev)
    g.adjoint_solve_dev(w.data_ptr(), lam.data_ptr(), 1e-10, 1000)   # builds the transposed operators
    times, ra = [], None
    for _ in range(steps):
        lam.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ra = g.adjoint_solve_dev(w.data_ptr(), lam.data_ptr(), 1e-10, 1000)
        g.grad_values_dev(lam.data_ptr(), x.data_ptr(), grad.data_ptr())
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    by = iteration_bytes(g, fa, s, False) * ra.iterations + grad_bytes(s)
    lam_ref = lam.clone()
    nref = torch.linalg.vector_norm(lam_ref).item()
    g.adjoint_solve_mixed_dev(w.data_ptr(), lam.data_ptr(), 1e-10)   # builds the FP32 transposes
    times_m, rm = [], None
    for _ in range(steps):
        lam.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rm = g.adjoint_solve_mixed_dev(w.data_ptr(), lam.data_ptr(), 1e-10)
        g.grad_values_dev(lam.data_ptr(), x.data_ptr(), grad.data_ptr())
        torch.cuda.synchronize()
        times_m.append(time.perf_counter() - t0)
    tm = float(np.median(times_m))
    em = torch.linalg.vector_norm(lam - lam_ref).item() / nref
    out = {"workload": "grid64x64x32", "block_rows": s.n, "blocks": s.nnzb, "time_s": t, "tol": 1e-10,
           "iterations": ra.iterations, "converged": ra.converged, "rel_residual": ra.rel_residual,
           "GBs_algorithmic": by / t / 1e9, "steps": steps,
           "mixed": {"time_s": tm, "outer": rm.outer_iterations, "inner_iterations": rm.iterations,
                     "converged": rm.converged, "lambda_rel_diff_vs_fp64": em, "speedup_vs_fp64": t / tm},
           "note": "175 MB resident set: cache-resident, GB/s above HBM peak is possible"}
    g.close()
    log(f"adjoint grid64x64x32: {t * 1e3:.2f} ms, {ra.iterations} iterations; precision form {tm * 1e3:.2f} ms")
    return out


def big_block(fa, torch, dev, steps):
    """The 256x256x64 grid with 4x4 blocks: FP64 forward and precision-form
    forward to 1e-6, adjoint backward to 1e-10 with the gradient; x errors
    against a 1e-10 FP64 solve."""
    from fpga_amd.configs import GRID_256x256x64
    s, g, st = setup(fa, GRID_256x256x64, torch)
    t0 = time.perf_counter()
    g.prepare_mixed()
    st["fp32_copies"] = time.perf_counter() - t0
    b = torch.from_numpy(s.b).to(dev)
    x = torch.zeros(s.N, dtype=torch.float64, device=dev)
    g.solve_dev(b.data_ptr(), x.data_ptr(), 1e-10, 2000)
    x_ref = x.clone()
    nref = torch.linalg.vector_norm(x_ref).item()

    def timed(fn):
        x.zero_()
        fn()
        times, r = [], None
        for _ in range(steps):
            x.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = fn()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        return float(np.median(times)), r

    t64, r64 = timed(lambda: g.solve_dev(b.data_ptr(), x.data_ptr(), 1e-6, 1000))
    e64 = torch.linalg.vector_norm(x - x_ref).item() / nref
    tmx, rmx = timed(lambda: g.solve_mixed_dev(b.data_ptr(), x.data_ptr(), 1e-6, 20, 1e-4, 1000, True))
    emx = torch.linalg.vector_norm(x - x_ref).item() / nref
    w = torch.from_numpy(np.random.default_rng(0).standard_normal(s.N)).to(dev)
    lam = torch.zeros_like(w)
    grad = torch.empty((s.nnzb, s.bs, s.bs), dtype=torch.float64, device=dev)

    def adj():
        lam.zero_()
        r = g.adjoint_solve_dev(w.data_ptr(), lam.data_ptr(), 1e-10, 1000)
        g.grad_values_dev(lam.data_ptr(), x_ref.data_ptr(), grad.data_ptr())
        return r

    tad, rad = timed(adj)
    lam_ref = lam.clone()
    nlam = torch.linalg.vector_norm(lam_ref).item()

    def adj_mixed():
        lam.zero_()
        r = g.adjoint_solve_mixed_dev(w.data_ptr(), lam.data_ptr(), 1e-10)
        g.grad_values_dev(lam.data_ptr(), x_ref.data_ptr(), grad.data_ptr())
        return r

    tam, ram = timed(adj_mixed)
    eam = torch.linalg.vector_norm(lam - lam_ref).item() / nlam
    it64, it32 = iteration_bytes(g, fa, s, False), iteration_bytes(g, fa, s, True)
    out = {"workload": "grid256x256x64", "block": s.bs, "block_rows": s.n, "blocks": s.nnzb, "unknowns": s.N,
           "colours": g.num_colours, "setup_s": st, "steps": steps,
           "fp64": {"time_s": t64, "iterations": r64.iterations, "x_rel_err_vs_1e-10": e64,
                    "GBs_algorithmic": it64 * r64.iterations / t64 / 1e9},
           "mixed": {"time_s": tmx, "outer": rmx.outer_iterations, "inner_iterations": rmx.iterations,
                     "x_rel_err_vs_1e-10": emx, "speedup_vs_fp64": t64 / tmx,
                     "GBs_algorithmic_approx": it32 * rmx.iterations / tmx / 1e9},
           "adjoint": {"time_s": tad, "tol": 1e-10, "iterations": rad.iterations, "converged": rad.converged,
                       "GBs_algorithmic": (it64 * rad.iterations + grad_bytes(s)) / tad / 1e9},
           "adjoint_mixed": {"time_s": tam, "tol": 1e-10, "outer": ram.outer_iterations,
                             "inner_iterations": ram.iterations, "converged": ram.converged,
                             "lambda_rel_diff_vs_fp64": eam, "speedup_vs_fp64_adjoint": tad / tam}}
    g.close()
    log(f"grid256x256x64: fp64 {t64 * 1e3:.1f} ms, mixed {tmx * 1e3:.1f} ms, adjoint {tad * 1e3:.1f} ms, "
        f"precision-form adjoint {tam * 1e3:.1f} ms")
    return out


def cpu_baseline(fa, s, args, tol, maxit):
    """The CPU oracle (the same algorithm, OpenMP over the rows of a colour)
    on this machine's host cores."""
    import hashlib

    import oracle
    from fpga_amd.configs import GRID_32x32x16
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    o = oracle.OracleSystem(s.bs, s.rowptr, s.colidx, s.values)
    o.factor()
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        xo, ro = o.solve(s.b, tol=tol, maxit=maxit)
        times.append(time.perf_counter() - t0)
    oracle.set_threads(1)
    t0 = time.perf_counter()
    _, r1 = o.solve(s.b, tol=tol, maxit=maxit)
    t1 = time.perf_counter() - t0
    del o
    # the CPU-runnable case: 32x32x16, rel tol 1e-2, maxit 200
    c1 = fa.generate(GRID_32x32x16)
    res1 = {}
    for nth in (threads, 1):
        oracle.set_threads(nth)
        t0 = time.perf_counter()
        o1 = oracle.OracleSystem(c1.bs, c1.rowptr, c1.colidx, c1.values)
        t_col = time.perf_counter() - t0
        t0 = time.perf_counter()
        o1.factor()
        t_f = time.perf_counter() - t0
        ts = []
        for _ in range(5):
            t0 = time.perf_counter()
            _, rc1 = o1.solve(c1.b, tol=1e-2, maxit=200)
            ts.append(time.perf_counter() - t0)
        colour, perm, _ = o1.colouring()
        res1[f"threads_{nth}"] = {"solve_s_median": float(np.median(ts)), "solve_s": ts,
                                  "colour_and_permute_s": t_col, "factor_s": t_f}
        res1["iterations"] = rc1["iterations"]
        res1["half_iterations"] = rc1["half_iterations"]
        res1["rel_residual"] = rc1["rel_residual"]
        res1["colour_sha256"] = hashlib.sha256(colour.tobytes()).hexdigest()
        res1["perm_sha256"] = hashlib.sha256(perm.tobytes()).hexdigest()
        del o1
    oracle.set_threads(threads)
    med = float(np.median(times))
    log(f"cpu oracle: {times} s ({ro['iterations']} iterations), 1 thread {t1:.2f} s; "
        f"grid32x32x16 {res1[f'threads_{threads}']['solve_s_median'] * 1e3:.2f} ms")
    return {"value": med, "unit": "s", "cores": threads, "kind": "port",
            "sample": f"full FP64 solve of {args.workload} to rel 1e-6 ({ro['iterations']} iterations), "
                      f"median of 3, OpenMP over the rows of a colour",
            "s_per_iteration": med / ro["iterations"],
            "nproc": os.cpu_count(), "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"),
            "one_thread": {"solve_s": t1, "iterations": r1["iterations"],
                           "sample": "one full FP64 solve to rel 1e-6 on 1 thread"},
            "grid32x32x16": dict(res1, sample="FP64 solve to rel 1e-2, maxit 200, median of 5, "
                                             "all threads and 1 thread; colouring SHA-256 of colour[] and perm[]")}


def in_solve_avg(times_ms):
    """Average of the event-timed launches of one solve.  Launches queued
    after the device-side stop return at once (a few microseconds); they are
    the ones below half the median and are left out."""
    t = np.asarray(times_ms, dtype=np.float64)
    if t.size == 0:
        return {"avg_ms": float("nan"), "launches": 0, "no_op_launches": 0}
    keep = t[t >= 0.5 * np.median(t)]
    return {"avg_ms": float(keep.mean()), "launches": int(keep.size), "no_op_launches": int(t.size - keep.size)}


def kernels_only(args, g, s, torch, dev, fa):
    """Repeated SpMV launches for profiling; prints one JSON line."""
    res = {"workload": args.workload, "reps": args.reps}
    for fp32 in (False, True):
        ms = g.time_spmv(args.reps, fp32)
        by = g.algorithmic_bytes(fa.KIND_SPMV, fp32)
        res["spmv_fp32" if fp32 else "spmv_fp64"] = {"avg_ms": ms, "algorithmic_bytes": by,
                                                     "GBs": by / (ms * 1e-3) / 1e9}
    for fp32 in (False, True):
        ms = g.time_precond(args.reps, fp32)
        by = g.algorithmic_bytes(fa.KIND_LOWER, fp32) + g.algorithmic_bytes(fa.KIND_UPPER, fp32)
        res["precond_fp32" if fp32 else "precond_fp64"] = {"avg_ms": ms, "algorithmic_bytes": by,
                                                           "GBs": by / (ms * 1e-3) / 1e9}
    # reference copy of known size (64 Mi doubles = 512 MiB read + 512 MiB written)
    src = torch.ones(1 << 26, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.reps
    res["copy_512MiB"] = {"avg_ms": ms, "GBs": 2 * src.numel() * 8 / (ms * 1e-3) / 1e9}
    print(json.dumps(res), flush=True)
    g.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
