"""HPCG PCG throughput on one MI355X.

A step is one PCG iteration (CG.cpp:100-130 of the reference): one V-cycle
(4 levels, 1 pre + 1 post symmetric Gauss-Seidel sweep, 8-colour ordering),
one SpMV, two dot products fused into their producers, the x/r/p updates.
The timed region runs K iterations as one CG call from x = 0 with tolerance
0, like the reference's timed sets (main.cpp:334-340), inputs resident in HBM.

Prints one JSON line:
  value      HPCG GFLOP/s with convergence overhead: the reference's flop model
             (ReportResults.cpp:75-99) over the timed iterations, times 50 / the
             iterations the 8-colour ordering needs to reach the residual that 50
             natural-order iterations reach (ReportResults.cpp:99,351); the raw
             figure is "raw_gflops"
  roofline   the finest level's colour-step Gauss-Seidel kernel: algorithmic
             bytes per launch / mean launch time (HIP events on its stream)
  convergence  FP64 PCG vs the mixed-precision PCG (FP32 V-cycle on FP16
             matrix values, FP64 outer loop) from x0 = 0 to ||r||/||r0|| <= 1e-9:
             iterations, time, error against the all-ones solution
  cpu_baseline  the reference itself (oracle/_ref/ref_time, OpenMP on the host
             cores), one CG_ref set on a bounded grid
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50, help="timed PCG iterations")
    ap.add_argument("--warmup", type=int, default=3, help="untimed PCG iterations")
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--nz", type=int, default=640)
    ap.add_argument("--ordering", default="multicolor", choices=["multicolor", "lexicographic"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-grid", type=int, default=256, help="grid edge of the CPU sample")
    ap.add_argument("--cpu-iters", type=int, default=10, help="CG_ref iterations of the CPU sample")
    ap.add_argument("--ref-iters", type=int, default=50,
                    help="natural-order (lexicographic) CG iterations whose residual the "
                         "optimized ordering must reach, for HPCG's convergence overhead "
                         "(main.cpp:212-224, ReportResults.cpp:99; 0 = skip and report the raw "
                         "GFLOP/s as the value)")
    ap.add_argument("--tol-compare", type=float, default=1e-9,
                    help="FP64 vs mixed-precision time to this residual (0 = skip)")
    ap.add_argument("--max-iters", type=int, default=2000,
                    help="iteration cap of the time-to-tolerance runs")
    ap.add_argument("--col16", type=int, default=1, choices=[0, 1],
                    help="16-bit column ids (tile base + offset) where built")
    ap.add_argument("--band-lines", type=int, default=0,
                    help="grid lines per band of the in-colour row order (0 = default, 1 = natural)")
    ap.add_argument("--sweep-order", type=int, default=0, choices=[0, 1],
                    help="colour visiting order of the multicolour sweeps (HPCGX_SWEEP_ORDER: "
                         "0 = default, 1 = colour id order)")
    ap.add_argument("--xcd-remap", type=int, default=0, choices=[0, 1])
    ap.add_argument("--spmv-dot-blocks", type=int, default=0, help="0 = library default")
    ap.add_argument("--stream-min-bytes", type=float, default=0,
                    help="non-temporal matrix loads above this matrix size (0 = library default)")
    ap.add_argument("--graph", type=int, default=0, choices=[0, 1],
                    help="replay a captured hipGraph of the CG iteration in the timed "
                         "region (the colour-step launches are then timed in a separate "
                         "eager pass of the same length)")
    ap.add_argument("--block", type=int, nargs="*", default=[8, 16],
                    help="also solve s right-hand sides at once (block PCG, FP32 V-cycle) "
                         "for each s given, to --tol-compare (none given = skip)")
    ap.add_argument("--block-grid", type=int, default=256,
                    help="grid edge of the block PCG problem (its own problem unless it "
                         "equals the main grid)")
    ap.add_argument("--fleet", type=int, default=32,
                    help="also run this many independent problems (grids drawn from "
                         "{32,48,64,96,128} per axis by splitmix64 seed 7) as one fleet, "
                         "--steps iterations, against the same problems one at a time "
                         "(0 = skip)")
    ap.add_argument("--fleet-only", action="store_true",
                    help="run only the fleet measurement and print its JSON")
    ap.add_argument("--block-only", action="store_true",
                    help="run only the block PCG measurement (--block, --block-grid) and "
                         "print its JSON")
    ap.add_argument("--block-iters", type=int, default=0,
                    help="block PCG: this many iterations at tolerance 0 instead of a solve to "
                         "--tol-compare, without the single-RHS comparison (for profiling)")
    ap.add_argument("--mixed-steps", type=int, default=0,
                    help="after the timed FP64 set, this many mixed-precision iterations at "
                         "tolerance 0 (for profiling the FP32 V-cycle on its own)")
    ap.add_argument("--fleet-compare", type=int, default=1, choices=[0, 1],
                    help="also solve the fleet's problems one at a time (0 = skip)")
    ap.add_argument("--autotune", type=int, default=1, choices=[0, 1],
                    help="pick the tuning knobs by timing (hpcgx_autotune, before the warm-up)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device (the known-size kernel that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters) and exit")
    return ap.parse_args()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(grid, iters, workload_rows):
    """Time the reference (compiled from its own sources into oracle/_ref)
    on the host cores; fall back to the single-threaded oracle port.  The
    threads are OMP_NUM_THREADS where it is set (the host's share of cores
    for this GPU), else every core the process sees."""
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_time")
    if os.path.exists(exe):
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread")
        out = subprocess.run([exe, "time", str(grid), str(grid), str(grid), str(iters)],
                             capture_output=True, text=True, env=env, timeout=600, check=True)
        r = json.loads(out.stdout.strip().splitlines()[-1])
        per_it = r["seconds"] / iters
        scale = workload_rows / float(grid ** 3)
        return {"value": round(r["gflops"], 4), "unit": "GFLOP/s", "cores": threads,
                "kind": "reference", "cpu_model": cpu_model(),
                "threads_source": "OMP_NUM_THREADS" if os.environ.get("OMP_NUM_THREADS")
                                  else "os.cpu_count()",
                "sample": "reference CG_ref (oracle/_ref/ref_time, -O3 -fopenmp), %d^3, "
                          "%d iterations, 4-level MG, %.1f s (HPCG flop model); its serial "
                          "SYMGS dominates" % (grid, iters, r["seconds"]),
                "seconds_per_iteration": round(per_it, 4),
                "seconds_per_iteration_scaled_to_workload": round(per_it * scale, 3),
                "scaling": "time per iteration x %.3f (the workload's rows over the "
                           "sample's; the reference's time grows with the rows)" % scale}
    sys.path.insert(0, ROOT)
    from oracle import pyoracle as O
    from hpcg_amd import model
    g = min(grid, 64)
    h = O.Hierarchy(g, g, g)
    t0 = time.perf_counter()
    h.cg(max_iter=iters, tol=0.0)
    dt = time.perf_counter() - t0
    lv = model.grid_levels(g, g, g)
    fl = model.hpcg_model([l["n"] for l in lv], [l["nnz"] for l in lv], 1, iters, iters)["total"]
    return {"value": round(fl / dt * 1e-9, 4), "unit": "GFLOP/s", "cores": 1, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": "oracle port CG, %d^3, %d iterations, %.1f s" % (g, iters, dt)}


def converge(P, b, lv, a, log):
    """FP64 PCG and the mixed-precision PCG (FP32 V-cycle on FP16 values, FP64
    outer loop, FP64 residual recomputed every 10 iterations), each from
    x0 = 0 to ||r||/||r0|| <= tol: iterations, time, error vs the all-ones
    solution, and the mixed solution's distance from the FP64 one."""
    import torch
    from hpcg_amd import model
    nrow = [l["n"] for l in lv]
    nnz = [l["nnz"] for l in lv]
    res = {"tolerance": a.tol_compare, "recompute_every": 10}
    xs = {}
    P.prepare_mixed()
    for mode in ("fp64", "mixed"):
        x = torch.zeros_like(b)
        if mode == "fp64":
            P.cg(b, x, max_iter=2, tolerance=0.0)  # warm
        else:
            P.cg_mixed(b, x, max_iter=2, tolerance=0.0)
        x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if mode == "fp64":
            r = P.cg(b, x, max_iter=a.max_iters, tolerance=a.tol_compare, graph=True)
        else:
            r = P.cg_mixed(b, x, max_iter=a.max_iters, tolerance=a.tol_compare, recompute=10)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        k = r["niters"]
        fl = model.hpcg_model(nrow, nnz, 1, k, k)["total"]
        res[mode] = {"iterations": k, "seconds": round(dt, 4),
                     "ms_per_iteration": round(dt / max(k, 1) * 1e3, 4),
                     "normr_over_normr0": r["normr"] / r["normr0"],
                     "max_abs_err_vs_ones": float((x - 1.0).abs().max().item()),
                     "hpcg_gflops_equiv": round(fl / dt * 1e-9, 3)}
        xs[mode] = x
        log("%s: %d iterations to %.1e in %.3f s" % (mode, k, a.tol_compare, dt))
    res["mixed"]["max_abs_diff_vs_fp64"] = float((xs["mixed"] - xs["fp64"]).abs().max().item())
    res["speedup_time_to_tolerance"] = round(res["fp64"]["seconds"] / res["mixed"]["seconds"], 4)
    return res


def splitmix_uniform(seed, n):
    """splitmix64 stream mapped to uniform [-1, 1) (vectorised)."""
    import numpy as np
    with np.errstate(over="ignore"):
        s = np.uint64(seed) + np.uint64(0x9E3779B97F4A7C15) * np.arange(1, n + 1, dtype=np.uint64)
        z = (s ^ (s >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def block_solve(P, s, a, log):
    """Block PCG over s right-hand sides (splitmix64 uniform(-1, 1) from seed
    42, read row-major into [n][s]) with the FP32 V-cycle, to ||r||/||r0|| <=
    tol for every system, against s single-vector mixed solves of the same
    systems one after another."""
    import torch
    n = P.nrow(0)
    B = P.to_device_order_multi(splitmix_uniform(42, n * s).reshape(n, s))
    X = torch.zeros_like(B)
    P.cg_block_mixed(B, X, max_iter=2, tolerance=0.0)  # warm
    X.zero_()
    if a.block_iters > 0:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = P.cg_block_mixed(B, X, max_iter=a.block_iters, tolerance=0.0, recompute=10)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        log("block s=%d: %d iterations in %.3f s" % (s, r["niters"], dt))
        return {"s": s, "iterations": r["niters"], "seconds": round(dt, 4),
                "ms_per_iteration": round(dt / max(r["niters"], 1) * 1e3, 4)}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = P.cg_block_mixed(B, X, max_iter=a.max_iters, tolerance=a.tol_compare, recompute=10)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    log("block s=%d: %d iterations in %.3f s" % (s, r["niters"], dt))
    seq, its = 0.0, []
    for k in range(s):
        b = B[:, k].contiguous()
        x = torch.zeros_like(b)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        q = P.cg_mixed(b, x, max_iter=a.max_iters, tolerance=a.tol_compare, recompute=10)
        torch.cuda.synchronize()
        seq += time.perf_counter() - t1
        its.append(q["niters"])
    Y = torch.zeros_like(B)
    P.spmv_block(X, Y)
    true_res = max(float((B[:, k] - Y[:, k]).norm() / B[:, k].norm()) for k in range(s))
    return {"s": s, "iterations": r["niters"], "seconds": round(dt, 4),
            "ms_per_iteration": round(dt / max(r["niters"], 1) * 1e3, 4),
            "max_true_relative_residual": true_res,
            "single_rhs_seconds_total": round(seq, 4), "single_rhs_iterations": its,
            "speedup_vs_single_rhs": round(seq / dt, 4)}


def fleet_solve(count, a, log):
    """`count` independent HPCG problems (hpcg_amd.fleet_dims(count, seed=7))
    as one fleet: setup, autotune, then a timed set of --steps PCG iterations
    of every problem (tolerance 0, captured iteration), against the same
    problems solved one after another with the single-problem CG (captured
    iteration as well)."""
    import torch
    import hpcg_amd
    from hpcg_amd import model
    dims = hpcg_amd.fleet_dims(count, seed=7)
    log("fleet of %d problems, %d rows" % (count, sum(x * y * z for x, y, z in dims)))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    F = hpcg_amd.Problem.generate_fleet(dims, nlevels=4)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    tuned = F.autotune(ordering=a.ordering) if a.autotune else None
    b, x, _ = F.vectors()
    F.cg_fleet(b, x, max_iter=3, ordering=a.ordering, graph=True)  # warm, capture
    x.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = F.cg_fleet(b, x, max_iter=a.steps, tolerance=0.0, ordering=a.ordering, graph=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    K = a.steps
    flops, canon = 0.0, 0.0
    for d in dims:
        lv = model.grid_levels(*d)
        flops += model.hpcg_model([l["n"] for l in lv], [l["nnz"] for l in lv], 1, K, K)["total"]
        canon += model.canonical_bytes(lv)["iteration"] * K
    log("fleet: %d iterations in %.3f s" % (K, dt))
    F.close()
    seq, worst = 0.0, 0.0
    for q, d in enumerate(dims if a.fleet_compare else []):
        P = hpcg_amd.Problem.generate(*d, nlevels=4)
        bq, xq, _ = P.vectors()
        P.cg(bq, xq, max_iter=3, ordering=a.ordering, graph=True)
        xq.zero_()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        s1 = P.cg(bq, xq, max_iter=K, tolerance=0.0, ordering=a.ordering, graph=True)
        torch.cuda.synchronize()
        seq += time.perf_counter() - t1
        rel_f = r["normr"][q] / r["normr0"][q]
        rel_s = s1["normr"] / s1["normr0"]
        worst = max(worst, abs(rel_f - rel_s) / rel_s)
        P.close()
    log("one at a time: %.3f s" % seq)
    out = {"problems": count, "grids": [list(d) for d in dims],
           "rows": sum(x * y * z for x, y, z in dims),
           "iterations": K, "seconds": round(dt, 4), "ms_per_iteration": round(dt / K * 1e3, 4),
           "hpcg_gflops": round(flops / dt * 1e-9, 3),
           "iteration_gbs_canonical": round(canon / dt * 1e-9, 1),
           "iteration_hbm_frac": round(canon / dt * 1e-9 / HBM_PEAK_GBS, 4),
           "setup_s": round(setup, 3)}
    if a.fleet_compare:
        out.update({"one_at_a_time_seconds": round(seq, 4),
                    "speedup_vs_one_at_a_time": round(seq / dt, 4),
                    "max_rel_diff_normr_vs_single": worst})
    if tuned is not None:
        out["autotune"] = [l for l in tuned if l.startswith("chosen")]
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    if a.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 8}))
        return
    import hpcg_amd
    from hpcg_amd import model

    def log(*m):
        print("[bench]", *m, file=sys.stderr, flush=True)

    torch.cuda.init()
    if a.fleet_only:
        print(json.dumps({"fleet": fleet_solve(max(a.fleet, 1), a, log)}), flush=True)
        return
    if a.block_only:
        g = a.block_grid
        log("block PCG problem %d^3" % g)
        PB = hpcg_amd.Problem.generate(g, g, g, nlevels=4)
        out = [dict(block_solve(PB, s, a, log), grid=[g, g, g]) for s in (a.block or [8])]
        PB.close()
        print(json.dumps({"block_pcg": out}), flush=True)
        return
    log("generate %dx%dx%d on device" % (a.nx, a.ny, a.nz))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    P = hpcg_amd.Problem.generate(a.nx, a.ny, a.nz, nlevels=4, band_lines=a.band_lines,
                                  sweep_order=a.sweep_order)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    b, x, _ = P.vectors()
    P.tune("col16", a.col16)
    P.tune("xcd_remap", a.xcd_remap)
    if a.spmv_dot_blocks:
        P.tune("spmv_dot_blocks", a.spmv_dot_blocks)
    if a.stream_min_bytes:
        P.tune("stream_min_bytes", a.stream_min_bytes)
    log("setup %.2f s, device memory %.1f GB" % (setup_s, P.memory_use() / 1e9))
    tuned = None
    if a.autotune:
        tuned = P.autotune(ordering=a.ordering)
        log("autotune: " + "; ".join(l for l in tuned if l.startswith("chosen")))

    if a.warmup > 0:
        P.cg(b, x, max_iter=a.warmup, tolerance=0.0, ordering=a.ordering, graph=bool(a.graph))
    x.zero_()
    if not a.graph:
        P.kernel_timing(True)  # HIP events around each fine colour step, inside the timed region
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = P.cg(b, x, max_iter=a.steps, tolerance=0.0, ordering=a.ordering, graph=bool(a.graph))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    if a.graph:  # graph replays carry no per-launch events: time the launches eagerly
        x.zero_()
        P.kernel_timing(True)
        P.cg(b, x, max_iter=a.steps, tolerance=0.0, ordering=a.ordering)
    P.kernel_timing(False)
    ks_ms, ks_n = P.kernel_timing_read()
    if a.mixed_steps > 0:
        P.prepare_mixed()
        x.zero_()
        P.cg_mixed(b, x, max_iter=a.mixed_steps, tolerance=0.0, recompute=10)
        torch.cuda.synchronize()
    log("timed %d iterations in %.3f s; normr/normr0 %.3e" % (r["niters"], dt, r["normr"] / r["normr0"]))
    conv = None
    if a.ref_iters > 0:
        # main.cpp:212-224 and :266-280: the residual ref_iters natural-order
        # iterations reach, then the optimized ordering's count to reach it
        x.zero_()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rr = P.cg(b, x, max_iter=a.ref_iters, tolerance=0.0, ordering="lexicographic", graph=True)
        torch.cuda.synchronize()
        t_ref = time.perf_counter() - t1
        ref_tol = rr["normr"] / rr["normr0"]
        x.zero_()
        ro = P.cg(b, x, max_iter=10 * a.ref_iters, tolerance=ref_tol, ordering=a.ordering,
                  graph=True)
        opt_iters = max(a.ref_iters, ro["niters"])
        conv = {"ref_iterations": a.ref_iters, "ref_ordering": "lexicographic",
                "ref_normr_over_normr0": ref_tol,
                "ref_ms_per_iteration": round(t_ref / a.ref_iters * 1e3, 3),
                "opt_iterations": opt_iters, "opt_ordering": a.ordering,
                "opt_normr_over_normr0": ro["normr"] / ro["normr0"]}
        log("convergence overhead: %d %s iterations reach %.4e (%d lexicographic, %.1f ms each)"
            % (opt_iters, a.ordering, ref_tol, a.ref_iters, t_ref / a.ref_iters * 1e3))

    lv = model.grid_levels(a.nx, a.ny, a.nz)
    nrow = [l["n"] for l in lv]
    nnz = [l["nnz"] for l in lv]
    K = r["niters"]
    flops = model.hpcg_model(nrow, nnz, 1, K, K)
    gflops = flops["total"] / dt * 1e-9
    cb = model.canonical_bytes(lv)
    pre_bytes = cb["outer_spmv"] + 7 * 8 * lv[0]["n"]  # preamble: copy, SpMV, r = b - Ap, r.r
    gbs = (K * cb["iteration"] + pre_bytes) / dt * 1e-9
    # dominant kernel: the finest level's multicolour Gauss-Seidel launches
    # (two symmetric sweeps per iteration: pre- and post-smoother)
    info0 = P.info(0)
    c16 = bool(a.col16 and info0["col16"])
    ntiles0 = (lv[0]["n"] + 63) // 64
    c8 = c16 and info0["byte_tiles"] == ntiles0 and not any(
        l.strip() == "chosen col8=0" for l in (tuned or []))
    ksym = ("k_gs_wave<true,true>" if info0["wave"] else
            "k_gs_range16<true,true> + k_gs_rz16<true>" if c16 else "k_gs_range<true,true>")
    kname = ksym + (" (finest-level wavefront sweep, one direction)" if info0["wave"]
                    else " (finest-level colour steps; k_gs_rz16 runs the 8 of the last sweep "
                    "of each V-cycle and also leaves the r.z partials)" if c16
                    else " (finest-level colour step)")
    # one colour step is one (2 x ncolors)-th of a symmetric sweep; the
    # logged launches are the full colour steps (the zero-x steps of each
    # V-cycle's first forward sweep are another kernel, k_gs_zero16)
    nsteps = 2 * info0["ncolors"] if not info0["wave"] else 2
    per_launch = cb["symgs_L0"] / nsteps
    # the same sweep in the stored format: every ELL slot (padding included)
    # with 8-byte values and 1- (8-bit offsets), 2- (16-bit) or 4-byte ids
    a_st = 8 + (1 if c8 else 2 if c16 else 4)
    stored = (2 * a_st * 27 * lv[0]["n"] + 6 * 8 * lv[0]["n"]) / nsteps
    launch_ms = ks_ms / max(ks_n, 1)
    achieved = per_launch / (launch_ms * 1e-3) * 1e-9 if ks_n else None
    # ReportResults.cpp:99,351: the flops of ref_iters reference iterations
    # over the time of the opt_iters the optimized ordering needs for the
    # same residual reduction
    conv_factor = conv["ref_iterations"] / conv["opt_iterations"] if conv else 1.0
    gflops_conv = gflops * conv_factor
    rating = gflops_conv * dt / (dt + setup_s / 10.0)  # ReportResults.cpp:353, one set
    # HBM bytes per launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE
    # (profiles/), when this run is the profiled workload and kernel
    traffic = traffic_note = None
    try:
        pm = json.load(open(os.path.join(ROOT, "profiles", "pmc_traffic.json")))
        k0 = ksym.split(" + ")[0]
        if (pm["workload"] == "%dx%dx%d fp64 %s" % (a.nx, a.ny, a.nz, a.ordering)
                and k0 in pm["kernels"]):
            traffic = pm["kernels"][k0]["traffic_bytes"]
            traffic_note = pm["kernels"][k0].get("note")
    except (OSError, ValueError, KeyError):
        traffic = None

    out = {
        "metric": ("HPCG GFLOP/s with convergence overhead" if conv else "HPCG GFLOP/s raw")
                  + " (FP64 PCG, 4-level MG, %s SYMGS)" % a.ordering,
        "value": round(gflops_conv, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": a.warmup,
        "ms_per_step": round(dt / K * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: HPCG 27-point problem generated on the device, b = A*1, x0 = 0",
        "config": {"workload": "HPCG PCG %dx%dx%d FP64, 4 MG levels, %s Gauss-Seidel"
                               % (a.nx, a.ny, a.nz, a.ordering),
                   "nx": a.nx, "ny": a.ny, "nz": a.nz, "mg_levels": 4,
                   "ordering": a.ordering, "sweep_order": a.sweep_order,
                   "rows": lv[0]["n"], "nnz": lv[0]["nnz"]},
        "value_is": ("HPCG's 'Total with convergence overhead' GFLOP/s (ReportResults.cpp:351): "
                     "the raw GFLOP/s x ref_iterations / opt_iterations" if conv else
                     "raw GFLOP/s (ReportResults.cpp:350)"),
        "raw_gflops": round(gflops, 3),
        **({"convergence_overhead": conv} if conv else {}),
        "iteration_gbs_canonical": round(gbs, 1),
        "iteration_hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
        "rating_with_setup_gflops": round(rating, 3),
        "setup_s": round(setup_s, 3),
        "normr_over_normr0": r["normr"] / r["normr0"],
        "roofline": {
            "kernel": kname,
            "bound": "hbm",
            "achieved": round(achieved, 1) if achieved else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": round(traffic) if traffic else None,
            **({"traffic_note": traffic_note} if traffic and traffic_note else {}),
            "bytes_per_launch": per_launch,
            "stored_format_bytes_per_launch": stored,
            "column_ids": "8-bit" if c8 else "16-bit" if c16 else "32-bit",
            "mean_launch_ms": round(launch_ms, 5),
            "launch_timing": "eager pass after the timed region" if a.graph else "timed region",
            "launches": ks_n,
        },
    }
    if tuned is not None:
        out["autotune"] = [l for l in tuned if l.startswith("chosen")]
    if a.tol_compare > 0:
        out["convergence"] = converge(P, b, lv, a, log)
    if a.block:
        g = a.block_grid
        if (g, g, g) == (a.nx, a.ny, a.nz):
            out["block_pcg"] = [block_solve(P, s, a, log) for s in a.block]
        else:
            log("block PCG problem %d^3" % g)
            PB = hpcg_amd.Problem.generate(g, g, g, nlevels=4)
            out["block_pcg"] = [dict(block_solve(PB, s, a, log), grid=[g, g, g])
                                for s in a.block]
            PB.close()
    if a.fleet:
        out["fleet"] = fleet_solve(a.fleet, a, log)
    if not a.no_cpu_baseline:
        log("cpu baseline")
        try:
            out["cpu_baseline"] = cpu_baseline(a.cpu_grid, a.cpu_iters, lv[0]["n"])
        except Exception as e:  # reported, never fatal to the GPU number
            out["cpu_baseline"] = {"value": None, "error": str(e)[:200]}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
