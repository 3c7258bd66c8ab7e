"""Benchmark: converged eigenpairs/s of mixed-precision filtered OFRR on one MI355X.

Workload (one step = one complete solve from the seeded start block):
  3D 7-point Dirichlet Laplacian on a 128^3 grid (n = 2,097,152, nnz = 14,581,760),
  block k = 64, basis dimension 192 (X plus two filtered correction blocks),
  the 50 largest eigenpairs to ||A x - theta x|| / (||A|| ||x||) <= 1e-8.
  Growth blocks stored in FP16 with FP32 accumulation (operator values in FP16,
  exact for this stencil); Ritz block, Gram products V^T V / V^T A V, truncated
  Rayleigh-Ritz and residuals in FP64.  Start block N(0,1), Philox seed 0.

Reported beside it, in the same JSON line:
  * the plain FP64 GPU path (same algorithm, FP64 growth blocks) and the ratio;
  * p50 / p99 solve latency over the timed steps;
  * roofline of the dominant kernel (the Chebyshev filter step, HBM-bound), with
    algorithmic bytes per launch / its average duration from HIP events;
  * cpu_baseline: the FP64 CPU restatement (oracle) of the same algorithm timed
    on a bounded sample on this host's cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "converged eigenpairs/sec at resid≤1e-8 (mixed vs FP64 GPU); p50/p99 latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def pct(xs, p):
    import numpy as np
    return float(np.percentile(np.asarray(xs), p)) if xs else None


def laplacian_top(dim, N, count):
    """The `count` largest eigenvalues of the Dirichlet stencil Laplacian on an
    N^dim grid, from the closed form sum_d 4 sin^2(p_d pi / (2 (N + 1)))."""
    import numpy as np
    l1 = 4.0 * np.sin(np.arange(1, N + 1) * np.pi / (2.0 * (N + 1))) ** 2
    L = l1
    for _ in range(dim - 1):
        L = (L[:, None] + l1[None, :]).ravel()
    return np.sort(L)[::-1][:count]


def filter_bytes(n, nnz, b_val, b_basis, launches, col_steps, first_cols):
    """Algorithmic HBM bytes of the Chebyshev-step launches of a solve: the
    operator (values + int32 columns + int64 row pointers) once per launch; per
    launch of width w (the columns filtered: k, or fewer once pairs are locked)
    the gathered block, the previous block, the residual block and the output
    block, n w elements each.  The first step of each sweep gathers the residual
    block itself and has no previous block (2 blocks).  col_steps is the sum of
    the widths over the launches, first_cols the sum over the first steps."""
    op = nnz * (b_val + 4) + 8 * (n + 1)
    return launches * op + n * b_basis * (4 * col_steps - 2 * first_cols)


PROFILE = os.path.join(ROOT, "profiles", "cheb_step_f16_laplacian128_k64.md")


def measured_traffic(args):
    """HBM bytes per launch of the filter kernel from the committed rocprofv3
    counter summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated); it covers the
    128^3, k = 64, FP16 workload only."""
    if args.grid != 128 or args.k != 64 or args.basis != "f16":
        return None
    try:
        import re
        m = re.search(r"traffic_bytes_per_launch=(\d+)", open(PROFILE).read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def wide_of(args):
    """The wide format of the mixed solve: FP32 for FP16 growth blocks when
    --wide-tol > 0 (None: every sweep in the growth format)."""
    import torch
    return torch.float32 if (args.wide_tol > 0 and args.basis == "f16") else None


def run_solver(api, ctx, A, X0, args, basis, label, wide=None):
    import torch
    st, th, X, res, stats = api.filtered_eig(
        ctx, A, X0, args.nwant, nd=args.nd, degree=args.degree, maxiter=args.maxiter,
        basis_dtype=basis, tol=args.tol, tau=args.tau, wide_dtype=wide,
        wide_tol=args.wide_tol if wide is not None else 0.0)
    torch.cuda.synchronize()
    return st, th, X, res, stats


def host_info():
    """The CPU the baseline runs on: lscpu's model, sockets, cores per socket and
    threads per core, and the CPUs this process may use (its affinity set)."""
    info = {"nproc_affinity": len(os.sched_getaffinity(0)),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
        for line in out.splitlines():
            name, _, val = line.partition(":")
            if name.strip() in keys:
                info[keys[name.strip()]] = val.strip()
    except (OSError, ValueError) as e:
        info["lscpu_error"] = repr(e)
    return info


def baseline_threads(info):
    """Physical cores among the CPUs this process may use: the affinity set
    divided by the threads per core.  A host that sets OMP_NUM_THREADS for the
    GPU's share of its cores caps it there."""
    try:
        tpc = max(1, int(info.get("threads_per_core", "1")))
    except ValueError:
        tpc = 1
    cores = max(1, info["nproc_affinity"] // tpc)
    env = info.get("omp_num_threads_env")
    if env and env.isdigit() and int(env) > 0:
        cores = min(cores, int(env))
    return cores


def cpu_baseline(args, gpu_iters, gpu_col_steps, A_dims, reps=3):
    """FP64 oracle on the same operator and start block: one outer iteration at
    two filter degrees (bounded sample, each timed `reps` times, the median
    taken), extrapolated to the GPU solve: the per-iteration remainder times its
    iterations plus the cost of a k-column filter step times its filtered
    columns / k (soft locking narrows the later sweeps)."""
    import numpy as np

    info = host_info()
    cores = baseline_threads(info)
    os.environ["OMP_NUM_THREADS"] = str(cores)
    from oracle import oracle as O
    dim, N = A_dims
    t0 = time.time()
    Ah = O.stencil(dim, N)
    X0 = O.randn(args.seed, Ah.n * args.k).reshape(Ah.n, args.k)  # row-major draws
    X0 = np.asfortranarray(X0)
    t_build = time.time() - t0
    samples = {2: [], 4: []}
    for r in range(reps):
        for deg in (2, 4):
            t = time.time()
            O.filtered_ofrr(Ah, X0, args.nwant, nd=min(args.nd, 1 if deg < 4 else args.nd),
                            degree=deg, maxiter=1, tol=1e-30, tau=args.tau)
            samples[deg].append(time.time() - t)
            log(f"cpu baseline rep {r} degree {deg}: {samples[deg][-1]:.1f} s")
    times = {d: float(np.median(v)) for d, v in samples.items()}
    step = max(0.0, (times[4] - times[2]) / 2.0)
    fixed = max(0.0, times[2] - step)          # initial RR + 1 iteration at degree 2 - 1 step
    col_steps = gpu_col_steps / float(args.k)
    solve_s = fixed * max(1, gpu_iters) + step * col_steps
    return {
        "value": args.nwant / solve_s,
        "unit": "eigenpairs/s",
        "cores": cores,
        "kind": "port",
        "host": info,
        "samples_s": {str(d): v for d, v in samples.items()},
        "sample": (f"FP64 oracle (oracle/ofrr_oracle.c, OpenMP, {cores} threads = the physical "
                   f"cores of this process's CPU set, capped by OMP_NUM_THREADS; same algorithm "
                   f"and start block, Jacobi Rayleigh-Ritz) on the full {N}^3 operator: one outer "
                   f"iteration at filter degree 2 and 4, median of {reps} runs each "
                   f"({times[2]:.1f} s, {times[4]:.1f} s) -> {step:.2f} s per filter step, "
                   f"{fixed:.2f} s per-iteration remainder; solve time = remainder x the GPU's "
                   f"{gpu_iters} iterations + step x its {col_steps:.1f} k-column-equivalent filter "
                   f"steps = {solve_s:.0f} s; operator build {t_build:.1f} s excluded"),
    }


def arnoldi_small(api, ctx, reps=5):
    """The plain FP64 OFRR case on the 2-D 64 x 64 Dirichlet Laplacian (n = 4096):
    Arnoldi with the Hessenberg-process basis, basis dimension 60, the 20 largest
    eigenvalues, N(0,1) start vector from seed 0, three restart cycles
    (maxiter 181).  The GPU path (ofrr_arnoldi_hess_eig) and the CPU oracle run
    the same operations; each is timed `reps` times (median), and their Ritz
    values and residuals are reported side by side."""
    import numpy as np
    import torch

    from oracle import oracle as O
    N, k, kdim, maxiter = 64, 20, 60, 1 + 60 * 3
    Ah = O.stencil(2, N)
    A = api.CSRMatrix.stencil(ctx, 2, N)
    v0 = torch.empty(Ah.n, dtype=torch.float64, device="cuda:0")
    api.fill_randn(ctx, v0, seed=0)
    v0h = v0.cpu().numpy()
    exact = laplacian_top(2, N, k)
    gpu_ms, cpu_ms = [], []
    for _ in range(reps + 1):  # the first GPU run is a warmup
        torch.cuda.synchronize()
        t = time.perf_counter()
        th, X, built = api.arnoldi_hess_eig(ctx, A, k, kdim, maxiter, v0)
        torch.cuda.synchronize()
        gpu_ms.append((time.perf_counter() - t) * 1e3)
    gpu_ms = gpu_ms[1:]
    for _ in range(reps):
        t = time.perf_counter()
        th_o, X_o, built_o = O.arnoldi_hess_eig(Ah, k, kdim, maxiter, v0h)
        cpu_ms.append((time.perf_counter() - t) * 1e3)
    res = O.residuals(Ah, X.cpu().numpy(), th, exact[0])
    res_o = O.residuals(Ah, X_o, th_o, exact[0])
    return {"workload": "plain FP64 OFRR (Hessenberg-process Arnoldi), 2-D 5-point Dirichlet Laplacian "
                        "64 x 64 (n = 4096), basis dimension 60, 20 largest, N(0,1) start from seed 0, "
                        "maxiter 181",
            "gpu_ms_median": float(np.median(gpu_ms)), "cpu_ms_median": float(np.median(cpu_ms)),
            "cpu_threads": int(os.environ.get("OMP_NUM_THREADS", "1")), "reps": reps,
            "gpu_converged": int(np.sum(res <= 1e-8)), "cpu_converged": int(np.sum(res_o <= 1e-8)),
            "gpu_max_residual": float(np.max(res)), "cpu_max_residual": float(np.max(res_o)),
            "max_ritz_rel_diff_gpu_vs_cpu": float(np.max(np.abs(th - th_o) / np.abs(th_o))),
            "max_eig_rel_err_vs_closed_form": float(np.max(np.abs(th - exact) / exact))}


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--grid", type=int, default=128)
    p.add_argument("--k", type=int, default=64)
    p.add_argument("--nd", type=int, default=2)
    p.add_argument("--nwant", type=int, default=50)
    p.add_argument("--degree", type=int, default=110)
    p.add_argument("--maxiter", type=int, default=200)
    p.add_argument("--tol", type=float, default=1e-8)
    p.add_argument("--tau", type=float, default=1e-12)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--fp64-steps", type=int, default=1)
    p.add_argument("--wide-tol", type=float, default=0.0,
                   help="FP16 basis: the sweeps after the first that start with a wanted residual "
                        "above this build their blocks in FP32 (0, the default: every sweep in FP16)")
    p.add_argument("--basis", choices=["f16", "bf16", "fp8"], default="f16",
                   help="growth-block format of the mixed path (the workload names FP16)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-fp64", action="store_true")
    p.add_argument("--workload", choices=["eig", "svd", "serve", "qr", "graphs", "syev"], default="eig",
                   help="eig: the 128^3 Laplacian eigensolve (default); svd: truncated SVD of "
                        "the dense 262,144 x 4,096 low-rank-plus-noise matrix; serve: a stream "
                        "of eigen requests with Poisson arrivals through the serving loop")
    p.add_argument("--serve-requests", type=int, default=512)
    p.add_argument("--serve-rate", type=float, default=4.0,
                   help="Poisson arrival rate of the serving stream (requests/s)")
    p.add_argument("--serve-slots", type=int, default=4)
    p.add_argument("--serve-seed", type=int, default=0)
    p.add_argument("--serve-graphs", choices=["on", "off"], default="off",
                   help="serve workload: filter sweeps as captured HIP graphs")
    p.add_argument("--serve-rr-group-us", type=int, default=0,
                   help="serve workload: grouped Rayleigh-Ritz (0: the jobs that meet, w > 0: "
                        "also wait up to w microseconds for every running solve, -1: off)")
    p.add_argument("--serve-big", action="store_true",
                   help="also submit the large FP8 solve (3-D variable-coefficient diffusion on "
                        "a 256^3 grid, 100 largest pairs, FP8 E4M3 growth blocks) a quarter of "
                        "the way into the stream")
    p.add_argument("--qr-sizes", default="25000x200,50000x200,50000x400",
                   help="qr workload: m x n shapes (CPP/examples/ex00.cu:160-164)")
    p.add_argument("--qr-dtypes", default="f64,f32,f16")
    p.add_argument("--svd-m", type=int, default=262144)
    p.add_argument("--svd-n", type=int, default=4096)
    p.add_argument("--svd-k", type=int, default=96,
                   help="svd workload: Ritz block width (64 wanted + guard columns); the "
                        "basis is 2 k (one correction block)")
    p.add_argument("--copy-probe", action="store_true",
                   help="run a plain 1 GiB device copy first (calibrates HBM byte counters)")
    args = p.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build runs on one GPU only: the OFRR path is measured "
                                   "at N=1; --gpus above 1 is not supported", "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    from mixedprecisionofrr_amd import api

    if args.workload == "svd":
        return svd_main(args)
    if args.workload == "qr":
        return qr_main(args)
    if args.workload == "graphs":
        return graphs_main(args)
    if args.workload == "syev":
        return syev_main(args)
    if args.workload == "serve":
        r = serve_stream(api, args.serve_requests, args.serve_rate, args.serve_slots,
                         args.serve_seed, big=args.serve_big, graphs=args.serve_graphs == "on",
                         rr_group_us=args.serve_rr_group_us)
        out = {"metric": "serving: converged eigenpairs/s and p50/p99 request latency",
               "value": r["pairs_per_s"], "unit": "eigenpairs/s", "n_gpus": 1,
               "steps": args.serve_requests, "warmup": 1, "ms_per_step": r["wall_s"] * 1e3 / max(1, args.serve_requests),
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "f16 storage / f32 accumulate (growth), f64 (Ritz block, Gram, RR, residual)",
               "data": "synthetic", "config": {"workload": r["workload"]}, "serving": r}
        print(json.dumps(out), flush=True)
        return
    dev = "cuda:0"
    ctx = api.Context(0)
    if args.copy_probe:
        # a plain copy of known size: 1 GiB read + 1 GiB written per launch
        src = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    N = args.grid
    log(f"building 3D 7-point Laplacian {N}^3 on the GPU")
    A = api.CSRMatrix.stencil(ctx, 3, N).with_low_precision(torch.float16)
    basis = {"f16": torch.float16, "bf16": torch.bfloat16, "fp8": torch.float8_e4m3fn}[args.basis]
    X0 = torch.empty(A.n, args.k, dtype=torch.float64, device=dev)
    api.fill_randn(ctx, X0, seed=args.seed)
    torch.cuda.synchronize()

    # ---- mixed path: warmup + timed steps
    wide = wide_of(args)
    for w in range(args.warmup):
        st, th, X, res, stats = run_solver(api, ctx, A, X0, args, basis, "mixed", wide)
        log(f"warmup {w}: status {st} iters {stats['iterations']} {stats['t_total_ms']:.1f} ms")
    lat, pairs, last = [], 0, None
    filt_ms = filt_steps = 0.0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        ts = time.perf_counter()
        st, th, X, res, stats = run_solver(api, ctx, A, X0, args, basis, "mixed", wide)
        lat.append((time.perf_counter() - ts) * 1e3)
        pairs += int(np.sum(res[:args.nwant] <= args.tol))
        filt_ms += stats["t_filter_ms"]
        filt_steps += stats["filter_steps"]
        last = (st, th, X, res, stats)
        log(f"step {s}: status {st} iters {stats['iterations']} {lat[-1]:.1f} ms "
            f"converged {stats['converged']}/{args.nwant}")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    st, th, X, res, stats = last
    value = pairs / elapsed
    exact = laplacian_top(3, N, args.nwant)
    api_exact = float(np.max(np.abs(th[:args.nwant] - exact) / exact))
    # independent residual check of the returned block (FP64, on the GPU)
    res_chk = api.residuals(ctx, A, X, th, float(exact[0]))

    # ---- every sweep in the growth format (untimed in the value): its own
    # figures, and the roofline of the growth-format step kernel
    plain = None
    rstats, rfilt_ms, rfilt_steps = stats, filt_ms, filt_steps
    if wide is not None:
        ts = time.perf_counter()
        stp, thp, Xp, resp, sp = run_solver(api, ctx, A, X0, args, basis, "plain")
        msp = (time.perf_counter() - ts) * 1e3
        log(f"{args.basis} every sweep: status {stp} iters {sp['iterations']} {msp:.1f} ms")
        plain = {"value": int(np.sum(resp[:args.nwant] <= args.tol)) / (msp * 1e-3),
                 "unit": "eigenpairs/s", "ms_per_solve": msp, "iterations": sp["iterations"],
                 "status": stp, "max_residual": float(np.max(resp[:args.nwant])),
                 "t_filter_ms": sp["t_filter_ms"], "filter_steps": sp["filter_steps"]}
        rstats, rfilt_ms, rfilt_steps = sp, sp["t_filter_ms"], sp["filter_steps"]

    # ---- roofline of the dominant kernel (Chebyshev step, growth format)
    n, nnz, k = A.n, A.nnz, args.k
    per_solve_steps = rstats["filter_steps"]
    b_basis = 1 if args.basis == "fp8" else 2
    bytes_solve = filter_bytes(n, nnz, 2, b_basis, per_solve_steps, rstats["filter_col_steps"],
                               rstats["filter_first_cols"])
    avg_launch_ms = rfilt_ms / max(1, rfilt_steps)
    bytes_per_launch = bytes_solve / max(1, per_solve_steps)
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9

    # ---- FP64 GPU path (same algorithm)
    fp64 = None
    if not args.no_fp64:
        lat64, pairs64 = [], 0
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for s in range(args.fp64_steps):
            ts = time.perf_counter()
            st64, th64, X64, res64, stats64 = run_solver(api, ctx, A, X0, args, torch.float64, "fp64")
            lat64.append((time.perf_counter() - ts) * 1e3)
            pairs64 += int(np.sum(res64[:args.nwant] <= args.tol))
            log(f"fp64 step {s}: status {st64} iters {stats64['iterations']} {lat64[-1]:.1f} ms")
        el64 = time.perf_counter() - t1
        fp64 = {"value": pairs64 / el64, "unit": "eigenpairs/s", "ms_per_solve": el64 * 1e3 / args.fp64_steps,
                "iterations": stats64["iterations"], "status": st64,
                "max_residual": float(np.max(res64[:args.nwant])),
                "t_filter_ms": stats64["t_filter_ms"], "t_gram_ms": stats64["t_gram_ms"],
                "t_aux_ms": stats64["t_aux_ms"],
                "t_rr_ms": stats64["t_rr_ms"], "t_update_ms": stats64["t_update_ms"],
                "t_resid_ms": stats64["t_resid_ms"], "t_zspmm_ms": stats64["t_zspmm_ms"]}
    # ---- CPU baseline
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(args, stats["iterations"], stats["filter_col_steps"], (3, N))
        except Exception as e:  # reported, never silently replaced
            cpu = {"value": None, "unit": "eigenpairs/s", "cores": 0, "kind": "port",
                   "sample": f"failed: {e!r}"}
    # ---- the plain FP64 Arnoldi case (64 x 64 Laplacian): GPU and CPU timings
    small = None
    if not args.no_cpu:
        try:
            small = arnoldi_small(api, ctx)
        except Exception as e:
            small = {"error": repr(e)}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "eigenpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": (f"{args.basis} storage" +
                  (f" (f32 for the sweeps after the first that start above residual {args.wide_tol:g}:"
                   f" {stats['wide_sweeps']} of {stats['iterations']})" if wide is not None else "") +
                  " / f32 accumulate (growth), f64 (Ritz block, Gram, RR, residual)"),
        "data": "synthetic",
        "config": {
            "workload": (f"mixed-precision filtered block subspace iteration + OFRR, 3D 7-point "
                         f"Dirichlet Laplacian {N}^3, block k={k}, basis dim {k * (1 + args.nd)}, "
                         f"{args.nwant} largest eigenpairs to residual {args.tol:g}, "
                         f"{args.basis.upper()} basis"),
            "n": n, "nnz": nnz, "k": k, "basis_dim": k * (1 + args.nd), "nwant": args.nwant,
            "filter_degree": args.degree, "tol": args.tol, "seed": args.seed,
            "wide_sweeps": ("f32 after the first sweep while the wanted residual > "
                            f"{args.wide_tol:g}") if wide is not None else None,
        },
        "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99), "arrival": "back-to-back solves"},
        "speedup_vs_fp64_gpu": (value / fp64["value"]) if fp64 and fp64["value"] else None,
        "fp64_gpu": fp64,
        "every_sweep_in_growth_format": plain,
        "solve": {"status": st, "iterations": stats["iterations"], "spmm_calls": stats["spmm_calls"],
                  "converged": stats["converged"], "max_residual": float(np.max(res[:args.nwant])),
                  "max_residual_recheck": float(np.max(res_chk[:args.nwant])),
                  "max_eig_rel_err_vs_closed_form": api_exact,
                  "t_total_ms": stats["t_total_ms"], "t_filter_ms": stats["t_filter_ms"],
                  "t_zspmm_ms": stats["t_zspmm_ms"], "t_gram_ms": stats["t_gram_ms"],
                  "t_aux_ms": stats["t_aux_ms"],
                  "t_rr_ms": stats["t_rr_ms"], "t_update_ms": stats["t_update_ms"],
                  "t_resid_ms": stats["t_resid_ms"], "rank_min": stats["rank_min"],
                  "wide_sweeps": stats["wide_sweeps"], "filter_steps": stats["filter_steps"]},
        "roofline": {"kernel": f"cheb_step_kernel (FP16 operator, {args.basis.upper()} blocks: CSR SpMM "
                               f"+ filter axpys)",
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args),
                     "traffic_source": "profiles/cheb_step_f16_laplacian128_k64.md",
                     "bytes_per_launch": bytes_per_launch, "avg_launch_us": avg_launch_ms * 1e3},
        "cpu_baseline": cpu,
        "arnoldi_64x64": small,
    }
    print(json.dumps(out), flush=True)


def serve_requests(api, count, seed):
    """The serving stream: 2-D 5-point or 3-D 7-point diffusion operators with
    face coefficients 10^(2u-1) (log-uniform in [0.1, 10], per-request seed),
    n log-uniform in [1e4, 4e6] rounded to a grid side, nwant uniform over
    {8, 16, 32, 64}, FP16 growth blocks, tolerance 1e-8."""
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    reqs, meta = [], []
    for i in range(count):
        dim = int(rng.integers(2, 4))
        n = float(np.exp(rng.uniform(np.log(1e4), np.log(4e6))))
        N = max(4, int(round(n ** (1.0 / dim))))
        nwant = int(rng.choice([8, 16, 32, 64]))
        reqs.append(api.request(dim, N, nwant, coef_seed=1000 + i, start_seed=i,
                                basis_dtype=torch.float16, tol=1e-8))
        meta.append((dim, N, nwant))
    return reqs, meta


def serve_stream(api, count, rate, slots, seed, big=False, graphs=False, rr_group_us=0):
    """Submit `count` requests at seeded Poisson arrival times through a
    `slots`-slot server; latency = completion - arrival on the server's clock.
    big: the large FP8 solve joins the stream after a quarter of the arrivals.
    graphs: filter sweeps as captured HIP graphs; rr_group_us: the grouped
    Rayleigh-Ritz window (0: the server's default, -1: off)."""
    import numpy as np
    import torch
    reqs, meta = serve_requests(api, count, seed)
    # the large FP8 solve as configured: basis dimension 300 = a 128-column Ritz
    # block (100 wanted + 28 guard columns), the 128-column FP8 correction block
    # and the trailing 44 columns of the FP8 mid-filter block; FP64 Gram,
    # truncated Rayleigh-Ritz and residuals; the precision ladder finishes the
    # growth blocks in FP16 once the wanted residuals fall below 1e-4 or stall
    big_req = api.request(3, 256, 100, coef_seed=5, start_seed=7, basis_dtype=torch.float8_e4m3fn,
                          tol=1e-8, k=128, nd=2, mid_cols=44, refine_dtype=torch.float16,
                          refine_tol=1e-4) if big else None
    big_id = big_out = None
    arr = np.cumsum(np.random.default_rng(seed + 1).exponential(1.0 / rate, size=count))
    arr -= arr[0]
    with api.Server(0, slots=slots, graphs=graphs, rr_group_us=rr_group_us) as srv:
        # first use of every kernel on the box, outside the stream
        srv.wait(srv.submit(api.request(3, 24, 8, coef_seed=1, start_seed=1)))
        log(f"serve: {count} requests, Poisson {rate:g}/s, {slots} slots")
        t0 = time.perf_counter()
        ids = []
        for i, r in enumerate(reqs):
            delay = arr[i] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit(r))
            if big_req is not None and i == count // 4:
                big_id = srv.submit(big_req)
                log("serve: large FP8 solve submitted")
            if (i + 1) % 64 == 0:
                st = srv.stats()
                log(f"serve: submitted {i + 1}, completed {st['completed'] - 1}, "
                    f"running {st['running']}, queued {st['queued']}")
        out = []
        for j, i in enumerate(ids):
            out.append(srv.wait(i))
            if (j + 1) % 64 == 0:
                log(f"serve: collected {j + 1}")
        wall = time.perf_counter() - t0
        if big_id is not None:
            b, t_last = None, time.perf_counter()
            while b is None:
                b = srv.poll(big_id)
                if b is None:
                    time.sleep(0.5)
                    if time.perf_counter() - t_last > 30:
                        t_last = time.perf_counter()
                        log("serve: large FP8 solve still running")
            log(f"serve: large FP8 solve status {b['status']} iterations {b['iterations']} "
                f"{(b['t_done'] - b['t_submit']):.1f} s")
            big_out = {"workload": "3-D 7-point variable-coefficient diffusion, 256^3 grid "
                                   "(n = 16,777,216), 100 largest pairs, basis dimension 300 "
                                   "(128-column FP64 Ritz block + 128-column FP8 E4M3 correction "
                                   "block + 44 columns of the FP8 mid-filter block, per-column "
                                   "power-of-two scales, FP32 accumulation; FP16 ladder below "
                                   "1e-4), FP64 Gram / truncated Rayleigh-Ritz / residuals, tol "
                                   "1e-8, submitted during the serving stream",
                       "basis_dim": 300, "k": b["k"], "nd": b["nd"],
                       "status": b["status"], "iterations": b["iterations"],
                       "filter_steps": b["filter_steps"], "converged": b["converged"],
                       "latency_s": b["t_done"] - b["t_submit"],
                       "queue_wait_s": b["t_admit"] - b["t_submit"],
                       "max_residual": float(np.max(b["res"])),
                       "theta_top": float(b["theta"][0]), "theta_100": float(b["theta"][99]),
                       "footprint_bytes": b["footprint"]}
        stats = srv.stats()
    lat = np.array([o["t_done"] - o["t_submit"] for o in out]) * 1e3
    wait = np.array([o["t_admit"] - o["t_submit"] for o in out]) * 1e3
    pairs = int(sum(int(np.sum(o["res"] <= 1e-8)) for o in out))
    ok = sum(1 for o in out if o["status"] == 0)
    ns = np.array([o["n"] for o in out])
    return {
        "workload": (f"serving stream: {count} eigen requests, Poisson arrivals at {rate:g}/s "
                     f"(seed {seed}), 2-D 5-point / 3-D 7-point variable-coefficient diffusion "
                     f"(face coefficients log-uniform in [0.1, 10]), n log-uniform in [1e4, 4e6], "
                     f"nwant in {{8,16,32,64}}, tol 1e-8, FP16 growth blocks, {slots} slots, "
                     f"HBM-footprint admission, filter sweeps as HIP graphs "
                     f"{'on' if graphs else 'off'}, grouped Rayleigh-Ritz "
                     f"{'off' if rr_group_us < 0 else 'on'}"),
        "arrival_rate_per_s": rate, "requests": count, "slots": slots,
        "graphs": bool(graphs), "rr_group_us": rr_group_us,
        "rr_groups": stats["rr_groups"], "rr_group_jobs": stats["rr_group_jobs"],
        "rr_group_max": stats["rr_group_max"],
        "completed_ok": ok, "converged_pairs": pairs, "wall_s": wall,
        "pairs_per_s": pairs / wall, "requests_per_s": count / wall,
        "latency_ms": {"p50": pct(list(lat), 50), "p99": pct(list(lat), 99),
                       "mean": float(lat.mean()), "max": float(lat.max())},
        "queue_wait_ms": {"p50": pct(list(wait), 50), "p99": pct(list(wait), 99)},
        "n_range": [int(ns.min()), int(ns.max())],
        "running_peak": stats["running_peak"], "bytes_peak": stats["bytes_peak"],
        "budget": stats["budget"], "large_fp8_solve": big_out,
    }


def svd_main(args):
    """Truncated SVD: A = U diag(sigma) W^T + 1e-6 N(0,1), M x N FP32, U and W from
    QR of seeded Gaussians (seeds 1, 2; noise seed 3), sigma_i = exp(-i/64); the
    64 largest singular triplets by filtered OFRR on A^T A (BF16 growth products,
    FP64 Gram / Rayleigh-Ritz / residual).  One step = one complete solve."""
    import numpy as np
    import torch

    from mixedprecisionofrr_amd import api

    dev = "cuda:0"
    ctx = api.Context(0)
    M, N, nwant = args.svd_m, args.svd_n, 64
    k = args.svd_k
    log(f"building the {M} x {N} test matrix on the GPU")
    with torch.no_grad():
        g = torch.Generator(device=dev)
        g.manual_seed(1)
        U = torch.linalg.qr(torch.randn(M, N, device=dev, generator=g))[0]
        g.manual_seed(2)
        W = torch.linalg.qr(torch.randn(N, N, device=dev, generator=g))[0]
        sig = torch.exp(-torch.arange(N, device=dev, dtype=torch.float32) / 64.0)
        U.mul_(sig[None, :])
        g.manual_seed(3)
        noise = torch.randn(N, M, device=dev, generator=g)  # column-major A^T storage
        At = torch.addmm(noise, W, U.t(), beta=1e-6)      # A^T = W diag(s) U^T + noise^T
        del U, W, noise
    A = At.t()                                             # M x N, column-major view
    torch.cuda.synchronize()
    V0 = torch.empty(N, k, dtype=torch.float64, device=dev)
    api.fill_randn(ctx, V0, seed=args.seed)

    def solve(bdt):
        return api.svd_filtered(ctx, A, V0, nwant, nd=1, degree=args.degree, maxiter=args.maxiter,
                                basis_dtype=bdt, tol=args.tol, want_u=True)

    for w in range(args.warmup):
        solve(torch.bfloat16)
    torch.cuda.synchronize()
    lat, pairs = [], 0
    t0 = time.perf_counter()
    for s in range(args.steps):
        ts = time.perf_counter()
        st, sv, Uo, Vo, res, stats = solve(torch.bfloat16)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - ts) * 1e3)
        pairs += int(np.sum(res[:nwant] <= args.tol))
        log(f"svd step {s}: status {st} iters {stats['iterations']} {lat[-1]:.1f} ms")
    elapsed = time.perf_counter() - t0
    value = pairs / elapsed
    # FP64-basis path of the same algorithm
    fp64 = None
    if not args.no_fp64:
        ts = time.perf_counter()
        st64, sv64, _, _, res64, stats64 = solve(torch.float64)
        torch.cuda.synchronize()
        t64 = time.perf_counter() - ts
        fp64 = {"value": int(np.sum(res64[:nwant] <= args.tol)) / t64, "ms_per_solve": t64 * 1e3,
                "iterations": stats64["iterations"], "status": st64}
    exact = np.exp(-np.arange(nwant) / 64.0)
    out = {
        "metric": "converged singular triplets/sec at resid<=1e-8 (BF16 growth vs FP64 GPU)",
        "value": value, "unit": "triplets/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 growth products (fp32 accumulate), fp64 Gram/RR/residual",
        "data": "synthetic",
        "config": {"workload": f"truncated SVD of a dense {M} x {N} low-rank-plus-noise FP32 matrix, "
                               f"top {nwant} triplets, block {k}, basis {2 * k}",
                   "M": M, "N": N, "k": k, "nwant": nwant, "filter_degree": args.degree},
        "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99)},
        "solve": {"status": st, "iterations": stats["iterations"], "filter_steps": stats["filter_steps"],
                  "max_residual": float(np.max(res[:nwant])),
                  "max_sigma_rel_err_vs_construction": float(np.max(np.abs(sv[:nwant] - exact) / exact)),
                  "t_filter_ms": stats["t_filter_ms"], "t_rr_ms": stats["t_rr_ms"],
                  "t_resid_ms": stats["t_resid_ms"], "t_gram_ms": stats["t_gram_ms"]},
        "fp64_gpu": fp64,
    }
    print(json.dumps(out), flush=True)


def graphs_main(args):
    """Filter sweeps as captured HIP graphs against plain launches: the same
    filtered solve (variable-coefficient diffusion, FP16 blocks, degree 120,
    soft locking) with and without opts.graph, twice each (the second graph
    solve replays the context's graphs without recording); device time per
    Chebyshev step from the solver's events.  Results are bitwise equal."""
    import numpy as np
    import torch

    from mixedprecisionofrr_amd import api
    ctx = api.Context(0)
    rows = []
    for dim, N, k, nwant in [(2, 100, 16, 8), (2, 300, 32, 16), (3, 64, 64, 32)]:
        A = api.CSRMatrix.stencil(ctx, dim, N, seed=3).with_low_precision(torch.float32)
        X0 = torch.empty(A.n, k, dtype=torch.float64, device="cuda:0")
        api.fill_randn(ctx, X0, seed=1)
        ref = None
        for graph in (False, True, False, True):
            torch.cuda.synchronize()
            t = time.perf_counter()
            st, th, X, res, s = api.filtered_eig(ctx, A, X0, nwant, nd=2, degree=120, maxiter=200,
                                                 basis_dtype=torch.float16, tol=1e-8, graph=graph)
            torch.cuda.synchronize()
            wall = (time.perf_counter() - t) * 1e3
            if ref is None:
                ref = th
            rows.append({"n": A.n, "k": k, "graph": graph, "status": st, "iterations": s["iterations"],
                         "wall_ms": wall, "t_filter_ms": s["t_filter_ms"], "filter_steps": s["filter_steps"],
                         "us_per_step": 1e3 * s["t_filter_ms"] / max(1, s["filter_steps"]),
                         "graph_replays": s["graph_replays"], "graph_captures": s["graph_captures"],
                         "same_as_plain": bool(np.array_equal(th, ref))})
            log(json.dumps(rows[-1]))
    print(json.dumps({"metric": "filter sweep device time per Chebyshev step, graph replay vs plain launches",
                      "unit": "us", "rows": rows}), flush=True)


def syev_main(args):
    """Device eigensolver (the Rayleigh-Ritz Jacobi) by size: a graded symmetric
    matrix per m (eigenvalues log-spaced over 1e-9 .. 60 with seeded random
    orthogonal eigenvectors, scaled to unit diagonal like the solver's Gram
    matrix), one ofrr_syev_batched call timed (best of 3 after a warmup).  The
    sweep count is the oracle's, whose Jacobi rotates identically."""
    import numpy as np
    import torch

    from mixedprecisionofrr_amd import api
    from oracle import oracle as O
    ctx = api.Context(0)
    rows = []
    for m in (96, 128, 160, 192, 256, 300, 384):
        rng = np.random.default_rng(m)
        Qm = np.linalg.qr(rng.standard_normal((m, m)))[0]
        A = (Qm * np.logspace(-9, np.log10(60.0), m)) @ Qm.T
        sc = 1.0 / np.sqrt(np.diag(A))
        A = A * sc[:, None] * sc[None, :]
        A = 0.5 * (A + A.T)
        At = torch.tensor(A, device="cuda:0")
        api.syev_batched(ctx, At)
        torch.cuda.synchronize()
        ts = []
        for _ in range(3):
            t = time.perf_counter()
            api.syev_batched(ctx, At)
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t) * 1e3)
        w = np.zeros(m)
        V = np.zeros((m, m), order="F")
        sweeps = O.lib().orc_syev(m, np.asfortranarray(A), m, w, V, m)
        ms = min(ts)
        rows.append({"m": m, "ms": ms, "sweeps": int(sweeps),
                     "us_per_round": 1e3 * ms / max(1, (sweeps + 1) * (m - 1 + (m & 1)))})
        log(json.dumps(rows[-1]))
    print(json.dumps({"metric": "symmetric eigensolve time by size (Jacobi, one workgroup)",
                      "unit": "ms", "rows": rows}), flush=True)


def qr_main(args):
    """The orthogonalizing baselines against the Hessenberg process: the comparison of
    CPP/examples/ex00.cu:55-150 (5 factorizations per method after a warm-up, time
    per factorization; A uniform [-1, 1), column-major).  MGS without and with
    reorthogonalization (reorth_tol = -1 and 1/sqrt(2)), MGS_V2, CGS, CGS2, Hessenberg.
    Besides the time, ||Q^T Q - I||_max over the kept columns and ||A - QR||_max."""
    import numpy as np
    import torch

    from mixedprecisionofrr_amd import api

    dev = "cuda:0"
    ctx = api.Context(0)
    dts = {"f64": torch.float64, "f32": torch.float32, "f16": torch.float16}
    methods = [("MGS", "mgs", -1.0), ("MGS R", "mgs", 2 ** -0.5), ("MGS V2", "mgs_v2", -1.0),
               ("CGS", "cgs", -1.0), ("CGS2", "cgs2", -1.0), ("Hessenberg", None, None)]
    ntests = max(1, args.steps)
    rows = []
    g = torch.Generator(device=dev)
    for shape in args.qr_sizes.split(","):
        m, n = (int(x) for x in shape.split("x"))
        g.manual_seed(m + n)
        A64 = (torch.rand(n, m, device=dev, dtype=torch.float64, generator=g) * 2 - 1).t()
        for dname in args.qr_dtypes.split(","):
            A = A64.to(dts[dname]).t().contiguous().t()
            for label, var, rtol in methods:
                def run():
                    if var is None:
                        Q, R, _, skip = api.hessenberg(ctx, A)
                    else:
                        Q, R, skip = api.gs_qr(ctx, A, var, reorth_tol=rtol)
                    return Q, R, skip
                run()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(ntests):
                    Q, R, skip = run()
                torch.cuda.synchronize()
                ms = (time.perf_counter() - t0) * 1e3 / ntests
                Qd = Q.double()
                keep = skip == 0
                Qk = Qd[:, keep]
                eye = torch.eye(Qk.shape[1], dtype=torch.float64, device=dev)
                loss = (Qk.t() @ Qk - eye).abs().max().item()
                res = (A.double() - Qd @ R).abs().max().item()
                rows.append({"m": m, "n": n, "dtype": dname, "method": label, "ms": ms,
                             "orth_loss": loss if var is not None else None,
                             "qr_residual": res, "dropped": int((~keep).sum().item())})
                ltxt = f"{loss:.2e}" if var is not None else "n/a (not orthogonal)"
                log(f"qr {m}x{n} {dname} {label}: {ms:.3f} ms, loss {ltxt}, |A-QR| {res:.2e}")
            del A
        del A64
    pick = [r for r in rows if r["method"] == "CGS2" and r["dtype"] == "f64"]
    ref = pick[-1] if pick else rows[-1]
    out = {"metric": "Gram-Schmidt QR baselines vs the Hessenberg process: ms per factorization",
           "value": ref["ms"], "unit": "ms", "n_gpus": 1, "steps": ntests, "warmup": 1,
           "ms_per_step": ref["ms"], "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64/f32/f16 storage, f64 dot products and norms",
           "data": "synthetic",
           "config": {"workload": f"QR of uniform [-1,1) matrices {args.qr_sizes}; value = CGS2 "
                                  f"FP64 at {ref['m']}x{ref['n']}"},
           "table": rows}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
