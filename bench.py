#!/usr/bin/env python3
"""bench.py -- time to 1e-6 relative KKT of the MI355X PDLP engine on one GPU.

Workload (default): the single large planted LP -- m = 200,000 rows (70 % '=',
30 % '<='), n = 500,000 columns, Zipf(1.2) nonzeros per column in [1, 200],
|a| log-uniform in [1e-2, 1e2], 0 <= x <= 10 -- solved to a relative KKT
score <= 1e-6 (the reference's score: max of per-row relative primal
violation, relative dual residual, relative gap; pdlp.cpp:62-135) through the
product C-ABI mc_solve_pdlp with Backend::gpu, mixed precision (FP32 SpMVs
with an FP64 iterate and FP64 updates until the FP32 phase stalls, then FP64
SpMVs; FP64 checks throughout) and the reflected Halpern restart scheme
(--scheme; the plain FP64 comparison uses the classic PDLP restarts to the
average, --plain-scheme kkt).

One "step" is one full solve: host packing + H2D + device Ruiz + PDHG
iterations + D2H (the reference's total_ms also starts before Ruiz,
pdlp.cpp:141).  The JSON line also carries: the returned pair re-scored in
FP64 on the host, the objective error against the planted optimum, the
roofline of the dominant fused kernel (device-event timing on the engine's
stream, each launch right after the other fused step as in the solve, on the
solve's layout; HBM traffic from the committed rocprofv3 counter summary when
present) and the CPU baseline (the compiled reference
solve_pdlp, or the oracle, on one core: ms per iteration on this LP times the
GPU's iteration count).  --plain adds the plain FP64 GPU solve of the same LP
to the same tolerance (several more minutes on this LP).

Other workloads: --workload batch (256 uneven LPs in one batched solve vs the
plain FP64 GPU solve per instance), --workload warm (learned warm start vs
cold start on a perturbed LP family), --workload cpu-lp (the reference's
CPU-runnable 2000 x 5000 LP), --workload scale (the 48-period block-angular
LP, 4.8M columns, ~48M nonzeros).  --profile-kernels KINDS only times fused
kernels, for rocprofv3 runs.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "markov-zip1_amd"))

HBM_PEAK_GBS = 8000.0
T_START = time.time()
BUDGET_S = [480.0]


def left():
    """Seconds left of the run's wall-clock budget."""
    return BUDGET_S[0] - (time.time() - T_START)

METRIC = "time to 1e-6 rel KKT + SpMV HBM GB/s; mixed/batched/warm vs plain PDLP, 1 GPU"

WORKLOADS = {
    "large": dict(family=2, spec={}, tol=1e-6, row_weighting=2,
                  name="single LP m=200000 n=500000, Zipf(1.2) nnz/col in [1,200], 70% '=' rows"),
    "cpu-lp": dict(family=1, spec={}, tol=1e-6, row_weighting=3,
                   name="single LP m=2000 n=5000, 5 nnz/col, '=' rows, 0<=x<=10"),
    "scale": dict(family=5, spec={}, tol=1e-6, row_weighting=3,
                  name="block-angular LP, 48 periods x (40000 rows x 100000 cols, 8 nnz/col) + 200000 "
                       "linking rows (2 nnz/col), |a| log-uniform [1e-2,1e2]"),
    "batch": dict(tol=1e-6, name="256 planted LPs, m log-uniform [500,20000], n=2.5m, 4-8 nnz/col, "
                                 "box [0,U[1,100]], one batched solve"),
    "warm": dict(tol=1e-6, name="LP family on one structure ({rows}x{cols}, 6 nnz/col), b/c/u perturbed "
                                "(lognormal 0.1, U[0.8,1.2]); learned warm start vs cold"),
}


SCHEMES = {"reference": 0, "kkt": 1, "halpern": 2}  # mcpdlp.SCHEME_*
PRECISIONS = {"fp64": 0, "mixed": 1, "packed": 2}  # mcpdlp.PRECISION_*
# the arithmetic of each precision's SpMVs (the iterate and the checks are FP64 in all)
DTYPES = {"fp64": "fp64", "mixed": "fp32 then packed fp32+8-bit values, fp64 sums",
          "packed": "packed fp32+8-bit values, fp64 sums"}


def max_drop(args, default):
    """--primal-weight-max-drop, or the workload's default."""
    return default if args.primal_weight_max_drop is None else args.primal_weight_max_drop


def row_weighting(args, default):
    """--row-weighting, or the workload's default."""
    return default if args.row_weighting is None else args.row_weighting


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1,
                    help="full solves to time; fewer run when they would not fit --budget-s (at least one)")
    ap.add_argument("--warmup", type=int, default=0,
                    help="untimed warm-up solves, each capped at --warmup-iterations iterations")
    ap.add_argument("--warmup-iterations", type=int, default=2000)
    ap.add_argument("--budget-s", type=float, default=480.0,
                    help="wall-clock budget of the whole run; the JSON line is printed within it")
    ap.add_argument("--no-fp64-compare", action="store_true",
                    help="skip the FP64 solve under the same scheme (mixed vs FP64 precision)")
    ap.add_argument("--workload", default="large", choices=sorted(WORKLOADS))
    ap.add_argument("--tolerance", type=float, default=None)
    ap.add_argument("--max-iterations", type=int, default=2_000_000)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "packed", "fp64"])
    ap.add_argument("--plain", action="store_true",
                    help="also time the plain FP64 GPU solve of the same LP to the same tolerance")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="approximate CPU sample budget")
    ap.add_argument("--graph-chunk", type=int, default=0, help="iterations per HIP graph (<0: eager launches)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--scheme", default="halpern", choices=sorted(SCHEMES),
                    help="restart scheme of the measured solves")
    ap.add_argument("--row-weighting", type=int, default=None, choices=range(9), metavar="0..8",
                    help="rows scaled by (1 + |bound|)^(-k/8) after Ruiz in the measured solves (0: the "
                         "reference's scaling; the plain FP64 comparison always uses 0). Default: 2 for the "
                         "Zipf LP (large), 3 for the uniform and block-angular LPs (cpu-lp, scale), the values "
                         "measured best on each; 0 for the batch and the warm family (uniform LPs whose "
                         "objectives cancel larger terms: the weights' looser absolute residuals on rows with "
                         "large bounds move their objectives by more than 1e-6)")
    ap.add_argument("--primal-weight-max-drop", type=int, default=None, metavar="K",
                    help="Halpern scheme: a restart lowers the primal weight by at most K (a power of two; 0: no "
                         "limit). Default 2 for the batch of small LPs, 0 for the single LPs and the warm family")
    ap.add_argument("--plain-scheme", default="kkt", choices=sorted(SCHEMES),
                    help="restart scheme of the plain FP64 comparison (classic PDLP restarts to the average)")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--plain-sample", type=int, default=16, help="batch: instances solved one by one (FP64)")
    ap.add_argument("--train", type=int, default=384, help="warm: training instances")
    ap.add_argument("--test", type=int, default=64, help="warm: test instances")
    ap.add_argument("--warm-rows", type=int, default=2000, help="warm: rows of the family's structure")
    ap.add_argument("--warm-cols", type=int, default=5000, help="warm: columns of the family's structure")
    ap.add_argument("--no-roofline", action="store_true",
                    help="skip the per-kernel timing after the solve (profiling runs of the solve itself)")
    ap.add_argument("--profile-mlp", action="store_true",
                    help="only run the warm-start network's MFMA kernels (forward + Adam training on synthetic "
                         "data) and print their GEMM flop count, for rocprofv3 runs")
    ap.add_argument("--profile-kernels", default=None, metavar="KINDS",
                    help="only time the listed fused kernels (comma-separated Engine.KERNELS names) on the "
                         "workload's layout and exit -- the short program rocprofv3 runs wrap")
    args = ap.parse_args()
    BUDGET_S[0] = args.budget_s
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested); the batched and "
              f"single-LP paths run on one MI355X -- rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import mcpdlp as mp

    if mp.device_count() < 1:
        raise SystemExit("no HIP device visible")
    if args.profile_mlp:
        return profile_mlp(args)
    if args.profile_kernels:
        return profile_kernels(args)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "warm":
        return bench_warm(args)
    wl = WORKLOADS[args.workload]
    tol = args.tolerance or wl["tol"]
    t = time.time()
    lp = mp.generate(wl["family"], **wl["spec"])
    log(f"[bench] generated {wl['name']}: m={lp.rows} n={lp.cols} nnz={lp.nnz} in {time.time() - t:.1f}s")

    prec = PRECISIONS[args.precision]
    common = dict(tolerance=tol, backend=mp.BACKEND_GPU, require_gpu=1, restart_scheme=SCHEMES[args.scheme],
                  graph_chunk=args.graph_chunk, autotune=0 if args.no_autotune else 1,
                  row_weighting=row_weighting(args, wl["row_weighting"]), primal_weight_max_drop=max_drop(args, 0))
    opts = mp.options(precision=prec, max_iterations=args.max_iterations, **common)

    def run(o):
        t0 = time.perf_counter()
        r = mp.solve(lp, o)
        wall = (time.perf_counter() - t0) * 1e3
        assert r["backend_actually_used"] == "hip", r["backend_actually_used"]
        return r, wall

    # Warm-up: iteration-capped solves of the same LP with the same options
    # (autotune cache, captured graphs, device buffers); never the full solve.
    wopts = mp.options(precision=prec, max_iterations=min(args.max_iterations, args.warmup_iterations), **common)
    warm_done = 0
    for k in range(args.warmup):
        if left() < 0.6 * args.budget_s:
            break
        r, wall = run(wopts)
        warm_done += 1
        log(f"[bench] warmup {k}: {r['status']} it={r['iterations']} wall={wall:.1f}ms")

    # Timed solves: the first always runs; more run while they fit the budget,
    # keeping room for the kernel timing, the CPU sample and (when it fits)
    # one FP64 solve under the same scheme, which isolates the precision.
    reserve = 25.0 + (0.0 if args.no_cpu else args.cpu_seconds * 1.3)
    walls, res = [], None
    fp64 = None
    while len(walls) < max(1, args.steps):
        if walls:
            per = max(walls) / 1e3
            want_fp64 = fp64 is None and prec != mp.PRECISION_FP64 and not args.no_fp64_compare
            slack = left() - reserve - (1.3 * per if want_fp64 else 0.0)
            if slack < 1.15 * per:
                break
        res, wall = run(opts)
        walls.append(wall)
        log(f"[bench] step {len(walls) - 1}: {res['status']} it={res['iterations']} fp32_it={res['fp32_iterations']} "
            f"packed_it={res['packed_iterations']} "
            f"wall={wall:.1f}ms total={res['total_ms']:.1f} kernel={res['kernel_ms']:.1f} "
            f"setup={res['setup_ms']:.1f} h2d={res['h2d_ms']:.1f} d2h={res['d2h_ms']:.1f} ({left():.0f}s left)")
    ms = float(np.mean(walls))

    if prec != mp.PRECISION_FP64 and not args.no_fp64_compare and left() - reserve > 1.3 * max(walls) / 1e3:
        # capped so that it ends inside the budget even if its trajectory is longer
        per_it_s = 1.2 * max(walls) / 1e3 / max(1, res["iterations"])
        cap = int(min(args.max_iterations, max(1000, (left() - reserve) / per_it_s)))
        fo = mp.options(precision=mp.PRECISION_FP64, max_iterations=cap, **common)
        fr, fwall = run(fo)
        fp64 = {"ms": round(fwall, 3), "status": fr["status"], "iterations": fr["iterations"],
                "restart_scheme": args.scheme, "kernel_ms": round(fr["kernel_ms"], 3), "iteration_cap": cap,
                "note": "same LP, scheme and tolerance with FP64 SpMVs throughout"}
        log(f"[bench] FP64 (same scheme): {fr['status']} it={fr['iterations']} wall={fwall:.1f}ms")

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as orc  # the FP64 checker and the CPU baseline

    kkt = orc.kkt(lp, res["primal"], res["dual"])
    obj_err = abs(res["objective"] - lp.planted_objective) / (1.0 + abs(lp.planted_objective))
    log(f"[bench] FP64 re-score: primal={kkt['primal']:.3e} dual={kkt['dual']:.3e} gap={kkt['gap']:.3e}; "
        f"objective rel err vs planted {obj_err:.2e}")

    plain = None
    if args.plain and left() > reserve + 60:
        po = mp.options(tolerance=tol, backend=mp.BACKEND_GPU, require_gpu=1, precision=mp.PRECISION_FP64,
                        restart_scheme=SCHEMES[args.plain_scheme], max_iterations=args.max_iterations,
                        graph_chunk=args.graph_chunk, autotune=0 if args.no_autotune else 1)
        pr, pwall = run(po)
        plain = {"ms": round(pwall, 3), "status": pr["status"], "iterations": pr["iterations"],
                 "restart_scheme": args.plain_scheme,
                 "kernel_ms": round(pr["kernel_ms"], 3)}
        log(f"[bench] plain FP64: {pr['status']} it={pr['iterations']} wall={pwall:.1f}ms")

    # roofline of the fused iteration kernels (device events on the engine stream),
    # on the layout the solve uses (rows/columns relabelled by length bucket), in
    # the precision most of the solve's iterations ran in
    kern, roof, ta, tt = {}, None, None, None
    if not args.no_roofline:
        kern, (ta, tt) = time_kernels(mp, lp, ("primal_fp32", "dual_fp32", "primal_packed", "dual_packed",
                                               "primal_fp64", "dual_fp64", "spmv_fp32", "spmv_t_fp32",
                                               "gather_fp64", "gather_local_fp64", "stream_fp64", "copy"),
                                      not args.no_autotune)
        kinds = dominant_kinds(res)
        dom = max(kinds, key=lambda k: kern[k]["us"])
        ach = kern[dom]["GBps"]
        roof = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic(args.workload, dom)}
    gather = gather_ceiling(kern, lp.nnz)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(orc, lp, tol, res["iterations"], min(args.cpu_seconds, max(3.0, left() - 15.0)))

    line = {
        "metric": METRIC,
        "value": round(ms, 3), "unit": "ms", "n_gpus": 1, "steps": len(walls), "warmup": warm_done,
        "ms_per_step": round(ms, 3), "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": DTYPES[args.precision], "data": "synthetic (planted-feasible, seeded)",
        "config": {"workload": wl["name"], "m": lp.rows, "n": lp.cols, "nnz": lp.nnz, "tolerance": tol,
                   "precision": args.precision, "restart_scheme": args.scheme,
                   "row_weighting": row_weighting(args, wl["row_weighting"]), "primal_weight_max_drop": max_drop(args, 0),
                   "parallelism": "single GPU"},
        "steps_requested": args.steps, "warmup_requested": args.warmup,
        "warmup_iterations": args.warmup_iterations, "budget_s": args.budget_s,
        "walls_ms": [round(w, 1) for w in walls],
        "status": res["status"], "iterations": res["iterations"], "fp32_iterations": res["fp32_iterations"],
        "packed_iterations": res["packed_iterations"],
        "kernel_ms": round(res["kernel_ms"], 3), "setup_ms": round(res["setup_ms"], 3),
        "h2d_ms": round(res["h2d_ms"], 3), "d2h_ms": round(res["d2h_ms"], 3),
        "kkt_fp64_rescored": {k: float(f"{kkt[k]:.4e}") for k in ("primal", "dual", "gap", "score")},
        "objective_rel_err_vs_planted": float(f"{obj_err:.3e}"),
        "fp64_same_scheme": fp64,
        "speedup_vs_fp64_same_scheme": round(fp64["ms"] / ms, 3) if fp64 else None,
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": round(plain["ms"] / ms, 3) if plain else None,
        "kernels": kern, "tiles": {"rows": list(ta), "cols": list(tt)} if ta else None,
        "roofline": roof, "gather_ceiling": gather, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def dominant_kinds(res):
    """The fused primal/dual kinds of the precision most of the solve's
    iterations ran in: FP32 SpMVs, the packed stream or FP64."""
    its, f32, pk = res["iterations"], res["fp32_iterations"], res.get("packed_iterations", 0)
    if 2 * f32 > its:
        return ("primal_fp32", "dual_fp32")
    if 2 * pk > its:
        return ("primal_packed", "dual_packed")
    return ("primal_fp64", "dual_fp64")


def gather_ceiling(kern, nnz):
    """The FP64 block SpMV A x_bar on the solve's tiles with random gathers, with
    all gathers inside a 256-entry window, and with none (mcp_time_kernel kinds
    7-9, each after a primal step as in the solve): the random-gather rate this
    matrix allows, and the fused steps' time against that plain SpMV."""
    need = ("gather_fp64", "gather_local_fp64", "stream_fp64", "dual_fp64", "primal_fp64")
    if not all(k in kern for k in need):
        return None
    g = kern["gather_fp64"]["us"]
    return {"spmv_random_gathers_us": g, "spmv_window_gathers_us": kern["gather_local_fp64"]["us"],
            "spmv_no_gathers_us": kern["stream_fp64"]["us"],
            "random_gathers_per_s": float(f"{nnz / g * 1e6:.4e}"),
            "dual_fp64_vs_plain_spmv": round(g / kern["dual_fp64"]["us"], 4),
            "primal_fp64_vs_plain_spmv": round(g / kern["primal_fp64"]["us"], 4),
            "dual_packed_vs_plain_spmv": round(g / kern["dual_packed"]["us"], 4) if "dual_packed" in kern else None,
            "primal_packed_vs_plain_spmv": round(g / kern["primal_packed"]["us"], 4) if "primal_packed" in kern else None,
            "note": "fraction = plain FP64 SpMV time with the same random gathers / fused step time; "
                    "above 1 when the fused step streams fewer bytes per nonzero (packed)"}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(orc, lp, tol, gpu_iterations, seconds):
    """ms/iteration of the CPU oracle (the C restatement of the reference's
    solve_pdlp, bit-identical to the compiled reference on the pinned cases)
    on one pinned core, from two bounded runs of the same LP (Ruiz + setup
    cancel in the difference), times the GPU's iteration count."""
    aff = os.sched_getaffinity(0) if hasattr(os, "sched_getaffinity") else None
    core = min(aff) if aff else None
    if core is not None:
        os.sched_setaffinity(0, {core})
    try:
        lo_it = 40
        t0 = time.perf_counter()
        orc.solve(lp, tolerance=tol, max_iterations=lo_it)
        t_lo = time.perf_counter() - t0
        per_guess = max(t_lo / (lo_it + 20), 1e-4)
        hi_it = lo_it + max(40, int(max(seconds - t_lo, 1.0) * 0.8 / per_guess) // 40 * 40)
        t0 = time.perf_counter()
        orc.solve(lp, tolerance=tol, max_iterations=hi_it)
        t_hi = time.perf_counter() - t0
    finally:
        if core is not None:
            os.sched_setaffinity(0, aff)
    ms_it = max(t_hi - t_lo, 1e-9) * 1e3 / (hi_it - lo_it)
    log(f"[bench] cpu baseline: {ms_it:.3f} ms/iteration (oracle, core {core}, {cpu_model()})")
    return {"value": round(ms_it * gpu_iterations, 1), "unit": "ms", "cores": 1, "kind": "port",
            "cpu": cpu_model(), "pinned_core": core, "ms_per_iteration": round(ms_it, 4),
            "sample": f"{hi_it - lo_it} PDHG iterations of this LP by the oracle (runs of {lo_it} and {hi_it} "
                      f"iterations, difference; reference restart scheme, FP64, one pinned core) -> "
                      f"{ms_it:.3f} ms/iteration x {gpu_iterations} GPU iterations (extrapolated; the "
                      f"reference's own restart rule does not reach {tol:g} on this LP)"}


KERNEL_SYMBOL = {"primal_fp32": "k_primal_blk", "dual_fp32": "k_dual_blk", "primal_fp64": "k_primal_blk",
                 "dual_fp64": "k_dual_blk", "primal_packed": "k_primal_blk", "dual_packed": "k_dual_blk",
                 "spmv_fp32": "k_spmv_plain", "spmv_t_fp32": "k_spmv_plain",
                 "copy": "k_copy", "gather_fp64": "k_gather_probe<.,0>", "gather_local_fp64": "k_gather_probe<.,1>",
                 "stream_fp64": "k_gather_probe<.,2>"}


def time_kernels(mp, lp, kinds, autotune, reps=50):
    """{kind: us/bytes/GBps} from HIP events on the engine's stream, and the tiling.
    `lp` may be a list: the LPs are packed block-diagonally as in a batched solve."""
    eng = mp.Engine(0)
    eng.load([mp.relabel_by_length(q) for q in lp] if isinstance(lp, list) else mp.relabel_by_length(lp))
    eng.setup(autotune=autotune)
    kern = {}
    for name in kinds:
        us, b = eng.time_kernel(name, reps=reps)
        kern[name] = {"us": round(us, 3), "GBps": round(b / us / 1e3, 1), "bytes": int(b)}
        log(f"[bench] kernel {name} ({KERNEL_SYMBOL[name]}): {us:.2f} us/launch, {b / 1e6:.2f} MB algorithmic "
            f"-> {b / us / 1e3:.0f} GB/s")
    tiling = eng.tiling()
    (ba, na), (bt, nt) = eng.block_tiling()
    tiling = (tiling[0] + (ba, na), tiling[1] + (bt, nt))  # wave policy + tiles, block budget + tiles
    eng.close()
    return kern, tiling


def measured_traffic(workload, kind):
    """HBM bytes per launch of `kind` on `workload` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (profiles/k_primal_k_dual_<workload>_traffic.txt: lines
    'traffic <kind> <bytes>'), or None when no such summary exists."""
    path = os.path.join(ROOT, "profiles", f"k_primal_k_dual_{workload}_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic" and parts[1] == kind:
                return float(parts[2])
    return None


def profile_kernels(args):
    """Time only the listed kernels on the workload's LP (for rocprofv3 runs)."""
    import mcpdlp as mp
    wl = WORKLOADS[args.workload]
    if args.workload == "batch":
        lp = mp.generate_batch(args.batch_count)
    elif "family" in wl:
        lp = mp.generate(wl["family"], **wl["spec"])
    else:
        raise SystemExit("--profile-kernels takes the large, scale, cpu-lp or batch workload")
    kinds = [k for k in args.profile_kernels.split(",") if k]
    for k in kinds:
        if k not in mp.Engine.KERNELS:
            raise SystemExit(f"unknown kernel kind {k!r}; choose from {sorted(mp.Engine.KERNELS)}")
    kern, tiling = time_kernels(mp, lp, kinds, not args.no_autotune)
    lps = lp if isinstance(lp, list) else [lp]
    print(json.dumps({"workload": wl["name"], "instances": len(lps), "m": sum(q.rows for q in lps),
                      "n": sum(q.cols for q in lps), "nnz": sum(q.nnz for q in lps), "kernels": kern,
                      "tiles": {"rows": list(tiling[0]), "cols": list(tiling[1])}}), flush=True)


def profile_mlp(args):
    """The 256-256-256-256-64 BF16 network on MFMA: forward of 8192 samples and
    20 epochs of Adam over 2048 samples in batches of 256 (synthetic data).
    Prints the k_gemm_bf16 launches and flops of that work (each GEMM is
    2 M N K flops; forward 4 per batch, backward 4 weight and 3 input
    gradients per batch)."""
    import numpy as np
    import mcpdlp as mp
    rng = np.random.default_rng(3)
    dims = [256, 256, 256, 256, 64]
    net = mp.MLP(in_dim=256, hidden=256, layers=3, out_dim=64, capacity=256, seed=1)
    x = rng.standard_normal((2048, 256)).astype(np.float32)
    t = rng.standard_normal((2048, 64)).astype(np.float32)
    xf = rng.standard_normal((8192, 256)).astype(np.float32)
    t0 = time.perf_counter()
    net.forward(xf)
    fwd_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss, ms = net.train(x, t, batch=256, epochs=20, lr=1e-3)
    train_s = time.perf_counter() - t0
    net.close()
    layer = [a * b for a, b in zip(dims[:-1], dims[1:])]
    fwd_flops_per_row = 2 * sum(layer)
    bwd_flops_per_row = 2 * sum(layer) + 2 * sum(layer[1:])  # weight gradients + input gradients (not layer 0)
    batches = 20 * (2048 // 256)
    gemm_fwd = 4 * (8192 // 256)  # forward in chunks of the capacity
    gemm_train = batches * (4 + 4 + 3)
    flops = 8192 * fwd_flops_per_row + batches * 256 * (fwd_flops_per_row + bwd_flops_per_row)
    print(json.dumps({"k_gemm_bf16_launches": gemm_fwd + gemm_train, "gemm_flops": flops,
                      "forward_s": round(fwd_s, 4), "train_s": round(train_s, 4), "final_loss": loss,
                      "note": "achieved MFMA rate = gemm_flops / total k_gemm_bf16 time of the kernel trace"}),
          flush=True)


def _line(value_ms, workload, cfg, extra, steps=1, warmup=0):
    line = {"metric": METRIC, "value": round(value_ms, 3), "unit": "ms", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": round(value_ms, 3), "higher_is_better": False, "scaling": "strong",
            "vs_baseline": None, "dtype": "fp32+fp64", "data": "synthetic (planted-feasible, seeded)",
            "config": {"workload": workload, **cfg, "parallelism": "single GPU"}}
    if cfg.get("precision") == "fp64":
        line["dtype"] = "fp64"
    line.update(extra)
    print(json.dumps(line), flush=True)


def bench_batch(args):
    """Batched PDLP (all instances in one packed problem, graph-replayed launches)
    vs the plain FP64 GPU PDLP run instance by instance."""
    import numpy as np
    import mcpdlp as mp

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as orc

    tol = args.tolerance or 1e-6
    t = time.time()
    lps = mp.generate_batch(args.batch_count)
    log(f"[bench] batch of {len(lps)}: sum m={sum(l.rows for l in lps)} sum nnz={sum(l.nnz for l in lps)} "
        f"generated in {time.time() - t:.1f}s")
    kw = dict(tolerance=tol, backend=mp.BACKEND_GPU, require_gpu=1,
              max_iterations=args.max_iterations, graph_chunk=args.graph_chunk)
    t0 = time.perf_counter()
    prec = PRECISIONS[args.precision]
    res = mp.solve_batch(lps, precision=prec, restart_scheme=SCHEMES[args.scheme], row_weighting=row_weighting(args, 0),
                         primal_weight_max_drop=max_drop(args, 2), **kw)
    wall = (time.perf_counter() - t0) * 1e3
    its = np.array([r["iterations"] for r in res])
    ok = sum(r["status"] == "optimal" for r in res)
    log(f"[bench] batched {args.precision}: {ok}/{len(res)} optimal, iterations mean={its.mean():.0f} max={its.max()} "
        f"wall={wall:.1f}ms kernel={res[0]['kernel_ms']:.1f}ms")
    # independent FP64 re-score and objective error of a few instances
    # every instance re-scored by the oracle's FP64 residuals and held against its planted optimum
    worst, worst_obj, over = 0.0, 0.0, 0
    for k in range(len(lps)):
        sc = orc.kkt(lps[k], res[k]["primal"], res[k]["dual"])["score"]
        oe = abs(res[k]["objective"] - lps[k].planted_objective) / (1.0 + abs(lps[k].planted_objective))
        worst, worst_obj, over = max(worst, sc), max(worst_obj, oe), over + (oe > 1e-6)
    log(f"[bench] FP64 re-score of all {len(lps)} instances: worst score {worst:.3e}, worst objective error "
        f"{worst_obj:.3e} ({over} above 1e-6)")
    # roofline of the dominant fused kernel on the packed batch
    kinds = dominant_kinds({"iterations": int(its.sum()),
                            "fp32_iterations": sum(r["fp32_iterations"] for r in res),
                            "packed_iterations": sum(r["packed_iterations"] for r in res)})
    kern, _ = time_kernels(mp, lps, kinds, True)
    dom = max(kinds, key=lambda k: kern[k]["us"])
    # what a persistent batch kernel would pay per phase boundary: a grid-wide
    # barrier in a cooperative kernel vs a dependent launch in a captured graph
    eng = mp.Engine(0)
    sync = {}
    for g in (8, 64, 256):
        bu, lu = eng.time_sync(g, 2000)
        sync[str(g)] = {"grid_barrier_us": round(bu, 3), "graph_launch_us": round(lu, 3)}
        log(f"[bench] sync probe, {g} workgroups: grid barrier {bu:.2f} us, graph launch {lu:.2f} us")
    eng.close()
    roof = {"bound": "hbm", "kernel": dom, "achieved": kern[dom]["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(kern[dom]["GBps"] / HBM_PEAK_GBS, 4), "traffic": measured_traffic("batch", dom)}
    cpu = None
    if not args.no_cpu:
        # compiled reference on 1 core: ms/iteration of three instances (smallest, median,
        # largest), fitted linearly in nnz and applied to every instance's GPU iteration count
        use_ref = orc.ref_available()
        fn = orc.ref_solve if use_ref else orc.solve
        order_nnz = np.argsort([l.nnz for l in lps])
        picks = [order_nnz[0], order_nnz[len(lps) // 2], order_nnz[-1]]
        pts = []
        for k in picks:
            lo_it = 40
            t0c = time.perf_counter()
            fn(lps[k], tolerance=tol, max_iterations=lo_it)
            t_lo = time.perf_counter() - t0c
            per_guess = max(t_lo / (lo_it + 20), 1e-5)
            hi_it = lo_it + max(40, int(args.cpu_seconds / 3 * 0.6 / per_guess) // 40 * 40)
            t0c = time.perf_counter()
            fn(lps[k], tolerance=tol, max_iterations=hi_it)
            ms_it = (time.perf_counter() - t0c - t_lo) * 1e3 / (hi_it - lo_it)
            pts.append((lps[k].nnz, ms_it))
            log(f"[bench] cpu baseline sample nnz={lps[k].nnz}: {ms_it:.3f} ms/iteration")
        a = sum(p[0] * p[1] for p in pts) / sum(p[0] * p[0] for p in pts)  # ms per iteration per nonzero
        est_cpu = float(sum(a * l.nnz * r["iterations"] for l, r in zip(lps, res)))
        cpu = {"value": round(est_cpu, 1), "unit": "ms", "cores": 1, "kind": "reference" if use_ref else "port",
               "sample": "ms/iteration of the smallest, median and largest instance "
                         + ", ".join(f"(nnz {n}: {v:.3f} ms)" for n, v in pts)
                         + "; fitted linearly in nnz and summed over all instances x their GPU iteration counts "
                           "(extrapolated)"}
    # plain FP64 GPU PDLP, one instance at a time, on an evenly spaced sample by size
    order = np.argsort([l.rows for l in lps])
    sample = order[np.linspace(0, len(lps) - 1, min(args.plain_sample, len(lps))).astype(int)]
    plain_ms, plain_it = [], []
    for k in sample:
        t1 = time.perf_counter()
        r = mp.solve(lps[k], precision=mp.PRECISION_FP64, restart_scheme=SCHEMES[args.plain_scheme], **kw)
        plain_ms.append((time.perf_counter() - t1) * 1e3)
        plain_it.append(r["iterations"])
        log(f"[bench] plain FP64 instance m={lps[k].rows}: {r['status']} it={r['iterations']} {plain_ms[-1]:.1f}ms")
    est = float(np.mean(plain_ms)) * len(lps)
    _line(wall, WORKLOADS["batch"]["name"],
          {"instances": len(lps), "sum_rows": int(sum(l.rows for l in lps)), "sum_nnz": int(sum(l.nnz for l in lps)),
           "tolerance": tol, "precision": args.precision, "restart_scheme": args.scheme,
           "row_weighting": row_weighting(args, 0), "primal_weight_max_drop": max_drop(args, 2),
           "plain_restart_scheme": args.plain_scheme},
          {"optimal": int(ok), "iterations_mean": float(its.mean()), "iterations_max": int(its.max()),
           "kkt_fp64_rescored_worst": float(f"{worst:.4e}"),
           "objective_rel_err_vs_planted_worst": float(f"{worst_obj:.3e}"),
           "objective_rel_err_above_1e-6": int(over),
           "plain_fp64_per_instance": {"sampled": int(len(sample)), "mean_ms": round(float(np.mean(plain_ms)), 3),
                                       "estimated_total_ms": round(est, 1),
                                       "iterations_mean": float(np.mean(plain_it)),
                                       "note": "evenly spaced by size; total = mean x instances (extrapolated)"},
           "speedup_vs_plain_fp64_per_instance": round(est / wall, 2),
           "kernels": kern, "roofline": roof, "cpu_baseline": cpu,
           "sync_probe": {"workgroups": sync,
                          "note": "per phase boundary: grid-wide barrier inside a cooperative persistent kernel "
                                  "vs dependent launch in a captured HIP graph (mcp_time_sync)"}})


def bench_warm(args):
    """Learned warm start (BF16 MLP on MFMA) vs cold start on a family of perturbed LPs."""
    import numpy as np
    import mcpdlp as mp
    import warmstart

    tol = args.tolerance or 1e-6
    ntr, nte = args.train, args.test
    t = time.time()
    base, fam = mp.generate_family(dict(rows=args.warm_rows, cols=args.warm_cols, nnz_min=6, nnz_max=6),
                                   seeds=[0x4D4350444C50 + 4000 + k for k in range(ntr + nte)], sigma=0.1)
    train, test = fam[:ntr], fam[ntr:]
    log(f"[bench] family: {ntr} train + {nte} test instances of {base.rows}x{base.cols} "
        f"({base.nnz} nnz) in {time.time() - t:.1f}s")
    kw = dict(tolerance=tol, backend=mp.BACKEND_GPU, require_gpu=1, restart_scheme=SCHEMES[args.scheme],
              max_iterations=args.max_iterations, precision=mp.PRECISION_MIXED, graph_chunk=args.graph_chunk,
              row_weighting=row_weighting(args, 0), primal_weight_max_drop=max_drop(args, 0))
    t0 = time.perf_counter()
    sols = []
    for k in range(0, ntr, 128):
        sols += mp.solve_batch(train[k:k + 128], **kw)
        log(f"[bench] training solutions {len(sols)}/{ntr}: {(time.perf_counter() - t0):.1f}s")
    gen_ms = (time.perf_counter() - t0) * 1e3
    log(f"[bench] training solutions: {sum(r['status'] == 'optimal' for r in sols)}/{ntr} optimal in {gen_ms:.0f}ms")
    # the last nval training instances validate the start primal weight
    nval = min(16, max(1, ntr // 8))
    fit_n = ntr - nval
    ws = warmstart.WarmStart(base)
    fit = ws.fit(train[:fit_n], [r["primal"] for r in sols[:fit_n]], [r["dual"] for r in sols[:fit_n]],
                 omegas=[r["primal_weight"] for r in sols[:fit_n]],
                 iterations=[r["iterations"] for r in sols[:fit_n]])
    log(f"[bench] MLP fit: {fit}")
    tune = ws.tune_omega(train[fit_n:], **kw)
    log(f"[bench] start primal weight on {nval} validation instances: {tune}")
    t1 = time.perf_counter()
    cold = mp.solve_batch(test, **kw)
    cold_ms = (time.perf_counter() - t1) * 1e3
    t2 = time.perf_counter()
    warm, fallbacks = ws.solve_batch(test, **kw)
    warm_ms = (time.perf_counter() - t2) * 1e3
    # ablation: every test instance started at the training solutions' mean
    t3 = time.perf_counter()
    mean_start, mean_fallbacks = ws.solve_batch(test, use_network=False, **kw)
    mean_ms = (time.perf_counter() - t3) * 1e3
    ci = np.array([r["iterations"] for r in cold])
    wi = np.array([r["iterations"] for r in warm])
    mi = np.array([r["iterations"] for r in mean_start])
    log(f"[bench] mean start (no network): it mean={mi.mean():.0f} max={mi.max()} {mean_ms:.0f}ms "
        f"({mean_fallbacks} cold fallbacks)")
    log(f"[bench] cold: it mean={ci.mean():.0f} max={ci.max()} {cold_ms:.0f}ms | warm: it mean={wi.mean():.0f} "
        f"max={wi.max()} {warm_ms:.0f}ms (incl. features+inference; {fallbacks} cold fallbacks)")
    _line(warm_ms, WORKLOADS["warm"]["name"].format(rows=base.rows, cols=base.cols),
          {"train": ntr, "test": nte, "m": base.rows, "n": base.cols, "nnz": base.nnz, "tolerance": tol,
           "mlp": "256-256-256-256-64 BF16 MFMA", "precision": "mixed", "restart_scheme": args.scheme,
           "row_weighting": row_weighting(args, 0), "primal_weight_max_drop": max_drop(args, 0)},
          {"cold_ms": round(cold_ms, 1), "warm_ms": round(warm_ms, 1),
           "iterations_cold_mean": float(ci.mean()), "iterations_warm_mean": float(wi.mean()),
           "iterations_cold_max": int(ci.max()), "iterations_warm_max": int(wi.max()),
           "optimal_cold": int(sum(r["status"] == "optimal" for r in cold)),
           "optimal_warm": int(sum(r["status"] == "optimal" for r in warm)),
           "warm_fallbacks": int(fallbacks), "start_primal_weight": tune,
           "mean_start": {"iterations_mean": float(mi.mean()), "iterations_max": int(mi.max()),
                          "ms": round(mean_ms, 1), "fallbacks": int(mean_fallbacks),
                          "note": "ablation: every instance started at the training solutions' mean in the "
                                  "reduced basis, same start weight and fallback"},
           "speedup_warm_vs_cold": round(cold_ms / warm_ms, 3), "fit": fit,
           "training_solutions_ms": round(gen_ms, 1)})


if __name__ == "__main__":
    main()
