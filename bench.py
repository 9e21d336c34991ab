#!/usr/bin/env python3
"""bench.py -- time to 1e-6 relative KKT of the MI355X PDLP engine on one GPU.

Workload (default): the single large planted LP -- m = 200,000 rows (70 % '=',
30 % '<='), n = 500,000 columns, Zipf(1.2) nonzeros per column in [1, 200],
|a| log-uniform in [1e-2, 1e2], 0 <= x <= 10 -- solved to a relative KKT
score <= 1e-6 (the reference's score: max of per-row relative primal
violation, relative dual residual, relative gap; pdlp.cpp:62-135) through the
product C-ABI mc_solve_pdlp with Backend::gpu, mixed precision (FP32 SpMV and
updates, FP64 checks, FP64 polish) and the KKT restart scheme.

One "step" is one full solve: host packing + H2D + device Ruiz + PDHG
iterations + D2H (the reference's total_ms also starts before Ruiz,
pdlp.cpp:141).  The JSON line also carries: the plain FP64 GPU solve of the
same LP, the returned pair re-scored in FP64 on the host, the objective error
against the planted optimum, the roofline of the dominant fused kernel
(device-event timing on the engine's stream) and the CPU baseline (the
compiled reference solve_pdlp, or the oracle, on one core: ms per iteration
on this LP times the GPU's iteration count).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "markov-zip1_amd"))

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "large": dict(family=2, spec={}, tol=1e-6,
                  name="single LP m=200000 n=500000, Zipf(1.2) nnz/col in [1,200], 70% '=' rows"),
    "cpu-lp": dict(family=1, spec={}, tol=1e-6,
                   name="single LP m=2000 n=5000, 5 nnz/col, '=' rows, 0<=x<=10"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="large", choices=sorted(WORKLOADS))
    ap.add_argument("--tolerance", type=float, default=None)
    ap.add_argument("--max-iterations", type=int, default=2_000_000)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 comparison solve")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="approximate CPU sample budget")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested); the batched and "
              f"single-LP paths run on one MI355X -- rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import mcpdlp as mp

    if mp.device_count() < 1:
        raise SystemExit("no HIP device visible")
    wl = WORKLOADS[args.workload]
    tol = args.tolerance or wl["tol"]
    t = time.time()
    lp = mp.generate(wl["family"], **wl["spec"])
    log(f"[bench] generated {wl['name']}: m={lp.rows} n={lp.cols} nnz={lp.nnz} in {time.time() - t:.1f}s")

    prec = mp.PRECISION_MIXED if args.precision == "mixed" else mp.PRECISION_FP64
    opts = mp.options(tolerance=tol, backend=mp.BACKEND_GPU, require_gpu=1, precision=prec,
                      restart_scheme=mp.SCHEME_KKT, max_iterations=args.max_iterations)

    def run(o):
        t0 = time.perf_counter()
        r = mp.solve(lp, o)
        wall = (time.perf_counter() - t0) * 1e3
        assert r["backend_actually_used"] == "hip", r["backend_actually_used"]
        return r, wall

    for k in range(args.warmup):
        r, wall = run(opts)
        log(f"[bench] warmup {k}: {r['status']} it={r['iterations']} wall={wall:.1f}ms")
    walls, res = [], None
    for k in range(args.steps):
        res, wall = run(opts)
        walls.append(wall)
        log(f"[bench] step {k}: {res['status']} it={res['iterations']} fp32_it={res['fp32_iterations']} "
            f"wall={wall:.1f}ms total={res['total_ms']:.1f} kernel={res['kernel_ms']:.1f} "
            f"setup={res['setup_ms']:.1f} h2d={res['h2d_ms']:.1f} d2h={res['d2h_ms']:.1f}")
    ms = float(np.mean(walls))

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as orc  # the FP64 checker and the CPU baseline

    kkt = orc.kkt(lp, res["primal"], res["dual"])
    obj_err = abs(res["objective"] - lp.planted_objective) / (1.0 + abs(lp.planted_objective))
    log(f"[bench] FP64 re-score: primal={kkt['primal']:.3e} dual={kkt['dual']:.3e} gap={kkt['gap']:.3e}; "
        f"objective rel err vs planted {obj_err:.2e}")

    plain = None
    if not args.no_plain:
        po = mp.options(tolerance=tol, backend=mp.BACKEND_GPU, require_gpu=1, precision=mp.PRECISION_FP64,
                        restart_scheme=mp.SCHEME_KKT, max_iterations=args.max_iterations)
        pr, pwall = run(po)
        plain = {"ms": round(pwall, 3), "status": pr["status"], "iterations": pr["iterations"],
                 "kernel_ms": round(pr["kernel_ms"], 3)}
        log(f"[bench] plain FP64: {pr['status']} it={pr['iterations']} wall={pwall:.1f}ms")

    # roofline of the fused iteration kernels (device events on the engine stream)
    eng = mp.Engine(0)
    eng.load(lp)
    eng.setup()
    kern = {}
    for name in ("primal_fp32", "dual_fp32", "primal_fp64", "dual_fp64", "spmv_fp32", "spmv_t_fp32"):
        us, b = eng.time_kernel(name, reps=50)
        kern[name] = {"us": round(us, 3), "GBps": round(b / us / 1e3, 1), "bytes": int(b)}
        log(f"[bench] kernel {name}: {us:.2f} us/launch, {b / 1e6:.1f} MB -> {b / us / 1e3:.0f} GB/s")
    (ta, tt), = [eng.tiling()]
    eng.close()
    dom = max(("primal_fp32", "dual_fp32"), key=lambda k: kern[k]["us"])
    ach = kern[dom]["GBps"]
    roof = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None}

    cpu = None
    if not args.no_cpu:
        use_ref = orc.ref_available()
        fn = orc.ref_solve if use_ref else orc.solve
        # ms per iteration from two short runs (Ruiz + setup cancel out)
        lo_it = 40
        t0 = time.perf_counter()
        fn(lp, tolerance=tol, max_iterations=lo_it)
        t_lo = time.perf_counter() - t0
        per_guess = max(t_lo / (lo_it + 20), 1e-4)
        hi_it = lo_it + max(40, int(args.cpu_seconds * 0.6 / per_guess) // 40 * 40)
        t0 = time.perf_counter()
        fn(lp, tolerance=tol, max_iterations=hi_it)
        t_hi = time.perf_counter() - t0
        ms_it = (t_hi - t_lo) * 1e3 / (hi_it - lo_it)
        cpu = {"value": round(ms_it * res["iterations"], 1), "unit": "ms", "cores": 1,
               "kind": "reference" if use_ref else "port",
               "sample": f"{hi_it - lo_it} PDHG iterations of this LP (runs of {lo_it} and {hi_it} "
                         f"iterations, difference) -> {ms_it:.2f} ms/iteration x {res['iterations']} GPU "
                         f"iterations (extrapolated; the reference's own restart rule does not reach "
                         f"{tol:g} on this LP)"}
        log(f"[bench] cpu baseline: {ms_it:.2f} ms/iteration ({cpu['kind']})")

    line = {
        "metric": "time to 1e-6 rel KKT + SpMV HBM GB/s; mixed/batched/warm vs plain PDLP, 1 GPU",
        "value": round(ms, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "fp32+fp64" if prec == mp.PRECISION_MIXED else "fp64", "data": "synthetic (planted-feasible, seeded)",
        "config": {"workload": wl["name"], "m": lp.rows, "n": lp.cols, "nnz": lp.nnz, "tolerance": tol,
                   "precision": args.precision, "restart_scheme": "kkt", "parallelism": "single GPU"},
        "status": res["status"], "iterations": res["iterations"], "fp32_iterations": res["fp32_iterations"],
        "kernel_ms": round(res["kernel_ms"], 3), "setup_ms": round(res["setup_ms"], 3),
        "h2d_ms": round(res["h2d_ms"], 3), "d2h_ms": round(res["d2h_ms"], 3),
        "kkt_fp64_rescored": {k: float(f"{kkt[k]:.4e}") for k in ("primal", "dual", "gap", "score")},
        "objective_rel_err_vs_planted": float(f"{obj_err:.3e}"),
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": round(plain["ms"] / ms, 3) if plain else None,
        "kernels": kern, "tiles": {"rows": list(ta), "cols": list(tt)},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
