"""GPU parity at the benchmark workloads' own sizes (reference path
src/lp/first_order/pdlp.cpp:139-276, SpMVs :22-44, Ruiz
src/scale/ruiz_scaling.cpp:152-247).

* single large LP, m = 200,000, n = 500,000, Zipf(1.2) nonzeros per column
  (11.1M nonzeros): device Ruiz bit-identical to the oracle; SpMV in both
  orientations on the tiles and block budgets the solve uses (autotuned, rows
  and columns relabelled as the solve relabels them): reference order
  bit-identical, fast order within 1e-13, the packed stream within 2^-30 and
  FP32 within 2e-6 of the row's absolute sum; 120 iterations in reference order bit-identical to the oracle
  (reference and Halpern schemes, and Halpern with the benchmark's row
  weighting); and the benchmark's own solve (mixed precision, Halpern scheme,
  row weighting k = 2, autotune on) to 1e-6, re-scored by the oracle's FP64
  residuals and held to 1e-6 x (1 + |p*|) of the planted optimum;
* 256-LP uneven batch (m log-uniform in [500, 20000], n = 2.5 m, 21.1M
  nonzeros): 200 iterations in reference order, every instance bit-identical
  to its own oracle solve (default options, and the benchmark's Halpern
  scheme with the primal-weight drop limit); then the
  batched fast-order solve with the benchmark's options to 1e-6 with every
  instance re-scored and held to its planted optimum;
* learned warm-start family on the 20,000 x 50,000 structure (6 nonzeros per
  column): a small train/test split; every warm-started instance converges,
  is re-scored by the oracle and matches its planted optimum, and the warm
  start does not cost iterations on average;
* 48-period block-angular LP (2.12M x 4.8M, 48M nonzeros): device Ruiz and
  SpMV parity as above, 80 iterations in reference order bit-identical
  (reference scheme; Halpern with row weighting).
"""
import numpy as np
import pytest

import mcpdlp as mp
import oracle as orc

pytestmark = pytest.mark.gpu

OBJ_BAR = 1e-6  # |objective - p*| <= OBJ_BAR * (1 + |p*|)


def solve_layout(lp):
    """The LP as the solve lays it out: natural order for local rows, else
    rows and columns relabelled by length bucket (host/pdlp_gpu.cpp)."""
    return lp if mp.rows_are_local(lp) else mp.relabel_by_length(lp)


def abs_lp(lp):
    q = lp.copy()
    q.value = np.abs(q.value)
    return q


def check_ruiz_and_spmv(lp):
    """Device Ruiz on the solve's layout == oracle; SpMVs of the scaled matrix
    on the solve's (autotuned) tiles against the oracle's."""
    q = solve_layout(lp)
    e = mp.Engine(0)
    try:
        e.load(q)
        e.setup(ruiz=True, ruiz_iterations=10, autotune=True)
        vals, rs, cs, passes, conv = e.scaling(q.nnz)
        ref = orc.equilibrate(q, 10)
        assert passes == ref["passes"] and bool(conv) == ref["converged"]
        assert np.array_equal(vals, ref["lp"].value)
        assert np.array_equal(rs, ref["row_scale"]) and np.array_equal(cs, ref["col_scale"])
        (ba, na), (bt, nt) = e.block_tiling()
        print("block budgets", ba, bt, "tiles", na, nt, flush=True)
        assert ba > 0 and bt > 0 and na > 0 and nt > 0
        s = ref["lp"]
        sa = abs_lp(s)
        rng = np.random.default_rng(3)
        x = rng.standard_normal(q.cols)
        x[rng.random(q.cols) < 0.2] = 0.0
        y = rng.standard_normal(q.rows)
        for t, v, want, bound in ((False, x, orc.spmv(s, x), orc.spmv(sa, np.abs(x))),
                                  (True, y, orc.spmv_t(s, y), orc.spmv_t(sa, np.abs(y)))):
            g, _ = e.spmv(v, transpose=t, exact=True)
            assert np.array_equal(g, want), ("reference order", t)
            f, _ = e.spmv(v, transpose=t)
            assert np.all(np.abs(f - want) <= 1e-13 * (np.abs(want) + bound)), ("fast order", t)
            f32, _ = e.spmv(v, transpose=t, fp32=True)
            assert np.all(np.abs(f32 - want) <= 2e-6 * bound + 1e-6 * np.abs(want) + 1e-30), ("fp32", t)
            # the packed stream the benchmark's solve runs on: within 2^-30 of the absolute sum
            pk, _ = e.spmv(v, transpose=t, packed=True)
            assert np.all(np.abs(pk - want) <= 2.0 ** -30 * bound + 1e-13 * (np.abs(want) + bound)), ("packed", t)
    finally:
        e.close()


# the benchmark's options beyond the scheme: row weighting k = 2 on the Zipf
# LP and k = 3 on the block-angular one, the primal-weight drop limit 2 (and
# no row weighting) on the batch
BENCH_LARGE = dict(row_weighting=2)
BENCH_SCALE = dict(row_weighting=3)
BENCH_BATCH = dict(primal_weight_max_drop=2)


def check_bit_exact(lp, its, scheme, **extra):
    kw = dict(tolerance=1e-6, max_iterations=its, restart_scheme=scheme, **extra)
    want = orc.solve(lp, **kw)
    got = mp.solve(lp, backend=mp.BACKEND_GPU, require_gpu=1, reduction_order=mp.ORDER_REFERENCE, **kw)
    assert got["backend_actually_used"] == "hip"
    assert got["status"] == want["status"] and got["iterations"] == want["iterations"] == its
    assert got["objective"] == want["objective"]
    assert got["primal_infeasibility"] == want["primal_infeasibility"]
    assert got["dual_infeasibility"] == want["dual_infeasibility"] and got["duality_gap"] == want["duality_gap"]
    assert np.array_equal(got["primal"], want["primal"]) and np.array_equal(got["dual"], want["dual"])


def check_solution(lp, r, tol=1e-6):
    assert r["status"] == "optimal", (r["status"], r["iterations"])
    k = orc.kkt(lp, r["primal"], r["dual"])
    assert k["score"] <= tol * 1.01, k
    p_star = lp.planted_objective
    err = abs(r["objective"] - p_star) / (1.0 + abs(p_star))
    assert err <= OBJ_BAR, (r["objective"], p_star, err)
    return k["score"], err


# ---- single large LP (200k x 500k, Zipf) -------------------------------------
@pytest.fixture(scope="module")
def large():
    lp = mp.generate(mp.FAMILY_ZIPF_MIXED)
    print("large LP", lp.rows, lp.cols, lp.nnz, flush=True)
    assert (lp.rows, lp.cols) == (200000, 500000) and lp.nnz > 10_000_000
    return lp


def test_large_ruiz_and_spmv(large):
    check_ruiz_and_spmv(large)


@pytest.mark.parametrize("scheme,extra", [(mp.SCHEME_REFERENCE, {}), (mp.SCHEME_HALPERN, {}),
                                          (mp.SCHEME_HALPERN, BENCH_LARGE)], ids=["reference", "halpern", "bench"])
def test_large_reference_order_120_iterations(large, scheme, extra):
    print("large reference order, scheme", scheme, extra, flush=True)
    check_bit_exact(large, 120, scheme, **extra)


@pytest.mark.timeout(600)
def test_large_benchmark_solve_to_1e6(large):
    """The benchmark's solve: mixed precision, Halpern scheme, row weighting,
    autotuned tiles."""
    print("large LP full solve (about 40 s)", flush=True)
    r = mp.solve(large, tolerance=1e-6, backend=mp.BACKEND_GPU, require_gpu=1, precision=mp.PRECISION_MIXED,
                 restart_scheme=mp.SCHEME_HALPERN, autotune=1, max_iterations=2_000_000, **BENCH_LARGE)
    assert r["backend_actually_used"] == "hip"
    # 11.1M nonzeros: an FP32 phase, then the packed stream
    assert r["fp32_iterations"] > 0 and r["packed_iterations"] > 0
    assert r["fp32_iterations"] + r["packed_iterations"] <= r["iterations"]
    score, err = check_solution(large, r)
    print("iterations", r["iterations"], "fp32", r["fp32_iterations"], "score", score, "objective error", err,
          "wall", r["total_ms"], flush=True)


# ---- small LP with the autotuner on (the benchmark's setting) ----------------
def test_autotuned_solve_matches_planted_optimum():
    lp = mp.generate(mp.FAMILY_UNIFORM_EQ, rows=2000, cols=5000, seed=31)
    r = mp.solve(lp, tolerance=1e-8, backend=mp.BACKEND_GPU, require_gpu=1, precision=mp.PRECISION_MIXED,
                 restart_scheme=mp.SCHEME_HALPERN, autotune=1, max_iterations=2_000_000)
    check_solution(lp, r, 1e-8)
    again = mp.solve(lp, tolerance=1e-8, backend=mp.BACKEND_GPU, require_gpu=1, precision=mp.PRECISION_MIXED,
                     restart_scheme=mp.SCHEME_HALPERN, autotune=1, max_iterations=2_000_000)
    # the tuned choice is cached per problem signature: the repeat takes the same path
    assert again["iterations"] == r["iterations"] and again["objective"] == r["objective"]


# ---- 256-LP uneven batch ------------------------------------------------------
@pytest.fixture(scope="module")
def batch():
    lps = mp.generate_batch(256)
    print("batch", len(lps), sum(q.rows for q in lps), sum(q.nnz for q in lps), flush=True)
    return lps


@pytest.mark.timeout(600)
@pytest.mark.parametrize("extra", [{}, dict(restart_scheme=mp.SCHEME_HALPERN, **BENCH_BATCH)], ids=["reference", "bench"])
def test_batch_reference_order_200_iterations(batch, extra):
    kw = dict(tolerance=1e-6, max_iterations=200, **extra)
    print("batch reference order", extra, flush=True)
    got = mp.solve_batch(batch, backend=mp.BACKEND_GPU, require_gpu=1, reduction_order=mp.ORDER_REFERENCE, **kw)
    for k, (lp, g) in enumerate(zip(batch, got)):
        w = orc.solve(lp, **kw)
        assert g["backend_actually_used"] == "hip"
        assert g["status"] == w["status"] and g["iterations"] == w["iterations"], k
        assert g["objective"] == w["objective"], k
        assert np.array_equal(g["primal"], w["primal"]) and np.array_equal(g["dual"], w["dual"]), k


@pytest.mark.timeout(600)
def test_batch_solve_to_1e6(batch):
    got = mp.solve_batch(batch, tolerance=1e-6, backend=mp.BACKEND_GPU, require_gpu=1,
                         precision=mp.PRECISION_MIXED, restart_scheme=mp.SCHEME_HALPERN, max_iterations=2_000_000,
                         **BENCH_BATCH)
    worst = [0.0, 0.0]
    for lp, r in zip(batch, got):
        score, err = check_solution(lp, r)
        worst = [max(worst[0], score), max(worst[1], err)]
    its = np.array([r["iterations"] for r in got])
    print("batch iterations mean", its.mean(), "max", its.max(), "worst score", worst[0], "worst objective error",
          worst[1], flush=True)


# ---- learned warm start on the 20,000 x 50,000 structure ------------------------
@pytest.mark.timeout(600)
def test_warm_start_family_at_full_structure():
    import warmstart
    ntr, nval, nte = 56, 8, 8
    base, fam = mp.generate_family(dict(rows=20000, cols=50000, nnz_min=6, nnz_max=6),
                                   seeds=[0x4D4350444C50 + 4000 + k for k in range(ntr + nval + nte)], sigma=0.1)
    train, val, test = fam[:ntr], fam[ntr:ntr + nval], fam[ntr + nval:]
    kw = dict(tolerance=1e-6, backend=mp.BACKEND_GPU, require_gpu=1, restart_scheme=mp.SCHEME_HALPERN,
              precision=mp.PRECISION_MIXED, max_iterations=2_000_000)
    print("warm family", base.rows, base.cols, base.nnz, "train", ntr, flush=True)
    sols = mp.solve_batch(train, **kw)
    for lp, r in zip(train, sols):
        check_solution(lp, r)
    ws = warmstart.WarmStart(base)
    fit = ws.fit(train, [r["primal"] for r in sols], [r["dual"] for r in sols],
                 omegas=[r["primal_weight"] for r in sols], iterations=[r["iterations"] for r in sols])
    tune = ws.tune_omega(val, **kw)
    print("fit", fit, "tune", tune, flush=True)
    cold = mp.solve_batch(test, **kw)
    warm, fallbacks = ws.solve_batch(test, **kw)
    for lp, c, w in zip(test, cold, warm):
        check_solution(lp, c)
        check_solution(lp, w)
    ci = np.mean([r["iterations"] for r in cold])
    wi = np.mean([r["iterations"] for r in warm])
    print(f"iterations to 1e-6: cold mean {ci:.0f}, warm mean {wi:.0f} ({100 * (wi / ci - 1):+.1f} %), "
          f"{fallbacks} cold fallbacks", flush=True)
    assert wi <= ci


# ---- 48-period block-angular LP ---------------------------------------------------
@pytest.fixture(scope="module")
def scale():
    lp = mp.generate(mp.FAMILY_BLOCK_ANGULAR)
    print("block-angular LP", lp.rows, lp.cols, lp.nnz, flush=True)
    assert lp.cols == 4_800_000 and lp.nnz >= 45_000_000
    return lp


def test_scale_ruiz_and_spmv(scale):
    check_ruiz_and_spmv(scale)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("scheme,extra", [(mp.SCHEME_REFERENCE, {}), (mp.SCHEME_HALPERN, BENCH_SCALE)],
                         ids=["reference", "bench"])
def test_scale_reference_order_80_iterations(scale, scheme, extra):
    print("scale reference order", scheme, extra, flush=True)
    check_bit_exact(scale, 80, scheme, **extra)
