"""Packed matrix stream (MC_PRECISION_PACKED, and MC_PRECISION_MIXED after its
FP32 phase): each nonzero is one uint32 word holding the index in its low 24
bits and a signed 8-bit correction k in its top byte, beside the FP32 value
a32; the SpMVs use a32 + k ulp(a32)/256 with FP64 gathers, products and sums
(reference SpMVs src/lp/first_order/pdlp.cpp:22-44, solve :139-276).

Bars:
* SpMV in both orientations, block tiles of every budget and wave tiles:
  within 2^-30 of the row's absolute sum (the stored value is within
  ulp(a32)/256 <= 2^-31 |a| of the FP64 value) plus the FP64 summation
  bound 1e-13.  An FP32 value without its correction misses this bar by a
  factor of about 100, and so does a correction with the wrong sign or scale;
* solves in packed and mixed precision reach the tolerance; the returned pair
  re-scored by the oracle's FP64 residuals is within 1.5x of it; the objective
  is within 1e-6 x (1 + |p*|) of the planted optimum (or the oracle's FP64
  objective at 1e-9); packed solves, and mixed ones on latency-bound LPs
  (which skip the FP32 phase), report their iterations on the packed stream
  or, near its floor, in FP64 (the FP32 -> packed hand-over runs at the
  benchmark sizes, tests/test_gpu_configs.py);
* once the score nears the packed floor (the largest change of a KKT
  component when the check's average is scored with the packed values; about
  5e-7 in the duality gap of the uniform planted LPs, whose objective cancels
  terms 1e4 times larger) an instance continues in FP64: a solve to 1e-9
  ends in FP64 and reaches it;
* a batch compacted as its instances finish, in packed precision.
"""
import numpy as np
import pytest

import mcpdlp as mp
import oracle as orc
from test_gpu_blocks import MAKERS, SOLVE_CASES, exact_objective, long_cols, long_rows

pytestmark = pytest.mark.gpu

PACK_BAR = 2.0 ** -30


def spmv_bound(lp, v, transpose):
    a = lp.copy()
    a.value = np.abs(a.value)
    return orc.spmv_t(a, np.abs(v)) if transpose else orc.spmv(a, np.abs(v))


@pytest.mark.parametrize("budget", [0, 256, 1024, 2048])
@pytest.mark.parametrize("maker", MAKERS)
def test_packed_spmv_parity(budget, maker):
    lp = maker()
    print("packed spmv", maker.__name__, budget, lp.rows, lp.cols, lp.nnz, flush=True)
    e = mp.Engine(0)
    try:
        e.load(lp)
        e.setup(ruiz=True)
        e.set_block_budget(budget, budget)
        vals, _, _, _, _ = e.scaling(lp.nnz)
        s = lp.copy()
        s.value = vals  # the scaled matrix the engine holds
        rng = np.random.default_rng(budget + 7)
        x = rng.standard_normal(lp.cols)
        y = rng.standard_normal(lp.rows)
        for t, v in ((False, x), (True, y)):
            want = orc.spmv_t(s, v) if t else orc.spmv(s, v)
            bound = spmv_bound(s, v, t)
            got, _ = e.spmv(v, transpose=t, packed=True)
            err = np.abs(got - want)
            assert np.all(err <= PACK_BAR * bound + 1e-13 * (np.abs(want) + bound)), (t, (err / bound).max())
            # the correction is applied: the FP32 values alone are off by far more
            f32, _ = e.spmv(v, transpose=t, fp32=True)
            if np.abs(f32 - want).max() > 0:
                assert err.max() < 0.05 * np.abs(f32 - want).max(), t
    finally:
        e.close()


@pytest.mark.parametrize("precision", [mp.PRECISION_PACKED, mp.PRECISION_MIXED], ids=["packed", "mixed"])
@pytest.mark.parametrize("maker,tol", SOLVE_CASES, ids=[c[0].__name__ for c in SOLVE_CASES])
def test_packed_solves_reach_the_optimum(maker, tol, precision):
    lp = maker()
    print("packed solve", maker.__name__, precision, flush=True)
    r = mp.solve(lp, tolerance=tol, backend=mp.BACKEND_GPU, require_gpu=1, precision=precision,
                 restart_scheme=mp.SCHEME_HALPERN, max_iterations=400000)
    assert r["backend_actually_used"] == "hip"
    assert r["status"] == "optimal", r["status"]
    assert orc.kkt(lp, r["primal"], r["dual"])["score"] <= 1.5 * tol
    p_star = exact_objective(lp)
    assert abs(r["objective"] - p_star) <= 1e-6 * (1.0 + abs(p_star)), (r["objective"], p_star)
    its, f32, pk = r["iterations"], r["fp32_iterations"], r["packed_iterations"]
    print("iterations", its, "fp32", f32, "packed", pk, flush=True)
    if precision == mp.PRECISION_PACKED:
        # the last iterations may run in FP64, once the score nears the packed floor
        assert f32 == 0 and 0 < pk <= its
    else:
        # these LPs are latency-bound (at most 2^21 nonzeros): mixed skips the FP32 phase
        assert f32 == 0 and 0 < pk <= its


def test_packed_hands_over_to_fp64_below_its_floor():
    """Tolerance 1e-9 lies far below this LP's packed floor (about 5e-7 in the
    duality gap): the instance leaves the packed stream and reaches 1e-9 in
    FP64."""
    lp = mp.generate(mp.FAMILY_UNIFORM_EQ, rows=200, cols=500, seed=11)
    r = mp.solve(lp, tolerance=1e-9, backend=mp.BACKEND_GPU, require_gpu=1, precision=mp.PRECISION_PACKED,
                 restart_scheme=mp.SCHEME_HALPERN, max_iterations=400000)
    print("iterations", r["iterations"], "packed", r["packed_iterations"], flush=True)
    assert r["status"] == "optimal"
    assert 0 < r["packed_iterations"] < r["iterations"]
    assert orc.kkt(lp, r["primal"], r["dual"])["score"] <= 1.5e-9
    assert abs(r["objective"] - lp.planted_objective) <= 1e-8 * (1.0 + abs(lp.planted_objective))


def test_packed_batch_with_compaction():
    lps = ([mp.generate(mp.FAMILY_UNIFORM_EQ, rows=100 + 37 * k, cols=250 + 90 * k, seed=100 + k) for k in range(5)]
           + [long_rows(), long_cols()])
    tol = 1e-8
    res = mp.solve_batch(lps, tolerance=tol, backend=mp.BACKEND_GPU, require_gpu=1, precision=mp.PRECISION_PACKED,
                         restart_scheme=mp.SCHEME_HALPERN, max_iterations=400000)
    for lp, r in zip(lps, res):
        assert r["status"] == "optimal"
        assert orc.kkt(lp, r["primal"], r["dual"])["score"] <= 1.5 * tol
        p_star = exact_objective(lp)
        assert abs(r["objective"] - p_star) <= 1e-6 * (1.0 + abs(p_star))
        assert 0 < r["packed_iterations"] <= r["iterations"]
