"""Row weighting and the primal-weight drop limit (extensions).

Row weighting (mc_pdlp_options::row_weighting = k): after the
Ruiz passes (reference src/scale/ruiz_scaling.cpp:152-247) row i of A and its
bounds are scaled by (1 + beta_i)^(-k/8), beta_i the largest magnitude of its
finite original bounds.  The solve (reference pdlp.cpp:139-276) then runs on
the weighted problem; every reported quantity is in the original space.

Bars:
* device scalers and scaled matrix bit-identical to the oracle's
  (or_apply_row_weights after or_equilibrate_model), with and without Ruiz;
* reference order: bit-identical to the oracle's solve (iterations,
  objective, primal and dual bits) under the reference, KKT and Halpern
  schemes, single and batched;
* fast order, mixed precision: reaches its tolerance (1e-6; 1e-8 for the
  uniform LP, whose objective cancels terms 1e4 times larger); the pair
  re-scored by the oracle's FP64 residuals is within 1.5x of it, the
  objective within 1e-6 x (1 + |p*|) of the planted optimum.

The drop limit (primal_weight_max_drop = K, Halpern scheme): a restart lowers
the primal weight by at most K.  Bars: reference order bit-identical to the
oracle, single and batched; fast order reaches the planted optimum.
"""
import numpy as np
import pytest

import mcpdlp as mp
import oracle as orc

pytestmark = pytest.mark.gpu


def uniform(seed=11, m=200, n=500):
    return mp.generate(mp.FAMILY_UNIFORM_EQ, rows=m, cols=n, seed=seed)


def zipf_mixed(seed=12, m=300, n=700):
    return mp.generate(mp.FAMILY_ZIPF_MIXED, rows=m, cols=n, nnz_max=40, seed=seed)


def ranged():
    """Equality, one-sided, ranged and free rows (beta = 3, 7, 5 and 0)."""
    inf = np.inf
    A = [[1.0, 2.0, 0.0, -1.0], [0.5, 0.0, 3.0, 1.0], [0.0, -1.0, 1.0, 2.0], [1.0, 1.0, 1.0, 1.0]]
    return mp.LPData.from_dense(A, [1.0, -1.0, 2.0, 0.5], [3.0, -inf, -2.0, -inf], [3.0, -7.0, 5.0, inf],
                                [0.0, -10.0, 0.0, 0.0], [10.0, 10.0, inf, 4.0])


@pytest.mark.parametrize("k,ruiz", [(3, True), (8, True), (1, True), (3, False)])
@pytest.mark.parametrize("maker", [uniform, zipf_mixed, ranged])
def test_device_row_weights_bit_exact(maker, k, ruiz):
    lp = maker()
    print("row weights", maker.__name__, k, ruiz, flush=True)
    e = mp.Engine(0)
    try:
        e.load(lp)
        e.setup(ruiz=ruiz, ruiz_iterations=10, row_weighting=k)
        vals, rs, cs, _, _ = e.scaling(lp.nnz)
        ref = orc.weighted_scaling(lp, k, ruiz=ruiz)
        assert np.array_equal(rs, ref["row_scale"]) and np.array_equal(cs, ref["col_scale"])
        assert np.array_equal(vals, ref["lp"].value)
        x = np.linspace(-1, 1, lp.cols)
        g, _ = e.spmv(x, exact=True)
        assert np.array_equal(g, orc.spmv(ref["lp"], x))
    finally:
        e.close()


SCHEMES = [(mp.SCHEME_REFERENCE, 400), (mp.SCHEME_KKT, 3000), (mp.SCHEME_HALPERN, 3000)]


@pytest.mark.parametrize("scheme,its", SCHEMES, ids=["reference", "kkt", "halpern"])
@pytest.mark.parametrize("maker", [uniform, zipf_mixed, ranged])
def test_row_weighted_reference_order_bit_exact(maker, scheme, its):
    lp = maker()
    kw = dict(tolerance=1e-6, max_iterations=its, restart_scheme=scheme, row_weighting=3)
    print("row weighted solve", maker.__name__, scheme, flush=True)
    want = orc.solve(lp, **kw)
    got = mp.solve(lp, backend=mp.BACKEND_GPU, require_gpu=1, reduction_order=mp.ORDER_REFERENCE, **kw)
    assert got["backend_actually_used"] == "hip"
    assert got["status"] == want["status"] and got["iterations"] == want["iterations"]
    assert got["objective"] == want["objective"]
    assert np.array_equal(got["primal"], want["primal"]) and np.array_equal(got["dual"], want["dual"])
    assert (got["primal_infeasibility"], got["dual_infeasibility"], got["duality_gap"]) == \
           (want["primal_infeasibility"], want["dual_infeasibility"], want["duality_gap"])


def test_row_weighted_batch_reference_order():
    lps = [uniform(seed=100 + k, m=100 + 37 * k, n=250 + 90 * k) for k in range(4)] + [zipf_mixed(seed=77, m=120, n=260),
                                                                                       ranged()]
    kw = dict(tolerance=1e-6, max_iterations=3000, restart_scheme=mp.SCHEME_HALPERN, row_weighting=3)
    got = mp.solve_batch(lps, backend=mp.BACKEND_GPU, require_gpu=1, reduction_order=mp.ORDER_REFERENCE, **kw)
    for lp, g in zip(lps, got):
        w = orc.solve(lp, **kw)
        assert g["iterations"] == w["iterations"] and g["objective"] == w["objective"]
        assert np.array_equal(g["primal"], w["primal"]) and np.array_equal(g["dual"], w["dual"])


# the uniform LP's objective (about 0.12) cancels terms of size 1e3, so it is
# solved to 1e-8 for its objective to be within 1e-6 of the planted one
@pytest.mark.parametrize("maker,tol", [(uniform, 1e-8), (zipf_mixed, 1e-6)], ids=["uniform", "zipf_mixed"])
def test_row_weighted_mixed_solve_reaches_the_planted_optimum(maker, tol):
    lp = maker()
    r = mp.solve(lp, tolerance=tol, backend=mp.BACKEND_GPU, require_gpu=1, precision=mp.PRECISION_MIXED,
                 restart_scheme=mp.SCHEME_HALPERN, row_weighting=3, max_iterations=400000, autotune=1)
    print("iterations", r["iterations"], "packed", r["packed_iterations"], flush=True)
    assert r["backend_actually_used"] == "hip" and r["status"] == "optimal"
    assert orc.kkt(lp, r["primal"], r["dual"])["score"] <= 1.5 * tol
    assert abs(r["objective"] - lp.planted_objective) <= 1e-6 * (1.0 + abs(lp.planted_objective))


@pytest.mark.parametrize("maker", [uniform, zipf_mixed, ranged])
def test_max_drop_reference_order_bit_exact(maker):
    lp = maker()
    kw = dict(tolerance=1e-6, max_iterations=5000, restart_scheme=mp.SCHEME_HALPERN, row_weighting=3,
              primal_weight_max_drop=2)
    print("max drop solve", maker.__name__, flush=True)
    want = orc.solve(lp, **kw)
    got = mp.solve(lp, backend=mp.BACKEND_GPU, require_gpu=1, reduction_order=mp.ORDER_REFERENCE, **kw)
    assert got["status"] == want["status"] and got["iterations"] == want["iterations"]
    assert got["objective"] == want["objective"]
    assert np.array_equal(got["primal"], want["primal"]) and np.array_equal(got["dual"], want["dual"])


def test_max_drop_batch_reference_order_and_fast_solve():
    lps = [uniform(seed=300 + k, m=150 + 41 * k, n=380 + 95 * k) for k in range(5)] + [zipf_mixed(seed=78, m=140, n=300)]
    kw = dict(tolerance=1e-6, max_iterations=5000, restart_scheme=mp.SCHEME_HALPERN, row_weighting=3,
              primal_weight_max_drop=2)
    got = mp.solve_batch(lps, backend=mp.BACKEND_GPU, require_gpu=1, reduction_order=mp.ORDER_REFERENCE, **kw)
    for lp, g in zip(lps, got):
        w = orc.solve(lp, **kw)
        assert g["iterations"] == w["iterations"] and g["objective"] == w["objective"]
        assert np.array_equal(g["primal"], w["primal"]) and np.array_equal(g["dual"], w["dual"])
    tol = 1e-8  # the uniform LPs' objectives cancel terms 1e4 times larger
    fast = mp.solve_batch(lps, backend=mp.BACKEND_GPU, require_gpu=1, precision=mp.PRECISION_MIXED,
                          **dict(kw, tolerance=tol, max_iterations=400000))
    for lp, r in zip(lps, fast):
        assert r["status"] == "optimal"
        assert orc.kkt(lp, r["primal"], r["dual"])["score"] <= 1.5 * tol
        assert abs(r["objective"] - lp.planted_objective) <= 1e-6 * (1.0 + abs(lp.planted_objective))


@pytest.mark.parametrize("k", [0, 3])
def test_warm_start_from_the_solution_stops_at_the_first_check(k):
    """On the device, a warm start at a solution found to 1e-9 is optimal at
    the first check (40 iterations), with and without row weighting: warm_x
    and warm_y enter the scaled space through the weighted row scales."""
    lp = uniform()
    kw = dict(tolerance=1e-9, backend=mp.BACKEND_GPU, require_gpu=1, restart_scheme=mp.SCHEME_HALPERN,
              row_weighting=k, max_iterations=200000)
    ref = mp.solve(lp, **kw)
    assert ref["status"] == "optimal"
    r = mp.solve(lp, **dict(kw, tolerance=1e-6), warm_primal=ref["primal"], warm_dual=ref["dual"],
                 initial_primal_weight=ref["primal_weight"])
    print("warm iterations", r["iterations"], flush=True)
    assert r["status"] == "optimal" and r["iterations"] == 40, r["iterations"]
