"""Block tiles of the fused fast-order steps (one 256-thread workgroup streams
a run of rows, the products are summed per row from LDS) against the oracle
and against the wave-task form of the same steps.

Bars:
* SpMV over block tiles of every budget (256, 512, 1024, 2048 nonzeros), including
  rows longer than the budget in both orientations: FP64 within 1e-13 of the
  oracle's serial sum, FP32 within 2e-6 of the row's absolute sum;
* full solves (Halpern scheme, FP64 and mixed precision, single LPs and a
  batch that is compacted as its instances finish) with the block tiles on
  and off: all reach their tolerance, the returned pairs re-scored by the
  oracle's FP64 residuals are <= 1.5x the tolerance, and every objective is
  within 1e-6 x (1 + |p*|) of the exact optimum p*: the planted c^T x* of a
  generated LP, or the oracle's FP64 objective at tolerance 1e-9 for the
  dense LPs.  A relative KKT score bounds residuals, not the objective error
  itself: on the uniform planted LPs, whose objective (~0.1) is a sum of
  terms of size 1e3, a score of 1e-6 leaves an objective error of ~3e-6, so
  those solves run to 1e-8 (error <= 4e-8 on the oracle).
"""
import numpy as np
import pytest

import mcpdlp as mp
import oracle as orc

pytestmark = pytest.mark.gpu


def dense_lp(m, n, seed):
    """Dense rows a_i x <= b_i with b = A x0 + slack for an interior x0, box
    0 <= x <= 10, random costs: feasible and bounded by construction."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(0.1, 10.0, (m, n)) * rng.choice([-1.0, 1.0], (m, n))
    x0 = rng.uniform(1.0, 9.0, n)
    b = A @ x0 + rng.uniform(0.5, 2.0, m)
    c = rng.standard_normal(n)
    return mp.LPData.from_dense(A, c, [-np.inf] * m, b, [0.0] * n, [10.0] * n)


def uniform():
    return mp.generate(mp.FAMILY_UNIFORM_EQ, rows=200, cols=500, seed=11)


def mixed():
    return mp.generate(mp.FAMILY_ZIPF_MIXED, rows=300, cols=700, nnz_max=40, seed=12)


def zipf():  # a Zipf LP the oracle solves to 1e-6 in 5600 iterations (Halpern, FP64)
    return mp.generate(mp.FAMILY_ZIPF_MIXED, rows=150, cols=400, nnz_max=40, seed=16)


def exact_objective(lp):
    """p* of the LP: the planted optimum, or (no planted pair) the oracle's FP64
    objective at tolerance 1e-9."""
    if lp.x_star is not None:
        return lp.planted_objective
    r = orc.solve(lp, tolerance=1e-9, restart_scheme=mp.SCHEME_HALPERN, max_iterations=400000)
    assert r["status"] == "optimal"
    return r["objective"]


def long_rows():  # every row of A holds 2500 nonzeros: longer than any budget
    return dense_lp(3, 2500, 1)


def long_cols():  # every column (row of A^T) holds 2500 nonzeros
    return dense_lp(2500, 6, 2)


MAKERS = [uniform, mixed, long_rows, long_cols]


@pytest.mark.parametrize("budget", [256, 512, 1024, 2048])
@pytest.mark.parametrize("maker", MAKERS)
def test_block_spmv_parity(budget, maker):
    lp = maker()
    print("block spmv", maker.__name__, budget, lp.rows, lp.cols, lp.nnz, flush=True)
    e = mp.Engine(0)
    e.load(lp)
    e.setup(ruiz=False)
    e.set_block_budget(budget, budget)
    (ba, na), (bt, nt) = e.block_tiling()
    assert ba == bt == budget and na > 0 and nt > 0
    rng = np.random.default_rng(budget)
    x = rng.standard_normal(lp.cols)
    x[rng.random(lp.cols) < 0.2] = 0.0
    y = rng.standard_normal(lp.rows)
    ax, at = orc.spmv(lp, x), orc.spmv_t(lp, y)
    f, _ = e.spmv(x)
    ft, _ = e.spmv(y, transpose=True)
    np.testing.assert_allclose(f, ax, rtol=1e-13, atol=1e-13 * np.abs(ax).max())
    np.testing.assert_allclose(ft, at, rtol=1e-13, atol=1e-13 * np.abs(at).max())
    A = np.abs(lp.dense())
    s32, _ = e.spmv(x, fp32=True)
    assert np.all(np.abs(s32 - ax) <= (A @ np.abs(x)) * 2e-6 + 1e-30)
    t32, _ = e.spmv(y, transpose=True, fp32=True)
    assert np.all(np.abs(t32 - at) <= (A.T @ np.abs(y)) * 2e-6 + 1e-30)
    e.close()


def engine_solve(lps, budget, precision, tol=1e-6):
    e = mp.Engine(0)
    e.load(lps)
    e.setup()
    e.set_block_budget(budget, budget)
    e.solve(tolerance=tol, max_iterations=200000, precision=precision, restart_scheme=mp.SCHEME_HALPERN)
    out, _ = e.fetch()
    e.close()
    return out


def check_pair(lp, r, tol=1e-6):
    assert r["status"] == "optimal", r["status"]
    k = orc.kkt(lp, r["x"], r["y"])
    assert k["score"] <= 1.5 * tol, k


# The oracle (Halpern scheme, FP64) converges on these in 5600 (uniform, 1e-8),
# 5600 (zipf, 1e-6), about 6000 (long_rows, 1e-8) and 4000 (long_cols, 1e-8) iterations.
SOLVE_CASES = [(uniform, 1e-8), (zipf, 1e-6), (long_rows, 1e-8), (long_cols, 1e-8)]


@pytest.mark.parametrize("precision", [mp.PRECISION_FP64, mp.PRECISION_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("maker,tol", SOLVE_CASES, ids=[c[0].__name__ for c in SOLVE_CASES])
def test_block_and_wave_steps_reach_the_same_optimum(maker, tol, precision):
    lp = maker()
    print("block solve", maker.__name__, precision, flush=True)
    runs = {b: engine_solve(lp, b, precision, tol)[0] for b in (0, 256, 512, 1024)}
    p_star = exact_objective(lp)
    for b, r in runs.items():
        check_pair(lp, r, tol)
        print("budget", b, "iterations", r["iterations"], "objective error", abs(r["objective"] - p_star), flush=True)
        assert abs(r["objective"] - p_star) <= 1e-6 * (1.0 + abs(p_star))


def test_block_batch_with_compaction():
    """Instances of different sizes finish at different iterations, so the
    batch is re-tiled over the running instances (block tiles included)
    several times; each instance must still converge on its own."""
    lps = ([mp.generate(mp.FAMILY_UNIFORM_EQ, rows=100 + 37 * k, cols=250 + 90 * k, seed=100 + k) for k in range(5)]
           + [long_rows(), long_cols()])
    print("block batch", len(lps), flush=True)
    tol = 1e-8
    on = engine_solve(lps, 1024, mp.PRECISION_FP64, tol)
    off = engine_solve(lps, 0, mp.PRECISION_FP64, tol)
    for lp, a, b in zip(lps, on, off):
        check_pair(lp, a, tol)
        check_pair(lp, b, tol)
        p_star = exact_objective(lp)
        for r in (a, b):
            assert abs(r["objective"] - p_star) <= 1e-6 * (1.0 + abs(p_star))


def permute_columns(lp, perm):
    """The same LP with its columns in the order `perm` (new j = old perm[j])."""
    starts = lp.column_start
    idx = np.concatenate([np.arange(starts[j], starts[j + 1]) for j in perm])
    cs = np.concatenate([[0], np.cumsum(np.diff(starts)[perm])]).astype(starts.dtype)
    return mp.LPData(lp.rows, lp.cols, cs, lp.row_index[idx], lp.value[idx], lp.objective[perm],
                     lp.row_lower.copy(), lp.row_upper.copy(), lp.var_lower[perm], lp.var_upper[perm],
                     lp.sense, lp.offset, None if lp.x_star is None else lp.x_star[perm], lp.y_star,
                     lp.planted_objective)


def test_local_rows_keep_their_order_and_reach_the_same_optimum():
    """The host library keeps the natural row and column order for an LP whose
    rows are local (block-angular: each row inside one or two periods'
    columns) and relabels by length bucket otherwise.  The same LP with its
    columns shuffled is not local and takes the relabelled path; both solves
    reach the tolerance on their own LP and agree on the objective, and the
    natural-order solve matches the planted optimum."""
    tol = 1e-6  # the oracle: 35,880 iterations, objective error 5.6e-9
    lp = mp.generate(mp.FAMILY_BLOCK_ANGULAR, rows=200, cols=500, periods=8, link_rows=400, seed=21)
    shuffled = permute_columns(lp, np.random.default_rng(4).permutation(lp.cols))
    assert mp.rows_are_local(lp) and not mp.rows_are_local(shuffled)
    kw = dict(tolerance=tol, backend=mp.BACKEND_GPU, require_gpu=1, restart_scheme=mp.SCHEME_HALPERN,
              precision=mp.PRECISION_MIXED, max_iterations=400000)
    print("local rows", lp.rows, lp.cols, lp.nnz, flush=True)
    runs = [(q, mp.solve(q, **kw)) for q in (lp, shuffled)]
    for q, r in runs:
        print("iterations", r["iterations"], "objective", r["objective"], flush=True)
        assert r["status"] == "optimal", r["status"]
        assert orc.kkt(q, r["primal"], r["dual"])["score"] <= 1.5 * tol
    p_star = lp.planted_objective
    for _, r in runs:
        assert abs(r["objective"] - p_star) <= 1e-6 * (1.0 + abs(p_star))


def test_block_budget_arguments():
    """Budgets are 256, 512, 1024 or 2048 nonzeros, or 0 for both (block
    tiles off); anything else is refused with an error and leaves the tiling
    as it was."""
    e = mp.Engine(0)
    e.load(uniform())
    e.setup(ruiz=False)
    e.set_block_budget(256, 2048)
    (ba, na), (bt, nt) = e.block_tiling()
    assert (ba, bt) == (256, 2048) and na > 0 and nt > 0
    e.set_block_budget(2048, 512)
    for rows, cols in [(300, 300), (0, 1024), (1024, 0), (4096, 1024), (128, 512), (-1, -1)]:
        with pytest.raises(RuntimeError):
            e.set_block_budget(rows, cols)
    (ba, na), (bt, nt) = e.block_tiling()
    assert (ba, bt) == (2048, 512) and na > 0 and nt > 0
    e.set_block_budget(0, 0)
    assert e.block_tiling() == ((0, 0), (0, 0))
    e.close()
