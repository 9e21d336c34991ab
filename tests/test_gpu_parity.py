"""GPU parity of the HIP engine against the oracle (CPU restatement of the
reference PDLP path).  Run on an MI355X: ``pytest -m gpu``.

Bars:
* reference reduction order (MC_ORDER_REFERENCE, FP64): bit-identical to the
  oracle -- same iteration count, same objective bits, same primal/dual bits;
* fast order (FP64): same status, objective within 1e-9 relative, residuals
  within 1e-6 relative (only the summation order of the reductions differs);
* SpMV: reference order bit-identical, fast order within 1e-13 relative
  (FP64) / 2e-6 relative to the row's absolute sum (FP32).
"""
import numpy as np
import pytest

import mcpdlp as mp
import oracle as orc

pytestmark = pytest.mark.gpu


def tiny_simplex():  # examples/phase3/tiny_simplex.mps; optimum -9 at (1, 3)
    inf = np.inf
    return mp.LPData.from_dense([[1, 1], [2, 1]], [-3, -2], [-inf, -inf], [4, 5], [0, 0], [inf, inf])


def blend():  # tests/pdlp_test.cpp:12-48; optimum -7 at (1, 3)
    inf = np.inf
    return mp.LPData.from_dense([[1, 1], [1, 0], [0, 1]], [-1, -2], [-inf] * 3, [4, 3, 3], [0, 0], [10, 10])


def equality():  # tests/pdlp_test.cpp:82-109; optimum 5
    return mp.LPData.from_dense([[1, 1]], [1, 1], [5], [5], [0, 0], [5, 5])


def small_uniform(seed=11, m=200, n=500):
    return mp.generate(mp.FAMILY_UNIFORM_EQ, rows=m, cols=n, seed=seed)


def small_mixed(seed=12, m=300, n=700):
    return mp.generate(mp.FAMILY_ZIPF_MIXED, rows=m, cols=n, nnz_max=40, seed=seed)


@pytest.fixture(scope="module")
def gpu():
    assert mp.device_count() >= 1, "no HIP device visible"
    return True


def gpu_solve(lp, **kw):
    kw.setdefault("backend", mp.BACKEND_GPU)
    kw.setdefault("require_gpu", 1)
    r = mp.solve(lp, **kw)
    assert r["backend_actually_used"] == "hip", r["backend_actually_used"]
    return r


def test_engine_visible(gpu):
    e = mp.Engine(0)
    e.close()


@pytest.mark.parametrize("maker", [small_uniform, small_mixed])
def test_spmv_parity(gpu, maker):
    lp = maker()
    print("spmv", maker.__name__, lp.rows, lp.cols, lp.nnz, flush=True)
    e = mp.Engine(0)
    e.load(lp)
    e.setup(ruiz=False)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(lp.cols)
    x[rng.random(lp.cols) < 0.2] = 0.0
    y = rng.standard_normal(lp.rows)
    ax, at = orc.spmv(lp, x), orc.spmv_t(lp, y)
    g, _ = e.spmv(x, exact=True)
    gt, _ = e.spmv(y, transpose=True, exact=True)
    assert np.array_equal(g, ax) and np.array_equal(gt, at)
    f, _ = e.spmv(x)
    ft, _ = e.spmv(y, transpose=True)
    np.testing.assert_allclose(f, ax, rtol=1e-13, atol=1e-13 * np.abs(ax).max())
    np.testing.assert_allclose(ft, at, rtol=1e-13, atol=1e-13 * np.abs(at).max())
    A = np.abs(lp.dense())
    bound = A @ np.abs(x) * 2e-6 + 1e-30
    s32, _ = e.spmv(x, fp32=True)
    assert np.all(np.abs(s32 - ax) <= bound + 1e-6 * np.abs(ax))
    e.close()


@pytest.mark.parametrize("maker", [small_uniform, small_mixed])
def test_device_ruiz_bit_exact(gpu, maker):
    lp = maker()
    e = mp.Engine(0)
    e.load(lp)
    e.setup(ruiz=True, ruiz_iterations=10)
    vals, rs, cs, passes, conv = e.scaling(lp.nnz)
    ref = orc.equilibrate(lp, 10)
    assert np.array_equal(vals, ref["lp"].value)
    assert np.array_equal(rs, ref["row_scale"]) and np.array_equal(cs, ref["col_scale"])
    assert passes == ref["passes"] and bool(conv) == ref["converged"]
    # the SpMV of the scaled matrix equals the oracle's on the oracle-scaled matrix
    x = np.linspace(-1, 1, lp.cols)
    g, _ = e.spmv(x, exact=True)
    assert np.array_equal(g, orc.spmv(ref["lp"], x))
    e.close()


CASES = [
    ("tiny_simplex@1e-6", tiny_simplex, dict(tolerance=1e-6)),
    ("blend@1e-4", blend, dict(tolerance=1e-4, max_iterations=200000)),
    ("equality@1e-4", equality, dict(tolerance=1e-4, max_iterations=200000)),
    ("uniform200x500@400it", small_uniform, dict(tolerance=1e-6, max_iterations=400)),
    ("mixed300x700@600it", small_mixed, dict(tolerance=1e-6, max_iterations=600)),
]


@pytest.mark.parametrize("name,maker,kw", CASES, ids=[c[0] for c in CASES])
def test_reference_order_bit_exact(gpu, name, maker, kw):
    lp = maker()
    print("case", name, flush=True)
    want = orc.solve(lp, **kw)
    got = gpu_solve(lp, reduction_order=mp.ORDER_REFERENCE, **kw)
    assert got["status"] == want["status"]
    assert got["iterations"] == want["iterations"]
    assert got["objective"] == want["objective"]
    assert np.array_equal(got["primal"], want["primal"])
    assert np.array_equal(got["dual"], want["dual"])
    assert got["primal_infeasibility"] == want["primal_infeasibility"]
    assert got["dual_infeasibility"] == want["dual_infeasibility"]
    assert got["duality_gap"] == want["duality_gap"]


@pytest.mark.parametrize("name,maker,kw", CASES, ids=[c[0] for c in CASES])
def test_fast_order_close(gpu, name, maker, kw):
    lp = maker()
    want = orc.solve(lp, **kw)
    got = gpu_solve(lp, **kw)
    assert got["status"] == want["status"]
    assert abs(got["iterations"] - want["iterations"]) <= 80
    assert abs(got["objective"] - want["objective"]) <= 1e-9 * (1 + abs(want["objective"]))
    if want["status"] == "optimal":
        assert max(got["primal_infeasibility"], got["dual_infeasibility"], got["duality_gap"]) <= kw["tolerance"]


def test_cfg1_full_shape_reference_order(gpu):
    """m=2000, n=5000, 5 nnz/col equality LP, 2000 iterations: bit-identical trajectory."""
    lp = mp.generate(mp.FAMILY_UNIFORM_EQ)
    want = orc.solve(lp, tolerance=1e-6, max_iterations=2000)
    got = gpu_solve(lp, tolerance=1e-6, max_iterations=2000, reduction_order=mp.ORDER_REFERENCE)
    assert got["iterations"] == want["iterations"] == 2000
    assert got["objective"] == want["objective"]
    assert np.array_equal(got["primal"], want["primal"]) and np.array_equal(got["dual"], want["dual"])


def test_batch_reference_order_matches_single_solves(gpu):
    """Instances finish at different iterations, so the batch is re-tiled
    over the running instances several times (8 -> 4 -> 2 -> 1); every
    instance still follows the oracle's trajectory bit for bit."""
    lps = ([small_uniform(seed=100 + k, m=100 + 37 * k, n=250 + 90 * k) for k in range(5)] + [tiny_simplex(), blend()]
           + [small_mixed(seed=77, m=120, n=260)])
    kw = dict(tolerance=1e-6, max_iterations=3000)
    got = mp.solve_batch(lps, backend=mp.BACKEND_GPU, require_gpu=1, reduction_order=mp.ORDER_REFERENCE, **kw)
    for lp, g in zip(lps, got):
        w = orc.solve(lp, **kw)
        assert g["backend_actually_used"] == "hip"
        assert g["iterations"] == w["iterations"] and g["objective"] == w["objective"]
        assert np.array_equal(g["primal"], w["primal"]) and np.array_equal(g["dual"], w["dual"])


def test_graphs_and_eager_agree(gpu):
    lp = small_mixed()
    kw = dict(tolerance=1e-6, max_iterations=500, reduction_order=mp.ORDER_REFERENCE)
    a = gpu_solve(lp, graph_chunk=0, **kw)
    b = gpu_solve(lp, graph_chunk=-1, **kw)
    c = gpu_solve(lp, graph_chunk=40, **kw)
    for r in (b, c):
        assert r["iterations"] == a["iterations"] and r["objective"] == a["objective"]
        assert np.array_equal(r["primal"], a["primal"])


KKT_CASES = [
    ("tiny_simplex@1e-6", tiny_simplex, dict(tolerance=1e-6)),
    ("blend@1e-6", blend, dict(tolerance=1e-6)),
    ("uniform200x500@1e-6", small_uniform, dict(tolerance=1e-6, max_iterations=3000)),
    ("mixed300x700@1e-6", small_mixed, dict(tolerance=1e-6, max_iterations=3000)),
]


@pytest.mark.parametrize("scheme", [mp.SCHEME_KKT, mp.SCHEME_HALPERN], ids=["kkt", "halpern"])
@pytest.mark.parametrize("name,maker,kw", KKT_CASES, ids=[c[0] for c in KKT_CASES])
def test_kkt_scheme_reference_order_bit_exact(gpu, name, maker, kw, scheme):
    """KKT and Halpern restart schemes (extensions): GPU reference order ==
    oracle, bit for bit."""
    lp = maker()
    print("scheme", scheme, "case", name, flush=True)
    want = orc.solve(lp, restart_scheme=scheme, **kw)
    got = gpu_solve(lp, restart_scheme=scheme, reduction_order=mp.ORDER_REFERENCE, **kw)
    assert got["status"] == want["status"] and got["iterations"] == want["iterations"]
    assert got["objective"] == want["objective"]
    assert np.array_equal(got["primal"], want["primal"]) and np.array_equal(got["dual"], want["dual"])


@pytest.mark.parametrize("scheme", [mp.SCHEME_KKT, mp.SCHEME_HALPERN], ids=["kkt", "halpern"])
@pytest.mark.parametrize("precision", [mp.PRECISION_FP64, mp.PRECISION_MIXED])
def test_kkt_scheme_converges_to_1e6(gpu, precision, scheme):
    """Fast order, FP64 or mixed: reaches 1e-6 on a planted LP; the returned pair
    is re-scored independently by the oracle's FP64 residuals on the host.  The
    objective is held to 1e-6 x (1 + |p*|) of the planted optimum p* = c^T x*
    on a solve to 1e-8: a KKT score of 1e-6 bounds the residuals, not |p - p*|
    itself, and this LP's objective (~0.12) is a sum of terms of size 1e3 (the
    oracle's error at 1e-6 is 2.6e-6, at 1e-8 at most 4e-8)."""
    lp = small_uniform()  # the oracle converges on it in 8720 iterations (FP64, KKT scheme)
    got = gpu_solve(lp, tolerance=1e-6, restart_scheme=scheme, precision=precision,
                    max_iterations=200000)
    assert got["status"] == "optimal", got["message"]
    k = orc.kkt(lp, got["primal"], got["dual"])
    assert k["score"] <= 1.5e-6, k
    if precision == mp.PRECISION_MIXED:
        # latency-bound (25k nonzeros): mixed starts on the packed stream, no FP32 phase
        assert got["fp32_iterations"] == 0 and got["packed_iterations"] > 0
    tight = gpu_solve(lp, tolerance=1e-8, restart_scheme=scheme, precision=precision, max_iterations=200000)
    assert tight["status"] == "optimal", tight["message"]
    assert orc.kkt(lp, tight["primal"], tight["dual"])["score"] <= 1.5e-8
    p_star = lp.planted_objective
    assert abs(tight["objective"] - p_star) <= 1e-6 * (1.0 + abs(p_star)), (tight["objective"], p_star)


# scripts/run_netlib.py:20-93 optima of the Netlib instances in tests/golden/netlib
NETLIB = {"afiro": -464.753142857143, "adlittle": 225494.96316238, "recipe": -266.616}


def netlib(name):
    import os
    return mp.read_mps(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "netlib", f"{name}.mps"))


@pytest.mark.parametrize("name,scheme", [("afiro", 0), ("recipe", 0), ("adlittle", 1), ("recipe", 1),
                                         ("adlittle", 2), ("recipe", 2)])
def test_netlib_reference_order_bit_exact(gpu, name, scheme):
    """Real LPs (free rows, ranges, bounds of every kind): the GPU reference
    order reproduces the oracle's whole solve bit for bit."""
    lp = netlib(name)
    print("netlib case", name, scheme, flush=True)
    kw = dict(tolerance=1e-6, restart_scheme=scheme, max_iterations=100000)
    want = orc.solve(lp, **kw)
    got = gpu_solve(lp, reduction_order=mp.ORDER_REFERENCE, **kw)
    assert got["status"] == want["status"] == "optimal" and got["iterations"] == want["iterations"]
    assert got["objective"] == want["objective"]
    assert np.array_equal(got["primal"], want["primal"]) and np.array_equal(got["dual"], want["dual"])


@pytest.mark.parametrize("scheme", [mp.SCHEME_KKT, mp.SCHEME_HALPERN], ids=["kkt", "halpern"])
@pytest.mark.parametrize("name", sorted(NETLIB))
@pytest.mark.parametrize("precision", [mp.PRECISION_FP64, mp.PRECISION_MIXED])
def test_netlib_fast_order_reaches_published_optimum(gpu, name, precision, scheme):
    lp = netlib(name)
    print("netlib fast", name, precision, scheme, flush=True)
    r = gpu_solve(lp, tolerance=1e-6, restart_scheme=scheme, precision=precision, max_iterations=200000)
    assert r["status"] == "optimal"
    assert abs(r["objective"] - NETLIB[name]) <= 1e-6 * abs(NETLIB[name])
    k = orc.kkt(lp, r["primal"], r["dual"])
    assert k["score"] <= 1e-6 * 1.01


def test_repeated_solves_follow_their_own_options(gpu):
    """Solving one LP again on the same engine with other options (tolerance,
    polish threshold) must use those options, not the ones the cached HIP
    graphs were captured with; a repeat of the first solve is bit-identical."""
    lp = small_uniform(seed=21)
    kw = dict(restart_scheme=mp.SCHEME_HALPERN, max_iterations=200000)
    a = gpu_solve(lp, tolerance=1e-3, **kw)
    b = gpu_solve(lp, tolerance=1e-6, **kw)
    c = gpu_solve(lp, tolerance=1e-3, **kw)
    assert a["status"] == b["status"] == c["status"] == "optimal"
    assert b["iterations"] > a["iterations"]
    assert orc.kkt(lp, b["primal"], b["dual"])["score"] <= 1.5e-6
    assert c["iterations"] == a["iterations"] and np.array_equal(c["primal"], a["primal"])
    m1 = gpu_solve(lp, tolerance=1e-6, precision=mp.PRECISION_PACKED, polish_threshold=1e-3, **kw)
    m2 = gpu_solve(lp, tolerance=1e-6, precision=mp.PRECISION_PACKED, polish_threshold=0.0, **kw)
    assert m1["status"] == m2["status"] == "optimal"
    # m1 leaves the packed stream at a score of 1e-3, m2 only near its floor
    assert m1["packed_iterations"] < m2["packed_iterations"]


def test_repeated_solves_follow_adaptive_step_option(gpu):
    """adaptive_step_size decides whether the captured iteration chunks hold the
    A dx / eta-update launches.  Flipping only that option between solves of one
    LP on the same engine (1, 0, 1) must follow each solve's own setting: every
    run is bit-identical to the oracle with the same options, in reference order."""
    lp = small_uniform(seed=23)
    for adaptive in (1, 0, 1):
        kw = dict(tolerance=1e-6, max_iterations=700, adaptive_step_size=adaptive)
        print("adaptive_step_size", adaptive, flush=True)
        want = orc.solve(lp, **kw)
        got = gpu_solve(lp, reduction_order=mp.ORDER_REFERENCE, **kw)
        assert got["iterations"] == want["iterations"] and got["objective"] == want["objective"]
        assert np.array_equal(got["primal"], want["primal"]) and np.array_equal(got["dual"], want["dual"])
        assert got["step_size"] == want["step_size"]


def _infeasibility_cases():
    import test_host
    return test_host._infeasibility_cases()


@pytest.mark.parametrize("precision", [mp.PRECISION_FP64, mp.PRECISION_MIXED])
@pytest.mark.parametrize("scheme", [mp.SCHEME_KKT, mp.SCHEME_HALPERN], ids=["kkt", "halpern"])
def test_gpu_infeasibility_detection(gpu, scheme, precision):
    """Extension: the device ray-certificate kernels stop infeasible and
    unbounded LPs with infeasible_or_unbounded; on a feasible LP they only
    read the iterate (the solve is bit-identical to one without them); in a
    batch each instance gets its own status."""
    cases, base = _infeasibility_cases()
    kw = dict(tolerance=1e-6, restart_scheme=scheme, precision=precision, max_iterations=100000)
    for name, lp in cases.items():
        print("infeasibility case", name, flush=True)
        r = gpu_solve(lp, detect_infeasibility=1, **kw)
        want = orc.solve(lp, detect_infeasibility=1, **kw)  # the oracle's restatement of the test
        assert r["status"] == want["status"] == "infeasible_or_unbounded", (name, r["status"], r["iterations"])
    a = gpu_solve(base, detect_infeasibility=1, **kw)
    b = gpu_solve(base, **kw)
    assert a["status"] == "optimal" and a["iterations"] == b["iterations"]
    assert np.array_equal(a["primal"], b["primal"])
    lps = [base] + list(cases.values()) + [small_uniform(seed=5)]
    got = mp.solve_batch(lps, backend=mp.BACKEND_GPU, require_gpu=1, detect_infeasibility=1, **kw)
    assert [g["status"] for g in got] == ["optimal"] + ["infeasible_or_unbounded"] * len(cases) + ["optimal"]


def _golden_runs():
    import test_oracle
    return test_oracle.INPUTS, test_oracle.GOLD["runs"]


@pytest.mark.parametrize("scheme", [mp.SCHEME_REFERENCE, mp.SCHEME_KKT, mp.SCHEME_HALPERN],
                         ids=["reference", "kkt", "halpern"])
def test_golden_inputs_reference_order_bit_exact(gpu, scheme):
    """Every input the compiled reference wrote golden fixtures for (maximize
    sense, ranged rows, free variables, open '>=' rows, one-variable and
    uniform/Zipf LPs), with each fixture's options: the GPU in reference order
    reproduces the oracle bit for bit -- and, for the reference scheme, the
    compiled reference's own iteration count and objective bits.

    One exception: a primal_weight_smoothing other than 0.5 makes the weight
    update call pow(r, theta), and the device pow and the host libm pow may
    round differently in the last place (theta = 0.5 uses the correctly
    rounded sqrt on both).  Those runs must keep the iteration count and agree
    to 1e-11 relative."""
    inputs, runs = _golden_runs()
    for run in runs:
        lp = inputs[run["case"]]
        kw = dict(run["options"], restart_scheme=scheme)
        print("golden", run["case"], kw, flush=True)
        want = orc.solve(lp, **kw)
        got = gpu_solve(lp, reduction_order=mp.ORDER_REFERENCE, **kw)
        assert got["status"] == want["status"] and got["iterations"] == want["iterations"], run["case"]
        if kw.get("primal_weight_smoothing", 0.5) != 0.5:
            assert abs(got["objective"] - want["objective"]) <= 1e-11 * abs(want["objective"]), run["case"]
            assert np.allclose(got["primal"], want["primal"], rtol=1e-9, atol=1e-9)
            continue
        assert got["objective"] == want["objective"], run["case"]
        assert np.array_equal(got["primal"], want["primal"]) and np.array_equal(got["dual"], want["dual"])
        if scheme == mp.SCHEME_REFERENCE:
            assert got["iterations"] == run["iterations"]
            assert got["objective"] == float.fromhex(run["objective"])


def test_cli_gpu_backend(gpu):
    """markov-cero-solve --engine pdlp --backend gpu: the JSON and the stderr
    verdict name the HIP path; reference order reproduces the CPU backend's
    answer to the bit (tests/golden/tiny_simplex.mps, ADR-P3-05 known answer)."""
    import json
    import os
    import subprocess
    mps = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tiny_simplex.mps")

    def run(*extra):
        p = subprocess.run([mp.CLI, mps, "--engine", "pdlp", "--tolerance", "1e-6", *extra],
                           capture_output=True, text=True, timeout=120)
        return p.returncode, json.loads(p.stdout.strip().splitlines()[-1]), p.stderr

    code, js, err = run("--backend", "gpu", "--order", "reference")
    assert code == 0 and js["status"] == "Optimal" and js["verified"]
    assert js["backend_actually_used"] == "hip" and "backend_actually_used=hip" in err
    code_c, js_c, _ = run("--backend", "cpu")
    assert js["objective"] == js_c["objective"] and js["primal"] == js_c["primal"]
    assert round(js["objective"], 7) == -9.0000013
    # the packed stream from the CLI: same optimum, every iteration counted on it or in FP64
    code_p, js_p, err_p = run("--backend", "gpu", "--precision", "packed", "--restart-scheme", "halpern")
    assert code_p == 0 and js_p["status"] == "Optimal" and js_p["verified"] and js_p["precision"] == "packed"
    assert js_p["backend_actually_used"] == "hip" and 0 < js_p["packed_iterations"] <= js_p["lp_iterations"]
    assert abs(js_p["objective"] - (-9.0)) <= 1e-5
