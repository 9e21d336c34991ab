"""Pins the oracle (oracle/pdlp_oracle.c) -- CPU only.

Each check is chosen so a plausible restatement mistake (dropped term, wrong
sign/index, transposed operand, changed operation order) fails it:
* golden fixtures written by the compiled reference (tests/golden): iteration
  counts, objective/residual bits, primal/dual bits, Ruiz scaler hashes;
* the compiled reference itself, live, where oracle/_ref is built;
* closed forms: dense long-double products, exact integer SpMV, the Ruiz
  identity A_s = D_r A D_c and its convergence property, a diagonal matrix
  equilibrated in one pass, zero KKT at the planted pair, hand residuals;
* known answers held by the reference: ADR-P3-05 (tiny_simplex @1e-6),
  tests/pdlp_test.cpp objectives, verify_primal verdicts, and the Netlib
  optima of scripts/run_netlib.py:20-93 for the instances its data/netlib
  holds (afiro, adlittle, recipe; files copied to tests/golden/netlib).
"""
import json
import os

import numpy as np
import pytest

import mcpdlp as mp
import oracle as orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_inputs():
    z = np.load(os.path.join(GOLDEN, "pdlp_inputs.npz"))
    names = sorted({k.split(".")[0] for k in z.files})
    out = {}
    for nm in names:
        m, n, sense = (int(v) for v in z[f"{nm}.meta"])
        b = {}
        for k in ("row_lower", "row_upper", "var_lower", "var_upper"):
            b[k] = mp.LPData.bounds(z[f"{nm}.{k}.kind"], z[f"{nm}.{k}.value"])
        out[nm] = mp.LPData(m, n, z[f"{nm}.column_start"].astype(np.int64), z[f"{nm}.row_index"].astype(np.int64),
                            z[f"{nm}.value"], z[f"{nm}.objective"], b["row_lower"], b["row_upper"],
                            b["var_lower"], b["var_upper"], sense, float(z[f"{nm}.offset"][0]))
    return out


INPUTS = load_inputs()
with open(os.path.join(GOLDEN, "pdlp_golden.json")) as _f:
    GOLD = json.load(_f)


def unhex(xs):
    return np.array([float.fromhex(v) for v in xs])


@pytest.mark.parametrize("run", GOLD["runs"], ids=[f"{r['case']}-{i}" for i, r in enumerate(GOLD["runs"])])
def test_oracle_matches_reference_golden(run):
    lp = INPUTS[run["case"]]
    got = orc.solve(lp, **run["options"])
    assert got["status"] == run["status"]
    assert got["iterations"] == run["iterations"]
    assert got["objective"] == float.fromhex(run["objective"])
    assert got["primal_infeasibility"] == float.fromhex(run["primal_infeasibility"])
    assert got["dual_infeasibility"] == float.fromhex(run["dual_infeasibility"])
    assert got["duality_gap"] == float.fromhex(run["duality_gap"])
    assert np.array_equal(got["primal"], unhex(run["primal"]))
    assert np.array_equal(got["dual"], unhex(run["dual"]))
    assert got["message"] == run["message"] and got["backend_actually_used"] == run["backend_actually_used"]
    if run["verify_1e-4"] is not None:
        v = orc.verify(lp, got["primal"], got["objective"], tol=1e-4)
        assert v["passed"] == run["verify_1e-4"]["passed"]
        assert v["recomputed_objective"] == float.fromhex(run["verify_1e-4"]["recomputed_objective"])


@pytest.mark.parametrize("name", sorted(GOLD["ruiz"]))
def test_oracle_ruiz_matches_reference_golden(name):
    import hashlib

    q = orc.equilibrate(INPUTS[name], 10)
    g = GOLD["ruiz"][name]
    h = lambda a: hashlib.sha256(np.ascontiguousarray(a, np.float64).tobytes()).hexdigest()
    assert q["passes"] == g["passes"] and q["converged"] == g["converged"]
    assert h(q["lp"].value) == g["values_sha256"]
    assert h(q["row_scale"]) == g["row_scale_sha256"] and h(q["col_scale"]) == g["col_scale_sha256"]


@pytest.mark.skipif(not orc.ref_available(), reason="compiled reference not built here")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_matches_compiled_reference_live(seed):
    lp = mp.generate(mp.FAMILY_ZIPF_MIXED, rows=150 + 20 * seed, cols=400, nnz_max=30, seed=seed)
    for kw in (dict(tolerance=1e-6, max_iterations=300), dict(tolerance=1e-3, max_iterations=20000)):
        a, b = orc.solve(lp, **kw), orc.ref_solve(lp, **kw)
        assert a["iterations"] == b["iterations"] and a["objective"] == b["objective"]
        assert np.array_equal(a["primal"], b["primal"]) and np.array_equal(a["dual"], b["dual"])
    qa, qb = orc.equilibrate(lp, 10), orc.equilibrate(lp, 10, use_ref=True)
    assert np.array_equal(qa["lp"].value, qb["lp"].value)
    assert np.array_equal(qa["row_scale"], qb["row_scale"]) and np.array_equal(qa["col_scale"], qb["col_scale"])


# ---- closed forms -----------------------------------------------------------
def random_lp(seed, m=7, n=5, integer=False):
    rng = np.random.default_rng(seed)
    A = np.where(rng.random((m, n)) < 0.45, rng.integers(-4, 5, (m, n)) if integer else rng.normal(size=(m, n)), 0.0)
    inf = np.inf
    return mp.LPData.from_dense(A, rng.normal(size=n), [-inf] * m, [inf] * m, [0] * n, [inf] * n), A


@pytest.mark.parametrize("seed", range(25))
def test_spmv_against_dense_long_double(seed):
    lp, A = random_lp(seed)
    rng = np.random.default_rng(1000 + seed)
    x = rng.normal(size=lp.cols)
    x[rng.random(lp.cols) < 0.3] = 0.0
    y = rng.normal(size=lp.rows)
    Al = A.astype(np.longdouble)
    np.testing.assert_allclose(orc.spmv(lp, x), (Al @ x.astype(np.longdouble)).astype(float), rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(orc.spmv_t(lp, y), (Al.T @ y.astype(np.longdouble)).astype(float), rtol=1e-14,
                               atol=1e-14)


@pytest.mark.parametrize("seed", range(10))
def test_spmv_exact_on_integers(seed):
    lp, A = random_lp(seed, 9, 6, integer=True)
    x = np.arange(lp.cols, dtype=float) - 2.0
    y = np.arange(lp.rows, dtype=float) * 3 - 7
    assert np.array_equal(orc.spmv(lp, x), A @ x)
    assert np.array_equal(orc.spmv_t(lp, y), A.T @ y)


def test_ruiz_identity_and_convergence():
    lp = mp.generate(mp.FAMILY_ZIPF_MIXED, rows=200, cols=500, nnz_max=30, seed=4)
    q = orc.equilibrate(lp, 60)
    cols = np.repeat(np.arange(lp.cols), np.diff(lp.column_start))
    # A_s = D_r A D_c (product of per-pass deltas; rounding of up to 60 products)
    np.testing.assert_allclose(q["lp"].value, lp.value * q["row_scale"][lp.row_index] * q["col_scale"][cols],
                               rtol=1e-12)
    assert q["converged"]
    a = np.abs(q["lp"].value)
    rmax = np.zeros(lp.rows)
    np.maximum.at(rmax, lp.row_index, a)
    cmax = np.maximum.reduceat(a, lp.column_start[:-1])
    assert np.all(np.abs(rmax[rmax > 0] - 1) < 1e-3) and np.all(np.abs(cmax - 1) < 1e-3)
    # bounds scaled by D_r (rows) and D_c^-1 (columns), objective by D_c
    fin = lp.var_upper["kind"] == mp.FINITE
    np.testing.assert_array_equal(q["lp"].var_upper["value"][fin],
                                  lp.var_upper["value"][fin] * (1.0 / q["col_scale"][fin]))
    np.testing.assert_array_equal(q["lp"].objective, lp.objective * q["col_scale"])


def test_ruiz_diagonal_one_pass():
    d = np.array([4.0, 0.25, 9.0, 1.0])
    lp = mp.LPData.from_dense(np.diag(d), np.ones(4), [1] * 4, [1] * 4, [0] * 4, [np.inf] * 4)
    q = orc.equilibrate(lp, 10)
    np.testing.assert_allclose(np.abs(q["lp"].value), 1.0, rtol=1e-15)
    assert q["passes"] == 1 and q["converged"]


def test_residuals_zero_at_planted_pair():
    for fam, kw in ((mp.FAMILY_UNIFORM_EQ, dict(rows=300, cols=800)),
                    (mp.FAMILY_ZIPF_MIXED, dict(rows=300, cols=800, nnz_max=50))):
        lp = mp.generate(fam, seed=9, **kw)
        k = orc.kkt(lp, lp.x_star, -lp.y_star)  # reported dual = -y* (g = c + A^T y convention)
        assert k["score"] < 1e-11, k
        assert abs(k["objective"] - lp.planted_objective) <= 1e-12 * abs(lp.planted_objective)
        bad = orc.kkt(lp, lp.x_star, lp.y_star)  # wrong sign must be caught
        assert bad["score"] > 1e-3


def test_residuals_hand_case():
    # min x1 + 2 x2  s.t.  x1 + x2 >= 2 (row 0), x1 <= 3 (row 1);  x1 in [0, inf), x2 in [0, 1]
    inf = np.inf
    lp = mp.LPData.from_dense([[1, 1], [1, 0]], [1, 2], [2, -inf], [inf, 3], [0, 0], [inf, 1])
    # x = (1, 0.5): row0 = 1.5 < 2 -> violation 0.5 / (1 + 2) ; y = (-1, 0)
    k = orc.kkt(lp, [1.0, 0.5], [-1.0, 0.0])
    assert k["primal"] == pytest.approx(0.5 / 3.0, rel=1e-15)
    # g = c + A^T y = (0, 1): x2 - clamp(x2 - 1, 0, 1) = 0.5 -> dual = 0.5 / max(1, |c|inf=2)
    assert k["dual"] == pytest.approx(0.25, rel=1e-15)
    # p = 1*1 + 2*0.5 = 2; d = -(-1)*2 [y<0 uses row lower] + g2>0 * lo2 = 2 ; gap 0
    assert k["objective"] == 2.0 and k["gap"] == 0.0
    # maximize quirk (pdlp.cpp:113-131): primal objective unsigned, dual signed
    lpm = mp.LPData.from_dense([[1.0]], [1.0], [-inf], [1.0], [0], [1.0], sense=mp.MAXIMIZE)
    km = orc.kkt(lpm, [1.0], [1.0])
    assert km["objective"] == 1.0 and km["gap"] == pytest.approx(2.0 / 3.0)


# ---- known answers held by the reference ------------------------------------
def test_adr_p3_05_tiny_simplex():
    """docs/decisions/ADR-P3-05-pdhg-pdlp.md: @1e-6 Optimal, obj ~ -9.0000013, x ~ (1.000008, 2.999989)."""
    r = orc.solve(INPUTS["tiny_simplex"], tolerance=1e-6)
    assert r["status"] == "optimal" and r["message"] == "PDLP converged"
    assert round(r["objective"], 7) == -9.0000013
    assert r["primal"][0] == pytest.approx(1.000008, abs=5e-7) and r["primal"][1] == pytest.approx(2.999989, abs=5e-7)
    assert orc.verify(INPUTS["tiny_simplex"], r["primal"], r["objective"], 1e-6)["passed"]


def test_pdlp_test_objectives():
    """tests/pdlp_test.cpp:62 (-7 +- 0.1) and :119 (5 +- 0.1), verify at 1e-4."""
    for name, want in (("blend", -7.0), ("equality", 5.0)):
        r = orc.solve(INPUTS[name], tolerance=1e-4, max_iterations=200000)
        assert abs(r["objective"] - want) < 0.1
        assert r["status"] == "optimal" and orc.verify(INPUTS[name], r["primal"], r["objective"], 1e-4)["passed"]


def test_kkt_scheme_reaches_planted_optimum():
    """Extension scheme in the oracle: converges to tolerance, objective = planted optimum."""
    lp = mp.generate(mp.FAMILY_UNIFORM_EQ, rows=200, cols=500, seed=11)
    r = orc.solve(lp, tolerance=1e-6, restart_scheme=mp.SCHEME_KKT, max_iterations=100000)
    assert r["status"] == "optimal"
    k = orc.kkt(lp, r["primal"], r["dual"])
    assert k["score"] <= 1e-6 * 1.01
    assert abs(r["objective"] - lp.planted_objective) <= 1e-5 * (1 + abs(lp.planted_objective))


# scripts/run_netlib.py:20-93 (published Netlib optima); the MPS files are the
# reference's data/netlib copies.
NETLIB = {"afiro": -464.753142857143, "adlittle": 225494.96316238, "recipe": -266.616}


def netlib(name):
    return mp.read_mps(os.path.join(GOLDEN, "netlib", f"{name}.mps"))


@pytest.mark.parametrize("name,scheme", [("afiro", 0), ("recipe", 0), ("afiro", 1), ("adlittle", 1), ("recipe", 1),
                                         ("afiro", 2), ("adlittle", 2), ("recipe", 2)])
def test_netlib_optima(name, scheme):
    """The oracle reaches each published optimum (1e-6 KKT, objective within
    1e-6 relative) -- under the reference's restart rule where that rule
    converges in 400k iterations, and under the KKT and Halpern schemes."""
    lp = netlib(name)
    r = orc.solve(lp, tolerance=1e-6, restart_scheme=scheme, max_iterations=400000)
    assert r["status"] == "optimal"
    assert abs(r["objective"] - NETLIB[name]) <= 1e-6 * abs(NETLIB[name])
    assert orc.verify(lp, r["primal"], r["objective"], 1e-6)["passed"]


def test_netlib_adlittle_reference_rule_stalls():
    """adlittle under the reference's own restart rule: no 1e-6 in 400k
    iterations (the dual residual stays near 2.6e-5) -- the behaviour the
    KKT restart scheme exists to fix."""
    r = orc.solve(netlib("adlittle"), tolerance=1e-6, max_iterations=400000)
    assert r["status"] == "iteration_limit" and r["dual_infeasibility"] > 1e-5


def test_halpern_scheme_reaches_planted_optimum():
    """Halpern scheme (extension) in the oracle: converges to tolerance on a
    uniform and a Zipf LP, the pair re-scored independently is within
    tolerance and the objective equals the planted optimum."""
    for lp in (mp.generate(mp.FAMILY_UNIFORM_EQ, rows=200, cols=500, seed=11),
               mp.generate(mp.FAMILY_ZIPF_MIXED, rows=300, cols=700, nnz_max=40, seed=31)):
        r = orc.solve(lp, tolerance=1e-6, restart_scheme=mp.SCHEME_HALPERN, max_iterations=200000)
        assert r["status"] == "optimal"
        k = orc.kkt(lp, r["primal"], r["dual"])
        assert k["score"] <= 1e-6 * 1.01
        assert abs(r["objective"] - lp.planted_objective) <= 1e-5 * (1 + abs(lp.planted_objective))


def test_halpern_iterates_by_hand():
    """Three Halpern steps worked by hand on min x s.t. x = 1, 0 <= x <= 2
    (no scaling, omega = 1, eta = 0.9 fixed, so tau = sigma = 0.9), from
    z0 = (0, 0); T is the PDHG step, z_k+1 = w (2 T(z_k) - z_k) + (1 - w) z0
    with w = (k+1)/(k+2):
      k=0: T = (0, -0.9),    w = 1/2: z1 = (0, -0.9)
      k=1: T = (0, -1.8),    w = 2/3: z2 = (0, -1.8)
      k=2: T = (0.72, -1.404), x_bar = 1.44, w = 3/4:
           z3 = (3/4 * 1.44, 3/4 * (2 * -1.404 + 1.8)) = (1.08, -0.756).
    z3 has the lower KKT score of the two candidates, so it is returned.  The
    plain PDHG schemes return other points here (a mistake in w, in the
    reflection or in the anchor changes z3)."""
    lp = mp.LPData.from_dense([[1.0]], [1.0], [1.0], [1.0], [0.0], [2.0])
    kw = dict(tolerance=1e-12, ruiz_scaling=0, initial_primal_weight=1.0, adaptive_step_size=0)
    want = {1: (0.0, -0.9), 2: (0.0, -1.8), 3: (1.08, -0.756)}
    for it, (x, y) in want.items():
        r = orc.solve(lp, restart_scheme=mp.SCHEME_HALPERN, max_iterations=it, **kw)
        assert r["primal"][0] == pytest.approx(x, rel=1e-14, abs=1e-15)
        assert r["dual"][0] == pytest.approx(y, rel=1e-14)
    r = orc.solve(lp, restart_scheme=mp.SCHEME_KKT, max_iterations=3, **kw)
    assert r["primal"][0] != pytest.approx(1.08)


def _csc_lp(A):
    A = np.asarray(A, dtype=float)
    m, n = A.shape
    inf = np.inf
    return mp.LPData.from_dense(A, np.zeros(n), [-inf] * m, [inf] * m, [-inf] * n, [inf] * n)


def test_step_cap_closed_forms():
    """or_step_cap: L = ||D_r^-1/2 A D_c^-1/2|| by the power method, cap 0.95 / L.
    A diagonal matrix of any signs and sizes maps to a signed identity (L = 1);
    [[1, 1], [1, -1]] maps to A / 2, of norm 1/sqrt(2); a positive rank-one
    matrix a b^T maps to sqrt(a_i b_j) / sqrt(sum a sum b), of norm 1.  No
    iterations (or an all-zero matrix) leaves the reference's cap 0.99."""
    assert orc.step_cap(_csc_lp(np.diag([3.0, -0.25, 7.0, -40.0])), 20) == pytest.approx(0.95, rel=1e-15)
    assert orc.step_cap(_csc_lp([[1.0, 1.0], [1.0, -1.0]]), 5) == pytest.approx(0.95 * np.sqrt(2.0), rel=1e-15)
    a, b = np.array([1.0, 2.0, 5.0]), np.array([0.5, 3.0, 1.0, 4.0])
    assert orc.step_cap(_csc_lp(np.outer(a, b)), 20) == pytest.approx(0.95, rel=1e-12)
    assert orc.step_cap(_csc_lp([[1.0, 1.0], [1.0, -1.0]]), 0) == 0.99


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_step_cap_matches_the_singular_value(seed):
    """On seeded LPs the estimate is the largest singular value of the
    preconditioned matrix (dense SVD) from below: within 1e-6 after 400
    iterations, within 10 % after 20 (the count the solve uses by default)."""
    lp = mp.generate(mp.FAMILY_ZIPF_MIXED, rows=60, cols=150, nnz_max=30, seed=seed)
    A = lp.dense()
    r = np.abs(A).sum(1)
    c = np.abs(A).sum(0)
    r[r < 1e-12] = 1.0
    c[c < 1e-12] = 1.0
    sigma = np.linalg.svd(A / np.sqrt(r)[:, None] / np.sqrt(c)[None, :], compute_uv=False)[0]
    L400 = 0.95 / orc.step_cap(lp, 400)
    L20 = 0.95 / orc.step_cap(lp, 20)
    assert abs(L400 - sigma) <= 1e-6 * sigma
    assert 0.9 * sigma <= L20 <= sigma * (1 + 1e-12)


def test_halpern_step_size_respects_the_cap():
    """Under the Halpern scheme eta grows up to 0.95 / L instead of 0.99 (the
    other schemes keep the reference's 0.99), and never above it."""
    lp = mp.generate(mp.FAMILY_UNIFORM_EQ, rows=200, cols=500, seed=11)
    q = orc.equilibrate(lp, 10)["lp"]
    cap = orc.step_cap(q, 20)
    assert cap > 0.99
    h = orc.solve(lp, tolerance=1e-6, restart_scheme=mp.SCHEME_HALPERN, trace_cap=4096)
    k = orc.solve(lp, tolerance=1e-6, restart_scheme=mp.SCHEME_KKT, trace_cap=4096)
    assert h["status"] == k["status"] == "optimal"
    assert 0.99 < h["trace"]["eta"].max() <= cap
    assert k["trace"]["eta"].max() <= 0.99


def _weight_closed_form(omega, dx, dy, p, d):
    L = np.longdouble
    om, ratio = L(omega), L(dy) / L(dx)
    move = ratio ** (L(127) / L(128)) * om ** (L(1) / L(128)) if dx > 1e-10 and dy > 1e-10 else om
    res = om * np.sqrt(np.clip(np.sqrt(L(p) / L(d)), L(0.05), L(20))) if p > 1e-12 and d > 1e-12 else om
    return move ** L(0.75) * res ** L(0.25)


@pytest.mark.parametrize("omega,dx,dy,p,d", [
    (0.3, 2.0, 0.05, 1e-4, 3e-5), (5.0, 1e-3, 7.0, 2e-6, 8e-6), (1.0, 1.0, 1.0, 1e-5, 1e-5),
    (0.01, 0.4, 0.4, 3e-3, 1e-4), (2.5, 1e-12, 3.0, 1e-6, 2e-6), (0.7, 3.0, 0.2, 0.0, 1e-6)])
def test_halpern_weight_closed_form(omega, dx, dy, p, d):
    """omega_move^(3/4) omega_res^(1/4) with omega_move = (dY/dX)^(127/128)
    omega^(1/128) (omega when a movement is below 1e-10) and omega_res =
    omega sqrt(clamp(sqrt(p/d), 0.05, 20)) (omega when a residual is below
    1e-12), evaluated in long double from the powers themselves."""
    want = float(_weight_closed_form(omega, dx, dy, p, d))
    assert orc.halpern_weight(omega, dx, dy, p, d) == pytest.approx(want, rel=1e-14)


def test_halpern_weight_guard_and_clamp():
    """No move against a residual imbalance of 100x or more; [1e-6, 1e6]."""
    # primal 1000x the dual and a movement that would lower omega: omega stays
    assert orc.halpern_weight(0.5, 1.0, 1e-3, 1e-3, 1e-6) == 0.5
    # dual 1000x the primal and a movement that would raise omega: omega stays
    assert orc.halpern_weight(0.5, 1e-3, 1.0, 1e-9, 1e-6) == 0.5
    # the same imbalance with the movement in its favour moves omega
    assert orc.halpern_weight(0.5, 1e-3, 1.0, 1e-3, 1e-6) > 0.5
    assert orc.halpern_weight(1e-6, 1.0, 1e-9, 1e-6, 1e-6) == 1e-6
    assert orc.halpern_weight(1e6, 1e-9, 1.0, 1e-6, 1e-6) == 1e6


def test_ray_certificate_hand_cases():
    """or_ray_certificate on LPs whose certificates are worked by hand.
    x1 + x2 = 1 and x1 + x2 = 2, x >= 0: dy = (1, -1) gives A^T dy = 0 and
    D = -1*1 - (-1)*2 = 1 > 0 (primal infeasible); dy = (-1, 1) gives D = -1.
    min -x1 s.t. x1 - x2 = 0, x >= 0: dx = (1, 1) keeps A dx = 0 and x >= 0
    with c^T dx = -1 (unbounded); dx = (1, 0) violates the row by 1; dx =
    (-1, -1) leaves the bounds.  A dual direction that would need a missing
    bound is rejected, and the maximize sense flips the cost."""
    inf = np.inf
    two = mp.LPData.from_dense([[1, 1], [1, 1]], [1, 1], [1, 2], [1, 2], [0, 0], [inf, inf])
    assert orc.ray_certificate(two, [0, 0], [1, -1]) == 1
    assert orc.ray_certificate(two, [0, 0], [-1, 1]) == 0
    assert orc.ray_certificate(two, [0, 0], [0, 0]) == 0
    unb = mp.LPData.from_dense([[1, -1]], [-1, 0], [0], [0], [0, 0], [inf, inf])
    assert orc.ray_certificate(unb, [1, 1], [0]) == 2
    assert orc.ray_certificate(unb, [1, 0], [0]) == 0
    assert orc.ray_certificate(unb, [-1, -1], [0]) == 0
    # maximize sense flips the cost: max -x1 along (1, 1) is bounded
    unb_max = mp.LPData.from_dense([[1, -1]], [-1, 0], [0], [0], [0, 0], [inf, inf], sense=mp.MAXIMIZE)
    assert orc.ray_certificate(unb_max, [1, 1], [0]) == 0
    # rows x1 + x2 <= 1 and x1 + x2 >= 2, x free: dy = (1, -1) has A^T dy = 0 and pays the
    # finite bounds only, D = -1*1 - (-1)*2 = 1; dy = (-1, 1) would need the missing lower
    # bound of row 0 and upper bound of row 1: rejected
    rows = mp.LPData.from_dense([[1, 1], [1, 1]], [0, 0], [-inf, 2], [1, inf], [-inf, -inf], [inf, inf])
    assert orc.ray_certificate(rows, [0, 0], [1, -1]) == 1
    assert orc.ray_certificate(rows, [0, 0], [-1, 1]) == 0


@pytest.mark.parametrize("seed", [1, 2])
def test_ray_certificate_never_fires_on_feasible_bounded_lps(seed):
    """By weak duality a feasible LP has no dual ray and a bounded one no primal
    ray: random movements on planted LPs are never certified."""
    lp = mp.generate(mp.FAMILY_ZIPF_MIXED, rows=80, cols=200, nnz_max=20, seed=seed)
    rng = np.random.default_rng(seed)
    for _ in range(200):
        dx = rng.standard_normal(lp.cols) * rng.choice([0.0, 1.0, 1e3])
        dy = rng.standard_normal(lp.rows) * rng.choice([0.0, 1.0, 1e3])
        assert orc.ray_certificate(lp, dx, dy) == 0


def test_oracle_detects_infeasibility():
    """With detect_infeasibility the oracle stops the constructed infeasible and
    unbounded LPs with infeasible_or_unbounded; on the feasible LP the test only
    reads the iterate, so the solve equals the one without it; the host path
    (product CPU backend) gives the same statuses."""
    import test_host
    cases, base = test_host._infeasibility_cases()
    for scheme in (mp.SCHEME_KKT, mp.SCHEME_HALPERN):
        kw = dict(tolerance=1e-6, restart_scheme=scheme, max_iterations=100000)
        for name, lp in cases.items():
            r = orc.solve(lp, detect_infeasibility=1, **kw)
            assert r["status"] == "infeasible_or_unbounded", (name, scheme, r["status"])
            assert r["message"] == "infeasibility certificate found"
            assert mp.solve(lp, detect_infeasibility=1, **kw)["status"] == r["status"]
        a = orc.solve(base, detect_infeasibility=1, **kw)
        b = orc.solve(base, **kw)
        assert a["status"] == "optimal" and a["iterations"] == b["iterations"] and a["objective"] == b["objective"]


# ---- row weighting (extension: mc_pdlp_options::row_weighting) -------------
@pytest.mark.parametrize("beta", [0.0, 1e-300, 0.37, 1.0, 3.0, 255.0, 1e6, 1e300])
def test_row_weight_closed_forms(beta):
    """or_row_weight(beta, k) = (1 + beta)^(-k/8): k = 0 is 1 exactly, k = 8
    the correctly rounded 1/(1 + beta), k = 4 1/sqrt(1 + beta); every k lies
    within 4 ulp of the long-double power (a wrong bit-to-root mapping moves
    the exponent by 1/8 and fails for every beta > 1)."""
    s0 = 1.0 + beta
    assert orc.row_weight(beta, 0) == 1.0
    assert orc.row_weight(beta, 8) == 1.0 / s0
    assert orc.row_weight(beta, 4) == 1.0 / np.sqrt(s0)
    for k in range(9):
        want = float(np.power(np.longdouble(s0), -np.longdouble(k) / 8))
        assert abs(orc.row_weight(beta, k) - want) <= 4 * np.spacing(want), k


def ranged_lp():
    """Equality, one-sided, ranged and free rows: beta = 3, 7, 5, 0."""
    inf = np.inf
    A = [[1.0, 2.0, 0.0, -1.0], [0.5, 0.0, 3.0, 1.0], [0.0, -1.0, 1.0, 2.0], [1.0, 1.0, 1.0, 1.0]]
    return mp.LPData.from_dense(A, [1.0, -1.0, 2.0, 0.5], [3.0, -inf, -2.0, -inf], [3.0, -7.0, 5.0, inf],
                                [0.0, -10.0, 0.0, 0.0], [10.0, 10.0, inf, 4.0])


@pytest.mark.parametrize("k", [1, 3, 8])
def test_row_weights_scale_rows_and_bounds(k):
    """After Ruiz, row i is multiplied by w_i = (1 + beta_i)^(-k/8), beta_i the
    largest |finite ORIGINAL bound| of the row (0 for a free row): the scaled
    matrix is W times the Ruiz-scaled one entry for entry, row_scale is the
    Ruiz scale times w, finite row bounds are the original ones times the
    final row_scale, and the columns, objective and variable bounds are those
    of the Ruiz pass."""
    for lp, beta in ((ranged_lp(), np.array([3.0, 7.0, 5.0, 0.0])),
                     (mp.generate(mp.FAMILY_ZIPF_MIXED, rows=200, cols=500, nnz_max=30, seed=4), None)):
        if beta is None:
            lo = np.where(lp.row_lower["kind"] == mp.FINITE, np.abs(lp.row_lower["value"]), 0.0)
            hi = np.where(lp.row_upper["kind"] == mp.FINITE, np.abs(lp.row_upper["value"]), 0.0)
            beta = np.maximum(lo, hi)
        w = np.array([orc.row_weight(b, k) for b in beta])
        eq = orc.equilibrate(lp, 10)
        q = orc.weighted_scaling(lp, k)
        np.testing.assert_array_equal(q["row_scale"], eq["row_scale"] * w)
        np.testing.assert_array_equal(q["lp"].value, eq["lp"].value * w[lp.row_index])
        np.testing.assert_array_equal(q["col_scale"], eq["col_scale"])
        np.testing.assert_array_equal(q["lp"].objective, eq["lp"].objective)
        np.testing.assert_array_equal(q["lp"].var_upper["value"], eq["lp"].var_upper["value"])
        for side in ("row_lower", "row_upper"):
            fin = getattr(lp, side)["kind"] == mp.FINITE
            np.testing.assert_array_equal(getattr(q["lp"], side)["value"][fin],
                                          getattr(lp, side)["value"][fin] * q["row_scale"][fin])


@pytest.mark.parametrize("scheme", [mp.SCHEME_KKT, mp.SCHEME_HALPERN])
@pytest.mark.parametrize("ruiz", [1, 0])
def test_row_weighted_solve_reaches_planted_optimum(scheme, ruiz):
    """With the weights (and with Ruiz off, where the weights are the only
    scaling) the returned pair, re-scored in the original space, is within
    tolerance and the objective is the planted optimum: every reported
    quantity undoes the weights.  (Without Ruiz the Zipf LP does not reach
    1e-6 in 200k iterations with or without the weights.)"""
    lps = [mp.generate(mp.FAMILY_UNIFORM_EQ, rows=200, cols=500, seed=11)]
    if ruiz:
        lps.append(mp.generate(mp.FAMILY_ZIPF_MIXED, rows=300, cols=700, nnz_max=40, seed=31))
    for lp in lps:
        r = orc.solve(lp, tolerance=1e-6, restart_scheme=scheme, row_weighting=3, ruiz_scaling=ruiz,
                      max_iterations=200000)
        assert r["status"] == "optimal"
        k = orc.kkt(lp, r["primal"], r["dual"])
        assert k["score"] <= 1e-6 * 1.01
        # the solver's own residuals (A_s x_s / row_scale) agree with the re-score;
        # a weight left in would move them by about (1 + beta)^(3/8)
        np.testing.assert_allclose([r["primal_infeasibility"], r["dual_infeasibility"], r["duality_gap"]],
                                   [k["primal"], k["dual"], k["gap"]], rtol=1e-4, atol=1e-12)
        assert abs(r["objective"] - lp.planted_objective) <= 1e-5 * (1 + abs(lp.planted_objective))


# ---- primal-weight drop limit (extension: primal_weight_max_drop) ----------
def test_halpern_weight_max_drop():
    """primal_weight_max_drop = K: a restart lowers the weight by at most K
    (to exactly omega / K, K a power of two); a smaller fall and any rise are
    what the rule gives without the limit."""
    omega = 0.37
    # movement ratio 1e-6 omega with balanced residuals: the rule falls ~3e4-fold
    free = orc.halpern_weight(omega, 1.0, 1e-6 * omega, 1e-5, 1e-5)
    assert free < omega / 1024
    assert orc.halpern_weight(omega, 1.0, 1e-6 * omega, 1e-5, 1e-5, 2) == omega / 2
    assert orc.halpern_weight(omega, 1.0, 1e-6 * omega, 1e-5, 1e-5, 1024) == omega / 1024
    for dy in (0.9 * omega, 10.0 * omega):
        assert orc.halpern_weight(omega, 1.0, dy, 1e-5, 1e-5, 2) == orc.halpern_weight(omega, 1.0, dy, 1e-5, 1e-5)
    # the guard still holds the weight while the primal residual is 100x the dual
    assert orc.halpern_weight(omega, 1.0, 1e-6 * omega, 1e-3, 1e-6, 2) == omega


def test_max_drop_solve_reaches_planted_optimum():
    """Halpern solves with the drop limit (and row weighting) reach the planted
    optimum of a uniform and a Zipf LP; re-scored pairs are within tolerance."""
    for lp in (mp.generate(mp.FAMILY_UNIFORM_EQ, rows=200, cols=500, seed=11),
               mp.generate(mp.FAMILY_ZIPF_MIXED, rows=300, cols=700, nnz_max=40, seed=31)):
        r = orc.solve(lp, tolerance=1e-6, restart_scheme=mp.SCHEME_HALPERN, row_weighting=3,
                      primal_weight_max_drop=2, max_iterations=200000)
        assert r["status"] == "optimal"
        assert orc.kkt(lp, r["primal"], r["dual"])["score"] <= 1e-6 * 1.01
        assert abs(r["objective"] - lp.planted_objective) <= 1e-5 * (1 + abs(lp.planted_objective))
