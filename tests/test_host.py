"""Host library, C-ABI and CLI -- CPU only (no compute on a GPU).

* every function declared in include/*.h is exported by the built libraries;
* Backend::cpu of the host library equals the oracle bit for bit (both restart
  schemes), and a GPU request without a device reports "cpu_fallback"
  (tests/regression_backend_actually_used.cpp:26-43) or fails with require_gpu;
* the CLI keeps the reference's pdlp branch: JSON keys, exit codes,
  10000 -> 100000 iteration-limit quirk, verification, stderr verdict;
* MPS / binary LP round trips; generator determinism and planted optimality.
"""
import json
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import mcpdlp as mp
import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(m[c]p?_\w+)\s*\(", text, flags=re.M)))


def exported(lib_path):
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True, check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


@pytest.mark.parametrize("header,lib", [("mcpdlp.h", mp.HOST_LIB), ("mcpdlp_engine.h", mp.ENGINE_LIB)])
def test_every_declared_symbol_is_exported(header, lib):
    names = declared_functions(header)
    assert len(names) >= 8
    missing = [n for n in names if n not in exported(lib)]
    assert not missing, missing


def test_libraries_load_without_gpu():
    mp.host_lib()
    mp.engine_lib()
    assert mp.device_count() >= 0


def tiny():
    inf = np.inf
    return mp.LPData.from_dense([[1, 1], [2, 1]], [-3, -2], [-inf, -inf], [4, 5], [0, 0], [inf, inf])


CASES = [
    (lambda: tiny(), dict(tolerance=1e-6)),
    (lambda: mp.generate(mp.FAMILY_UNIFORM_EQ, rows=200, cols=500, seed=11), dict(tolerance=1e-6, max_iterations=400)),
    (lambda: mp.generate(mp.FAMILY_ZIPF_MIXED, rows=300, cols=700, nnz_max=40, seed=12),
     dict(tolerance=1e-6, max_iterations=600)),
    (lambda: mp.generate(mp.FAMILY_UNIFORM_EQ, rows=200, cols=500, seed=11),
     dict(tolerance=1e-6, restart_scheme=mp.SCHEME_KKT, max_iterations=20000)),
    (lambda: mp.generate(mp.FAMILY_ZIPF_MIXED, rows=300, cols=700, nnz_max=40, seed=12),
     dict(tolerance=1e-6, restart_scheme=mp.SCHEME_KKT, max_iterations=3000)),
    (lambda: mp.generate(mp.FAMILY_UNIFORM_EQ, rows=200, cols=500, seed=11),
     dict(tolerance=1e-6, restart_scheme=mp.SCHEME_HALPERN, max_iterations=20000)),
    (lambda: mp.generate(mp.FAMILY_ZIPF_MIXED, rows=300, cols=700, nnz_max=40, seed=12),
     dict(tolerance=1e-6, restart_scheme=mp.SCHEME_HALPERN, max_iterations=3000)),
    (lambda: mp.generate(mp.FAMILY_ZIPF_MIXED, rows=300, cols=700, nnz_max=40, seed=12),
     dict(tolerance=1e-6, restart_scheme=mp.SCHEME_HALPERN, row_weighting=3, max_iterations=3000)),
    (lambda: mp.generate(mp.FAMILY_UNIFORM_EQ, rows=200, cols=500, seed=11),
     dict(tolerance=1e-6, restart_scheme=mp.SCHEME_KKT, row_weighting=5, ruiz_scaling=0, max_iterations=3000)),
    (lambda: mp.generate(mp.FAMILY_ZIPF_MIXED, rows=300, cols=700, nnz_max=40, seed=12),
     dict(tolerance=1e-6, restart_scheme=mp.SCHEME_HALPERN, row_weighting=3, primal_weight_max_drop=2,
          max_iterations=20000)),
]


@pytest.mark.parametrize("k", range(len(CASES)))
def test_host_cpu_backend_equals_oracle(k):
    make, kw = CASES[k]
    lp = make()
    a, b = mp.solve(lp, **kw), orc.solve(lp, **kw)
    assert a["backend_actually_used"] == "cpu"
    assert a["iterations"] == b["iterations"] and a["objective"] == b["objective"]
    assert np.array_equal(a["primal"], b["primal"]) and np.array_equal(a["dual"], b["dual"])
    assert (a["primal_infeasibility"], a["dual_infeasibility"], a["duality_gap"]) == \
           (b["primal_infeasibility"], b["dual_infeasibility"], b["duality_gap"])


def test_backend_labels_and_trivial():
    lp = tiny()
    assert mp.solve(lp, tolerance=1e-4)["backend_actually_used"] == "cpu"
    if mp.device_count() == 0:
        r = mp.solve(lp, tolerance=1e-4, backend=mp.BACKEND_GPU)
        assert r["backend_actually_used"] == "cpu_fallback" and r["status"] == "optimal"
        with pytest.raises(RuntimeError, match="GPU required"):
            mp.solve(lp, tolerance=1e-4, backend=mp.BACKEND_GPU, require_gpu=1)
    empty = mp.LPData(0, 0, np.zeros(1, np.int64), np.zeros(0, np.int64), np.zeros(0), np.zeros(0),
                      np.zeros(0, mp.BOUND_DTYPE), np.zeros(0, mp.BOUND_DTYPE), np.zeros(0, mp.BOUND_DTYPE),
                      np.zeros(0, mp.BOUND_DTYPE))
    r = mp.solve(empty)
    assert r["status"] == "optimal" and r["message"] == "trivial" and r["primal"].size == 0  # pdlp.cpp:144-149


def test_precision_option_on_the_cpu_backend():
    """The precision extension (fp64, mixed, packed) selects the GPU engine's
    SpMV arithmetic; the host restatement is FP64 whatever it says, so every
    value gives the same bits there and reports no reduced-precision
    iterations."""
    lp = tiny()
    runs = [mp.solve(lp, tolerance=1e-6, precision=p) for p in (mp.PRECISION_FP64, mp.PRECISION_MIXED,
                                                               mp.PRECISION_PACKED)]
    for r in runs:
        assert r["backend_actually_used"] == "cpu" and r["status"] == "optimal"
        assert r["fp32_iterations"] == 0 and r["packed_iterations"] == 0
        assert r["iterations"] == runs[0]["iterations"] and r["objective"] == runs[0]["objective"]
        assert np.array_equal(r["primal"], runs[0]["primal"])


def test_iteration_limit_zero_reports_start_point():
    lp = mp.generate(mp.FAMILY_UNIFORM_EQ, rows=50, cols=120, seed=2)
    a, b = mp.solve(lp, max_iterations=0), orc.solve(lp, max_iterations=0)
    assert a["status"] == b["status"] == "iteration_limit" and a["iterations"] == 0
    assert a["objective"] == b["objective"]


def test_invalid_model_is_an_error_code():
    lp = tiny()
    bad = lp.copy()
    bad.row_index = bad.row_index.copy()
    bad.row_index[0] = 99
    with pytest.raises(RuntimeError, match="row index out of range"):
        mp.solve(bad)


def test_verify_matches_oracle_and_reference_semantics():
    lp = tiny()
    r = orc.solve(lp, tolerance=1e-6)
    a = mp.verify(lp, r["primal"], r["objective"], 1e-6)
    b = orc.verify(lp, r["primal"], r["objective"], 1e-6)
    assert a["passed"] == b["passed"] and a["recomputed_objective"] == b["recomputed_objective"]
    assert not mp.verify(lp, [3.0, 3.0], -15.0, 1e-6)["passed"]
    with pytest.raises(RuntimeError, match="tolerances must be finite"):
        mp.verify(lp, r["primal"], r["objective"], 1e-3)  # primal_verifier.cpp:39-43 rejects > 1e-4


# ---- CLI ------------------------------------------------------------------------
def run_cli(args):
    p = subprocess.run([mp.CLI, *args], capture_output=True, text=True, timeout=120)
    return p.returncode, json.loads(p.stdout.strip().splitlines()[-1]), p.stderr


def test_cli_tiny_simplex_adr():
    code, js, err = run_cli([os.path.join(GOLDEN, "tiny_simplex.mps"), "--engine", "pdlp", "--tolerance", "1e-6"])
    assert code == 0 and js["status"] == "Optimal" and js["verified"]
    assert round(js["objective"], 7) == -9.0000013 and js["backend"] == "cpu"
    # extension key: the dual in the reference's convention g = c + A^T y; at the
    # optimum (1, 3) both rows are tight and c + A^T y = 0 gives y = (1, 1)
    assert js["dual"] == pytest.approx([1.0, 1.0], abs=1e-5)
    for key in ("relative_primal_residual", "relative_dual_residual", "relative_duality_gap", "h2d_ms", "kernel_ms",
                "d2h_ms", "total_ms", "lp_iterations", "pdlp_tolerance", "primal"):
        assert key in js
    # json_number prints 17 significant digits (apps/json_output.hpp:69-78)
    assert "Optimal [tol=9.9999999999999995e-07] [cpu total=" in err and "VERIFIED" in err


def test_cli_exit_codes():
    mps = os.path.join(GOLDEN, "tiny_simplex.mps")
    code, js, _ = run_cli([mps, "--engine", "pdlp", "--tolerance", "1e-9", "--iteration-limit", "80"])
    assert code == 6 and js["status"] == "IterationLimit" and js["lp_iterations"] == 80
    with tempfile.NamedTemporaryFile("w", suffix=".mps", delete=False) as f:
        f.write("NAME BAD\nROWS\n N COST\nCOLUMNS\n X1 NOPE 1.0\nENDATA\n")
    code, js, _ = run_cli([f.name, "--engine", "pdlp"])
    os.unlink(f.name)
    assert code == 3 and js["status"] == "InvalidModel"
    if mp.device_count() == 0:
        code, js, err = run_cli([mps, "--engine", "pdlp", "--backend", "gpu", "--tolerance", "1e-6"])
        assert code == 0 and js["backend_actually_used"] == "cpu_fallback" and "[gpu H2D=" in err


def test_cli_iteration_limit_quirk():
    """--iteration-limit 10000 (the simplex default) means 100000 (markov_cero_solve.cpp:141-144)."""
    lp_path = os.path.join(GOLDEN, "tiny_simplex.mps")
    code, js, _ = run_cli([lp_path, "--engine", "pdlp", "--tolerance", "1e-6", "--iteration-limit", "10000"])
    assert code == 0 and js["lp_iterations"] == 200


# ---- ingest -----------------------------------------------------------------------
def test_mps_and_binary_round_trip(tmp_path):
    lp = mp.generate(mp.FAMILY_ZIPF_MIXED, rows=60, cols=150, nnz_max=12, seed=3)
    ref = orc.solve(lp, tolerance=1e-6, max_iterations=400)
    # the CLI reads the binary LP format and free MPS; both must give the oracle's run
    path = tmp_path / "lp.mclp"
    write_binary(lp, str(path))
    code, js, _ = run_cli([str(path), "--engine", "pdlp", "--tolerance", "1e-6", "--iteration-limit", "400"])
    assert js["lp_iterations"] == 400 and js["objective"] == ref["objective"]
    mps_path = tmp_path / "lp.mps"
    write_mps(lp, str(mps_path))
    code, js2, _ = run_cli([str(mps_path), "--engine", "pdlp", "--tolerance", "1e-6", "--iteration-limit", "400"])
    assert js2["objective"] == pytest.approx(ref["objective"], rel=1e-12)


def write_binary(lp, path):
    with open(path, "wb") as f:
        f.write(b"MCLP0001")
        np.array([lp.rows, lp.cols, lp.nnz], np.int64).tofile(f)
        np.array([lp.sense, 0], np.int32).tofile(f)
        np.array([lp.offset]).tofile(f)
        for a in (lp.column_start.astype(np.int64), lp.row_index.astype(np.int64), lp.value, lp.objective,
                  lp.row_lower, lp.row_upper, lp.var_lower, lp.var_upper):
            a.tofile(f)


def write_mps(lp, path):
    A = lp.dense()
    with open(path, "w") as f:
        f.write("NAME RT\nROWS\n N COST\n")
        for i in range(lp.rows):
            lo, hi = lp.row_lower[i], lp.row_upper[i]
            t = "E" if lo["kind"] == mp.FINITE and hi["kind"] == mp.FINITE and lo["value"] == hi["value"] else (
                "L" if hi["kind"] == mp.FINITE else "G")
            f.write(f" {t} R{i}\n")
        f.write("COLUMNS\n")
        for j in range(lp.cols):
            f.write(f"    C{j} COST {float(lp.objective[j])!r}\n")
            for i in np.nonzero(A[:, j])[0]:
                f.write(f"    C{j} R{i} {float(A[i, j])!r}\n")
        f.write("RHS\n")
        for i in range(lp.rows):
            hi = lp.row_upper[i]
            v = hi["value"] if hi["kind"] == mp.FINITE else lp.row_lower[i]["value"]
            f.write(f"    RHS R{i} {float(v)!r}\n")
        f.write("BOUNDS\n")
        for j in range(lp.cols):
            f.write(f" UP BND C{j} {float(lp.var_upper[j]['value'])!r}\n")
        f.write("ENDATA\n")


def test_generators_deterministic_and_planted():
    a = mp.generate(mp.FAMILY_ZIPF_MIXED, rows=400, cols=900, nnz_max=50, seed=77)
    b = mp.generate(mp.FAMILY_ZIPF_MIXED, rows=400, cols=900, nnz_max=50, seed=77)
    assert np.array_equal(a.value, b.value) and np.array_equal(a.row_index, b.row_index)
    assert orc.kkt(a, a.x_star, -a.y_star)["score"] < 1e-11
    eq = (a.row_lower["kind"] == mp.FINITE).mean()
    assert 0.6 < eq < 0.8  # 70% equality rows
    blk = mp.generate(mp.FAMILY_BLOCK_ANGULAR, rows=40, cols=100, periods=4, link_rows=30, seed=5)
    assert blk.rows == 4 * 40 + 30 and blk.cols == 400 and blk.nnz == 400 * (8 + 2)
    assert orc.kkt(blk, blk.x_star, -blk.y_star)["score"] < 1e-11
    base, fam = mp.generate_family(dict(rows=100, cols=250, nnz_min=6, nnz_max=6), seeds=[1, 2, 3])
    for q in fam:
        assert np.array_equal(q.row_index, base.row_index) and np.array_equal(q.value, base.value)
        assert not np.array_equal(q.objective, base.objective)
        assert orc.kkt(q, q.x_star, -q.y_star)["score"] < 1e-11


def test_cli_restart_scheme_flag():
    """--restart-scheme halpern solves Netlib adlittle (optimum 225494.96316238,
    scripts/run_netlib.py) to 1e-6; the reference rule does not within the
    CLI's 100000-iteration default; an unknown scheme is a usage error."""
    mps = os.path.join(GOLDEN, "netlib", "adlittle.mps")
    code, js, err = run_cli([mps, "--engine", "pdlp", "--tolerance", "1e-6", "--restart-scheme", "halpern"])
    assert code == 0 and js["status"] == "Optimal" and js["verified"]
    assert abs(js["objective"] - 225494.96316238) <= 1e-6 * 225494.96316238
    code, js, err = run_cli([mps, "--engine", "pdlp", "--tolerance", "1e-6"])
    assert code == 6 and js["status"] == "IterationLimit"
    p = subprocess.run([mp.CLI, mps, "--restart-scheme", "bogus"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 4 and "invalid restart scheme" in p.stderr


def _infeasibility_cases():
    inf = np.inf
    base = mp.generate(mp.FAMILY_UNIFORM_EQ, rows=200, cols=500, seed=11)
    A, rl = base.dense(), base.row_lower["value"]
    lo, hi = base.var_lower["value"], base.var_upper["value"]
    # a copy of row 0 with its right-hand side moved by 1: no x satisfies both
    contra = mp.LPData.from_dense(np.vstack([A, A[0]]), base.objective, np.append(rl, rl[0] + 1.0),
                                  np.append(rl, rl[0] + 1.0), lo, hi)
    # an extra column in no row, cost -1, no upper bound: unbounded along it
    free = mp.LPData.from_dense(np.hstack([A, np.zeros((A.shape[0], 1))]), np.append(base.objective, -1.0),
                                rl, rl, np.append(lo, 0.0), np.append(hi, inf))
    return {
        "x1+x2=1 and x1+x2=2": mp.LPData.from_dense([[1, 1], [1, 1]], [1, 1], [1, 2], [1, 2], [0, 0], [inf, inf]),
        "min -x1, x1=x2, x>=0": mp.LPData.from_dense([[1, -1]], [-1, 0], [0], [0], [0, 0], [inf, inf]),
        "planted + contradictory row": contra,
        "planted + free column": free,
    }, base


@pytest.mark.parametrize("scheme", [mp.SCHEME_KKT, mp.SCHEME_HALPERN])
def test_cpu_infeasibility_detection(scheme):
    """Extension: with detect_infeasibility the host path stops infeasible and
    unbounded LPs with infeasible_or_unbounded; on a feasible LP the test only
    reads the iterate, so the solve is bit-identical to one without it."""
    cases, base = _infeasibility_cases()
    for name, lp in cases.items():
        r = mp.solve(lp, tolerance=1e-6, restart_scheme=scheme, max_iterations=100000, detect_infeasibility=1)
        assert r["status"] == "infeasible_or_unbounded", (name, r["status"], r["iterations"])
        assert r["message"] == "infeasibility certificate found"
        off = mp.solve(lp, tolerance=1e-6, restart_scheme=scheme, max_iterations=2000)
        assert off["status"] == "iteration_limit"  # the reference's behaviour: no such status
    a = mp.solve(base, tolerance=1e-6, restart_scheme=scheme, max_iterations=100000, detect_infeasibility=1)
    b = mp.solve(base, tolerance=1e-6, restart_scheme=scheme, max_iterations=100000)
    assert a["status"] == "optimal" and a["iterations"] == b["iterations"]
    assert np.array_equal(a["primal"], b["primal"]) and np.array_equal(a["dual"], b["dual"])


def test_relabel_by_length_is_a_permutation():
    """mcpdlp.relabel_by_length (the host library's length-bucket relabelling,
    for Engine users): same LP up to a row/column permutation -- the relabelled
    dense matrix is a permutation of the original, columns and rows come in
    descending length bucket, every column's rows are strictly increasing, and
    the permuted planted pair keeps its objective and KKT score."""
    lp = mp.generate(mp.FAMILY_ZIPF_MIXED, rows=300, cols=700, nnz_max=40, seed=12)
    q = mp.relabel_by_length(lp)
    A, Q = lp.dense(), q.dense()
    clen, rlen = np.diff(lp.column_start), np.bincount(lp.row_index, minlength=lp.rows)
    cperm = np.argsort(-mp._length_bucket(clen), kind="stable")
    rperm = np.argsort(-mp._length_bucket(rlen), kind="stable")
    assert np.array_equal(Q, A[np.ix_(rperm, cperm)])
    assert np.all(np.diff(mp._length_bucket(np.diff(q.column_start))) <= 0)
    for j in range(q.cols):
        assert np.all(np.diff(q.row_index[q.column_start[j]:q.column_start[j + 1]]) > 0)
    x, y = lp.x_star[cperm], -lp.y_star[rperm]
    assert np.isclose(q.objective @ x, lp.objective @ lp.x_star, rtol=1e-14)
    assert orc.kkt(q, x, y)["score"] == pytest.approx(orc.kkt(lp, lp.x_star, -lp.y_star)["score"], abs=1e-12)


def test_rows_are_local_matches_a_dense_span():
    """mcpdlp.rows_are_local (the host library's test for keeping the natural
    order) against the span read off the dense matrix: block-angular LPs (rows
    inside a period's columns) are local and relabel_by_length leaves them as
    they are; random-column LPs are not; a hand-made matrix sits on each side
    of the 1/4 threshold."""
    for lp, want in [(mp.generate(mp.FAMILY_BLOCK_ANGULAR, rows=400, cols=1000, periods=8, link_rows=800), True),
                     (mp.generate(mp.FAMILY_ZIPF_MIXED, rows=300, cols=700, nnz_max=40, seed=12), False),
                     (mp.generate(mp.FAMILY_UNIFORM_EQ, rows=200, cols=500, seed=3), False)]:
        A = lp.dense() != 0
        used = A.any(axis=1)
        first = np.argmax(A, axis=1)
        last = A.shape[1] - 1 - np.argmax(A[:, ::-1], axis=1)
        span = ((last - first + 1) / A.shape[1])[used].mean()
        assert (span < 0.25) == want == mp.rows_are_local(lp), span
        if want:
            assert mp.relabel_by_length(lp) is lp
    # 4 rows x 8 columns: spans 1/8, 1/8, 1/8 and s/8 (row 3 covers columns 0..s-1)
    for s_, want in [(4, True), (5, False)]:
        rows = [[0], [1], [2], list(range(s_))]
        cs, ri = [0], []
        for j in range(8):
            ri += [i for i, r in enumerate(rows) if j in r]
            cs.append(len(ri))
        lp = mp.LPData(rows=4, cols=8, column_start=np.array(cs, np.int64), row_index=np.array(ri, np.int64),
                       value=np.ones(len(ri)), objective=np.zeros(8), var_lower=np.zeros(8),
                       var_upper=np.ones(8), row_lower=np.zeros(4), row_upper=np.ones(4))
        assert mp.rows_are_local(lp) == want  # mean span (3 + s) / 32


def test_row_weighting_out_of_range_is_an_error():
    lp = tiny()
    for k in (-1, 9):
        with pytest.raises(RuntimeError, match="row_weighting must be in 0..8"):
            mp.solve(lp, tolerance=1e-6, row_weighting=k)
    p = subprocess.run([mp.CLI, os.path.join(GOLDEN, "tiny_simplex.mps"), "--row-weighting", "9"],
                       capture_output=True, text=True, timeout=60)
    assert p.returncode == 4 and "row weighting must be in 0..8" in p.stderr


def test_cli_row_weighting_flag():
    """--row-weighting 3 with the Halpern scheme solves Netlib adlittle (optimum
    225494.96316238, scripts/run_netlib.py) to 1e-6, verified."""
    mps = os.path.join(GOLDEN, "netlib", "adlittle.mps")
    code, js, err = run_cli([mps, "--engine", "pdlp", "--tolerance", "1e-6", "--restart-scheme", "halpern",
                             "--row-weighting", "3"])
    assert code == 0 and js["status"] == "Optimal" and js["verified"]
    assert abs(js["objective"] - 225494.96316238) <= 1e-6 * 225494.96316238


def test_primal_weight_max_drop_must_be_a_power_of_two():
    lp = tiny()
    for k in (1, 3, -2, 1 << 21):
        with pytest.raises(RuntimeError, match="primal_weight_max_drop must be 0 or a power of two"):
            mp.solve(lp, tolerance=1e-6, restart_scheme=mp.SCHEME_HALPERN, primal_weight_max_drop=k)
    assert mp.solve(lp, tolerance=1e-6, restart_scheme=mp.SCHEME_HALPERN, primal_weight_max_drop=4)["status"] == "optimal"
    p = subprocess.run([mp.CLI, os.path.join(GOLDEN, "tiny_simplex.mps"), "--primal-weight-max-drop", "3"],
                       capture_output=True, text=True, timeout=60)
    assert p.returncode == 4 and "power of two" in p.stderr


@pytest.mark.parametrize("k", [0, 3])
def test_warm_start_from_the_solution_stops_at_the_first_check(k):
    """A warm start (warm_primal / warm_dual, original space) at a solution
    found to 1e-9 is optimal at the first check (40 iterations) on the host
    path, with and without row weighting: the start is mapped into the scaled
    space through the same row scales (Ruiz times the weights) that map the
    answer back."""
    lp = mp.generate(mp.FAMILY_UNIFORM_EQ, rows=200, cols=500, seed=11)
    kw = dict(tolerance=1e-9, restart_scheme=mp.SCHEME_HALPERN, row_weighting=k, max_iterations=200000)
    ref = mp.solve(lp, **kw)
    assert ref["status"] == "optimal"
    r = mp.solve(lp, **dict(kw, tolerance=1e-6), warm_primal=ref["primal"], warm_dual=ref["dual"],
                 initial_primal_weight=ref["primal_weight"])
    assert r["status"] == "optimal" and r["iterations"] == 40, r["iterations"]
