"""ctypes binding of the CPU checkers under oracle/ (test infrastructure).

* ``liboracle.so``        -- oracle/pdlp_oracle.c, the C restatement of the
                             reference PDLP path (always built by oracle/Makefile).
* ``_ref/libref_pdlp.so`` -- the reference's own sources compiled by
                             oracle/Makefile where the reference checkout exists;
                             ``ref_available()`` says whether it was built.

Only tests, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg use this.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "markov-zip1_amd"))
import mcpdlp as mp  # noqa: E402

ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_pdlp.so")

P_i64, P_f64, P_i32 = mp.P_i64, mp.P_f64, mp.P_i32


class Trace(C.Structure):
    _fields_ = [("capacity", C.c_int64), ("count", C.c_int64), ("score", P_f64), ("omega", P_f64),
                ("eta", P_f64), ("restart", P_i32)]


_or = None
_ref = None


def oracle_lib():
    global _or
    if _or is None:
        if not os.path.exists(ORACLE_LIB):
            raise OSError(f"{ORACLE_LIB} not built (make -C oracle)")
        lib = C.CDLL(ORACLE_LIB)
        lib.or_spmv.argtypes = [C.c_int64, C.c_int64, P_i64, P_i64, P_f64, P_f64, P_f64]
        lib.or_spmv_t.argtypes = [C.c_int64, C.c_int64, P_i64, P_i64, P_f64, P_f64, P_f64]
        lib.or_equilibrate_model.argtypes = [C.c_int64, C.c_int64, P_i64, P_i64, P_f64, P_f64,
                                             C.POINTER(mp.Bound), C.POINTER(mp.Bound), C.POINTER(mp.Bound),
                                             C.POINTER(mp.Bound), C.c_int64, C.c_double, P_f64, P_f64, P_i32]
        lib.or_equilibrate_model.restype = C.c_int64
        lib.or_solve_pdlp.argtypes = [C.POINTER(mp.LP), C.POINTER(mp.Options), C.POINTER(mp.Result),
                                      C.POINTER(Trace)]
        lib.or_verify_primal.argtypes = [C.POINTER(mp.LP), P_f64, C.c_double, C.c_double, C.c_double,
                                         C.c_double, C.c_double, C.POINTER(mp.VerifyReport)]
        lib.or_kkt.argtypes = [C.POINTER(mp.LP), P_f64, P_f64, P_f64]
        lib.or_step_cap.argtypes = [C.c_int64, C.c_int64, P_i64, P_i64, P_f64, P_f64, P_f64, C.c_int64]
        lib.or_step_cap.restype = C.c_double
        lib.or_ray_certificate.argtypes = [C.POINTER(mp.LP), P_f64, P_f64, C.c_double]
        lib.or_halpern_weight.argtypes = [C.c_double] * 5 + [C.c_int32]
        lib.or_halpern_weight.restype = C.c_double
        lib.or_row_weight.argtypes = [C.c_double, C.c_int32]
        lib.or_row_weight.restype = C.c_double
        lib.or_apply_row_weights.argtypes = [C.POINTER(mp.LP), C.c_int32, P_f64, C.POINTER(mp.Bound),
                                             C.POINTER(mp.Bound), P_f64]
        _or = lib
    return _or


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref_lib():
    global _ref
    if _ref is None:
        lib = C.CDLL(REF_LIB)
        lib.ref_solve_pdlp.argtypes = [C.POINTER(mp.LP), C.POINTER(mp.Options), C.POINTER(mp.Result)]
        lib.ref_equilibrate_model.argtypes = [C.POINTER(mp.LP), C.c_int64, P_f64, P_f64, C.POINTER(mp.Bound),
                                              C.POINTER(mp.Bound), C.POINTER(mp.Bound), C.POINTER(mp.Bound),
                                              P_f64, P_f64, P_i64, P_i32]
        lib.ref_verify_primal.argtypes = [C.POINTER(mp.LP), P_f64, C.c_double, C.c_double, C.c_double,
                                          C.c_double, C.c_double, C.POINTER(mp.VerifyReport)]
        lib.ref_last_error.restype = C.c_char_p
        _ref = lib
    return _ref


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def spmv(lp: mp.LPData, x) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(max(lp.rows, 1))
    oracle_lib().or_spmv(lp.rows, lp.cols, _p(lp.column_start, C.c_int64), _p(lp.row_index, C.c_int64),
                         _p(lp.value, C.c_double), _p(x, C.c_double), _p(y, C.c_double))
    return y[: lp.rows]


def spmv_t(lp: mp.LPData, y) -> np.ndarray:
    y = np.ascontiguousarray(y, dtype=np.float64)
    z = np.zeros(max(lp.cols, 1))
    oracle_lib().or_spmv_t(lp.rows, lp.cols, _p(lp.column_start, C.c_int64), _p(lp.row_index, C.c_int64),
                           _p(lp.value, C.c_double), _p(y, C.c_double), _p(z, C.c_double))
    return z[: lp.cols]


def equilibrate(lp: mp.LPData, iterations=10, use_ref=False) -> dict:
    """Scaled copy of lp's data + scalers from the oracle (or the compiled reference)."""
    q = lp.copy()
    rs, cs = np.ones(max(lp.rows, 1)), np.ones(max(lp.cols, 1))
    conv = C.c_int32(0)
    if use_ref:
        passes = C.c_int64(0)
        s = lp.struct()
        rc = ref_lib().ref_equilibrate_model(
            C.byref(s), iterations, _p(q.value, C.c_double), _p(q.objective, C.c_double),
            _p(q.row_lower, mp.Bound), _p(q.row_upper, mp.Bound), _p(q.var_lower, mp.Bound),
            _p(q.var_upper, mp.Bound), _p(rs, C.c_double), _p(cs, C.c_double), C.byref(passes), C.byref(conv))
        assert rc == 0, ref_lib().ref_last_error()
        npass = passes.value
    else:
        npass = oracle_lib().or_equilibrate_model(
            lp.rows, lp.cols, _p(lp.column_start, C.c_int64), _p(lp.row_index, C.c_int64),
            _p(q.value, C.c_double), _p(q.objective, C.c_double), _p(q.row_lower, mp.Bound),
            _p(q.row_upper, mp.Bound), _p(q.var_lower, mp.Bound), _p(q.var_upper, mp.Bound), iterations, 1e-3,
            _p(rs, C.c_double), _p(cs, C.c_double), C.byref(conv))
    return {"lp": q, "row_scale": rs[: lp.rows], "col_scale": cs[: lp.cols], "passes": int(npass),
            "converged": bool(conv.value)}


def _run(fn, lp, opts, trace_cap=0):
    prim = np.zeros(max(lp.cols, 1))
    dual = np.zeros(max(lp.rows, 1))
    r = mp.Result()
    r.primal = _p(prim, C.c_double)
    r.dual = _p(dual, C.c_double)
    s = lp.struct()
    tr = None
    if trace_cap:
        bufs = [np.zeros(trace_cap), np.zeros(trace_cap), np.zeros(trace_cap), np.zeros(trace_cap, np.int32)]
        tr = Trace(trace_cap, 0, _p(bufs[0], C.c_double), _p(bufs[1], C.c_double), _p(bufs[2], C.c_double),
                   _p(bufs[3], C.c_int32))
        rc = fn(C.byref(s), C.byref(opts), C.byref(r), C.byref(tr))
    elif fn is getattr(oracle_lib(), "or_solve_pdlp"):
        rc = fn(C.byref(s), C.byref(opts), C.byref(r), None)
    else:
        rc = fn(C.byref(s), C.byref(opts), C.byref(r))
    assert rc == 0
    out = mp._result_dict(r, prim, dual)
    if tr is not None:
        k = tr.count
        out["trace"] = {"score": bufs[0][:k].copy(), "omega": bufs[1][:k].copy(), "eta": bufs[2][:k].copy(),
                        "restart": bufs[3][:k].copy()}
    return out


def solve(lp: mp.LPData, opts=None, trace_cap=0, **kw) -> dict:
    o = opts if opts is not None else mp.options(**kw)
    return _run(oracle_lib().or_solve_pdlp, lp, o, trace_cap)


def ref_solve(lp: mp.LPData, opts=None, **kw) -> dict:
    o = opts if opts is not None else mp.options(**kw)
    return _run(ref_lib().ref_solve_pdlp, lp, o)


def verify(lp: mp.LPData, x, objective, tol=1e-4, use_ref=False) -> dict:
    rep = mp.VerifyReport()
    x = np.ascontiguousarray(x, dtype=np.float64)
    s = lp.struct()
    fn = ref_lib().ref_verify_primal if use_ref else oracle_lib().or_verify_primal
    rc = fn(C.byref(s), _p(x, C.c_double), objective, tol, tol, tol, tol, C.byref(rep))
    assert rc == 0
    return {"passed": bool(rep.passed), "violations": rep.violation_count,
            "recomputed_objective": rep.recomputed_objective,
            "max_row_violation": rep.maximum_row_violation,
            "max_variable_violation": rep.maximum_variable_violation}


def kkt(lp: mp.LPData, x, y) -> dict:
    """Reference relative-KKT residuals of an unscaled (x, y) on the original model."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros(5)
    s = lp.struct()
    oracle_lib().or_kkt(C.byref(s), _p(x, C.c_double), _p(y, C.c_double), _p(out, C.c_double))
    return {"primal": out[0], "dual": out[1], "gap": out[2], "score": out[3], "objective": out[4]}


def step_cap(lp: mp.LPData, iters: int) -> float:
    """Halpern-scheme step cap 0.95 / L of the oracle, L the power-method estimate
    of ||D_r^-1/2 A D_c^-1/2|| (D = L1 row / column norms of lp's matrix, floored
    at 1 below 1e-12) after `iters` iterations."""
    a = np.abs(lp.value)
    cn = np.array([a[lp.column_start[j]:lp.column_start[j + 1]].sum() for j in range(lp.cols)])
    rn = np.zeros(lp.rows)
    np.add.at(rn, lp.row_index, a)
    cn[cn < 1e-12] = 1.0
    rn[rn < 1e-12] = 1.0
    return oracle_lib().or_step_cap(lp.rows, lp.cols, _p(lp.column_start, C.c_int64), _p(lp.row_index, C.c_int64),
                                    _p(lp.value, C.c_double), _p(rn, C.c_double), _p(cn, C.c_double), iters)


def row_weight(beta: float, k: int) -> float:
    """(1 + beta)^(-k/8) as the oracle forms it (nested square roots)."""
    return oracle_lib().or_row_weight(float(beta), int(k))


def weighted_scaling(lp: mp.LPData, k: int, ruiz=True) -> dict:
    """equilibrate() (or unit scalers without Ruiz) followed by the oracle's
    row weighting of row_weighting = k: scaled copy of lp and its scalers."""
    eq = equilibrate(lp) if ruiz else {"lp": lp.copy(), "row_scale": np.ones(lp.rows), "col_scale": np.ones(lp.cols)}
    q = eq["lp"]
    rs = np.ascontiguousarray(eq["row_scale"], dtype=np.float64).copy()
    s = lp.struct()
    oracle_lib().or_apply_row_weights(C.byref(s), int(k), _p(q.value, C.c_double), _p(q.row_lower, mp.Bound),
                                      _p(q.row_upper, mp.Bound), _p(rs, C.c_double))
    return {"lp": q, "row_scale": rs, "col_scale": eq["col_scale"]}


def halpern_weight(omega, dx, dy, primal, dual, max_drop=0) -> float:
    """The Halpern scheme's primal weight after a restart (oracle)."""
    return oracle_lib().or_halpern_weight(omega, dx, dy, primal, dual, int(max_drop))


def ray_certificate(lp: mp.LPData, dx, dy, tol=1e-8) -> int:
    """Oracle verdict on a movement (dx, dy) of the iterate in the original
    space: 1 dual ray (primal infeasible), 2 primal ray (unbounded), 0 none."""
    dx = np.ascontiguousarray(dx, dtype=np.float64)
    dy = np.ascontiguousarray(dy, dtype=np.float64)
    s = lp.struct()
    return oracle_lib().or_ray_certificate(C.byref(s), _p(dx, C.c_double), _p(dy, C.c_double), tol)
