"""ctypes binding of the MI355X PDLP engine (include/mcpdlp.h, include/mcpdlp_engine.h).

This is the Python-side FFI a user (or the test-suite and bench.py) drives the
C-ABI with: plain structs mirroring the C headers, numpy arrays for the model,
and thin wrappers that raise on any non-zero status.  The shared libraries are
the in-tree builds under ``lib/``; importing this module never falls back to
anything else -- a missing library raises ``OSError`` naming the path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
HOST_LIB = os.path.join(LIB_DIR, "libmarkov_cero_pdlp.so")
ENGINE_LIB = os.path.join(LIB_DIR, "libmcpdlp_hip.so")
CLI = os.path.join(HERE, "bin", "markov-cero-solve")

# ---- constants (mcpdlp.h) -------------------------------------------------
NEG_INF, FINITE, POS_INF = 0, 1, 2
MINIMIZE, MAXIMIZE = 0, 1
BACKEND_CPU, BACKEND_GPU = 0, 1
RESTART_NONE, RESTART_ADAPTIVE, RESTART_FIXED = 0, 1, 2
PRECISION_FP64, PRECISION_MIXED, PRECISION_PACKED = 0, 1, 2
ORDER_FAST, ORDER_REFERENCE = 0, 1
SCHEME_REFERENCE, SCHEME_KKT, SCHEME_HALPERN = 0, 1, 2
STATUS = {0: "optimal", 1: "iteration_limit", 2: "infeasible_or_unbounded", 3: "numerical_failure"}
FAMILY_UNIFORM_EQ, FAMILY_ZIPF_MIXED, FAMILY_BLOCK_ANGULAR = 1, 2, 5

BOUND_DTYPE = np.dtype([("kind", "<i4"), ("pad", "<i4"), ("value", "<f8")])

P_i64 = C.POINTER(C.c_int64)
P_i32 = C.POINTER(C.c_int32)
P_f64 = C.POINTER(C.c_double)


class Bound(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad_", C.c_int32), ("value", C.c_double)]


class LP(C.Structure):
    _fields_ = [
        ("rows", C.c_int64), ("cols", C.c_int64),
        ("column_start", P_i64), ("row_index", P_i64), ("value", P_f64), ("objective", P_f64),
        ("row_lower", C.POINTER(Bound)), ("row_upper", C.POINTER(Bound)),
        ("var_lower", C.POINTER(Bound)), ("var_upper", C.POINTER(Bound)),
        ("sense", C.c_int32), ("pad_", C.c_int32), ("objective_offset", C.c_double),
    ]


class Options(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_uint64), ("restart_every", C.c_uint64),
        ("primal_tolerance", C.c_double), ("dual_tolerance", C.c_double), ("gap_tolerance", C.c_double),
        ("step_size_reduction", C.c_double), ("power_iterations", C.c_uint64),
        ("backend", C.c_int32), ("restart_strategy", C.c_int32),
        ("restart_reduction_factor", C.c_double),
        ("adaptive_step_size", C.c_int32), ("adaptive_primal_weight", C.c_int32),
        ("initial_primal_weight", C.c_double), ("primal_weight_smoothing", C.c_double),
        ("ruiz_scaling", C.c_int32), ("autotune", C.c_int32), ("ruiz_iterations", C.c_uint64),
        ("precision", C.c_int32), ("reduction_order", C.c_int32),
        ("polish_threshold", C.c_double),
        ("graph_chunk", C.c_int32), ("device", C.c_int32), ("require_gpu", C.c_int32), ("restart_scheme", C.c_int32),
        ("warm_primal", P_f64), ("warm_dual", P_f64),
        ("fp32_stall_checks", C.c_int32), ("detect_infeasibility", C.c_int32),
        ("infeasibility_tolerance", C.c_double), ("row_weighting", C.c_int32),
        ("primal_weight_max_drop", C.c_int32),
    ]


class Result(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("pad_", C.c_int32),
        ("primal", P_f64), ("dual", P_f64), ("primal_len", C.c_int64), ("dual_len", C.c_int64),
        ("objective", C.c_double), ("primal_infeasibility", C.c_double),
        ("dual_infeasibility", C.c_double), ("duality_gap", C.c_double), ("tolerance", C.c_double),
        ("iterations", C.c_uint64),
        ("h2d_ms", C.c_double), ("kernel_ms", C.c_double), ("d2h_ms", C.c_double), ("total_ms", C.c_double),
        ("message", C.c_char * 128), ("backend_actually_used", C.c_char * 32),
        ("fp32_iterations", C.c_uint64), ("setup_ms", C.c_double),
        ("primal_weight", C.c_double), ("step_size", C.c_double),
        ("packed_iterations", C.c_uint64),
    ]


class VerifyReport(C.Structure):
    _fields_ = [
        ("passed", C.c_int32), ("violation_count", C.c_int32),
        ("recomputed_objective", C.c_double), ("objective_difference", C.c_double),
        ("maximum_row_violation", C.c_double), ("maximum_variable_violation", C.c_double),
        ("first_violation", C.c_char * 160),
    ]


class GenSpec(C.Structure):
    _fields_ = [
        ("family", C.c_int32), ("pad_", C.c_int32), ("seed", C.c_uint64),
        ("rows", C.c_int64), ("cols", C.c_int64), ("nnz_min", C.c_int32), ("nnz_max", C.c_int32),
        ("zipf_s", C.c_double), ("upper_lo", C.c_double), ("upper_hi", C.c_double),
        ("equality_fraction", C.c_double), ("periods", C.c_int32), ("link_nnz", C.c_int32),
        ("link_rows", C.c_int64), ("at_bound_fraction", C.c_double),
    ]


# ---- engine-level structs (mcpdlp_engine.h) -------------------------------
class EngProblem(C.Structure):
    _fields_ = [
        ("instances", C.c_int32), ("rows", C.c_int32), ("cols", C.c_int32), ("nnz", C.c_int64),
        ("inst_row", P_i32), ("inst_col", P_i32),
        ("row_ptr", P_i32), ("col_idx", P_i32), ("row_val", P_f64),
        ("col_ptr", P_i32), ("row_idx", P_i32), ("col_val", P_f64),
        ("row_lower", P_f64), ("row_upper", P_f64), ("var_lower", P_f64), ("var_upper", P_f64),
        ("objective", P_f64), ("sense", P_f64), ("offset", P_f64),
    ]


class EngSetup(C.Structure):
    _fields_ = [("ruiz", C.c_int32), ("ruiz_iterations", C.c_int32), ("exact", C.c_int32), ("autotune", C.c_int32),
                ("row_weighting", C.c_int32), ("pad_", C.c_int32)]


class EngSolve(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_uint64), ("restart_every", C.c_uint64),
        ("primal_tolerance", C.c_double), ("dual_tolerance", C.c_double), ("gap_tolerance", C.c_double),
        ("step_size_reduction", C.c_double),
        ("restart_strategy", C.c_int32), ("adaptive_step_size", C.c_int32),
        ("restart_reduction_factor", C.c_double),
        ("adaptive_primal_weight", C.c_int32), ("precision", C.c_int32),
        ("initial_primal_weight", C.c_double), ("primal_weight_smoothing", C.c_double),
        ("polish_threshold", C.c_double),
        ("reduction_order", C.c_int32), ("graph_chunk", C.c_int32),
        ("restart_scheme", C.c_int32), ("fp32_stall_checks", C.c_int32),
        ("warm_primal", P_f64), ("warm_dual", P_f64),
        ("detect_infeasibility", C.c_int32), ("pad2_", C.c_int32), ("infeasibility_tolerance", C.c_double),
        ("power_iterations", C.c_uint64), ("primal_weight_max_drop", C.c_int32), ("pad3_", C.c_int32),
    ]


class EngStats(C.Structure):
    _fields_ = [
        ("kernel_ms", C.c_double), ("setup_ms", C.c_double), ("h2d_ms", C.c_double),
        ("iterations", C.c_uint64), ("graphs_built", C.c_int32), ("chunks", C.c_int32),
    ]


class EngInstResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("mode", C.c_int32), ("iterations", C.c_uint64), ("fp32_iterations", C.c_uint64),
        ("objective", C.c_double), ("primal_infeasibility", C.c_double), ("dual_infeasibility", C.c_double),
        ("duality_gap", C.c_double), ("omega", C.c_double), ("eta", C.c_double),
        ("packed_iterations", C.c_uint64),
    ]


class TilePolicy(C.Structure):
    _fields_ = [("nnz_budget", C.c_int32), ("per_lane", C.c_int32)]


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None and a.size else C.cast(None, C.POINTER(ct))


_host = None
_engine = None


def host_lib():
    """The C++20 host library (mcpdlp.h).  Raises OSError if it is not built."""
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB):
            raise OSError(f"{HOST_LIB} is not built (run __graft_entry__.build())")
        lib = C.CDLL(HOST_LIB)
        lib.mc_pdlp_options_default.argtypes = [C.POINTER(Options)]
        lib.mc_solve_pdlp.argtypes = [C.POINTER(LP), C.POINTER(Options), C.POINTER(Result)]
        lib.mc_solve_pdlp_batch.argtypes = [C.POINTER(LP), C.c_int32, C.POINTER(Options), C.POINTER(Result)]
        lib.mc_last_error.restype = C.c_char_p
        lib.mc_verify_primal.argtypes = [C.POINTER(LP), P_f64, C.c_double, C.c_double, C.c_double,
                                         C.c_double, C.c_double, C.POINTER(VerifyReport)]
        lib.mc_gen_spec_default.argtypes = [C.POINTER(GenSpec), C.c_int32]
        lib.mc_lp_generate.argtypes = [C.POINTER(GenSpec), C.POINTER(C.c_void_p)]
        lib.mc_lp_view.argtypes = [C.c_void_p, C.POINTER(LP)]
        lib.mc_lp_planted.argtypes = [C.c_void_p, C.POINTER(P_f64), C.POINTER(P_f64), P_f64]
        lib.mc_lp_free.argtypes = [C.c_void_p]
        lib.mc_lp_perturb.argtypes = [C.c_void_p, C.c_uint64, C.c_double, C.POINTER(C.c_void_p)]
        lib.mc_lp_read_mps.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
        _host = lib
    return _host


def engine_lib():
    """The HIP engine library (mcpdlp_engine.h).  Raises OSError if it is not built."""
    global _engine
    if _engine is None:
        if not os.path.exists(ENGINE_LIB):
            raise OSError(f"{ENGINE_LIB} is not built (run __graft_entry__.build())")
        lib = C.CDLL(ENGINE_LIB)
        lib.mcp_device_count.argtypes = [P_i32]
        lib.mcp_engine_create.argtypes = [C.c_int32, C.POINTER(C.c_void_p)]
        lib.mcp_engine_destroy.argtypes = [C.c_void_p]
        lib.mcp_last_error.restype = C.c_char_p
        lib.mcp_load.argtypes = [C.c_void_p, C.POINTER(EngProblem)]
        lib.mcp_setup.argtypes = [C.c_void_p, C.POINTER(EngSetup)]
        lib.mcp_solve.argtypes = [C.c_void_p, C.POINTER(EngSolve), C.POINTER(EngStats)]
        lib.mcp_fetch.argtypes = [C.c_void_p, P_f64, P_f64, C.POINTER(EngInstResult), P_f64]
        lib.mcp_spmv.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, P_f64, P_f64, C.c_int32, P_f64]
        lib.mcp_get_scaling.argtypes = [C.c_void_p, P_f64, P_f64, P_f64, P_i32, P_i32]
        lib.mcp_get_tiling.argtypes = [C.c_void_p, C.POINTER(TilePolicy), C.POINTER(TilePolicy), P_i32, P_i32]
        lib.mcp_time_kernel.argtypes = [C.c_void_p, C.c_int32, C.c_int32, P_f64, P_f64]
        lib.mcp_time_sync.argtypes = [C.c_void_p, C.c_int32, C.c_int32, P_f64, P_f64]
        lib.mcp_set_tile_policy.argtypes = [C.c_void_p, C.POINTER(TilePolicy), C.POINTER(TilePolicy)]
        lib.mcp_set_block_budget.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        lib.mcp_get_block_tiling.argtypes = [C.c_void_p, P_i32, P_i32, P_i32, P_i32]
        _engine = lib
    return _engine


def _check_host(rc):
    if rc != 0:
        raise RuntimeError(f"mcpdlp error {rc}: {host_lib().mc_last_error().decode()}")


def _check_eng(rc):
    if rc != 0:
        raise RuntimeError(f"mcpdlp engine error {rc}: {engine_lib().mcp_last_error().decode()}")


def device_count() -> int:
    c = C.c_int32(0)
    engine_lib().mcp_device_count(C.byref(c))
    return int(c.value)


# ---- model -----------------------------------------------------------------
@dataclass
class LPData:
    """An LP in the reference's storage (CSC, int64 indices, kinded bounds)."""
    rows: int
    cols: int
    column_start: np.ndarray
    row_index: np.ndarray
    value: np.ndarray
    objective: np.ndarray
    row_lower: np.ndarray
    row_upper: np.ndarray
    var_lower: np.ndarray
    var_upper: np.ndarray
    sense: int = MINIMIZE
    offset: float = 0.0
    x_star: np.ndarray | None = None
    y_star: np.ndarray | None = None
    planted_objective: float | None = None
    _keep: list = field(default_factory=list, repr=False)

    @property
    def nnz(self) -> int:
        return int(self.column_start[-1]) if self.cols else 0

    def struct(self) -> LP:
        s = LP()
        s.rows, s.cols = self.rows, self.cols
        s.column_start = _ptr(self.column_start, C.c_int64)
        s.row_index = _ptr(self.row_index, C.c_int64)
        s.value = _ptr(self.value, C.c_double)
        s.objective = _ptr(self.objective, C.c_double)
        s.row_lower = _ptr(self.row_lower, Bound)
        s.row_upper = _ptr(self.row_upper, Bound)
        s.var_lower = _ptr(self.var_lower, Bound)
        s.var_upper = _ptr(self.var_upper, Bound)
        s.sense = self.sense
        s.objective_offset = self.offset
        return s

    def copy(self) -> "LPData":
        return LPData(self.rows, self.cols, self.column_start.copy(), self.row_index.copy(), self.value.copy(),
                      self.objective.copy(), self.row_lower.copy(), self.row_upper.copy(), self.var_lower.copy(),
                      self.var_upper.copy(), self.sense, self.offset, self.x_star, self.y_star,
                      self.planted_objective)

    def dense(self) -> np.ndarray:
        A = np.zeros((self.rows, self.cols))
        for j in range(self.cols):
            for k in range(self.column_start[j], self.column_start[j + 1]):
                A[self.row_index[k], j] += self.value[k]
        return A

    @staticmethod
    def bounds(kinds, values) -> np.ndarray:
        b = np.zeros(len(kinds), dtype=BOUND_DTYPE)
        b["kind"] = kinds
        b["value"] = values
        return b

    @staticmethod
    def from_dense(A, objective, row_lower, row_upper, var_lower, var_upper, sense=MINIMIZE, offset=0.0):
        """Bounds given as floats; +-inf marks an infinite kind."""
        A = np.asarray(A, dtype=np.float64)
        m, n = A.shape
        cs = [0]
        ri, val = [], []
        for j in range(n):
            for i in range(m):
                if A[i, j] != 0.0:
                    ri.append(i)
                    val.append(A[i, j])
            cs.append(len(ri))

        def kb(v, lower):
            v = np.asarray(v, dtype=np.float64)
            kinds = np.where(np.isinf(v), NEG_INF if lower else POS_INF, FINITE)
            return LPData.bounds(kinds, np.where(np.isinf(v), 0.0, v))
        return LPData(m, n, np.array(cs, np.int64), np.array(ri, np.int64), np.array(val, np.float64),
                      np.asarray(objective, np.float64), kb(row_lower, True), kb(row_upper, False),
                      kb(var_lower, True), kb(var_upper, False), sense, offset)


def _lp_from_handle(h: C.c_void_p, planted: bool = True) -> LPData:
    lib = host_lib()
    v = LP()
    _check_host(lib.mc_lp_view(h, C.byref(v)))
    m, n = int(v.rows), int(v.cols)
    cs = np.ctypeslib.as_array(v.column_start, shape=(n + 1,)).copy()
    nnz = int(cs[-1])

    def arr(p, k, dt=np.float64):
        return np.ctypeslib.as_array(p, shape=(k,)).copy() if k else np.zeros(0, dt)

    def barr(p, k):
        if not k:
            return np.zeros(0, BOUND_DTYPE)
        raw = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(16 * k,)).copy()
        return raw.view(BOUND_DTYPE)
    x_star = y_star = planted_objective = None
    if planted:
        xs, ys, po = P_f64(), P_f64(), C.c_double()
        _check_host(lib.mc_lp_planted(h, C.byref(xs), C.byref(ys), C.byref(po)))
        x_star, y_star, planted_objective = arr(xs, n), arr(ys, m), float(po.value)
    return LPData(m, n, cs, arr(v.row_index, nnz, np.int64), arr(v.value, nnz), arr(v.objective, n),
                  barr(v.row_lower, m), barr(v.row_upper, m), barr(v.var_lower, n), barr(v.var_upper, n),
                  int(v.sense), float(v.objective_offset), x_star, y_star, planted_objective)


def read_mps(path: str) -> LPData:
    """LP read from an MPS file by the host library's reader (mc_lp_read_mps)."""
    lib = host_lib()
    h = C.c_void_p()
    _check_host(lib.mc_lp_read_mps(os.fsencode(path), C.byref(h)))
    try:
        return _lp_from_handle(h, planted=False)
    finally:
        lib.mc_lp_free(h)


def generate(family: int = FAMILY_UNIFORM_EQ, **spec) -> LPData:
    """Planted-feasible LP of a family; keyword arguments override GenSpec fields."""
    lib = host_lib()
    sp = GenSpec()
    lib.mc_gen_spec_default(C.byref(sp), family)
    for k, v in spec.items():
        setattr(sp, k, v)
    h = C.c_void_p()
    _check_host(lib.mc_lp_generate(C.byref(sp), C.byref(h)))
    try:
        return _lp_from_handle(h)
    finally:
        lib.mc_lp_free(h)


def generate_batch(count: int = 256, m_lo: int = 500, m_hi: int = 20000, seed: int = 0x4D4350444C50 + 300):
    """Uneven batch of planted LPs: m log-uniform in [m_lo, m_hi], n = 2.5 m,
    4-8 nonzeros per column (+-U[0.1,10]), equality rows, 0 <= x <= U[1,100]."""
    rng = np.random.default_rng(seed)
    ms = np.rint(np.exp(rng.uniform(np.log(m_lo), np.log(m_hi), size=count))).astype(int)
    return [generate(FAMILY_UNIFORM_EQ, rows=int(m), cols=int(round(2.5 * m)), nnz_min=4, nnz_max=8,
                     upper_lo=1.0, upper_hi=100.0, seed=int(seed + k)) for k, m in enumerate(ms)]


def generate_family(base_spec: dict, seeds, sigma: float = 0.1, family: int = FAMILY_UNIFORM_EQ):
    """Base planted LP plus perturbed siblings (learned-warm-start family)."""
    lib = host_lib()
    sp = GenSpec()
    lib.mc_gen_spec_default(C.byref(sp), family)
    for k, v in base_spec.items():
        setattr(sp, k, v)
    h = C.c_void_p()
    _check_host(lib.mc_lp_generate(C.byref(sp), C.byref(h)))
    try:
        out = []
        for s in seeds:
            ph = C.c_void_p()
            _check_host(lib.mc_lp_perturb(h, int(s), float(sigma), C.byref(ph)))
            try:
                out.append(_lp_from_handle(ph))
            finally:
                lib.mc_lp_free(ph)
        return _lp_from_handle(h), out
    finally:
        lib.mc_lp_free(h)


# ---- options / solve -------------------------------------------------------
def options(**kw) -> Options:
    o = Options()
    host_lib().mc_pdlp_options_default(C.byref(o))
    tol = kw.pop("tolerance", None)
    if tol is not None:
        o.primal_tolerance = o.dual_tolerance = o.gap_tolerance = tol
    keep = []
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        if k in ("warm_primal", "warm_dual") and v is not None and not isinstance(v, C._Pointer):
            arr = np.ascontiguousarray(v, dtype=np.float64)
            keep.append(arr)  # the options borrow the array; keep it alive with them
            v = _ptr(arr, C.c_double)
        setattr(o, k, v)
    o._keep = keep
    return o


def _result_dict(r: Result, prim, dual) -> dict:
    return {
        "status": STATUS[r.status], "objective": r.objective,
        "primal": prim[: r.primal_len].copy(), "dual": dual[: r.dual_len].copy(),
        "primal_infeasibility": r.primal_infeasibility, "dual_infeasibility": r.dual_infeasibility,
        "duality_gap": r.duality_gap, "tolerance": r.tolerance, "iterations": int(r.iterations),
        "h2d_ms": r.h2d_ms, "kernel_ms": r.kernel_ms, "d2h_ms": r.d2h_ms, "total_ms": r.total_ms,
        "message": r.message.decode(), "backend_actually_used": r.backend_actually_used.decode(),
        "fp32_iterations": int(r.fp32_iterations), "packed_iterations": int(r.packed_iterations),
        "setup_ms": r.setup_ms,
        "primal_weight": r.primal_weight, "step_size": r.step_size,
    }


def solve(lp: LPData, opts: Options | None = None, **kw) -> dict:
    """solve_pdlp through the C-ABI; returns the PdlpResult fields as a dict."""
    o = opts if opts is not None else options(**kw)
    prim = np.zeros(max(lp.cols, 1))
    dual = np.zeros(max(lp.rows, 1))
    r = Result()
    r.primal = _ptr(prim, C.c_double)
    r.dual = _ptr(dual, C.c_double)
    s = lp.struct()
    _check_host(host_lib().mc_solve_pdlp(C.byref(s), C.byref(o), C.byref(r)))
    return _result_dict(r, prim, dual)


def solve_batch(lps, opts: Options | None = None, **kw) -> list:
    o = opts if opts is not None else options(**kw)
    k = len(lps)
    structs = (LP * k)(*[lp.struct() for lp in lps])
    res = (Result * k)()
    bufs = []
    for i, lp in enumerate(lps):
        p, d = np.zeros(max(lp.cols, 1)), np.zeros(max(lp.rows, 1))
        bufs.append((p, d))
        res[i].primal = _ptr(p, C.c_double)
        res[i].dual = _ptr(d, C.c_double)
    _check_host(host_lib().mc_solve_pdlp_batch(structs, k, C.byref(o), res))
    return [_result_dict(res[i], *bufs[i]) for i in range(k)]


def verify(lp: LPData, x, objective: float, tol: float = 1e-4) -> dict:
    rep = VerifyReport()
    x = np.ascontiguousarray(x, dtype=np.float64)
    s = lp.struct()
    _check_host(host_lib().mc_verify_primal(C.byref(s), _ptr(x, C.c_double), objective, tol, tol, tol, tol,
                                            C.byref(rep)))
    return {"passed": bool(rep.passed), "violations": rep.violation_count,
            "recomputed_objective": rep.recomputed_objective,
            "max_row_violation": rep.maximum_row_violation,
            "max_variable_violation": rep.maximum_variable_violation,
            "first_violation": rep.first_violation.decode()}


# ---- engine (device-level) access ------------------------------------------
def _length_bucket(lens: np.ndarray) -> np.ndarray:
    """Exact below 32 nonzeros, log2 classes above (host/pdlp_gpu.cpp length_bucket)."""
    lens = np.asarray(lens, np.int64)
    big = lens >= 32
    out = lens.copy()
    out[big] = 32 + np.floor(np.log2(lens[big] / 32.0)).astype(np.int64)
    return out


def rows_are_local(lp: LPData) -> bool:
    """The host library's locality test (host/pdlp_gpu.cpp): the mean over
    non-empty rows of (largest - smallest column index + 1) / columns is
    below 1/4."""
    if lp.rows == 0 or lp.cols == 0 or lp.nnz == 0:
        return False
    col_of = np.repeat(np.arange(lp.cols), np.diff(lp.column_start))
    lo = np.full(lp.rows, lp.cols, np.int64)
    hi = np.full(lp.rows, -1, np.int64)
    np.minimum.at(lo, lp.row_index, col_of)
    np.maximum.at(hi, lp.row_index, col_of)
    used = hi >= 0
    return float(np.mean((hi[used] - lo[used] + 1) / lp.cols)) < 0.25


def relabel_by_length(lp: LPData) -> LPData:
    """The same LP with rows and columns stably sorted by descending length
    bucket -- the relabelling the host library applies before mcp_load in the
    fast reduction order, so Engine users can time kernels on the product's
    layout.  Like the host, it keeps the natural order (returns `lp` itself)
    when the rows are local (rows_are_local).  Otherwise returns an LP without
    a planted pair."""
    if rows_are_local(lp):
        return lp
    clen = np.diff(lp.column_start)
    rlen = np.bincount(lp.row_index, minlength=lp.rows)
    cperm = np.argsort(-_length_bucket(clen), kind="stable")  # new -> old
    rperm = np.argsort(-_length_bucket(rlen), kind="stable")
    rinv = np.empty(lp.rows, np.int64)
    rinv[rperm] = np.arange(lp.rows)
    cs = np.zeros(lp.cols + 1, np.int64)
    cs[1:] = np.cumsum(clen[cperm])
    ri = np.empty(lp.nnz, np.int64)
    val = np.empty(lp.nnz)
    col_of = np.repeat(np.arange(lp.cols), clen)          # old column of each entry
    newc = np.empty(lp.cols, np.int64)
    newc[cperm] = np.arange(lp.cols)
    order = np.lexsort((rinv[lp.row_index], newc[col_of]))
    ri[:] = rinv[lp.row_index[order]]
    val[:] = lp.value[order]
    return LPData(lp.rows, lp.cols, cs, ri, val, lp.objective[cperm].copy(), lp.row_lower[rperm].copy(),
                  lp.row_upper[rperm].copy(), lp.var_lower[cperm].copy(), lp.var_upper[cperm].copy(),
                  lp.sense, lp.offset)


def _csc_to_csr(m, n, cs, ri, val):
    cols = np.repeat(np.arange(n, dtype=np.int64), np.diff(cs))
    order = np.lexsort((cols, ri))  # by row, then column
    rp = np.zeros(m + 1, np.int64)
    np.add.at(rp, ri + 1, 1)
    return np.cumsum(rp).astype(np.int32), cols[order].astype(np.int32), val[order].copy()


class Engine:
    """Direct handle on the HIP engine (mcpdlp_engine.h) for kernel-level tests."""

    def __init__(self, device: int = 0):
        self.lib = engine_lib()
        self.h = C.c_void_p()
        _check_eng(self.lib.mcp_engine_create(device, C.byref(self.h)))
        self._keep = []

    def close(self):
        if self.h:
            self.lib.mcp_engine_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, lps):
        lps = lps if isinstance(lps, (list, tuple)) else [lps]
        ir, ic = [0], [0]
        rps, cis, rvs, cps, ris, cvs = [], [], [], [], [], []
        rlo, rhi, vlo, vhi, obj, sense, off = [], [], [], [], [], [], []
        p0 = r0 = c0 = 0
        for lp in lps:
            rp, ci, rv = _csc_to_csr(lp.rows, lp.cols, lp.column_start, lp.row_index, lp.value)
            rps.append(rp[:-1] + p0) if lp.rows else None
            cis.append(ci + c0)
            rvs.append(rv)
            cps.append(lp.column_start[:-1].astype(np.int32) + p0)
            ris.append(lp.row_index.astype(np.int32) + r0)
            cvs.append(lp.value)

            def fb(b, lower):
                v = b["value"].astype(np.float64).copy()
                v[b["kind"] != FINITE] = -np.inf if lower else np.inf
                return v
            rlo.append(fb(lp.row_lower, True))
            rhi.append(fb(lp.row_upper, False))
            vlo.append(fb(lp.var_lower, True))
            vhi.append(fb(lp.var_upper, False))
            obj.append(lp.objective)
            sense.append(-1.0 if lp.sense == MAXIMIZE else 1.0)
            off.append(lp.offset)
            p0 += lp.nnz
            r0 += lp.rows
            c0 += lp.cols
            ir.append(r0)
            ic.append(c0)
        cat = lambda xs, dt: np.ascontiguousarray(np.concatenate(xs) if xs else np.zeros(0), dtype=dt)
        arrays = dict(
            inst_row=np.array(ir, np.int32), inst_col=np.array(ic, np.int32),
            row_ptr=np.append(cat(rps, np.int32), np.int32(p0)), col_idx=cat(cis, np.int32),
            row_val=cat(rvs, np.float64),
            col_ptr=np.append(cat(cps, np.int32), np.int32(p0)), row_idx=cat(ris, np.int32),
            col_val=cat(cvs, np.float64),
            row_lower=cat(rlo, np.float64), row_upper=cat(rhi, np.float64),
            var_lower=cat(vlo, np.float64), var_upper=cat(vhi, np.float64),
            objective=cat(obj, np.float64), sense=np.array(sense), offset=np.array(off))
        p = EngProblem()
        p.instances, p.rows, p.cols, p.nnz = len(lps), r0, c0, p0
        for k, a in arrays.items():
            setattr(p, k, _ptr(a, C.c_int32 if a.dtype == np.int32 else C.c_double))
        self._keep = [arrays]
        self.m, self.n, self.B = r0, c0, len(lps)
        self.inst_row, self.inst_col = arrays["inst_row"], arrays["inst_col"]
        _check_eng(self.lib.mcp_load(self.h, C.byref(p)))

    def setup(self, ruiz=True, ruiz_iterations=10, exact=False, autotune=False, row_weighting=0):
        s = EngSetup(int(ruiz), int(ruiz_iterations), int(exact), int(autotune), int(row_weighting), 0)
        _check_eng(self.lib.mcp_setup(self.h, C.byref(s)))

    def spmv(self, v, transpose=False, fp32=False, exact=False, reps=1, packed=False):
        """A v (or A^T v) of the scaled matrix: FP64, FP32 (fp32) or on the packed
        stream (packed: FP32 value + 8-bit correction, FP64 vector and sums)."""
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros(self.n if transpose else self.m)
        us = C.c_double(0.0)
        prec = PRECISION_PACKED if packed else (PRECISION_MIXED if fp32 else PRECISION_FP64)
        _check_eng(self.lib.mcp_spmv(self.h, int(transpose), prec,
                                     ORDER_REFERENCE if exact else ORDER_FAST, _ptr(v, C.c_double),
                                     _ptr(out, C.c_double), int(reps), C.byref(us)))
        return out, us.value

    def scaling(self, nnz):
        cv, rs, cs = np.zeros(max(nnz, 1)), np.zeros(max(self.m, 1)), np.zeros(max(self.n, 1))
        passes, conv = C.c_int32(), C.c_int32()
        _check_eng(self.lib.mcp_get_scaling(self.h, _ptr(cv, C.c_double), _ptr(rs, C.c_double),
                                            _ptr(cs, C.c_double), C.byref(passes), C.byref(conv)))
        return cv[:nnz], rs[: self.m], cs[: self.n], passes.value, conv.value

    def set_tiling(self, rows=(1024, 4), cols=(1024, 4)):
        a, t = TilePolicy(*rows), TilePolicy(*cols)
        _check_eng(self.lib.mcp_set_tile_policy(self.h, C.byref(a), C.byref(t)))

    def tiling(self):
        a, t = TilePolicy(), TilePolicy()
        na, nt = C.c_int32(), C.c_int32()
        _check_eng(self.lib.mcp_get_tiling(self.h, C.byref(a), C.byref(t), C.byref(na), C.byref(nt)))
        return (a.nnz_budget, a.per_lane, na.value), (t.nnz_budget, t.per_lane, nt.value)

    def set_block_budget(self, rows=1024, cols=1024):
        """Nonzeros per block tile of the fused fast-order steps (256/512/1024/2048),
        or 0 for both to run those steps on the wave tiles."""
        _check_eng(self.lib.mcp_set_block_budget(self.h, int(rows), int(cols)))

    def block_tiling(self):
        """((rows budget, row tiles), (cols budget, col tiles)) of the block tiles."""
        ba, bt, na, nt = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        _check_eng(self.lib.mcp_get_block_tiling(self.h, C.byref(ba), C.byref(bt), C.byref(na), C.byref(nt)))
        return (ba.value, na.value), (bt.value, nt.value)

    # gather_fp64 / gather_local_fp64 / stream_fp64: the FP64 block SpMV A x_bar on the
    # solve's tiles with random gathers, with every gather inside a 256-entry window
    # (cache hits), and with no gathers -- what the random gathers cost on this matrix
    KERNELS = {"primal_fp32": 0, "dual_fp32": 1, "primal_fp64": 2, "dual_fp64": 3,
               "spmv_fp32": 4, "spmv_t_fp32": 5, "copy": 6,
               "gather_fp64": 7, "gather_local_fp64": 8, "stream_fp64": 9,
               "primal_packed": 10, "dual_packed": 11}

    def time_kernel(self, kind: str, reps: int = 20):
        """(mean device us per launch, algorithmic bytes per launch) of one fused kernel."""
        us, b = C.c_double(), C.c_double()
        _check_eng(self.lib.mcp_time_kernel(self.h, self.KERNELS[kind], int(reps), C.byref(us), C.byref(b)))
        return us.value, b.value

    def time_sync(self, workgroups, rounds=2000):
        """(us per grid-wide barrier in a cooperative persistent kernel of
        `workgroups` workgroups, us per dependent launch in a captured graph)."""
        bu, lu = C.c_double(), C.c_double()
        _check_eng(self.lib.mcp_time_sync(self.h, int(workgroups), int(rounds), C.byref(bu), C.byref(lu)))
        return bu.value, lu.value

    def solve(self, tolerance=1e-4, max_iterations=100000, precision=PRECISION_FP64, order=ORDER_FAST,
              graph_chunk=0, **kw):
        o = EngSolve()
        d = options()
        o.max_iterations, o.restart_every = max_iterations, d.restart_every
        o.primal_tolerance = o.dual_tolerance = o.gap_tolerance = tolerance
        o.step_size_reduction = d.step_size_reduction
        o.restart_strategy, o.adaptive_step_size = RESTART_ADAPTIVE, 1
        o.restart_reduction_factor = d.restart_reduction_factor
        o.adaptive_primal_weight, o.precision = 1, precision
        o.initial_primal_weight, o.primal_weight_smoothing = 0.0, 0.5
        o.polish_threshold, o.reduction_order, o.graph_chunk = 0.0, order, graph_chunk
        o.power_iterations = d.power_iterations
        for k, v in kw.items():
            setattr(o, k, v)
        st = EngStats()
        _check_eng(self.lib.mcp_solve(self.h, C.byref(o), C.byref(st)))
        return {"kernel_ms": st.kernel_ms, "setup_ms": st.setup_ms, "h2d_ms": st.h2d_ms,
                "iterations": int(st.iterations), "graphs_built": st.graphs_built, "chunks": st.chunks}

    def fetch(self):
        x, y = np.zeros(max(self.n, 1)), np.zeros(max(self.m, 1))
        res = (EngInstResult * self.B)()
        d2h = C.c_double()
        _check_eng(self.lib.mcp_fetch(self.h, _ptr(x, C.c_double), _ptr(y, C.c_double), res, C.byref(d2h)))
        out = []
        for b in range(self.B):
            r = res[b]
            out.append({"status": STATUS[r.status], "mode": r.mode, "iterations": int(r.iterations),
                        "fp32_iterations": int(r.fp32_iterations),
                        "packed_iterations": int(r.packed_iterations), "objective": r.objective,
                        "primal_infeasibility": r.primal_infeasibility,
                        "dual_infeasibility": r.dual_infeasibility, "duality_gap": r.duality_gap,
                        "omega": r.omega, "eta": r.eta,
                        "x": x[self.inst_col[b]:self.inst_col[b + 1]].copy(),
                        "y": y[self.inst_row[b]:self.inst_row[b + 1]].copy()})
        return out, d2h.value


# ---- learned warm start network (mcpdlp_mlp.h) ------------------------------
def _mlp_lib():
    lib = engine_lib()
    if not getattr(lib, "_mlp_bound", False):
        lib.mcp_mlp_create.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64,
                                       C.POINTER(C.c_void_p)]
        lib.mcp_mlp_destroy.argtypes = [C.c_void_p]
        lib.mcp_mlp_last_error.restype = C.c_char_p
        lib.mcp_mlp_param_count.argtypes = [C.c_void_p, P_i64]
        lib.mcp_mlp_get_params.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
        lib.mcp_mlp_set_params.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
        lib.mcp_mlp_forward.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_int32, C.POINTER(C.c_float)]
        lib.mcp_mlp_loss_grad.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int32,
                                          C.POINTER(C.c_float), C.POINTER(C.c_float)]
        lib.mcp_mlp_train.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int32, C.c_int32,
                                      C.c_int32, C.c_float, C.c_uint64, C.POINTER(C.c_float), P_f64]
        lib.mcp_gemm_bf16_test.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_float),
                                           C.POINTER(C.c_float), C.c_int32, C.c_int32, C.POINTER(C.c_float)]
        lib._mlp_bound = True
    return lib


def _chk_mlp(rc):
    if rc != 0:
        raise RuntimeError(f"mcpdlp mlp error {rc}: {_mlp_lib().mcp_mlp_last_error().decode()}")


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def gemm_bf16(A, B, transpose_a=False, transpose_b=False, device=0):
    """op(A) @ op(B) on MFMA with A, B rounded to BF16 (FP32 accumulation)."""
    A, B = _f32(A), _f32(B)
    M = A.shape[1] if transpose_a else A.shape[0]
    K = A.shape[0] if transpose_a else A.shape[1]
    N = B.shape[0] if transpose_b else B.shape[1]
    out = np.zeros((M, N), np.float32)
    _chk_mlp(_mlp_lib().mcp_gemm_bf16_test(device, M, N, K, _ptr(A, C.c_float), _ptr(B, C.c_float),
                                           int(transpose_a), int(transpose_b), _ptr(out, C.c_float)))
    return out


class MLP:
    """BF16 MLP on MFMA: in -> (hidden, ReLU) x layers -> out (widths multiples of 32)."""

    def __init__(self, in_dim=256, hidden=256, layers=3, out_dim=64, capacity=256, seed=1, device=0):
        self.lib = _mlp_lib()
        self.h = C.c_void_p()
        _chk_mlp(self.lib.mcp_mlp_create(device, in_dim, hidden, layers, out_dim, capacity, seed, C.byref(self.h)))
        n = C.c_int64()
        self.lib.mcp_mlp_param_count(self.h, C.byref(n))
        self.nparam = n.value
        self.dims = [in_dim] + [hidden] * layers + [out_dim]
        self.capacity = capacity

    def close(self):
        if self.h:
            self.lib.mcp_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.zeros(self.nparam, np.float32)
        _chk_mlp(self.lib.mcp_mlp_get_params(self.h, _ptr(p, C.c_float)))
        return p

    def set_params(self, p):
        p = _f32(p)
        _chk_mlp(self.lib.mcp_mlp_set_params(self.h, _ptr(p, C.c_float)))

    def unpack(self, p):
        """List of (W [in x out], b [out]) per layer from a flat parameter vector."""
        out, off = [], 0
        for a, b in zip(self.dims[:-1], self.dims[1:]):
            W = p[off:off + a * b].reshape(a, b)
            off += a * b
            out.append((W, p[off:off + b]))
            off += b
        return out

    def forward(self, x):
        x = _f32(x)
        y = np.zeros((x.shape[0], self.dims[-1]), np.float32)
        _chk_mlp(self.lib.mcp_mlp_forward(self.h, _ptr(x, C.c_float), x.shape[0], _ptr(y, C.c_float)))
        return y

    def loss_grad(self, x, t):
        x, t = _f32(x), _f32(t)
        loss = C.c_float()
        g = np.zeros(self.nparam, np.float32)
        _chk_mlp(self.lib.mcp_mlp_loss_grad(self.h, _ptr(x, C.c_float), _ptr(t, C.c_float), x.shape[0],
                                            C.byref(loss), _ptr(g, C.c_float)))
        return loss.value, g

    def train(self, x, t, batch=64, epochs=50, lr=1e-3, seed=7):
        x, t = _f32(x), _f32(t)
        loss, ms = C.c_float(), C.c_double()
        _chk_mlp(self.lib.mcp_mlp_train(self.h, _ptr(x, C.c_float), _ptr(t, C.c_float), x.shape[0], batch, epochs,
                                        lr, seed, C.byref(loss), C.byref(ms)))
        return loss.value, ms.value
