/*
 * mcpdlp.h -- C-ABI of the MI355X PDLP engine for markov-cero.
 *
 * This is the drop-in boundary for `markov_cero::lp::first_order::solve_pdlp`
 * (reference: include/markov_cero/lp/first_order/pdlp.hpp:67).  The C++20
 * host library (libmarkov_cero_pdlp.so) exports every function declared here;
 * the structs are plain data with the same fields, meanings and defaults as the
 * reference's C++ types, so an FFI (ctypes, cgo, JNI) can drive the solver with
 * plain pointers and sizes.  No function throws across this boundary: each
 * returns an int status (0 = success, see MC_ERR_*) and stores a message
 * readable through mc_last_error() (thread-local).
 */
#ifndef MCPDLP_H_
#define MCPDLP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
#define MC_OK 0
#define MC_ERR_INVALID_ARGUMENT 1 /* malformed model / options (std::invalid_argument) */
#define MC_ERR_OUT_OF_MEMORY 2    /* host or device allocation failed             */
#define MC_ERR_DEVICE 3           /* HIP runtime error or no usable device        */
#define MC_ERR_INTERNAL 4         /* any other exception                          */

/* ---- model --------------------------------------------------------------- */
/* BoundKind (reference: include/markov_cero/model/model.hpp:13). */
#define MC_BOUND_NEGATIVE_INFINITY 0
#define MC_BOUND_FINITE 1
#define MC_BOUND_POSITIVE_INFINITY 2

/* Bound (reference: model.hpp:15-23).  Same layout as model::Bound:
 * a 4-byte kind, 4 bytes of padding, an 8-byte value; 16 bytes in total. */
typedef struct mc_bound {
    int32_t kind;
    int32_t pad_;
    double value;
} mc_bound;

/* ObjectiveSense (reference: model.hpp:11). */
#define MC_MINIMIZE 0
#define MC_MAXIMIZE 1

/* A linear program in the reference's storage: CSC matrix with 64-bit
 * indices (reference SparseMatrixCSC, model.hpp:25-34), objective, four
 * bound vectors and the objective sense/offset (Model, model.hpp:53-70).
 * Names and variable types are not needed by PDLP and are not carried.
 * All pointers are borrowed; the solver never writes through them. */
typedef struct mc_lp {
    int64_t rows;                 /* m */
    int64_t cols;                 /* n */
    const int64_t* column_start;  /* [n+1], column_start[0] == 0 */
    const int64_t* row_index;     /* [nnz], strictly increasing inside a column */
    const double* value;          /* [nnz], finite and nonzero */
    const double* objective;      /* [n] */
    const mc_bound* row_lower;    /* [m] */
    const mc_bound* row_upper;    /* [m] */
    const mc_bound* var_lower;    /* [n] */
    const mc_bound* var_upper;    /* [n] */
    int32_t sense;                /* MC_MINIMIZE / MC_MAXIMIZE */
    int32_t pad_;
    double objective_offset;
} mc_lp;

/* ---- options / result ---------------------------------------------------- */
/* Backend (reference: pdlp.hpp:15). */
#define MC_BACKEND_CPU 0
#define MC_BACKEND_GPU 1
/* RestartStrategy (reference: pdlp.hpp:16). */
#define MC_RESTART_NONE 0
#define MC_RESTART_ADAPTIVE 1
#define MC_RESTART_FIXED 2
/* Arithmetic of the GPU engine (extension; no reference counterpart). */
#define MC_PRECISION_FP64 0  /* every iteration in FP64                                   */
#define MC_PRECISION_MIXED 1 /* FP32 SpMVs (values, gathered vectors, products and sums); the
                              * iterate, its updates and every check stay FP64; all-FP64
                              * iterations once the FP32-SpMV phase stalls              */
#define MC_PRECISION_PACKED 2 /* SpMV values packed in 8 bytes per nonzero with the index:
                               * the FP32 value plus an 8-bit correction in the index word's
                               * top byte (about 31 significant bits); FP64 gathered vectors,
                               * products and sums, FP64 iterate and checks; all-FP64
                               * iterations if this phase stalls.  Needs m, n < 2^24.    */
/* Restart scheme (extension; both backends).
 * REFERENCE: the reference's rule (pdlp.cpp:246-260): restart to the average when
 *   its score fell to 0.368 x the last restart's, primal weight from sqrt(primal/dual).
 * KKT: the adaptive restarts of Applegate et al. (2021, "Practical large-scale linear
 *   programming using primal-dual hybrid gradient", sec. 3): candidate = the better of
 *   the current iterate and the average by the same relative KKT score; restart on
 *   sufficient decay (<= 0.2 x the score at the last restart), on necessary decay
 *   (<= 0.8 x) without progress since the previous check, or when the epoch reaches
 *   0.36 x all iterations; primal weight from the movement since the last restart,
 *   omega <- exp(theta log(|dy|/|dx|) + (1-theta) log omega), theta = primal_weight_smoothing.
 * HALPERN: the KKT scheme's restarts on reflected Halpern iterations (Halpern 1967;
 *   the reflected PDHG form of Lu & Yang, 2024-25): with T the PDHG step
 *   (x, y) -> (x+, y+) and z0 the point of the last restart,
 *   z <- w (2 T(z) - z) + (1 - w) z0, w = (k+1)/(k+2) after k iterations since
 *   the restart; the check compares T(z) (in place of the average) with z.  The
 *   primal weight at a restart is the geometric blend
 *   omega <- omega_move^(3/4) omega_res^(1/4) of the movement rule above and the
 *   reference's residual-balance rule omega_res = omega sqrt(clamp(sqrt(p/d), 0.05, 20)),
 *   except that it never moves against a residual imbalance of 100x or more: while
 *   p > 100 d it does not decrease, while d > 100 p it does not increase.
 * Termination, tolerances and the reported score are the reference's in all three. */
#define MC_SCHEME_REFERENCE 0
#define MC_SCHEME_KKT 1
#define MC_SCHEME_HALPERN 2
/* Reduction order of the GPU engine (extension). */
#define MC_ORDER_FAST 0      /* tile-parallel sums (deterministic, not reference order)   */
#define MC_ORDER_REFERENCE 1 /* every sum in the reference's serial order: bit-exact FP64 */

/* PdlpOptions (reference: pdlp.hpp:18-41), same defaults.  Fields after
 * `ruiz_iterations` are extensions for the HIP engine. */
typedef struct mc_pdlp_options {
    uint64_t max_iterations;          /* 100000 */
    uint64_t restart_every;           /* 40 */
    double primal_tolerance;          /* 1e-4 */
    double dual_tolerance;            /* 1e-4 */
    double gap_tolerance;             /* 1e-4 */
    double step_size_reduction;       /* 0.9 */
    uint64_t power_iterations;        /* 20 (never read, as in the reference) */
    int32_t backend;                  /* MC_BACKEND_CPU */
    int32_t restart_strategy;         /* MC_RESTART_ADAPTIVE */
    double restart_reduction_factor;  /* 0.368 */
    int32_t adaptive_step_size;       /* 1 */
    int32_t adaptive_primal_weight;   /* 1 */
    double initial_primal_weight;     /* 0.0 => automatic */
    double primal_weight_smoothing;   /* 0.5 */
    int32_t ruiz_scaling;             /* 1 */
    int32_t autotune;                 /* extension: time SpMV tile shapes on the device first (0) */
    uint64_t ruiz_iterations;         /* 10 */
    /* --- extensions: no counterpart in the reference's PdlpOptions (pdlp.hpp:19-34); precision,
       reduction_order, polish_threshold, graph_chunk, device, require_gpu and fp32_stall_checks act on
       the HIP engine only, the others on both backends --- */
    int32_t precision;                /* MC_PRECISION_FP64 */
    int32_t reduction_order;          /* MC_ORDER_FAST */
    double polish_threshold;          /* mixed: switch to FP64 once the FP64-checked score <= this (0 => only on stall) */
    int32_t graph_chunk;              /* iterations per captured HIP graph (0 => 64; < 0 => eager launches) */
    int32_t device;                   /* HIP device ordinal */
    int32_t require_gpu;              /* 1: a GPU request fails with MC_ERR_DEVICE instead of the cpu_fallback path */
    int32_t restart_scheme;           /* MC_SCHEME_REFERENCE (default), MC_SCHEME_KKT or MC_SCHEME_HALPERN */
    const double* warm_primal;        /* optional [n] start point in the ORIGINAL space (NULL => cold x=proj(0));
                                         a warm-started solve runs no FP32-SpMV phase */
    const double* warm_dual;          /* optional [m] start dual in the ORIGINAL space (reported sign convention) */
    int32_t fp32_stall_checks;        /* mixed: checks without a new best score before the FP64 switch (0 => 50) */
    int32_t detect_infeasibility;     /* 1: report MC_PDLP_INFEASIBLE_OR_UNBOUNDED when the iterate's movement
                                         since the last restart is a primal or dual ray certificate (0: never,
                                         as the reference) */
    double infeasibility_tolerance;   /* certificate tolerance (0 => 1e-8) */
    int32_t row_weighting;            /* k in 0..8: after the Ruiz passes, row i of A and its bounds are scaled by
                                         (1 + beta_i)^(-k/8), beta_i the largest magnitude of its finite original
                                         bounds; the reported residuals and vectors stay those of the original
                                         model (0: the reference's scaling); both backends */
    int32_t primal_weight_max_drop;   /* Halpern scheme: a restart lowers the primal weight by at most this factor
                                         (0: no limit; else a power of two in 2..2^20); both backends */
} mc_pdlp_options;

/* PdlpStatus (reference: pdlp.hpp:43). */
#define MC_PDLP_OPTIMAL 0
#define MC_PDLP_ITERATION_LIMIT 1
#define MC_PDLP_INFEASIBLE_OR_UNBOUNDED 2
#define MC_PDLP_NUMERICAL_FAILURE 3

/* PdlpResult (reference: pdlp.hpp:45-63).  `primal`/`dual` are caller-owned
 * buffers of at least `cols`/`rows` doubles (may be NULL to skip the copy);
 * primal_len/dual_len report how many were written (0 for the trivial m==0 or
 * n==0 exit, as in the reference). */
typedef struct mc_pdlp_result {
    int32_t status;
    int32_t pad_;
    double* primal;
    double* dual;
    int64_t primal_len;
    int64_t dual_len;
    double objective;
    double primal_infeasibility;
    double dual_infeasibility;
    double duality_gap;
    double tolerance;
    uint64_t iterations;
    double h2d_ms;
    double kernel_ms;
    double d2h_ms;
    double total_ms;
    char message[128];
    char backend_actually_used[32]; /* "cpu", "hip" or "cpu_fallback" */
    /* --- extensions --- */
    uint64_t fp32_iterations;       /* iterations with FP32 SpMVs (mixed)                */
    double setup_ms;                /* Ruiz + transpose + norms on the device          */
    double primal_weight;           /* final primal weight omega (scaled space)        */
    double step_size;               /* final step-size factor eta                      */
    uint64_t packed_iterations;     /* iterations on the packed stream (MC_PRECISION_PACKED, or
                                     * mixed after its FP32 phase)                     */
} mc_pdlp_result;

/* Fill `opt` with the reference defaults (pdlp.hpp:19-34) and the extension defaults. */
void mc_pdlp_options_default(mc_pdlp_options* opt);

/* solve_pdlp (reference: pdlp.hpp:67, src/lp/first_order/pdlp.cpp:139-276).
 * Backend::cpu runs the host restatement; Backend::gpu runs the HIP engine and
 * reports "hip", or -- when no device is usable and require_gpu == 0 -- runs
 * the host path and reports "cpu_fallback" (reference behaviour, pdlp.cpp:237,268). */
int mc_solve_pdlp(const mc_lp* lp, const mc_pdlp_options* opt, mc_pdlp_result* res);

/* Batched extension: solve `count` independent LPs; with Backend::gpu they
 * advance together (one launch per phase for the whole batch, per-instance
 * state).  res[k].primal/dual are caller buffers as for mc_solve_pdlp. */
int mc_solve_pdlp_batch(const mc_lp* lps, int32_t count, const mc_pdlp_options* opt,
                        mc_pdlp_result* res);

/* Message of the last failed call on this thread ("" if none). */
const char* mc_last_error(void);

/* ---- verification (reference: src/verify/primal_verifier.cpp:44-104) ----- */
typedef struct mc_verify_report {
    int32_t passed;
    int32_t violation_count;
    double recomputed_objective;
    double objective_difference;
    double maximum_row_violation;
    double maximum_variable_violation;
    char first_violation[160];
} mc_verify_report;

/* verify_primal with Tolerance{abs,rel} for feasibility and objective; variables
 * are continuous (PDLP models), so the integrality check never fires. */
int mc_verify_primal(const mc_lp* lp, const double* primal, double claimed_objective,
                     double feas_abs, double feas_rel, double obj_abs, double obj_rel,
                     mc_verify_report* report);

/* ---- synthetic planted-feasible LP families (extension) --------------------
 * Workload generators for the benchmark shapes.  Every instance is feasible by
 * construction and its planted optimum c^T x* is exact.  The generator owns
 * the arrays; mc_lp_view() borrows them until mc_lp_free(). */
typedef struct mc_lp_handle mc_lp_handle;

#define MC_FAMILY_UNIFORM_EQ 1   /* equality rows, k nnz/col, +-U[0.1,10], box bounds      */
#define MC_FAMILY_ZIPF_MIXED 2   /* 70% '=' / 30% '<=' rows, Zipf(1.2) nnz/col, log-uniform |a| */
#define MC_FAMILY_BLOCK_ANGULAR 5 /* periods x blocks + linking rows                       */

typedef struct mc_gen_spec {
    int32_t family;
    int32_t pad_;
    uint64_t seed;
    int64_t rows;          /* uniform/zipf: m;  block-angular: rows per period block   */
    int64_t cols;          /* uniform/zipf: n;  block-angular: cols per period block   */
    int32_t nnz_min;       /* uniform: nnz per column drawn from [nnz_min, nnz_max]     */
    int32_t nnz_max;       /* zipf: truncation max (min = 1); block: nnz/col in block   */
    double zipf_s;         /* zipf exponent (1.2)                                        */
    double upper_lo;       /* variable upper bound drawn from U[upper_lo, upper_hi]      */
    double upper_hi;       /*   (equal values => fixed upper bound, e.g. 10)             */
    double equality_fraction; /* zipf: share of '=' rows (0.7)                           */
    int32_t periods;       /* block-angular: number of periods (48)                      */
    int32_t link_nnz;      /* block-angular: nnz per column in linking rows (2)          */
    int64_t link_rows;     /* block-angular: total linking rows (200000)                 */
    double at_bound_fraction; /* share of x* at a bound (0.6: 30% lower, 30% upper)      */
} mc_gen_spec;

void mc_gen_spec_default(mc_gen_spec* spec, int32_t family);
int mc_lp_generate(const mc_gen_spec* spec, mc_lp_handle** out);
int mc_lp_view(const mc_lp_handle* h, mc_lp* out);
/* Planted optimum: x*[n], y*[m] in the sign convention c = A^T y* + s. */
int mc_lp_planted(const mc_lp_handle* h, const double** x_star, const double** y_star,
                  double* planted_objective);
void mc_lp_free(mc_lp_handle* h);

/* Copy of an mc_lp into a new handle whose b (row bounds), c and variable
 * upper bounds are perturbed for the learned-warm-start family: the planted
 * pair is re-drawn around the base one (x* interior entries and upper bounds
 * scaled by U[0.8,1.2], y* and s scaled by lognormal(0, sigma)) and b, c are
 * recomputed so the instance stays feasible with an exact planted optimum. */
int mc_lp_perturb(const mc_lp_handle* base, uint64_t seed, double sigma, mc_lp_handle** out);

/* Reads a fixed- or free-format MPS file (reference io::read_mps_file,
 * include/markov_cero/io/mps.hpp and src/io/mps.cpp) into a new handle --
 * the CLI's input format, for FFI callers.  The reference parser's size caps
 * (mps.hpp:13-20) are not applied.  A handle read this way has no planted
 * pair: mc_lp_planted() returns MC_ERR_INVALID_ARGUMENT for it. */
int mc_lp_read_mps(const char* path, mc_lp_handle** out);

#ifdef __cplusplus
}
#endif
#endif /* MCPDLP_H_ */
