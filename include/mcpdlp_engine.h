/*
 * mcpdlp_engine.h -- thin C-ABI between the C++20 host library and the HIP
 * PDHG engine (libmcpdlp_hip.so).
 *
 * The reference has no device layer: its solve_pdlp runs the PDHG loop on the
 * host (src/lp/first_order/pdlp.cpp:139-276).  These entry points replace the
 * pieces of that function one for one:
 *   mcp_load        -- the model copy (pdlp.cpp:150) + host->device upload
 *   mcp_setup       -- equilibrate_model (pdlp.cpp:152 -> src/scale/ruiz_scaling.cpp:152-247)
 *                      and the norm/sentinel setup (pdlp.cpp:153-185)
 *   mcp_solve       -- start point + the iteration loop and exits (pdlp.cpp:186-275)
 *   mcp_fetch       -- the unscaled averaged iterate and residuals the result carries
 *                      (pdlp.cpp:236-244, 267-274)
 *   mcp_spmv        -- spmv / spmv_t (pdlp.cpp:22-44) on the loaded, scaled matrix
 * A "problem" is B >= 1 independent LPs packed block-diagonally; B == 1 is
 * the reference's single solve, B > 1 the batched extension.  Plain pointers
 * and sizes only; every call returns 0 or an MC_ERR_* code (mcpdlp.h) and
 * never throws; mcp_last_error() holds the message (thread-local).
 * An engine owns one HIP stream and its device buffers; use one engine per
 * host thread (calls on one engine are not synchronized).
 */
#ifndef MCPDLP_ENGINE_H_
#define MCPDLP_ENGINE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mcp_engine mcp_engine;

/* Number of visible HIP devices (0 when none; never an error). */
int mcp_device_count(int32_t* count);
int mcp_engine_create(int32_t device, mcp_engine** out);
void mcp_engine_destroy(mcp_engine* e);
const char* mcp_last_error(void);

/* Packed problem in host memory (int32 indices; nnz < 2^31).  The matrix is
 * given in both orientations: CSR of A (rows) and CSR of A^T (= CSC of A),
 * entries sorted by index inside each row/column.  Bounds use IEEE +-inf for
 * infinite kinds.  `objective` is the model's (unsigned) objective; `sense`
 * holds +1 (minimize) / -1 (maximize) per instance. */
typedef struct mcp_problem {
    int32_t instances;
    int32_t rows, cols;
    int64_t nnz;
    const int32_t* inst_row;   /* [B+1] */
    const int32_t* inst_col;   /* [B+1] */
    const int32_t* row_ptr;    /* [m+1] CSR(A) */
    const int32_t* col_idx;    /* [nnz] */
    const double* row_val;     /* [nnz] */
    const int32_t* col_ptr;    /* [n+1] CSR(A^T) */
    const int32_t* row_idx;    /* [nnz] */
    const double* col_val;     /* [nnz] */
    const double* row_lower;   /* [m] */
    const double* row_upper;   /* [m] */
    const double* var_lower;   /* [n] */
    const double* var_upper;   /* [n] */
    const double* objective;   /* [n] */
    const double* sense;       /* [B] */
    const double* offset;      /* [B] */
} mcp_problem;

/* Tile shape policy of the wave tasks (autotuner knobs): the kernels of the
 * reference reduction order, Ruiz and setup, the ray certificates, and the
 * fused fast-order steps when block tiles are switched off. */
typedef struct mcp_tile_policy {
    int32_t nnz_budget;     /* nonzeros per wave task (tile)                    */
    int32_t per_lane;       /* target nonzeros per lane when picking lanes/row  */
} mcp_tile_policy;

/* Upload + tiling.  Reuses device buffers when sizes allow.  The default tile
 * policy depends on the size: {128 nnz, 2 per lane} (short wave tasks, many
 * waves) up to 2M nonzeros, {1024, 8} above; mcp_setup's autotuner or
 * mcp_set_tile_policy may replace it. */
int mcp_load(mcp_engine* e, const mcp_problem* p);

typedef struct mcp_setup_opts {
    int32_t ruiz;              /* PdlpOptions::ruiz_scaling */
    int32_t ruiz_iterations;   /* PdlpOptions::ruiz_iterations */
    int32_t exact;             /* MC_ORDER_REFERENCE tiling (1 lane per row) */
    int32_t autotune;          /* time candidate tile policies on the device */
    int32_t row_weighting;     /* mc_pdlp_options::row_weighting (0..8) */
    int32_t pad_;
} mcp_setup_opts;
int mcp_setup(mcp_engine* e, const mcp_setup_opts* o);

typedef struct mcp_solve_opts {
    uint64_t max_iterations;
    uint64_t restart_every;
    double primal_tolerance, dual_tolerance, gap_tolerance;
    double step_size_reduction;
    int32_t restart_strategy;
    int32_t adaptive_step_size;
    double restart_reduction_factor;
    int32_t adaptive_primal_weight;
    int32_t precision;          /* MC_PRECISION_* */
    double initial_primal_weight;
    double primal_weight_smoothing;
    double polish_threshold;
    int32_t reduction_order;    /* MC_ORDER_* */
    int32_t graph_chunk;        /* iterations per HIP graph; 0 => default; <0 => no graphs */
    int32_t restart_scheme;     /* MC_SCHEME_* */
    int32_t fp32_stall_checks;  /* mixed: checks without a new best score before the FP64 switch; 0 => 50 */
    const double* warm_primal;  /* optional [n] host, original space; a warm start (either vector)
                                   begins in FP64 SpMVs even with MC_PRECISION_MIXED */
    const double* warm_dual;    /* optional [m] host, original space (reported convention) */
    int32_t detect_infeasibility;     /* 1: stop on a ray certificate (status infeasible_or_unbounded) */
    int32_t pad2_;
    double infeasibility_tolerance;   /* certificate tolerance; 0 => 1e-8 */
    uint64_t power_iterations;        /* Halpern scheme with adaptive steps: power iterations that
                                         estimate L = ||D_r^-1/2 A D_c^-1/2|| per instance, and the
                                         step factor eta is capped at 0.95 / L instead of 0.99
                                         (0 => 0.99; reference option pdlp.hpp:25, unread there) */
    int32_t primal_weight_max_drop;   /* mc_pdlp_options::primal_weight_max_drop */
    int32_t pad3_;
} mcp_solve_opts;

typedef struct mcp_solve_stats {
    double kernel_ms;           /* device time of the iteration loop */
    double setup_ms;            /* device time of mcp_setup           */
    double h2d_ms;              /* time of the last mcp_load upload   */
    uint64_t iterations;        /* global iterations executed          */
    int32_t graphs_built;
    int32_t chunks;
} mcp_solve_stats;
/* Runs the iteration loop in captured 64-iteration HIP graphs (reused while the
 * problem view and the options are unchanged).  In a batch, once half of the
 * instances running at the last tiling have finished, the tiles are rebuilt
 * over the running instances only (the latency policy once they hold <= 2M
 * nonzeros); the full tiling is restored before returning. */
int mcp_solve(mcp_engine* e, const mcp_solve_opts* o, mcp_solve_stats* stats);

/* Per-instance outcome. */
typedef struct mcp_inst_result {
    int32_t status;             /* MC_PDLP_* */
    int32_t mode;               /* SpMV precision at the end: 0 FP64, 1 FP32, 2 packed */
    uint64_t iterations;
    uint64_t fp32_iterations;   /* iterations of the FP32-SpMV phase */
    double objective;
    double primal_infeasibility, dual_infeasibility, duality_gap;
    double omega, eta;
    uint64_t packed_iterations; /* iterations of the packed-stream phase */
} mcp_inst_result;

/* Copy the unscaled averaged iterate (x[n], y[m]; either may be NULL) and the
 * per-instance results ([B]; may be NULL) to the host.  *d2h_ms = copy time. */
int mcp_fetch(mcp_engine* e, double* x, double* y, mcp_inst_result* res, double* d2h_ms);

/* SpMV on the loaded problem's (scaled, after mcp_setup) matrix: transpose=0
 * computes out[m] = A v[n], transpose=1 out[n] = A^T v[m].  precision selects
 * the FP64 or FP32 value copy (vectors are converted at the boundary);
 * order=MC_ORDER_REFERENCE sums each row serially.  reps > 1 repeats the
 * launch; *avg_us receives the mean device time per launch. */
int mcp_spmv(mcp_engine* e, int32_t transpose, int32_t precision, int32_t order,
             const double* v, double* out, int32_t reps, double* avg_us);

/* Scaled matrix values/scales after mcp_setup (for parity tests of the
 * device Ruiz): CSC(A)-ordered values [nnz] and row/col scales. */
int mcp_get_scaling(mcp_engine* e, double* col_val, double* row_scale, double* col_scale,
                    int32_t* passes, int32_t* converged);

/* Micro-benchmark of one fused iteration kernel on the loaded, set-up problem,
 * launched `reps` times on the engine's stream.  The fused steps (kinds 0-3,
 * 10, 11) and the gather probes (7-9) are timed as the solve runs them: each
 * timed launch follows a launch of the other step (untimed), which streams
 * the other matrix through the caches; all launches are queued before one
 * wait, and each timed launch records a HIP event pair as part of its own
 * dispatch (hipExtLaunchKernelGGL), so the mean is the kernel's own duration.
 * The other kinds are timed back to back between two events.
 * kind: 0 primal (A^T y + primal update) FP32, 1 dual (A x_bar + dual update)
 * FP32, 2 primal FP64, 3 dual FP64, 4 plain SpMV A x FP32, 5 plain A^T y FP32,
 * 6 plain streaming copy of the FP32 values of A (the calibration kernel for
 * HBM byte counters: it moves exactly 8 bytes per copied value), 7-9 the FP64
 * block SpMV A x_bar with random gathers, with every gather inside a window of
 * at most 256 entries, and with no gathers, 10 primal and 11 dual on the
 * packed streams (MC_PRECISION_PACKED).
 * *avg_us = mean device time per launch; *bytes = the kernel's algorithmic
 * HBM bytes per launch (matrix and pointer streams once, gathered vector once,
 * per-row/column operands once).  Leaves the iterate in an unspecified state;
 * the next mcp_solve re-initializes it. */
int mcp_time_kernel(mcp_engine* e, int32_t kind, int32_t reps, double* avg_us, double* bytes);

/* Synchronization probe (extension; no reference counterpart -- a
 * measurement).  *barrier_us: mean cost of one grid-wide barrier inside a
 * cooperative persistent kernel of `workgroups` 256-thread workgroups (all
 * resident; more than can be is refused), each barrier an agent-scope release,
 * one atomic arrival and an acquire-spin on a generation word, over `rounds`
 * rounds of two barriers that also check that every workgroup sees its
 * neighbour's writes.  *launch_us: mean cost of one dependent kernel launch of
 * the same workgroups in a captured HIP graph.  A persistent batch kernel that
 * replaced the iteration's launches by barriers would pay barrier_us where the
 * graph pays launch_us.  Every wait is bounded: a barrier that runs out of
 * spins is reported as an error, never left hanging. */
int mcp_time_sync(mcp_engine* e, int32_t workgroups, int32_t rounds, double* barrier_us, double* launch_us);

/* Pin the tile policies of the loaded problem (rows of A, rows of A^T) and
 * re-tile; the autotuner's choice is overwritten.  For tuning experiments and
 * for callers that know their matrix. */
int mcp_set_tile_policy(mcp_engine* e, const mcp_tile_policy* rows_policy, const mcp_tile_policy* cols_policy);

/* Block tiles of the fused fast-order primal/dual steps (one 256-thread
 * workgroup streams a contiguous run of rows holding at most `budget`
 * nonzeros; the products are summed per row from LDS).  Budgets are 256,
 * 512, 1024 or 2048 nonzeros per tile; 0 for both runs those steps on the wave
 * tiles of mcp_set_tile_policy instead.  Default: 256 up to 128k nonzeros,
 * 512 up to 2M, 1024 above (a batch re-tiled as its instances finish takes
 * the budget of the nonzeros still running); the autotuner may replace it.  Reference order always uses the
 * wave tiles.  No counterpart in the reference (its SpMV is a serial loop,
 * src/lp/first_order/pdlp.cpp:22-44): a tuning knob of this engine. */
int mcp_set_block_budget(mcp_engine* e, int32_t rows_budget, int32_t cols_budget);
int mcp_get_block_tiling(mcp_engine* e, int32_t* rows_budget, int32_t* cols_budget, int32_t* row_tiles,
                         int32_t* col_tiles);

/* Current tile policies (after mcp_setup / autotune) and tile counts. */
int mcp_get_tiling(mcp_engine* e, mcp_tile_policy* rows_policy, mcp_tile_policy* cols_policy,
                   int32_t* row_tiles, int32_t* col_tiles);

#ifdef __cplusplus
}
#endif
#endif /* MCPDLP_ENGINE_H_ */
