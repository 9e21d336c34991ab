// dev.hpp -- device pointers of one packed problem + iterate state, passed
// by value to every kernel, and the host-side launch helpers of kernels.hip.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace mcp {

// A kernel launch that, given events e0/e1, records them as part of the
// kernel's dispatch packet (device timestamps of the kernel alone).
#define MCP_LAUNCH_TIMED(K, g, b, s, e0, e1, ...)                        \
    do {                                                                  \
        if (e0)                                                           \
            hipExtLaunchKernelGGL(K, g, b, 0, s, e0, e1, 0, __VA_ARGS__); \
        else                                                              \
            hipLaunchKernelGGL(K, g, b, 0, s, __VA_ARGS__);               \
    } while (0)

struct Dev {
    int32_t B;      // instances
    int32_t m, n;   // packed rows / columns
    int32_t ntA, ntT;
    // structure
    const int32_t* rp;  // CSR(A) row pointers [m+1]
    const int32_t* ci;  // CSR(A) column ids [nnz]
    const int32_t* cp;  // CSR(A^T) pointers [n+1]
    const int32_t* ri;  // CSR(A^T) row ids [nnz]
    double* av64;       // CSR(A) values (scaled in place by the setup)
    double* tv64;       // CSR(A^T) values
    float* av32;
    float* tv32;
    // packed streams of A and A^T (kModePacked): index | correction << 24 per
    // nonzero, beside av32 / tv32; null when an index does not fit 24 bits
    uint32_t* cix;
    uint32_t* rix;
    const int32_t* inst_row;   // [B+1] packed row offsets
    const int32_t* inst_col;   // [B+1]
    const int32_t* row_inst;   // [m] instance of each row
    const int32_t* col_inst;   // [n]
    const Tile* tA;            // tiles over rows of A
    const Tile* tT;            // tiles over rows of A^T
    const int32_t* tA_begin;   // [B+1] tile range of each instance in tA
    const int32_t* tT_begin;   // [B+1]
    // original model (residuals are evaluated on it); infinite bounds = IEEE +-inf
    double* orow_lo;
    double* orow_hi;
    double* ovar_lo;
    double* ovar_hi;
    const double* obj;     // unsigned original objective [n]
    const double* sense;   // [B] +1 minimize / -1 maximize
    const double* offset;  // [B] objective offset
    // scaling
    double* row_scale;  // [m]
    double* col_scale;  // [n]
    double* rtmp;       // [m] setup scratch (Ruiz row max / delta)
    double* ctmp;       // [n]
    // scaled iteration data
    double* blo;
    double* bhi;
    double* rnorm;   // [m] row L1 norm of scaled A (floored)
    double* c;
    double* slo;
    double* shi;
    double* cnorm;   // [n]
    // iterates (FP64 in both precision modes)
    double *x, *xbar, *xavg, *dx;
    double *y, *yavg;
    double *xlast, *ylast;      // KKT/Halpern schemes: iterate at the last restart
    // FP32 copies that FP32-mode SpMVs gather (written by the step that
    // produces the FP64 value)
    float *xbar32, *dx32, *y32;
    double* ax;    // [m] reference-order mode: A*v results (average candidate)
    double* aty;   // [n] reference-order mode: A^T*v results
    double* ax2;   // [m] reference-order mode, current-iterate candidate
    double* aty2;  // [n]
    // reduction partials (one slot group per tile)
    double* part_dxn;   // [ntT], or [nbT] when the primal step runs on block tiles
    double* part_adx;   // [ntA], or [nbA] when A dx runs on block tiles
    // check partials per candidate c (0 average, 1 current iterate):
    // (tiles = ntA / ntT, or nbA / nbT when the check runs on block tiles)
    double* part_rows;  // [(c*3 + k)*tiles + tile]: max violation, dual-objective part, |y - ylast|^2
    double* part_cols;  // [(c*5 + k)*tiles + tile]: max violation, dual residual^2, primal obj,
                        //                            dual-objective part, |x - xlast|^2
    // infeasibility detection: the candidate's movement since the last
    // restart and its per-tile certificate partials
    double* ray_x;      // [n]
    double* ray_y;      // [m]
    double* part_ray_rows;  // [tile*5 + k]: primal-ray row violation, dual-ray objective part and
                            //   its magnitude, dual-ray row violation, |dy|_inf
    double* part_ray_cols;  // [tile*7 + k]: primal-ray bound violation, c^T dx and its magnitude,
                            //   dual-ray objective part and its magnitude, dual-ray column
                            //   violation, |dx|_inf
    double* pw;         // [2B] step-cap power method: S_b = |v_b|^2, and 1 once S_b stopped being > 0
    InstState* st;      // [B]
    int64_t* iter_base; // global iteration counter at the start of the current chunk
    int32_t* counters;  // [0] running instances, [1] running instances in FP32 or packed mode
    // block tiles of the fused fast-order primal/dual kernels (stream.hip):
    // one 256-thread workgroup per tile
    const Tile* bA;
    const Tile* bT;
    const int32_t* bA_begin;   // [B+1] tile range of each instance in bA
    const int32_t* bT_begin;   // [B+1]
    int32_t nbA, nbT;
    int32_t budA, budT;        // nonzeros per block tile: 512, 1024 or 2048
    int32_t halpern;    // reflected Halpern iterates: the average slots hold T(z), the
                        // last-restart slots the anchor z0
    int32_t nt_rows;    // 1: the FP64 streams of A (dual step, A dx, the check's rows) use
                        // non-temporal loads, so that A^T stays in the 256 MB last-level
                        // cache when the pair of FP64 matrices does not fit but one does
    int32_t blocks;     // 1: the fast-order primal/dual steps, A dx and the check's
                        // residual terms run on the block tiles (part_dxn, part_adx,
                        // part_rows, part_cols then hold slots per tile of bT / bA);
                        // 0: on tA / tT
};

// ---- launch helpers (kernels.hip) ----
// Iteration kernels (each instance runs in its own FP32/FP64 mode).  `t` is
// the iteration's offset inside the chunk; exact = reference reduction order.
// With e0/e1 (timing only, never inside a captured graph) the launch records
// the two events as part of the kernel's own dispatch (hipExtLaunchKernelGGL).
void launch_primal(const Dev& d, int t, bool adapt, bool exact, hipStream_t s, hipEvent_t e0 = nullptr,
                   hipEvent_t e1 = nullptr);
void launch_dual(const Dev& d, int t, bool exact, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void launch_adx(const Dev& d, int t, bool exact, hipStream_t s);
void launch_adapt_finalize(const Dev& d, bool exact, hipStream_t s);
void launch_resid(const Dev& d, bool exact, bool current, hipStream_t s);
// The check's residual terms: the average and, with `both`, the current
// iterate (block tiles in one pass when the engine uses them); with
// `packed_floor` also the average scored with the packed stream's values.
void launch_check_resid(const Dev& d, bool exact, bool both, bool packed_floor, hipStream_t s);
void launch_check_finalize(const Dev& d, const IterOpts& o, int t, bool exact, bool final_pass,
                           hipStream_t s);
void launch_post_check(const Dev& d, bool kkt, hipStream_t s);
// Infeasibility certificates from the candidate's movement since the last
// restart (after the check, before the restart copy).
void launch_ray_check(const Dev& d, const IterOpts& o, int t, hipStream_t s);
void launch_advance(const Dev& d, int len, hipStream_t s);
// Halpern-scheme step cap: `iters` power iterations per instance on M^T M,
// M = D_r^-1/2 A D_c^-1/2 (scaled A, D the floored L1 norms), then
// st[b].eta_cap = 0.95 / L_b (0.99 if the estimate fails); exact => every sum
// in the serial order of the oracle.  Uses ctmp, rtmp and aty2 as scratch.
void launch_step_cap(const Dev& d, int64_t iters, bool exact, hipStream_t s);
// Setup kernels (run on the current tiles).
void launch_ruiz_pass(const Dev& d, double tol, hipStream_t s);
// rows scaled by (1 + beta_i)^(-k8/8) after the Ruiz passes (mcp_setup_opts::row_weighting)
void launch_row_weights(const Dev& d, int k8, hipStream_t s);
void launch_setup_data(const Dev& d, hipStream_t s);
void launch_init_iterates(const Dev& d, const double* warm_x, const double* warm_y, hipStream_t s);
// start_mode: the SpMV mode every instance starts in (kModeFp64 / kModeFp32 / kModePacked).
void launch_setup_state(const Dev& d, double initial_weight, double step_reduction, int start_mode,
                        hipStream_t s);
void launch_fill(double* p, double v, int64_t n, hipStream_t s);
// Synchronization probe (mcp_time_sync): the cooperative barrier kernel, the
// number of its workgroups that can be resident at once, and one round of
// the barrier-free form.
int barrier_probe_capacity();
hipError_t launch_barrier_probe(unsigned* bar, double* buf, int words, int rounds, int workgroups, hipStream_t s);
void launch_barrier_step(double* buf, int words, int round, int workgroups, hipStream_t s);
// Plain streaming copy of n floats (n a multiple of 4, 16-byte aligned
// buffers): the reference kernel that calibrates the HBM byte counters.
void launch_copy(const float* src, float* dst, int64_t n, hipStream_t s);
void launch_unscale(const Dev& d, double* x_out, double* y_out, hipStream_t s);
// Plain SpMV over a tile set: out[r] = sum val * vec (tests and micro-benchmarks);
// tiles with one lane per row sum in the reference's serial order.
void launch_spmv_plain(const Tile* tiles, int ntiles, const int32_t* ptr, const int32_t* idx,
                       const void* val, const void* vec, void* out, bool fp32, hipStream_t s);
// Block-tile forms (stream.hip): the fused fast-order primal/dual steps and a
// plain SpMV over block tiles of `budget` (512, 1024 or 2048) nonzeros.
void launch_primal_blk(const Dev& d, int t, bool adapt, hipStream_t s, hipEvent_t e0 = nullptr,
                       hipEvent_t e1 = nullptr);
void launch_dual_blk(const Dev& d, int t, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void launch_adx_blk(const Dev& d, int t, hipStream_t s);
// Residual terms of the check's candidates over block tiles in one pass:
// the average only, or (both) the average and the current iterate.
void launch_resid_blk(const Dev& d, bool both, bool packed_floor, hipStream_t s);
// Packed-stream SpMV (index words with the value correction, FP32 values,
// FP64 vector, products and sums), block and wave forms.
void launch_spmv_blk_packed(const Tile* tiles, int ntiles, int budget, const int32_t* ptr, const uint32_t* idxp,
                            const float* val32, const double* vec, double* out, hipStream_t s);
void launch_spmv_plain_packed(const Tile* tiles, int ntiles, const int32_t* ptr, const uint32_t* idxp,
                              const float* val32, const double* vec, double* out, hipStream_t s);
void launch_spmv_blk(const Tile* tiles, int ntiles, int budget, const int32_t* ptr, const int32_t* idx,
                     const void* val, const void* vec, void* out, bool fp32, hipStream_t s);
// Gather probe (FP64 block SpMV, stream.hip): mode 0 gathers vec[idx], mode 1
// vec[idx & mask] (mask + 1 entries, all cache hits), mode 2 no gather.
void launch_gather_probe(const Tile* tiles, int ntiles, int budget, const int32_t* ptr, const int32_t* idx,
                         const double* val, const double* vec, double* out, int mode, int mask, hipStream_t s,
                         hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);

}  // namespace mcp
