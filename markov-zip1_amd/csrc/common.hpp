// common.hpp -- device-side building blocks shared by the PDHG engine kernels.
//
// Layout in HBM (one packed problem holding B >= 1 independent LPs placed
// block-diagonally; a single LP is B = 1):
//   CSR(A)   rows = constraints: rp[m+1], ci[nnz] (int32, global column ids)
//   CSR(A^T) rows = variables:   cp[n+1], ri[nnz] (int32, global row ids)
//   values in FP64 (checks, FP64 iterations) and FP32 (mixed-precision iterations)
// Work is cut into Tiles: runs of consecutive rows of one matrix that belong to
// one instance and whose nonzeros fit the LDS staging buffer (or one long row).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcp {

constexpr int kBlock = 256;        // threads per workgroup: 4 wave64s
constexpr int kWaves = kBlock / 64;
constexpr int kNnzTile = 2048;     // LDS staging capacity (products) per tile
constexpr int kXcd = 8;            // MI355X: 8 XCDs, workgroups dealt round-robin

// Tile.shape: bits 0..2 = log2(lanes per row in the reduce phase, 1..64);
// bit 8 = long-row tile (a single row with more than kNnzTile nonzeros).
constexpr int32_t kShapeLong = 1 << 8;

struct Tile {
    int32_t r0, r1;  // rows [r0, r1)
    int32_t inst;    // instance id
    int32_t shape;
};

// Instance status values.
constexpr int32_t kRunning = 0;
constexpr int32_t kOptimal = 1;
constexpr int32_t kIterLimit = 2;
constexpr int32_t kNumFail = 3;
// Arithmetic mode of an instance's iterations.
constexpr int32_t kModeFp64 = 0;
constexpr int32_t kModeFp32 = 1;

// Per-instance scalar PDHG state (device resident, one per LP).
struct InstState {
    double eta, omega;
    double last_restart_score;
    double c_scale;          // max(1, ||c_orig||_inf): the dual-residual denominator
    double primal, dual, gap, score, objective;  // last FP64 check
    double best_score;       // mixed precision: best checked score while in FP32
    int64_t restart_iter;    // global iteration count at the last restart
    int64_t done_iter;       // iteration count when the instance stopped
    int64_t fp32_iters;      // iterations spent in FP32 before switching to FP64
    int32_t status;
    int32_t mode;            // kModeFp64 / kModeFp32
    int32_t restart_flag;    // set by a check, consumed by k_post_check
    int32_t switch_flag;     // set by a check: convert FP32 iterate to FP64
    int32_t mode_at_check;   // mode in which the last check's iterate lives
    int32_t stall;           // checks without a 10% score improvement (FP32)
    int32_t cand_cur;        // KKT scheme: this check's candidate is the current iterate
    int32_t answer_cur;      // the reported iterate is the current one (else the average)
    double prev_cand;        // KKT scheme: candidate score at the previous check of the epoch
    double ruiz_err_r, ruiz_err_c;  // setup: max |1 - norm| of the current pass
    int32_t ruiz_done, ruiz_passes, ruiz_converged, pad_;
    double b_norm_inf, c_norm_inf;  // setup: norms of the scaled b / c (floored at 1)
};

// Options every iteration kernel reads (passed by value).
struct IterOpts {
    double primal_tol, dual_tol, gap_tol;
    double restart_factor;
    double smoothing;
    double polish_threshold;
    int64_t restart_every;
    int64_t max_iterations;
    int32_t restart_strategy;    // 0 none, 1 adaptive, 2 fixed
    int32_t adaptive_weight;
    int32_t ruiz;
    int32_t mixed;
    int32_t kkt;                 // restart scheme: 0 reference, 1 KKT (Applegate et al.)
    int32_t pad_;
};

// Map a workgroup id to a tile so that each XCD works on a contiguous range
// of tiles (neighbouring tiles touch neighbouring vector segments, which then
// share that XCD's L2).  grid = kXcd * per, per = ceil(ntiles / kXcd).
__device__ __forceinline__ int xcd_tile(int ntiles) {
    const int per = (ntiles + kXcd - 1) / kXcd;
    const int b = blockIdx.x;
    return (b % kXcd) * per + b / kXcd;
}

template <typename T>
__device__ __forceinline__ T clampv(T v, T lo, T hi) {  // std::clamp order of tests
    return (v < lo) ? lo : ((hi < v) ? hi : v);
}
template <typename T>
__device__ __forceinline__ T maxv(T a, T b) { return (a < b) ? b : a; }  // std::max
template <typename T>
__device__ __forceinline__ T minv(T a, T b) { return (b < a) ? b : a; }  // std::min

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = maxv(v, __shfl_xor(v, off, 64));
    return v;
}

// Deterministic block reductions: waves reduce by xor-shuffle, wave results
// are combined in wave order by thread 0.  Result valid in thread 0.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T* scratch /* >= kWaves */) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) scratch[w] = v;
    __syncthreads();
    T r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kWaves; ++i) r += scratch[i];
    return r;
}
template <typename T>
__device__ __forceinline__ T block_max(T v, T* scratch) {
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) scratch[w] = v;
    __syncthreads();
    T r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kWaves; ++i) r = maxv(r, scratch[i]);
    return r;
}

// Row sums of one tile of a CSR matrix:
//   rsum[r - tl.r0] = sum_{p in row r} val[p] * vec[idx[p]]
// Regular tiles stream the tile's nonzeros once, fully coalesced, into LDS
// (phase 1) and then sum each row from LDS with 2^shape lanes per row
// (phase 2).  With 1 lane per row the sum runs in increasing column order,
// one rounding per product and per add -- the reference's serial order.
// Long tiles (one row) stride the row with the whole block and reduce
// (EXACT: thread 0 sums serially instead).  Returns the tile's row count;
// rsum is valid for all threads after the call.
template <bool EXACT, typename TS, typename TA, typename TV>
__device__ __forceinline__ int tile_rowsums(const Tile& tl, const int32_t* __restrict__ rp,
                                           const int32_t* __restrict__ idx,
                                           const TA* __restrict__ val,
                                           const TV* __restrict__ vec, TS* prod, TS* rsum,
                                           TS* red) {
    const int tid = threadIdx.x;
    const int p0 = rp[tl.r0];
    const int p1 = rp[tl.r1];
    if (tl.shape & kShapeLong) {
        TS s = 0;
        if (EXACT) {
            if (tid == 0)
                for (int p = p0; p < p1; ++p) {
                    const TS pr = (TS)val[p] * (TS)vec[idx[p]];
                    s += pr;
                }
            __syncthreads();
            if (tid == 0) rsum[0] = s;
        } else {
            for (int p = p0 + tid; p < p1; p += kBlock) {
                const TS pr = (TS)val[p] * (TS)vec[idx[p]];
                s += pr;
            }
            s = block_sum(s, red);
            if (tid == 0) rsum[0] = s;
        }
        __syncthreads();
        return 1;
    }
    for (int p = p0 + tid; p < p1; p += kBlock) prod[p - p0] = (TS)val[p] * (TS)vec[idx[p]];
    __syncthreads();
    const int lg = tl.shape & 7;
    const int G = 1 << lg;
    const int lane = tid & (G - 1);
    const int row = tl.r0 + (tid >> lg);
    TS s = 0;
    if (row < tl.r1) {
        const int q1 = rp[row + 1] - p0;
        for (int q = rp[row] - p0 + lane; q < q1; q += G) s += prod[q];
    }
    for (int off = G >> 1; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0 && row < tl.r1) rsum[tid >> lg] = s;
    __syncthreads();
    return tl.r1 - tl.r0;
}

}  // namespace mcp
