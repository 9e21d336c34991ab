// common.hpp -- device-side building blocks shared by the PDHG engine kernels.
//
// Layout in HBM (one packed problem holding B >= 1 independent LPs placed
// block-diagonally; a single LP is B = 1):
//   CSR(A)   rows = constraints: rp[m+1], ci[nnz] (int32, global column ids)
//   CSR(A^T) rows = variables:   cp[n+1], ri[nnz] (int32, global row ids)
//   values in FP64 (checks, FP64 iterations) and FP32 (mixed-precision iterations)
// Work is cut into Tiles: a run of consecutive rows of one matrix and one
// instance.  A wave tile is the task of ONE wave64 with 2^lg lanes per row; a
// block tile (stream.hip) is the task of one 256-thread workgroup.  The host
// orders rows by length bucket (fast order), so a tile's rows have similar
// lengths.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcp {

constexpr int kBlock = 256;        // threads per workgroup: 4 wave64s, 4 tiles
constexpr int kWaves = kBlock / 64;
constexpr int kFinal = 1024;       // threads of the per-instance finalize kernels
constexpr int kMaxSteps = 8;       // row groups one wave sweeps per tile
constexpr int kXcd = 8;            // MI355X: 8 XCDs, workgroups dealt round-robin

struct Tile {
    int32_t r0, r1;  // rows [r0, r1)
    int32_t inst;    // instance id
    int32_t shape;   // bits 0..2: lg = log2(lanes per row), 0..6; bit 3 (block
                     // tiles): one row longer than the budget, strided by the workgroup
};

// Instance status values.
constexpr int32_t kRunning = 0;
constexpr int32_t kOptimal = 1;
constexpr int32_t kIterLimit = 2;
constexpr int32_t kNumFail = 3;
constexpr int32_t kInfeasible = 4;  // a primal or dual ray certificate was found
// Arithmetic mode of an instance's SpMVs (the iterate, its updates and the
// checks are FP64 in every mode).
constexpr int32_t kModeFp64 = 0;
constexpr int32_t kModeFp32 = 1;  // FP32 values, FP32 gathered vectors, FP32 products and sums
constexpr int32_t kModePacked = 2;  // packed stream (FP32 value + 8-bit correction in the index
                                    // word): FP64 gathered vectors, products and sums

// Per-instance scalar PDHG state (device resident, one per LP).
struct InstState {
    double eta, omega;
    double last_restart_score;
    double c_scale;          // max(1, ||c_orig||_inf): the dual-residual denominator
    double primal, dual, gap, score, objective;  // last FP64 check
    double best_score;       // mixed precision: best checked score while in FP32
    int64_t restart_iter;    // global iteration count at the last restart
    int64_t done_iter;       // iteration count when the instance stopped
    int64_t fp32_iters;      // iteration count when the FP32-SpMV phase ended
    int32_t status;
    int32_t mode;            // kModeFp64 / kModeFp32 / kModePacked (the iterate and every
                             // update stay FP64 in all modes)
    int32_t restart_flag;    // set by a check, consumed by k_post_check
    int32_t switch_flag;     // set by the check that moved the instance to another SpMV mode
    int32_t mode_at_check;   // mode of the iterations before the last check
    int32_t stall;           // FP32: checks since the last new best score
    int32_t cand_cur;        // KKT scheme: this check's candidate is the current iterate
    int32_t answer_cur;      // the reported iterate is the current one (else the average)
    double prev_cand;        // KKT scheme: candidate score at the previous check of the epoch
    double ruiz_err_r, ruiz_err_c;  // setup: max |1 - norm| of the current pass
    int32_t ruiz_done, ruiz_passes, ruiz_converged, pad_;
    double b_norm_inf, c_norm_inf;  // setup: norms of the scaled b / c (floored at 1)
    double eta_cap;          // largest step factor of the adaptive rule: 0.99, or 0.95 / L
                             // (Halpern scheme, L the power-method norm of the scaled matrix)
    int64_t packed_end;      // iteration count when the packed-stream phase ended
};

// Options every iteration kernel reads (passed by value).
struct IterOpts {
    double primal_tol, dual_tol, gap_tol;
    double restart_factor;
    double smoothing;
    double polish_threshold;
    double infeas_tol;           // ray certificates: violation <= infeas_tol x |ray objective|
    int64_t restart_every;
    int64_t max_iterations;
    int32_t restart_strategy;    // 0 none, 1 adaptive, 2 fixed
    int32_t adaptive_weight;
    int32_t ruiz;
    int32_t mixed;
    int32_t kkt;                 // restart scheme: 0 reference, 1 KKT (Applegate et al.)
    int32_t stall_checks;        // mixed: checks without a new best before the FP64 switch
    int32_t halpern;             // Halpern scheme: blended primal weight at restarts
    int32_t detect;              // test the iterate's movement for infeasibility certificates
    int32_t adaptive_step;       // A dx + eta update every 10 iterations (part of the captured
                                 // kernel sequence, so a change must drop the cached graphs)
    int32_t packed_next;         // the FP32-SpMV phase hands over to the packed stream (else FP64)
    int32_t packed_floor;        // the check also scores the average with the packed values
    int32_t omega_max_drop;      // Halpern scheme: a restart lowers omega by at most this factor (0: none)
};

// Map (workgroup, wave) to a tile so that each XCD works on a contiguous
// range of tiles (neighbouring tiles touch neighbouring vector segments,
// which then share that XCD's L2).  grid = kXcd * per workgroups,
// per = ceil(ceil(ntiles / kWaves) / kXcd).
__device__ __forceinline__ int wave_tile(int ntiles) {
    const int nblk = (ntiles + kWaves - 1) / kWaves;
    const int per = (nblk + kXcd - 1) / kXcd;
    const int b = blockIdx.x;
    return ((b % kXcd) * per + b / kXcd) * kWaves + (threadIdx.x >> 6);
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

// Deterministic block reductions over NW waves: waves reduce by
// xor-shuffle, wave results are combined in wave order by thread 0.  Result
// valid in thread 0; every thread of the block must call.
template <int NW, typename T>
__device__ __forceinline__ T block_sum(T v, T* scratch /* >= NW */) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) scratch[w] = v;
    __syncthreads();
    T r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < NW; ++i) r += scratch[i];
    return r;
}
// One maximum and NS sums reduced together (one barrier pair): results valid
// in thread 0, combined in wave order (deterministic).  scratch >= NW*(NS+1).
template <int NW, int NS>
__device__ __forceinline__ void block_max_sums(double& mx, double (&s)[NS], double* scratch) {
    mx = wave_max(mx);
#pragma unroll
    for (int k = 0; k < NS; ++k) s[k] = wave_sum(s[k]);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
        scratch[w * (NS + 1)] = mx;
#pragma unroll
        for (int k = 0; k < NS; ++k) scratch[w * (NS + 1) + 1 + k] = s[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        mx = 0.0;
#pragma unroll
        for (int k = 0; k < NS; ++k) s[k] = 0.0;
        for (int i = 0; i < NW; ++i) {
            mx = maxv(mx, scratch[i * (NS + 1)]);
#pragma unroll
            for (int k = 0; k < NS; ++k) s[k] += scratch[i * (NS + 1) + 1 + k];
        }
    }
}
template <int NW, typename T>
__device__ __forceinline__ T block_max(T v, T* scratch) {
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) scratch[w] = v;
    __syncthreads();
    T r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < NW; ++i) r = maxv(r, scratch[i]);
    return r;
}

// Packed matrix stream (MC_PRECISION_PACKED): each nonzero is one uint32
// word holding the column (row) index in its low 24 bits and a signed 8-bit
// correction k in its top byte, beside the FP32 value a32 = fl32(a).  The
// value used is a32 + k ulp(a32)/256 = a32 + k 2^(e-31) (e the exponent of
// a32), exact in FP64 and within ulp(a32)/256 of the FP64 value a: about 31
// significant bits for the 8 bytes per nonzero of an FP32 stream.
constexpr uint32_t kPackIndexMask = 0xFFFFFFu;
constexpr int32_t kPackMaxIndex = 1 << 24;  // indices must stay below this
// The packed stream hands over to FP64 once the check's score is within
// kFloorMargin of the packed floor -- the largest change of a KKT component
// when the check's candidate is scored with the packed values instead of the
// FP64 ones -- or below kPackedFloor.
constexpr double kFloorMargin = 8.0;
constexpr double kPackedFloor = 1e-8;
__device__ __forceinline__ double packed_value(float a32, uint32_t w) {
    const int k = (int)(int8_t)(uint8_t)(w >> 24);
    const float p2 = __int_as_float(__float_as_int(a32) & 0x7F800000);  // 2^e; 0 for zero/subnormal
    return (double)a32 + (double)k * ((double)p2 * 0x1p-31);
}
// The correction k of an FP64 value a with a32 = fl32(a):
// round((a - a32) / 2^(e-31)) clamped to [-127, 127].
__device__ __forceinline__ int pack_correction(double a, float a32) {
    const float p2 = __int_as_float(__float_as_int(a32) & 0x7F800000);
    if (p2 == 0.0f) return 0;
    const double r = (a - (double)a32) / ((double)p2 * 0x1p-31);
    return (int)rint(r < -127.0 ? -127.0 : (r > 127.0 ? 127.0 : r));
}
// The value the packed stream holds for the FP64 value a (what the SpMVs of
// the packed mode multiply by), recomputed from a alone.
__device__ __forceinline__ double packed_of(double a) {
    const float a32 = (float)a;
    const float p2 = __int_as_float(__float_as_int(a32) & 0x7F800000);
    return (double)a32 + (double)pack_correction(a, a32) * ((double)p2 * 0x1p-31);
}
// Row weight (1 + beta)^(-k/8), k in 0..8, from the square roots of 1 + beta
// that the bits of k select, multiplied in the order 8, 4, 2, 1: every
// operation is correctly rounded, so host and device agree bit for bit.
constexpr int kMaxRowWeighting = 8;
__device__ __forceinline__ double row_weight(double beta, int k) {
    const double s0 = 1.0 + beta;
    const double s1 = sqrt(s0), s2 = sqrt(s1), s3 = sqrt(s2);
    double p = 1.0;
    if (k & 8) p *= s0;
    if (k & 4) p *= s1;
    if (k & 2) p *= s2;
    if (k & 1) p *= s3;
    return 1.0 / p;
}
// Index and value of one stream element: plain (int32 index, value as
// stored) or packed (uint32 word, FP32 value).
__device__ __forceinline__ int ix_of(int32_t i) { return i; }
__device__ __forceinline__ int ix_of(uint32_t w) { return (int)(w & kPackIndexMask); }
template <typename TS, typename TA>
__device__ __forceinline__ TS val_of(TA a, int32_t) { return (TS)a; }
template <typename TS>
__device__ __forceinline__ TS val_of(float a, uint32_t w) { return (TS)packed_value(a, w); }

// Dot product of one CSR row [p0, p1) with a gathered vector, by the G lanes
// of a row group (gl = lane within the group).  Each lane keeps four
// (index, value) loads and four gathers in flight; the group then reduces by
// xor-shuffle.  With G = 1 the lane adds the products in increasing index
// order, one rounding per product and per add -- the reference's serial
// order.  Every lane of the wave must call (the shuffles need their partners).
// RP: multiply by packed_of(value) (the packed stream's value of an FP64 entry).
template <typename TS, typename TA, typename TV, typename IX = int32_t, bool RP = false>
__device__ __forceinline__ TS row_dot(int p0, int p1, int gl, int G, const IX* __restrict__ idx,
                                      const TA* __restrict__ val, const TV* __restrict__ vec) {
    // Every round issues four index/value loads and four gathers per lane
    // before the first product is used, the tail round included: a lane
    // past its row's end re-reads its first entry of the round (always in
    // bounds) and skips the sum, so a row costs ceil(len / 4G) memory
    // round trips.  Products are added in index order (bit-exact with a
    // serial sum when G == 1).
    TS s = 0;
    for (int q = p0 + gl; q < p1; q += 4 * G) {
        const bool m1 = q + G < p1, m2 = q + 2 * G < p1, m3 = q + 3 * G < p1;
        const int q1 = m1 ? q + G : q, q2 = m2 ? q + 2 * G : q, q3 = m3 ? q + 3 * G : q;
        const IX w0 = idx[q], w1 = idx[q1], w2 = idx[q2], w3 = idx[q3];
        const TA a0 = val[q], a1 = val[q1], a2 = val[q2], a3 = val[q3];
        const TS t0 = (RP ? (TS)packed_of((double)a0) : val_of<TS>(a0, w0)) * (TS)vec[ix_of(w0)];
        const TS t1 = (RP ? (TS)packed_of((double)a1) : val_of<TS>(a1, w1)) * (TS)vec[ix_of(w1)];
        const TS t2 = (RP ? (TS)packed_of((double)a2) : val_of<TS>(a2, w2)) * (TS)vec[ix_of(w2)];
        const TS t3 = (RP ? (TS)packed_of((double)a3) : val_of<TS>(a3, w3)) * (TS)vec[ix_of(w3)];
        s += t0;
        if (m1) s += t1;
        if (m2) s += t2;
        if (m3) s += t3;
    }
    for (int off = G >> 1; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

}  // namespace mcp
