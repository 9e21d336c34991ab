// stream.hip -- block-tile forms of the fast-order iteration kernels (the
// fused primal and dual steps, A dx of the adaptive step size and the FP64
// residual terms of the check).
//
// A block tile is the task of one 256-thread workgroup: a run of consecutive
// rows of one instance holding at most NB nonzeros (512, 1024 or 2048).  The
// workgroup streams the tile's nonzeros as one contiguous range -- thread k
// takes entries k, k + 256, ... -- so every index and value load is
// coalesced, and all of a thread's loads and gathers are in flight together
// (two dependent memory round trips per tile, whatever the row lengths).  The
// products go to LDS; then G = 2^lg threads per row sum that row's products
// (lane gl takes every G-th, the group reduces by xor-shuffle) and the
// group's first lane applies the row's update.  The update's own operands are
// loaded before the products, so their latency hides under the gathers.  A
// row longer than NB is a tile of its own that the whole workgroup strides.
//
// The random gathers of x_bar / y are what bounds these kernels on sparse
// matrices without locality: each one is a separate L2 request.  The wave
// tasks of kernels.hip spend several dependent round trips per row group on
// top of that; this form does not.  Reference reduction order keeps the wave
// tasks (one lane per row sums in index order).
//
// The updates are those of k_primal / k_dual in kernels.hip (reference
// src/lp/first_order/pdlp.cpp:195-206, 220-227, and the Halpern form).
#include "dev.hpp"

namespace mcp {

constexpr int kBT = kBlock;  // threads per block tile

// Workgroup -> tile so that each XCD (workgroups are dealt round-robin to the
// 8 XCDs) takes a contiguous range of tiles; grid = kXcd * ceil(nt / kXcd).
__device__ __forceinline__ int block_tile_id(int ntiles) {
    const int per = (ntiles + kXcd - 1) / kXcd;
    const int b = blockIdx.x;
    return (b % kXcd) * per + b / kXcd;
}

// One index/value pair of the matrix stream; nt: non-temporal loads, which
// leave the last-level cache (the 256 MB Infinity Cache) to the other
// matrix (see Dev::nt_rows).
// NT is a template parameter: the compiler merges a run-time choice between
// the two load forms into one plain load.
template <bool NT, typename TA, typename IX>
__device__ __forceinline__ void stream_pair(const IX* __restrict__ idx, const TA* __restrict__ val, int p, IX& i,
                                            TA& a) {
    if (NT) {
        i = __builtin_nontemporal_load(idx + p);
        a = __builtin_nontemporal_load(val + p);
    } else {
        i = idx[p];
        a = val[p];
    }
}

// Products val[p] * vec[idx[p]] of the nonzeros [p0, p0 + len) into prod[0, len),
// len <= K * kBT.  All K index/value loads are issued, then all K gathers.
// IX = uint32_t: the packed stream (index word with the value correction).
template <int K, typename TS, typename TA, typename TV, bool NT = false, typename IX>
__device__ __forceinline__ void stage_products(int p0, int len, const IX* __restrict__ idx,
                                               const TA* __restrict__ val, const TV* __restrict__ vec,
                                               TS* prod) {
    IX ii[K];
    TA aa[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int q = (int)threadIdx.x + k * kBT;
        ii[k] = 0;
        aa[k] = (TA)0;
        if (q < len) stream_pair<NT>(idx, val, p0 + q, ii[k], aa[k]);
    }
    TV vv[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int q = (int)threadIdx.x + k * kBT;
        vv[k] = q < len ? vec[ix_of(ii[k])] : (TV)0;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int q = (int)threadIdx.x + k * kBT;
        if (q < len) prod[q] = val_of<TS>(aa[k], ii[k]) * (TS)vv[k];
    }
}

// Sum of one row's staged products [q0, q1) by its G lanes.  Groups are
// aligned runs of G <= 64 threads, so a group lies inside one wave.  Every
// thread must call (the shuffles need their partners); a thread with no row
// passes q0 == q1.
template <typename TS>
__device__ __forceinline__ TS group_sum(const TS* prod, int q0, int q1, int gl, int G) {
    TS s = 0;
    for (int q = q0 + gl; q < q1; q += G) s += prod[q];
    for (int off = G >> 1; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

// Dot product of one long row [p0, p1) by the whole workgroup (four
// accumulators per thread, fixed-order block reduction); valid in thread 0.
// RP: multiply by packed_of(value) (the packed stream's value of an FP64 entry).
template <typename TS, bool NT = false, bool RP = false, typename TA, typename TV, typename IX>
__device__ __forceinline__ double long_row_dot(int p0, int p1, const IX* __restrict__ idx,
                                               const TA* __restrict__ val, const TV* __restrict__ vec,
                                               double* red) {
    TS s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int p = p0 + (int)threadIdx.x;
    for (; p + 3 * kBT < p1; p += 4 * kBT) {
        IX i0, i1, i2, i3;
        TA a0, a1, a2, a3;
        stream_pair<NT>(idx, val, p, i0, a0);
        stream_pair<NT>(idx, val, p + kBT, i1, a1);
        stream_pair<NT>(idx, val, p + 2 * kBT, i2, a2);
        stream_pair<NT>(idx, val, p + 3 * kBT, i3, a3);
        s0 += (RP ? (TS)packed_of((double)a0) : val_of<TS>(a0, i0)) * (TS)vec[ix_of(i0)];
        s1 += (RP ? (TS)packed_of((double)a1) : val_of<TS>(a1, i1)) * (TS)vec[ix_of(i1)];
        s2 += (RP ? (TS)packed_of((double)a2) : val_of<TS>(a2, i2)) * (TS)vec[ix_of(i2)];
        s3 += (RP ? (TS)packed_of((double)a3) : val_of<TS>(a3, i3)) * (TS)vec[ix_of(i3)];
    }
    for (; p < p1; p += kBT) {
        IX i;
        TA a;
        stream_pair<NT>(idx, val, p, i, a);
        s0 += (RP ? (TS)packed_of((double)a) : val_of<TS>(a, i)) * (TS)vec[ix_of(i)];
    }
    return block_sum<kBT / 64>((double)((s0 + s1) + (s2 + s3)), red);
}

// Row sum of a block tile in the instance's mode: FP32 values and gathers with
// FP32 products and sums (kModeFp32), the packed stream (index words with
// the value correction, FP32 values) with FP64 gathers, products and sums
// (kModePacked), or FP64.  Every thread must call.
// NT64: the FP64 matrix stream uses non-temporal loads (FP32 and packed
// streams never do: the pair of 8-byte-per-nonzero matrices is small enough
// to share the cache, and on the 200k x 500k LP a non-temporal packed row
// stream made the dual step 83 instead of 63 us).
template <int NB, bool NT64 = false, typename TA32, typename TV32>
__device__ __forceinline__ double tile_row_sum(int mode, int p0, int len, int q0, int q1, int gl, int G,
                                               const int32_t* __restrict__ idx, const uint32_t* __restrict__ idxp,
                                               const double* __restrict__ val64, const double* __restrict__ vec64,
                                               const TA32* __restrict__ val32, const TV32* __restrict__ vec32,
                                               double* prod) {
    if (mode == kModeFp32) {
        float* pr = reinterpret_cast<float*>(prod);
        stage_products<NB / kBT, float>(p0, len, idx, val32, vec32, pr);
        __syncthreads();
        return (double)group_sum<float>(pr, q0, q1, gl, G);
    }
    if (mode == kModePacked)
        stage_products<NB / kBT, double>(p0, len, idxp, val32, vec64, prod);
    else
        stage_products<NB / kBT, double, double, double, NT64>(p0, len, idx, val64, vec64, prod);
    __syncthreads();
    return group_sum<double>(prod, q0, q1, gl, G);
}

// Dot product of one long row in the instance's mode (see tile_row_sum);
// valid in thread 0.
template <bool NT64 = false, typename TA32, typename TV32>
__device__ __forceinline__ double long_row_mode(int mode, int p0, int p1, const int32_t* __restrict__ idx,
                                                const uint32_t* __restrict__ idxp, const double* __restrict__ val64,
                                                const double* __restrict__ vec64, const TA32* __restrict__ val32,
                                                const TV32* __restrict__ vec32, double* red) {
    if (mode == kModeFp32) return long_row_dot<float>(p0, p1, idx, val32, vec32, red);
    if (mode == kModePacked) return long_row_dot<double>(p0, p1, idxp, val32, vec64, red);
    return long_row_dot<double, NT64>(p0, p1, idx, val64, vec64, red);
}

// ---------------------------------------------------------------------------
// Primal step over block tiles of A^T: z = A^T y, x+ = proj(x - tau (c + z)),
// x_bar = 2 x+ - x, running average (or Halpern update); with ADAPT also dx
// and the tile's partial of |dx|^2_colnorm.
template <int NB, bool ADAPT>
__global__ __launch_bounds__(kBT) void k_primal_blk(Dev d, int t) {
    __shared__ double prod[NB];
    __shared__ double red[kBT / 64];
    const int tile = block_tile_id(d.nbT);
    if (tile >= d.nbT) return;
    // Instance state, iteration counter and the tile's row pointers are read
    // together, and a stopped instance's tile is turned into an empty one
    // rather than an early exit: a launch over a small instance is a chain of
    // dependent round trips, and this keeps it at tile -> (state, row
    // pointers) -> nonzeros -> gathers.
    const int32_t* __restrict__ ptr = d.cp;
    const int64_t* __restrict__ ib = d.iter_base;
    const Tile tl = d.bT[tile];
    const InstState& s = d.st[tl.inst];
    const int32_t status = s.status, mode = s.mode;
    const double eta = s.eta, omega = s.omega;
    const int64_t r_iter = s.restart_iter;
    const int64_t iter = *ib + t;
    const int p0 = ptr[tl.r0], pe = ptr[tl.r1];  // both loaded whatever the status
    const bool run = status == kRunning;  // uniform over the workgroup
    const bool f32 = mode == kModeFp32;
    const int p1 = run ? pe : p0;
    const double kavg = (double)(iter - r_iter + 1);
    const bool halp = d.halpern != 0;
    const double w = kavg / (kavg + 1.0);
    const double* xsrc = halp ? d.xlast : d.xavg;
    const int tid = threadIdx.x;
    double part = 0.0;
    double c = 0, cn = 1, lo = 0, hi = 0, xo = 0, xa = 0;
    auto update = [&](int j, double aty) {
        const double g = c + aty;
        const double tau = (eta / omega) / cn;
        const double xp = clampv(xo - tau * g, lo, hi);
        const double dd = xp - xo;
        if (ADAPT) {
            d.dx[j] = dd;
            if (f32) d.dx32[j] = (float)dd;
            part += cn * dd * dd;
        }
        const double xb = 2.0 * xp - xo;
        d.xbar[j] = xb;
        if (f32) d.xbar32[j] = (float)xb;
        if (halp) {
            d.xavg[j] = xp;
            d.x[j] = w * xb + (1.0 - w) * xa;
        } else {
            d.x[j] = xp;
            d.xavg[j] = xa + (xp - xa) / kavg;
        }
    };
    auto load = [&](int j) {
        c = d.c[j];
        cn = d.cnorm[j];
        lo = d.slo[j];
        hi = d.shi[j];
        xo = d.x[j];
        xa = xsrc[j];  // running average, or the Halpern anchor
    };
    if (tl.shape & 8) {  // one long column (r1 = r0 + 1), strided by the whole workgroup
        const int j = tl.r0;
        const double aty = long_row_mode(mode, p0, p1, d.ri, d.rix, d.tv64, d.y, d.tv32, d.y32, red);
        if (tid == 0 && run) {
            load(j);
            update(j, aty);
            if (ADAPT) d.part_dxn[tile] = part;
        }
        return;
    }
    const int lg = tl.shape & 7, G = 1 << lg;
    const int j = tl.r0 + (tid >> lg), gl = tid & (G - 1);
    const bool own = run && j < tl.r1;
    const int len = p1 - p0;
    int q0 = 0, q1 = 0;
    if (own) {
        q0 = d.cp[j] - p0;
        q1 = d.cp[j + 1] - p0;
        if (gl == 0) load(j);
    }
    const double aty = tile_row_sum<NB>(mode, p0, len, q0, q1, gl, G, d.ri, d.rix, d.tv64, d.y, d.tv32, d.y32, prod);
    if (own && gl == 0) update(j, aty);
    if (ADAPT) {
        part = block_sum<kBT / 64>(part, red);
        if (tid == 0 && run) d.part_dxn[tile] = part;
    }
}

// ---------------------------------------------------------------------------
// Dual step over block tiles of A: A x_bar, Moreau prox, running average (or
// Halpern update); FP32 mode also refreshes the FP32 copy of y.
template <int NB, bool NT>
__global__ __launch_bounds__(kBT) void k_dual_blk(Dev d, int t) {
    __shared__ double prod[NB];
    __shared__ double red[kBT / 64];
    const int tile = block_tile_id(d.nbA);
    if (tile >= d.nbA) return;
    // State, iteration counter and row pointers read together; a stopped
    // instance's tile is empty (as in k_primal_blk).
    const int32_t* __restrict__ ptr = d.rp;
    const int64_t* __restrict__ ib = d.iter_base;
    const Tile tl = d.bA[tile];
    const InstState& s = d.st[tl.inst];
    const int32_t status = s.status, mode = s.mode;
    const double eta = s.eta, omega = s.omega;
    const int64_t r_iter = s.restart_iter;
    const int64_t iter = *ib + t;
    const int p0 = ptr[tl.r0], pe = ptr[tl.r1];  // both loaded whatever the status
    const bool run = status == kRunning;
    const bool f32 = mode == kModeFp32;
    const int p1 = run ? pe : p0;
    const double kavg = (double)(iter - r_iter + 1);
    const bool halp = d.halpern != 0;
    const double w = kavg / (kavg + 1.0);
    const double* ysrc = halp ? d.ylast : d.yavg;
    const int tid = threadIdx.x;
    double rn = 1, bl = 0, bh = 0, yo = 0, ya = 0;
    auto load = [&](int i) {
        rn = d.rnorm[i];
        bl = d.blo[i];
        bh = d.bhi[i];
        yo = d.y[i];
        ya = ysrc[i];
    };
    auto update = [&](int i, double ax) {
        const double sigma = (eta * omega) / rn;
        const double v = yo + sigma * ax;
        const double yn = v - sigma * clampv(v / sigma, bl, bh);
        double ynext;
        if (halp) {
            d.yavg[i] = yn;
            ynext = w * (2.0 * yn - yo) + (1.0 - w) * ya;
        } else {
            ynext = yn;
            d.yavg[i] = ya + (yn - ya) / kavg;
        }
        d.y[i] = ynext;
        if (f32) d.y32[i] = (float)ynext;
    };
    if (tl.shape & 8) {  // one long row (r1 = r0 + 1)
        const int i = tl.r0;
        const double ax = long_row_mode<NT>(mode, p0, p1, d.ci, d.cix, d.av64, d.xbar, d.av32, d.xbar32, red);
        if (tid == 0 && run) {
            load(i);
            update(i, ax);
        }
        return;
    }
    const int lg = tl.shape & 7, G = 1 << lg;
    const int i = tl.r0 + (tid >> lg), gl = tid & (G - 1);
    const bool own = run && i < tl.r1;
    const int len = p1 - p0;
    int q0 = 0, q1 = 0;
    if (own) {
        q0 = d.rp[i] - p0;
        q1 = d.rp[i + 1] - p0;
        if (gl == 0) load(i);
    }
    const double ax = tile_row_sum<NB, NT>(mode, p0, len, q0, q1, gl, G, d.ci, d.cix, d.av64, d.xbar, d.av32, d.xbar32, prod);
    if (own && gl == 0) update(i, ax);
}

// ---------------------------------------------------------------------------
// A dx for the adaptive step size (pdlp.cpp:208-210) over block tiles of A:
// the tile's partial of |A dx|^2_{1/rownorm}.
template <int NB, bool NT>
__global__ __launch_bounds__(kBT) void k_adx_blk(Dev d, int) {
    __shared__ double prod[NB];
    __shared__ double red[kBT / 64];
    const int tile = block_tile_id(d.nbA);
    if (tile >= d.nbA) return;
    // State and row pointers read together; a stopped instance's tile is
    // empty (as in k_primal_blk).
    const int32_t* __restrict__ ptr = d.rp;
    const Tile tl = d.bA[tile];
    const InstState& s = d.st[tl.inst];
    const int32_t status = s.status, mode = s.mode;
    const int p0 = ptr[tl.r0], pe = ptr[tl.r1];  // both loaded whatever the status
    const bool run = status == kRunning;
    const int p1 = run ? pe : p0;
    const int tid = threadIdx.x;
    if (tl.shape & 8) {  // one long row (r1 = r0 + 1)
        const int i = tl.r0;
        const double a = long_row_mode<NT>(mode, p0, p1, d.ci, d.cix, d.av64, d.dx, d.av32, d.dx32, red);
        if (tid == 0 && run) d.part_adx[tile] = (a * a) / d.rnorm[i];
        return;
    }
    const int lg = tl.shape & 7, G = 1 << lg;
    const int i = tl.r0 + (tid >> lg), gl = tid & (G - 1);
    const bool own = run && i < tl.r1;
    const int len = p1 - p0;
    int q0 = 0, q1 = 0;
    double rn = 1.0;
    if (own) {
        q0 = d.rp[i] - p0;
        q1 = d.rp[i + 1] - p0;
        if (gl == 0) rn = d.rnorm[i];
    }
    const double a = tile_row_sum<NB, NT>(mode, p0, len, q0, q1, gl, G, d.ci, d.cix, d.av64, d.dx, d.av32, d.dx32, prod);
    const double part = block_sum<kBT / 64>((own && gl == 0) ? (a * a) / rn : 0.0, red);
    if (tid == 0 && run) d.part_adx[tile] = part;
}

// ---------------------------------------------------------------------------
// FP64 check over block tiles: the residual terms of NC candidates -- the
// running average (Halpern: T(z)) and, with NC == 2, the current iterate --
// from one pass over the matrix (one index/value stream, one gather per
// candidate), on the unscaled model (pdlp.cpp:76-131).  The tile partials go
// to the same field-major slots as the wave-task check kernels of kernels.hip,
// [(c * fields + k) * tiles + tile], with tiles = nbA / nbT.

// Products of the tile's nonzeros with NC gathered vectors into prod[c * NB + q];
// with PK also the packed stream's values times the first vector into
// prod[NC * NB + q] (the packed floor of the check).
template <int K, int NC, bool NT = false, bool PK = false>
__device__ __forceinline__ void stage_products_multi(int p0, int len, const int32_t* __restrict__ idx,
                                                     const double* __restrict__ val, const double* const* vecs,
                                                     double* prod) {
    constexpr int NB = K * kBT;
    int ii[K];
    double aa[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int q = (int)threadIdx.x + k * kBT;
        ii[k] = 0;
        aa[k] = 0.0;
        if (q < len) stream_pair<NT>(idx, val, p0 + q, ii[k], aa[k]);
    }
    double vv[NC][K];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int q = (int)threadIdx.x + k * kBT;
            vv[c][k] = q < len ? vecs[c][ii[k]] : 0.0;
        }
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int q = (int)threadIdx.x + k * kBT;
            if (q < len) prod[c * NB + q] = aa[k] * vv[c][k];
        }
    if (PK) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int q = (int)threadIdx.x + k * kBT;
            if (q < len) prod[NC * NB + q] = packed_of(aa[k]) * vv[0][k];
        }
    }
}

// Rows: A x_c, row violation, dual-objective row part, |y_c - y_last|^2.
// PK: a further candidate, the average with the packed stream's values, in
// slot 2 (the packed floor).
template <int NB, int NC, bool NT, bool PK>
__global__ __launch_bounds__(kBT) void k_resid_rows_blk(Dev d) {
    constexpr int NP = NC + (PK ? 1 : 0);  // products staged per nonzero
    __shared__ double prod[NP * NB];
    __shared__ double red[(kBT / 64) * 3];
    const int tile = block_tile_id(d.nbA);
    if (tile >= d.nbA) return;
    const Tile tl = d.bA[tile];
    const InstState& s = d.st[tl.inst];
    if (s.status != kRunning) return;
    const double* xs[2] = {d.xavg, d.x};
    const double* ys[2] = {d.yavg, d.y};
    const int tid = threadIdx.x;
    double rs = 1, lo = 0, hi = 0, yl = 0, yc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) yc[c] = 0.0;
    auto load = [&](int i) {
        rs = d.row_scale[i];
        lo = d.orow_lo[i];
        hi = d.orow_hi[i];
        yl = d.ylast[i];
#pragma unroll
        for (int c = 0; c < NC; ++c) yc[c] = ys[c][i];
    };
    // residual terms of one row for candidate c: (violation, dual-objective part, movement^2)
    auto terms = [&](int c, double ax, double& viol, double& dob, double& mv) {
        const double axu = ax / rs;
        const double proj = clampv(axu, lo, hi);
        const double r = fabs(axu - proj);
        const double sc = 1.0 + maxv(fabs(axu), fabs(proj) < 1e299 ? fabs(proj) : 0.0);
        viol = r / sc;
        const double yu = yc[c] * rs;
        dob = (yu > 0.0 && hi < 1e299) ? -(yu * hi) : ((yu < 0.0 && lo > -1e299) ? -(yu * lo) : 0.0);
        const double dy = yc[c] - yl;
        mv = dy * dy;
    };
    auto store = [&](int c, double viol, double dob, double mv) {
        double* out = d.part_rows + c * 3 * d.nbA + tile;
        out[0] = viol;
        out[d.nbA] = dob;
        out[2 * d.nbA] = mv;
    };
    if (tl.shape & 8) {
        const int i = tl.r0;
        double ax[NP];
#pragma unroll
        for (int c = 0; c < NC; ++c)
            ax[c] = long_row_dot<double, NT>(d.rp[i], d.rp[i + 1], d.ci, d.av64, xs[c], red);
        if (PK) ax[NP - 1] = long_row_dot<double, NT, true>(d.rp[i], d.rp[i + 1], d.ci, d.av64, xs[0], red);
        if (tid == 0) {
            load(i);
            for (int c = 0; c < NP; ++c) {
                double v, ob, mv;
                terms(c < NC ? c : 0, ax[c], v, ob, mv);
                store(c < NC ? c : 2, v, ob, mv);
            }
        }
        return;
    }
    const int lg = tl.shape & 7, G = 1 << lg;
    const int i = tl.r0 + (tid >> lg), gl = tid & (G - 1);
    const bool own = i < tl.r1;
    const int p0 = d.rp[tl.r0], len = d.rp[tl.r1] - p0;
    int q0 = 0, q1 = 0;
    if (own) {
        q0 = d.rp[i] - p0;
        q1 = d.rp[i + 1] - p0;
        if (gl == 0) load(i);
    }
    stage_products_multi<NB / kBT, NC, NT, PK>(p0, len, d.ci, d.av64, xs, prod);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NP; ++c) {
        const double ax = group_sum<double>(prod + c * NB, q0, q1, gl, G);
        double v = 0.0, ob[2] = {0.0, 0.0};
        if (own && gl == 0) terms(c < NC ? c : 0, ax, v, ob[0], ob[1]);
        block_max_sums<kBT / 64, 2>(v, ob, red);
        if (tid == 0) store(c < NC ? c : 2, v, ob[0], ob[1]);
    }
}

// Columns: A^T y_c, bound violation, dual residual, primal/dual objective
// parts, |x_c - x_last|^2.
template <int NB, int NC, bool PK>
__global__ __launch_bounds__(kBT) void k_resid_cols_blk(Dev d) {
    constexpr int NP = NC + (PK ? 1 : 0);  // products staged per nonzero (slot 2: packed floor)
    __shared__ double prod[NP * NB];
    __shared__ double red[(kBT / 64) * 5];
    const int tile = block_tile_id(d.nbT);
    if (tile >= d.nbT) return;
    const Tile tl = d.bT[tile];
    const InstState& s = d.st[tl.inst];
    if (s.status != kRunning) return;
    const double* xs[2] = {d.xavg, d.x};
    const double* ys[2] = {d.yavg, d.y};
    const double sense = d.sense[tl.inst];
    const int tid = threadIdx.x;
    double cs = 1, lo = 0, hi = 0, ob = 0, xl = 0, xc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) xc[c] = 0.0;
    auto load = [&](int j) {
        cs = d.col_scale[j];
        lo = d.ovar_lo[j];
        hi = d.ovar_hi[j];
        ob = d.obj[j];
        xl = d.xlast[j];
#pragma unroll
        for (int c = 0; c < NC; ++c) xc[c] = xs[c][j];
    };
    // (violation; dual residual^2, primal objective, dual-objective part, movement^2)
    auto terms = [&](int c, double aty, double& viol, double (&t)[4]) {
        const double xu = xc[c] * cs;
        const double proj = clampv(xu, lo, hi);
        const double r = fabs(xu - proj);
        const double sc = 1.0 + maxv(fabs(xu), fabs(proj) < 1e299 ? fabs(proj) : 0.0);
        viol = r / sc;
        const double g = sense * ob + aty / cs;
        const double xp = clampv(xu - g, lo, hi);
        const double df = xu - xp;
        t[0] = df * df;
        t[1] = ob * xu;
        t[2] = (g > 0.0 && lo > -1e299) ? g * lo : ((g < 0.0 && hi < 1e299) ? g * hi : 0.0);
        const double dxm = xc[c] - xl;
        t[3] = dxm * dxm;
    };
    auto store = [&](int c, double viol, const double (&t)[4]) {
        double* out = d.part_cols + c * 5 * d.nbT + tile;
        out[0] = viol;
        out[d.nbT] = t[0];
        out[2 * d.nbT] = t[1];
        out[3 * d.nbT] = t[2];
        out[4 * d.nbT] = t[3];
    };
    if (tl.shape & 8) {
        const int j = tl.r0;
        double aty[NP];
#pragma unroll
        for (int c = 0; c < NC; ++c) aty[c] = long_row_dot<double>(d.cp[j], d.cp[j + 1], d.ri, d.tv64, ys[c], red);
        if (PK) aty[NP - 1] = long_row_dot<double, false, true>(d.cp[j], d.cp[j + 1], d.ri, d.tv64, ys[0], red);
        if (tid == 0) {
            load(j);
            for (int c = 0; c < NP; ++c) {
                double v, t[4];
                terms(c < NC ? c : 0, aty[c], v, t);
                store(c < NC ? c : 2, v, t);
            }
        }
        return;
    }
    const int lg = tl.shape & 7, G = 1 << lg;
    const int j = tl.r0 + (tid >> lg), gl = tid & (G - 1);
    const bool own = j < tl.r1;
    const int p0 = d.cp[tl.r0], len = d.cp[tl.r1] - p0;
    int q0 = 0, q1 = 0;
    if (own) {
        q0 = d.cp[j] - p0;
        q1 = d.cp[j + 1] - p0;
        if (gl == 0) load(j);
    }
    stage_products_multi<NB / kBT, NC, false, PK>(p0, len, d.ri, d.tv64, ys, prod);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NP; ++c) {
        const double aty = group_sum<double>(prod + c * NB, q0, q1, gl, G);
        double v = 0.0, t[4] = {0.0, 0.0, 0.0, 0.0};
        if (own && gl == 0) terms(c < NC ? c : 0, aty, v, t);
        block_max_sums<kBT / 64, 4>(v, t, red);
        if (tid == 0) store(c < NC ? c : 2, v, t);
    }
}

// ---------------------------------------------------------------------------
// Plain SpMV over block tiles: out[r] = sum val * vec (tests, autotuning).
// T: the type of the vectors, products and sums; TA/IX: the stored values and
// index words (T/int32, or float/uint32 for the packed stream with T double).
template <typename T, int NB, typename TA = T, typename IX = int32_t>
__global__ __launch_bounds__(kBT) void k_spmv_blk(const Tile* tiles, int ntiles, const int32_t* __restrict__ ptr,
                                                  const IX* __restrict__ idx, const TA* __restrict__ val,
                                                  const T* __restrict__ vec, T* __restrict__ out) {
    __shared__ T prod[NB];
    __shared__ double red[kBT / 64];
    const int tile = block_tile_id(ntiles);
    if (tile >= ntiles) return;
    const Tile tl = tiles[tile];
    if (tl.shape & 8) {
        const double v = long_row_dot<T>(ptr[tl.r0], ptr[tl.r0 + 1], idx, val, vec, red);
        if (threadIdx.x == 0) out[tl.r0] = (T)v;
        return;
    }
    const int lg = tl.shape & 7, G = 1 << lg;
    const int r = tl.r0 + ((int)threadIdx.x >> lg), gl = (int)threadIdx.x & (G - 1);
    const bool own = r < tl.r1;
    const int p0 = ptr[tl.r0], len = ptr[tl.r1] - p0;
    const int q0 = own ? ptr[r] - p0 : 0, q1 = own ? ptr[r + 1] - p0 : 0;
    stage_products<NB / kBT, T>(p0, len, idx, val, vec, prod);
    __syncthreads();
    const T v = group_sum<T>(prod, q0, q1, gl, G);
    if (own && gl == 0) out[r] = v;
}

// ---------------------------------------------------------------------------
// Gather probe: the FP64 block SpMV above with its gathers redirected, to
// measure what the random gathers of x_bar / y cost on a given tiling.
// GM 0: vec[idx] (the SpMV itself); GM 1: vec[idx & mask] (a window of
// mask + 1 <= 256 entries, so every gather hits in L1/L2); GM 2: no gather
// (the product takes the index as its operand).  The matrix stream, the LDS
// staging and the row sums are those of k_spmv_blk in all three.
template <int GM>
__device__ __forceinline__ double probe_operand(const double* __restrict__ vec, int i, int mask) {
    if (GM == 0) return vec[i];
    if (GM == 1) return vec[i & mask];
    return (double)i;
}

template <int NB, int GM>
__global__ __launch_bounds__(kBT) void k_gather_probe(const Tile* tiles, int ntiles, const int32_t* __restrict__ ptr,
                                                      const int32_t* __restrict__ idx,
                                                      const double* __restrict__ val,
                                                      const double* __restrict__ vec, double* __restrict__ out,
                                                      int mask) {
    constexpr int K = NB / kBT;
    __shared__ double prod[NB];
    __shared__ double red[kBT / 64];
    const int tile = block_tile_id(ntiles);
    if (tile >= ntiles) return;
    const Tile tl = tiles[tile];
    if (tl.shape & 8) {
        double s = 0.0;
        for (int p = ptr[tl.r0] + (int)threadIdx.x; p < ptr[tl.r0 + 1]; p += kBT)
            s += val[p] * probe_operand<GM>(vec, idx[p], mask);
        const double v = block_sum<kBT / 64>(s, red);
        if (threadIdx.x == 0) out[tl.r0] = v;
        return;
    }
    const int lg = tl.shape & 7, G = 1 << lg;
    const int r = tl.r0 + ((int)threadIdx.x >> lg), gl = (int)threadIdx.x & (G - 1);
    const bool own = r < tl.r1;
    const int p0 = ptr[tl.r0], len = ptr[tl.r1] - p0;
    const int q0 = own ? ptr[r] - p0 : 0, q1 = own ? ptr[r + 1] - p0 : 0;
    int ii[K];
    double aa[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int q = (int)threadIdx.x + k * kBT;
        const bool in = q < len;
        ii[k] = in ? idx[p0 + q] : 0;
        aa[k] = in ? val[p0 + q] : 0.0;
    }
    double vv[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int q = (int)threadIdx.x + k * kBT;
        vv[k] = q < len ? probe_operand<GM>(vec, ii[k], mask) : 0.0;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int q = (int)threadIdx.x + k * kBT;
        if (q < len) prod[q] = aa[k] * vv[k];
    }
    __syncthreads();
    const double v = group_sum<double>(prod, q0, q1, gl, G);
    if (own && gl == 0) out[r] = v;
}

// ---------------------------------------------------------------------------
static inline int block_grid(int ntiles) { return kXcd * ((ntiles + kXcd - 1) / kXcd); }

template <int GM>
static void gather_probe(const Tile* tiles, int ntiles, int budget, const int32_t* ptr, const int32_t* idx,
                         const double* val, const double* vec, double* out, int mask, hipStream_t s, hipEvent_t e0,
                         hipEvent_t e1) {
    const dim3 g(block_grid(ntiles)), b(kBT);
    if (budget == 256)
        MCP_LAUNCH_TIMED((k_gather_probe<256, GM>), g, b, s, e0, e1, tiles, ntiles, ptr, idx, val, vec, out, mask);
    else if (budget == 512)
        MCP_LAUNCH_TIMED((k_gather_probe<512, GM>), g, b, s, e0, e1, tiles, ntiles, ptr, idx, val, vec, out, mask);
    else if (budget == 2048)
        MCP_LAUNCH_TIMED((k_gather_probe<2048, GM>), g, b, s, e0, e1, tiles, ntiles, ptr, idx, val, vec, out, mask);
    else
        MCP_LAUNCH_TIMED((k_gather_probe<1024, GM>), g, b, s, e0, e1, tiles, ntiles, ptr, idx, val, vec, out, mask);
}

void launch_gather_probe(const Tile* tiles, int ntiles, int budget, const int32_t* ptr, const int32_t* idx,
                         const double* val, const double* vec, double* out, int mode, int mask, hipStream_t s,
                         hipEvent_t e0, hipEvent_t e1) {
    if (ntiles == 0) return;
    if (mode == 1)
        gather_probe<1>(tiles, ntiles, budget, ptr, idx, val, vec, out, mask, s, e0, e1);
    else if (mode == 2)
        gather_probe<2>(tiles, ntiles, budget, ptr, idx, val, vec, out, mask, s, e0, e1);
    else
        gather_probe<0>(tiles, ntiles, budget, ptr, idx, val, vec, out, mask, s, e0, e1);
}

template <bool ADAPT>
static void primal_blk(const Dev& d, int t, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    const dim3 g(block_grid(d.nbT)), b(kBT);
    if (d.budT == 256)
        MCP_LAUNCH_TIMED((k_primal_blk<256, ADAPT>), g, b, s, e0, e1, d, t);
    else if (d.budT == 512)
        MCP_LAUNCH_TIMED((k_primal_blk<512, ADAPT>), g, b, s, e0, e1, d, t);
    else if (d.budT == 2048)
        MCP_LAUNCH_TIMED((k_primal_blk<2048, ADAPT>), g, b, s, e0, e1, d, t);
    else
        MCP_LAUNCH_TIMED((k_primal_blk<1024, ADAPT>), g, b, s, e0, e1, d, t);
}

void launch_primal_blk(const Dev& d, int t, bool adapt, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    if (d.nbT == 0) return;
    if (adapt)
        primal_blk<true>(d, t, s, e0, e1);
    else
        primal_blk<false>(d, t, s, e0, e1);
}

template <bool NT>
static void dual_blk(const Dev& d, int t, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    const dim3 g(block_grid(d.nbA)), b(kBT);
    if (d.budA == 256)
        MCP_LAUNCH_TIMED((k_dual_blk<256, NT>), g, b, s, e0, e1, d, t);
    else if (d.budA == 512)
        MCP_LAUNCH_TIMED((k_dual_blk<512, NT>), g, b, s, e0, e1, d, t);
    else if (d.budA == 2048)
        MCP_LAUNCH_TIMED((k_dual_blk<2048, NT>), g, b, s, e0, e1, d, t);
    else
        MCP_LAUNCH_TIMED((k_dual_blk<1024, NT>), g, b, s, e0, e1, d, t);
}

void launch_dual_blk(const Dev& d, int t, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    if (d.nbA == 0) return;
    if (d.nt_rows)
        dual_blk<true>(d, t, s, e0, e1);
    else
        dual_blk<false>(d, t, s, e0, e1);
}

template <bool NT>
static void adx_blk(const Dev& d, int t, hipStream_t s) {
    const dim3 g(block_grid(d.nbA)), b(kBT);
    if (d.budA == 256)
        hipLaunchKernelGGL((k_adx_blk<256, NT>), g, b, 0, s, d, t);
    else if (d.budA == 512)
        hipLaunchKernelGGL((k_adx_blk<512, NT>), g, b, 0, s, d, t);
    else if (d.budA == 2048)
        hipLaunchKernelGGL((k_adx_blk<2048, NT>), g, b, 0, s, d, t);
    else
        hipLaunchKernelGGL((k_adx_blk<1024, NT>), g, b, 0, s, d, t);
}

void launch_adx_blk(const Dev& d, int t, hipStream_t s) {
    if (d.nbA == 0) return;
    if (d.nt_rows)
        adx_blk<true>(d, t, s);
    else
        adx_blk<false>(d, t, s);
}

template <int NC, bool NT, bool PK>
static void resid_rows_blk(const Dev& d, hipStream_t s) {
    const dim3 g(block_grid(d.nbA)), b(kBT);
    if (d.budA == 256)
        hipLaunchKernelGGL((k_resid_rows_blk<256, NC, NT, PK>), g, b, 0, s, d);
    else if (d.budA == 512)
        hipLaunchKernelGGL((k_resid_rows_blk<512, NC, NT, PK>), g, b, 0, s, d);
    else if (d.budA == 2048)
        hipLaunchKernelGGL((k_resid_rows_blk<2048, NC, NT, PK>), g, b, 0, s, d);
    else
        hipLaunchKernelGGL((k_resid_rows_blk<1024, NC, NT, PK>), g, b, 0, s, d);
}

template <int NC, bool PK>
static void resid_blk(const Dev& d, hipStream_t s) {
    const dim3 b(kBT);
    if (d.nbA > 0) {
        if (d.nt_rows)
            resid_rows_blk<NC, true, PK>(d, s);
        else
            resid_rows_blk<NC, false, PK>(d, s);
    }
    if (d.nbT > 0) {
        const dim3 g(block_grid(d.nbT));
        if (d.budT == 256)
            hipLaunchKernelGGL((k_resid_cols_blk<256, NC, PK>), g, b, 0, s, d);
        else if (d.budT == 512)
            hipLaunchKernelGGL((k_resid_cols_blk<512, NC, PK>), g, b, 0, s, d);
        else if (d.budT == 2048)
            hipLaunchKernelGGL((k_resid_cols_blk<2048, NC, PK>), g, b, 0, s, d);
        else
            hipLaunchKernelGGL((k_resid_cols_blk<1024, NC, PK>), g, b, 0, s, d);
    }
}

void launch_resid_blk(const Dev& d, bool both, bool packed_floor, hipStream_t s) {
    if (both && packed_floor)
        resid_blk<2, true>(d, s);
    else if (both)
        resid_blk<2, false>(d, s);
    else if (packed_floor)
        resid_blk<1, true>(d, s);
    else
        resid_blk<1, false>(d, s);
}

template <typename T, typename TA = T, typename IX = int32_t>
static void spmv_blk(const Tile* tiles, int ntiles, int budget, const int32_t* ptr, const IX* idx,
                     const TA* val, const T* vec, T* out, hipStream_t s) {
    const dim3 g(block_grid(ntiles)), b(kBT);
    if (budget == 256)
        hipLaunchKernelGGL((k_spmv_blk<T, 256, TA, IX>), g, b, 0, s, tiles, ntiles, ptr, idx, val, vec, out);
    else if (budget == 512)
        hipLaunchKernelGGL((k_spmv_blk<T, 512, TA, IX>), g, b, 0, s, tiles, ntiles, ptr, idx, val, vec, out);
    else if (budget == 2048)
        hipLaunchKernelGGL((k_spmv_blk<T, 2048, TA, IX>), g, b, 0, s, tiles, ntiles, ptr, idx, val, vec, out);
    else
        hipLaunchKernelGGL((k_spmv_blk<T, 1024, TA, IX>), g, b, 0, s, tiles, ntiles, ptr, idx, val, vec, out);
}

void launch_spmv_blk(const Tile* tiles, int ntiles, int budget, const int32_t* ptr, const int32_t* idx,
                     const void* val, const void* vec, void* out, bool fp32, hipStream_t s) {
    if (ntiles == 0) return;
    if (fp32)
        spmv_blk<float>(tiles, ntiles, budget, ptr, idx, (const float*)val, (const float*)vec, (float*)out, s);
    else
        spmv_blk<double>(tiles, ntiles, budget, ptr, idx, (const double*)val, (const double*)vec, (double*)out, s);
}

void launch_spmv_blk_packed(const Tile* tiles, int ntiles, int budget, const int32_t* ptr, const uint32_t* idxp,
                            const float* val32, const double* vec, double* out, hipStream_t s) {
    if (ntiles == 0) return;
    spmv_blk<double, float, uint32_t>(tiles, ntiles, budget, ptr, idxp, val32, vec, out, s);
}

}  // namespace mcp
