// kernels.hip -- PDHG iteration, check and setup kernels for gfx950.
//
// One PDHG iteration (reference src/lp/first_order/pdlp.cpp:194-262) is two
// fused SpMV kernels over the packed problem:
//   k_primal : A^T y (CSR of A^T) + projected primal step, extrapolation
//              x_bar = 2x+ - x, running average, (every 10th) dx and its
//              weighted norm partials                        (pdlp.cpp:195-206, 226)
//   k_dual   : A x_bar (CSR of A) + Moreau-prox dual step + running average
//                                                              (pdlp.cpp:220-227)
// plus, every 10th iteration, k_adx (A dx) and k_adapt_finalize (eta update,
// :207-219), and every restart_every iterations the FP64 check: k_resid
// (A x_avg, A^T y_avg and the residual terms on the unscaled model, :62-135),
// k_check_finalize (termination, restart, primal weight, :229-260) and
// k_post_check (restart copy / FP32->FP64 switch).  Every kernel gates on the
// instance's status and arithmetic mode, so finished instances cost nothing
// and a batch advances with one launch per phase.
//
// Step sizes are recomputed per element from (eta, omega) and the scaled L1
// norms exactly as the reference's update_step_sizes() forms them
// (tau = (eta/omega)/colnorm, sigma = (eta*omega)/rownorm), so no tau/sigma
// vectors are stored or rewritten.  Built with -ffp-contract=off: every
// product and sum is rounded once, as on the host.
#include <math.h>

#include "dev.hpp"

namespace mcp {

template <typename T>
struct Prec;
template <>
struct Prec<double> {
    static constexpr int32_t mode = kModeFp64;
};
template <>
struct Prec<float> {
    static constexpr int32_t mode = kModeFp32;
};

template <typename T>
__device__ __forceinline__ T* pick(double* p64, float* p32);
template <>
__device__ __forceinline__ double* pick<double>(double* p64, float*) { return p64; }
template <>
__device__ __forceinline__ float* pick<float>(double*, float* p32) { return p32; }

__device__ __forceinline__ bool gate(const InstState& s, int32_t mode) {
    return s.status == kRunning && s.mode == mode;
}

// ---------------------------------------------------------------------------
// k_primal: z = A^T y fused with the primal update of every column.
template <typename T, bool ADAPT, bool EXACT>
__global__ __launch_bounds__(kBlock) void k_primal(Dev d, int t) {
    const int tile = xcd_tile(d.ntT);
    if (tile >= d.ntT) return;
    const Tile tl = d.tT[tile];
    const InstState& s = d.st[tl.inst];
    if (!gate(s, Prec<T>::mode)) return;
    __shared__ T prod[kNnzTile];
    __shared__ T rsum[kBlock];
    __shared__ T redT[kWaves];
    __shared__ double redD[kWaves];
    const int64_t iter = *d.iter_base + t;
    const T* y = pick<T>(d.y, d.y32);
    const T* tv = pick<T>(d.tv64, d.tv32);
    const int nr = tile_rowsums<EXACT, T>(tl, d.cp, d.ri, tv, y, prod, rsum, redT);
    const int tid = threadIdx.x;
    const double eta = s.eta, omega = s.omega;
    const T kavg = (T)(double)(iter - s.restart_iter + 1);
    double part = 0.0;
    if (tid < nr) {
        const int j = tl.r0 + tid;
        if constexpr (sizeof(T) == 8) {
            const double g = d.c[j] + rsum[tid];
            const double tau = (eta / omega) / d.cnorm[j];
            const double xo = d.x[j];
            const double xp = clampv(xo - tau * g, d.slo[j], d.shi[j]);
            const double dd = xp - xo;
            if (ADAPT) {
                d.dx[j] = dd;
                part = d.cnorm[j] * dd * dd;
            }
            d.x[j] = xp;
            d.xbar[j] = 2.0 * xp - xo;
            const double xa = d.xavg[j];
            d.xavg[j] = xa + (xp - xa) / kavg;
        } else {
            const float4 pk = d.colpk[j];  // c, lo, hi, 1/colnorm
            const float g = pk.x + rsum[tid];
            const float tau = (float)(eta / omega) * pk.w;
            const float xo = d.x32[j];
            const float xp = clampv(xo - tau * g, pk.y, pk.z);
            const float dd = xp - xo;
            if (ADAPT) {
                d.dx32[j] = dd;
                part = d.cnorm[j] * (double)dd * (double)dd;
            }
            d.x32[j] = xp;
            d.xbar32[j] = 2.0f * xp - xo;
            const float xa = d.xavg32[j];
            d.xavg32[j] = xa + (xp - xa) / kavg;
        }
    }
    if (ADAPT && !EXACT) {
        part = block_sum(part, redD);
        if (tid == 0) d.part_dxn[tile] = part;
    }
}

// ---------------------------------------------------------------------------
// k_dual: A x_bar fused with the dual prox step and the dual running average.
template <typename T, bool EXACT>
__global__ __launch_bounds__(kBlock) void k_dual(Dev d, int t) {
    const int tile = xcd_tile(d.ntA);
    if (tile >= d.ntA) return;
    const Tile tl = d.tA[tile];
    const InstState& s = d.st[tl.inst];
    if (!gate(s, Prec<T>::mode)) return;
    __shared__ T prod[kNnzTile];
    __shared__ T rsum[kBlock];
    __shared__ T redT[kWaves];
    const int64_t iter = *d.iter_base + t;
    const T* xb = pick<T>(d.xbar, d.xbar32);
    const T* av = pick<T>(d.av64, d.av32);
    const int nr = tile_rowsums<EXACT, T>(tl, d.rp, d.ci, av, xb, prod, rsum, redT);
    const int tid = threadIdx.x;
    const double eta = s.eta, omega = s.omega;
    const T kavg = (T)(double)(iter - s.restart_iter + 1);
    if (tid < nr) {
        const int i = tl.r0 + tid;
        if constexpr (sizeof(T) == 8) {
            const double sigma = (eta * omega) / d.rnorm[i];
            const double v = d.y[i] + sigma * rsum[tid];
            const double yn = v - sigma * clampv(v / sigma, d.blo[i], d.bhi[i]);
            d.y[i] = yn;
            const double ya = d.yavg[i];
            d.yavg[i] = ya + (yn - ya) / kavg;
        } else {
            const float4 pk = d.rowpk[i];  // blo, bhi, 1/rownorm
            const float sigma = (float)(eta * omega) * pk.z;
            const float v = d.y32[i] + sigma * rsum[tid];
            const float yn = v - sigma * clampv(v / sigma, pk.x, pk.y);
            d.y32[i] = yn;
            const float ya = d.yavg32[i];
            d.yavg32[i] = ya + (yn - ya) / kavg;
        }
    }
}

// ---------------------------------------------------------------------------
// k_adx: A dx for the adaptive step-size estimate (pdlp.cpp:208-210).
template <typename T, bool EXACT>
__global__ __launch_bounds__(kBlock) void k_adx(Dev d, int) {
    const int tile = xcd_tile(d.ntA);
    if (tile >= d.ntA) return;
    const Tile tl = d.tA[tile];
    const InstState& s = d.st[tl.inst];
    if (!gate(s, Prec<T>::mode)) return;
    __shared__ T prod[kNnzTile];
    __shared__ T rsum[kBlock];
    __shared__ T redT[kWaves];
    __shared__ double redD[kWaves];
    const T* dx = pick<T>(d.dx, d.dx32);
    const T* av = pick<T>(d.av64, d.av32);
    const int nr = tile_rowsums<EXACT, T>(tl, d.rp, d.ci, av, dx, prod, rsum, redT);
    const int tid = threadIdx.x;
    double part = 0.0;
    if (tid < nr) {
        const int i = tl.r0 + tid;
        const double a = (double)rsum[tid];
        if (EXACT)
            d.ax[i] = a;
        else
            part = (a * a) / d.rnorm[i];
    }
    if (!EXACT) {
        part = block_sum(part, redD);
        if (tid == 0) d.part_adx[tile] = part;
    }
}

// eta update from L = sqrt(|A dx|^2_{1/rownorm} / |dx|^2_{colnorm}) (pdlp.cpp:207-219).
__device__ void adapt_decide(InstState& s, double dxn, double adxn) {
    if (!(dxn > 1e-14)) return;
    if (adxn > 1e-14) {
        const double L = sqrt(adxn / dxn);
        if (L > 1e-6) {
            const double target = 0.95 / L;
            if (target < s.eta)
                s.eta = maxv(0.1, maxv(target, s.eta * 0.8));
            else if (target > 1.05 * s.eta && s.eta < 0.99)
                s.eta = minv(0.99, s.eta * 1.05);
        }
    }
}

template <bool EXACT>
__global__ void k_adapt_finalize(Dev d) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.B) return;
    InstState& s = d.st[b];
    if (s.status != kRunning) return;
    double dxn = 0.0, adxn = 0.0;
    if (EXACT) {  // serial sums in the reference's index order (FP64 only)
        for (int j = d.inst_col[b]; j < d.inst_col[b + 1]; ++j) dxn += d.cnorm[j] * d.dx[j] * d.dx[j];
        for (int i = d.inst_row[b]; i < d.inst_row[b + 1]; ++i) adxn += (d.ax[i] * d.ax[i]) / d.rnorm[i];
    } else {
        for (int k = d.tT_begin[b]; k < d.tT_begin[b + 1]; ++k) dxn += d.part_dxn[k];
        for (int k = d.tA_begin[b]; k < d.tA_begin[b + 1]; ++k) adxn += d.part_adx[k];
    }
    adapt_decide(s, dxn, adxn);
}

// ---------------------------------------------------------------------------
// FP64 residual terms of a candidate iterate on the unscaled model
// (pdlp.cpp:76-131).  CUR = false: the running average (the reference's only
// candidate); CUR = true: the current iterate (KKT restart scheme).  Rows:
// A x_c, row violation, dual-objective row part, |y_c - y_last|^2.
template <typename TV, bool EXACT, bool CUR>
__global__ __launch_bounds__(kBlock) void k_resid_rows(Dev d) {
    const int tile = xcd_tile(d.ntA);
    if (tile >= d.ntA) return;
    const Tile tl = d.tA[tile];
    const InstState& s = d.st[tl.inst];
    if (!gate(s, Prec<TV>::mode)) return;
    __shared__ double prod[kNnzTile];
    __shared__ double rsum[kBlock];
    __shared__ double red[kWaves];
    const TV* xc = CUR ? pick<TV>(d.x, d.x32) : pick<TV>(d.xavg, d.xavg32);
    const TV* yc = CUR ? pick<TV>(d.y, d.y32) : pick<TV>(d.yavg, d.yavg32);
    const TV* yl = pick<TV>(d.ylast, d.ylast32);
    const int nr = tile_rowsums<EXACT, double>(tl, d.rp, d.ci, d.av64, xc, prod, rsum, red);
    const int tid = threadIdx.x;
    double viol = 0.0, dob = 0.0, mv = 0.0;
    if (tid < nr) {
        const int i = tl.r0 + tid;
        if (EXACT) {
            (CUR ? d.ax2 : d.ax)[i] = rsum[tid];
        } else {
            const double rs = d.row_scale[i];
            const double axu = rsum[tid] / rs;
            const double lo = d.orow_lo[i], hi = d.orow_hi[i];
            const double proj = clampv(axu, lo, hi);
            const double r = fabs(axu - proj);
            const double sc = 1.0 + maxv(fabs(axu), fabs(proj) < 1e299 ? fabs(proj) : 0.0);
            viol = r / sc;
            const double yu = (double)yc[i] * rs;
            if (yu > 0.0 && hi < 1e299)
                dob = -(yu * hi);
            else if (yu < 0.0 && lo > -1e299)
                dob = -(yu * lo);
            const double dy = (double)yc[i] - (double)yl[i];
            mv = dy * dy;
        }
    }
    if (!EXACT) {
        double* out = d.part_rows + 3 * ((CUR ? d.ntA : 0) + tile);
        viol = block_max(viol, red);
        if (tid == 0) out[0] = viol;
        dob = block_sum(dob, red);
        if (tid == 0) out[1] = dob;
        mv = block_sum(mv, red);
        if (tid == 0) out[2] = mv;
    }
}

// Columns: A^T y_c, bound violation, dual residual, primal/dual objective
// parts, |x_c - x_last|^2.
template <typename TV, bool EXACT, bool CUR>
__global__ __launch_bounds__(kBlock) void k_resid_cols(Dev d) {
    const int tile = xcd_tile(d.ntT);
    if (tile >= d.ntT) return;
    const Tile tl = d.tT[tile];
    const InstState& s = d.st[tl.inst];
    if (!gate(s, Prec<TV>::mode)) return;
    __shared__ double prod[kNnzTile];
    __shared__ double rsum[kBlock];
    __shared__ double red[kWaves];
    const TV* xc = CUR ? pick<TV>(d.x, d.x32) : pick<TV>(d.xavg, d.xavg32);
    const TV* yc = CUR ? pick<TV>(d.y, d.y32) : pick<TV>(d.yavg, d.yavg32);
    const TV* xl = pick<TV>(d.xlast, d.xlast32);
    const int nr = tile_rowsums<EXACT, double>(tl, d.cp, d.ri, d.tv64, yc, prod, rsum, red);
    const int tid = threadIdx.x;
    double viol = 0.0, dres = 0.0, pob = 0.0, dob = 0.0, mv = 0.0;
    if (tid < nr) {
        const int j = tl.r0 + tid;
        if (EXACT) {
            (CUR ? d.aty2 : d.aty)[j] = rsum[tid];
        } else {
            const double cs = d.col_scale[j];
            const double xu = (double)xc[j] * cs;
            const double lo = d.ovar_lo[j], hi = d.ovar_hi[j];
            const double proj = clampv(xu, lo, hi);
            const double r = fabs(xu - proj);
            const double sc = 1.0 + maxv(fabs(xu), fabs(proj) < 1e299 ? fabs(proj) : 0.0);
            viol = r / sc;
            const double atyu = rsum[tid] / cs;
            const double corig = d.sense[tl.inst] * d.obj[j];
            const double g = corig + atyu;
            const double xp = clampv(xu - g, lo, hi);
            const double df = xu - xp;
            dres = df * df;
            pob = d.obj[j] * xu;
            if (g > 0.0 && lo > -1e299)
                dob = g * lo;
            else if (g < 0.0 && hi < 1e299)
                dob = g * hi;
            const double dxm = (double)xc[j] - (double)xl[j];
            mv = dxm * dxm;
        }
    }
    if (!EXACT) {
        double* out = d.part_cols + 5 * ((CUR ? d.ntT : 0) + tile);
        viol = block_max(viol, red);
        if (tid == 0) out[0] = viol;
        dres = block_sum(dres, red);
        if (tid == 0) out[1] = dres;
        pob = block_sum(pob, red);
        if (tid == 0) out[2] = pob;
        dob = block_sum(dob, red);
        if (tid == 0) out[3] = dob;
        mv = block_sum(mv, red);
        if (tid == 0) out[4] = mv;
    }
}

struct CandResid {
    double primal, dual, gap, score, objective, xmove2, ymove2;
};

// Reference-order residuals (pdlp.cpp:80-133), serial, of candidate
// (xs, ys) of instance b from the stored products ax = A_s xs, aty = A_s^T ys;
// movement sums |xs - x_last|^2, |ys - y_last|^2 in index order.
__device__ void serial_residuals(const Dev& d, int b, const double* xs, const double* ys,
                                 const double* ax, const double* aty, double c_scale,
                                 CandResid& out) {
    const int r0 = d.inst_row[b], r1 = d.inst_row[b + 1];
    const int c0 = d.inst_col[b], c1 = d.inst_col[b + 1];
    const double sense = d.sense[b], off = d.offset[b];
    double mp = 0.0;
    for (int i = r0; i < r1; ++i) {
        const double axu = ax[i] / d.row_scale[i];
        const double lo = d.orow_lo[i], hi = d.orow_hi[i];
        const double proj = clampv(axu, lo, hi);
        const double r = fabs(axu - proj);
        const double sc = 1.0 + maxv(fabs(axu), fabs(proj) < 1e299 ? fabs(proj) : 0.0);
        mp = maxv(mp, r / sc);
    }
    for (int j = c0; j < c1; ++j) {
        const double xu = xs[j] * d.col_scale[j];
        const double lo = d.ovar_lo[j], hi = d.ovar_hi[j];
        const double proj = clampv(xu, lo, hi);
        const double r = fabs(xu - proj);
        const double sc = 1.0 + maxv(fabs(xu), fabs(proj) < 1e299 ? fabs(proj) : 0.0);
        mp = maxv(mp, r / sc);
    }
    out.primal = mp;
    double dsq = 0.0;
    for (int j = c0; j < c1; ++j) {
        const double xu = xs[j] * d.col_scale[j];
        const double g = sense * d.obj[j] + aty[j] / d.col_scale[j];
        const double xp = clampv(xu - g, d.ovar_lo[j], d.ovar_hi[j]);
        const double df = xu - xp;
        dsq += df * df;
    }
    out.dual = sqrt(dsq) / c_scale;
    double po = off;
    for (int j = c0; j < c1; ++j) po += d.obj[j] * (xs[j] * d.col_scale[j]);
    out.objective = po;
    double dob = off;
    for (int i = r0; i < r1; ++i) {
        const double yu = ys[i] * d.row_scale[i];
        const double lo = d.orow_lo[i], hi = d.orow_hi[i];
        if (yu > 0.0 && hi < 1e299)
            dob -= yu * hi;
        else if (yu < 0.0 && lo > -1e299)
            dob -= yu * lo;
    }
    for (int j = c0; j < c1; ++j) {
        const double g = sense * d.obj[j] + aty[j] / d.col_scale[j];
        const double lo = d.ovar_lo[j], hi = d.ovar_hi[j];
        if (g > 0.0 && lo > -1e299)
            dob += g * lo;
        else if (g < 0.0 && hi < 1e299)
            dob += g * hi;
    }
    out.gap = fabs(po - dob) / (1.0 + fabs(po) + fabs(dob));
    double sc = out.primal;
    if (sc < out.dual) sc = out.dual;
    if (sc < out.gap) sc = out.gap;
    out.score = sc;
    double xm = 0.0, ym = 0.0;
    for (int j = c0; j < c1; ++j) xm += (xs[j] - d.xlast[j]) * (xs[j] - d.xlast[j]);
    for (int i = r0; i < r1; ++i) ym += (ys[i] - d.ylast[i]) * (ys[i] - d.ylast[i]);
    out.xmove2 = xm;
    out.ymove2 = ym;
}

// Candidate residuals from the tile partials (fast order).
__device__ void partial_residuals(const Dev& d, int b, int cand, double c_scale, CandResid& out) {
    double mp = 0.0, dsq = 0.0, po = 0.0, dr = 0.0, dc = 0.0, xm = 0.0, ym = 0.0;
    const double* pr = d.part_rows + 3 * (cand ? d.ntA : 0);
    const double* pc = d.part_cols + 5 * (cand ? d.ntT : 0);
    for (int k = d.tA_begin[b]; k < d.tA_begin[b + 1]; ++k) {
        mp = maxv(mp, pr[3 * k]);
        dr += pr[3 * k + 1];
        ym += pr[3 * k + 2];
    }
    for (int k = d.tT_begin[b]; k < d.tT_begin[b + 1]; ++k) {
        mp = maxv(mp, pc[5 * k]);
        dsq += pc[5 * k + 1];
        po += pc[5 * k + 2];
        dc += pc[5 * k + 3];
        xm += pc[5 * k + 4];
    }
    out.primal = mp;
    out.dual = sqrt(dsq) / c_scale;
    out.objective = d.offset[b] + po;
    const double dobj = d.offset[b] + dr + dc;
    out.gap = fabs(out.objective - dobj) / (1.0 + fabs(out.objective) + fabs(dobj));
    double sc = out.primal;
    if (sc < out.dual) sc = out.dual;
    if (sc < out.gap) sc = out.gap;
    out.score = sc;
    out.xmove2 = xm;
    out.ymove2 = ym;
}

__device__ __forceinline__ void stop_instance(const Dev& d, InstState& s, int32_t status, int64_t it) {
    s.status = status;
    s.done_iter = it;
    atomicSub(&d.counters[0], 1);
    if (s.mode == kModeFp32) atomicSub(&d.counters[1], 1);
}

// Termination and restart decision of one instance (pdlp.cpp:229-260; the
// KKT scheme as documented for MC_SCHEME_KKT).  FINAL: the iteration limit
// was reached -- record the better candidate's residuals and stop.
template <bool EXACT, bool FINAL>
__global__ void k_check_finalize(Dev d, IterOpts o, int t) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.B) return;
    InstState& s = d.st[b];
    s.restart_flag = 0;
    s.switch_flag = 0;
    if (s.status != kRunning) return;
    const int64_t iter_after = *d.iter_base + t + 1;
    CandResid r, rc;
    if (EXACT)
        serial_residuals(d, b, d.xavg, d.yavg, d.ax, d.aty, s.c_scale, r);
    else
        partial_residuals(d, b, 0, s.c_scale, r);
    int use_cur = 0;
    if (o.kkt) {
        if (EXACT)
            serial_residuals(d, b, d.x, d.y, d.ax2, d.aty2, s.c_scale, rc);
        else
            partial_residuals(d, b, 1, s.c_scale, rc);
        if (rc.score < r.score) {
            r = rc;
            use_cur = 1;
        }
    }
    s.primal = r.primal;
    s.dual = r.dual;
    s.gap = r.gap;
    s.score = r.score;
    s.objective = r.objective;
    s.answer_cur = use_cur;
    s.cand_cur = use_cur;
    s.mode_at_check = s.mode;
    if (FINAL) {
        stop_instance(d, s, kIterLimit, iter_after);
        return;
    }
    if (r.primal <= o.primal_tol && r.dual <= o.dual_tol && r.gap <= o.gap_tol) {
        stop_instance(d, s, kOptimal, iter_after);
        return;
    }
    if (!(r.score == r.score)) {  // NaN: the iterate diverged
        stop_instance(d, s, kNumFail, iter_after);
        return;
    }
    const double score = r.score;
    const int64_t since = iter_after - s.restart_iter;
    bool do_restart;
    if (!o.kkt) {
        do_restart = (o.restart_strategy == 2) ||
                     (o.restart_strategy == 1 &&
                      (score <= o.restart_factor * s.last_restart_score ||
                       (since >= 5 * o.restart_every && score < s.last_restart_score)));
    } else {
        do_restart = (o.restart_strategy == 2) ||
                     (o.restart_strategy == 1 &&
                      (score <= 0.2 * s.last_restart_score ||
                       (score <= 0.8 * s.last_restart_score && score > s.prev_cand) ||
                       (double)since >= 0.36 * (double)iter_after));
        s.prev_cand = score;
    }
    if (do_restart) {
        if (!o.kkt) {
            if (o.adaptive_weight && r.primal > 1e-12 && r.dual > 1e-12) {
                const double ratio = clampv(sqrt(r.primal / r.dual), 0.05, 20.0);
                const double f = (o.smoothing == 0.5) ? sqrt(ratio) : pow(ratio, o.smoothing);
                s.omega = clampv(s.omega * f, 1e-6, 1e6);
            }
        } else {
            if (o.adaptive_weight) {
                const double dX = sqrt(r.xmove2), dY = sqrt(r.ymove2);
                if (dX > 1e-10 && dY > 1e-10) {
                    const double th = o.smoothing, ratio = dY / dX;
                    const double w = (th == 0.5) ? sqrt(s.omega * ratio)
                                                 : pow(ratio, th) * pow(s.omega, 1.0 - th);
                    s.omega = clampv(w, 1e-6, 1e6);
                }
            }
            s.prev_cand = 1e300;
        }
        s.restart_flag = 1;
        s.restart_iter = iter_after;
        s.last_restart_score = score;
    }
    if (s.mode == kModeFp32) {
        // FP32 iterations stop paying once the score is near tolerance or has
        // stopped improving; continue from this iterate in FP64.
        if (score < 0.9 * s.best_score) {
            s.best_score = score;
            s.stall = 0;
        } else {
            s.stall += 1;
        }
        if (score <= o.polish_threshold || s.stall >= 12) {
            s.switch_flag = 1;
            s.mode = kModeFp64;
            s.fp32_iters = iter_after;
            atomicSub(&d.counters[1], 1);
        }
    }
}

// Restart to the candidate (pdlp.cpp:257; KKT scheme: the current iterate or
// the average, which also becomes the new x_last/y_last) and FP32 -> FP64
// conversion of an instance's iterate.  Grid: column tiles, then row tiles.
__global__ __launch_bounds__(kBlock) void k_post_check(Dev d, int kkt) {
    const int g = blockIdx.x;
    const bool cols = g < d.ntT;
    const Tile tl = cols ? d.tT[g] : d.tA[g - d.ntT];
    const InstState& s = d.st[tl.inst];
    if (!(s.restart_flag || s.switch_flag)) return;
    const bool r32 = s.mode_at_check == kModeFp32;
    const bool to_avg = !s.cand_cur;
    double* v64 = cols ? d.x : d.y;
    double* a64 = cols ? d.xavg : d.yavg;
    double* l64 = cols ? d.xlast : d.ylast;
    float* v32 = cols ? d.x32 : d.y32;
    float* a32 = cols ? d.xavg32 : d.yavg32;
    float* l32 = cols ? d.xlast32 : d.ylast32;
    for (int k = tl.r0 + threadIdx.x; k < tl.r1; k += kBlock) {
        if (s.restart_flag) {
            if (r32) {
                if (to_avg) v32[k] = a32[k];
                if (kkt) l32[k] = v32[k];
            } else {
                if (to_avg) v64[k] = a64[k];
                if (kkt) l64[k] = v64[k];
            }
        }
        if (s.switch_flag) {
            v64[k] = (double)v32[k];
            a64[k] = (double)a32[k];
            l64[k] = (double)l32[k];
        }
    }
}

__global__ void k_advance(Dev d, int len) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *d.iter_base += len;
}

// ---------------------------------------------------------------------------
// Setup: device Ruiz equilibration (ruiz_scaling.cpp:152-247).  All maxima
// are order-independent, so the result equals the serial host pass bit for
// bit.  Nonnegative doubles order like their bit patterns, so atomicMax on
// the 64-bit pattern is an exact max.
__device__ __forceinline__ void atomic_max_nonneg(double* addr, double v) {
    atomicMax((unsigned long long*)addr, (unsigned long long)__double_as_longlong(v));
}

__global__ void k_ruiz_colmax(Dev d) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.n) return;
    const int b = d.col_inst[j];
    if (d.st[b].ruiz_done) return;
    double mx = 0.0;
    for (int p = d.cp[j]; p < d.cp[j + 1]; ++p) mx = maxv(mx, fabs(d.tv64[p]));
    if (mx == 0.0) mx = 1.0;
    d.ctmp[j] = mx;
    atomic_max_nonneg(&d.st[b].ruiz_err_c, fabs(1.0 - mx));
}

__global__ void k_ruiz_rowmax(Dev d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.m) return;
    const int b = d.row_inst[i];
    if (d.st[b].ruiz_done) return;
    double mx = 0.0;
    for (int p = d.rp[i]; p < d.rp[i + 1]; ++p) mx = maxv(mx, fabs(d.av64[p]));
    if (mx == 0.0) mx = 1.0;
    d.rtmp[i] = mx;
    atomic_max_nonneg(&d.st[b].ruiz_err_r, fabs(1.0 - mx));
}

__global__ void k_ruiz_decide(Dev d, double tol) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.B) return;
    InstState& s = d.st[b];
    if (s.ruiz_done) return;
    if (s.ruiz_err_r < tol && s.ruiz_err_c < tol) {
        s.ruiz_done = 1;
        s.ruiz_converged = 1;
    } else {
        s.ruiz_passes += 1;
    }
    s.ruiz_err_r = 0.0;
    s.ruiz_err_c = 0.0;
}

// delta = 1/sqrt(norm); scale *= delta.  rtmp/ctmp then hold the deltas.
__global__ void k_ruiz_deltas(Dev d) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < d.m) {
        if (!d.st[d.row_inst[k]].ruiz_done) {
            const double dr = 1.0 / sqrt(d.rtmp[k]);
            d.rtmp[k] = dr;
            d.row_scale[k] *= dr;
        }
    } else if (k < d.m + d.n) {
        const int j = k - d.m;
        if (!d.st[d.col_inst[j]].ruiz_done) {
            const double dc = 1.0 / sqrt(d.ctmp[j]);
            d.ctmp[j] = dc;
            d.col_scale[j] *= dc;
        }
    }
}

// A <- D_r A D_c on both stored orientations (value *= dr[i] * dc[j]).
__global__ void k_ruiz_apply(Dev d) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < d.n) {
        if (d.st[d.col_inst[k]].ruiz_done) return;
        const double dc = d.ctmp[k];
        for (int p = d.cp[k]; p < d.cp[k + 1]; ++p) d.tv64[p] *= d.rtmp[d.ri[p]] * dc;
    } else if (k < d.n + d.m) {
        const int i = k - d.n;
        if (d.st[d.row_inst[i]].ruiz_done) return;
        const double dr = d.rtmp[i];
        for (int p = d.rp[i]; p < d.rp[i + 1]; ++p) d.av64[p] *= dr * d.ctmp[d.ci[p]];
    }
}

// Scaled bounds/objective (ruiz_scaling.cpp:221-244), sentinels and norms
// (pdlp.cpp:153-176) and FP32 copies; once per loaded problem.  Original
// bounds stay as IEEE +-inf for infinite kinds: every test the residuals make
// on them (clamp, |proj| < 1e299, hi < 1e299) reads the same on inf as on the
// reference's 1e300 sentinel.
__global__ void k_setup_rows(Dev d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.m) return;
    const int b = d.row_inst[i];
    const double rs = d.row_scale[i];
    const double lo = d.orow_lo[i], hi = d.orow_hi[i];
    const bool flo = !isinf(lo), fhi = !isinf(hi);
    const double bl = flo ? lo * rs : -1e300;
    const double bh = fhi ? hi * rs : 1e300;
    d.blo[i] = bl;
    d.bhi[i] = bh;
    double bn = 0.0;
    if (fabs(bl) < 1e299) bn = fabs(bl);
    if (fabs(bh) < 1e299) bn = maxv(bn, fabs(bh));
    atomic_max_nonneg(&d.st[b].b_norm_inf, bn);
    double nrm = 0.0;
    for (int p = d.rp[i]; p < d.rp[i + 1]; ++p) {
        nrm += fabs(d.av64[p]);
        d.av32[p] = (float)d.av64[p];
    }
    if (nrm < 1e-12) nrm = 1.0;
    d.rnorm[i] = nrm;
    d.rowpk[i] = make_float4(flo ? (float)bl : -INFINITY, fhi ? (float)bh : INFINITY,
                             (float)(1.0 / nrm), 0.0f);
}

__global__ void k_setup_cols(Dev d) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.n) return;
    const int b = d.col_inst[j];
    const double cs = d.col_scale[j];
    const double inv = 1.0 / cs;
    const double lo = d.ovar_lo[j], hi = d.ovar_hi[j];
    const bool flo = !isinf(lo), fhi = !isinf(hi);
    const double sl = flo ? lo * inv : -1e300;
    const double sh = fhi ? hi * inv : 1e300;
    d.slo[j] = sl;
    d.shi[j] = sh;
    const double sense = d.sense[b];
    const double cj = sense * (d.obj[j] * cs);
    d.c[j] = cj;
    atomic_max_nonneg(&d.st[b].c_norm_inf, fabs(cj));
    atomic_max_nonneg(&d.st[b].c_scale, fabs(sense * d.obj[j]));
    double nrm = 0.0;
    for (int p = d.cp[j]; p < d.cp[j + 1]; ++p) {
        nrm += fabs(d.tv64[p]);
        d.tv32[p] = (float)d.tv64[p];
    }
    if (nrm < 1e-12) nrm = 1.0;
    d.cnorm[j] = nrm;
    d.colpk[j] = make_float4((float)cj, flo ? (float)sl : -INFINITY, fhi ? (float)sh : INFINITY,
                             (float)(1.0 / nrm));
}

// Start point of a solve (pdlp.cpp:186-191): x = proj(0) (or the projected
// warm start mapped into the scaled space), y = 0 (or the warm dual).
__global__ void k_init_iterates(Dev d, const double* warm_x, const double* warm_y) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < d.n) {
        const double x0 = clampv(warm_x ? warm_x[k] / d.col_scale[k] : 0.0, d.slo[k], d.shi[k]);
        d.x[k] = x0;
        d.xbar[k] = x0;
        d.xavg[k] = x0;
        d.x32[k] = (float)x0;
        d.xbar32[k] = (float)x0;
        d.xavg32[k] = (float)x0;
        d.xlast[k] = x0;
        d.xlast32[k] = (float)x0;
    } else if (k < d.n + d.m) {
        const int i = k - d.n;
        const double y0 = warm_y ? warm_y[i] / d.row_scale[i] : 0.0;
        d.y[i] = y0;
        d.yavg[i] = y0;
        d.y32[i] = (float)y0;
        d.yavg32[i] = (float)y0;
        d.ylast[i] = y0;
        d.ylast32[i] = (float)y0;
    }
}

__global__ void k_setup_state(Dev d, double initial_weight, double step_reduction, int mixed) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.B) return;
    InstState& s = d.st[b];
    s.omega = (initial_weight > 0.0) ? initial_weight
                                     : clampv(sqrt(s.c_norm_inf / s.b_norm_inf), 0.01, 100.0);
    s.eta = clampv(step_reduction, 0.1, 0.99);
    s.last_restart_score = 1e300;
    s.best_score = 1e300;
    s.restart_iter = 0;
    s.done_iter = 0;
    s.fp32_iters = 0;
    s.status = kRunning;
    s.mode = mixed ? kModeFp32 : kModeFp64;
    s.mode_at_check = s.mode;
    s.restart_flag = 0;
    s.switch_flag = 0;
    s.stall = 0;
    s.cand_cur = 0;
    s.answer_cur = 0;
    s.prev_cand = 1e300;
    s.primal = s.dual = s.gap = s.score = 1e300;
    s.objective = 0.0;
}

// Unscaled reported iterate (pdlp.cpp:76-79): x = x_c * col_scale, y = y_c * row_scale
// for the candidate of the last check (the average under the reference scheme).
__global__ void k_unscale(Dev d, double* xo, double* yo) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < d.n) {
        const InstState& s = d.st[d.col_inst[k]];
        const bool f32 = s.mode == kModeFp32;
        const double v = s.answer_cur ? (f32 ? (double)d.x32[k] : d.x[k])
                                      : (f32 ? (double)d.xavg32[k] : d.xavg[k]);
        xo[k] = v * d.col_scale[k];
    } else if (k < d.n + d.m) {
        const int i = k - d.n;
        const InstState& s = d.st[d.row_inst[i]];
        const bool f32 = s.mode == kModeFp32;
        const double v = s.answer_cur ? (f32 ? (double)d.y32[i] : d.y[i])
                                      : (f32 ? (double)d.yavg32[i] : d.yavg[i]);
        yo[i] = v * d.row_scale[i];
    }
}

__global__ void k_fill(double* p, double v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <typename T, bool EXACT>
__global__ __launch_bounds__(kBlock) void k_spmv_plain(const Tile* tiles, int ntiles,
                                                       const int32_t* ptr, const int32_t* idx,
                                                       const T* val, const T* vec, T* out) {
    const int tile = xcd_tile(ntiles);
    if (tile >= ntiles) return;
    const Tile tl = tiles[tile];
    __shared__ T prod[kNnzTile];
    __shared__ T rsum[kBlock];
    __shared__ T red[kWaves];
    const int nr = tile_rowsums<EXACT, T>(tl, ptr, idx, val, vec, prod, rsum, red);
    if ((int)threadIdx.x < nr) out[tl.r0 + threadIdx.x] = rsum[threadIdx.x];
}

// ---------------------------------------------------------------------------
// Host launch helpers.
static inline int tile_grid(int ntiles) { return kXcd * ((ntiles + kXcd - 1) / kXcd); }
static inline int elem_grid(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

void launch_primal(const Dev& d, int t, bool fp32, bool adapt, bool exact, hipStream_t s) {
    if (d.ntT == 0) return;
    const dim3 g(tile_grid(d.ntT)), b(kBlock);
    if (fp32) {
        if (adapt)
            hipLaunchKernelGGL((k_primal<float, true, false>), g, b, 0, s, d, t);
        else
            hipLaunchKernelGGL((k_primal<float, false, false>), g, b, 0, s, d, t);
    } else if (exact) {
        if (adapt)
            hipLaunchKernelGGL((k_primal<double, true, true>), g, b, 0, s, d, t);
        else
            hipLaunchKernelGGL((k_primal<double, false, true>), g, b, 0, s, d, t);
    } else {
        if (adapt)
            hipLaunchKernelGGL((k_primal<double, true, false>), g, b, 0, s, d, t);
        else
            hipLaunchKernelGGL((k_primal<double, false, false>), g, b, 0, s, d, t);
    }
}

void launch_dual(const Dev& d, int t, bool fp32, bool exact, hipStream_t s) {
    if (d.ntA == 0) return;
    const dim3 g(tile_grid(d.ntA)), b(kBlock);
    if (fp32)
        hipLaunchKernelGGL((k_dual<float, false>), g, b, 0, s, d, t);
    else if (exact)
        hipLaunchKernelGGL((k_dual<double, true>), g, b, 0, s, d, t);
    else
        hipLaunchKernelGGL((k_dual<double, false>), g, b, 0, s, d, t);
}

void launch_adx(const Dev& d, int t, bool fp32, bool exact, hipStream_t s) {
    if (d.ntA == 0) return;
    const dim3 g(tile_grid(d.ntA)), b(kBlock);
    if (fp32)
        hipLaunchKernelGGL((k_adx<float, false>), g, b, 0, s, d, t);
    else if (exact)
        hipLaunchKernelGGL((k_adx<double, true>), g, b, 0, s, d, t);
    else
        hipLaunchKernelGGL((k_adx<double, false>), g, b, 0, s, d, t);
}

void launch_adapt_finalize(const Dev& d, bool exact, hipStream_t s) {
    const dim3 g(elem_grid(d.B)), b(kBlock);
    if (exact)
        hipLaunchKernelGGL((k_adapt_finalize<true>), g, b, 0, s, d);
    else
        hipLaunchKernelGGL((k_adapt_finalize<false>), g, b, 0, s, d);
}

template <bool CUR>
static void resid_pair(const Dev& d, bool fp32, bool exact, hipStream_t s) {
    const dim3 b(kBlock);
    if (d.ntA > 0) {
        const dim3 g(tile_grid(d.ntA));
        if (fp32)
            hipLaunchKernelGGL((k_resid_rows<float, false, CUR>), g, b, 0, s, d);
        else if (exact)
            hipLaunchKernelGGL((k_resid_rows<double, true, CUR>), g, b, 0, s, d);
        else
            hipLaunchKernelGGL((k_resid_rows<double, false, CUR>), g, b, 0, s, d);
    }
    if (d.ntT > 0) {
        const dim3 g(tile_grid(d.ntT));
        if (fp32)
            hipLaunchKernelGGL((k_resid_cols<float, false, CUR>), g, b, 0, s, d);
        else if (exact)
            hipLaunchKernelGGL((k_resid_cols<double, true, CUR>), g, b, 0, s, d);
        else
            hipLaunchKernelGGL((k_resid_cols<double, false, CUR>), g, b, 0, s, d);
    }
}

void launch_resid(const Dev& d, bool fp32, bool exact, bool current, hipStream_t s) {
    if (current)
        resid_pair<true>(d, fp32, exact, s);
    else
        resid_pair<false>(d, fp32, exact, s);
}

void launch_check_finalize(const Dev& d, const IterOpts& o, int t, bool exact, bool final_pass,
                           hipStream_t s) {
    const dim3 g(elem_grid(d.B)), b(kBlock);
    if (exact) {
        if (final_pass)
            hipLaunchKernelGGL((k_check_finalize<true, true>), g, b, 0, s, d, o, t);
        else
            hipLaunchKernelGGL((k_check_finalize<true, false>), g, b, 0, s, d, o, t);
    } else {
        if (final_pass)
            hipLaunchKernelGGL((k_check_finalize<false, true>), g, b, 0, s, d, o, t);
        else
            hipLaunchKernelGGL((k_check_finalize<false, false>), g, b, 0, s, d, o, t);
    }
}

void launch_post_check(const Dev& d, bool kkt, hipStream_t s) {
    if (d.ntT + d.ntA == 0) return;
    hipLaunchKernelGGL(k_post_check, dim3(d.ntT + d.ntA), dim3(kBlock), 0, s, d, kkt ? 1 : 0);
}

void launch_advance(const Dev& d, int len, hipStream_t s) {
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, s, d, len);
}

void launch_ruiz_pass(const Dev& d, double tol, hipStream_t s) {
    const dim3 b(kBlock);
    hipLaunchKernelGGL(k_ruiz_colmax, dim3(elem_grid(d.n)), b, 0, s, d);
    hipLaunchKernelGGL(k_ruiz_rowmax, dim3(elem_grid(d.m)), b, 0, s, d);
    hipLaunchKernelGGL(k_ruiz_decide, dim3(elem_grid(d.B)), b, 0, s, d, tol);
    hipLaunchKernelGGL(k_ruiz_deltas, dim3(elem_grid((int64_t)d.m + d.n)), b, 0, s, d);
    hipLaunchKernelGGL(k_ruiz_apply, dim3(elem_grid((int64_t)d.m + d.n)), b, 0, s, d);
}

void launch_setup_data(const Dev& d, hipStream_t s) {
    const dim3 b(kBlock);
    if (d.m > 0) hipLaunchKernelGGL(k_setup_rows, dim3(elem_grid(d.m)), b, 0, s, d);
    if (d.n > 0) hipLaunchKernelGGL(k_setup_cols, dim3(elem_grid(d.n)), b, 0, s, d);
}

void launch_init_iterates(const Dev& d, const double* warm_x, const double* warm_y, hipStream_t s) {
    const int64_t tot = (int64_t)d.m + d.n;
    if (tot == 0) return;
    hipLaunchKernelGGL(k_init_iterates, dim3(elem_grid(tot)), dim3(kBlock), 0, s, d, warm_x, warm_y);
}

void launch_setup_state(const Dev& d, double initial_weight, double step_reduction, bool mixed,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_setup_state, dim3(elem_grid(d.B)), dim3(kBlock), 0, s, d, initial_weight,
                       step_reduction, mixed ? 1 : 0);
}

void launch_unscale(const Dev& d, double* x_out, double* y_out, hipStream_t s) {
    const int64_t tot = (int64_t)d.m + d.n;
    if (tot == 0) return;
    hipLaunchKernelGGL(k_unscale, dim3(elem_grid(tot)), dim3(kBlock), 0, s, d, x_out, y_out);
}

void launch_fill(double* p, double v, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill, dim3(elem_grid(n)), dim3(kBlock), 0, s, p, v, n);
}

void launch_spmv_plain(const Tile* tiles, int ntiles, const int32_t* ptr, const int32_t* idx,
                       const void* val, const void* vec, void* out, bool fp32, bool exact,
                       hipStream_t s) {
    if (ntiles == 0) return;
    const dim3 g(tile_grid(ntiles)), b(kBlock);
    if (fp32)
        hipLaunchKernelGGL((k_spmv_plain<float, false>), g, b, 0, s, tiles, ntiles, ptr, idx,
                           (const float*)val, (const float*)vec, (float*)out);
    else if (exact)
        hipLaunchKernelGGL((k_spmv_plain<double, true>), g, b, 0, s, tiles, ntiles, ptr, idx,
                           (const double*)val, (const double*)vec, (double*)out);
    else
        hipLaunchKernelGGL((k_spmv_plain<double, false>), g, b, 0, s, tiles, ntiles, ptr, idx,
                           (const double*)val, (const double*)vec, (double*)out);
}

}  // namespace mcp
