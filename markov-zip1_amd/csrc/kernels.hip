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
// (A x_c, A^T y_c and the residual terms on the unscaled model, :62-135),
// k_check_finalize (termination, restart, primal weight, :229-260),
// optionally the infeasibility-certificate kernels k_ray_*, and k_post_check
// (restart copy).
//
// The kernels here are wave tasks (one Tile per wave64, no LDS, no workgroup
// barrier): a wave gates on its instance's status, picks FP32 or FP64 SpMV
// operands from the instance's mode at run time (so mixed precision costs no
// extra launches; the iterate is FP64 in both), sweeps its rows 64/G at a time
// and lets lane 0 of each row group apply the update.  They run the reference
// reduction order (one lane per row), the setup, and the ray certificates; in
// the fast order the launch helpers below hand the primal/dual steps, A dx and
// the check's residual terms to the block-tile kernels of stream.hip when the
// engine has block tiles.  Step sizes are recomputed per element from
// (eta, omega) and the scaled L1 norms exactly as the reference's
// update_step_sizes() forms them (tau = (eta/omega)/colnorm,
// sigma = (eta*omega)/rownorm).  Built with -ffp-contract=off: every product
// and sum is rounded once, as on the host.
#include <math.h>

#include <algorithm>

#include "dev.hpp"

namespace mcp {

struct Lanes {
    int lane, lg, G, gl, step;
};
__device__ __forceinline__ Lanes lanes_of(const Tile& tl) {
    Lanes l;
    l.lane = threadIdx.x & 63;
    l.lg = tl.shape & 7;
    l.G = 1 << l.lg;
    l.gl = l.lane & (l.G - 1);
    l.step = 64 >> l.lg;
    return l;
}

// ---------------------------------------------------------------------------
// k_primal: z = A^T y fused with the primal update of every column.
// The iterate is FP64 in both modes; FP32 mode gathers the FP32 copy of y
// through the FP32 values of A^T (FP32 products and sums) and also writes the
// FP32 copies of x_bar (and dx) that the next FP32 SpMVs gather.
template <bool ADAPT, bool EXACT>
__global__ __launch_bounds__(kBlock) void k_primal(Dev d, int t) {
    const int tile = wave_tile(d.ntT);
    if (tile >= d.ntT) return;
    const Tile tl = d.tT[tile];
    const InstState& s = d.st[tl.inst];
    if (s.status != kRunning) return;
    const Lanes L = lanes_of(tl);
    const int64_t iter = *d.iter_base + t;
    const double eta = s.eta, omega = s.omega;
    const double kavg = (double)(iter - s.restart_iter + 1);
    const int mode = EXACT ? kModeFp64 : s.mode;
    const bool f32 = mode == kModeFp32;
    // Halpern scheme: x_avg <- T(z)_x = x+, x <- w x_bar + (1 - w) x0 with
    // x_bar = 2 x+ - x (the reflection), w = k/(k+1), k = kavg.
    const bool halp = d.halpern != 0;
    const double w = kavg / (kavg + 1.0);
    const double* xsrc = halp ? d.xlast : d.xavg;
    double part = 0.0;
    for (int base = tl.r0; base < tl.r1; base += L.step) {
        const int j = base + (L.lane >> L.lg);
        const bool own = j < tl.r1;
        int p0 = 0, p1 = 0;
        double c = 0, cn = 1, lo = 0, hi = 0, xo = 0, xa = 0;
        if (own) {  // operands in flight while the column streams
            p0 = d.cp[j];
            p1 = d.cp[j + 1];
            c = d.c[j];
            cn = d.cnorm[j];
            lo = d.slo[j];
            hi = d.shi[j];
            xo = d.x[j];
            xa = xsrc[j];  // running average, or the Halpern anchor
        }
        const double aty = f32 ? (double)row_dot<float>(p0, p1, L.gl, L.G, d.ri, d.tv32, d.y32)
                           : mode == kModePacked ? row_dot<double>(p0, p1, L.gl, L.G, d.rix, d.tv32, d.y)
                                               : row_dot<double>(p0, p1, L.gl, L.G, d.ri, d.tv64, d.y);
        if (own && L.gl == 0) {
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
        }
    }
    if (ADAPT && !EXACT) {
        part = wave_sum(part);
        if (L.lane == 0) d.part_dxn[tile] = part;
    }
}

// ---------------------------------------------------------------------------
// k_dual: A x_bar fused with the dual prox step and the dual running average
// (FP64 iterate; FP32 mode gathers the FP32 x_bar through the FP32 values of
// A and refreshes the FP32 copy of y that the next primal step gathers).
template <bool EXACT>
__global__ __launch_bounds__(kBlock) void k_dual(Dev d, int t) {
    const int tile = wave_tile(d.ntA);
    if (tile >= d.ntA) return;
    const Tile tl = d.tA[tile];
    const InstState& s = d.st[tl.inst];
    if (s.status != kRunning) return;
    const Lanes L = lanes_of(tl);
    const int64_t iter = *d.iter_base + t;
    const double eta = s.eta, omega = s.omega;
    const double kavg = (double)(iter - s.restart_iter + 1);
    const int mode = EXACT ? kModeFp64 : s.mode;
    const bool f32 = mode == kModeFp32;
    // Halpern scheme: y_avg <- T(z)_y = y+, y <- w (2 y+ - y) + (1 - w) y0.
    const bool halp = d.halpern != 0;
    const double w = kavg / (kavg + 1.0);
    const double* ysrc = halp ? d.ylast : d.yavg;
    for (int base = tl.r0; base < tl.r1; base += L.step) {
        const int i = base + (L.lane >> L.lg);
        const bool own = i < tl.r1;
        int p0 = 0, p1 = 0;
        double rn = 1, bl = 0, bh = 0, yo = 0, ya = 0;
        if (own) {
            p0 = d.rp[i];
            p1 = d.rp[i + 1];
            rn = d.rnorm[i];
            bl = d.blo[i];
            bh = d.bhi[i];
            yo = d.y[i];
            ya = ysrc[i];
        }
        const double ax = f32 ? (double)row_dot<float>(p0, p1, L.gl, L.G, d.ci, d.av32, d.xbar32)
                          : mode == kModePacked ? row_dot<double>(p0, p1, L.gl, L.G, d.cix, d.av32, d.xbar)
                                              : row_dot<double>(p0, p1, L.gl, L.G, d.ci, d.av64, d.xbar);
        if (own && L.gl == 0) {
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
        }
    }
}

// ---------------------------------------------------------------------------
// k_adx: A dx for the adaptive step-size estimate (pdlp.cpp:208-210).
template <bool EXACT>
__global__ __launch_bounds__(kBlock) void k_adx(Dev d, int) {
    const int tile = wave_tile(d.ntA);
    if (tile >= d.ntA) return;
    const Tile tl = d.tA[tile];
    const InstState& s = d.st[tl.inst];
    if (s.status != kRunning) return;
    const Lanes L = lanes_of(tl);
    const int mode = EXACT ? kModeFp64 : s.mode;
    const bool f32 = mode == kModeFp32;
    double part = 0.0;
    for (int base = tl.r0; base < tl.r1; base += L.step) {
        const int i = base + (L.lane >> L.lg);
        const bool own = i < tl.r1;
        int p0 = 0, p1 = 0;
        double rn = 1.0;
        if (own) {
            p0 = d.rp[i];
            p1 = d.rp[i + 1];
            rn = d.rnorm[i];
        }
        const double a = f32 ? (double)row_dot<float>(p0, p1, L.gl, L.G, d.ci, d.av32, d.dx32)
                         : mode == kModePacked ? row_dot<double>(p0, p1, L.gl, L.G, d.cix, d.av32, d.dx)
                                             : row_dot<double>(p0, p1, L.gl, L.G, d.ci, d.av64, d.dx);
        if (own && L.gl == 0) {
            if (EXACT)
                d.ax[i] = a;
            else
                part += (a * a) / rn;
        }
    }
    if (!EXACT) {
        part = wave_sum(part);
        if (L.lane == 0) d.part_adx[tile] = part;
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
            else if (target > 1.05 * s.eta && s.eta < s.eta_cap)
                s.eta = minv(s.eta_cap, s.eta * 1.05);
        }
    }
}

// One workgroup per instance: the tile partials are summed by the whole
// block (fixed assignment of partials to threads, four independent loads per
// thread, fixed-order block reduction -- deterministic); EXACT sums the
// vectors serially in the reference's order.
__device__ __forceinline__ double sum_strided(const double* p, int beg, int end, int stride) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int k = beg + (int)threadIdx.x;
    for (; k + 3 * kFinal < end; k += 4 * kFinal) {
        a0 += p[(size_t)k * stride];
        a1 += p[(size_t)(k + kFinal) * stride];
        a2 += p[(size_t)(k + 2 * kFinal) * stride];
        a3 += p[(size_t)(k + 3 * kFinal) * stride];
    }
    for (; k < end; k += kFinal) a0 += p[(size_t)k * stride];
    return (a0 + a1) + (a2 + a3);
}
__device__ __forceinline__ double max_strided(const double* p, int beg, int end, int stride) {
    double a0 = 0.0, a1 = 0.0;
    int k = beg + (int)threadIdx.x;
    for (; k + kFinal < end; k += 2 * kFinal) {
        a0 = maxv(a0, p[(size_t)k * stride]);
        a1 = maxv(a1, p[(size_t)(k + kFinal) * stride]);
    }
    for (; k < end; k += kFinal) a0 = maxv(a0, p[(size_t)k * stride]);
    return maxv(a0, a1);
}

template <bool EXACT>
__global__ __launch_bounds__(kFinal) void k_adapt_finalize(Dev d) {
    const int b = blockIdx.x;
    InstState& s = d.st[b];
    if (s.status != kRunning) return;  // uniform over the block
    __shared__ double red[kFinal / 64];
    double dxn = 0.0, adxn = 0.0;
    if (EXACT) {
        if (threadIdx.x != 0) return;
        for (int j = d.inst_col[b]; j < d.inst_col[b + 1]; ++j) dxn += d.cnorm[j] * d.dx[j] * d.dx[j];
        for (int i = d.inst_row[b]; i < d.inst_row[b + 1]; ++i) adxn += (d.ax[i] * d.ax[i]) / d.rnorm[i];
    } else {
        const int32_t* pb = d.blocks ? d.bT_begin : d.tT_begin;  // part_dxn slots: tiles of the primal step
        dxn = block_sum<kFinal / 64>(sum_strided(d.part_dxn, pb[b], pb[b + 1], 1), red);
        const int32_t* pa = d.blocks ? d.bA_begin : d.tA_begin;  // part_adx slots: tiles of A dx
        adxn = block_sum<kFinal / 64>(sum_strided(d.part_adx, pa[b], pa[b + 1], 1), red);
        if (threadIdx.x != 0) return;
    }
    adapt_decide(s, dxn, adxn);
}

// ---------------------------------------------------------------------------
// FP64 residual terms of a candidate iterate on the unscaled model
// (pdlp.cpp:76-131).  CAND 0: the running average (the reference's only
// candidate); 1: the current iterate (KKT restart scheme); 2: the average
// scored with the packed stream's values (the packed floor).  The iterate and
// the matrix are read in FP64 in all of them.
// Rows: A x_c, row violation, dual-objective row part, |y_c - y_last|^2.
template <bool EXACT, int CAND>
__global__ __launch_bounds__(kBlock) void k_resid_rows(Dev d) {
    constexpr bool CUR = CAND == 1;
    const int tile = wave_tile(d.ntA);
    if (tile >= d.ntA) return;
    const Tile tl = d.tA[tile];
    const InstState& s = d.st[tl.inst];
    if (s.status != kRunning) return;
    const Lanes L = lanes_of(tl);
    const double* x64 = CUR ? d.x : d.xavg;
    const double* y64 = CUR ? d.y : d.yavg;
    double viol = 0.0, dob = 0.0, mv = 0.0;
    for (int base = tl.r0; base < tl.r1; base += L.step) {
        const int i = base + (L.lane >> L.lg);
        const bool own = i < tl.r1;
        int p0 = 0, p1 = 0;
        double rs = 1, lo = 0, hi = 0, yc = 0, yl = 0;
        if (own) {
            p0 = d.rp[i];
            p1 = d.rp[i + 1];
            rs = d.row_scale[i];
            lo = d.orow_lo[i];
            hi = d.orow_hi[i];
            yc = y64[i];
            yl = d.ylast[i];
        }
        const double ax = row_dot<double, double, double, int32_t, CAND == 2>(p0, p1, L.gl, L.G, d.ci, d.av64, x64);
        if (own && L.gl == 0) {
            if (EXACT) {
                (CUR ? d.ax2 : d.ax)[i] = ax;
            } else {
                const double axu = ax / rs;
                const double proj = clampv(axu, lo, hi);
                const double r = fabs(axu - proj);
                const double sc = 1.0 + maxv(fabs(axu), fabs(proj) < 1e299 ? fabs(proj) : 0.0);
                viol = maxv(viol, r / sc);
                const double yu = yc * rs;
                if (yu > 0.0 && hi < 1e299)
                    dob += -(yu * hi);
                else if (yu < 0.0 && lo > -1e299)
                    dob += -(yu * lo);
                const double dy = yc - yl;
                mv += dy * dy;
            }
        }
    }
    if (!EXACT) {
        viol = wave_max(viol);
        dob = wave_sum(dob);
        mv = wave_sum(mv);
        if (L.lane == 0) {  // field-major: [field][tile], one block per candidate
            double* out = d.part_rows + CAND * 3 * d.ntA + tile;
            out[0] = viol;
            out[d.ntA] = dob;
            out[2 * d.ntA] = mv;
        }
    }
}

// Columns: A^T y_c, bound violation, dual residual, primal/dual objective
// parts, |x_c - x_last|^2.
template <bool EXACT, int CAND>
__global__ __launch_bounds__(kBlock) void k_resid_cols(Dev d) {
    constexpr bool CUR = CAND == 1;
    const int tile = wave_tile(d.ntT);
    if (tile >= d.ntT) return;
    const Tile tl = d.tT[tile];
    const InstState& s = d.st[tl.inst];
    if (s.status != kRunning) return;
    const Lanes L = lanes_of(tl);
    const double* x64 = CUR ? d.x : d.xavg;
    const double* y64 = CUR ? d.y : d.yavg;
    const double sense = d.sense[tl.inst];
    double viol = 0.0, dres = 0.0, pob = 0.0, dob = 0.0, mv = 0.0;
    for (int base = tl.r0; base < tl.r1; base += L.step) {
        const int j = base + (L.lane >> L.lg);
        const bool own = j < tl.r1;
        int p0 = 0, p1 = 0;
        double cs = 1, lo = 0, hi = 0, xc = 0, xl = 0, ob = 0;
        if (own) {
            p0 = d.cp[j];
            p1 = d.cp[j + 1];
            cs = d.col_scale[j];
            lo = d.ovar_lo[j];
            hi = d.ovar_hi[j];
            ob = d.obj[j];
            xc = x64[j];
            xl = d.xlast[j];
        }
        const double aty = row_dot<double, double, double, int32_t, CAND == 2>(p0, p1, L.gl, L.G, d.ri, d.tv64, y64);
        if (own && L.gl == 0) {
            if (EXACT) {
                (CUR ? d.aty2 : d.aty)[j] = aty;
            } else {
                const double xu = xc * cs;
                const double proj = clampv(xu, lo, hi);
                const double r = fabs(xu - proj);
                const double sc = 1.0 + maxv(fabs(xu), fabs(proj) < 1e299 ? fabs(proj) : 0.0);
                viol = maxv(viol, r / sc);
                const double g = sense * ob + aty / cs;
                const double xp = clampv(xu - g, lo, hi);
                const double df = xu - xp;
                dres += df * df;
                pob += ob * xu;
                if (g > 0.0 && lo > -1e299)
                    dob += g * lo;
                else if (g < 0.0 && hi < 1e299)
                    dob += g * hi;
                const double dxm = xc - xl;
                mv += dxm * dxm;
            }
        }
    }
    if (!EXACT) {
        viol = wave_max(viol);
        dres = wave_sum(dres);
        pob = wave_sum(pob);
        dob = wave_sum(dob);
        mv = wave_sum(mv);
        if (L.lane == 0) {  // field-major: [field][tile], one block per candidate
            double* out = d.part_cols + CAND * 5 * d.ntT + tile;
            out[0] = viol;
            out[d.ntT] = dres;
            out[2 * d.ntT] = pob;
            out[3 * d.ntT] = dob;
            out[4 * d.ntT] = mv;
        }
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

// Candidate residuals from the tile partials (fast order), reduced by the
// whole block; the result is valid in thread 0.
__device__ void partial_residuals(const Dev& d, int b, int cand, double c_scale, CandResid& out,
                                  double* red) {
    // One coalesced pass over the field-major tile partials of this
    // instance, every field accumulated in registers, one combined reduction.
    // slots per tile of the tiles the check ran on (block or wave tiles)
    const int ntA = d.blocks ? d.nbA : d.ntA, ntT = d.blocks ? d.nbT : d.ntT;
    const int32_t* ba = d.blocks ? d.bA_begin : d.tA_begin;
    const int32_t* bt = d.blocks ? d.bT_begin : d.tT_begin;
    const double* pr = d.part_rows + cand * 3 * ntA;
    const double* pc = d.part_cols + cand * 5 * ntT;
    const int a0 = ba[b], a1 = ba[b + 1], t0 = bt[b], t1 = bt[b + 1];
    constexpr int NW = kFinal / 64;
    double mp = 0.0, s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // dr, ym, dsq, po, dc, xm
    for (int i = a0 + (int)threadIdx.x; i < a1; i += kFinal) {
        mp = maxv(mp, pr[i]);
        s[0] += pr[ntA + i];
        s[1] += pr[2 * ntA + i];
    }
    for (int i = t0 + (int)threadIdx.x; i < t1; i += kFinal) {
        mp = maxv(mp, pc[i]);
        s[2] += pc[ntT + i];
        s[3] += pc[2 * ntT + i];
        s[4] += pc[3 * ntT + i];
        s[5] += pc[4 * ntT + i];
    }
    block_max_sums<NW, 6>(mp, s, red);
    const double dr = s[0], ym = s[1], dsq = s[2], po = s[3], dc = s[4], xm = s[5];
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
    if (s.mode != kModeFp64) {  // finished without leaving the reduced-precision phase
        if (s.mode == kModeFp32)
            s.fp32_iters = it;
        else
            s.packed_end = it;
        atomicSub(&d.counters[1], 1);
    }
}

// Termination and restart decision of one instance (pdlp.cpp:229-260; the
// KKT scheme as documented for MC_SCHEME_KKT).  FINAL: the iteration limit
// was reached -- record the better candidate's residuals and stop.
template <bool EXACT, bool FINAL>
__global__ __launch_bounds__(kFinal) void k_check_finalize(Dev d, IterOpts o, int t) {
    const int b = blockIdx.x;  // one workgroup per instance
    InstState& s = d.st[b];
    __shared__ double red[(kFinal / 64) * 7];
    __shared__ int running;
    if (threadIdx.x == 0) running = s.status == kRunning;
    __syncthreads();
    if (threadIdx.x == 0) {
        s.restart_flag = 0;
        s.switch_flag = 0;
    }
    if (!running) return;
    const int64_t iter_after = *d.iter_base + t + 1;
    CandResid r, rc, rpk;
    if (!EXACT) {
        partial_residuals(d, b, 0, s.c_scale, r, red);
        if (o.kkt) partial_residuals(d, b, 1, s.c_scale, rc, red);
        if (o.packed_floor) partial_residuals(d, b, 2, s.c_scale, rpk, red);
    }
    if (threadIdx.x != 0) return;
    if (EXACT) {
        serial_residuals(d, b, d.xavg, d.yavg, d.ax, d.aty, s.c_scale, r);
        if (o.kkt) serial_residuals(d, b, d.x, d.y, d.ax2, d.aty2, s.c_scale, rc);
    }
    // the packed floor: how far the packed values move the average's scores
    const double pk_floor = (!EXACT && o.packed_floor)
                                ? maxv(fabs(r.primal - rpk.primal), maxv(fabs(r.dual - rpk.dual), fabs(r.gap - rpk.gap)))
                                : 0.0;
    int use_cur = 0;
    if (o.kkt) {
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
                double w = s.omega;
                bool upd = false;
                if (dX > 1e-10 && dY > 1e-10) {
                    const double th = o.smoothing, ratio = dY / dX;
                    if (o.halpern) {  // ratio^(127/128) omega^(1/128) in square roots
                        double f = s.omega / ratio;
                        for (int q = 0; q < 7; ++q) f = sqrt(f);
                        w = ratio * f;
                    } else {
                        w = (th == 0.5) ? sqrt(s.omega * ratio) : pow(ratio, th) * pow(s.omega, 1.0 - th);
                    }
                    upd = true;
                }
                if (o.halpern) {  // omega_move^(3/4) omega_res^(1/4), in square roots
                    double wres = s.omega;
                    if (r.primal > 1e-12 && r.dual > 1e-12)
                        wres = s.omega * sqrt(clampv(sqrt(r.primal / r.dual), 0.05, 20.0));
                    w = sqrt(w * sqrt(w * wres));
                    upd = true;
                    // no move against a residual imbalance of 100x or more
                    if (r.primal > 1e-12 && r.dual > 1e-12) {
                        if (r.primal > 100.0 * r.dual && w < s.omega) w = s.omega;
                        if (r.dual > 100.0 * r.primal && w > s.omega) w = s.omega;
                    }
                    // at most a fall by omega_max_drop (a power of two: exact)
                    if (o.omega_max_drop > 0) w = maxv(w, s.omega / (double)o.omega_max_drop);
                }
                if (upd) s.omega = clampv(w, 1e-6, 1e6);
            }
            s.prev_cand = 1e300;
        }
        s.restart_flag = 1;
        s.restart_iter = iter_after;
        s.last_restart_score = score;
    }
    if (s.mode != kModeFp64) {
        // FP32 products stop paying once the score is near tolerance or has
        // not set a new best for stall_checks checks (their rounding floor,
        // 1e-2 to 1e-4 on planted LPs); the instance moves to the packed
        // stream (about 31 significant bits of A, FP64 sums) or to FP64.  The
        // packed stream's floor is measured at every check (pk_floor); it is
        // usually far below the scores where restarted PDHG has plateaus of
        // tens of thousands of iterations without a new best, so the packed
        // stream leaves at the polish threshold, within kFloorMargin of its
        // floor, below kPackedFloor, or after 100x as many checks without a
        // new best.
        if (score < s.best_score) {
            s.best_score = score;
            s.stall = 0;
        } else {
            s.stall += 1;
        }
        const bool fp32 = s.mode == kModeFp32;
        const bool leave = score <= o.polish_threshold ||
                           (fp32 ? s.stall >= o.stall_checks
                                 : (s.stall >= 100 * o.stall_checks || score <= kPackedFloor ||
                                    score <= kFloorMargin * pk_floor));
        if (leave) {
            s.switch_flag = 1;
            if (fp32) s.fp32_iters = iter_after;
            if (fp32 && o.packed_next && score > o.polish_threshold) {
                s.mode = kModePacked;
                s.stall = 0;
                s.best_score = score;
            } else {
                if (!fp32) s.packed_end = iter_after;
                s.mode = kModeFp64;
                atomicSub(&d.counters[1], 1);
            }
        }
    }
}

// Restart to the candidate (pdlp.cpp:257; KKT/Halpern schemes: the current
// iterate or the average slot, which also becomes the new x_last/y_last).
// The iterate is FP64 in both precision modes, so a switch to FP64 needs no
// conversion.  Grid: column tiles, then row tiles.
__global__ __launch_bounds__(kBlock) void k_post_check(Dev d, int kkt) {
    const int g = blockIdx.x * kWaves + (threadIdx.x >> 6);  // one wave per tile
    if (g >= d.ntT + d.ntA) return;
    const bool cols = g < d.ntT;
    const Tile tl = cols ? d.tT[g] : d.tA[g - d.ntT];
    const InstState& s = d.st[tl.inst];
    if (!s.restart_flag) return;
    const bool to_avg = !s.cand_cur;
    // FP32 gathers read y32: refresh it after the restart moved y.  (x_bar32
    // is rewritten by the next primal step before it is gathered.)
    const bool refresh32 = !cols && s.mode == kModeFp32;
    double* v64 = cols ? d.x : d.y;
    const double* a64 = cols ? d.xavg : d.yavg;
    double* l64 = cols ? d.xlast : d.ylast;
    for (int k = tl.r0 + (int)(threadIdx.x & 63); k < tl.r1; k += 64) {
        if (to_avg) v64[k] = a64[k];
        if (kkt) l64[k] = v64[k];
        if (refresh32) d.y32[k] = (float)v64[k];
    }
}

// ---------------------------------------------------------------------------
// Infeasibility detection (extension; no reference counterpart: the
// reference declares infeasible_or_unbounded but never reports it).  A
// diverging PDHG iterate moves along a certificate ray (Applegate et al.
// 2021, sec. 5).  At a check, the candidate's movement since the last
// restart, (dx, dy), is tested in the unscaled space as
//  * a dual ray certificate of primal infeasibility: the homogeneous dual
//    objective D(dy) = sum_i -support_[lo_i,hi_i](dy_i) + sum_j (g_j > 0 ?
//    g_j lo_j : g_j hi_j), g = A^T dy, is positive beyond rounding (more than
//    1e-9 of the sum of its terms' magnitudes: for a feasible LP D <= 0 by
//    weak duality), and the parts that need an infinite bound (|dy_i| or |g_j|
//    toward it) sum to at most tol x D;
//  * a primal ray certificate of dual infeasibility: c^T dx < 0 beyond
//    rounding, and the recession violations of A dx against the row bounds
//    and of dx against the variable bounds are at most tol x |c^T dx|.
// Either stops the instance with status infeasible_or_unbounded.  A false
// positive needs an optimal dual (primal) of norm above 1/tol.
__global__ void k_ray_prep(Dev d) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < d.n) {
        const InstState& s = d.st[d.col_inst[k]];
        if (s.status != kRunning) return;
        d.ray_x[k] = (s.cand_cur ? d.x[k] : d.xavg[k]) - d.xlast[k];
    } else if (k < d.n + d.m) {
        const int i = k - d.n;
        const InstState& s = d.st[d.row_inst[i]];
        if (s.status != kRunning) return;
        d.ray_y[i] = (s.cand_cur ? d.y[i] : d.yavg[i]) - d.ylast[i];
    }
}

__global__ __launch_bounds__(kBlock) void k_ray_rows(Dev d) {
    const int tile = wave_tile(d.ntA);
    if (tile >= d.ntA) return;
    const Tile tl = d.tA[tile];
    if (d.st[tl.inst].status != kRunning) return;
    const Lanes L = lanes_of(tl);
    double vp = 0.0, dob = 0.0, dab = 0.0, vd = 0.0, ny = 0.0;
    for (int base = tl.r0; base < tl.r1; base += L.step) {
        const int i = base + (L.lane >> L.lg);
        const bool own = i < tl.r1;
        int p0 = 0, p1 = 0;
        if (own) {
            p0 = d.rp[i];
            p1 = d.rp[i + 1];
        }
        const double a = row_dot<double>(p0, p1, L.gl, L.G, d.ci, d.av64, d.ray_x);
        if (own && L.gl == 0) {
            const double rs = d.row_scale[i];
            const double lo = d.orow_lo[i], hi = d.orow_hi[i];
            const double u = a / rs;  // (A dx)_i, unscaled
            if (lo > -INFINITY && u < 0.0) vp = maxv(vp, -u);
            if (hi < INFINITY && u > 0.0) vp = maxv(vp, u);
            const double yu = d.ray_y[i] * rs;
            double term = 0.0;
            if (yu > 0.0) {
                if (hi < INFINITY) term = -(yu * hi); else vd += yu;
            } else if (yu < 0.0) {
                if (lo > -INFINITY) term = -(yu * lo); else vd -= yu;
            }
            dob += term;
            dab += fabs(term);
            ny = maxv(ny, fabs(yu));
        }
    }
    vp = wave_max(vp);
    dob = wave_sum(dob);
    dab = wave_sum(dab);
    vd = wave_sum(vd);
    ny = wave_max(ny);
    if (L.lane == 0) {
        double* out = d.part_ray_rows + 5 * tile;
        out[0] = vp;
        out[1] = dob;
        out[2] = dab;
        out[3] = vd;
        out[4] = ny;
    }
}

__global__ __launch_bounds__(kBlock) void k_ray_cols(Dev d) {
    const int tile = wave_tile(d.ntT);
    if (tile >= d.ntT) return;
    const Tile tl = d.tT[tile];
    if (d.st[tl.inst].status != kRunning) return;
    const Lanes L = lanes_of(tl);
    const double sense = d.sense[tl.inst];
    double vp = 0.0, cdx = 0.0, cab = 0.0, dob = 0.0, dab = 0.0, vd = 0.0, nx = 0.0;
    for (int base = tl.r0; base < tl.r1; base += L.step) {
        const int j = base + (L.lane >> L.lg);
        const bool own = j < tl.r1;
        int p0 = 0, p1 = 0;
        if (own) {
            p0 = d.cp[j];
            p1 = d.cp[j + 1];
        }
        const double a = row_dot<double>(p0, p1, L.gl, L.G, d.ri, d.tv64, d.ray_y);
        if (own && L.gl == 0) {
            const double cs = d.col_scale[j];
            const double lo = d.ovar_lo[j], hi = d.ovar_hi[j];
            const double g = a / cs;  // (A^T dy)_j, unscaled
            double term = 0.0;
            if (g > 0.0) {
                if (lo > -INFINITY) term = g * lo; else vd += g;
            } else if (g < 0.0) {
                if (hi < INFINITY) term = g * hi; else vd -= g;
            }
            dob += term;
            dab += fabs(term);
            const double xu = d.ray_x[j] * cs;
            if (lo > -INFINITY && xu < 0.0) vp = maxv(vp, -xu);
            if (hi < INFINITY && xu > 0.0) vp = maxv(vp, xu);
            const double cterm = sense * d.obj[j] * xu;
            cdx += cterm;
            cab += fabs(cterm);
            nx = maxv(nx, fabs(xu));
        }
    }
    vp = wave_max(vp);
    cdx = wave_sum(cdx);
    cab = wave_sum(cab);
    dob = wave_sum(dob);
    dab = wave_sum(dab);
    vd = wave_sum(vd);
    nx = wave_max(nx);
    if (L.lane == 0) {
        double* out = d.part_ray_cols + 7 * tile;
        out[0] = vp;
        out[1] = cdx;
        out[2] = cab;
        out[3] = dob;
        out[4] = dab;
        out[5] = vd;
        out[6] = nx;
    }
}

__global__ __launch_bounds__(kFinal) void k_ray_finalize(Dev d, IterOpts o, int t) {
    const int b = blockIdx.x;
    InstState& s = d.st[b];
    __shared__ double red[kFinal / 64];
    __shared__ int running;
    if (threadIdx.x == 0) running = s.status == kRunning;
    __syncthreads();
    if (!running) return;
    const double* pr = d.part_ray_rows;
    const double* pc = d.part_ray_cols;
    const int a0 = d.tA_begin[b], a1 = d.tA_begin[b + 1], t0 = d.tT_begin[b], t1 = d.tT_begin[b + 1];
    constexpr int NW = kFinal / 64;
    const double vp = block_max<NW>(maxv(max_strided(pr, a0, a1, 5), max_strided(pc, t0, t1, 7)), red);
    const double cdx = block_sum<NW>(sum_strided(pc + 1, t0, t1, 7), red);
    const double cab = block_sum<NW>(sum_strided(pc + 2, t0, t1, 7), red);
    const double dob = block_sum<NW>(sum_strided(pr + 1, a0, a1, 5), red) +
                       block_sum<NW>(sum_strided(pc + 3, t0, t1, 7), red);
    const double dab = block_sum<NW>(sum_strided(pr + 2, a0, a1, 5), red) +
                       block_sum<NW>(sum_strided(pc + 4, t0, t1, 7), red);
    const double vd = block_sum<NW>(sum_strided(pr + 3, a0, a1, 5), red) +
                      block_sum<NW>(sum_strided(pc + 5, t0, t1, 7), red);
    const double ny = block_max<NW>(max_strided(pr + 4, a0, a1, 5), red);
    const double nx = block_max<NW>(max_strided(pc + 6, t0, t1, 7), red);
    if (threadIdx.x != 0) return;
    const bool primal_infeasible = ny > 0.0 && dob > 1e-9 * dab && vd <= o.infeas_tol * dob;
    const bool dual_infeasible = nx > 0.0 && cdx < -1e-9 * cab && vp <= o.infeas_tol * -cdx;
    if (primal_infeasible || dual_infeasible) stop_instance(d, s, kInfeasible, *d.iter_base + t + 1);
}

__global__ void k_advance(Dev d, int len) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *d.iter_base += len;
}

// ---------------------------------------------------------------------------
// Setup: device Ruiz equilibration (ruiz_scaling.cpp:152-247) on the same
// wave tasks.  All maxima are order-independent and every scaling is a
// product, so the result equals the serial host pass bit for bit.
// Nonnegative doubles order like their bit patterns, so atomicMax on the
// 64-bit pattern is an exact max.
__device__ __forceinline__ void atomic_max_nonneg(double* addr, double v) {
    atomicMax((unsigned long long*)addr, (unsigned long long)__double_as_longlong(v));
}

// Group max of |val| over one row; valid in every lane of the group.
__device__ __forceinline__ double row_absmax(int p0, int p1, int gl, int G, const double* val) {
    double m = 0.0;
    for (int q = p0 + gl; q < p1; q += G) m = maxv(m, fabs(val[q]));
    for (int off = G >> 1; off >= 1; off >>= 1) m = maxv(m, __shfl_xor(m, off, 64));
    return m;
}

// ROWS = true: rows of A (row max -> rtmp, error -> ruiz_err_r);
// false: rows of A^T = columns of A (-> ctmp, ruiz_err_c).
template <bool ROWS>
__global__ __launch_bounds__(kBlock) void k_ruiz_max(Dev d) {
    const int nt = ROWS ? d.ntA : d.ntT;
    const int tile = wave_tile(nt);
    if (tile >= nt) return;
    const Tile tl = (ROWS ? d.tA : d.tT)[tile];
    InstState& s = d.st[tl.inst];
    if (s.ruiz_done) return;
    const Lanes L = lanes_of(tl);
    const int32_t* ptr = ROWS ? d.rp : d.cp;
    const double* val = ROWS ? d.av64 : d.tv64;
    double* out = ROWS ? d.rtmp : d.ctmp;
    double err = 0.0;
    for (int base = tl.r0; base < tl.r1; base += L.step) {
        const int r = base + (L.lane >> L.lg);
        const bool own = r < tl.r1;
        const double m = row_absmax(own ? ptr[r] : 0, own ? ptr[r + 1] : 0, L.gl, L.G, val);
        if (own && L.gl == 0) {
            const double mm = (m == 0.0) ? 1.0 : m;
            out[r] = mm;
            err = maxv(err, fabs(1.0 - mm));
        }
    }
    err = wave_max(err);
    if (L.lane == 0) atomic_max_nonneg(ROWS ? &s.ruiz_err_r : &s.ruiz_err_c, err);
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
// A <- D_r A D_c on both stored orientations (value *= dr[i] * dc[j]).
// Tiles of A^T first, then tiles of A; one wave per tile.  ALL: every
// instance, also those whose Ruiz passes have ended (the row weights).
template <bool ALL>
__global__ __launch_bounds__(kBlock) void k_ruiz_apply(Dev d) {
    const int g = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (g >= d.ntT + d.ntA) return;
    const bool cols = g < d.ntT;
    const Tile tl = cols ? d.tT[g] : d.tA[g - d.ntT];
    if (!ALL && d.st[tl.inst].ruiz_done) return;
    const Lanes L = lanes_of(tl);
    for (int base = tl.r0; base < tl.r1; base += L.step) {
        const int r = base + (L.lane >> L.lg);
        if (r >= tl.r1) continue;
        if (cols) {
            const double dc = d.ctmp[r];
            for (int p = d.cp[r] + L.gl; p < d.cp[r + 1]; p += L.G) d.tv64[p] *= d.rtmp[d.ri[p]] * dc;
        } else {
            const double dr = d.rtmp[r];
            for (int p = d.rp[r] + L.gl; p < d.rp[r + 1]; p += L.G) d.av64[p] *= dr * d.ctmp[d.ci[p]];
        }
    }
}

// Row weights after the Ruiz passes: w_i = (1 + beta_i)^(-k/8) (row_weight),
// beta_i the largest magnitude of row i's finite original bounds (0 for a
// free row).  rtmp <- w, ctmp <- 1, row_scale *= w; k_ruiz_apply<true> then
// scales the rows of both stored orientations, and k_setup_rows scales the
// original bounds by the final row_scale.  The relative primal residual
// divides row i's violation by 1 + |a_i x|, about 1 + beta_i at an active
// bound, so rows with small bounds must be met more tightly; with the L1-norm
// step sizes, weighting row i by w_i multiplies its dual step by w_i.
__global__ void k_row_weights(Dev d, int k8) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < d.m) {
        const double lo = d.orow_lo[k], hi = d.orow_hi[k];
        double beta = 0.0;
        if (!isinf(lo)) beta = maxv(beta, fabs(lo));
        if (!isinf(hi)) beta = maxv(beta, fabs(hi));
        const double w = row_weight(beta, k8);
        d.rtmp[k] = w;
        d.row_scale[k] *= w;
    } else if (k < d.m + d.n) {
        d.ctmp[k - d.m] = 1.0;
    }
}

// Packed stream word of one nonzero (see packed_value): index | k << 24 with
// k = round((a - fl32(a)) / (ulp(fl32(a)) / 256)) clamped to [-127, 127].
__device__ __forceinline__ uint32_t pack_word(int32_t index, double a) {
    const int k = pack_correction(a, (float)a);
    return (uint32_t)index | ((uint32_t)(k & 0xFF) << 24);
}

// Scaled bounds/objective (ruiz_scaling.cpp:221-244), sentinels and norms
// (pdlp.cpp:153-176) and FP32 copies; once per loaded problem.  The L1 norms
// are summed serially by the group's first lane in the reference's order.
// Original bounds stay as IEEE +-inf for infinite kinds: every test the
// residuals make on them (clamp, |proj| < 1e299, hi < 1e299) reads the same
// on inf as on the reference's 1e300 sentinel.
__global__ __launch_bounds__(kBlock) void k_setup_rows(Dev d) {
    const int tile = wave_tile(d.ntA);
    if (tile >= d.ntA) return;
    const Tile tl = d.tA[tile];
    const Lanes L = lanes_of(tl);
    double bn_max = 0.0;
    for (int base = tl.r0; base < tl.r1; base += L.step) {
        const int i = base + (L.lane >> L.lg);
        if (i >= tl.r1) continue;
        const int p0 = d.rp[i], p1 = d.rp[i + 1];
        for (int p = p0 + L.gl; p < p1; p += L.G) {
            d.av32[p] = (float)d.av64[p];
            if (d.cix) d.cix[p] = pack_word(d.ci[p], d.av64[p]);
        }
        if (L.gl != 0) continue;
        const double rs = d.row_scale[i];
        const double lo = d.orow_lo[i], hi = d.orow_hi[i];
        const bool flo = !isinf(lo), fhi = !isinf(hi);
        const double bl = flo ? lo * rs : -1e300;
        const double bh = fhi ? hi * rs : 1e300;
        d.blo[i] = bl;
        d.bhi[i] = bh;
        if (fabs(bl) < 1e299) bn_max = maxv(bn_max, fabs(bl));
        if (fabs(bh) < 1e299) bn_max = maxv(bn_max, fabs(bh));
        double nrm = 0.0;
        for (int p = p0; p < p1; ++p) nrm += fabs(d.av64[p]);
        if (nrm < 1e-12) nrm = 1.0;
        d.rnorm[i] = nrm;
    }
    bn_max = wave_max(bn_max);
    if (L.lane == 0) atomic_max_nonneg(&d.st[tl.inst].b_norm_inf, bn_max);
}

__global__ __launch_bounds__(kBlock) void k_setup_cols(Dev d) {
    const int tile = wave_tile(d.ntT);
    if (tile >= d.ntT) return;
    const Tile tl = d.tT[tile];
    const Lanes L = lanes_of(tl);
    const double sense = d.sense[tl.inst];
    double cmax = 0.0, cs_max = 0.0;
    for (int base = tl.r0; base < tl.r1; base += L.step) {
        const int j = base + (L.lane >> L.lg);
        if (j >= tl.r1) continue;
        const int p0 = d.cp[j], p1 = d.cp[j + 1];
        for (int p = p0 + L.gl; p < p1; p += L.G) {
            d.tv32[p] = (float)d.tv64[p];
            if (d.rix) d.rix[p] = pack_word(d.ri[p], d.tv64[p]);
        }
        if (L.gl != 0) continue;
        const double cs = d.col_scale[j];
        const double inv = 1.0 / cs;
        const double lo = d.ovar_lo[j], hi = d.ovar_hi[j];
        const bool flo = !isinf(lo), fhi = !isinf(hi);
        const double sl = flo ? lo * inv : -1e300;
        const double sh = fhi ? hi * inv : 1e300;
        d.slo[j] = sl;
        d.shi[j] = sh;
        const double cj = sense * (d.obj[j] * cs);
        d.c[j] = cj;
        cmax = maxv(cmax, fabs(cj));
        cs_max = maxv(cs_max, fabs(sense * d.obj[j]));
        double nrm = 0.0;
        for (int p = p0; p < p1; ++p) nrm += fabs(d.tv64[p]);
        if (nrm < 1e-12) nrm = 1.0;
        d.cnorm[j] = nrm;
    }
    cmax = wave_max(cmax);
    cs_max = wave_max(cs_max);
    if (L.lane == 0) {
        atomic_max_nonneg(&d.st[tl.inst].c_norm_inf, cmax);
        atomic_max_nonneg(&d.st[tl.inst].c_scale, cs_max);
    }
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
        d.xbar32[k] = (float)x0;
        d.xlast[k] = x0;
    } else if (k < d.n + d.m) {
        const int i = k - d.n;
        const double y0 = warm_y ? warm_y[i] / d.row_scale[i] : 0.0;
        d.y[i] = y0;
        d.yavg[i] = y0;
        d.y32[i] = (float)y0;
        d.ylast[i] = y0;
    }
}

__global__ void k_setup_state(Dev d, double initial_weight, double step_reduction, int start_mode) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.B) return;
    InstState& s = d.st[b];
    s.omega = (initial_weight > 0.0) ? initial_weight
                                     : clampv(sqrt(s.c_norm_inf / s.b_norm_inf), 0.01, 100.0);
    s.eta = clampv(step_reduction, 0.1, 0.99);
    s.eta_cap = 0.99;  // launch_step_cap may raise it (Halpern scheme)
    s.last_restart_score = 1e300;
    s.best_score = 1e300;
    s.restart_iter = 0;
    s.done_iter = 0;
    s.fp32_iters = 0;
    s.packed_end = 0;
    s.status = kRunning;
    s.mode = start_mode;
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
        xo[k] = (s.answer_cur ? d.x[k] : d.xavg[k]) * d.col_scale[k];
    } else if (k < d.n + d.m) {
        const int i = k - d.n;
        const InstState& s = d.st[d.row_inst[i]];
        yo[i] = (s.answer_cur ? d.y[i] : d.yavg[i]) * d.row_scale[i];
    }
}

__global__ void k_fill(double* p, double v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <typename T, typename TA = T, typename IX = int32_t>
__global__ __launch_bounds__(kBlock) void k_spmv_plain(const Tile* tiles, int ntiles, const int32_t* ptr,
                                                       const IX* idx, const TA* val, const T* vec, T* out) {
    const int tile = wave_tile(ntiles);
    if (tile >= ntiles) return;
    const Tile tl = tiles[tile];
    const Lanes L = lanes_of(tl);
    for (int base = tl.r0; base < tl.r1; base += L.step) {
        const int r = base + (L.lane >> L.lg);
        const bool own = r < tl.r1;
        const T v = row_dot<T>(own ? ptr[r] : 0, own ? ptr[r + 1] : 0, L.gl, L.G, idx, val, vec);
        if (own && L.gl == 0) out[r] = v;
    }
}

// ---------------------------------------------------------------------------
// Synchronization probe (measurement only, mcp_time_sync): what a persistent
// kernel would pay for a grid-wide barrier between the phases of an
// iteration, against the boundary between two dependent launches in a
// captured graph.  bar[0]: arrivals, bar[1]: generation, bar[2]: set when a
// wait ran out of spins (the barrier is then abandoned, never hung), bar[3]:
// reads after a barrier that saw a stale value.
constexpr unsigned kBarrierSpinLimit = 1u << 20;

__device__ __forceinline__ void grid_barrier(unsigned* bar, unsigned nblocks) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned g = __hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned a = __hip_atomic_fetch_add(bar, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (a == nblocks - 1) {
            __hip_atomic_store(bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(bar + 1, g + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            unsigned spins = 0;
            while (__hip_atomic_load(bar + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g &&
                   __hip_atomic_load(bar + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > kBarrierSpinLimit) {
                    __hip_atomic_store(bar + 2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// `rounds` rounds of: write this workgroup's slot of `words` doubles,
// barrier, check the next workgroup's slot, barrier.  Cooperative launch
// only (every workgroup resident).
__global__ __launch_bounds__(kBlock) void k_barrier_probe(unsigned* bar, double* buf, int words, int rounds) {
    const unsigned nb = gridDim.x;
    double* mine = buf + (size_t)blockIdx.x * words;
    const double* next = buf + (size_t)((blockIdx.x + 1) % nb) * words;
    for (int r = 0; r < rounds; ++r) {
        for (int k = threadIdx.x; k < words; k += blockDim.x) mine[k] = (double)(r + 1);
        grid_barrier(bar, nb);
        int stale = 0;
        for (int k = threadIdx.x; k < words; k += blockDim.x) stale += next[k] != (double)(r + 1);
        if (stale) atomicAdd(bar + 3, 1u);
        grid_barrier(bar, nb);
    }
}

// One round's write of the same slots as a kernel of its own (a graph of
// `rounds` dependent launches is the barrier-free alternative).
__global__ __launch_bounds__(kBlock) void k_barrier_step(double* buf, int words, int round) {
    double* mine = buf + (size_t)blockIdx.x * words;
    for (int k = threadIdx.x; k < words; k += blockDim.x) mine[k] = (double)(round + 1);
}

int barrier_probe_capacity() {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)k_barrier_probe, kBlock, 0) != hipSuccess)
        return 0;
    return cus * per;
}

hipError_t launch_barrier_probe(unsigned* bar, double* buf, int words, int rounds, int workgroups, hipStream_t s) {
    void* args[] = {&bar, &buf, &words, &rounds};
    return hipLaunchCooperativeKernel((const void*)k_barrier_probe, dim3(workgroups), dim3(kBlock), args, 0, s);
}

void launch_barrier_step(double* buf, int words, int round, int workgroups, hipStream_t s) {
    hipLaunchKernelGGL(k_barrier_step, dim3(workgroups), dim3(kBlock), 0, s, buf, words, round);
}

// ---------------------------------------------------------------------------
// Step cap of the Halpern scheme (the oracle's or_step_cap, per instance):
// v_j = 1 + (j mod 7)/1000 (j local to the instance), S = |v|^2; per iteration
// t = (v / sqrt(S)) / sqrt(cnorm), u = (A t) / rnorm, v = (A^T u) / sqrt(cnorm),
// S = |v|^2; L = sqrt(sqrt(S)), cap = 0.95 / L.  v lives in ctmp, u in rtmp,
// t in aty2; pw[b] = S, pw[B + b] = 1 once an S that starts an iteration was
// not > 0 (the oracle stops there and keeps 0.99).
__global__ void k_pw_init(Dev d) {  // grid covers max(n, B) threads
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < d.B) d.pw[d.B + j] = 0.0;
    if (j >= d.n) return;
    const int jl = j - d.inst_col[d.col_inst[j]];
    d.ctmp[j] = 1.0 + (double)(jl % 7) / 1000.0;
}

// S_b = sum of v_j^2 over the instance's columns; `check`: S_b starts an
// iteration, so a non-positive (or NaN) S_b marks the estimate failed.
template <bool EXACT>
__global__ __launch_bounds__(kFinal) void k_pw_sum(Dev d, int check) {
    const int b = blockIdx.x;
    const int j0 = d.inst_col[b], j1 = d.inst_col[b + 1];
    double S = 0.0;
    if (EXACT) {
        if (threadIdx.x != 0) return;
        for (int j = j0; j < j1; ++j) S += d.ctmp[j] * d.ctmp[j];
    } else {
        __shared__ double red[kFinal / 64];
        double a0 = 0.0, a1 = 0.0;
        int j = j0 + (int)threadIdx.x;
        for (; j + kFinal < j1; j += 2 * kFinal) {
            a0 += d.ctmp[j] * d.ctmp[j];
            a1 += d.ctmp[j + kFinal] * d.ctmp[j + kFinal];
        }
        for (; j < j1; j += kFinal) a0 += d.ctmp[j] * d.ctmp[j];
        S = block_sum<kFinal / 64>(a0 + a1, red);
        if (threadIdx.x != 0) return;
    }
    d.pw[b] = S;
    if (check && !(S > 0.0)) d.pw[d.B + b] = 1.0;
}

__global__ void k_pw_cols_in(Dev d) {  // t = (v / sqrt(S)) / sqrt(cnorm)
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.n) return;
    const double nv = sqrt(d.pw[d.col_inst[j]]);
    d.aty2[j] = (d.ctmp[j] / nv) / sqrt(d.cnorm[j]);
}

__global__ void k_pw_rows(Dev d) {  // u = u / rnorm
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d.m) d.rtmp[i] = d.rtmp[i] / d.rnorm[i];
}

__global__ void k_pw_cols_out(Dev d) {  // v = v / sqrt(cnorm)
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < d.n) d.ctmp[j] = d.ctmp[j] / sqrt(d.cnorm[j]);
}

__global__ void k_pw_finish(Dev d) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.B) return;
    const double L = sqrt(sqrt(d.pw[b]));
    d.st[b].eta_cap = (d.pw[d.B + b] == 0.0 && L > 0.0 && L < 1e300) ? 0.95 / L : 0.99;
}

// ---------------------------------------------------------------------------
// Host launch helpers.
static inline int tile_grid(int ntiles) {
    const int nblk = (ntiles + kWaves - 1) / kWaves;
    return kXcd * ((nblk + kXcd - 1) / kXcd);
}
static inline int elem_grid(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

void launch_primal(const Dev& d, int t, bool adapt, bool exact, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    if (!exact && d.blocks) {
        launch_primal_blk(d, t, adapt, s, e0, e1);
        return;
    }
    if (d.ntT == 0) return;
    const dim3 g(tile_grid(d.ntT)), b(kBlock);
    if (exact) {
        if (adapt)
            MCP_LAUNCH_TIMED((k_primal<true, true>), g, b, s, e0, e1, d, t);
        else
            MCP_LAUNCH_TIMED((k_primal<false, true>), g, b, s, e0, e1, d, t);
    } else {
        if (adapt)
            MCP_LAUNCH_TIMED((k_primal<true, false>), g, b, s, e0, e1, d, t);
        else
            MCP_LAUNCH_TIMED((k_primal<false, false>), g, b, s, e0, e1, d, t);
    }
}

void launch_dual(const Dev& d, int t, bool exact, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    if (!exact && d.blocks) {
        launch_dual_blk(d, t, s, e0, e1);
        return;
    }
    if (d.ntA == 0) return;
    const dim3 g(tile_grid(d.ntA)), b(kBlock);
    if (exact)
        MCP_LAUNCH_TIMED((k_dual<true>), g, b, s, e0, e1, d, t);
    else
        MCP_LAUNCH_TIMED((k_dual<false>), g, b, s, e0, e1, d, t);
}

void launch_adx(const Dev& d, int t, bool exact, hipStream_t s) {
    if (!exact && d.blocks) {
        launch_adx_blk(d, t, s);
        return;
    }
    if (d.ntA == 0) return;
    const dim3 g(tile_grid(d.ntA)), b(kBlock);
    if (exact)
        hipLaunchKernelGGL((k_adx<true>), g, b, 0, s, d, t);
    else
        hipLaunchKernelGGL((k_adx<false>), g, b, 0, s, d, t);
}

void launch_adapt_finalize(const Dev& d, bool exact, hipStream_t s) {
    const dim3 g(d.B), b(kFinal);
    if (exact)
        hipLaunchKernelGGL((k_adapt_finalize<true>), g, b, 0, s, d);
    else
        hipLaunchKernelGGL((k_adapt_finalize<false>), g, b, 0, s, d);
}

template <int CAND>
static void resid_pair(const Dev& d, bool exact, hipStream_t s) {
    const dim3 b(kBlock);
    if (d.ntA > 0) {
        const dim3 g(tile_grid(d.ntA));
        if (exact)
            hipLaunchKernelGGL((k_resid_rows<true, CAND>), g, b, 0, s, d);
        else
            hipLaunchKernelGGL((k_resid_rows<false, CAND>), g, b, 0, s, d);
    }
    if (d.ntT > 0) {
        const dim3 g(tile_grid(d.ntT));
        if (exact)
            hipLaunchKernelGGL((k_resid_cols<true, CAND>), g, b, 0, s, d);
        else
            hipLaunchKernelGGL((k_resid_cols<false, CAND>), g, b, 0, s, d);
    }
}

void launch_resid(const Dev& d, bool exact, bool current, hipStream_t s) {
    if (current)
        resid_pair<1>(d, exact, s);
    else
        resid_pair<0>(d, exact, s);
}

void launch_check_resid(const Dev& d, bool exact, bool both, bool packed_floor, hipStream_t s) {
    if (!exact && d.blocks) {
        launch_resid_blk(d, both, packed_floor, s);
        return;
    }
    launch_resid(d, exact, false, s);
    if (both) launch_resid(d, exact, true, s);
    if (packed_floor && !exact) resid_pair<2>(d, false, s);
}

void launch_check_finalize(const Dev& d, const IterOpts& o, int t, bool exact, bool final_pass,
                           hipStream_t s) {
    const dim3 g(d.B), b(kFinal);
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

void launch_ray_check(const Dev& d, const IterOpts& o, int t, hipStream_t s) {
    const int64_t tot = (int64_t)d.m + d.n;
    if (tot == 0) return;
    hipLaunchKernelGGL(k_ray_prep, dim3(elem_grid(tot)), dim3(kBlock), 0, s, d);
    if (d.ntA > 0) hipLaunchKernelGGL(k_ray_rows, dim3(tile_grid(d.ntA)), dim3(kBlock), 0, s, d);
    if (d.ntT > 0) hipLaunchKernelGGL(k_ray_cols, dim3(tile_grid(d.ntT)), dim3(kBlock), 0, s, d);
    hipLaunchKernelGGL(k_ray_finalize, dim3(d.B), dim3(kFinal), 0, s, d, o, t);
}

void launch_post_check(const Dev& d, bool kkt, hipStream_t s) {
    const int nt = d.ntT + d.ntA;
    if (nt == 0) return;
    hipLaunchKernelGGL(k_post_check, dim3((nt + kWaves - 1) / kWaves), dim3(kBlock), 0, s, d, kkt ? 1 : 0);
}

void launch_advance(const Dev& d, int len, hipStream_t s) {
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, s, d, len);
}

void launch_ruiz_pass(const Dev& d, double tol, hipStream_t s) {
    const dim3 b(kBlock);
    if (d.ntT) hipLaunchKernelGGL((k_ruiz_max<false>), dim3(tile_grid(d.ntT)), b, 0, s, d);
    if (d.ntA) hipLaunchKernelGGL((k_ruiz_max<true>), dim3(tile_grid(d.ntA)), b, 0, s, d);
    hipLaunchKernelGGL(k_ruiz_decide, dim3(elem_grid(d.B)), b, 0, s, d, tol);
    hipLaunchKernelGGL(k_ruiz_deltas, dim3(elem_grid((int64_t)d.m + d.n)), b, 0, s, d);
    const int nt = d.ntT + d.ntA;
    if (nt) hipLaunchKernelGGL(k_ruiz_apply<false>, dim3((nt + kWaves - 1) / kWaves), b, 0, s, d);
}

void launch_row_weights(const Dev& d, int k8, hipStream_t s) {
    const dim3 b(kBlock);
    hipLaunchKernelGGL(k_row_weights, dim3(elem_grid((int64_t)d.m + d.n)), b, 0, s, d, k8);
    const int nt = d.ntT + d.ntA;
    if (nt) hipLaunchKernelGGL(k_ruiz_apply<true>, dim3((nt + kWaves - 1) / kWaves), b, 0, s, d);
}

void launch_setup_data(const Dev& d, hipStream_t s) {
    const dim3 b(kBlock);
    if (d.ntA) hipLaunchKernelGGL(k_setup_rows, dim3(tile_grid(d.ntA)), b, 0, s, d);
    if (d.ntT) hipLaunchKernelGGL(k_setup_cols, dim3(tile_grid(d.ntT)), b, 0, s, d);
}

void launch_init_iterates(const Dev& d, const double* warm_x, const double* warm_y, hipStream_t s) {
    const int64_t tot = (int64_t)d.m + d.n;
    if (tot == 0) return;
    hipLaunchKernelGGL(k_init_iterates, dim3(elem_grid(tot)), dim3(kBlock), 0, s, d, warm_x, warm_y);
}

void launch_setup_state(const Dev& d, double initial_weight, double step_reduction, int start_mode,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_setup_state, dim3(elem_grid(d.B)), dim3(kBlock), 0, s, d, initial_weight,
                       step_reduction, start_mode);
}

void launch_step_cap(const Dev& d, int64_t iters, bool exact, hipStream_t s) {
    if (iters <= 0 || d.n == 0 || d.B == 0) return;
    const dim3 b(kBlock), gn(elem_grid(d.n)), gm(elem_grid(d.m > 0 ? d.m : 1));
    auto sum = [&](int check) {
        if (exact)
            hipLaunchKernelGGL((k_pw_sum<true>), dim3(d.B), dim3(kFinal), 0, s, d, check);
        else
            hipLaunchKernelGGL((k_pw_sum<false>), dim3(d.B), dim3(kFinal), 0, s, d, check);
    };
    hipLaunchKernelGGL(k_pw_init, dim3(elem_grid(d.n > d.B ? d.n : d.B)), b, 0, s, d);
    sum(1);
    for (int64_t k = 0; k < iters; ++k) {
        hipLaunchKernelGGL(k_pw_cols_in, gn, b, 0, s, d);
        launch_spmv_plain(d.tA, d.ntA, d.rp, d.ci, d.av64, d.aty2, d.rtmp, false, s);
        if (d.m) hipLaunchKernelGGL(k_pw_rows, gm, b, 0, s, d);
        launch_spmv_plain(d.tT, d.ntT, d.cp, d.ri, d.tv64, d.rtmp, d.ctmp, false, s);
        hipLaunchKernelGGL(k_pw_cols_out, gn, b, 0, s, d);
        sum(k + 1 < iters ? 1 : 0);
    }
    hipLaunchKernelGGL(k_pw_finish, dim3(elem_grid(d.B)), b, 0, s, d);
}

void launch_unscale(const Dev& d, double* x_out, double* y_out, hipStream_t s) {
    const int64_t tot = (int64_t)d.m + d.n;
    if (tot == 0) return;
    hipLaunchKernelGGL(k_unscale, dim3(elem_grid(tot)), dim3(kBlock), 0, s, d, x_out, y_out);
}

__global__ __launch_bounds__(kBlock) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst,
                                                  int64_t n4) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n4; k += stride) dst[k] = src[k];
}

void launch_copy(const float* src, float* dst, int64_t n, hipStream_t s) {
    const int64_t n4 = n / 4;
    if (n4 <= 0) return;
    const int grid = (int)std::min<int64_t>(elem_grid(n4), 256 * 16);
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kBlock), 0, s, reinterpret_cast<const float4*>(src),
                       reinterpret_cast<float4*>(dst), n4);
}

void launch_fill(double* p, double v, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill, dim3(elem_grid(n)), dim3(kBlock), 0, s, p, v, n);
}

void launch_spmv_plain(const Tile* tiles, int ntiles, const int32_t* ptr, const int32_t* idx, const void* val,
                       const void* vec, void* out, bool fp32, hipStream_t s) {
    if (ntiles == 0) return;
    const dim3 g(tile_grid(ntiles)), b(kBlock);
    if (fp32)
        hipLaunchKernelGGL((k_spmv_plain<float>), g, b, 0, s, tiles, ntiles, ptr, idx, (const float*)val,
                           (const float*)vec, (float*)out);
    else
        hipLaunchKernelGGL((k_spmv_plain<double>), g, b, 0, s, tiles, ntiles, ptr, idx, (const double*)val,
                           (const double*)vec, (double*)out);
}

void launch_spmv_plain_packed(const Tile* tiles, int ntiles, const int32_t* ptr, const uint32_t* idxp,
                              const float* val32, const double* vec, double* out, hipStream_t s) {
    if (ntiles == 0) return;
    hipLaunchKernelGGL((k_spmv_plain<double, float, uint32_t>), dim3(tile_grid(ntiles)), dim3(kBlock), 0, s, tiles,
                       ntiles, ptr, idxp, val32, vec, out);
}

}  // namespace mcp
