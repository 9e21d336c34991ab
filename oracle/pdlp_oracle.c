/*
 * pdlp_oracle.c -- CPU restatement of markov-cero's PDLP path, in C99.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * liboracle.so.  Nothing in the product links or calls it.
 *
 * Every function restates one piece of the reference in the SAME operation
 * order (serial sums in index order, no FMA contraction: build with
 * -ffp-contract=off), so that on x86-64 it reproduces the reference bit for
 * bit.  Pinning (tests/test_oracle.py):
 *   or_spmv / or_spmv_t      -- dense long-double products on seeded matrices,
 *                               exact integer matrices, compiled reference
 *   or_equilibrate_model     -- A_s = D_r A D_c identity, norms within 1e-3 of 1
 *                               on convergence, one-pass diagonal case,
 *                               compiled reference scalers (bit-exact)
 *   or_residuals             -- zero at planted (x*, -y*), hand 2x2 cases,
 *                               compiled reference
 *   or_solve_pdlp            -- compiled reference (bit-exact), golden
 *                               fixtures (iterations,
 *                               objective, residuals, vector hashes), ADR-P3-05
 *                               known answer, pdlp_test objectives, planted
 *                               optimum c^T x*, Netlib optima (afiro, adlittle,
 *                               recipe); the KKT and Halpern extension schemes
 *                               by planted and Netlib optima, and the Halpern
 *                               update by three iterations worked by hand
 *   or_verify_primal         -- hand cases + compiled reference
 *   or_halpern_weight        -- closed forms (long-double powers), both sides
 *                               of the guard, the clamp, the drop limit
 *   or_row_weight,           -- closed forms of (1 + beta)^(-k/8); the weighted
 *   or_apply_row_weights        matrix, scales and bounds entry for entry;
 *                               weighted solves reach planted optima
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/mcpdlp.h"

#define OR_BIG 1e300

static double or_clamp(double v, double lo, double hi) { /* std::clamp */
    return (v < lo) ? lo : ((hi < v) ? hi : v);
}
static double or_max(double a, double b) { return (a < b) ? b : a; } /* std::max */
static double or_min(double a, double b) { return (b < a) ? b : a; } /* std::min */

static double lo_val(const mc_bound* b) { /* pdlp.cpp:47 */
    return b->kind == MC_BOUND_NEGATIVE_INFINITY ? -OR_BIG : b->value;
}
static double hi_val(const mc_bound* b) { /* pdlp.cpp:48 */
    return b->kind == MC_BOUND_POSITIVE_INFINITY ? +OR_BIG : b->value;
}

/* spmv, reference src/lp/first_order/pdlp.cpp:22-32: column scatter, skips
 * x[col] == 0, accumulates y[row] in increasing column order. */
void or_spmv(int64_t m, int64_t n, const int64_t* cs, const int64_t* ri, const double* v,
             const double* x, double* y) {
    for (int64_t i = 0; i < m; ++i) y[i] = 0.0;
    for (int64_t col = 0; col < n; ++col) {
        const double xc = x[col];
        if (xc == 0.0) continue;
        for (int64_t p = cs[col]; p < cs[col + 1]; ++p) y[ri[p]] += v[p] * xc;
    }
}

/* spmv_t, reference pdlp.cpp:34-44: per-column dot in increasing row order. */
void or_spmv_t(int64_t m, int64_t n, const int64_t* cs, const int64_t* ri, const double* v,
               const double* y, double* z) {
    (void)m;
    for (int64_t col = 0; col < n; ++col) {
        double s = 0.0;
        for (int64_t p = cs[col]; p < cs[col + 1]; ++p) s += v[p] * y[ri[p]];
        z[col] = s;
    }
}

/* equilibrate_model, reference src/scale/ruiz_scaling.cpp:152-247.
 * Scales `val` (nnz), the four bound vectors and the objective in place and
 * writes row_scale[m], col_scale[n].  Returns passes executed; *converged
 * set as RuizScalers::converged. */
int64_t or_equilibrate_model(int64_t m, int64_t n, const int64_t* cs, const int64_t* ri,
                             double* val, double* objective, mc_bound* row_lower,
                             mc_bound* row_upper, mc_bound* var_lower, mc_bound* var_upper,
                             int64_t max_iterations, double tolerance, double* row_scale,
                             double* col_scale, int32_t* converged) {
    *converged = 0;
    for (int64_t i = 0; i < m; ++i) row_scale[i] = 1.0;
    for (int64_t j = 0; j < n; ++j) col_scale[j] = 1.0;
    const int64_t nnz = (n > 0) ? cs[n] : 0;
    if (m == 0 || n == 0 || nnz == 0) return 0;
    double* r_norm = (double*)malloc(sizeof(double) * (size_t)m);
    double* c_norm = (double*)malloc(sizeof(double) * (size_t)n);
    double* d_r = (double*)malloc(sizeof(double) * (size_t)m);
    double* d_c = (double*)malloc(sizeof(double) * (size_t)n);
    int64_t iter = 0;
    for (; iter < max_iterations; ++iter) {
        for (int64_t i = 0; i < m; ++i) r_norm[i] = 0.0;
        for (int64_t j = 0; j < n; ++j) c_norm[j] = 0.0;
        for (int64_t col = 0; col < n; ++col)
            for (int64_t p = cs[col]; p < cs[col + 1]; ++p) {
                const double a = fabs(val[p]);
                c_norm[col] = or_max(c_norm[col], a);
                r_norm[ri[p]] = or_max(r_norm[ri[p]], a);
            }
        for (int64_t i = 0; i < m; ++i)
            if (r_norm[i] == 0.0) r_norm[i] = 1.0;
        for (int64_t j = 0; j < n; ++j)
            if (c_norm[j] == 0.0) c_norm[j] = 1.0;
        double max_r_err = 0.0, max_c_err = 0.0;
        for (int64_t i = 0; i < m; ++i) max_r_err = or_max(max_r_err, fabs(1.0 - r_norm[i]));
        for (int64_t j = 0; j < n; ++j) max_c_err = or_max(max_c_err, fabs(1.0 - c_norm[j]));
        if (max_r_err < tolerance && max_c_err < tolerance) {
            *converged = 1;
            break;
        }
        for (int64_t i = 0; i < m; ++i) {
            d_r[i] = 1.0 / sqrt(r_norm[i]);
            row_scale[i] *= d_r[i];
        }
        for (int64_t j = 0; j < n; ++j) {
            d_c[j] = 1.0 / sqrt(c_norm[j]);
            col_scale[j] *= d_c[j];
        }
        for (int64_t col = 0; col < n; ++col) {
            const double dc = d_c[col];
            for (int64_t p = cs[col]; p < cs[col + 1]; ++p) val[p] *= d_r[ri[p]] * dc;
        }
    }
    /* ruiz_scaling.cpp:221-244: bounds by D_r / D_c^{-1}, objective by D_c. */
    for (int64_t i = 0; i < m; ++i) {
        if (row_lower[i].kind == MC_BOUND_FINITE) row_lower[i].value *= row_scale[i];
        if (row_upper[i].kind == MC_BOUND_FINITE) row_upper[i].value *= row_scale[i];
    }
    for (int64_t j = 0; j < n; ++j) {
        const double inv = 1.0 / col_scale[j];
        if (var_lower[j].kind == MC_BOUND_FINITE) var_lower[j].value *= inv;
        if (var_upper[j].kind == MC_BOUND_FINITE) var_upper[j].value *= inv;
        objective[j] *= col_scale[j];
    }
    free(r_norm);
    free(c_norm);
    free(d_r);
    free(d_c);
    return iter;
}

typedef struct or_resid {
    double primal, dual, gap, score, objective;
} or_resid;

/* compute_unscaled_residuals, reference pdlp.cpp:62-135, on the ORIGINAL
 * model.  x_out[n], y_out[m] receive the unscaled average iterate. */
void or_residuals(const mc_lp* lp, const double* x_avg, const double* y_avg,
                  const double* Ax_avg, const double* At_y_avg, const double* row_scale,
                  const double* col_scale, int32_t ruiz, double* x_out, double* y_out,
                  or_resid* r) {
    const int64_t m = lp->rows, n = lp->cols;
    const double obj_sign = (lp->sense == MC_MAXIMIZE) ? -1.0 : 1.0;
    for (int64_t j = 0; j < n; ++j) x_out[j] = ruiz ? (x_avg[j] * col_scale[j]) : x_avg[j];
    for (int64_t i = 0; i < m; ++i) y_out[i] = ruiz ? (y_avg[i] * row_scale[i]) : y_avg[i];
    double max_prim = 0.0;
    for (int64_t i = 0; i < m; ++i) {
        const double ax = ruiz ? (Ax_avg[i] / row_scale[i]) : Ax_avg[i];
        const double lo = lo_val(&lp->row_lower[i]), hi = hi_val(&lp->row_upper[i]);
        const double proj = or_clamp(ax, lo, hi);
        const double res = fabs(ax - proj);
        const double sc = 1.0 + or_max(fabs(ax), fabs(proj) < 1e299 ? fabs(proj) : 0.0);
        max_prim = or_max(max_prim, res / sc);
    }
    for (int64_t j = 0; j < n; ++j) {
        const double lo = lo_val(&lp->var_lower[j]), hi = hi_val(&lp->var_upper[j]);
        const double proj = or_clamp(x_out[j], lo, hi);
        const double res = fabs(x_out[j] - proj);
        const double sc = 1.0 + or_max(fabs(x_out[j]), fabs(proj) < 1e299 ? fabs(proj) : 0.0);
        max_prim = or_max(max_prim, res / sc);
    }
    r->primal = max_prim;
    double c_scale = 1.0, dual_sq = 0.0;
    for (int64_t j = 0; j < n; ++j) {
        const double aty = ruiz ? (At_y_avg[j] / col_scale[j]) : At_y_avg[j];
        const double c_orig = obj_sign * lp->objective[j];
        c_scale = or_max(c_scale, fabs(c_orig));
        const double g = c_orig + aty;
        const double lo = lo_val(&lp->var_lower[j]), hi = hi_val(&lp->var_upper[j]);
        const double xp = or_clamp(x_out[j] - g, lo, hi);
        const double d = x_out[j] - xp;
        dual_sq += d * d;
    }
    r->dual = sqrt(dual_sq) / c_scale;
    double pobj = lp->objective_offset;
    for (int64_t j = 0; j < n; ++j) pobj += lp->objective[j] * x_out[j];
    r->objective = pobj;
    double dobj = lp->objective_offset;
    for (int64_t i = 0; i < m; ++i) {
        const double lo = lo_val(&lp->row_lower[i]), hi = hi_val(&lp->row_upper[i]);
        if (y_out[i] > 0.0 && hi < 1e299)
            dobj -= y_out[i] * hi;
        else if (y_out[i] < 0.0 && lo > -1e299)
            dobj -= y_out[i] * lo;
    }
    for (int64_t j = 0; j < n; ++j) {
        const double aty = ruiz ? (At_y_avg[j] / col_scale[j]) : At_y_avg[j];
        const double c_orig = obj_sign * lp->objective[j];
        const double g = c_orig + aty;
        const double lo = lo_val(&lp->var_lower[j]), hi = hi_val(&lp->var_upper[j]);
        if (g > 0.0 && lo > -1e299)
            dobj += g * lo;
        else if (g < 0.0 && hi < 1e299)
            dobj += g * hi;
    }
    r->gap = fabs(pobj - dobj) / (1.0 + fabs(pobj) + fabs(dobj));
    /* std::max({p, d, g}) */
    double s = r->primal;
    if (s < r->dual) s = r->dual;
    if (s < r->gap) s = r->gap;
    r->score = s;
}

/* Trace of the scalar state at every check, for trajectory comparisons. */
typedef struct or_trace {
    int64_t capacity;
    int64_t count;
    double* score;   /* [capacity] score at each check            */
    double* omega;   /* [capacity] omega after the check          */
    double* eta;     /* [capacity] eta at the check               */
    int32_t* restart;/* [capacity] 1 if the check restarted       */
} or_trace;

/* Step-size cap of the Halpern scheme (extension; no counterpart in the
 * reference).  The reference's steps are tau_j = (eta/omega)/colnorm_j and
 * sigma_i = (eta omega)/rownorm_i (pdlp.cpp:186-191) with eta capped at 0.99
 * (:214-217): that assumes ||M||_2 <= 1 for M = D_r^-1/2 A D_c^-1/2 (D the L1
 * norms), which always holds, while PDHG needs only eta ||M|| < 1.  K power
 * iterations on M^T M from v_j = 1 + (j mod 7)/1000 estimate L <= ||M||
 * (within 4 % after 20 iterations on the benchmark LPs), and the cap becomes
 * 0.95 / L.  The reference's power_iterations option (pdlp.hpp:25, unread by
 * the reference) gives K.  Every sum is serial in index order. */
double or_step_cap(int64_t m, int64_t n, const int64_t* cs, const int64_t* ri, const double* val,
                          const double* row_norms, const double* col_norms, int64_t iters) {
    double* v = (double*)malloc(sizeof(double) * (size_t)n);
    double* t = (double*)malloc(sizeof(double) * (size_t)n);
    double* u = (double*)malloc(sizeof(double) * (size_t)m);
    double s = 0.0, cap = 0.99;
    for (int64_t j = 0; j < n; ++j) {
        v[j] = 1.0 + (double)(j % 7) / 1000.0;
        s += v[j] * v[j];
    }
    int64_t k = 0;
    for (; k < iters && s > 0.0; ++k) {
        const double nv = sqrt(s);
        for (int64_t j = 0; j < n; ++j) t[j] = (v[j] / nv) / sqrt(col_norms[j]);
        or_spmv(m, n, cs, ri, val, t, u);
        for (int64_t i = 0; i < m; ++i) u[i] = u[i] / row_norms[i];
        or_spmv_t(m, n, cs, ri, val, u, v);
        s = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            v[j] = v[j] / sqrt(col_norms[j]);
            s += v[j] * v[j];
        }
    }
    const double L = sqrt(sqrt(s));
    if (iters > 0 && k == iters && L > 0.0 && L < 1e300) cap = 0.95 / L;
    free(v);
    free(t);
    free(u);
    return cap;
}

/* Infeasibility certificates (extension; the reference declares
 * PdlpStatus::infeasible_or_unbounded, pdlp.hpp:43, but never returns it).
 * (dx, dy) is a movement of the iterate in the ORIGINAL space (Applegate et
 * al. 2021, sec. 5: a diverging PDHG iterate moves along a ray).
 *  1: dy is a dual ray -- the LP is primal infeasible: the homogeneous dual
 *     objective D = sum_i -support_[lo_i, hi_i](dy_i) + sum_j (g_j > 0 ?
 *     g_j lo_j : g_j < 0 ? g_j hi_j : 0), g = A^T dy, exceeds 1e-9 times the
 *     sum of its terms' magnitudes, and the parts that would need an
 *     infinite bound sum to at most tol D;
 *  2: dx is a primal ray -- the LP is unbounded: c^T dx (c with the
 *     objective's sense) is below -1e-9 times the sum of |c_j dx_j|, and the
 *     recession violations of A dx against the finite row bounds and of dx
 *     against the finite variable bounds are at most tol |c^T dx|;
 *  0: neither.  A feasible LP has D <= 0 along every dual ray and a bounded
 *  one c^T dx >= 0 along every primal ray (weak duality). */
int or_ray_certificate(const mc_lp* lp, const double* dx, const double* dy, double tol) {
    const int64_t m = lp->rows, n = lp->cols;
    const double sense = (lp->sense == MC_MAXIMIZE) ? -1.0 : 1.0;
    double* adx = (double*)calloc((size_t)(m > 0 ? m : 1), sizeof(double));
    double* g = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    for (int64_t j = 0; j < n; ++j)
        for (int64_t p = lp->column_start[j]; p < lp->column_start[j + 1]; ++p) {
            adx[lp->row_index[p]] += lp->value[p] * dx[j];
            g[j] += lp->value[p] * dy[lp->row_index[p]];
        }
    double dobj = 0.0, dmag = 0.0, dinf = 0.0, ymax = 0.0;
    double cdx = 0.0, cmag = 0.0, pviol = 0.0, xmax = 0.0;
    for (int64_t i = 0; i < m; ++i) {
        const int has_lo = lp->row_lower[i].kind == MC_BOUND_FINITE, has_hi = lp->row_upper[i].kind == MC_BOUND_FINITE;
        const double lo = lp->row_lower[i].value, hi = lp->row_upper[i].value;
        /* dual ray: -support of the row interval along dy_i */
        double t = 0.0;
        if (dy[i] > 0.0) {
            if (has_hi) t = -dy[i] * hi; else dinf += dy[i];
        } else if (dy[i] < 0.0) {
            if (has_lo) t = -dy[i] * lo; else dinf += -dy[i];
        }
        dobj += t;
        dmag += fabs(t);
        ymax = or_max(ymax, fabs(dy[i]));
        /* primal ray: A dx in the recession cone of the row interval */
        if (has_lo && adx[i] < 0.0) pviol = or_max(pviol, -adx[i]);
        if (has_hi && adx[i] > 0.0) pviol = or_max(pviol, adx[i]);
    }
    for (int64_t j = 0; j < n; ++j) {
        const int has_lo = lp->var_lower[j].kind == MC_BOUND_FINITE, has_hi = lp->var_upper[j].kind == MC_BOUND_FINITE;
        const double lo = lp->var_lower[j].value, hi = lp->var_upper[j].value;
        double t = 0.0;
        if (g[j] > 0.0) {
            if (has_lo) t = g[j] * lo; else dinf += g[j];
        } else if (g[j] < 0.0) {
            if (has_hi) t = g[j] * hi; else dinf += -g[j];
        }
        dobj += t;
        dmag += fabs(t);
        if (has_lo && dx[j] < 0.0) pviol = or_max(pviol, -dx[j]);
        if (has_hi && dx[j] > 0.0) pviol = or_max(pviol, dx[j]);
        const double ct = sense * lp->objective[j] * dx[j];
        cdx += ct;
        cmag += fabs(ct);
        xmax = or_max(xmax, fabs(dx[j]));
    }
    free(adx);
    free(g);
    if (ymax > 0.0 && dobj > 1e-9 * dmag && dinf <= tol * dobj) return 1;
    if (xmax > 0.0 && cdx < -1e-9 * cmag && pviol <= tol * -cdx) return 2;
    return 0;
}

/* Primal weight of the Halpern scheme at a restart (extension): from the
 * movement since the last restart (dX = |x - x_last|, dY = |y - y_last|) the
 * movement rule omega_move = ratio^(127/128) omega^(1/128), ratio = dY / dX
 * (omega follows the primal/dual distance balance almost fully, the
 * proportional gain 0.99 of cuPDLPx, Lu & Yang 2025); from the residuals
 * omega_res = omega sqrt(clamp(sqrt(p/d), 0.05, 20)) (the reference's rule,
 * pdlp.cpp:247-253); the result omega_move^(3/4) omega_res^(1/4), but never a
 * move against a residual imbalance of 100x or more, clamped to [1e-6, 1e6].
 * Every power is a chain of correctly rounded square roots, so a device
 * evaluation in the same order is bit-identical. */
double or_halpern_weight(double omega, double dX, double dY, double primal, double dual, int32_t max_drop) {
    double wmove = omega;
    if (dX > 1e-10 && dY > 1e-10) {
        const double ratio = dY / dX;
        double f = omega / ratio;
        for (int q = 0; q < 7; ++q) f = sqrt(f);
        wmove = ratio * f;
    }
    double wres = omega;
    if (primal > 1e-12 && dual > 1e-12) wres = omega * sqrt(or_clamp(sqrt(primal / dual), 0.05, 20.0));
    double w = sqrt(wmove * sqrt(wmove * wres));
    if (primal > 1e-12 && dual > 1e-12) {
        if (primal > 100.0 * dual && w < omega) w = omega;
        if (dual > 100.0 * primal && w > omega) w = omega;
    }
    /* extension: at most a fall by max_drop (0: none; a power of two, so exact) */
    if (max_drop > 0) w = or_max(w, omega / (double)max_drop);
    return or_clamp(w, 1e-6, 1e6);
}

/* The check's certificate test: the candidate's movement since the last
 * restart (the start point under the reference scheme), scaled back to the
 * original space, tested by or_ray_certificate. */
static int or_movement_certifies(const mc_lp* lp, const double* xc, const double* yc, const double* x_last,
                                 const double* y_last, const double* col_scale, const double* row_scale, int ruiz,
                                 double tol) {
    const int64_t m = lp->rows, n = lp->cols;
    double* dx = (double*)malloc(sizeof(double) * (size_t)n);
    double* dy = (double*)malloc(sizeof(double) * (size_t)m);
    for (int64_t j = 0; j < n; ++j) dx[j] = (xc[j] - x_last[j]) * (ruiz ? col_scale[j] : 1.0);
    for (int64_t i = 0; i < m; ++i) dy[i] = (yc[i] - y_last[i]) * (ruiz ? row_scale[i] : 1.0);
    const int r = or_ray_certificate(lp, dx, dy, tol > 0.0 ? tol : 1e-8);
    free(dx);
    free(dy);
    return r != 0;
}

/* solve_pdlp, reference pdlp.cpp:139-276.  With o->restart_scheme ==
 * MC_SCHEME_KKT the restart decision follows the extension documented in
 * include/mcpdlp.h (Applegate et al. 2021 adaptive restarts); with
 * MC_SCHEME_HALPERN the iterate also follows the reflected Halpern update
 * documented there (the running average slot holds T(z)); everything else is
 * unchanged.  `res->primal`/`res->dual` must hold
 * n / m doubles (or be NULL).  Timings are wall-clock of this restatement.
 * Optional `trace` records the per-check scalar sequence. */
/* Row weight (1 + beta)^(-k/8), k in 0..8: the square roots s1 = (1+beta)^(1/2),
 * s2 = s1^(1/2), s3 = s2^(1/2) that the bits of k select (8: 1 + beta itself),
 * multiplied in the order 8, 4, 2, 1, then inverted.  Every step is correctly
 * rounded, so the HIP engine reproduces it bit for bit. */
double or_row_weight(double beta, int32_t k) {
    const double s0 = 1.0 + beta;
    const double s1 = sqrt(s0), s2 = sqrt(s1), s3 = sqrt(s2);
    double p = 1.0;
    if (k & 8) p *= s0;
    if (k & 4) p *= s1;
    if (k & 2) p *= s2;
    if (k & 1) p *= s3;
    return 1.0 / p;
}

/* Row weighting after the Ruiz passes (extension; no reference counterpart):
 * beta_i is the largest magnitude of row i's finite ORIGINAL bounds (0 for a
 * free row), w_i = or_row_weight(beta_i, k); row_scale[i] *= w_i, every
 * nonzero of row i *= w_i, and the scaled row bounds become the original
 * bounds times the final row_scale[i] (the order in which the engine forms
 * them).  The reported residuals divide row i's violation by 1 + |a_i x|, about
 * 1 + beta_i at an active bound: rows with small bounds must be met more
 * tightly, and with L1-norm step sizes the weight multiplies row i's dual
 * step by w_i. */
void or_apply_row_weights(const mc_lp* lp, int32_t k, double* val, mc_bound* rl, mc_bound* ru,
                          double* row_scale) {
    const int64_t m = lp->rows, n = lp->cols;
    double* w = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    for (int64_t i = 0; i < m; ++i) {
        double beta = 0.0;
        if (lp->row_lower[i].kind == MC_BOUND_FINITE) beta = or_max(beta, fabs(lp->row_lower[i].value));
        if (lp->row_upper[i].kind == MC_BOUND_FINITE) beta = or_max(beta, fabs(lp->row_upper[i].value));
        w[i] = or_row_weight(beta, k);
        row_scale[i] *= w[i];
        if (rl[i].kind == MC_BOUND_FINITE) rl[i].value = lp->row_lower[i].value * row_scale[i];
        if (ru[i].kind == MC_BOUND_FINITE) ru[i].value = lp->row_upper[i].value * row_scale[i];
    }
    for (int64_t col = 0; col < n; ++col)
        for (int64_t p = lp->column_start[col]; p < lp->column_start[col + 1]; ++p) val[p] *= w[lp->row_index[p]];
    free(w);
}

int or_solve_pdlp(const mc_lp* lp, const mc_pdlp_options* o, mc_pdlp_result* res,
                  or_trace* trace) {
    const int64_t m = lp->rows, n = lp->cols;
    const int requested_gpu = (o->backend == MC_BACKEND_GPU);
    memset(res->message, 0, sizeof(res->message));
    memset(res->backend_actually_used, 0, sizeof(res->backend_actually_used));
    strcpy(res->backend_actually_used, requested_gpu ? "cpu_fallback" : "cpu");
    res->h2d_ms = res->d2h_ms = 0.0;
    res->kernel_ms = res->total_ms = 0.0;
    res->tolerance = o->primal_tolerance;
    if (n == 0 || m == 0) { /* pdlp.cpp:144-149 */
        res->status = MC_PDLP_OPTIMAL;
        res->objective = res->primal_infeasibility = res->dual_infeasibility = 0.0;
        res->duality_gap = 0.0;
        res->iterations = 0;
        res->primal_len = res->dual_len = 0;
        strcpy(res->message, "trivial");
        return 0;
    }
    const int64_t nnz = lp->column_start[n];
    /* model copy (pdlp.cpp:150) */
    double* val = (double*)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    double* obj = (double*)malloc(sizeof(double) * (size_t)n);
    mc_bound* rl = (mc_bound*)malloc(sizeof(mc_bound) * (size_t)m);
    mc_bound* ru = (mc_bound*)malloc(sizeof(mc_bound) * (size_t)m);
    mc_bound* vl = (mc_bound*)malloc(sizeof(mc_bound) * (size_t)n);
    mc_bound* vu = (mc_bound*)malloc(sizeof(mc_bound) * (size_t)n);
    memcpy(val, lp->value, sizeof(double) * (size_t)nnz);
    memcpy(obj, lp->objective, sizeof(double) * (size_t)n);
    memcpy(rl, lp->row_lower, sizeof(mc_bound) * (size_t)m);
    memcpy(ru, lp->row_upper, sizeof(mc_bound) * (size_t)m);
    memcpy(vl, lp->var_lower, sizeof(mc_bound) * (size_t)n);
    memcpy(vu, lp->var_upper, sizeof(mc_bound) * (size_t)n);
    double* row_scale = (double*)malloc(sizeof(double) * (size_t)m);
    double* col_scale = (double*)malloc(sizeof(double) * (size_t)n);
    for (int64_t i = 0; i < m; ++i) row_scale[i] = 1.0;
    for (int64_t j = 0; j < n; ++j) col_scale[j] = 1.0;
    int32_t conv = 0;
    if (o->ruiz_scaling)
        or_equilibrate_model(m, n, lp->column_start, lp->row_index, val, obj, rl, ru, vl, vu,
                             (int64_t)o->ruiz_iterations, 1e-3, row_scale, col_scale, &conv);
    /* Row weights (extension, mc_pdlp_options::row_weighting = k): after the
     * Ruiz passes, row i and its bounds are scaled by (1 + beta_i)^(-k/8). */
    const int32_t scaled = o->ruiz_scaling || o->row_weighting > 0;
    if (o->row_weighting > 0)
        or_apply_row_weights(lp, o->row_weighting, val, rl, ru, row_scale);
    const int64_t* cs = lp->column_start;
    const int64_t* ri = lp->row_index;
    const double obj_sign = (lp->sense == MC_MAXIMIZE) ? -1.0 : 1.0;
    double* c = (double*)malloc(sizeof(double) * (size_t)n);
    double c_norm_inf = 1.0;
    for (int64_t j = 0; j < n; ++j) {
        c[j] = obj_sign * obj[j];
        c_norm_inf = or_max(c_norm_inf, fabs(c[j]));
    }
    double* b_lo = (double*)malloc(sizeof(double) * (size_t)m);
    double* b_hi = (double*)malloc(sizeof(double) * (size_t)m);
    double b_norm_inf = 1.0;
    for (int64_t i = 0; i < m; ++i) {
        b_lo[i] = lo_val(&rl[i]);
        b_hi[i] = hi_val(&ru[i]);
        if (fabs(b_lo[i]) < 1e299) b_norm_inf = or_max(b_norm_inf, fabs(b_lo[i]));
        if (fabs(b_hi[i]) < 1e299) b_norm_inf = or_max(b_norm_inf, fabs(b_hi[i]));
    }
    double* row_norms = (double*)calloc((size_t)m, sizeof(double));
    double* col_norms = (double*)calloc((size_t)n, sizeof(double));
    for (int64_t col = 0; col < n; ++col)
        for (int64_t p = cs[col]; p < cs[col + 1]; ++p) {
            const double a = fabs(val[p]);
            col_norms[col] += a;
            row_norms[ri[p]] += a;
        }
    for (int64_t j = 0; j < n; ++j)
        if (col_norms[j] < 1e-12) col_norms[j] = 1.0;
    for (int64_t i = 0; i < m; ++i)
        if (row_norms[i] < 1e-12) row_norms[i] = 1.0;
    double omega = (o->initial_primal_weight > 0.0)
                       ? o->initial_primal_weight
                       : or_clamp(sqrt(c_norm_inf / b_norm_inf), 0.01, 100.0);
    double eta = or_clamp(o->step_size_reduction, 0.1, 0.99);
    const double eta_cap = (o->restart_scheme == MC_SCHEME_HALPERN && o->adaptive_step_size && o->power_iterations > 0)
                               ? or_step_cap(m, n, lp->column_start, lp->row_index, val, row_norms, col_norms,
                                             (int64_t)o->power_iterations)
                               : 0.99;
    double* tau = (double*)malloc(sizeof(double) * (size_t)n);
    double* sigma = (double*)malloc(sizeof(double) * (size_t)m);
#define OR_UPDATE_STEPS()                                                         \
    do {                                                                          \
        for (int64_t j_ = 0; j_ < n; ++j_) tau[j_] = (eta / omega) / col_norms[j_]; \
        for (int64_t i_ = 0; i_ < m; ++i_) sigma[i_] = (eta * omega) / row_norms[i_]; \
    } while (0)
    OR_UPDATE_STEPS();
    double* x = (double*)calloc((size_t)n, sizeof(double));
    double* y = (double*)calloc((size_t)m, sizeof(double));
    double* x_bar = (double*)calloc((size_t)n, sizeof(double));
    double* x_avg = (double*)calloc((size_t)n, sizeof(double));
    double* y_avg = (double*)calloc((size_t)m, sizeof(double));
    double* dx = (double*)calloc((size_t)n, sizeof(double));
    double* x_prev = (double*)malloc(sizeof(double) * (size_t)n);
    double* At_y = (double*)malloc(sizeof(double) * (size_t)n);
    double* Ax = (double*)malloc(sizeof(double) * (size_t)m);
    double* xo = (double*)malloc(sizeof(double) * (size_t)n);
    double* yo = (double*)malloc(sizeof(double) * (size_t)m);
    for (int64_t j = 0; j < n; ++j) {
        x[j] = or_clamp(0.0, lo_val(&vl[j]), hi_val(&vu[j]));
        x_bar[j] = x[j];
        x_avg[j] = x[j];
    }
    const int halpern = (o->restart_scheme == MC_SCHEME_HALPERN);
    const int kkt = (o->restart_scheme == MC_SCHEME_KKT) || halpern;
    double* y_prev = (double*)calloc((size_t)m, sizeof(double));
    double* x_last = (double*)malloc(sizeof(double) * (size_t)n);
    double* y_last = (double*)calloc((size_t)m, sizeof(double));
    double* Ax2 = (double*)malloc(sizeof(double) * (size_t)m);
    double* At_y2 = (double*)malloc(sizeof(double) * (size_t)n);
    double* xo2 = (double*)malloc(sizeof(double) * (size_t)n);
    double* yo2 = (double*)malloc(sizeof(double) * (size_t)m);
    memcpy(x_last, x, sizeof(double) * (size_t)n);
    uint64_t iter = 0, avg_count = 0, since = 0;
    double last_restart_score = 1e300, prev_cand = 1e300;
    or_resid cur, cur2;
    int optimal = 0, use_cur = 0, infeasible = 0;
    /* Residuals of the average (and, for the KKT scheme, of the current
     * iterate); `cur`/xo/yo end up holding the chosen candidate. */
#define OR_EVALUATE()                                                                             \
    do {                                                                                          \
        or_spmv(m, n, cs, ri, val, x_avg, Ax);                                                    \
        or_spmv_t(m, n, cs, ri, val, y_avg, At_y);                                                \
        or_residuals(lp, x_avg, y_avg, Ax, At_y, row_scale, col_scale, scaled, xo, yo,            \
                     &cur);                                                                       \
        use_cur = 0;                                                                              \
        if (kkt) {                                                                                \
            or_spmv(m, n, cs, ri, val, x, Ax2);                                                   \
            or_spmv_t(m, n, cs, ri, val, y, At_y2);                                               \
            or_residuals(lp, x, y, Ax2, At_y2, row_scale, col_scale, scaled, xo2, yo2,            \
                         &cur2);                                                                  \
            if (cur2.score < cur.score) {                                                         \
                double* t_;                                                                       \
                use_cur = 1;                                                                      \
                cur = cur2;                                                                       \
                t_ = xo; xo = xo2; xo2 = t_;                                                      \
                t_ = yo; yo = yo2; yo2 = t_;                                                      \
            }                                                                                     \
        }                                                                                         \
    } while (0)
    while (iter < o->max_iterations) {
        or_spmv_t(m, n, cs, ri, val, y, At_y); /* :195 */
        memcpy(x_prev, x, sizeof(double) * (size_t)n);
        double dxn = 0.0;
        for (int64_t j = 0; j < n; ++j) { /* :198-205 */
            const double g = c[j] + At_y[j];
            const double px = or_clamp(x[j] - tau[j] * g, lo_val(&vl[j]), hi_val(&vu[j]));
            const double d = px - x[j];
            dx[j] = d;
            dxn += col_norms[j] * d * d;
            x[j] = px;
        }
        for (int64_t j = 0; j < n; ++j) x_bar[j] = 2.0 * x[j] - x_prev[j];
        if (o->adaptive_step_size && iter % 10 == 0 && dxn > 1e-14) { /* :207-219 */
            or_spmv(m, n, cs, ri, val, dx, Ax);
            double adxn = 0.0;
            for (int64_t i = 0; i < m; ++i) adxn += (Ax[i] * Ax[i]) / row_norms[i];
            if (adxn > 1e-14) {
                const double L = sqrt(adxn / dxn);
                if (L > 1e-6) {
                    const double target = 0.95 / L;
                    if (target < eta) {
                        eta = or_max(0.1, or_max(target, eta * 0.8));
                        OR_UPDATE_STEPS();
                    } else if (target > 1.05 * eta && eta < eta_cap) {
                        eta = or_min(eta_cap, eta * 1.05);
                        OR_UPDATE_STEPS();
                    }
                }
            }
        }
        or_spmv(m, n, cs, ri, val, x_bar, Ax); /* :220-224 */
        if (halpern) memcpy(y_prev, y, sizeof(double) * (size_t)m);
        for (int64_t i = 0; i < m; ++i) {
            const double v = y[i] + sigma[i] * Ax[i];
            y[i] = v - sigma[i] * or_clamp(v / sigma[i], b_lo[i], b_hi[i]);
        }
        ++avg_count; /* :225-227 */
        if (halpern) { /* T(z) to the average slot; z <- w (2 T(z) - z) + (1 - w) z0 */
            const double w = (double)avg_count / ((double)avg_count + 1.0);
            for (int64_t j = 0; j < n; ++j) {
                x_avg[j] = x[j];
                x[j] = w * x_bar[j] + (1.0 - w) * x_last[j];
            }
            for (int64_t i = 0; i < m; ++i) {
                y_avg[i] = y[i];
                y[i] = w * (2.0 * y[i] - y_prev[i]) + (1.0 - w) * y_last[i];
            }
        } else {
            for (int64_t j = 0; j < n; ++j) x_avg[j] += (x[j] - x_avg[j]) / (double)avg_count;
            for (int64_t i = 0; i < m; ++i) y_avg[i] += (y[i] - y_avg[i]) / (double)avg_count;
        }
        ++iter;
        ++since;
        if (iter % o->restart_every == 0) { /* :229-261 */
            OR_EVALUATE();
            int restarted = 0;
            if (cur.primal <= o->primal_tolerance && cur.dual <= o->dual_tolerance &&
                cur.gap <= o->gap_tolerance) {
                optimal = 1;
            } else if (o->detect_infeasibility && or_movement_certifies(lp, use_cur ? x : x_avg, use_cur ? y : y_avg,
                                                                        x_last, y_last, col_scale, row_scale,
                                                                        scaled, o->infeasibility_tolerance)) {
                infeasible = 1;
                break;
            } else if (!kkt) {
                const double score = cur.score;
                const int do_restart =
                    (o->restart_strategy == MC_RESTART_FIXED) ||
                    (o->restart_strategy == MC_RESTART_ADAPTIVE &&
                     (score <= o->restart_reduction_factor * last_restart_score ||
                      (since >= 5 * o->restart_every && score < last_restart_score)));
                if (do_restart) {
                    if (o->adaptive_primal_weight && cur.primal > 1e-12 && cur.dual > 1e-12) {
                        const double ratio = or_clamp(sqrt(cur.primal / cur.dual), 0.05, 20.0);
                        omega = or_clamp(omega * pow(ratio, o->primal_weight_smoothing), 1e-6, 1e6);
                        OR_UPDATE_STEPS();
                    }
                    memcpy(x, x_avg, sizeof(double) * (size_t)n);
                    memcpy(y, y_avg, sizeof(double) * (size_t)m);
                    for (int64_t j = 0; j < n; ++j) x_bar[j] = x[j];
                    avg_count = 0;
                    since = 0;
                    last_restart_score = score;
                    restarted = 1;
                }
            } else {
                /* KKT scheme (extension, see MC_SCHEME_KKT in include/mcpdlp.h) */
                const double score = cur.score;
                const int do_restart =
                    (o->restart_strategy == MC_RESTART_FIXED) ||
                    (o->restart_strategy == MC_RESTART_ADAPTIVE &&
                     (score <= 0.2 * last_restart_score ||
                      (score <= 0.8 * last_restart_score && score > prev_cand) ||
                      (double)since >= 0.36 * (double)iter));
                prev_cand = score;
                if (do_restart) {
                    const double* xc = use_cur ? x : x_avg;
                    const double* yc = use_cur ? y : y_avg;
                    if (o->adaptive_primal_weight) {
                        double dxs = 0.0, dys = 0.0;
                        for (int64_t j = 0; j < n; ++j) dxs += (xc[j] - x_last[j]) * (xc[j] - x_last[j]);
                        for (int64_t i = 0; i < m; ++i) dys += (yc[i] - y_last[i]) * (yc[i] - y_last[i]);
                        const double dX = sqrt(dxs), dY = sqrt(dys);
                        if (halpern) {
                            omega = or_halpern_weight(omega, dX, dY, cur.primal, cur.dual, o->primal_weight_max_drop);
                            OR_UPDATE_STEPS();
                        } else if (dX > 1e-10 && dY > 1e-10) {
                            const double th = o->primal_weight_smoothing;
                            const double ratio = dY / dX;
                            const double wmove =
                                (th == 0.5) ? sqrt(omega * ratio) : pow(ratio, th) * pow(omega, 1.0 - th);
                            omega = or_clamp(wmove, 1e-6, 1e6);
                            OR_UPDATE_STEPS();
                        }
                    }
                    if (!use_cur) {
                        memcpy(x, x_avg, sizeof(double) * (size_t)n);
                        memcpy(y, y_avg, sizeof(double) * (size_t)m);
                    }
                    memcpy(x_last, x, sizeof(double) * (size_t)n);
                    memcpy(y_last, y, sizeof(double) * (size_t)m);
                    for (int64_t j = 0; j < n; ++j) x_bar[j] = x[j];
                    avg_count = 0;
                    since = 0;
                    last_restart_score = score;
                    prev_cand = 1e300;
                    restarted = 1;
                }
            }
            if (trace && trace->count < trace->capacity) {
                trace->score[trace->count] = cur.score;
                trace->omega[trace->count] = omega;
                trace->eta[trace->count] = eta;
                trace->restart[trace->count] = restarted + 2 * use_cur;
                trace->count++;
            }
            if (optimal) break;
        }
    }
    if (!optimal && !infeasible) { /* :263-265 */
        OR_EVALUATE();
    }
#undef OR_EVALUATE
    res->status = optimal ? MC_PDLP_OPTIMAL : infeasible ? MC_PDLP_INFEASIBLE_OR_UNBOUNDED : MC_PDLP_ITERATION_LIMIT;
    strcpy(res->message, optimal ? "PDLP converged"
                         : infeasible ? "infeasibility certificate found" : "iteration limit reached");
    res->objective = cur.objective;
    res->primal_infeasibility = cur.primal;
    res->dual_infeasibility = cur.dual;
    res->duality_gap = cur.gap;
    res->iterations = iter;
    res->primal_weight = omega;
    res->step_size = eta;
    res->primal_len = n;
    res->dual_len = m;
    if (res->primal) memcpy(res->primal, xo, sizeof(double) * (size_t)n);
    if (res->dual) memcpy(res->dual, yo, sizeof(double) * (size_t)m);
    free(val); free(obj); free(rl); free(ru); free(vl); free(vu);
    free(row_scale); free(col_scale); free(c); free(b_lo); free(b_hi);
    free(row_norms); free(col_norms); free(tau); free(sigma);
    free(x); free(y); free(x_bar); free(x_avg); free(y_avg); free(dx); free(x_prev);
    free(At_y); free(Ax); free(xo); free(yo);
    free(x_last); free(y_last); free(Ax2); free(At_y2); free(xo2); free(yo2); free(y_prev);
#undef OR_UPDATE_STEPS
    return 0;
}

/* verify_primal for continuous variables, reference
 * src/verify/primal_verifier.cpp:9-104 (allowance, check_lower/upper, long
 * double row activity and objective). */
static double or_allowance(double actual, double expected, double a, double r) {
    double mx = 1.0;
    if (mx < fabs(actual)) mx = fabs(actual);
    if (mx < fabs(expected)) mx = fabs(expected);
    return a + r * mx;
}

int or_verify_primal(const mc_lp* lp, const double* x, double claimed, double fa, double fr,
                     double oa, double orel, mc_verify_report* rep) {
    memset(rep, 0, sizeof(*rep));
    const int64_t m = lp->rows, n = lp->cols;
    int32_t viol = 0;
    for (int64_t j = 0; j < n; ++j) {
        if (lp->var_lower[j].kind == MC_BOUND_FINITE) {
            const double b = lp->var_lower[j].value;
            const double mag = or_max(0.0, b - x[j]);
            rep->maximum_variable_violation = or_max(rep->maximum_variable_violation, mag);
            if (mag > or_allowance(x[j], b, fa, fr)) viol++;
        }
        if (lp->var_upper[j].kind == MC_BOUND_FINITE) {
            const double b = lp->var_upper[j].value;
            const double mag = or_max(0.0, x[j] - b);
            rep->maximum_variable_violation = or_max(rep->maximum_variable_violation, mag);
            if (mag > or_allowance(x[j], b, fa, fr)) viol++;
        }
    }
    long double* work = (long double*)calloc((size_t)(m > 0 ? m : 1), sizeof(long double));
    for (int64_t col = 0; col < n; ++col)
        for (int64_t p = lp->column_start[col]; p < lp->column_start[col + 1]; ++p)
            work[lp->row_index[p]] += (long double)lp->value[p] * x[col];
    for (int64_t i = 0; i < m; ++i) {
        const double act = (double)work[i];
        if (lp->row_lower[i].kind == MC_BOUND_FINITE) {
            const double b = lp->row_lower[i].value;
            const double mag = or_max(0.0, b - act);
            rep->maximum_row_violation = or_max(rep->maximum_row_violation, mag);
            if (mag > or_allowance(act, b, fa, fr)) viol++;
        }
        if (lp->row_upper[i].kind == MC_BOUND_FINITE) {
            const double b = lp->row_upper[i].value;
            const double mag = or_max(0.0, act - b);
            rep->maximum_row_violation = or_max(rep->maximum_row_violation, mag);
            if (mag > or_allowance(act, b, fa, fr)) viol++;
        }
    }
    free(work);
    long double objv = lp->objective_offset;
    for (int64_t j = 0; j < n; ++j) objv += (long double)lp->objective[j] * x[j];
    rep->recomputed_objective = (double)objv;
    rep->objective_difference = fabs(rep->recomputed_objective - claimed);
    if (rep->objective_difference > or_allowance(rep->recomputed_objective, claimed, oa, orel)) viol++;
    rep->violation_count = viol;
    rep->passed = (viol == 0);
    return 0;
}

/* Relative KKT score of an unscaled primal/dual pair (x[n], y[m] in the
 * reported sign convention) on the original model: the residuals of
 * pdlp.cpp:80-133 with A x and A^T y formed here.  out = {primal, dual, gap,
 * score, objective}.  Used to check any returned solution independently. */
int or_kkt(const mc_lp* lp, const double* x, const double* y, double* out) {
    const int64_t m = lp->rows, n = lp->cols;
    double* ax = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    double* aty = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* xo = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* yo = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    or_spmv(m, n, lp->column_start, lp->row_index, lp->value, x, ax);
    or_spmv_t(m, n, lp->column_start, lp->row_index, lp->value, y, aty);
    or_resid r;
    or_residuals(lp, x, y, ax, aty, NULL, NULL, 0, xo, yo, &r);
    out[0] = r.primal;
    out[1] = r.dual;
    out[2] = r.gap;
    out[3] = r.score;
    out[4] = r.objective;
    free(ax);
    free(aty);
    free(xo);
    free(yo);
    return 0;
}
