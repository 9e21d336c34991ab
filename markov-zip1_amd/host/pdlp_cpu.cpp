// pdlp_cpu.cpp -- Backend::cpu of solve_pdlp: the host PDHG loop
// (reference src/lp/first_order/pdlp.cpp:139-276) in the reference's
// arithmetic order, so that both backends of this library and the reference
// agree on the same inputs.  Also the cpu_fallback path when no GPU exists.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>

#include "markov_cero/lp/first_order/pdlp.hpp"
#include "markov_cero/scale/ruiz_scaling.hpp"
#include "pdlp_internal.hpp"

namespace markov_cero::lp::first_order::detail {
namespace {

constexpr double kOpen = 1e300;  // bound sentinel of the reference (pdlp.cpp:47-48)

double lower_of(const model::Bound& b) { return b.kind == model::BoundKind::negative_infinity ? -kOpen : b.value; }
double upper_of(const model::Bound& b) { return b.kind == model::BoundKind::positive_infinity ? kOpen : b.value; }

// (1 + beta)^(-k/8) from the square roots of 1 + beta that the bits of k
// select, multiplied in the order 8, 4, 2, 1 (each step correctly rounded,
// so the HIP engine forms the same bits).
double row_weight(double beta, int k) {
    const double s0 = 1.0 + beta;
    const double s1 = std::sqrt(s0), s2 = std::sqrt(s1), s3 = std::sqrt(s2);
    double p = 1.0;
    if (k & 8) p *= s0;
    if (k & 4) p *= s1;
    if (k & 2) p *= s2;
    if (k & 1) p *= s3;
    return 1.0 / p;
}

// PdlpOptions::row_weighting after the Ruiz passes: row i of the scaled model
// s and its row scale are multiplied by row_weight(beta_i, k), beta_i the
// largest magnitude of row i's finite bounds in the original model, and the
// scaled row bounds become the original ones times the final row scale.
void apply_row_weights(const model::Model& orig, model::Model& s, scale::RuizScalers& sc, int k) {
    const std::size_t m = s.matrix.row_count;
    if (sc.row_scale.size() != m) {  // no Ruiz passes: unit scalers
        sc.row_scale.assign(m, 1.0);
        sc.inv_row_scale.assign(m, 1.0);
        sc.col_scale.assign(s.matrix.column_count, 1.0);
        sc.inv_col_scale.assign(s.matrix.column_count, 1.0);
    }
    std::vector<double> w(m);
    for (std::size_t i = 0; i < m; ++i) {
        double beta = 0.0;
        const auto& lo = orig.row_lower[i];
        const auto& hi = orig.row_upper[i];
        if (lo.kind == model::BoundKind::finite) beta = std::max(beta, std::abs(lo.value));
        if (hi.kind == model::BoundKind::finite) beta = std::max(beta, std::abs(hi.value));
        w[i] = row_weight(beta, k);
        sc.row_scale[i] *= w[i];
        sc.inv_row_scale[i] = 1.0 / sc.row_scale[i];
        if (lo.kind == model::BoundKind::finite) s.row_lower[i].value = lo.value * sc.row_scale[i];
        if (hi.kind == model::BoundKind::finite) s.row_upper[i].value = hi.value * sc.row_scale[i];
    }
    for (std::size_t k2 = 0; k2 < s.matrix.value.size(); ++k2) s.matrix.value[k2] *= w[s.matrix.row_index[k2]];
}

// Scaled problem data + iterate of one solve.
struct Pdhg {
    const model::SparseMatrixCSC& A;
    std::size_t m, n;
    std::vector<double> c, lo, hi, blo, bhi, rnorm, cnorm;
    std::vector<double> x, y, xbar, xavg, yavg, dx, aty, ax;
    double eta{0.9}, omega{1.0};

    explicit Pdhg(const model::Model& s)
        : A(s.matrix), m(s.matrix.row_count), n(s.matrix.column_count) {}

    // z = A^T v, per column in increasing row order (pdlp.cpp:34-44).
    void at_times(const std::vector<double>& v, std::vector<double>& z) const {
        for (std::size_t j = 0; j < n; ++j) {
            double s = 0.0;
            for (std::size_t k = A.column_start[j]; k < A.column_start[j + 1]; ++k) s += A.value[k] * v[A.row_index[k]];
            z[j] = s;
        }
    }
    // z = A v, column scatter skipping zero entries of v (pdlp.cpp:22-32).
    void a_times(const std::vector<double>& v, std::vector<double>& z) const {
        std::fill(z.begin(), z.end(), 0.0);
        for (std::size_t j = 0; j < n; ++j) {
            const double vj = v[j];
            if (vj == 0.0) continue;
            for (std::size_t k = A.column_start[j]; k < A.column_start[j + 1]; ++k) z[A.row_index[k]] += A.value[k] * vj;
        }
    }
    // Halpern-scheme step cap: power iterations on M^T M, M = D_r^-1/2 A D_c^-1/2,
    // from v_j = 1 + (j mod 7)/1000, give L <= ||M||; cap = 0.95 / L (0.99, the
    // reference's cap, when the estimate fails).  Serial sums in index order.
    double step_cap(std::size_t iters) const {
        std::vector<double> v(n), t(n), u(m);
        double s = 0.0;
        for (std::size_t j = 0; j < n; ++j) {
            v[j] = 1.0 + static_cast<double>(j % 7) / 1000.0;
            s += v[j] * v[j];
        }
        std::size_t k = 0;
        for (; k < iters && s > 0.0; ++k) {
            const double nv = std::sqrt(s);
            for (std::size_t j = 0; j < n; ++j) t[j] = (v[j] / nv) / std::sqrt(cnorm[j]);
            a_times(t, u);
            for (std::size_t i = 0; i < m; ++i) u[i] = u[i] / rnorm[i];
            at_times(u, v);
            s = 0.0;
            for (std::size_t j = 0; j < n; ++j) {
                v[j] = v[j] / std::sqrt(cnorm[j]);
                s += v[j] * v[j];
            }
        }
        const double L = std::sqrt(std::sqrt(s));
        return (iters > 0 && k == iters && L > 0.0 && L < 1e300) ? 0.95 / L : 0.99;
    }
    double tau(std::size_t j) const { return (eta / omega) / cnorm[j]; }
    double sigma(std::size_t i) const { return (eta * omega) / rnorm[i]; }
};

struct Residuals {
    double primal{0}, dual{0}, gap{0}, score{0}, objective{0};
};

// compute_unscaled_residuals (pdlp.cpp:62-135) on the ORIGINAL model.
Residuals evaluate(const model::Model& orig, const std::vector<double>& xs, const std::vector<double>& ys,
                   const std::vector<double>& axs, const std::vector<double>& atys, const scale::RuizScalers& sc,
                   bool ruiz, std::vector<double>& xu, std::vector<double>& yu) {
    const std::size_t m = orig.matrix.row_count, n = orig.matrix.column_count;
    const double sense = orig.objective_sense == model::ObjectiveSense::maximize ? -1.0 : 1.0;
    xu.resize(n);
    yu.resize(m);
    for (std::size_t j = 0; j < n; ++j) xu[j] = ruiz ? xs[j] * sc.col_scale[j] : xs[j];
    for (std::size_t i = 0; i < m; ++i) yu[i] = ruiz ? ys[i] * sc.row_scale[i] : ys[i];
    auto rel_violation = [](double v, double lo, double hi) {
        const double p = std::clamp(v, lo, hi);
        return std::abs(v - p) / (1.0 + std::max(std::abs(v), std::abs(p) < 1e299 ? std::abs(p) : 0.0));
    };
    Residuals r;
    double worst = 0.0;
    for (std::size_t i = 0; i < m; ++i) {
        const double a = ruiz ? axs[i] / sc.row_scale[i] : axs[i];
        worst = std::max(worst, rel_violation(a, lower_of(orig.row_lower[i]), upper_of(orig.row_upper[i])));
    }
    for (std::size_t j = 0; j < n; ++j)
        worst = std::max(worst, rel_violation(xu[j], lower_of(orig.variable_lower[j]), upper_of(orig.variable_upper[j])));
    r.primal = worst;
    std::vector<double> g(n);
    double cmax = 1.0, sq = 0.0;
    for (std::size_t j = 0; j < n; ++j) {
        const double corig = sense * orig.objective[j];
        cmax = std::max(cmax, std::abs(corig));
        g[j] = corig + (ruiz ? atys[j] / sc.col_scale[j] : atys[j]);
        const double d = xu[j] - std::clamp(xu[j] - g[j], lower_of(orig.variable_lower[j]), upper_of(orig.variable_upper[j]));
        sq += d * d;
    }
    r.dual = std::sqrt(sq) / cmax;
    double pobj = orig.objective_offset;
    for (std::size_t j = 0; j < n; ++j) pobj += orig.objective[j] * xu[j];
    r.objective = pobj;
    double dobj = orig.objective_offset;
    for (std::size_t i = 0; i < m; ++i) {
        const double lo = lower_of(orig.row_lower[i]), hi = upper_of(orig.row_upper[i]);
        if (yu[i] > 0.0 && hi < 1e299)
            dobj -= yu[i] * hi;
        else if (yu[i] < 0.0 && lo > -1e299)
            dobj -= yu[i] * lo;
    }
    for (std::size_t j = 0; j < n; ++j) {
        const double lo = lower_of(orig.variable_lower[j]), hi = upper_of(orig.variable_upper[j]);
        if (g[j] > 0.0 && lo > -1e299)
            dobj += g[j] * lo;
        else if (g[j] < 0.0 && hi < 1e299)
            dobj += g[j] * hi;
    }
    r.gap = std::abs(pobj - dobj) / (1.0 + std::abs(pobj) + std::abs(dobj));
    r.score = std::max({r.primal, r.dual, r.gap});
    return r;
}

}  // namespace

// Infeasibility certificate test (extension) on the scaled movement (dx, dy)
// of the candidate since the last restart, in the unscaled space: a dual ray
// with positive homogeneous dual objective (primal infeasible) or a primal
// ray with negative objective inside the recession cones (dual infeasible),
// violations at most tol x |ray objective|, objective beyond 1e-9 of its terms'
// magnitudes.  Same tests as the device engine's k_ray_* kernels.
bool ray_certificate(const model::Model& orig, const Pdhg& p, const std::vector<double>& dx,
                     const std::vector<double>& dy, const scale::RuizScalers& sc, bool ruiz, double tol) {
    const std::size_t m = p.m, n = p.n;
    const double sense = orig.objective_sense == model::ObjectiveSense::maximize ? -1.0 : 1.0;
    std::vector<double> adx(m), atdy(n);
    p.a_times(dx, adx);
    p.at_times(dy, atdy);
    double vp = 0.0, cdx = 0.0, cab = 0.0, dob = 0.0, dab = 0.0, vd = 0.0, nx = 0.0, ny = 0.0;
    const double inf = std::numeric_limits<double>::infinity();
    for (std::size_t i = 0; i < m; ++i) {
        const double rs = ruiz ? sc.row_scale[i] : 1.0;
        const double lo = orig.row_lower[i].kind == model::BoundKind::finite ? orig.row_lower[i].value : -inf;
        const double hi = orig.row_upper[i].kind == model::BoundKind::finite ? orig.row_upper[i].value : inf;
        const double u = adx[i] / rs;
        if (lo > -inf && u < 0.0) vp = std::max(vp, -u);
        if (hi < inf && u > 0.0) vp = std::max(vp, u);
        const double yu = dy[i] * rs;
        double term = 0.0;
        if (yu > 0.0) {
            if (hi < inf) term = -(yu * hi); else vd += yu;
        } else if (yu < 0.0) {
            if (lo > -inf) term = -(yu * lo); else vd -= yu;
        }
        dob += term;
        dab += std::abs(term);
        ny = std::max(ny, std::abs(yu));
    }
    for (std::size_t j = 0; j < n; ++j) {
        const double cs = ruiz ? sc.col_scale[j] : 1.0;
        const double lo =
            orig.variable_lower[j].kind == model::BoundKind::finite ? orig.variable_lower[j].value : -inf;
        const double hi =
            orig.variable_upper[j].kind == model::BoundKind::finite ? orig.variable_upper[j].value : inf;
        const double g = atdy[j] / cs;
        double term = 0.0;
        if (g > 0.0) {
            if (lo > -inf) term = g * lo; else vd += g;
        } else if (g < 0.0) {
            if (hi < inf) term = g * hi; else vd -= g;
        }
        dob += term;
        dab += std::abs(term);
        const double xu = dx[j] * cs;
        if (lo > -inf && xu < 0.0) vp = std::max(vp, -xu);
        if (hi < inf && xu > 0.0) vp = std::max(vp, xu);
        const double cterm = sense * orig.objective[j] * xu;
        cdx += cterm;
        cab += std::abs(cterm);
        nx = std::max(nx, std::abs(xu));
    }
    const bool primal_infeasible = ny > 0.0 && dob > 1e-9 * dab && vd <= tol * dob;
    const bool dual_infeasible = nx > 0.0 && cdx < -1e-9 * cab && vp <= tol * -cdx;
    return primal_infeasible || dual_infeasible;
}

PdlpResult solve_cpu(const model::Model& orig, const PdlpOptions& opt, const char* backend_label) {
    const auto t0 = std::chrono::steady_clock::now();
    const std::size_t m = orig.matrix.row_count, n = orig.matrix.column_count;
    PdlpResult res;
    res.backend_actually_used = backend_label;
    res.tolerance = opt.primal_tolerance;
    if (n == 0 || m == 0) {
        res.status = PdlpStatus::optimal;
        res.message = "trivial";
        return res;
    }
    model::Model s = orig;
    scale::RuizScalers sc;
    if (opt.ruiz_scaling) sc = scale::equilibrate_model(s, {opt.ruiz_iterations});
    const bool scaled = opt.ruiz_scaling || opt.row_weighting > 0;
    if (opt.row_weighting > 0) apply_row_weights(orig, s, sc, opt.row_weighting);
    Pdhg p(s);
    const double sense = s.objective_sense == model::ObjectiveSense::maximize ? -1.0 : 1.0;
    p.c.resize(n);
    p.lo.resize(n);
    p.hi.resize(n);
    double cinf = 1.0, binf = 1.0;
    for (std::size_t j = 0; j < n; ++j) {
        p.c[j] = sense * s.objective[j];
        cinf = std::max(cinf, std::abs(p.c[j]));
        p.lo[j] = lower_of(s.variable_lower[j]);
        p.hi[j] = upper_of(s.variable_upper[j]);
    }
    p.blo.resize(m);
    p.bhi.resize(m);
    for (std::size_t i = 0; i < m; ++i) {
        p.blo[i] = lower_of(s.row_lower[i]);
        p.bhi[i] = upper_of(s.row_upper[i]);
        if (std::abs(p.blo[i]) < 1e299) binf = std::max(binf, std::abs(p.blo[i]));
        if (std::abs(p.bhi[i]) < 1e299) binf = std::max(binf, std::abs(p.bhi[i]));
    }
    p.rnorm.assign(m, 0.0);
    p.cnorm.assign(n, 0.0);
    for (std::size_t j = 0; j < n; ++j)
        for (std::size_t k = s.matrix.column_start[j]; k < s.matrix.column_start[j + 1]; ++k) {
            const double a = std::abs(s.matrix.value[k]);
            p.cnorm[j] += a;
            p.rnorm[s.matrix.row_index[k]] += a;
        }
    for (auto& v : p.cnorm)
        if (v < 1e-12) v = 1.0;
    for (auto& v : p.rnorm)
        if (v < 1e-12) v = 1.0;
    p.omega = opt.initial_primal_weight > 0.0 ? opt.initial_primal_weight
                                              : std::clamp(std::sqrt(cinf / binf), 0.01, 100.0);
    p.eta = std::clamp(opt.step_size_reduction, 0.1, 0.99);
    p.x.resize(n);
    for (std::size_t j = 0; j < n; ++j) {
        const double start = opt.warm_primal.size() == n && scaled ? opt.warm_primal[j] / sc.col_scale[j]
                             : opt.warm_primal.size() == n                 ? opt.warm_primal[j]
                                                                           : 0.0;
        p.x[j] = std::clamp(start, p.lo[j], p.hi[j]);
    }
    p.y.assign(m, 0.0);
    if (opt.warm_dual.size() == m)
        for (std::size_t i = 0; i < m; ++i) p.y[i] = scaled ? opt.warm_dual[i] / sc.row_scale[i] : opt.warm_dual[i];
    p.xbar = p.x;
    p.xavg = p.x;
    p.yavg = p.y;
    p.dx.assign(n, 0.0);
    p.aty.assign(n, 0.0);
    p.ax.assign(m, 0.0);
    std::vector<double> xprev(n), xu, yu;

    const bool halpern = opt.restart_scheme == RestartScheme::halpern;
    const bool kkt = opt.restart_scheme == RestartScheme::kkt || halpern;
    const double eta_cap =
        (halpern && opt.adaptive_step_size && opt.power_iterations > 0) ? p.step_cap(opt.power_iterations) : 0.99;
    std::vector<double> yprev(halpern ? m : 0);
    std::vector<double> xlast = p.x, ylast = p.y, ax2(m), aty2(n), xu2, yu2;
    std::size_t iter = 0, k = 0, since = 0;
    double last_score = 1e300, prev_cand = 1e300;
    Residuals cur;
    bool converged = false, use_cur = false, infeasible = false;
    // Average's residuals; for the KKT scheme also the current iterate's, and
    // the better of the two becomes the candidate (cur, xu, yu).
    auto evaluate_candidates = [&] {
        p.a_times(p.xavg, p.ax);
        p.at_times(p.yavg, p.aty);
        cur = evaluate(orig, p.xavg, p.yavg, p.ax, p.aty, sc, scaled, xu, yu);
        use_cur = false;
        if (!kkt) return;
        p.a_times(p.x, ax2);
        p.at_times(p.y, aty2);
        const Residuals alt = evaluate(orig, p.x, p.y, ax2, aty2, sc, scaled, xu2, yu2);
        if (alt.score < cur.score) {
            use_cur = true;
            cur = alt;
            std::swap(xu, xu2);
            std::swap(yu, yu2);
        }
    };
    while (iter < opt.max_iterations) {
        p.at_times(p.y, p.aty);
        xprev = p.x;
        double dxn = 0.0;
        for (std::size_t j = 0; j < n; ++j) {
            const double xn = std::clamp(p.x[j] - p.tau(j) * (p.c[j] + p.aty[j]), p.lo[j], p.hi[j]);
            const double d = xn - p.x[j];
            p.dx[j] = d;
            dxn += p.cnorm[j] * d * d;
            p.x[j] = xn;
        }
        for (std::size_t j = 0; j < n; ++j) p.xbar[j] = 2.0 * p.x[j] - xprev[j];
        if (opt.adaptive_step_size && iter % 10 == 0 && dxn > 1e-14) {
            p.a_times(p.dx, p.ax);
            double adxn = 0.0;
            for (std::size_t i = 0; i < m; ++i) adxn += (p.ax[i] * p.ax[i]) / p.rnorm[i];
            if (adxn > 1e-14) {
                const double L = std::sqrt(adxn / dxn);
                if (L > 1e-6) {
                    const double target = 0.95 / L;
                    if (target < p.eta)
                        p.eta = std::max(0.1, std::max(target, p.eta * 0.8));
                    else if (target > 1.05 * p.eta && p.eta < eta_cap)
                        p.eta = std::min(eta_cap, p.eta * 1.05);
                }
            }
        }
        p.a_times(p.xbar, p.ax);
        if (halpern) yprev = p.y;
        for (std::size_t i = 0; i < m; ++i) {
            const double sg = p.sigma(i);
            const double v = p.y[i] + sg * p.ax[i];
            p.y[i] = v - sg * std::clamp(v / sg, p.blo[i], p.bhi[i]);
        }
        ++k;
        const double kd = static_cast<double>(k);
        if (halpern) {  // T(z) into the average slot; z <- w (2 T(z) - z) + (1 - w) z0
            const double w = kd / (kd + 1.0);
            for (std::size_t j = 0; j < n; ++j) {
                p.xavg[j] = p.x[j];
                p.x[j] = w * p.xbar[j] + (1.0 - w) * xlast[j];
            }
            for (std::size_t i = 0; i < m; ++i) {
                p.yavg[i] = p.y[i];
                p.y[i] = w * (2.0 * p.y[i] - yprev[i]) + (1.0 - w) * ylast[i];
            }
        } else {
            for (std::size_t j = 0; j < n; ++j) p.xavg[j] += (p.x[j] - p.xavg[j]) / kd;
            for (std::size_t i = 0; i < m; ++i) p.yavg[i] += (p.y[i] - p.yavg[i]) / kd;
        }
        ++iter;
        ++since;
        if (iter % opt.restart_every != 0) continue;
        evaluate_candidates();
        if (cur.primal <= opt.primal_tolerance && cur.dual <= opt.dual_tolerance && cur.gap <= opt.gap_tolerance) {
            converged = true;
            break;
        }
        if (opt.detect_infeasibility) {
            const auto& xc = use_cur ? p.x : p.xavg;
            const auto& yc = use_cur ? p.y : p.yavg;
            std::vector<double> rdx(n), rdy(m);
            for (std::size_t j = 0; j < n; ++j) rdx[j] = xc[j] - xlast[j];
            for (std::size_t i = 0; i < m; ++i) rdy[i] = yc[i] - ylast[i];
            if (ray_certificate(orig, p, rdx, rdy, sc, scaled, opt.infeasibility_tolerance)) {
                infeasible = true;
                break;
            }
        }
        bool restart;
        if (!kkt) {
            restart = opt.restart_strategy == RestartStrategy::fixed ||
                      (opt.restart_strategy == RestartStrategy::adaptive &&
                       (cur.score <= opt.restart_reduction_factor * last_score ||
                        (since >= 5 * opt.restart_every && cur.score < last_score)));
        } else {
            restart = opt.restart_strategy == RestartStrategy::fixed ||
                      (opt.restart_strategy == RestartStrategy::adaptive &&
                       (cur.score <= 0.2 * last_score || (cur.score <= 0.8 * last_score && cur.score > prev_cand) ||
                        static_cast<double>(since) >= 0.36 * static_cast<double>(iter)));
            prev_cand = cur.score;
        }
        if (!restart) continue;
        if (!kkt) {
            if (opt.adaptive_primal_weight && cur.primal > 1e-12 && cur.dual > 1e-12) {
                const double ratio = std::clamp(std::sqrt(cur.primal / cur.dual), 0.05, 20.0);
                p.omega = std::clamp(p.omega * std::pow(ratio, opt.primal_weight_smoothing), 1e-6, 1e6);
            }
            p.x = p.xavg;
            p.y = p.yavg;
        } else {
            const auto& xc = use_cur ? p.x : p.xavg;
            const auto& yc = use_cur ? p.y : p.yavg;
            if (opt.adaptive_primal_weight) {
                double dxs = 0.0, dys = 0.0;
                for (std::size_t j = 0; j < n; ++j) dxs += (xc[j] - xlast[j]) * (xc[j] - xlast[j]);
                for (std::size_t i = 0; i < m; ++i) dys += (yc[i] - ylast[i]) * (yc[i] - ylast[i]);
                const double dX = std::sqrt(dxs), dY = std::sqrt(dys);
                double w = p.omega;
                bool upd = false;
                if (dX > 1e-10 && dY > 1e-10) {
                    const double th = opt.primal_weight_smoothing, ratio = dY / dX;
                    if (halpern) {  // ratio^(127/128) omega^(1/128) in square roots
                        double f = p.omega / ratio;
                        for (int q = 0; q < 7; ++q) f = std::sqrt(f);
                        w = ratio * f;
                    } else {
                        w = th == 0.5 ? std::sqrt(p.omega * ratio) : std::pow(ratio, th) * std::pow(p.omega, 1.0 - th);
                    }
                    upd = true;
                }
                if (halpern) {  // omega_move^(3/4) omega_res^(1/4), in square roots
                    double wres = p.omega;
                    if (cur.primal > 1e-12 && cur.dual > 1e-12)
                        wres = p.omega * std::sqrt(std::clamp(std::sqrt(cur.primal / cur.dual), 0.05, 20.0));
                    w = std::sqrt(w * std::sqrt(w * wres));
                    upd = true;
                    // no move against a residual imbalance of 100x or more
                    if (cur.primal > 1e-12 && cur.dual > 1e-12) {
                        if (cur.primal > 100.0 * cur.dual && w < p.omega) w = p.omega;
                        if (cur.dual > 100.0 * cur.primal && w > p.omega) w = p.omega;
                    }
                    // at most a fall by primal_weight_max_drop (a power of two: exact)
                    if (opt.primal_weight_max_drop > 0) w = std::max(w, p.omega / opt.primal_weight_max_drop);
                }
                if (upd) p.omega = std::clamp(w, 1e-6, 1e6);
            }
            if (!use_cur) {
                p.x = p.xavg;
                p.y = p.yavg;
            }
            xlast = p.x;
            ylast = p.y;
            prev_cand = 1e300;
        }
        p.xbar = p.x;
        k = 0;
        since = 0;
        last_score = cur.score;
    }
    if (!converged && !infeasible) evaluate_candidates();
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    res.status = converged ? PdlpStatus::optimal
                 : infeasible ? PdlpStatus::infeasible_or_unbounded
                              : PdlpStatus::iteration_limit;
    res.message = converged ? "PDLP converged" : infeasible ? "infeasibility certificate found" : "iteration limit reached";
    res.primal = std::move(xu);
    res.dual = std::move(yu);
    res.objective = cur.objective;
    res.primal_infeasibility = cur.primal;
    res.dual_infeasibility = cur.dual;
    res.duality_gap = cur.gap;
    res.iterations = iter;
    res.primal_weight = p.omega;
    res.step_size = p.eta;
    res.kernel_ms = ms;
    res.total_ms = ms;
    return res;
}

}  // namespace markov_cero::lp::first_order::detail
