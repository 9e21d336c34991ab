// generate.cpp -- seeded planted-feasible LP families for the benchmark
// shapes.  Every instance is built around a planted primal/dual pair
// (x*, y*, s) satisfying the KKT conditions of
//     min c^T x  s.t.  row_lower <= A x <= row_upper,  lo <= x <= hi
// by construction:  b = A x* on active rows, c = A^T y* + s with s >= 0 where
// x* sits at its lower bound, s <= 0 at its upper bound, s = 0 inside;
// y* free on equality rows, y* <= 0 on active '<=' rows, 0 on slack rows.
// So c^T x* is the optimum (up to the rounding of b).  Distributions are
// written out (53-bit uniforms, Box-Muller normals, inverse-CDF Zipf) on top
// of std::mt19937_64, whose output sequence the standard fixes.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mcpdlp.h"
#include "markov_cero/io/mps.hpp"

struct mc_lp_handle {
    int64_t m = 0, n = 0;
    std::vector<int64_t> cs, ri;
    std::vector<double> val, obj;
    std::vector<mc_bound> rlo, rhi, vlo, vhi;
    int32_t sense = MC_MINIMIZE;
    double offset = 0.0;
    // planted data
    std::vector<double> xs, ys, s;
    std::vector<int8_t> at;        // -1 lower, 0 interior, +1 upper
    std::vector<int8_t> row_kind;  // 0 equality, 1 active '<=', 2 slack '<='
    std::vector<double> slack;
    double planted_obj = 0.0;
};

namespace {

class Rng {
  public:
    explicit Rng(uint64_t seed) : g_(seed) {}
    double u01() { return (double)(g_() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double a, double b) { return a + (b - a) * u01(); }
    double normal() {
        const double u1 = 1.0 - u01();  // (0, 1]
        const double u2 = u01();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    int64_t below(int64_t k) { return (int64_t)(g_() % (uint64_t)k); }
    double sign() { return (g_() >> 63) ? -1.0 : 1.0; }

  private:
    std::mt19937_64 g_;
};

mc_bound fin(double v) { return mc_bound{MC_BOUND_FINITE, 0, v}; }
mc_bound ninf() { return mc_bound{MC_BOUND_NEGATIVE_INFINITY, 0, 0.0}; }

// k distinct rows of [lo, lo+span), sorted.
void pick_rows(Rng& r, int64_t lo, int64_t span, int k, std::vector<int64_t>& out) {
    out.clear();
    if (k >= span) {
        for (int64_t i = 0; i < span; ++i) out.push_back(lo + i);
        return;
    }
    while ((int)out.size() < k) {
        const int64_t c = lo + r.below(span);
        if (std::find(out.begin(), out.end(), c) == out.end()) out.push_back(c);
    }
    std::sort(out.begin(), out.end());
}

// Plant x* inside [0, hi] and the reduced cost s; fills at/xs/s.
void plant_columns(Rng& r, mc_lp_handle& h, double at_bound) {
    const int64_t n = h.n;
    h.xs.resize(n);
    h.s.resize(n);
    h.at.resize(n);
    for (int64_t j = 0; j < n; ++j) {
        const double u = r.u01();
        const double hi = h.vhi[j].value;
        if (u < 0.5 * at_bound) {
            h.at[j] = -1;
            h.xs[j] = 0.0;
            h.s[j] = std::abs(r.normal());
        } else if (u < at_bound) {
            h.at[j] = 1;
            h.xs[j] = hi;
            h.s[j] = -std::abs(r.normal());
        } else {
            h.at[j] = 0;
            h.xs[j] = r.uniform(0.0, hi);
            if (h.xs[j] <= 0.0) h.xs[j] = 0.5 * hi;
            h.s[j] = 0.0;
        }
    }
}

// b from x* (row kinds), c = A^T y* + s, planted objective.
void finish_planted(mc_lp_handle& h) {
    const int64_t m = h.m, n = h.n;
    std::vector<long double> ax(m, 0.0L);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t k = h.cs[j]; k < h.cs[j + 1]; ++k) ax[h.ri[k]] += (long double)h.val[k] * h.xs[j];
    h.rlo.resize(m);
    h.rhi.resize(m);
    for (int64_t i = 0; i < m; ++i) {
        const double a = (double)ax[i];
        if (h.row_kind[i] == 0) {
            h.rlo[i] = fin(a);
            h.rhi[i] = fin(a);
        } else if (h.row_kind[i] == 1) {
            h.rlo[i] = ninf();
            h.rhi[i] = fin(a);
        } else {
            h.rlo[i] = ninf();
            h.rhi[i] = fin(a + h.slack[i] * (1.0 + std::abs(a)));
        }
    }
    h.obj.assign(n, 0.0);
    long double po = 0.0L;
    for (int64_t j = 0; j < n; ++j) {
        long double cj = h.s[j];
        for (int64_t k = h.cs[j]; k < h.cs[j + 1]; ++k) cj += (long double)h.val[k] * h.ys[h.ri[k]];
        h.obj[j] = (double)cj;
        po += (long double)h.obj[j] * h.xs[j];
    }
    h.planted_obj = (double)po;
}

void set_rows(Rng& r, mc_lp_handle& h, double eq_fraction) {
    h.row_kind.assign(h.m, 0);
    h.slack.assign(h.m, 0.0);
    h.ys.assign(h.m, 0.0);
    for (int64_t i = 0; i < h.m; ++i) {
        if (eq_fraction >= 1.0 || r.u01() < eq_fraction) {
            h.row_kind[i] = 0;
            h.ys[i] = r.normal();
        } else if (r.u01() < 0.5) {
            h.row_kind[i] = 1;
            h.ys[i] = -std::abs(r.normal());
        } else {
            h.row_kind[i] = 2;
            h.slack[i] = 1.0 - r.u01();  // (0, 1]
            h.ys[i] = 0.0;
        }
    }
}

void set_uppers(Rng& r, mc_lp_handle& h, double ulo, double uhi) {
    h.vlo.assign(h.n, fin(0.0));
    h.vhi.resize(h.n);
    for (int64_t j = 0; j < h.n; ++j) h.vhi[j] = fin(uhi > ulo ? r.uniform(ulo, uhi) : ulo);
}

void gen_uniform(const mc_gen_spec& sp, mc_lp_handle& h) {
    Rng r(sp.seed);
    h.m = sp.rows;
    h.n = sp.cols;
    set_uppers(r, h, sp.upper_lo, sp.upper_hi);
    h.cs.assign(h.n + 1, 0);
    std::vector<int64_t> rows;
    const int lo = std::max(1, sp.nnz_min), hi = std::max(lo, sp.nnz_max);
    for (int64_t j = 0; j < h.n; ++j) {
        const int k = lo + (int)r.below(hi - lo + 1);
        pick_rows(r, 0, h.m, k, rows);
        for (int64_t i : rows) {
            h.ri.push_back(i);
            h.val.push_back(r.sign() * r.uniform(0.1, 10.0));
        }
        h.cs[j + 1] = (int64_t)h.ri.size();
    }
    plant_columns(r, h, sp.at_bound_fraction);
    set_rows(r, h, sp.equality_fraction);
    finish_planted(h);
}

double log_uniform_value(Rng& r) { return r.sign() * std::pow(10.0, r.uniform(-2.0, 2.0)); }

void gen_zipf(const mc_gen_spec& sp, mc_lp_handle& h) {
    Rng r(sp.seed);
    h.m = sp.rows;
    h.n = sp.cols;
    set_uppers(r, h, sp.upper_lo, sp.upper_hi);
    const int kmax = std::max(1, sp.nnz_max);
    std::vector<double> cdf(kmax);
    double tot = 0.0;
    for (int k = 1; k <= kmax; ++k) {
        tot += std::pow((double)k, -sp.zipf_s);
        cdf[k - 1] = tot;
    }
    h.cs.assign(h.n + 1, 0);
    std::vector<int64_t> rows;
    for (int64_t j = 0; j < h.n; ++j) {
        const double u = r.u01() * tot;
        const int k = 1 + (int)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        pick_rows(r, 0, h.m, std::min(k, kmax), rows);
        for (int64_t i : rows) {
            h.ri.push_back(i);
            h.val.push_back(log_uniform_value(r));
        }
        h.cs[j + 1] = (int64_t)h.ri.size();
    }
    plant_columns(r, h, sp.at_bound_fraction);
    set_rows(r, h, sp.equality_fraction);
    finish_planted(h);
}

// periods x (rows x cols) blocks, nnz_max per column inside its block, plus
// link_rows '<=' rows coupling periods t and t+1 (link_nnz per column).
void gen_block(const mc_gen_spec& sp, mc_lp_handle& h) {
    Rng r(sp.seed);
    const int P = std::max(1, sp.periods);
    const int64_t mb = sp.rows, nb = sp.cols, L = P > 1 ? sp.link_rows : 0;
    h.m = P * mb + L;
    h.n = P * nb;
    set_uppers(r, h, sp.upper_lo, sp.upper_hi);
    h.cs.assign(h.n + 1, 0);
    h.ri.reserve((size_t)(h.n * (sp.nnz_max + sp.link_nnz)));
    h.val.reserve(h.ri.capacity());
    std::vector<int64_t> rows, link;
    for (int t = 0; t < P; ++t)
        for (int64_t jj = 0; jj < nb; ++jj) {
            const int64_t j = t * nb + jj;
            pick_rows(r, t * mb, mb, sp.nnz_max, rows);
            link.clear();
            if (L > 0) {
                int bnd = t == 0 ? 0 : (t == P - 1 ? P - 2 : t - 1 + (int)r.below(2));
                const int64_t l0 = L * bnd / (P - 1), l1 = L * (bnd + 1) / (P - 1);
                pick_rows(r, P * mb + l0, l1 - l0, sp.link_nnz, link);
            }
            for (int64_t i : rows) {
                h.ri.push_back(i);
                h.val.push_back(log_uniform_value(r));
            }
            for (int64_t i : link) {
                h.ri.push_back(i);
                h.val.push_back(log_uniform_value(r));
            }
            h.cs[j + 1] = (int64_t)h.ri.size();
        }
    plant_columns(r, h, sp.at_bound_fraction);
    // block rows are equalities, linking rows '<=' (half active)
    h.row_kind.assign(h.m, 0);
    h.slack.assign(h.m, 0.0);
    h.ys.assign(h.m, 0.0);
    for (int64_t i = 0; i < h.m; ++i) {
        if (i < P * mb) {
            h.ys[i] = r.normal();
        } else if (r.u01() < 0.5) {
            h.row_kind[i] = 1;
            h.ys[i] = -std::abs(r.normal());
        } else {
            h.row_kind[i] = 2;
            h.slack[i] = 1.0 - r.u01();
        }
    }
    finish_planted(h);
}

}  // namespace

void mc_internal_set_error(const std::string& msg);

extern "C" {

void mc_gen_spec_default(mc_gen_spec* sp, int32_t family) {
    std::memset(sp, 0, sizeof(*sp));
    sp->family = family;
    sp->seed = 0x4D4350444C50ULL;
    sp->at_bound_fraction = 0.6;
    sp->upper_lo = sp->upper_hi = 10.0;
    sp->zipf_s = 1.2;
    switch (family) {
        case MC_FAMILY_ZIPF_MIXED:  // single large LP
            sp->rows = 200000;
            sp->cols = 500000;
            sp->nnz_min = 1;
            sp->nnz_max = 200;
            sp->equality_fraction = 0.7;
            break;
        case MC_FAMILY_BLOCK_ANGULAR:
            sp->rows = 40000;
            sp->cols = 100000;
            sp->nnz_max = 8;
            sp->periods = 48;
            sp->link_rows = 200000;
            sp->link_nnz = 2;
            sp->equality_fraction = 1.0;
            break;
        default:  // MC_FAMILY_UNIFORM_EQ: the m=2000, n=5000 CPU-runnable LP
            sp->family = MC_FAMILY_UNIFORM_EQ;
            sp->rows = 2000;
            sp->cols = 5000;
            sp->nnz_min = sp->nnz_max = 5;
            sp->equality_fraction = 1.0;
            break;
    }
}

int mc_lp_generate(const mc_gen_spec* sp, mc_lp_handle** out) {
    *out = nullptr;
    try {
        if (!sp || sp->rows < 0 || sp->cols < 0) throw std::invalid_argument("bad generator spec");
        auto h = std::make_unique<mc_lp_handle>();
        switch (sp->family) {
            case MC_FAMILY_UNIFORM_EQ: gen_uniform(*sp, *h); break;
            case MC_FAMILY_ZIPF_MIXED: gen_zipf(*sp, *h); break;
            case MC_FAMILY_BLOCK_ANGULAR: gen_block(*sp, *h); break;
            default: throw std::invalid_argument("unknown generator family");
        }
        *out = h.release();
        return MC_OK;
    } catch (const std::invalid_argument& e) {
        mc_internal_set_error(e.what());
        return MC_ERR_INVALID_ARGUMENT;
    } catch (const std::bad_alloc&) {
        mc_internal_set_error("out of memory");
        return MC_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        mc_internal_set_error(e.what());
        return MC_ERR_INTERNAL;
    }
}

int mc_lp_view(const mc_lp_handle* h, mc_lp* v) {
    if (!h || !v) return MC_ERR_INVALID_ARGUMENT;
    std::memset(v, 0, sizeof(*v));
    v->rows = h->m;
    v->cols = h->n;
    v->column_start = h->cs.data();
    v->row_index = h->ri.data();
    v->value = h->val.data();
    v->objective = h->obj.data();
    v->row_lower = h->rlo.data();
    v->row_upper = h->rhi.data();
    v->var_lower = h->vlo.data();
    v->var_upper = h->vhi.data();
    v->sense = h->sense;
    v->objective_offset = h->offset;
    return MC_OK;
}

int mc_lp_planted(const mc_lp_handle* h, const double** x_star, const double** y_star, double* planted) {
    if (!h || h->at.size() != (size_t)h->n) {
        mc_internal_set_error("mc_lp_planted needs a generated (planted) instance");
        return MC_ERR_INVALID_ARGUMENT;
    }
    if (x_star) *x_star = h->xs.data();
    if (y_star) *y_star = h->ys.data();
    if (planted) *planted = h->planted_obj;
    return MC_OK;
}

void mc_lp_free(mc_lp_handle* h) { delete h; }

int mc_lp_perturb(const mc_lp_handle* base, uint64_t seed, double sigma, mc_lp_handle** out) {
    *out = nullptr;
    if (!base || base->at.size() != (size_t)base->n) {
        mc_internal_set_error("mc_lp_perturb needs a generated (planted) base instance");
        return MC_ERR_INVALID_ARGUMENT;
    }
    try {
        auto h = std::make_unique<mc_lp_handle>(*base);
        Rng r(seed);
        auto ln = [&] { return std::exp(sigma * r.normal()); };
        for (int64_t j = 0; j < h->n; ++j) {
            const double u = base->vhi[j].value * r.uniform(0.8, 1.2);
            h->vhi[j] = fin(u);
            if (h->at[j] == -1) {
                h->xs[j] = 0.0;
            } else if (h->at[j] == 1) {
                h->xs[j] = u;
            } else {
                double x = base->xs[j] * ln();
                if (!(x > 0.0 && x < u)) x = 0.5 * u;
                h->xs[j] = x;
            }
            h->s[j] = base->s[j] * ln();
        }
        for (int64_t i = 0; i < h->m; ++i) h->ys[i] = base->ys[i] * ln();
        finish_planted(*h);
        *out = h.release();
        return MC_OK;
    } catch (const std::exception& e) {
        mc_internal_set_error(e.what());
        return MC_ERR_INTERNAL;
    }
}

int mc_lp_read_mps(const char* path, mc_lp_handle** out) {
    if (!out) return MC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!path) {
        mc_internal_set_error("mc_lp_read_mps: null path");
        return MC_ERR_INVALID_ARGUMENT;
    }
    try {
        const auto md = markov_cero::io::read_mps_file(path);
        auto h = std::make_unique<mc_lp_handle>();
        h->m = (int64_t)md.matrix.row_count;
        h->n = (int64_t)md.matrix.column_count;
        h->cs.assign(md.matrix.column_start.begin(), md.matrix.column_start.end());
        h->ri.assign(md.matrix.row_index.begin(), md.matrix.row_index.end());
        h->val = md.matrix.value;
        h->obj = md.objective;
        auto conv = [](const std::vector<markov_cero::model::Bound>& v) {
            std::vector<mc_bound> o(v.size());
            for (size_t k = 0; k < v.size(); ++k) {
                o[k].kind = static_cast<int32_t>(v[k].kind);
                o[k].pad_ = 0;
                o[k].value = v[k].value;
            }
            return o;
        };
        h->rlo = conv(md.row_lower);
        h->rhi = conv(md.row_upper);
        h->vlo = conv(md.variable_lower);
        h->vhi = conv(md.variable_upper);
        h->sense = md.objective_sense == markov_cero::model::ObjectiveSense::maximize ? MC_MAXIMIZE : MC_MINIMIZE;
        h->offset = md.objective_offset;
        *out = h.release();
        return MC_OK;
    } catch (const std::bad_alloc&) {
        mc_internal_set_error("out of memory");
        return MC_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        mc_internal_set_error(std::string("mc_lp_read_mps: ") + e.what());
        return MC_ERR_INVALID_ARGUMENT;
    }
}

}  // extern "C"
