// capi.cpp -- the C-ABI of include/mcpdlp.h over the C++ API.  No exception
// crosses it: each entry point returns MC_OK or an MC_ERR_* code and keeps
// the message for mc_last_error() (thread-local).
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/mcpdlp.h"
#include "markov_cero/lp/first_order/pdlp.hpp"
#include "markov_cero/verify/primal_verifier.hpp"
#include "pdlp_internal.hpp"

namespace fo = markov_cero::lp::first_order;

namespace {
thread_local std::string g_error;

template <class F>
int guarded(F&& f) {
    try {
        g_error.clear();
        f();
        return MC_OK;
    } catch (const std::invalid_argument& e) {
        g_error = e.what();
        return MC_ERR_INVALID_ARGUMENT;
    } catch (const std::bad_alloc&) {
        g_error = "out of memory";
        return MC_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        g_error = e.what();
        return std::string(e.what()).find("GPU") != std::string::npos ? MC_ERR_DEVICE : MC_ERR_INTERNAL;
    }
}

fo::PdlpOptions to_options(const mc_pdlp_options& o, int64_t n, int64_t m) {
    fo::PdlpOptions p;
    p.max_iterations = o.max_iterations;
    p.restart_every = o.restart_every;
    p.primal_tolerance = o.primal_tolerance;
    p.dual_tolerance = o.dual_tolerance;
    p.gap_tolerance = o.gap_tolerance;
    p.step_size_reduction = o.step_size_reduction;
    p.power_iterations = o.power_iterations;
    p.backend = o.backend == MC_BACKEND_GPU ? fo::Backend::gpu : fo::Backend::cpu;
    p.restart_strategy = o.restart_strategy == MC_RESTART_NONE    ? fo::RestartStrategy::none
                         : o.restart_strategy == MC_RESTART_FIXED ? fo::RestartStrategy::fixed
                                                                  : fo::RestartStrategy::adaptive;
    p.restart_reduction_factor = o.restart_reduction_factor;
    p.adaptive_step_size = o.adaptive_step_size != 0;
    p.adaptive_primal_weight = o.adaptive_primal_weight != 0;
    p.initial_primal_weight = o.initial_primal_weight;
    p.primal_weight_smoothing = o.primal_weight_smoothing;
    p.ruiz_scaling = o.ruiz_scaling != 0;
    p.ruiz_iterations = o.ruiz_iterations;
    p.autotune = o.autotune != 0;
    p.precision = o.precision == MC_PRECISION_MIXED        ? fo::Precision::mixed
                  : o.precision == MC_PRECISION_PACKED ? fo::Precision::packed
                                                           : fo::Precision::fp64;
    p.reduction_order = o.reduction_order == MC_ORDER_REFERENCE ? fo::ReductionOrder::reference
                                                                : fo::ReductionOrder::fast;
    p.restart_scheme = o.restart_scheme == MC_SCHEME_KKT       ? fo::RestartScheme::kkt
                       : o.restart_scheme == MC_SCHEME_HALPERN ? fo::RestartScheme::halpern
                                                               : fo::RestartScheme::reference;
    p.polish_threshold = o.polish_threshold;
    p.fp32_stall_checks = o.fp32_stall_checks;
    p.detect_infeasibility = o.detect_infeasibility != 0;
    p.infeasibility_tolerance = o.infeasibility_tolerance > 0.0 ? o.infeasibility_tolerance : 1e-8;
    p.row_weighting = o.row_weighting;
    p.primal_weight_max_drop = o.primal_weight_max_drop;
    p.graph_chunk = o.graph_chunk;
    p.device = o.device;
    p.require_gpu = o.require_gpu != 0;
    if (o.warm_primal) p.warm_primal.assign(o.warm_primal, o.warm_primal + n);
    if (o.warm_dual) p.warm_dual.assign(o.warm_dual, o.warm_dual + m);
    return p;
}

void to_result(const fo::PdlpResult& r, mc_pdlp_result* out) {
    out->status = static_cast<int32_t>(r.status);
    out->objective = r.objective;
    out->primal_infeasibility = r.primal_infeasibility;
    out->dual_infeasibility = r.dual_infeasibility;
    out->duality_gap = r.duality_gap;
    out->tolerance = r.tolerance;
    out->iterations = r.iterations;
    out->h2d_ms = r.h2d_ms;
    out->kernel_ms = r.kernel_ms;
    out->d2h_ms = r.d2h_ms;
    out->total_ms = r.total_ms;
    std::memset(out->message, 0, sizeof(out->message));
    std::strncpy(out->message, r.message.c_str(), sizeof(out->message) - 1);
    std::memset(out->backend_actually_used, 0, sizeof(out->backend_actually_used));
    std::strncpy(out->backend_actually_used, r.backend_actually_used.c_str(), sizeof(out->backend_actually_used) - 1);
    out->primal_len = (int64_t)r.primal.size();
    out->dual_len = (int64_t)r.dual.size();
    if (out->primal && !r.primal.empty()) std::memcpy(out->primal, r.primal.data(), sizeof(double) * r.primal.size());
    if (out->dual && !r.dual.empty()) std::memcpy(out->dual, r.dual.data(), sizeof(double) * r.dual.size());
    out->fp32_iterations = r.fp32_iterations;
    out->packed_iterations = r.packed_iterations;
    out->setup_ms = r.setup_ms;
    out->primal_weight = r.primal_weight;
    out->step_size = r.step_size;
}

void check_lp(const mc_lp* lp) {
    if (!lp) throw std::invalid_argument("null model");
    if (lp->rows < 0 || lp->cols < 0) throw std::invalid_argument("negative model dimensions");
    if (lp->cols > 0 && (!lp->column_start || lp->column_start[0] != 0))
        throw std::invalid_argument("model: column_start[0] must be 0");
    const int64_t nnz = lp->cols > 0 ? lp->column_start[lp->cols] : 0;
    for (int64_t j = 0; j < lp->cols; ++j)
        if (lp->column_start[j] > lp->column_start[j + 1]) throw std::invalid_argument("model: column_start decreases");
    for (int64_t k = 0; k < nnz; ++k)
        if (lp->row_index[k] < 0 || lp->row_index[k] >= lp->rows)
            throw std::invalid_argument("model: row index out of range");
}

}  // namespace

void mc_internal_set_error(const std::string& msg) { g_error = msg; }

extern "C" {

void mc_pdlp_options_default(mc_pdlp_options* o) {
    std::memset(o, 0, sizeof(*o));
    const fo::PdlpOptions d;
    o->max_iterations = d.max_iterations;
    o->restart_every = d.restart_every;
    o->primal_tolerance = d.primal_tolerance;
    o->dual_tolerance = d.dual_tolerance;
    o->gap_tolerance = d.gap_tolerance;
    o->step_size_reduction = d.step_size_reduction;
    o->power_iterations = d.power_iterations;
    o->backend = MC_BACKEND_CPU;
    o->restart_strategy = MC_RESTART_ADAPTIVE;
    o->restart_reduction_factor = d.restart_reduction_factor;
    o->adaptive_step_size = 1;
    o->adaptive_primal_weight = 1;
    o->initial_primal_weight = d.initial_primal_weight;
    o->primal_weight_smoothing = d.primal_weight_smoothing;
    o->ruiz_scaling = 1;
    o->ruiz_iterations = d.ruiz_iterations;
    o->precision = MC_PRECISION_FP64;
    o->reduction_order = MC_ORDER_FAST;
}

const char* mc_last_error(void) { return g_error.c_str(); }

int mc_solve_pdlp(const mc_lp* lp, const mc_pdlp_options* opt, mc_pdlp_result* res) {
    return guarded([&] {
        check_lp(lp);
        if (!opt || !res) throw std::invalid_argument("null argument");
        const auto o = to_options(*opt, lp->cols, lp->rows);
        fo::detail::validate_options(o);
        fo::PdlpResult r;
        if (o.backend == fo::Backend::cpu) {
            r = fo::detail::solve_cpu(fo::detail::model_of(*lp), o, "cpu");
        } else {
            try {
                r = fo::detail::solve_gpu({*lp}, o)[0];
            } catch (const fo::detail::DeviceUnavailable& e) {
                if (o.require_gpu) throw std::runtime_error(std::string("GPU required but unavailable: ") + e.what());
                r = fo::detail::solve_cpu(fo::detail::model_of(*lp), o, "cpu_fallback");
            }
        }
        to_result(r, res);
    });
}

int mc_solve_pdlp_batch(const mc_lp* lps, int32_t count, const mc_pdlp_options* opt, mc_pdlp_result* res) {
    return guarded([&] {
        if (count < 0 || (count > 0 && (!lps || !res)) || !opt) throw std::invalid_argument("null argument");
        std::vector<mc_lp> v(lps, lps + count);
        for (const auto& lp : v) check_lp(&lp);
        int64_t totn = 0, totm = 0;
        for (const auto& lp : v) {
            totn += lp.cols;
            totm += lp.rows;
        }
        const auto o = to_options(*opt, totn, totm);
        fo::detail::validate_options(o);
        std::vector<fo::PdlpResult> out;
        auto host = [&](const char* label) {
            out.clear();
            for (const auto& lp : v) out.push_back(fo::detail::solve_cpu(fo::detail::model_of(lp), o, label));
        };
        if (o.backend == fo::Backend::cpu) {
            host("cpu");
        } else {
            try {
                out = fo::detail::solve_gpu(v, o);
            } catch (const fo::detail::DeviceUnavailable& e) {
                if (o.require_gpu) throw std::runtime_error(std::string("GPU required but unavailable: ") + e.what());
                host("cpu_fallback");
            }
        }
        for (int32_t k = 0; k < count; ++k) to_result(out[k], &res[k]);
    });
}

int mc_verify_primal(const mc_lp* lp, const double* primal, double claimed, double fa, double fr, double oa,
                     double orel, mc_verify_report* rep) {
    return guarded([&] {
        check_lp(lp);
        if (!rep || (!primal && lp->cols > 0)) throw std::invalid_argument("null argument");
        const auto md = fo::detail::model_of(*lp);
        markov_cero::verify::Candidate c{std::vector<double>(primal, primal + lp->cols), claimed};
        const auto r = markov_cero::verify::verify_primal(md, c, {fa, fr}, {oa, orel}, 1e-6);
        std::memset(rep, 0, sizeof(*rep));
        rep->passed = r.passed ? 1 : 0;
        rep->violation_count = (int32_t)r.violations.size();
        rep->recomputed_objective = r.recomputed_objective;
        rep->objective_difference = r.objective_difference;
        rep->maximum_row_violation = r.maximum_row_violation;
        rep->maximum_variable_violation = r.maximum_variable_violation;
        if (!r.violations.empty()) {
            const auto& v = r.violations.front();
            const std::string s = v.category + " idx=" + std::to_string(v.index) + " diff=" + std::to_string(v.magnitude);
            std::strncpy(rep->first_violation, s.c_str(), sizeof(rep->first_violation) - 1);
        }
    });
}

}  // extern "C"
