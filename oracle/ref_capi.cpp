// ref_capi.cpp -- extern "C" shim over the COMPILED REFERENCE (test infrastructure).
//
// oracle/Makefile compiles this file together with the reference's own
// sources (src/model/model.cpp, src/scale/ruiz_scaling.cpp,
// src/lp/first_order/pdlp.cpp, src/verify/primal_verifier.cpp) where they lie
// under the reference checkout, into oracle/_ref/libref_pdlp.so.  Tests use it
// to pin oracle/pdlp_oracle.c bit for bit and bench.py times it as the CPU
// baseline.  Only the struct layouts of include/mcpdlp.h are shared.
#include "markov_cero/lp/first_order/pdlp.hpp"
#include "markov_cero/scale/ruiz_scaling.hpp"
#include "markov_cero/verify/primal_verifier.hpp"

#include <cstring>
#include <exception>
#include <string>

#include "../include/mcpdlp.h"

namespace mc = markov_cero;

namespace {

mc::model::Bound to_bound(const mc_bound& b) {
    mc::model::Bound out;
    out.kind = b.kind == MC_BOUND_NEGATIVE_INFINITY ? mc::model::BoundKind::negative_infinity
               : b.kind == MC_BOUND_POSITIVE_INFINITY ? mc::model::BoundKind::positive_infinity
                                                      : mc::model::BoundKind::finite;
    out.value = b.value;
    return out;
}

mc::model::Model to_model(const mc_lp* lp) {
    mc::model::Model md;
    md.name = "ref";
    md.objective_sense =
        lp->sense == MC_MAXIMIZE ? mc::model::ObjectiveSense::maximize : mc::model::ObjectiveSense::minimize;
    md.objective_offset = lp->objective_offset;
    const auto m = static_cast<std::size_t>(lp->rows), n = static_cast<std::size_t>(lp->cols);
    md.matrix.row_count = m;
    md.matrix.column_count = n;
    md.matrix.column_start.assign(lp->column_start, lp->column_start + n + 1);
    const auto nnz = static_cast<std::size_t>(lp->column_start[n]);
    md.matrix.row_index.assign(lp->row_index, lp->row_index + nnz);
    md.matrix.value.assign(lp->value, lp->value + nnz);
    md.objective.assign(lp->objective, lp->objective + n);
    for (std::size_t i = 0; i < m; ++i) {
        md.row_lower.push_back(to_bound(lp->row_lower[i]));
        md.row_upper.push_back(to_bound(lp->row_upper[i]));
        md.row_name.push_back("R" + std::to_string(i));
    }
    for (std::size_t j = 0; j < n; ++j) {
        md.variable_lower.push_back(to_bound(lp->var_lower[j]));
        md.variable_upper.push_back(to_bound(lp->var_upper[j]));
        md.variable_type.push_back(mc::model::VariableType::continuous);
        md.variable_name.push_back("C" + std::to_string(j));
    }
    return md;
}

mc::lp::first_order::PdlpOptions to_options(const mc_pdlp_options* o) {
    mc::lp::first_order::PdlpOptions p;
    p.max_iterations = o->max_iterations;
    p.restart_every = o->restart_every;
    p.primal_tolerance = o->primal_tolerance;
    p.dual_tolerance = o->dual_tolerance;
    p.gap_tolerance = o->gap_tolerance;
    p.step_size_reduction = o->step_size_reduction;
    p.power_iterations = o->power_iterations;
    p.backend = o->backend == MC_BACKEND_GPU ? mc::lp::first_order::Backend::gpu
                                             : mc::lp::first_order::Backend::cpu;
    p.restart_strategy = o->restart_strategy == MC_RESTART_NONE    ? mc::lp::first_order::RestartStrategy::none
                         : o->restart_strategy == MC_RESTART_FIXED ? mc::lp::first_order::RestartStrategy::fixed
                                                                   : mc::lp::first_order::RestartStrategy::adaptive;
    p.restart_reduction_factor = o->restart_reduction_factor;
    p.adaptive_step_size = o->adaptive_step_size != 0;
    p.adaptive_primal_weight = o->adaptive_primal_weight != 0;
    p.initial_primal_weight = o->initial_primal_weight;
    p.primal_weight_smoothing = o->primal_weight_smoothing;
    p.ruiz_scaling = o->ruiz_scaling != 0;
    p.ruiz_iterations = o->ruiz_iterations;
    return p;
}

thread_local std::string g_err;

}  // namespace

extern "C" {

const char* ref_last_error(void) { return g_err.c_str(); }

// Runs the reference solve_pdlp (pdlp.hpp:67) on `lp`.
int ref_solve_pdlp(const mc_lp* lp, const mc_pdlp_options* o, mc_pdlp_result* r) {
    try {
        const auto md = to_model(lp);
        const auto res = mc::lp::first_order::solve_pdlp(md, to_options(o));
        r->status = static_cast<int32_t>(res.status);
        r->objective = res.objective;
        r->primal_infeasibility = res.primal_infeasibility;
        r->dual_infeasibility = res.dual_infeasibility;
        r->duality_gap = res.duality_gap;
        r->tolerance = res.tolerance;
        r->iterations = res.iterations;
        r->h2d_ms = res.h2d_ms;
        r->kernel_ms = res.kernel_ms;
        r->d2h_ms = res.d2h_ms;
        r->total_ms = res.total_ms;
        std::memset(r->message, 0, sizeof(r->message));
        std::strncpy(r->message, res.message.c_str(), sizeof(r->message) - 1);
        std::memset(r->backend_actually_used, 0, sizeof(r->backend_actually_used));
        std::strncpy(r->backend_actually_used, res.backend_actually_used.c_str(),
                     sizeof(r->backend_actually_used) - 1);
        r->primal_len = static_cast<int64_t>(res.primal.size());
        r->dual_len = static_cast<int64_t>(res.dual.size());
        if (r->primal) std::memcpy(r->primal, res.primal.data(), sizeof(double) * res.primal.size());
        if (r->dual) std::memcpy(r->dual, res.dual.data(), sizeof(double) * res.dual.size());
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// Runs the reference equilibrate_model (ruiz_scaling.hpp:33) on a copy of
// `lp`; writes the scaled values/objective/bounds and the scalers.
int ref_equilibrate_model(const mc_lp* lp, int64_t iterations, double* val, double* objective,
                          mc_bound* row_lower, mc_bound* row_upper, mc_bound* var_lower,
                          mc_bound* var_upper, double* row_scale, double* col_scale,
                          int64_t* passes, int32_t* converged) {
    try {
        auto md = to_model(lp);
        mc::scale::RuizOptions ro;
        ro.max_iterations = static_cast<std::size_t>(iterations);
        const auto sc = mc::scale::equilibrate_model(md, ro);
        std::memcpy(val, md.matrix.value.data(), sizeof(double) * md.matrix.value.size());
        std::memcpy(objective, md.objective.data(), sizeof(double) * md.objective.size());
        auto put = [](const std::vector<mc::model::Bound>& src, mc_bound* dst) {
            for (std::size_t k = 0; k < src.size(); ++k) {
                dst[k].kind = static_cast<int32_t>(src[k].kind);
                dst[k].pad_ = 0;
                dst[k].value = src[k].value;
            }
        };
        put(md.row_lower, row_lower);
        put(md.row_upper, row_upper);
        put(md.variable_lower, var_lower);
        put(md.variable_upper, var_upper);
        std::memcpy(row_scale, sc.row_scale.data(), sizeof(double) * sc.row_scale.size());
        std::memcpy(col_scale, sc.col_scale.data(), sizeof(double) * sc.col_scale.size());
        *passes = static_cast<int64_t>(sc.iterations_executed);
        *converged = sc.converged ? 1 : 0;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// Runs the reference verify_primal (primal_verifier.hpp:41-45).
int ref_verify_primal(const mc_lp* lp, const double* x, double claimed, double fa, double fr,
                      double oa, double orel, mc_verify_report* rep) {
    try {
        const auto md = to_model(lp);
        mc::verify::Candidate cand{std::vector<double>(x, x + lp->cols), claimed};
        const auto r = mc::verify::verify_primal(md, cand, {fa, fr}, {oa, orel}, 1e-6);
        std::memset(rep, 0, sizeof(*rep));
        rep->passed = r.passed ? 1 : 0;
        rep->violation_count = static_cast<int32_t>(r.violations.size());
        rep->recomputed_objective = r.recomputed_objective;
        rep->objective_difference = r.objective_difference;
        rep->maximum_row_violation = r.maximum_row_violation;
        rep->maximum_variable_violation = r.maximum_variable_violation;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

}  // extern "C"
