// pdlp.cpp -- solve_pdlp / solve_pdlp_batch dispatch between the host path
// and the HIP engine (reference src/lp/first_order/pdlp.cpp:139-149 for the
// trivial exit and the backend labels).
#include "markov_cero/lp/first_order/pdlp.hpp"

#include <stdexcept>
#include <string>

#include "pdlp_internal.hpp"

namespace markov_cero::lp::first_order {

model::Model detail::model_of(const mc_lp& lp) {
    model::Model md;
    md.name = "lp";
    md.objective_sense = lp.sense == MC_MAXIMIZE ? model::ObjectiveSense::maximize : model::ObjectiveSense::minimize;
    md.objective_offset = lp.objective_offset;
    const auto m = (std::size_t)lp.rows, n = (std::size_t)lp.cols;
    md.matrix.row_count = m;
    md.matrix.column_count = n;
    md.matrix.column_start.assign(lp.column_start, lp.column_start + n + 1);
    const auto nnz = (std::size_t)lp.column_start[n];
    md.matrix.row_index.assign(lp.row_index, lp.row_index + nnz);
    md.matrix.value.assign(lp.value, lp.value + nnz);
    md.objective.assign(lp.objective, lp.objective + n);
    auto bound = [](const mc_bound& b) {
        return b.kind == MC_BOUND_NEGATIVE_INFINITY   ? model::Bound::negative_infinity()
               : b.kind == MC_BOUND_POSITIVE_INFINITY ? model::Bound::positive_infinity()
                                                      : model::Bound{model::BoundKind::finite, b.value};
    };
    for (std::size_t i = 0; i < m; ++i) {
        md.row_lower.push_back(bound(lp.row_lower[i]));
        md.row_upper.push_back(bound(lp.row_upper[i]));
        md.row_name.push_back("R" + std::to_string(i));
    }
    for (std::size_t j = 0; j < n; ++j) {
        md.variable_lower.push_back(bound(lp.var_lower[j]));
        md.variable_upper.push_back(bound(lp.var_upper[j]));
        md.variable_type.push_back(model::VariableType::continuous);
        md.variable_name.push_back("C" + std::to_string(j));
    }
    return md;
}

void detail::validate_options(const PdlpOptions& opt) {
    if (opt.row_weighting < 0 || opt.row_weighting > 8) throw std::invalid_argument("row_weighting must be in 0..8");
    const int d = opt.primal_weight_max_drop;
    if (d != 0 && (d < 2 || d > (1 << 20) || (d & (d - 1)) != 0))
        throw std::invalid_argument("primal_weight_max_drop must be 0 or a power of two in 2..2^20");
}

PdlpResult solve_pdlp(const model::Model& md, const PdlpOptions& opt) {
    detail::validate_options(opt);
    if (opt.backend == Backend::cpu) return detail::solve_cpu(md, opt, "cpu");
    try {
        return detail::solve_gpu({detail::view_of(md)}, opt)[0];
    } catch (const detail::DeviceUnavailable& e) {
        if (opt.require_gpu) throw std::runtime_error(std::string("GPU required but unavailable: ") + e.what());
        return detail::solve_cpu(md, opt, "cpu_fallback");
    }
}

std::vector<PdlpResult> solve_pdlp_batch(const std::vector<const model::Model*>& models, const PdlpOptions& opt) {
    detail::validate_options(opt);
    std::vector<PdlpResult> out;
    auto host_all = [&](const char* label) {
        out.clear();
        for (const auto* md : models) out.push_back(detail::solve_cpu(*md, opt, label));
        return out;
    };
    if (opt.backend == Backend::cpu) return host_all("cpu");
    std::vector<mc_lp> views;
    views.reserve(models.size());
    for (const auto* md : models) views.push_back(detail::view_of(*md));
    try {
        return detail::solve_gpu(views, opt);
    } catch (const detail::DeviceUnavailable& e) {
        if (opt.require_gpu) throw std::runtime_error(std::string("GPU required but unavailable: ") + e.what());
        return host_all("cpu_fallback");
    }
}

}  // namespace markov_cero::lp::first_order
