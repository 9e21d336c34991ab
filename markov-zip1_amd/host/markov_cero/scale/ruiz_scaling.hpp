// scale.hpp -- Ruiz l-inf equilibration as PDLP applies it
// (reference include/markov_cero/scale/ruiz_scaling.hpp:12-39).
#pragma once

#include <cstddef>
#include <vector>

#include "markov_cero/model/model.hpp"

namespace markov_cero::scale {

struct RuizOptions {
    std::size_t max_iterations{10};
    double tolerance{1e-3};
    double min_scale{1e-4};  // not applied by equilibrate_model (as in the reference)
    double max_scale{1e4};
};

struct RuizScalers {
    std::vector<double> row_scale;      // D_R
    std::vector<double> col_scale;      // D_C
    std::vector<double> inv_row_scale;  // D_R^{-1}
    std::vector<double> inv_col_scale;  // D_C^{-1}
    std::size_t iterations_executed{};
    bool converged{false};
};

// A <- D_R A D_C with up to max_iterations passes of 1/sqrt(inf-norm) scaling;
// row bounds scaled by D_R, variable bounds by D_C^{-1}, objective by D_C.
[[nodiscard]] RuizScalers equilibrate_model(model::Model& model, const RuizOptions& options = {});

void unscale_model_solution(const RuizScalers& scalers, std::vector<double>& primal,
                            std::vector<double>& dual);

}  // namespace markov_cero::scale
