// scale.cpp -- host Ruiz equilibration for the CPU path
// (reference src/scale/ruiz_scaling.cpp:152-258).  The GPU path runs the same
// passes on the device (kernels.hip k_ruiz_*); both use only maxima,
// 1/sqrt and products, so they agree bit for bit.
#include "markov_cero/scale/ruiz_scaling.hpp"

#include <algorithm>
#include <cmath>

namespace markov_cero::scale {

RuizScalers equilibrate_model(model::Model& md, const RuizOptions& options) {
    const std::size_t m = md.matrix.row_count, n = md.matrix.column_count;
    auto& A = md.matrix;
    RuizScalers sc;
    sc.row_scale.assign(m, 1.0);
    sc.col_scale.assign(n, 1.0);
    sc.inv_row_scale.assign(m, 1.0);
    sc.inv_col_scale.assign(n, 1.0);
    if (m == 0 || n == 0 || A.value.empty()) return sc;

    std::vector<double> rmax(m), cmax(n);
    std::size_t pass = 0;
    while (pass < options.max_iterations) {
        std::fill(rmax.begin(), rmax.end(), 0.0);
        for (std::size_t j = 0; j < n; ++j) {
            double cm = 0.0;
            for (std::size_t k = A.column_start[j]; k < A.column_start[j + 1]; ++k) {
                const double a = std::abs(A.value[k]);
                cm = std::max(cm, a);
                rmax[A.row_index[k]] = std::max(rmax[A.row_index[k]], a);
            }
            cmax[j] = cm == 0.0 ? 1.0 : cm;
        }
        double err = 0.0, errc = 0.0;
        for (auto& r : rmax) {
            if (r == 0.0) r = 1.0;
            err = std::max(err, std::abs(1.0 - r));
        }
        for (const double c : cmax) errc = std::max(errc, std::abs(1.0 - c));
        if (err < options.tolerance && errc < options.tolerance) {
            sc.converged = true;
            break;
        }
        for (std::size_t i = 0; i < m; ++i) {
            rmax[i] = 1.0 / std::sqrt(rmax[i]);  // now the row delta
            sc.row_scale[i] *= rmax[i];
        }
        for (std::size_t j = 0; j < n; ++j) {
            cmax[j] = 1.0 / std::sqrt(cmax[j]);  // now the column delta
            sc.col_scale[j] *= cmax[j];
        }
        for (std::size_t j = 0; j < n; ++j)
            for (std::size_t k = A.column_start[j]; k < A.column_start[j + 1]; ++k)
                A.value[k] *= rmax[A.row_index[k]] * cmax[j];
        ++pass;
    }
    sc.iterations_executed = pass;
    for (std::size_t i = 0; i < m; ++i) sc.inv_row_scale[i] = 1.0 / sc.row_scale[i];
    for (std::size_t j = 0; j < n; ++j) sc.inv_col_scale[j] = 1.0 / sc.col_scale[j];
    for (std::size_t i = 0; i < m; ++i) {
        if (md.row_lower[i].is_finite()) md.row_lower[i].value *= sc.row_scale[i];
        if (md.row_upper[i].is_finite()) md.row_upper[i].value *= sc.row_scale[i];
    }
    for (std::size_t j = 0; j < n; ++j) {
        if (md.variable_lower[j].is_finite()) md.variable_lower[j].value *= sc.inv_col_scale[j];
        if (md.variable_upper[j].is_finite()) md.variable_upper[j].value *= sc.inv_col_scale[j];
        md.objective[j] *= sc.col_scale[j];
    }
    return sc;
}

void unscale_model_solution(const RuizScalers& sc, std::vector<double>& primal, std::vector<double>& dual) {
    for (std::size_t j = 0; j < primal.size() && j < sc.col_scale.size(); ++j) primal[j] *= sc.col_scale[j];
    for (std::size_t i = 0; i < dual.size() && i < sc.row_scale.size(); ++i) dual[i] *= sc.row_scale[i];
}

}  // namespace markov_cero::scale
