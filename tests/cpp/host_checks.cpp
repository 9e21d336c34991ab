// Checks of the host library's large-model paths that the C-ABI cannot reach:
// the threaded name audit of Model::validate and the threaded row activities of
// verify_primal (compiled and run by tests/test_host_cpp.py).
#include <cstdio>
#include <random>
#include <stdexcept>
#include <string>

#include "markov_cero/model/model.hpp"
#include "markov_cero/verify/primal_verifier.hpp"

using namespace markov_cero;

static int failures = 0;
static void expect(bool ok, const char* what) {
    if (!ok) {
        std::printf("FAIL: %s\n", what);
        ++failures;
    }
}

// m x n model with `per` nonzeros per column (rows strictly increasing), names R<i> / C<j>.
static model::Model make(std::size_t m, std::size_t n, std::size_t per, unsigned seed) {
    model::Model md;
    md.matrix.row_count = m;
    md.matrix.column_count = n;
    md.matrix.column_start.assign(n + 1, 0);
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> val(0.1, 10.0);
    for (std::size_t j = 0; j < n; ++j) {
        md.matrix.column_start[j] = md.matrix.row_index.size();
        std::size_t r = rng() % (m / per);
        for (std::size_t k = 0; k < per && r < m; ++k) {
            md.matrix.row_index.push_back(r);
            md.matrix.value.push_back((rng() & 1 ? 1.0 : -1.0) * val(rng));
            r += 1 + rng() % (m / per);
        }
    }
    md.matrix.column_start[n] = md.matrix.row_index.size();
    md.objective.assign(n, 1.0);
    md.variable_lower.assign(n, model::Bound::finite(-100.0));
    md.variable_upper.assign(n, model::Bound::finite(100.0));
    md.variable_type.assign(n, model::VariableType::continuous);
    md.row_lower.assign(m, model::Bound::negative_infinity());
    md.row_upper.assign(m, model::Bound::positive_infinity());
    for (std::size_t i = 0; i < m; ++i) md.row_name.push_back("R" + std::to_string(i));
    for (std::size_t j = 0; j < n; ++j) md.variable_name.push_back("C" + std::to_string(j));
    return md;
}

static std::string validate_error(const model::Model& md) {
    try {
        md.validate();
    } catch (const std::invalid_argument& e) {
        return e.what();
    }
    return "";
}

int main() {
    // names: the first repeat in index order is reported, also across threads
    model::Model md = make(1'500'000, 1'200'000, 4, 1);
    expect(validate_error(md).empty(), "valid model validates");
    md.row_name[1'300'000] = "R17";   // repeats row 17
    md.row_name[1'100'000] = "R900";  // an earlier repeat (of row 900)
    expect(validate_error(md) == "row name repeated: R900", "first repeated row name");
    md.row_name[1'100'000] = "R1100000";
    md.row_name[1'200'000] = "";       // an empty name before the remaining repeat
    expect(validate_error(md) == "row 1200000 has an empty name", "empty name before a later repeat");
    md.row_name[1'200'000] = "R1200000";
    expect(validate_error(md) == "row name repeated: R17", "remaining repeat");
    md.row_name[1'300'000] = "R1300000";
    md.variable_name[5] = "C1199999";  // repeated later at index 1199999
    expect(validate_error(md) == "variable name repeated: C1199999", "variable repeat found at its second index");
    md.variable_name[5] = "C5";

    // row activities: verify_primal's threaded long-double pass equals multiply() bit for bit
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> u(-3.0, 3.0);
    std::vector<double> x(md.matrix.column_count);
    for (auto& v : x) v = u(rng);
    const std::vector<double> ax = md.matrix.multiply(x);
    for (std::size_t i = 0; i < ax.size(); ++i) {
        md.row_lower[i] = model::Bound::finite(ax[i]);
        md.row_upper[i] = model::Bound::finite(ax[i]);
    }
    const auto rep = verify::verify_primal(md, {x, 0.0}, {0.0, 0.0}, {1e-4, 1e-4}, 1e-6);
    expect(rep.maximum_row_violation == 0.0, "threaded row activities equal multiply()");
    md.row_lower[777] = model::Bound::finite(ax[777] - 2e-3);
    md.row_upper[777] = model::Bound::finite(ax[777] - 1e-3);
    const auto rep2 = verify::verify_primal(md, {x, 0.0}, {0.0, 0.0}, {1e-4, 1e-4}, 1e-6);
    expect(!rep2.violations.empty() && rep2.violations[0].category == "row_upper" && rep2.violations[0].index == 777,
           "a moved row bound is reported");
    if (failures == 0) std::printf("PASS host checks\n");
    return failures == 0 ? 0 : 1;
}
