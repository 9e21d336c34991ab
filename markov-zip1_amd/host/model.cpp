// model.cpp -- validation, long-double product and triplet aggregation for
// the markov-cero model types (reference src/model/model.cpp).
#include "markov_cero/model.hpp"

#include <algorithm>
#include <cmath>
#include <unordered_set>

namespace markov_cero::model {
namespace {

void finite_or_throw(double v, const std::string& what) {
    if (!std::isfinite(v)) throw std::invalid_argument(what + " must be finite");
}

void check_bound(const Bound& b, const char* what) {
    if (b.kind == BoundKind::finite) finite_or_throw(b.value, what);
}

void check_names(const std::vector<std::string>& names, std::size_t expected, const std::string& what) {
    if (names.size() != expected) throw std::invalid_argument(what + " count mismatch");
    std::unordered_set<std::string> seen;
    seen.reserve(names.size());
    for (const auto& nm : names) {
        if (nm.empty()) throw std::invalid_argument(what + " cannot be empty");
        if (!seen.insert(nm).second) throw std::invalid_argument("duplicate " + what + ": " + nm);
    }
}

}  // namespace

Bound Bound::negative_infinity() noexcept { return Bound{BoundKind::negative_infinity, 0.0}; }
Bound Bound::finite(double v) {
    finite_or_throw(v, "bound");
    return Bound{BoundKind::finite, v};
}
Bound Bound::positive_infinity() noexcept { return Bound{BoundKind::positive_infinity, 0.0}; }
bool Bound::is_finite() const noexcept { return kind == BoundKind::finite; }

void SparseMatrixCSC::validate() const {
    if (column_start.size() != column_count + 1) throw std::invalid_argument("CSC column_start size mismatch");
    if (column_start.front() != 0) throw std::invalid_argument("CSC must start at zero");
    if (row_index.size() != value.size()) throw std::invalid_argument("CSC index/value size mismatch");
    if (column_start.back() != value.size()) throw std::invalid_argument("CSC terminal pointer mismatch");
    for (std::size_t col = 0; col < column_count; ++col) {
        const std::size_t b = column_start[col], e = column_start[col + 1];
        if (b > e) throw std::invalid_argument("CSC pointers must be monotone");
        for (std::size_t k = b; k < e; ++k) {
            if (row_index[k] >= row_count) throw std::invalid_argument("CSC row index out of range");
            if (k > b && row_index[k] <= row_index[k - 1])
                throw std::invalid_argument("CSC rows must be strictly increasing per column");
            if (!std::isfinite(value[k]) || value[k] == 0.0)
                throw std::invalid_argument("CSC values must be finite and nonzero");
        }
    }
}

std::vector<double> SparseMatrixCSC::multiply(const std::vector<double>& x) const {
    validate();
    if (x.size() != column_count) throw std::invalid_argument("matrix/vector dimension mismatch");
    std::vector<long double> acc(row_count, 0.0L);
    for (std::size_t col = 0; col < column_count; ++col) {
        if (!std::isfinite(x[col])) throw std::invalid_argument("vector contains non-finite value");
        for (std::size_t k = column_start[col]; k < column_start[col + 1]; ++k)
            acc[row_index[k]] += static_cast<long double>(value[k]) * x[col];
    }
    std::vector<double> out(row_count);
    for (std::size_t r = 0; r < row_count; ++r) {
        out[r] = static_cast<double>(acc[r]);
        if (!std::isfinite(out[r])) throw std::overflow_error("matrix/vector product is non-finite");
    }
    return out;
}

SparseMatrixBuilder::SparseMatrixBuilder(std::size_t row_count, std::size_t column_count)
    : rows_(row_count), columns_(column_count) {}

void SparseMatrixBuilder::add(std::size_t row, std::size_t column, double value) {
    if (row >= rows_ || column >= columns_) throw std::out_of_range("sparse entry index out of range");
    finite_or_throw(value, "sparse value");
    triplets_.push_back({row, column, value});
}

SparseMatrixCSC SparseMatrixBuilder::build() const {
    // Order triplets by (column, row); a stable sort keeps insertion order
    // among duplicates, which are then summed in long double.
    std::vector<std::size_t> order(triplets_.size());
    for (std::size_t k = 0; k < order.size(); ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](std::size_t a, std::size_t b) {
        const auto& ta = triplets_[a];
        const auto& tb = triplets_[b];
        return ta.column != tb.column ? ta.column < tb.column : ta.row < tb.row;
    });
    SparseMatrixCSC out;
    out.row_count = rows_;
    out.column_count = columns_;
    out.column_start.assign(columns_ + 1, 0);
    std::size_t k = 0;
    while (k < order.size()) {
        const auto& first = triplets_[order[k]];
        long double sum = 0.0L;
        std::size_t e = k;
        while (e < order.size() && triplets_[order[e]].column == first.column &&
               triplets_[order[e]].row == first.row) {
            sum += triplets_[order[e]].value;
            ++e;
        }
        const double v = static_cast<double>(sum);
        if (!std::isfinite(v)) throw std::overflow_error("aggregated sparse value is non-finite");
        if (v != 0.0) {
            out.row_index.push_back(first.row);
            out.value.push_back(v);
            ++out.column_start[first.column + 1];
        }
        k = e;
    }
    for (std::size_t c = 0; c < columns_; ++c) out.column_start[c + 1] += out.column_start[c];
    out.validate();
    return out;
}

void Model::validate() const {
    matrix.validate();
    const std::size_t m = matrix.row_count, n = matrix.column_count;
    finite_or_throw(objective_offset, "objective offset");
    if (objective.size() != n || variable_lower.size() != n || variable_upper.size() != n ||
        variable_type.size() != n)
        throw std::invalid_argument("variable metadata dimension mismatch");
    if (row_lower.size() != m || row_upper.size() != m)
        throw std::invalid_argument("row metadata dimension mismatch");
    check_names(row_name, m, "row name");
    check_names(variable_name, n, "variable name");
    for (std::size_t j = 0; j < n; ++j) {
        finite_or_throw(objective[j], "objective coefficient");
        check_bound(variable_lower[j], "variable lower bound");
        check_bound(variable_upper[j], "variable upper bound");
        if (variable_lower[j].kind == BoundKind::positive_infinity ||
            variable_upper[j].kind == BoundKind::negative_infinity)
            throw std::invalid_argument("invalid variable infinity direction");
        if (variable_lower[j].is_finite() && variable_upper[j].is_finite() &&
            variable_lower[j].value > variable_upper[j].value)
            throw std::invalid_argument("inconsistent variable bounds");
        if (variable_type[j] == VariableType::binary) {
            if (!variable_lower[j].is_finite() || !variable_upper[j].is_finite())
                throw std::invalid_argument("binary bounds must be finite");
            if (variable_lower[j].value < 0.0 || variable_upper[j].value > 1.0)
                throw std::invalid_argument("binary bounds must stay within zero and one");
        }
    }
    for (std::size_t i = 0; i < m; ++i) {
        check_bound(row_lower[i], "row lower bound");
        check_bound(row_upper[i], "row upper bound");
        if (row_lower[i].kind == BoundKind::positive_infinity ||
            row_upper[i].kind == BoundKind::negative_infinity)
            throw std::invalid_argument("invalid row infinity direction");
        if (row_lower[i].is_finite() && row_upper[i].is_finite() && row_lower[i].value > row_upper[i].value)
            throw std::invalid_argument("inconsistent row bounds");
    }
    if (has_quadratic_objective) {
        quadratic_matrix.validate();
        if (quadratic_matrix.column_count != n || quadratic_matrix.row_count != n)
            throw std::invalid_argument("quadratic matrix dimension mismatch");
    }
}

const char* to_string(ObjectiveSense sense) noexcept {
    return sense == ObjectiveSense::minimize ? "minimize" : "maximize";
}

const char* to_string(VariableType type) noexcept {
    switch (type) {
        case VariableType::continuous: return "continuous";
        case VariableType::integer: return "integer";
        case VariableType::binary: return "binary";
    }
    return "unknown";
}

}  // namespace markov_cero::model
