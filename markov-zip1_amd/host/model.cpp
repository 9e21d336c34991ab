// model.cpp -- contract checks, the long-double product and triplet
// aggregation of the markov-cero model types (the contract of reference
// src/model/model.cpp; the structure and messages here are this library's).
#include "markov_cero/model/model.hpp"

#include <algorithm>
#include <cmath>
#include <string_view>
#include <thread>
#include <unordered_set>
#include <vector>

namespace markov_cero::model {
namespace {

[[noreturn]] void reject(const std::string& why) { throw std::invalid_argument(why); }

std::string at(const char* what, std::size_t k) { return std::string(what) + " " + std::to_string(k); }

// Bound pair of one row or variable: finite values must be finite numbers, a
// lower bound may not be +inf nor an upper bound -inf, and finite lower <= upper.
void audit_bound_pair(const Bound& lo, const Bound& hi, const char* what, std::size_t k) {
    if ((lo.kind == BoundKind::finite && !std::isfinite(lo.value)) ||
        (hi.kind == BoundKind::finite && !std::isfinite(hi.value)))
        reject(at(what, k) + ": a finite bound holds a non-finite number");
    if (lo.kind == BoundKind::positive_infinity || hi.kind == BoundKind::negative_infinity)
        reject(at(what, k) + ": lower bound at +inf or upper bound at -inf");
    if (lo.is_finite() && hi.is_finite() && hi.value < lo.value)
        reject(at(what, k) + ": lower bound above upper bound");
}

// Names are optional metadata, but when present there is one per entity,
// each non-empty and unique; the first offending index (in entity order) is
// reported.  Long lists are checked by several threads: every name is hashed
// once, and thread t keeps the set of the names whose hash is t modulo the
// thread count, so equal names always meet in the same set; each thread's
// first repeat is its smallest index, and the smallest over the threads is the
// serial scan's answer.
void audit_names(const std::vector<std::string>& names, std::size_t count, const char* what) {
    if (names.size() != count)
        reject(std::string(what) + " names: " + std::to_string(names.size()) + " given for " + std::to_string(count));
    const std::size_t none = names.size();
    const std::size_t parts =
        names.size() < ((std::size_t)1 << 20) ? 1 : std::min<std::size_t>(16, std::max(1u, std::thread::hardware_concurrency()));
    std::vector<std::size_t> hash(names.size());
    std::size_t first_empty = none;
    for (std::size_t k = 0; k < names.size() && first_empty == none; ++k)
        if (names[k].empty()) first_empty = k;
    std::vector<std::size_t> first_repeat(parts, none);
    auto scan = [&](std::size_t t) {
        std::unordered_set<std::string_view> seen;  // views into `names`: no copies
        seen.reserve(2 * names.size() / parts + 16);
        for (std::size_t k = 0; k < names.size(); ++k)
            if (hash[k] % parts == t && !seen.insert(names[k]).second) {
                first_repeat[t] = k;
                return;
            }
    };
    auto hash_range = [&](std::size_t a, std::size_t b) {
        for (std::size_t k = a; k < b; ++k) hash[k] = std::hash<std::string_view>{}(names[k]);
    };
    if (parts == 1) {
        hash_range(0, names.size());
        scan(0);
    } else {
        std::vector<std::thread> pool;
        for (std::size_t t = 0; t < parts; ++t)
            pool.emplace_back(hash_range, names.size() * t / parts, names.size() * (t + 1) / parts);
        for (auto& th : pool) th.join();
        pool.clear();
        for (std::size_t t = 0; t < parts; ++t) pool.emplace_back(scan, t);
        for (auto& th : pool) th.join();
    }
    const std::size_t dup = *std::min_element(first_repeat.begin(), first_repeat.end());
    if (first_empty < dup) reject(at(what, first_empty) + " has an empty name");
    if (dup != none) reject(std::string(what) + " name repeated: " + names[dup]);
}

}  // namespace

Bound Bound::negative_infinity() noexcept { return Bound{BoundKind::negative_infinity, 0.0}; }
Bound Bound::positive_infinity() noexcept { return Bound{BoundKind::positive_infinity, 0.0}; }
Bound Bound::finite(double v) {
    if (!std::isfinite(v)) reject("Bound::finite needs a finite value");
    return Bound{BoundKind::finite, v};
}
bool Bound::is_finite() const noexcept { return kind == BoundKind::finite; }

// Structure of the CSC arrays (reference src/model/model.cpp:41-66): n + 1
// column pointers from 0 to nnz, non-decreasing; as many values as row ids;
// inside a column, row ids strictly increasing and below m; every value a
// finite nonzero.
void SparseMatrixCSC::validate() const {
    const std::size_t n = column_count;
    if (column_start.size() != n + 1)
        reject("sparse matrix: " + std::to_string(column_start.size()) + " column pointers for " + std::to_string(n) +
               " columns");
    if (row_index.size() != value.size()) reject("sparse matrix: row ids and values differ in length");
    if (column_start[0] != 0) reject("sparse matrix: the first column pointer is not 0");
    if (column_start[n] != value.size()) reject("sparse matrix: the last column pointer is not the nonzero count");
    for (std::size_t j = 0; j < n; ++j) {
        const std::size_t first = column_start[j], last = column_start[j + 1];
        if (last < first) reject(at("sparse matrix: column pointers decrease at column", j));
        std::size_t prev_row = 0;
        for (std::size_t k = first; k < last; ++k) {
            const std::size_t r = row_index[k];
            if (r >= row_count) reject(at("sparse matrix: row index out of range in column", j));
            if (k > first && r <= prev_row) reject(at("sparse matrix: row ids not strictly increasing in column", j));
            prev_row = r;
            const double v = value[k];
            if (v == 0.0 || !std::isfinite(v)) reject(at("sparse matrix: zero or non-finite value in column", j));
        }
    }
}

// y = A x with each row's terms accumulated in long double, column by column
// (reference model.cpp:68-86; the verifier's row activities depend on this
// order).
std::vector<double> SparseMatrixCSC::multiply(const std::vector<double>& x) const {
    validate();
    if (x.size() != column_count)
        reject("multiply: vector of " + std::to_string(x.size()) + " for " + std::to_string(column_count) + " columns");
    std::vector<long double> row_sum(row_count, 0.0L);
    for (std::size_t j = 0; j < column_count; ++j) {
        const double xj = x[j];
        if (!std::isfinite(xj)) reject(at("multiply: non-finite entry", j));
        for (std::size_t k = column_start[j]; k < column_start[j + 1]; ++k)
            row_sum[row_index[k]] += static_cast<long double>(value[k]) * xj;
    }
    std::vector<double> y(row_sum.begin(), row_sum.end());
    for (std::size_t i = 0; i < row_count; ++i)
        if (!std::isfinite(y[i])) throw std::overflow_error(at("multiply: overflow in row", i));
    return y;
}

SparseMatrixBuilder::SparseMatrixBuilder(std::size_t row_count, std::size_t column_count)
    : rows_(row_count), columns_(column_count) {}

void SparseMatrixBuilder::add(std::size_t row, std::size_t column, double value) {
    if (row >= rows_ || column >= columns_) throw std::out_of_range("builder entry outside the matrix");
    if (!std::isfinite(value)) reject("builder entry is not finite");
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
        if (!std::isfinite(v)) throw std::overflow_error("builder: duplicate entries sum to a non-finite value");
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

// Model contract (reference model.cpp:120-165): a valid matrix; objective,
// bounds, types and names sized to it; finite objective data; consistent
// bound pairs; binary variables boxed inside [0, 1]; a quadratic term, if
// any, n x n.
void Model::validate() const {
    matrix.validate();
    const std::size_t m = matrix.row_count, n = matrix.column_count;
    if (!std::isfinite(objective_offset)) reject("objective offset is not finite");
    for (const auto* v : {&variable_lower, &variable_upper})
        if (v->size() != n) reject("variable bounds are not sized to the matrix columns");
    if (objective.size() != n || variable_type.size() != n)
        reject("objective or variable types are not sized to the matrix columns");
    if (row_lower.size() != m || row_upper.size() != m) reject("row bounds are not sized to the matrix rows");
    audit_names(row_name, m, "row");
    audit_names(variable_name, n, "variable");
    for (std::size_t j = 0; j < n; ++j) {
        if (!std::isfinite(objective[j])) reject(at("objective coefficient of variable", j) + " is not finite");
        audit_bound_pair(variable_lower[j], variable_upper[j], "variable", j);
        if (variable_type[j] == VariableType::binary) {
            const Bound& lo = variable_lower[j];
            const Bound& hi = variable_upper[j];
            if (!lo.is_finite() || !hi.is_finite() || lo.value < 0.0 || hi.value > 1.0)
                reject(at("binary variable", j) + " needs finite bounds inside [0, 1]");
        }
    }
    for (std::size_t i = 0; i < m; ++i) audit_bound_pair(row_lower[i], row_upper[i], "row", i);
    if (has_quadratic_objective) {
        quadratic_matrix.validate();
        if (quadratic_matrix.row_count != n || quadratic_matrix.column_count != n)
            reject("quadratic objective matrix is not n x n");
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
