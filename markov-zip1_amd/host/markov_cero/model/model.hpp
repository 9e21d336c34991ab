// model.hpp -- LP model types of markov-cero, as the PDLP engine accepts them.
//
// Same names, fields and semantics as the reference's
// include/markov_cero/model/model.hpp:11-70 so that code written against the
// reference compiles unchanged: bounds carry an explicit kind (infinities are
// kinds, not IEEE values), the matrix is CSC with size_t indices and strictly
// increasing rows per column, values finite and nonzero.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace markov_cero::model {

enum class ObjectiveSense { minimize, maximize };
enum class VariableType { continuous, integer, binary };
enum class BoundKind { negative_infinity, finite, positive_infinity };

struct Bound final {
    BoundKind kind{BoundKind::finite};
    double value{0.0};

    [[nodiscard]] static Bound negative_infinity() noexcept;
    [[nodiscard]] static Bound finite(double value);  // throws std::invalid_argument if not finite
    [[nodiscard]] static Bound positive_infinity() noexcept;
    [[nodiscard]] bool is_finite() const noexcept;
};

struct SparseMatrixCSC final {
    std::size_t row_count{};
    std::size_t column_count{};
    std::vector<std::size_t> column_start;  // [column_count + 1]
    std::vector<std::size_t> row_index;     // [nnz]
    std::vector<double> value;              // [nnz]

    void validate() const;  // throws std::invalid_argument
    // y = A x accumulated in long double (reference model.cpp:68-86).
    [[nodiscard]] std::vector<double> multiply(const std::vector<double>& x) const;
};

// Collects (row, column, value) triplets; build() sums duplicates in long
// double, drops entries that sum to zero and returns a validated CSC matrix
// (reference model.cpp:88-118).
class SparseMatrixBuilder final {
  public:
    SparseMatrixBuilder(std::size_t row_count, std::size_t column_count);
    void add(std::size_t row, std::size_t column, double value);
    [[nodiscard]] SparseMatrixCSC build() const;

  private:
    struct Triplet final {
        std::size_t row;
        std::size_t column;
        double value;
    };
    std::size_t rows_;
    std::size_t columns_;
    std::vector<Triplet> triplets_;
};

struct Model final {
    std::string name;
    ObjectiveSense objective_sense{ObjectiveSense::minimize};
    double objective_offset{0.0};
    SparseMatrixCSC matrix;
    std::vector<double> objective;
    std::vector<Bound> row_lower;
    std::vector<Bound> row_upper;
    std::vector<Bound> variable_lower;
    std::vector<Bound> variable_upper;
    std::vector<VariableType> variable_type;
    std::vector<std::string> row_name;
    std::vector<std::string> variable_name;
    bool has_quadratic_objective{false};
    SparseMatrixCSC quadratic_matrix;

    void validate() const;  // throws std::invalid_argument
};

[[nodiscard]] const char* to_string(ObjectiveSense sense) noexcept;
[[nodiscard]] const char* to_string(VariableType type) noexcept;

}  // namespace markov_cero::model
