// mps.hpp -- model ingest for the PDLP path: free-format MPS (reader and
// writer) and a raw binary LP format for instances past MPS's practical size.
// Row/bound semantics follow the reference reader (src/io/mps.cpp:390-440):
// N row = objective, E/L/G rows with RANGES, LO/UP/FX/FR/MI/PL/BV/LI/UI bounds,
// default variable bounds [0, +inf).
//
// Size limits: MpsLimits has the reference's fields and defaults
// (include/markov_cero/io/mps.hpp:13-20) and is applied when a caller passes
// it.  parse_mps(in) without limits reads any size: large instances are the
// point of the GPU path (the reference applies its defaults there).
#pragma once

#include <cstddef>
#include <istream>
#include <ostream>
#include <stdexcept>
#include <string>
#include <string_view>

#include "markov_cero/model/model.hpp"

namespace markov_cero::io {

struct MpsLimits final {
    std::size_t maximum_bytes{16U * 1024U * 1024U};
    std::size_t maximum_lines{1'000'000U};
    std::size_t maximum_rows{1'000'000U};
    std::size_t maximum_columns{1'000'000U};
    std::size_t maximum_nonzeros{20'000'000U};
    std::size_t maximum_name_bytes{255U};
};

class MpsError final : public std::runtime_error {
  public:
    MpsError(std::size_t line, const std::string& what)
        : std::runtime_error("MPS line " + std::to_string(line) + ": " + what), line_(line) {}
    [[nodiscard]] std::size_t line() const noexcept { return line_; }

  private:
    std::size_t line_;
};

[[nodiscard]] model::Model parse_mps(std::istream& in);
[[nodiscard]] model::Model parse_mps(std::istream& in, const MpsLimits& limits);
[[nodiscard]] model::Model parse_mps_string(std::string_view text, const MpsLimits& limits = {});
[[nodiscard]] model::Model read_mps_file(const std::string& path);
void write_mps(const model::Model& md, std::ostream& out);

// Binary LP: "MCLP0001", int64 m, n, nnz, int32 sense, pad, double offset,
// then column_start[n+1], row_index[nnz] (int64), value[nnz], objective[n],
// row_lower[m], row_upper[m], var_lower[n], var_upper[n] (16-byte bounds).
void write_binary_lp(const model::Model& md, const std::string& path);
[[nodiscard]] model::Model read_binary_lp(const std::string& path);

}  // namespace markov_cero::io
