// mps.hpp -- model ingest for the PDLP path: free-format MPS (reader and
// writer) and a raw binary LP format for instances past MPS's practical size.
// Row/bound semantics follow the reference reader (src/io/mps.cpp:390-440):
// N row = objective, E/L/G rows with RANGES, LO/UP/FX/FR/MI/PL/BV/LI/UI bounds,
// default variable bounds [0, +inf).  Unlike the reference there are no size
// limits: large instances are the point of the GPU path.
#pragma once

#include <istream>
#include <ostream>
#include <stdexcept>
#include <string>

#include "markov_cero/model.hpp"

namespace markov_cero::io {

struct MpsError : std::runtime_error {
    MpsError(std::size_t line, const std::string& what)
        : std::runtime_error("MPS line " + std::to_string(line) + ": " + what) {}
};

[[nodiscard]] model::Model parse_mps(std::istream& in);
[[nodiscard]] model::Model read_mps_file(const std::string& path);
void write_mps(const model::Model& md, std::ostream& out);

// Binary LP: "MCLP0001", int64 m, n, nnz, int32 sense, pad, double offset,
// then column_start[n+1], row_index[nnz] (int64), value[nnz], objective[n],
// row_lower[m], row_upper[m], var_lower[n], var_upper[n] (16-byte bounds).
void write_binary_lp(const model::Model& md, const std::string& path);
[[nodiscard]] model::Model read_binary_lp(const std::string& path);

}  // namespace markov_cero::io
