// pdlp_internal.hpp -- pieces shared by the two backends and the C-ABI.
#pragma once

#include <cstddef>
#include <stdexcept>
#include <vector>

#include "../../include/mcpdlp.h"
#include "markov_cero/model/model.hpp"
#include "markov_cero/lp/first_order/pdlp.hpp"

namespace markov_cero::lp::first_order::detail {

// Thrown when Backend::gpu cannot get a device (no HIP device, runtime init
// failure); solve_pdlp turns it into the cpu_fallback path unless require_gpu.
struct DeviceUnavailable : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// model::Bound and mc_bound share one layout (int kind, pad, double value), and
// size_t equals int64_t in width on the LP64 targets this builds for, so a
// Model's arrays can be viewed as an mc_lp without copying.
mc_lp view_of(const model::Model& md);
// Rejects option values outside their documented ranges (std::invalid_argument).
void validate_options(const PdlpOptions& opt);
model::Model model_of(const mc_lp& lp);

PdlpResult solve_cpu(const model::Model& md, const PdlpOptions& opt, const char* backend_label);
std::vector<PdlpResult> solve_gpu(const std::vector<mc_lp>& lps, const PdlpOptions& opt);

// CSC -> CSR (int32), entries of a row in increasing column order;
// `threads` workers split the columns.
template <class I>
void csc_to_csr(int64_t m, int64_t n, const I* col_start, const I* row_index, const double* val,
                int32_t* row_ptr, int32_t* col_idx, double* row_val, int threads);

}  // namespace markov_cero::lp::first_order::detail
