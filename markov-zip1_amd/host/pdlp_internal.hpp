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

// Block-diagonal layout of a batch for mcp_load: instance b holds rows
// [inst_row[b], inst_row[b+1]) and columns [inst_col[b], inst_col[b+1]);
// rp/ci/rv is CSR(A), cp/ri/cv CSR(A^T), row_perm[new] / col_perm[new] the
// original (global) index of each packed row / column.
struct PackedBatch {
    int64_t m = 0, n = 0, nnz = 0;
    std::vector<int32_t> inst_row, inst_col, rp, ci, cp, ri, row_perm, col_perm;
    std::vector<double> rv, cv, rlo, rhi, vlo, vhi, obj, sense, offset;
};
// Packs LPs with m, n > 0 on up to `threads` host threads (one instance per
// thread in a batch; a single LP threads inside its own packing), relabelled
// into row/column length buckets when `relabel` (the fast reduction order).
PackedBatch pack_batch(const std::vector<mc_lp>& lps, bool relabel, int threads);
std::vector<PdlpResult> solve_gpu(const std::vector<mc_lp>& lps, const PdlpOptions& opt);

// CSC -> CSR (int32), entries of a row in increasing column order;
// `threads` workers split the columns.
template <class I>
void csc_to_csr(int64_t m, int64_t n, const I* col_start, const I* row_index, const double* val,
                int32_t* row_ptr, int32_t* col_idx, double* row_val, int threads);

}  // namespace markov_cero::lp::first_order::detail
