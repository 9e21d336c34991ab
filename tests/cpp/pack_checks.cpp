// Checks of the batch packing that the GPU solve loads (detail::pack_batch,
// compiled and run by tests/test_host_cpp.py on the CPU):
// * packing on 16 threads equals packing on one, array for array;
// * both stored orientations are well formed: row/column pointers start at 0,
//   never decrease, end at nnz, and meet at every instance boundary; indices
//   stay inside their instance and increase within a row (column);
// * through row_perm / col_perm every packed entry is the original entry of
//   its instance's matrix, every original entry appears once, and the bounds,
//   objective, sense and offset are the original ones;
// on a batch of 40 uniform and Zipf LPs and on a single Zipf LP, in both
// reduction orders (relabelled by length bucket, and in the original labels).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <vector>

#include "pdlp_internal.hpp"

namespace fo = markov_cero::lp::first_order;

static int failures = 0;
static void expect(bool ok, const char* what) {
    if (!ok) {
        std::printf("FAIL: %s\n", what);
        ++failures;
    }
}

static double bound_of(const mc_bound& b, double open) { return b.kind == MC_BOUND_FINITE ? b.value : open; }

// A(i, j) of an LP in CSC (rows strictly increasing within a column); NaN if absent.
static double entry(const mc_lp& lp, int64_t i, int64_t j) {
    const int64_t* b = lp.row_index + lp.column_start[j];
    const int64_t* e = lp.row_index + lp.column_start[j + 1];
    const int64_t* it = std::lower_bound(b, e, i);
    return (it != e && *it == i) ? lp.value[lp.column_start[j] + (it - lp.row_index - lp.column_start[j])]
                                 : std::numeric_limits<double>::quiet_NaN();
}

static bool same(const fo::detail::PackedBatch& a, const fo::detail::PackedBatch& b) {
    return a.m == b.m && a.n == b.n && a.nnz == b.nnz && a.inst_row == b.inst_row && a.inst_col == b.inst_col &&
           a.rp == b.rp && a.ci == b.ci && a.cp == b.cp && a.ri == b.ri && a.row_perm == b.row_perm &&
           a.col_perm == b.col_perm && a.rv == b.rv && a.cv == b.cv && a.rlo == b.rlo && a.rhi == b.rhi &&
           a.vlo == b.vlo && a.vhi == b.vhi && a.obj == b.obj && a.sense == b.sense && a.offset == b.offset;
}

static void check_layout(const std::vector<mc_lp>& lps, const fo::detail::PackedBatch& pk, const char* label) {
    const int B = (int)lps.size();
    const double inf = std::numeric_limits<double>::infinity();
    bool ok_ptr = pk.rp.size() == (std::size_t)pk.m + 1 && pk.cp.size() == (std::size_t)pk.n + 1 && pk.rp[0] == 0 &&
                  pk.cp[0] == 0 && pk.rp[pk.m] == pk.nnz && pk.cp[pk.n] == pk.nnz;
    for (int64_t i = 0; ok_ptr && i < pk.m; ++i) ok_ptr = pk.rp[i] <= pk.rp[i + 1];
    for (int64_t j = 0; ok_ptr && j < pk.n; ++j) ok_ptr = pk.cp[j] <= pk.cp[j + 1];
    expect(ok_ptr, label);
    if (!ok_ptr) return;
    bool ok_inst = true, ok_csr = true, ok_csc = true, ok_bounds = true;
    int64_t r0 = 0, c0 = 0, p0 = 0;
    for (int b = 0; b < B; ++b) {
        const mc_lp& lp = lps[b];
        ok_inst = ok_inst && pk.inst_row[b] == r0 && pk.inst_col[b] == c0 && pk.rp[r0] == p0 && pk.cp[c0] == p0;
        const int64_t lnz = lp.column_start[lp.cols];
        int64_t seen_csr = 0, seen_csc = 0;
        for (int64_t i = r0; i < r0 + lp.rows; ++i) {
            const int64_t oi = pk.row_perm[i] - r0;
            ok_inst = ok_inst && oi >= 0 && oi < lp.rows;
            for (int32_t p = pk.rp[i]; p < pk.rp[i + 1]; ++p) {
                const int64_t j = pk.ci[p];
                ok_csr = ok_csr && j >= c0 && j < c0 + lp.cols && (p == pk.rp[i] || pk.ci[p - 1] < j);
                if (!ok_csr) break;
                ok_csr = entry(lp, oi, pk.col_perm[j] - c0) == pk.rv[p];
                ++seen_csr;
            }
            ok_bounds = ok_bounds && pk.rlo[i] == bound_of(lp.row_lower[oi], -inf) &&
                        pk.rhi[i] == bound_of(lp.row_upper[oi], inf);
        }
        for (int64_t j = c0; j < c0 + lp.cols; ++j) {
            const int64_t oj = pk.col_perm[j] - c0;
            ok_inst = ok_inst && oj >= 0 && oj < lp.cols;
            for (int32_t p = pk.cp[j]; p < pk.cp[j + 1]; ++p) {
                const int64_t i = pk.ri[p];
                ok_csc = ok_csc && i >= r0 && i < r0 + lp.rows && (p == pk.cp[j] || pk.ri[p - 1] < i);
                if (!ok_csc) break;
                ok_csc = entry(lp, pk.row_perm[i] - r0, oj) == pk.cv[p];
                ++seen_csc;
            }
            ok_bounds = ok_bounds && pk.vlo[j] == bound_of(lp.var_lower[oj], -inf) &&
                        pk.vhi[j] == bound_of(lp.var_upper[oj], inf) && pk.obj[j] == lp.objective[oj];
        }
        // every original entry once: as many entries as the original, each found in it
        ok_csr = ok_csr && seen_csr == lnz;
        ok_csc = ok_csc && seen_csc == lnz;
        ok_bounds = ok_bounds && pk.sense[b] == (lp.sense == MC_MAXIMIZE ? -1.0 : 1.0) &&
                    pk.offset[b] == lp.objective_offset;
        r0 += lp.rows;
        c0 += lp.cols;
        p0 += lnz;
    }
    ok_inst = ok_inst && pk.inst_row[B] == pk.m && pk.inst_col[B] == pk.n && r0 == pk.m && c0 == pk.n;
    expect(ok_inst, label);
    expect(ok_csr, label);
    expect(ok_csc, label);
    expect(ok_bounds, label);
}

int main() {
    std::vector<mc_lp_handle*> handles;
    std::vector<mc_lp> batch;
    for (int k = 0; k < 40; ++k) {
        mc_gen_spec spec;
        mc_gen_spec_default(&spec, k % 2 ? MC_FAMILY_UNIFORM_EQ : MC_FAMILY_ZIPF_MIXED);
        spec.rows = 50 + 37 * k;
        spec.cols = (spec.rows * 5) / 2;
        if (spec.family == MC_FAMILY_ZIPF_MIXED) spec.nnz_max = 60;
        spec.seed = 1000 + k;
        mc_lp_handle* h = nullptr;
        if (mc_lp_generate(&spec, &h) != MC_OK) {
            std::printf("FAIL: generate %d\n", k);
            return 1;
        }
        mc_lp lp;
        mc_lp_view(h, &lp);
        handles.push_back(h);
        batch.push_back(lp);
    }
    mc_gen_spec big;
    mc_gen_spec_default(&big, MC_FAMILY_ZIPF_MIXED);
    big.rows = 20000;
    big.cols = 50000;
    big.seed = 77;
    mc_lp_handle* hb = nullptr;
    if (mc_lp_generate(&big, &hb) != MC_OK) {
        std::printf("FAIL: generate single\n");
        return 1;
    }
    mc_lp single;
    mc_lp_view(hb, &single);
    handles.push_back(hb);

    for (bool relabel : {true, false}) {
        const auto one = fo::detail::pack_batch(batch, relabel, 1);
        const auto many = fo::detail::pack_batch(batch, relabel, 16);
        expect(same(one, many), relabel ? "batch, relabelled: 16 threads == 1" : "batch: 16 threads == 1");
        check_layout(batch, many, relabel ? "batch layout, relabelled" : "batch layout");
        const auto s1 = fo::detail::pack_batch({single}, relabel, 1);
        const auto s16 = fo::detail::pack_batch({single}, relabel, 16);
        expect(same(s1, s16), relabel ? "single, relabelled: 16 threads == 1" : "single: 16 threads == 1");
        check_layout({single}, s16, relabel ? "single layout, relabelled" : "single layout");
    }
    for (auto* h : handles) mc_lp_free(h);
    if (failures == 0) std::printf("PASS\n");
    return failures == 0 ? 0 : 1;
}
