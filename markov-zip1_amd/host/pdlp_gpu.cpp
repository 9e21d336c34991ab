// pdlp_gpu.cpp -- Backend::gpu of solve_pdlp: packs one or more LPs for the
// HIP engine (include/mcpdlp_engine.h), runs it and maps the outcome back to
// PdlpResult (reference pdlp.hpp:45-63 semantics, backend "hip").
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <exception>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>

#include "../../include/mcpdlp_engine.h"
#include "markov_cero/lp/first_order/pdlp.hpp"
#include "pdlp_internal.hpp"

namespace markov_cero::lp::first_order::detail {

static_assert(sizeof(model::Bound) == sizeof(mc_bound), "Bound layout");
static_assert(offsetof(model::Bound, value) == offsetof(mc_bound, value), "Bound layout");
static_assert(sizeof(std::size_t) == sizeof(int64_t), "LP64 index width");

mc_lp view_of(const model::Model& md) {
    mc_lp v{};
    v.rows = (int64_t)md.matrix.row_count;
    v.cols = (int64_t)md.matrix.column_count;
    v.column_start = reinterpret_cast<const int64_t*>(md.matrix.column_start.data());
    v.row_index = reinterpret_cast<const int64_t*>(md.matrix.row_index.data());
    v.value = md.matrix.value.data();
    v.objective = md.objective.data();
    v.row_lower = reinterpret_cast<const mc_bound*>(md.row_lower.data());
    v.row_upper = reinterpret_cast<const mc_bound*>(md.row_upper.data());
    v.var_lower = reinterpret_cast<const mc_bound*>(md.variable_lower.data());
    v.var_upper = reinterpret_cast<const mc_bound*>(md.variable_upper.data());
    v.sense = md.objective_sense == model::ObjectiveSense::maximize ? MC_MAXIMIZE : MC_MINIMIZE;
    v.objective_offset = md.objective_offset;
    return v;
}

template <class I>
void csc_to_csr(int64_t m, int64_t n, const I* cs, const I* ri, const double* val, int32_t* rp, int32_t* ci,
                double* rv, int threads) {
    const int64_t nnz = n > 0 ? cs[n] : 0;
    threads = std::max(1, std::min<int>(threads, (int)std::max<int64_t>(1, nnz / (1 << 20))));
    // counts[t][i]: entries of row i in worker t's column range; worker t's
    // entries of a row follow those of workers < t, so every row ends up in
    // increasing column order.
    std::vector<std::vector<int32_t>> counts(threads, std::vector<int32_t>(m, 0));
    std::vector<int64_t> cut(threads + 1);
    for (int t = 0; t <= threads; ++t) cut[t] = n * t / threads;
    auto run = [&](auto&& fn) {
        std::vector<std::thread> pool;
        for (int t = 1; t < threads; ++t) pool.emplace_back(fn, t);
        fn(0);
        for (auto& th : pool) th.join();
    };
    run([&](int t) {
        auto& c = counts[t];
        for (int64_t k = cs[cut[t]]; k < cs[cut[t + 1]]; ++k) ++c[ri[k]];
    });
    rp[0] = 0;
    for (int64_t i = 0; i < m; ++i) {
        int32_t acc = rp[i];
        for (int t = 0; t < threads; ++t) {
            const int32_t c = counts[t][i];
            counts[t][i] = acc;  // becomes worker t's next slot in row i
            acc += c;
        }
        rp[i + 1] = acc;
    }
    run([&](int t) {
        auto& next = counts[t];
        for (int64_t j = cut[t]; j < cut[t + 1]; ++j)
            for (int64_t k = cs[j]; k < cs[j + 1]; ++k) {
                const int32_t pos = next[ri[k]]++;
                ci[pos] = (int32_t)j;
                rv[pos] = val[k];
            }
    });
}

template void csc_to_csr<int64_t>(int64_t, int64_t, const int64_t*, const int64_t*, const double*, int32_t*,
                                  int32_t*, double*, int);
template void csc_to_csr<int32_t>(int64_t, int64_t, const int32_t*, const int32_t*, const double*, int32_t*,
                                  int32_t*, double*, int);

namespace {

template <class F>
void parallel_for(int64_t n, int threads, F&& fn) {
    threads = std::max(1, std::min<int>(threads, (int)std::max<int64_t>(1, n / 4096)));
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back([&, t] { fn(n * t / threads, n * (t + 1) / threads); });
    fn(0, n / threads);
    for (auto& th : pool) th.join();
}

// Length bucket of a row/column: exact below 32 nonzeros, log2 classes above.
inline int length_bucket(int64_t len) {
    if (len < 32) return (int)len;
    int b = 32;
    while (len >= 64) {
        len >>= 1;
        ++b;
    }
    return b;
}

// Stable order of [0, k) by descending length bucket: neighbouring rows of a
// tile then have similar lengths, and rows inside a bucket keep their order.
std::vector<int32_t> bucket_order(const std::vector<int64_t>& len) {
    const int64_t k = (int64_t)len.size();
    std::vector<int32_t> order(k);
    std::vector<int64_t> count(64, 0);
    for (int64_t i = 0; i < k; ++i) ++count[63 - length_bucket(len[i])];
    int64_t acc = 0;
    for (auto& c : count) {
        const int64_t t = c;
        c = acc;
        acc += t;
    }
    for (int64_t i = 0; i < k; ++i) order[count[63 - length_bucket(len[i])]++] = (int32_t)i;
    return order;
}

// True when the rows of `lp` are local: the mean over non-empty rows of
// (largest - smallest column index + 1) / columns is below 1/4.  Block-angular
// LPs (each row inside its period's columns) are; rows with random columns
// (span ~ all columns) are not.
bool rows_are_local(const mc_lp& lp) {
    const int64_t m = lp.rows, n = lp.cols;
    if (m == 0 || n == 0) return false;
    std::vector<int64_t> lo(m, n), hi(m, -1);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t k = lp.column_start[j]; k < lp.column_start[j + 1]; ++k) {
            const int64_t i = lp.row_index[k];
            lo[i] = std::min(lo[i], j);
            hi[i] = std::max(hi[i], j);
        }
    double span = 0.0;
    int64_t rows = 0;
    for (int64_t i = 0; i < m; ++i)
        if (hi[i] >= 0) {
            span += (double)(hi[i] - lo[i] + 1) / (double)n;
            ++rows;
        }
    return rows > 0 && span / (double)rows < 0.25;
}

struct EngineDeleter {
    void operator()(mcp_engine* e) const { mcp_engine_destroy(e); }
};

// One engine per (host thread, device): the engine owns a stream and device
// buffers that are reused by the next solve of the same thread.
mcp_engine* engine_for(int device) {
    thread_local std::map<int, std::unique_ptr<mcp_engine, EngineDeleter>> engines;
    auto it = engines.find(device);
    if (it != engines.end()) return it->second.get();
    int32_t count = 0;
    mcp_device_count(&count);
    if (count <= 0) throw DeviceUnavailable("no HIP device available");
    mcp_engine* e = nullptr;
    const int rc = mcp_engine_create(device, &e);
    if (rc == MC_ERR_DEVICE) throw DeviceUnavailable(mcp_last_error());
    if (rc != MC_OK) throw std::runtime_error(std::string("mcp_engine_create: ") + mcp_last_error());
    engines.emplace(device, std::unique_ptr<mcp_engine, EngineDeleter>(e));
    return e;
}

void check(int rc, const char* what) {
    if (rc == MC_OK) return;
    const std::string msg = std::string(what) + ": " + mcp_last_error();
    if (rc == MC_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

double bound_value(const mc_bound& b, double open) { return b.kind == MC_BOUND_FINITE ? b.value : open; }

}  // namespace

// Block-diagonal device layout of a batch of LPs (every one with m, n > 0):
// CSR(A) and CSR(A^T) with int32 indices, bounds with IEEE infinities, each
// instance's rows and columns contiguous.  In the fast reduction order
// (relabel) rows and columns are relabelled into length buckets
// (row_perm[new] = old, col_perm[new] = old, both global), so each SpMV tile
// holds rows of similar length -- unless the instance's rows already have
// column locality (rows_are_local): then a tile of consecutive rows gathers
// from a narrow window of x_bar, and the bucket order would interleave rows
// from all over the matrix.  The reference order keeps the model's labels:
// its sums must run in the reference's index order.
PackedBatch pack_batch(const std::vector<mc_lp>& lps, bool relabel, int threads) {
    const int B = (int)lps.size();
    int64_t m = 0, n = 0, nnz = 0;
    for (const auto& lp : lps) {
        m += lp.rows;
        n += lp.cols;
        nnz += lp.column_start[lp.cols];
    }
    constexpr int64_t kMax = std::numeric_limits<int32_t>::max();
    if (m >= kMax || n >= kMax || nnz >= kMax) throw std::invalid_argument("problem exceeds int32 indexing");
    PackedBatch pk;
    pk.m = m;
    pk.n = n;
    pk.nnz = nnz;
    auto& inst_row = pk.inst_row;
    auto& inst_col = pk.inst_col;
    auto& rp = pk.rp;
    auto& ci = pk.ci;
    auto& cp = pk.cp;
    auto& ri = pk.ri;
    auto& rv = pk.rv;
    auto& cv = pk.cv;
    auto& rlo = pk.rlo;
    auto& rhi = pk.rhi;
    auto& vlo = pk.vlo;
    auto& vhi = pk.vhi;
    auto& obj = pk.obj;
    auto& sense = pk.sense;
    auto& offset = pk.offset;
    auto& row_perm = pk.row_perm;
    auto& col_perm = pk.col_perm;
    inst_row.resize(B + 1);
    inst_col.resize(B + 1);
    rp.resize(m + 1);
    ci.resize(nnz);
    cp.resize(n + 1);
    ri.resize(nnz);
    rv.resize(nnz);
    cv.resize(nnz);
    rlo.resize(m);
    rhi.resize(m);
    vlo.resize(n);
    vhi.resize(n);
    obj.resize(n);
    sense.resize(B);
    offset.resize(B);
    row_perm.resize(m);
    col_perm.resize(n);
    const double inf = std::numeric_limits<double>::infinity();
    const int hw = std::max(1, threads);
    // Instance b occupies rows [r0s[b], r0s[b+1]), columns [c0s[b], c0s[b+1])
    // and nonzeros [p0s[b], p0s[b+1]) of the packed arrays; instances write
    // disjoint ranges, so a batch packs its instances on parallel threads
    // (largest first) and a single LP threads inside its own packing.
    std::vector<int64_t> r0s(B + 1, 0), c0s(B + 1, 0), p0s(B + 1, 0);
    for (int b = 0; b < B; ++b) {
        const mc_lp& lp = lps[b];
        r0s[b + 1] = r0s[b] + lp.rows;
        c0s[b + 1] = c0s[b] + lp.cols;
        p0s[b + 1] = p0s[b] + lp.column_start[lp.cols];
    }
    auto pack_one = [&](int b, int th) {
        const mc_lp& lp = lps[b];
        const int64_t r0 = r0s[b], c0 = c0s[b], p0 = p0s[b];
        inst_row[b] = (int32_t)r0;
        inst_col[b] = (int32_t)c0;
        const int64_t lm = lp.rows, ln = lp.cols, lnz = lp.column_start[ln];
        std::vector<int32_t> cperm(ln), rinv(lm);
        if (relabel && !rows_are_local(lp)) {
            std::vector<int64_t> clen(ln), rlen(lm, 0);
            for (int64_t j = 0; j < ln; ++j) clen[j] = lp.column_start[j + 1] - lp.column_start[j];
            for (int64_t k = 0; k < lnz; ++k) ++rlen[lp.row_index[k]];
            cperm = bucket_order(clen);
            const auto rperm = bucket_order(rlen);
            for (int64_t i = 0; i < lm; ++i) {
                rinv[rperm[i]] = (int32_t)i;
                row_perm[r0 + i] = (int32_t)(r0 + rperm[i]);
            }
        } else {
            for (int64_t j = 0; j < ln; ++j) cperm[j] = (int32_t)j;
            for (int64_t i = 0; i < lm; ++i) {
                rinv[i] = (int32_t)i;
                row_perm[r0 + i] = (int32_t)(r0 + i);
            }
        }
        // local CSC in the new labels (entries of a column sorted by new row)
        std::vector<int32_t> lcs(ln + 1), lri(lnz);
        std::vector<double> lcv(lnz);
        lcs[0] = 0;
        for (int64_t j = 0; j < ln; ++j) {
            const int64_t oj = cperm[j];
            lcs[j + 1] = lcs[j] + (int32_t)(lp.column_start[oj + 1] - lp.column_start[oj]);
        }
        parallel_for(ln, th, [&](int64_t j0, int64_t j1) {
            std::vector<std::pair<int32_t, double>> tmp;
            for (int64_t j = j0; j < j1; ++j) {
                const int64_t oj = cperm[j];
                const int64_t a = lp.column_start[oj], e = lp.column_start[oj + 1];
                tmp.clear();
                for (int64_t k = a; k < e; ++k) tmp.emplace_back(rinv[lp.row_index[k]], lp.value[k]);
                if (relabel) std::sort(tmp.begin(), tmp.end(), [](auto& u, auto& v) { return u.first < v.first; });
                for (int64_t k = 0; k < e - a; ++k) {
                    lri[lcs[j] + k] = tmp[k].first;
                    lcv[lcs[j] + k] = tmp[k].second;
                }
            }
        });
        // entries [r0, r0 + lm) of rp and [c0, c0 + ln) of cp only: the next
        // instance owns the boundary entry (rp[m] and cp[n] are set after all)
        for (int64_t j = 0; j < ln; ++j) cp[c0 + j] = (int32_t)(p0 + lcs[j]);
        for (int64_t k = 0; k < lnz; ++k) ri[p0 + k] = (int32_t)(r0 + lri[k]);
        std::memcpy(cv.data() + p0, lcv.data(), sizeof(double) * lnz);
        std::vector<int32_t> lrp(lm + 1);
        csc_to_csr<int32_t>(lm, ln, lcs.data(), lri.data(), lcv.data(), lrp.data(), ci.data() + p0, rv.data() + p0,
                            th);
        for (int64_t i = 0; i < lm; ++i) rp[r0 + i] = lrp[i] + (int32_t)p0;
        for (int64_t k = 0; k < lnz; ++k) ci[p0 + k] += (int32_t)c0;
        for (int64_t i = 0; i < lm; ++i) {
            const int64_t oi = row_perm[r0 + i] - r0;
            rlo[r0 + i] = bound_value(lp.row_lower[oi], -inf);
            rhi[r0 + i] = bound_value(lp.row_upper[oi], inf);
        }
        for (int64_t j = 0; j < ln; ++j) {
            const int64_t oj = cperm[j];
            col_perm[c0 + j] = (int32_t)(c0 + oj);
            vlo[c0 + j] = bound_value(lp.var_lower[oj], -inf);
            vhi[c0 + j] = bound_value(lp.var_upper[oj], inf);
            obj[c0 + j] = lp.objective[oj];
        }
        sense[b] = lp.sense == MC_MAXIMIZE ? -1.0 : 1.0;
        offset[b] = lp.objective_offset;
    };
    if (B == 1) {
        pack_one(0, hw);
    } else {
        std::vector<int> order(B);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int u, int v) { return p0s[u + 1] - p0s[u] > p0s[v + 1] - p0s[v]; });
        std::atomic<int> next{0};
        std::exception_ptr failure;
        std::mutex failure_mu;
        auto worker = [&] {
            for (int q = next.fetch_add(1); q < B; q = next.fetch_add(1)) {
                try {
                    pack_one(order[q], 1);
                } catch (...) {
                    std::lock_guard<std::mutex> lock(failure_mu);
                    if (!failure) failure = std::current_exception();
                }
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < std::min(hw, B); ++t) pool.emplace_back(worker);
        worker();
        for (auto& t : pool) t.join();
        if (failure) std::rethrow_exception(failure);
    }
    inst_row[B] = (int32_t)r0s[B];
    inst_col[B] = (int32_t)c0s[B];
    rp[m] = (int32_t)nnz;
    cp[n] = (int32_t)nnz;
    return pk;
}

std::vector<PdlpResult> solve_gpu(const std::vector<mc_lp>& lps, const PdlpOptions& opt) {
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<PdlpResult> out(lps.size());
    std::vector<int> live;  // instances that go to the device (m > 0 and n > 0)
    for (std::size_t k = 0; k < lps.size(); ++k) {
        auto& r = out[k];
        r.tolerance = opt.primal_tolerance;
        if (lps[k].rows == 0 || lps[k].cols == 0) {
            r.status = PdlpStatus::optimal;
            r.message = "trivial";
            r.backend_actually_used = "cpu_fallback";  // nothing ran on the device
        } else {
            live.push_back((int)k);
        }
    }
    if (live.empty()) return out;
    mcp_engine* eng = engine_for(opt.device);
    if (opt.reduction_order == ReductionOrder::reference && opt.precision != Precision::fp64)
        throw std::invalid_argument("reference reduction order requires FP64 precision");

    // ---- pack (block-diagonal) ----
    const int B = (int)live.size();
    std::vector<mc_lp> live_lps;
    live_lps.reserve(B);
    for (int k : live) live_lps.push_back(lps[k]);
    // host threads for packing: at most 16, like the verifier's and the name audit's pools
    const int hw = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    PackedBatch pk = pack_batch(live_lps, opt.reduction_order != ReductionOrder::reference, hw);
    const int64_t m = pk.m, n = pk.n, nnz = pk.nnz;
    auto& inst_row = pk.inst_row;
    auto& inst_col = pk.inst_col;
    auto& row_perm = pk.row_perm;
    auto& col_perm = pk.col_perm;
    auto& rp = pk.rp;
    auto& ci = pk.ci;
    auto& rv = pk.rv;
    auto& cp = pk.cp;
    auto& ri = pk.ri;
    auto& cv = pk.cv;
    auto& rlo = pk.rlo;
    auto& rhi = pk.rhi;
    auto& vlo = pk.vlo;
    auto& vhi = pk.vhi;
    auto& obj = pk.obj;
    auto& sense = pk.sense;
    auto& offset = pk.offset;
    mcp_problem pb{};
    pb.instances = B;
    pb.rows = (int32_t)m;
    pb.cols = (int32_t)n;
    pb.nnz = nnz;
    pb.inst_row = inst_row.data();
    pb.inst_col = inst_col.data();
    pb.row_ptr = rp.data();
    pb.col_idx = ci.data();
    pb.row_val = rv.data();
    pb.col_ptr = cp.data();
    pb.row_idx = ri.data();
    pb.col_val = cv.data();
    pb.row_lower = rlo.data();
    pb.row_upper = rhi.data();
    pb.var_lower = vlo.data();
    pb.var_upper = vhi.data();
    pb.objective = obj.data();
    pb.sense = sense.data();
    pb.offset = offset.data();
    check(mcp_load(eng, &pb), "mcp_load");

    mcp_setup_opts so{};
    so.ruiz = opt.ruiz_scaling ? 1 : 0;
    so.ruiz_iterations = (int32_t)opt.ruiz_iterations;
    so.exact = opt.reduction_order == ReductionOrder::reference ? 1 : 0;
    so.autotune = opt.autotune ? 1 : 0;
    so.row_weighting = opt.row_weighting;
    check(mcp_setup(eng, &so), "mcp_setup");

    // warm start: vectors concatenated over all input LPs in input order
    std::vector<double> wx, wy;
    int64_t totn = 0, totm = 0;
    std::vector<int64_t> ocol(lps.size()), orow(lps.size());
    for (std::size_t k = 0; k < lps.size(); ++k) {
        ocol[k] = totn;
        orow[k] = totm;
        totn += lps[k].cols;
        totm += lps[k].rows;
    }
    const bool warm_x = !opt.warm_primal.empty() && opt.warm_primal.size() == (std::size_t)totn;
    const bool warm_y = !opt.warm_dual.empty() && opt.warm_dual.size() == (std::size_t)totm;
    if ((!opt.warm_primal.empty() && !warm_x) || (!opt.warm_dual.empty() && !warm_y))
        throw std::invalid_argument("warm start vectors must hold the columns/rows of all models");
    mcp_solve_opts o{};
    o.max_iterations = opt.max_iterations;
    o.restart_every = opt.restart_every;
    o.primal_tolerance = opt.primal_tolerance;
    o.dual_tolerance = opt.dual_tolerance;
    o.gap_tolerance = opt.gap_tolerance;
    o.step_size_reduction = opt.step_size_reduction;
    o.restart_strategy = opt.restart_strategy == RestartStrategy::none    ? MC_RESTART_NONE
                         : opt.restart_strategy == RestartStrategy::fixed ? MC_RESTART_FIXED
                                                                          : MC_RESTART_ADAPTIVE;
    o.adaptive_step_size = opt.adaptive_step_size ? 1 : 0;
    o.restart_reduction_factor = opt.restart_reduction_factor;
    o.adaptive_primal_weight = opt.adaptive_primal_weight ? 1 : 0;
    o.precision = opt.precision == Precision::mixed        ? MC_PRECISION_MIXED
                  : opt.precision == Precision::packed ? MC_PRECISION_PACKED
                                                           : MC_PRECISION_FP64;
    o.initial_primal_weight = opt.initial_primal_weight;
    o.primal_weight_smoothing = opt.primal_weight_smoothing;
    o.polish_threshold = opt.polish_threshold;
    o.fp32_stall_checks = opt.fp32_stall_checks;
    o.detect_infeasibility = opt.detect_infeasibility ? 1 : 0;
    o.infeasibility_tolerance = opt.infeasibility_tolerance;
    o.power_iterations = opt.power_iterations;
    o.primal_weight_max_drop = opt.primal_weight_max_drop;
    o.reduction_order = so.exact ? MC_ORDER_REFERENCE : MC_ORDER_FAST;
    o.graph_chunk = opt.graph_chunk;
    o.restart_scheme = opt.restart_scheme == RestartScheme::kkt       ? MC_SCHEME_KKT
                       : opt.restart_scheme == RestartScheme::halpern ? MC_SCHEME_HALPERN
                                                                      : MC_SCHEME_REFERENCE;
    if (warm_x) {
        wx.resize(n);
        for (int b = 0; b < B; ++b)
            for (int64_t p = inst_col[b]; p < inst_col[b + 1]; ++p)
                wx[p] = opt.warm_primal[ocol[live[b]] + (col_perm[p] - inst_col[b])];
    }
    if (warm_y) {
        wy.resize(m);
        for (int b = 0; b < B; ++b)
            for (int64_t p = inst_row[b]; p < inst_row[b + 1]; ++p)
                wy[p] = opt.warm_dual[orow[live[b]] + (row_perm[p] - inst_row[b])];
    }
    o.warm_primal = warm_x ? wx.data() : nullptr;
    o.warm_dual = warm_y ? wy.data() : nullptr;
    mcp_solve_stats st{};
    check(mcp_solve(eng, &o, &st), "mcp_solve");
    std::vector<double> xp(n), yp(m), x(n), y(m);
    std::vector<mcp_inst_result> ir(B);
    double d2h = 0.0;
    check(mcp_fetch(eng, xp.data(), yp.data(), ir.data(), &d2h), "mcp_fetch");
    for (int64_t j = 0; j < n; ++j) x[col_perm[j]] = xp[j];
    for (int64_t i = 0; i < m; ++i) y[row_perm[i]] = yp[i];
    const double total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (int b = 0; b < B; ++b) {
        auto& r = out[live[b]];
        const auto& q = ir[b];
        r.backend_actually_used = "hip";
        r.status = q.status == MC_PDLP_OPTIMAL                   ? PdlpStatus::optimal
                   : q.status == MC_PDLP_NUMERICAL_FAILURE        ? PdlpStatus::numerical_failure
                   : q.status == MC_PDLP_INFEASIBLE_OR_UNBOUNDED ? PdlpStatus::infeasible_or_unbounded
                                                                  : PdlpStatus::iteration_limit;
        r.message = r.status == PdlpStatus::optimal                   ? "PDLP converged"
                    : r.status == PdlpStatus::numerical_failure         ? "numerical failure (non-finite residual)"
                    : r.status == PdlpStatus::infeasible_or_unbounded ? "infeasibility certificate found"
                                                                        : "iteration limit reached";
        r.primal.assign(x.begin() + inst_col[b], x.begin() + inst_col[b + 1]);
        r.dual.assign(y.begin() + inst_row[b], y.begin() + inst_row[b + 1]);
        r.objective = q.objective;
        r.primal_infeasibility = q.primal_infeasibility;
        r.dual_infeasibility = q.dual_infeasibility;
        r.duality_gap = q.duality_gap;
        r.iterations = (std::size_t)q.iterations;
        r.fp32_iterations = (std::size_t)q.fp32_iterations;
        r.packed_iterations = (std::size_t)q.packed_iterations;
        r.primal_weight = q.omega;
        r.step_size = q.eta;
        r.h2d_ms = st.h2d_ms;
        r.setup_ms = st.setup_ms;
        r.kernel_ms = st.kernel_ms;
        r.d2h_ms = d2h;
        r.total_ms = total;
    }
    return out;
}

}  // namespace markov_cero::lp::first_order::detail
