// engine.hip -- host side of the HIP PDHG engine: device buffers, tiling,
// the chunked iteration loop (HIP-graph captured) and the C-ABI of
// include/mcpdlp_engine.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/mcpdlp.h"
#include "../../include/mcpdlp_engine.h"
#include "dev.hpp"

namespace {

thread_local std::string g_err;

struct EngineError : std::runtime_error {
    int code;
    EngineError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        const hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess)                                                            \
            throw EngineError(e_ == hipErrorOutOfMemory ? MC_ERR_OUT_OF_MEMORY : MC_ERR_DEVICE, \
                              std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

void require(bool ok, const std::string& msg) {
    if (!ok) throw EngineError(MC_ERR_INVALID_ARGUMENT, msg);
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return MC_OK;
    } catch (const EngineError& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return MC_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        g_err = e.what();
        return MC_ERR_INTERNAL;
    }
}

// Tile policies: throughput shape (long tasks, few waves) for problems that
// fill the chip, latency shape (one short sweep per wave, many waves) for
// problems of up to kLatencyNnz nonzeros, where a launch is a handful of
// dependent memory round trips rather than a stream.
constexpr mcp_tile_policy kThroughputPolicy{1024, 8};
constexpr mcp_tile_policy kLatencyPolicy{128, 2};
constexpr int64_t kLatencyNnz = 2 * 1024 * 1024;

int lanes_log2(int len, int per_lane) {
    int lg = 0;
    while (lg < 6 && len > (per_lane << lg)) ++lg;
    return lg;
}

// Cut rows [inst[b], inst[b+1]) of every instance into wave tasks:
// consecutive rows whose nonzeros fit `budget`, at most (64 >> lg) *
// kMaxSteps rows, lanes per row 2^lg from the tile's longest row and the
// per-lane target; a row longer than the budget is a tile of its own with 64
// lanes.  exact => 1 lane per row (the reference's serial order).
void build_tiles(const std::vector<int32_t>& ptr, const std::vector<int32_t>& inst, int B,
                 mcp_tile_policy pol, bool exact, std::vector<mcp::Tile>& tiles,
                 std::vector<int32_t>& begin, const std::vector<char>* keep = nullptr) {
    const int budget = std::max(1, pol.nnz_budget);
    const int per = std::max(1, pol.per_lane);
    tiles.clear();
    begin.assign(B + 1, 0);
    for (int b = 0; b < B; ++b) {
        begin[b] = (int32_t)tiles.size();
        if (keep && !(*keep)[b]) continue;  // instance left out: empty tile range
        int r = inst[b];
        const int re = inst[b + 1];
        while (r < re) {
            const int len0 = ptr[r + 1] - ptr[r];
            if (len0 > budget) {
                tiles.push_back({r, r + 1, b, exact ? 0 : 6});
                ++r;
                continue;
            }
            const int r0 = r;
            int nnz = 0, maxlen = 0, lg = 0;
            while (r < re) {
                const int len = ptr[r + 1] - ptr[r];
                if (len > budget) break;
                const int nm = std::max(maxlen, len);
                const int lgn = exact ? 0 : lanes_log2(nm, per);
                if (nnz + len > budget || (r - r0 + 1) > ((64 >> lgn) * mcp::kMaxSteps)) break;
                nnz += len;
                maxlen = nm;
                lg = lgn;
                ++r;
            }
            tiles.push_back({r0, r, b, lg});
        }
    }
    begin[B] = (int32_t)tiles.size();
    // invariants the kernels rely on
    for (const auto& t : tiles) {
        const int lg = t.shape & 7;
        const bool single = t.r1 == t.r0 + 1;
        if (t.r1 <= t.r0 || lg > 6 || (t.shape & ~7) != 0 ||
            (!single && (t.r1 - t.r0) > ((64 >> lg) * mcp::kMaxSteps)) || (exact && lg != 0))
            throw EngineError(MC_ERR_INTERNAL, "tile invariant violated");
    }
}

// Block tiles of the fused fast-order primal/dual steps (stream.hip): runs of
// consecutive rows of one instance holding at most `budget` nonzeros and at
// most kBlock rows, one workgroup each.  Lanes per row G = 2^lg: as many as
// the workgroup has per row (G * rows <= kBlock), at most 64 and no more than
// the tile's longest row needs.  A row longer than the budget is a tile of
// its own (shape bit 3) that the whole workgroup strides.
// Budget by size (FP64 us per iteration, budgets 256 / 512 / 1024: 7.77 /
// 8.18 / 9.08 at 25k nonzeros, 7.85 / 8.37 / 9.53 at 52k, 10.46 / 9.95 /
// 10.34 at 209k): launches over small problems are latency chains, and
// smaller tiles shorten them until the tile count outgrows the gain.
constexpr int kBlockBudget = 1024;
constexpr int kBlockBudgetLatency = 512;   // up to kLatencyNnz nonzeros
constexpr int kBlockBudgetTiny = 256;      // up to kTinyNnz nonzeros
constexpr int64_t kTinyNnz = 128 * 1024;

int block_budget_for(int64_t nnz) {
    return nnz <= kTinyNnz ? kBlockBudgetTiny : nnz <= kLatencyNnz ? kBlockBudgetLatency : kBlockBudget;
}

int floor_log2(int v) {
    int l = 0;
    while ((2 << l) <= v) ++l;
    return l;
}

void build_block_tiles(const std::vector<int32_t>& ptr, const std::vector<int32_t>& inst, int B, int budget,
                       std::vector<mcp::Tile>& tiles, std::vector<int32_t>& begin,
                       const std::vector<char>* keep = nullptr) {
    tiles.clear();
    begin.assign(B + 1, 0);
    for (int b = 0; b < B; ++b) {
        begin[b] = (int32_t)tiles.size();
        if (keep && !(*keep)[b]) continue;
        int r = inst[b];
        const int re = inst[b + 1];
        while (r < re) {
            const int len0 = ptr[r + 1] - ptr[r];
            if (len0 > budget) {
                tiles.push_back({r, r + 1, b, 8});
                ++r;
                continue;
            }
            const int r0 = r;
            int nnz = 0, maxlen = 0;
            while (r < re && (r - r0) < mcp::kBlock) {
                const int len = ptr[r + 1] - ptr[r];
                if (len > budget || nnz + len > budget) break;
                nnz += len;
                maxlen = std::max(maxlen, len);
                ++r;
            }
            const int rows = r - r0;
            int lg = std::min(6, floor_log2(mcp::kBlock / rows));
            while (lg > 0 && (1 << (lg - 1)) >= maxlen) --lg;  // G no larger than the longest row needs
            tiles.push_back({r0, r, b, lg});
        }
    }
    begin[B] = (int32_t)tiles.size();
    for (const auto& t : tiles) {  // invariants the block kernels rely on
        const int lg = t.shape & 7;
        const int nz = ptr[t.r1] - ptr[t.r0];
        const bool lng = (t.shape & 8) != 0;
        if (t.r1 <= t.r0 || (t.shape & ~15) != 0 || (lng && (t.r1 != t.r0 + 1 || lg != 0)) ||
            (!lng && (nz > budget || ((t.r1 - t.r0) << lg) > mcp::kBlock || lg > 6)))
            throw EngineError(MC_ERR_INTERNAL, "block tile invariant violated");
    }
}

double ms_between(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, a, b));
    return (double)ms;
}

}  // namespace

struct mcp_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    struct Buf {
        void* p = nullptr;
        size_t bytes = 0;
    };
    std::map<std::string, Buf> bufs;
    mcp::Dev d{};
    int32_t B = 0, m = 0, n = 0;
    int64_t nnz = 0;
    std::vector<int32_t> inst_row, inst_col, rp, cp;
    std::vector<mcp::Tile> tA, tT;
    std::vector<int32_t> tA_begin, tT_begin;
    mcp_tile_policy polA{1024, 8}, polT{1024, 8};
    // block tiles of the fused fast-order steps; budget 0 => those steps run
    // on the wave tiles
    std::vector<mcp::Tile> bA, bT;
    std::vector<int32_t> bA_begin, bT_begin;
    int budA = kBlockBudget, budT = kBlockBudget;
    bool loaded = false, setup_done = false, exact_tiles = false;
    bool compacted = false;  // tiles cover only the instances still running (inside mcp_solve)
    int ruiz_passes_req = -1, ruiz_on_req = -1, row_weighting_req = -1;
    double h2d_ms = 0.0, setup_ms = 0.0;
    std::map<std::tuple<int64_t, int, int, int, int>, hipGraphExec_t> graphs;
    mcp::Dev graph_dev{};      // kernel arguments the cached graphs were captured with:
    mcp::IterOpts graph_io{};  //   the problem's device view and the solve options
    // autotuned tile policies per problem signature (instances, rows, columns,
    // nonzeros, pointer checksum), so repeated solves of one shape tune once
    struct Tuned {
        mcp_tile_policy polA, polT;
        int budA, budT;
    };
    std::map<std::tuple<int, int, int, int64_t, uint64_t>, Tuned> tuned;
    int32_t* pinned = nullptr;  // [2 slots x 2 counters]
    hipEvent_t ev[4]{};

    template <class T>
    T* buf(const std::string& name, size_t count) {
        const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        Buf& b = bufs[name];
        if (b.bytes < bytes) {
            if (b.p) HIPCHK(hipFree(b.p));
            b.p = nullptr;
            b.bytes = 0;
            HIPCHK(hipMalloc(&b.p, bytes));
            b.bytes = bytes;
        }
        return static_cast<T*>(b.p);
    }
    void drop_graphs() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }
    ~mcp_engine() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graphs();
        for (auto& kv : bufs)
            if (kv.second.p) (void)hipFree(kv.second.p);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (pinned) (void)hipHostFree(pinned);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void upload_tiles() {
        auto* dA = buf<mcp::Tile>("tA", tA.size());
        auto* dT = buf<mcp::Tile>("tT", tT.size());
        auto* tbA = buf<int32_t>("tA_begin", B + 1);
        auto* tbT = buf<int32_t>("tT_begin", B + 1);
        if (!tA.empty())
            HIPCHK(hipMemcpyAsync(dA, tA.data(), tA.size() * sizeof(mcp::Tile), hipMemcpyHostToDevice, stream));
        if (!tT.empty())
            HIPCHK(hipMemcpyAsync(dT, tT.data(), tT.size() * sizeof(mcp::Tile), hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(tbA, tA_begin.data(), (B + 1) * sizeof(int32_t), hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(tbT, tT_begin.data(), (B + 1) * sizeof(int32_t), hipMemcpyHostToDevice, stream));
        auto* dbA = buf<mcp::Tile>("bA", bA.size());
        auto* dbT = buf<mcp::Tile>("bT", bT.size());
        auto* bbA = buf<int32_t>("bA_begin", B + 1);
        auto* bbT = buf<int32_t>("bT_begin", B + 1);
        const bool blk = !bA_begin.empty();
        if (!bA.empty())
            HIPCHK(hipMemcpyAsync(dbA, bA.data(), bA.size() * sizeof(mcp::Tile), hipMemcpyHostToDevice, stream));
        if (!bT.empty())
            HIPCHK(hipMemcpyAsync(dbT, bT.data(), bT.size() * sizeof(mcp::Tile), hipMemcpyHostToDevice, stream));
        if (blk) {
            HIPCHK(hipMemcpyAsync(bbA, bA_begin.data(), (B + 1) * sizeof(int32_t), hipMemcpyHostToDevice, stream));
            HIPCHK(hipMemcpyAsync(bbT, bT_begin.data(), (B + 1) * sizeof(int32_t), hipMemcpyHostToDevice, stream));
        }
        d.bA = dbA;
        d.bT = dbT;
        d.bA_begin = bbA;
        d.bT_begin = bbT;
        d.nbA = (int32_t)bA.size();
        d.nbT = (int32_t)bT.size();
        d.budA = blk ? budA : 0;
        d.budT = blk ? budT : 0;
        d.blocks = blk ? 1 : 0;
        // partial-sum slots follow the tile counts
        d.part_dxn = buf<double>("part_dxn", std::max(tT.size(), bT.size()));
        d.part_adx = buf<double>("part_adx", std::max(tA.size(), bA.size()));
        // three candidates per check: the average, the current iterate, the
        // average on the packed stream's values
        d.part_rows = buf<double>("part_rows", 9 * std::max(tA.size(), bA.size()));
        d.part_cols = buf<double>("part_cols", 15 * std::max(tT.size(), bT.size()));
        d.part_ray_rows = buf<double>("part_ray_rows", 5 * tA.size());
        d.part_ray_cols = buf<double>("part_ray_cols", 7 * tT.size());
        d.tA = dA;
        d.tT = dT;
        d.tA_begin = tbA;
        d.tT_begin = tbT;
        d.ntA = (int32_t)tA.size();
        d.ntT = (int32_t)tT.size();
        HIPCHK(hipStreamSynchronize(stream));
    }

    // keep == nullptr: every instance; else only the instances with keep[b].
    void retile(bool exact, const std::vector<char>* keep = nullptr) {
        build_tiles(rp, inst_row, B, polA, exact, tA, tA_begin, keep);
        build_tiles(cp, inst_col, B, polT, exact, tT, tT_begin, keep);
        if (!exact && budA > 0 && budT > 0) {
            build_block_tiles(rp, inst_row, B, budA, bA, bA_begin, keep);
            build_block_tiles(cp, inst_col, B, budT, bT, bT_begin, keep);
        } else {
            bA.clear();
            bT.clear();
            bA_begin.clear();
            bT_begin.clear();
        }
        exact_tiles = exact;
        compacted = keep != nullptr;
        // Non-temporal FP64 streams of A when an FP64 iteration's working set
        // (both matrices at 12 B per nonzero and the vectors: 24 nnz + 100 n +
        // 84 m bytes over the instances tiled) exceeds the 256 MiB Infinity
        // Cache: A^T then stays resident for the primal step (on the 200k x
        // 500k LP the FP64 primal step takes 71 instead of 95 us, the dual step
        // is unchanged); where everything fits, or one matrix alone does not,
        // the plain loads are as fast.
        double live = 0.0;
        for (int b = 0; b < B; ++b)
            if (!keep || (*keep)[b])
                live += 24.0 * (rp[inst_row[b + 1]] - rp[inst_row[b]]) + 100.0 * (inst_col[b + 1] - inst_col[b]) +
                        84.0 * (inst_row[b + 1] - inst_row[b]);
        d.nt_rows = live > 256.0 * 1024.0 * 1024.0 ? 1 : 0;
        upload_tiles();
    }

    // Time the plain SpMV of one orientation under each candidate policy and
    // keep the fastest (FP32 values: the mixed-precision iteration's shape).
    // Each candidate's time is the best of three runs of five launches, and a
    // candidate replaces the incumbent (the size-based default policy, timed
    // first) only when it is at least 2 % faster, so run-to-run noise does not
    // flip the choice between nearly equal shapes.
    void autotune_one(bool rows) {
        const mcp_tile_policy cands[] = {rows ? polA : polT, {128, 2},  {256, 4},   {256, 8},  {256, 16},
                                         {256, 32},          {1024, 4}, {1024, 8},  {1024, 16}, {1024, 32},
                                         {4096, 4},          {4096, 8}, {4096, 16}, {4096, 32}};
        const auto& ptr = rows ? rp : cp;
        const auto& inst = rows ? inst_row : inst_col;
        const int nrows = rows ? m : n;
        const int ncols = rows ? n : m;
        if (nrows == 0) return;
        float* vin = buf<float>("tune_in", ncols);
        float* vout = buf<float>("tune_out", nrows);
        HIPCHK(hipMemsetAsync(vin, 0, sizeof(float) * std::max(ncols, 1), stream));
        const int32_t* dptr = rows ? d.rp : d.cp;
        const int32_t* didx = rows ? d.ci : d.ri;
        const float* dval = rows ? d.av32 : d.tv32;
        std::vector<mcp::Tile> tl;
        std::vector<int32_t> tb;
        double best = std::numeric_limits<double>::infinity();
        mcp_tile_policy bestp = rows ? polA : polT;
        for (const mcp_tile_policy& p : cands) {
            build_tiles(ptr, inst, B, p, false, tl, tb);
            auto* dt = buf<mcp::Tile>("tune_tiles", tl.size());
            HIPCHK(hipMemcpyAsync(dt, tl.data(), tl.size() * sizeof(mcp::Tile), hipMemcpyHostToDevice, stream));
            mcp::launch_spmv_plain(dt, (int)tl.size(), dptr, didx, dval, vin, vout, true, stream);
            double t = std::numeric_limits<double>::infinity();
            for (int trial = 0; trial < 3; ++trial) {
                HIPCHK(hipEventRecord(ev[2], stream));
                const int reps = 5;
                for (int r = 0; r < reps; ++r)
                    mcp::launch_spmv_plain(dt, (int)tl.size(), dptr, didx, dval, vin, vout, true, stream);
                HIPCHK(hipEventRecord(ev[3], stream));
                HIPCHK(hipEventSynchronize(ev[3]));
                t = std::min(t, ms_between(ev[2], ev[3]) / reps);
            }
            if (t < 0.98 * best) {
                best = t;
                bestp = p;
            }
        }
        (rows ? polA : polT) = bestp;
    }

    // Same for the block-tile budget of the fused steps (plain block SpMV,
    // FP32 values): candidates 512, 1024, 2048; the incumbent is the
    // size-based default.  Budget 0 (block tiles off) is left alone.  The
    // budget fixes the summation order of every fast-order reduction, and the
    // iteration count of a solve is sensitive to that order, so a candidate
    // must be at least 10 % faster to replace the default: timing noise (a few
    // per cent between these budgets) must not make repeated solves of one LP
    // take different paths.
    void autotune_block(bool rows) {
        int& bud = rows ? budA : budT;
        if (bud <= 0) return;
        const auto& ptr = rows ? rp : cp;
        const auto& inst = rows ? inst_row : inst_col;
        const int nrows = rows ? m : n;
        const int ncols = rows ? n : m;
        if (nrows == 0) return;
        float* vin = buf<float>("tune_in", ncols);
        float* vout = buf<float>("tune_out", nrows);
        HIPCHK(hipMemsetAsync(vin, 0, sizeof(float) * std::max(ncols, 1), stream));
        const int32_t* dptr = rows ? d.rp : d.cp;
        const int32_t* didx = rows ? d.ci : d.ri;
        const float* dval = rows ? d.av32 : d.tv32;
        std::vector<mcp::Tile> tl;
        std::vector<int32_t> tb;
        double best = std::numeric_limits<double>::infinity();
        int bestb = bud;
        for (const int cand : {bud, 512, 1024, 2048}) {
            build_block_tiles(ptr, inst, B, cand, tl, tb);
            auto* dt = buf<mcp::Tile>("tune_tiles", tl.size());
            HIPCHK(hipMemcpyAsync(dt, tl.data(), tl.size() * sizeof(mcp::Tile), hipMemcpyHostToDevice, stream));
            mcp::launch_spmv_blk(dt, (int)tl.size(), cand, dptr, didx, dval, vin, vout, true, stream);
            double t = std::numeric_limits<double>::infinity();
            for (int trial = 0; trial < 3; ++trial) {
                HIPCHK(hipEventRecord(ev[2], stream));
                const int reps = 5;
                for (int r = 0; r < reps; ++r)
                    mcp::launch_spmv_blk(dt, (int)tl.size(), cand, dptr, didx, dval, vin, vout, true, stream);
                HIPCHK(hipEventRecord(ev[3], stream));
                HIPCHK(hipEventSynchronize(ev[3]));
                t = std::min(t, ms_between(ev[2], ev[3]) / reps);
            }
            if (t < 0.90 * best) {
                best = t;
                bestb = cand;
            }
        }
        bud = bestb;
    }
};

namespace {

void make_opts(const mcp_solve_opts* o, mcp::IterOpts& io) {
    io.primal_tol = o->primal_tolerance;
    io.dual_tol = o->dual_tolerance;
    io.gap_tol = o->gap_tolerance;
    io.restart_factor = o->restart_reduction_factor;
    io.smoothing = o->primal_weight_smoothing;
    io.restart_every = (int64_t)o->restart_every;
    io.max_iterations = (int64_t)o->max_iterations;
    io.restart_strategy = o->restart_strategy;
    io.adaptive_weight = o->adaptive_primal_weight;
    io.ruiz = 1;
    io.mixed = o->precision != MC_PRECISION_FP64;
    io.packed_next = 0;  // set by mcp_solve when the packed streams exist
    io.packed_floor = 0;
    io.kkt = (o->restart_scheme == MC_SCHEME_KKT || o->restart_scheme == MC_SCHEME_HALPERN) ? 1 : 0;
    io.halpern = o->restart_scheme == MC_SCHEME_HALPERN ? 1 : 0;
    io.stall_checks = o->fp32_stall_checks > 0 ? o->fp32_stall_checks : 50;
    io.detect = o->detect_infeasibility != 0 ? 1 : 0;
    io.adaptive_step = o->adaptive_step_size != 0 ? 1 : 0;
    io.omega_max_drop = o->primal_weight_max_drop;
    io.infeas_tol = o->infeasibility_tolerance > 0.0 ? o->infeasibility_tolerance : 1e-8;
    const double tol = std::max({o->primal_tolerance, o->dual_tolerance, o->gap_tolerance});
    // 0 => no score threshold: FP32 runs until it converges or stalls
    io.polish_threshold = o->polish_threshold > 0.0 ? o->polish_threshold : 0.0;
    (void)tol;
}

// Kernels of iterations [iter0, iter0 + len): the sequence depends only on
// iter0 modulo lcm(10, restart_every), so a captured chunk is reusable.
void enqueue_chunk(mcp_engine* e, const mcp::IterOpts& io, bool adaptive, bool exact, int64_t iter0,
                   int len) {
    const auto& d = e->d;
    hipStream_t s = e->stream;
    for (int t = 0; t < len; ++t) {
        const int64_t it = iter0 + t;
        const bool adapt = adaptive && (it % 10 == 0);
        mcp::launch_primal(d, t, adapt, exact, s);
        if (adapt) {
            mcp::launch_adx(d, t, exact, s);
            mcp::launch_adapt_finalize(d, exact, s);
        }
        mcp::launch_dual(d, t, exact, s);
        if ((it + 1) % io.restart_every == 0) {
            mcp::launch_check_resid(d, exact, io.kkt != 0, io.packed_floor != 0, s);
            mcp::launch_check_finalize(d, io, t, exact, false, s);
            if (io.detect) mcp::launch_ray_check(d, io, t, s);
            mcp::launch_post_check(d, io.kkt != 0, s);
        }
    }
    mcp::launch_advance(d, len, s);
}

int64_t gcd64(int64_t a, int64_t b) {
    while (b) {
        const int64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

}  // namespace

extern "C" {

const char* mcp_last_error(void) { return g_err.c_str(); }

int mcp_device_count(int32_t* count) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return MC_OK;
}

int mcp_engine_create(int32_t device, mcp_engine** out) {
    *out = nullptr;
    return guarded([&] {
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess || c <= 0)
            throw EngineError(MC_ERR_DEVICE, "no HIP device available");
        require(device >= 0 && device < c, "device ordinal out of range");
        HIPCHK(hipSetDevice(device));
        auto e = std::make_unique<mcp_engine>();
        e->device = device;
        HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
        for (auto& ev : e->ev) HIPCHK(hipEventCreate(&ev));
        HIPCHK(hipHostMalloc((void**)&e->pinned, 8 * sizeof(int32_t), hipHostMallocDefault));
        *out = e.release();
    });
}

void mcp_engine_destroy(mcp_engine* e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    delete e;
}

int mcp_load(mcp_engine* e, const mcp_problem* p) {
    return guarded([&] {
        require(e && p, "null argument");
        HIPCHK(hipSetDevice(e->device));
        require(p->instances >= 1 && p->rows >= 0 && p->cols >= 0, "bad problem sizes");
        require(p->nnz >= 0 && p->nnz < (int64_t)std::numeric_limits<int32_t>::max(), "nnz must be < 2^31");
        const int B = p->instances, m = p->rows, n = p->cols;
        const int64_t nnz = p->nnz;
        require(p->inst_row[0] == 0 && p->inst_row[B] == m && p->inst_col[0] == 0 && p->inst_col[B] == n,
                "instance tables must span the packed problem");
        for (int b = 0; b < B; ++b)
            require(p->inst_row[b] <= p->inst_row[b + 1] && p->inst_col[b] <= p->inst_col[b + 1],
                    "instance tables must be monotone");
        require(p->row_ptr[0] == 0 && p->row_ptr[m] == nnz && p->col_ptr[0] == 0 && p->col_ptr[n] == nnz,
                "row/col pointers must span nnz");
        for (int i = 0; i < m; ++i) require(p->row_ptr[i] <= p->row_ptr[i + 1], "row_ptr not monotone");
        for (int j = 0; j < n; ++j) require(p->col_ptr[j] <= p->col_ptr[j + 1], "col_ptr not monotone");
        for (int64_t k = 0; k < nnz; ++k)
            require((uint32_t)p->col_idx[k] < (uint32_t)n && (uint32_t)p->row_idx[k] < (uint32_t)m,
                    "index out of range");
        e->B = B;
        e->m = m;
        e->n = n;
        e->nnz = nnz;
        e->inst_row.assign(p->inst_row, p->inst_row + B + 1);
        e->inst_col.assign(p->inst_col, p->inst_col + B + 1);
        e->rp.assign(p->row_ptr, p->row_ptr + m + 1);
        e->cp.assign(p->col_ptr, p->col_ptr + n + 1);
        std::vector<int32_t> row_inst(m), col_inst(n);
        for (int b = 0; b < B; ++b) {
            for (int i = p->inst_row[b]; i < p->inst_row[b + 1]; ++i) row_inst[i] = b;
            for (int j = p->inst_col[b]; j < p->inst_col[b + 1]; ++j) col_inst[j] = b;
        }
        auto& d = e->d;
        d = mcp::Dev{};
        d.B = B;
        d.m = m;
        d.n = n;
        hipStream_t s = e->stream;
        HIPCHK(hipEventRecord(e->ev[0], s));
        auto up = [&](const std::string& name, const auto* src, size_t count) {
            using T = std::remove_cv_t<std::remove_pointer_t<decltype(src)>>;
            T* dst = e->buf<T>(name, count);
            if (count) HIPCHK(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, s));
            return dst;
        };
        d.rp = up("rp", p->row_ptr, (size_t)m + 1);
        d.ci = up("ci", p->col_idx, (size_t)nnz);
        d.av64 = up("av64", p->row_val, (size_t)nnz);
        d.cp = up("cp", p->col_ptr, (size_t)n + 1);
        d.ri = up("ri", p->row_idx, (size_t)nnz);
        d.tv64 = up("tv64", p->col_val, (size_t)nnz);
        d.inst_row = up("inst_row", p->inst_row, (size_t)B + 1);
        d.inst_col = up("inst_col", p->inst_col, (size_t)B + 1);
        d.row_inst = up("row_inst", row_inst.data(), (size_t)m);
        d.col_inst = up("col_inst", col_inst.data(), (size_t)n);
        d.orow_lo = up("orow_lo", p->row_lower, (size_t)m);
        d.orow_hi = up("orow_hi", p->row_upper, (size_t)m);
        d.ovar_lo = up("ovar_lo", p->var_lower, (size_t)n);
        d.ovar_hi = up("ovar_hi", p->var_upper, (size_t)n);
        d.obj = up("obj", p->objective, (size_t)n);
        d.sense = up("sense", p->sense, (size_t)B);
        d.offset = up("offset", p->offset, (size_t)B);
        HIPCHK(hipEventRecord(e->ev[1], s));
        d.av32 = e->buf<float>("av32", nnz);
        d.tv32 = e->buf<float>("tv32", nnz);
        // packed streams (index | value correction << 24) when every index fits 24 bits
        const bool packable = m < mcp::kPackMaxIndex && n < mcp::kPackMaxIndex;
        d.cix = packable ? e->buf<uint32_t>("cix", nnz) : nullptr;
        d.rix = packable ? e->buf<uint32_t>("rix", nnz) : nullptr;
        d.row_scale = e->buf<double>("row_scale", m);
        d.col_scale = e->buf<double>("col_scale", n);
        d.rtmp = e->buf<double>("rtmp", m);
        d.ctmp = e->buf<double>("ctmp", n);
        d.blo = e->buf<double>("blo", m);
        d.bhi = e->buf<double>("bhi", m);
        d.rnorm = e->buf<double>("rnorm", m);
        d.c = e->buf<double>("c", n);
        d.slo = e->buf<double>("slo", n);
        d.shi = e->buf<double>("shi", n);
        d.cnorm = e->buf<double>("cnorm", n);
        d.x = e->buf<double>("x", n);
        d.xbar = e->buf<double>("xbar", n);
        d.xavg = e->buf<double>("xavg", n);
        d.dx = e->buf<double>("dx", n);
        d.y = e->buf<double>("y", m);
        d.yavg = e->buf<double>("yavg", m);
        d.xbar32 = e->buf<float>("xbar32", n);
        d.dx32 = e->buf<float>("dx32", n);
        d.y32 = e->buf<float>("y32", m);
        d.xlast = e->buf<double>("xlast", n);
        d.ylast = e->buf<double>("ylast", m);
        d.ray_x = e->buf<double>("ray_x", n);
        d.ray_y = e->buf<double>("ray_y", m);
        d.ax = e->buf<double>("ax", m);
        d.aty = e->buf<double>("aty", n);
        d.ax2 = e->buf<double>("ax2", m);
        d.aty2 = e->buf<double>("aty2", n);
        d.st = e->buf<mcp::InstState>("st", B);
        d.pw = e->buf<double>("pw", 2 * (size_t)B);
        d.iter_base = e->buf<int64_t>("iter_base", 1);
        d.counters = e->buf<int32_t>("counters", 4);
        HIPCHK(hipStreamSynchronize(s));
        e->h2d_ms = ms_between(e->ev[0], e->ev[1]);
        e->loaded = true;
        e->setup_done = false;
        e->polA = e->polT = e->nnz <= kLatencyNnz ? kLatencyPolicy : kThroughputPolicy;
        e->budA = e->budT = block_budget_for(e->nnz);
        e->retile(false);  // setup kernels run on these tiles
    });
}

int mcp_setup(mcp_engine* e, const mcp_setup_opts* o) {
    return guarded([&] {
        require(e && o, "null argument");
        require(e->loaded, "mcp_setup before mcp_load");
        require(o->row_weighting >= 0 && o->row_weighting <= mcp::kMaxRowWeighting, "row_weighting must be in 0..8");
        HIPCHK(hipSetDevice(e->device));
        if (e->setup_done) {
            require(e->ruiz_on_req == o->ruiz && e->ruiz_passes_req == o->ruiz_iterations &&
                        e->row_weighting_req == o->row_weighting,
                    "problem already equilibrated with other Ruiz or row-weighting options; reload it");
            if ((o->exact != 0) != e->exact_tiles) e->retile(o->exact != 0);
            return;
        }
        auto& d = e->d;
        hipStream_t s = e->stream;
        std::vector<mcp::InstState> st(e->B);
        for (auto& x : st) {
            std::memset(&x, 0, sizeof(x));
            x.c_scale = 1.0;
            x.b_norm_inf = 1.0;
            x.c_norm_inf = 1.0;
            x.ruiz_done = o->ruiz ? 0 : 1;
            x.status = mcp::kRunning;
        }
        HIPCHK(hipMemcpyAsync(d.st, st.data(), st.size() * sizeof(mcp::InstState), hipMemcpyHostToDevice, s));
        HIPCHK(hipEventRecord(e->ev[0], s));
        mcp::launch_fill(d.row_scale, 1.0, e->m, s);
        mcp::launch_fill(d.col_scale, 1.0, e->n, s);
        if (o->ruiz && e->nnz > 0)
            for (int k = 0; k < o->ruiz_iterations; ++k) mcp::launch_ruiz_pass(d, 1e-3, s);
        if (o->row_weighting > 0 && e->m > 0) mcp::launch_row_weights(d, o->row_weighting, s);
        mcp::launch_setup_data(d, s);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(e->ev[1], s));
        HIPCHK(hipStreamSynchronize(s));
        e->setup_ms = ms_between(e->ev[0], e->ev[1]);
        e->ruiz_on_req = o->ruiz;
        e->ruiz_passes_req = o->ruiz_iterations;
        e->row_weighting_req = o->row_weighting;
        if (o->autotune && !o->exact) {
            uint64_t h = 1469598103934665603ULL;
            for (size_t k = 0; k < e->rp.size(); k += 1 + e->rp.size() / 4096) h = (h ^ (uint64_t)e->rp[k]) * 1099511628211ULL;
            for (size_t k = 0; k < e->cp.size(); k += 1 + e->cp.size() / 4096) h = (h ^ (uint64_t)e->cp[k]) * 1099511628211ULL;
            const auto key = std::make_tuple(e->B, e->m, e->n, e->nnz, h);
            auto it = e->tuned.find(key);
            if (it != e->tuned.end()) {
                e->polA = it->second.polA;
                e->polT = it->second.polT;
                e->budA = it->second.budA;
                e->budT = it->second.budT;
            } else {
                e->autotune_one(true);
                e->autotune_one(false);
                e->autotune_block(true);
                e->autotune_block(false);
                e->tuned[key] = {e->polA, e->polT, e->budA, e->budT};
            }
        }
        e->retile(o->exact != 0);
        e->setup_done = true;
    });
}

int mcp_solve(mcp_engine* e, const mcp_solve_opts* o, mcp_solve_stats* stats) {
    return guarded([&] {
        require(e && o, "null argument");
        require(e->setup_done, "mcp_solve before mcp_setup");
        require(o->restart_every >= 1, "restart_every must be >= 1");
        const bool exact = o->reduction_order == MC_ORDER_REFERENCE;
        require(o->precision >= MC_PRECISION_FP64 && o->precision <= MC_PRECISION_PACKED, "unknown precision");
        require(o->precision != MC_PRECISION_PACKED || e->d.cix, "packed precision needs fewer than 2^24 rows and columns");
        const bool mixed = o->precision != MC_PRECISION_FP64;
        require(!(exact && mixed), "reference reduction order requires FP64 precision");
        require(!exact || e->exact_tiles, "reference order needs mcp_setup(exact=1)");
        require(!(exact && o->detect_infeasibility), "infeasibility detection runs in the fast reduction order");
        HIPCHK(hipSetDevice(e->device));
        auto& d = e->d;
        hipStream_t s = e->stream;
        require(o->restart_scheme >= MC_SCHEME_REFERENCE && o->restart_scheme <= MC_SCHEME_HALPERN,
                "unknown restart scheme");
        require(o->primal_weight_max_drop == 0 ||
                    (o->primal_weight_max_drop >= 2 && o->primal_weight_max_drop <= (1 << 20) &&
                     (o->primal_weight_max_drop & (o->primal_weight_max_drop - 1)) == 0),
                "primal_weight_max_drop must be 0 or a power of two in 2..2^20");
        mcp::IterOpts io{};
        make_opts(o, io);
        io.packed_next = (o->precision == MC_PRECISION_MIXED && d.cix) ? 1 : 0;
        io.packed_floor = (!exact && (io.packed_next || o->precision == MC_PRECISION_PACKED)) ? 1 : 0;
        d.halpern = io.halpern;
        // start point
        const double* wx = nullptr;
        const double* wy = nullptr;
        if (o->warm_primal) {
            double* t = e->buf<double>("warm_x", e->n);
            HIPCHK(hipMemcpyAsync(t, o->warm_primal, sizeof(double) * e->n, hipMemcpyHostToDevice, s));
            wx = t;
        }
        if (o->warm_dual) {
            double* t = e->buf<double>("warm_y", e->m);
            HIPCHK(hipMemcpyAsync(t, o->warm_dual, sizeof(double) * e->m, hipMemcpyHostToDevice, s));
            wy = t;
        }
        mcp::launch_init_iterates(d, wx, wy, s);
        // The FP32-SpMV phase serves the start far from the optimum; its floor
        // (1e-2 to 1e-4 relative KKT on planted LPs) lies above a warm start's
        // accuracy, so a warm-started mixed solve begins on the packed stream
        // (whose floor is measured at every check), or in FP64 without one.
        // On problems of up to kLatencyNnz nonzeros an iteration is a chain of
        // launch and memory latencies whatever the arithmetic (the 2000 x 5000
        // LP: 5.2-5.8 us per step in FP32, packed and FP64 alike), so the FP32
        // phase saves nothing there and only moves the trajectory (that LP:
        // 137,640 iterations after an FP32 start, 50,920 from a packed one).
        const bool warm = wx || wy;
        const bool fp32_pays = e->nnz > kLatencyNnz || !d.cix;
        int start_mode = mcp::kModeFp64;
        if (o->precision == MC_PRECISION_PACKED)
            start_mode = mcp::kModePacked;
        else if (o->precision == MC_PRECISION_MIXED)
            start_mode = (!warm && fp32_pays) ? mcp::kModeFp32 : (d.cix ? mcp::kModePacked : mcp::kModeFp64);
        const bool reduced_start = start_mode != mcp::kModeFp64;
        mcp::launch_setup_state(d, o->initial_primal_weight, o->step_size_reduction, start_mode, s);
        if (io.halpern && io.adaptive_step && o->power_iterations > 0)
            mcp::launch_step_cap(d, (int64_t)o->power_iterations, exact, s);
        int32_t* pin = e->pinned;
        pin[4] = e->B;
        pin[5] = reduced_start ? e->B : 0;
        HIPCHK(hipMemcpyAsync(d.counters, pin + 4, 2 * sizeof(int32_t), hipMemcpyHostToDevice, s));
        int64_t zero = 0;
        HIPCHK(hipMemcpyAsync(d.iter_base, &zero, sizeof(int64_t), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));

        const int64_t max_it = (int64_t)o->max_iterations;
        const int64_t R = io.restart_every;
        const bool adaptive = io.adaptive_step != 0;
        const int64_t period = adaptive ? (10 / gcd64(10, R)) * R : R;
        int chunk = o->graph_chunk == 0 ? 64 : std::abs(o->graph_chunk);
        const bool use_graphs = o->graph_chunk >= 0;
        int graphs_built = 0, chunks = 0;
        // Captured chunks bake the kernel arguments (buffer pointers, tile
        // counts, tolerances and scheme options); they stay valid while those
        // are unchanged.
        if (std::memcmp(&e->graph_dev, &d, sizeof(d)) != 0 || std::memcmp(&e->graph_io, &io, sizeof(io)) != 0) {
            e->drop_graphs();
            e->graph_dev = d;
            e->graph_io = io;
        }
        HIPCHK(hipEventRecord(e->ev[0], s));
        // MCPDLP_PROGRESS=<seconds>: report the solve's own progress on stderr.
        const char* prog_env = std::getenv("MCPDLP_PROGRESS");
        const double prog_every = prog_env ? std::atof(prog_env) : 0.0;
        const auto prog_start = std::chrono::steady_clock::now();
        auto prog_last = prog_start;
        int64_t iter = 0;
        bool finished = false;
        // Batch compaction: once a sixteenth of the instances that were
        // running at the last (re)tiling have finished, rebuild the tiles over
        // the running instances only, so finished instances stop costing
        // waves.  Measured on the 256-LP batch (same trajectories): retiling
        // at half took 21.75 s, at 3/4 20.56 s, at 15/16 20.18 s, at 0.99
        // 20.39 s (the retiling and graph captures start to cost more).
        int32_t compact_at = e->B > 1 ? e->B * 15 / 16 : -1;
        const mcp_tile_policy polA0 = e->polA, polT0 = e->polT;
        const int budA0 = e->budA, budT0 = e->budT;
        while (iter < max_it) {
            const int len = (int)std::min<int64_t>(chunk, max_it - iter);
            if (use_graphs && len == chunk) {
                const auto key = std::make_tuple(iter % period, len, 0, 0, (int)exact);
                auto it = e->graphs.find(key);
                if (it == e->graphs.end()) {
                    hipGraph_t g;
                    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    enqueue_chunk(e, io, adaptive, exact, iter, len);
                    HIPCHK(hipStreamEndCapture(s, &g));
                    hipGraphExec_t ex;
                    HIPCHK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
                    HIPCHK(hipGraphDestroy(g));
                    it = e->graphs.emplace(key, ex).first;
                    ++graphs_built;
                }
                HIPCHK(hipGraphLaunch(it->second, s));
            } else {
                enqueue_chunk(e, io, adaptive, exact, iter, len);
            }
            HIPCHK(hipGetLastError());
            const int slot = chunks & 1;
            HIPCHK(hipMemcpyAsync(pin + 2 * slot, d.counters, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
            HIPCHK(hipEventRecord(e->ev[2 + slot], s));
            iter += len;
            ++chunks;
            if (chunks >= 2) {
                const int prev = (chunks - 2) & 1;
                HIPCHK(hipEventSynchronize(e->ev[2 + prev]));
                if (pin[2 * prev] == 0) {
                    finished = true;
                    break;
                }
                if (pin[2 * prev] <= compact_at) {
                    HIPCHK(hipStreamSynchronize(s));
                    std::vector<mcp::InstState> stv(e->B);
                    HIPCHK(hipMemcpy(stv.data(), d.st, sizeof(mcp::InstState) * e->B, hipMemcpyDeviceToHost));
                    std::vector<char> keep(e->B, 0);
                    int32_t nk = 0;
                    for (int b = 0; b < e->B; ++b) {
                        keep[b] = stv[b].status == mcp::kRunning ? 1 : 0;
                        nk += keep[b];
                    }
                    if (nk == 0) {
                        finished = true;
                        break;
                    }
                    int64_t live_nnz = 0;
                    for (int b = 0; b < e->B; ++b)
                        if (keep[b]) live_nnz += e->rp[e->inst_row[b + 1]] - e->rp[e->inst_row[b]];
                    if (live_nnz <= kLatencyNnz) {
                        e->polA = e->polT = kLatencyPolicy;
                        if (e->budA > 0 && e->budT > 0) e->budA = e->budT = block_budget_for(live_nnz);
                    }
                    e->retile(e->exact_tiles, &keep);
                    e->drop_graphs();
                    e->graph_dev = d;
                    compact_at = nk * 15 / 16;
                }
                if (prog_every > 0.0) {
                    const auto now = std::chrono::steady_clock::now();
                    if (std::chrono::duration<double>(now - prog_last).count() >= prog_every) {
                        prog_last = now;
                        mcp::InstState s0;
                        HIPCHK(hipMemcpy(&s0, d.st, sizeof(s0), hipMemcpyDeviceToHost));
                        std::fprintf(stderr,
                                     "[mcpdlp] %.1fs iterations=%lld running=%d reduced=%d | inst0 score=%.3e "
                                     "(p=%.2e d=%.2e g=%.2e) omega=%.3g eta=%.3g mode=%s\n",
                                     std::chrono::duration<double>(now - prog_start).count(), (long long)iter,
                                     pin[2 * prev], pin[2 * prev + 1], s0.score, s0.primal, s0.dual, s0.gap,
                                     s0.omega, s0.eta, s0.mode == mcp::kModeFp32 ? "fp32" : s0.mode == mcp::kModePacked ? "packed" : "fp64");
                        std::fflush(stderr);
                    }
                }
            }
        }
        HIPCHK(hipStreamSynchronize(s));
        const int last = (chunks - 1) & 1;
        const int32_t still_running = chunks > 0 ? pin[2 * last] : e->B;
        if (!finished && still_running > 0) {
            // iteration limit: residuals of the final average (pdlp.cpp:263-265)
            mcp::launch_check_resid(d, exact, io.kkt != 0, io.packed_floor != 0, s);
            mcp::launch_check_finalize(d, io, -1, exact, true, s);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipEventRecord(e->ev[1], s));
        HIPCHK(hipStreamSynchronize(s));
        if (e->compacted) {  // outside mcp_solve the tiles always cover every instance
            e->polA = polA0;
            e->polT = polT0;
            e->budA = budA0;
            e->budT = budT0;
            e->retile(e->exact_tiles);
            e->drop_graphs();
            e->graph_dev = d;
        }
        if (stats) {
            stats->kernel_ms = ms_between(e->ev[0], e->ev[1]);
            stats->setup_ms = e->setup_ms;
            stats->h2d_ms = e->h2d_ms;
            stats->iterations = (uint64_t)iter;
            stats->graphs_built = graphs_built;
            stats->chunks = chunks;
        }
    });
}

int mcp_fetch(mcp_engine* e, double* x, double* y, mcp_inst_result* res, double* d2h_ms) {
    return guarded([&] {
        require(e && e->setup_done, "mcp_fetch before a solve");
        HIPCHK(hipSetDevice(e->device));
        auto& d = e->d;
        hipStream_t s = e->stream;
        double* xo = e->buf<double>("x_out", e->n);
        double* yo = e->buf<double>("y_out", e->m);
        HIPCHK(hipEventRecord(e->ev[0], s));
        mcp::launch_unscale(d, xo, yo, s);
        if (x && e->n) HIPCHK(hipMemcpyAsync(x, xo, sizeof(double) * e->n, hipMemcpyDeviceToHost, s));
        if (y && e->m) HIPCHK(hipMemcpyAsync(y, yo, sizeof(double) * e->m, hipMemcpyDeviceToHost, s));
        std::vector<mcp::InstState> st(e->B);
        HIPCHK(hipMemcpyAsync(st.data(), d.st, sizeof(mcp::InstState) * e->B, hipMemcpyDeviceToHost, s));
        HIPCHK(hipEventRecord(e->ev[1], s));
        HIPCHK(hipStreamSynchronize(s));
        if (d2h_ms) *d2h_ms = ms_between(e->ev[0], e->ev[1]);
        if (res)
            for (int b = 0; b < e->B; ++b) {
                const auto& q = st[b];
                auto& r = res[b];
                r.status = q.status == mcp::kOptimal       ? MC_PDLP_OPTIMAL
                           : q.status == mcp::kNumFail     ? MC_PDLP_NUMERICAL_FAILURE
                           : q.status == mcp::kInfeasible ? MC_PDLP_INFEASIBLE_OR_UNBOUNDED
                                                           : MC_PDLP_ITERATION_LIMIT;
                r.mode = q.mode;
                r.iterations = (uint64_t)q.done_iter;
                r.fp32_iterations = (uint64_t)q.fp32_iters;
                r.packed_iterations = q.packed_end > q.fp32_iters ? (uint64_t)(q.packed_end - q.fp32_iters) : 0;
                r.objective = q.objective;
                r.primal_infeasibility = q.primal;
                r.dual_infeasibility = q.dual;
                r.duality_gap = q.gap;
                r.omega = q.omega;
                r.eta = q.eta;
            }
    });
}

int mcp_spmv(mcp_engine* e, int32_t transpose, int32_t precision, int32_t order, const double* v,
             double* out, int32_t reps, double* avg_us) {
    return guarded([&] {
        require(e && e->setup_done, "mcp_spmv before mcp_setup");
        HIPCHK(hipSetDevice(e->device));
        auto& d = e->d;
        hipStream_t s = e->stream;
        const bool t = transpose != 0;
        const bool f32 = precision == MC_PRECISION_MIXED;
        const bool pk = precision == MC_PRECISION_PACKED;
        const bool exact = order == MC_ORDER_REFERENCE;
        require(!(exact && (f32 || pk)), "reference order is FP64 only");
        require(!pk || d.cix, "the packed streams need indices below 2^24");
        const int nin = t ? e->m : e->n, nout = t ? e->n : e->m;
        std::vector<mcp::Tile> tl;
        std::vector<int32_t> tb;
        // fast order: the block tiles of the fused steps (when enabled)
        const int bud = t ? e->budT : e->budA;
        const bool blk = !exact && e->budA > 0 && e->budT > 0;
        if (blk)
            build_block_tiles(t ? e->cp : e->rp, t ? e->inst_col : e->inst_row, e->B, bud, tl, tb);
        else
            build_tiles(t ? e->cp : e->rp, t ? e->inst_col : e->inst_row, e->B, t ? e->polT : e->polA,
                        exact, tl, tb);
        auto* dt = e->buf<mcp::Tile>("spmv_tiles", tl.size());
        if (!tl.empty())
            HIPCHK(hipMemcpyAsync(dt, tl.data(), tl.size() * sizeof(mcp::Tile), hipMemcpyHostToDevice, s));
        void* din = nullptr;
        void* dout = nullptr;
        std::vector<float> vf, of;
        if (f32) {
            vf.assign(v, v + nin);
            din = e->buf<float>("spmv_in32", nin);
            dout = e->buf<float>("spmv_out32", nout);
            if (nin) HIPCHK(hipMemcpyAsync(din, vf.data(), sizeof(float) * nin, hipMemcpyHostToDevice, s));
        } else {
            din = e->buf<double>("spmv_in", nin);
            dout = e->buf<double>("spmv_out", nout);
            if (nin) HIPCHK(hipMemcpyAsync(din, v, sizeof(double) * nin, hipMemcpyHostToDevice, s));
        }
        const int32_t* ptr = t ? d.cp : d.rp;
        const int32_t* idx = t ? d.ri : d.ci;
        const void* val = f32 ? (const void*)(t ? d.tv32 : d.av32) : (const void*)(t ? d.tv64 : d.av64);
        const uint32_t* idxp = t ? d.rix : d.cix;
        const float* val32 = t ? d.tv32 : d.av32;
        auto launch = [&] {
            if (pk && blk)
                mcp::launch_spmv_blk_packed(dt, (int)tl.size(), bud, ptr, idxp, val32, (const double*)din,
                                            (double*)dout, s);
            else if (pk)
                mcp::launch_spmv_plain_packed(dt, (int)tl.size(), ptr, idxp, val32, (const double*)din, (double*)dout,
                                              s);
            else if (blk)
                mcp::launch_spmv_blk(dt, (int)tl.size(), bud, ptr, idx, val, din, dout, f32, s);
            else
                mcp::launch_spmv_plain(dt, (int)tl.size(), ptr, idx, val, din, dout, f32, s);
        };
        launch();
        const int nrep = std::max(1, reps);
        HIPCHK(hipEventRecord(e->ev[0], s));
        for (int r = 0; r < nrep; ++r) launch();
        HIPCHK(hipEventRecord(e->ev[1], s));
        HIPCHK(hipGetLastError());
        if (f32) {
            of.resize(nout);
            if (nout) HIPCHK(hipMemcpyAsync(of.data(), dout, sizeof(float) * nout, hipMemcpyDeviceToHost, s));
        } else if (nout) {
            HIPCHK(hipMemcpyAsync(out, dout, sizeof(double) * nout, hipMemcpyDeviceToHost, s));
        }
        HIPCHK(hipStreamSynchronize(s));
        if (f32)
            for (int k = 0; k < nout; ++k) out[k] = of[k];
        if (avg_us) *avg_us = 1000.0 * ms_between(e->ev[0], e->ev[1]) / nrep;
    });
}

int mcp_get_scaling(mcp_engine* e, double* col_val, double* row_scale, double* col_scale,
                    int32_t* passes, int32_t* converged) {
    return guarded([&] {
        require(e && e->setup_done, "mcp_get_scaling before mcp_setup");
        HIPCHK(hipSetDevice(e->device));
        auto& d = e->d;
        hipStream_t s = e->stream;
        if (col_val && e->nnz)
            HIPCHK(hipMemcpyAsync(col_val, d.tv64, sizeof(double) * e->nnz, hipMemcpyDeviceToHost, s));
        if (row_scale && e->m)
            HIPCHK(hipMemcpyAsync(row_scale, d.row_scale, sizeof(double) * e->m, hipMemcpyDeviceToHost, s));
        if (col_scale && e->n)
            HIPCHK(hipMemcpyAsync(col_scale, d.col_scale, sizeof(double) * e->n, hipMemcpyDeviceToHost, s));
        mcp::InstState st0;
        HIPCHK(hipMemcpyAsync(&st0, d.st, sizeof(st0), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (passes) *passes = st0.ruiz_passes;
        if (converged) *converged = st0.ruiz_converged;
    });
}

int mcp_time_kernel(mcp_engine* e, int32_t kind, int32_t reps, double* avg_us, double* bytes) {
    return guarded([&] {
        require(e && e->setup_done, "mcp_time_kernel before mcp_setup");
        require(kind >= 0 && kind <= 11, "unknown kernel kind");
        const bool probe = kind >= 7 && kind <= 9;  // gather probes over the block tiles of A (FP64)
        // kinds 10, 11: the fused primal / dual step on the packed streams
        // (FP32 value + 8-bit correction) with FP64 gathers, products and sums
        const bool packed = kind == 10 || kind == 11;
        require(!packed || e->d.cix, "the packed streams need indices below 2^24");
        require(!probe || (e->d.blocks && e->d.nbA > 0 && e->n > 0), "gather probes need block tiles of A");
        HIPCHK(hipSetDevice(e->device));
        auto& d = e->d;
        hipStream_t s = e->stream;
        const double m = e->m, n = e->n, nnz = (double)e->nnz;
        const bool f32 = kind == 0 || kind == 1 || (kind >= 4 && kind <= 6);
        const double w = f32 ? 4.0 : 8.0;  // bytes per value
        // Matrix (index + value of width w) and pointers once, the gathered
        // vector (width w) once.  Per element, FP64 in both modes: the primal
        // step reads c/lo/hi/colnorm (32 B) and x, x_avg (16 B) and writes x,
        // x_bar, x_avg (24 B), plus the FP32 copy of x_bar (4 B) in FP32 mode;
        // the dual step reads blo/bhi/rownorm (24 B) and y, y_avg (16 B) and
        // writes y, y_avg (16 B), plus the FP32 copy of y in FP32 mode.
        double b = 0.0;
        if (kind == 0 || kind == 2)
            b = (4.0 + w) * nnz + 4.0 * (n + 1) + w * m + (72.0 + (f32 ? 4.0 : 0.0)) * n;
        else if (kind == 1 || kind == 3)
            b = (4.0 + w) * nnz + 4.0 * (m + 1) + w * n + (56.0 + (f32 ? 4.0 : 0.0)) * m;
        else if (kind == 10)
            b = 8.0 * nnz + 4.0 * (n + 1) + 8.0 * m + 72.0 * n;
        else if (kind == 11)
            b = 8.0 * nnz + 4.0 * (m + 1) + 8.0 * n + 56.0 * m;
        else if (kind == 4)
            b = 8.0 * nnz + 4.0 * (m + 1) + 4.0 * n + 4.0 * m;
        else if (kind == 5)
            b = 8.0 * nnz + 4.0 * (n + 1) + 4.0 * m + 4.0 * n;
        // kinds 7-9: the FP64 block SpMV A x_bar on the solve's tiles with random
        // gathers (7), gathers confined to a window of mask + 1 <= 256 entries (8),
        // no gathers (9); the stream is that of k_dual_blk in FP64 mode
        int mask = 1;
        while (2 * mask <= std::min(256, e->n)) mask *= 2;
        mask -= 1;
        if (kind == 7) b = 12.0 * nnz + 4.0 * (m + 1) + 8.0 * n + 8.0 * m;
        if (kind == 8) b = 12.0 * nnz + 4.0 * (m + 1) + 8.0 * (mask + 1) + 8.0 * m;
        if (kind == 9) b = 12.0 * nnz + 4.0 * (m + 1) + 8.0 * m;
        const int64_t ncopy = e->nnz / 4 * 4;  // kind 6: plain copy of the FP32 values of A
        float* copy_dst = kind == 6 ? e->buf<float>("copy_dst", ncopy) : nullptr;
        if (kind == 6) b = 8.0 * (double)ncopy;
        // every instance running in the kernel's precision, average count 1
        std::vector<mcp::InstState> st(e->B);
        HIPCHK(hipMemcpy(st.data(), d.st, sizeof(mcp::InstState) * e->B, hipMemcpyDeviceToHost));
        for (auto& q : st) {
            q.status = mcp::kRunning;
            q.mode = f32 ? mcp::kModeFp32 : packed ? mcp::kModePacked : mcp::kModeFp64;
            q.restart_iter = 0;
            if (!(q.eta > 0.0)) q.eta = 0.9;
            if (!(q.omega > 0.0)) q.omega = 1.0;
        }
        HIPCHK(hipMemcpy(d.st, st.data(), sizeof(mcp::InstState) * e->B, hipMemcpyHostToDevice));
        int64_t zero = 0;
        HIPCHK(hipMemcpy(d.iter_base, &zero, sizeof(zero), hipMemcpyHostToDevice));
        auto launch = [&] {
            switch (kind) {
                case 0:
                case 2:
                case 10: mcp::launch_primal(d, 1, false, false, s); break;
                case 1:
                case 3:
                case 11: mcp::launch_dual(d, 1, false, s); break;
                case 4: mcp::launch_spmv_plain(d.tA, d.ntA, d.rp, d.ci, d.av32, d.xbar32, d.y32, true, s); break;
                case 5: mcp::launch_spmv_plain(d.tT, d.ntT, d.cp, d.ri, d.tv32, d.y32, d.dx32, true, s); break;
                case 6: mcp::launch_copy(d.av32, copy_dst, ncopy, s); break;
                default:
                    mcp::launch_gather_probe(d.bA, d.nbA, d.budA, d.rp, d.ci, d.av64, d.xbar, d.ax, kind - 7, mask, s);
                    break;
            }
        };
        launch();
        const int nrep = std::max(1, reps);
        if (kind <= 3 || probe || packed) {
            // The fused steps are timed as the solve runs them: each launch
            // follows the other step, which streams the other matrix through
            // the caches in between.
            auto partner = [&] {
                if (kind == 0 || kind == 2 || kind == 10)
                    mcp::launch_dual(d, 1, false, s);
                else
                    mcp::launch_primal(d, 1, false, false, s);
            };
            // All launches are queued before the one wait; each timed launch
            // records its event pair as part of its own dispatch (no separate
            // event packets between the kernels, as between graph nodes).
            std::vector<hipEvent_t> evs(2 * (size_t)nrep, nullptr);
            auto destroy = [&] {
                for (auto& ev : evs)
                    if (ev) (void)hipEventDestroy(ev);
            };
            try {
                for (auto& ev : evs) HIPCHK(hipEventCreate(&ev));
                for (int r = 0; r < nrep; ++r) {
                    partner();
                    if (kind == 0 || kind == 2 || kind == 10)
                        mcp::launch_primal(d, 1, false, false, s, evs[2 * r], evs[2 * r + 1]);
                    else if (probe)
                        mcp::launch_gather_probe(d.bA, d.nbA, d.budA, d.rp, d.ci, d.av64, d.xbar, d.ax, kind - 7,
                                                 mask, s, evs[2 * r], evs[2 * r + 1]);
                    else
                        mcp::launch_dual(d, 1, false, s, evs[2 * r], evs[2 * r + 1]);
                }
                HIPCHK(hipGetLastError());
                HIPCHK(hipEventSynchronize(evs.back()));
                double tot = 0.0;
                for (int r = 0; r < nrep; ++r) tot += ms_between(evs[2 * r], evs[2 * r + 1]);
                *avg_us = 1000.0 * tot / nrep;
            } catch (...) {
                destroy();
                throw;
            }
            destroy();
        } else {
            HIPCHK(hipEventRecord(e->ev[0], s));
            for (int r = 0; r < nrep; ++r) launch();
            HIPCHK(hipEventRecord(e->ev[1], s));
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventSynchronize(e->ev[1]));
            *avg_us = 1000.0 * ms_between(e->ev[0], e->ev[1]) / nrep;
        }
        *bytes = b;
    });
}

int mcp_time_sync(mcp_engine* e, int32_t workgroups, int32_t rounds, double* barrier_us, double* launch_us) {
    return guarded([&] {
        require(e && barrier_us && launch_us, "null argument");
        require(workgroups >= 1 && rounds >= 1 && rounds <= 100000, "workgroups >= 1 and rounds in [1, 100000]");
        HIPCHK(hipSetDevice(e->device));
        const int cap = mcp::barrier_probe_capacity();
        require(workgroups <= cap, "more workgroups than can be resident at once (" + std::to_string(cap) + ")");
        hipStream_t s = e->stream;
        const int words = 64;
        unsigned* bar = e->buf<unsigned>("sync_bar", 4);
        double* slots = e->buf<double>("sync_slots", (size_t)workgroups * words);
        HIPCHK(hipMemsetAsync(bar, 0, 4 * sizeof(unsigned), s));
        HIPCHK(mcp::launch_barrier_probe(bar, slots, words, 4, workgroups, s));  // warm
        HIPCHK(hipEventRecord(e->ev[0], s));
        HIPCHK(mcp::launch_barrier_probe(bar, slots, words, rounds, workgroups, s));
        HIPCHK(hipEventRecord(e->ev[1], s));
        HIPCHK(hipEventSynchronize(e->ev[1]));
        unsigned hb[4];
        HIPCHK(hipMemcpy(hb, bar, sizeof(hb), hipMemcpyDeviceToHost));
        require(hb[2] == 0, "grid barrier: a wait ran out of spins");
        require(hb[3] == 0, "grid barrier: a read after the barrier saw a stale value");
        *barrier_us = 1000.0 * ms_between(e->ev[0], e->ev[1]) / (2.0 * rounds);
        // the same rounds as dependent launches in a captured graph (at most
        // 256 per graph, replayed)
        const int per = std::min(rounds, 256);
        const int reps = (rounds + per - 1) / per;
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int r = 0; r < per; ++r) mcp::launch_barrier_step(slots, words, r, workgroups, s);
        HIPCHK(hipStreamEndCapture(s, &g));
        hipGraphExec_t ex;
        const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(ie);
        auto run = [&] {
            HIPCHK(hipGraphLaunch(ex, s));  // warm
            HIPCHK(hipEventRecord(e->ev[0], s));
            for (int r = 0; r < reps; ++r) HIPCHK(hipGraphLaunch(ex, s));
            HIPCHK(hipEventRecord(e->ev[1], s));
            HIPCHK(hipEventSynchronize(e->ev[1]));
        };
        try {
            run();
        } catch (...) {
            (void)hipGraphExecDestroy(ex);
            throw;
        }
        (void)hipGraphExecDestroy(ex);
        *launch_us = 1000.0 * ms_between(e->ev[0], e->ev[1]) / ((double)reps * per);
    });
}

int mcp_set_tile_policy(mcp_engine* e, const mcp_tile_policy* rows_policy, const mcp_tile_policy* cols_policy) {
    return guarded([&] {
        require(e && e->loaded && rows_policy && cols_policy, "mcp_set_tile_policy needs a loaded problem");
        require(rows_policy->nnz_budget >= 1 && cols_policy->nnz_budget >= 1 && rows_policy->per_lane >= 1 &&
                    cols_policy->per_lane >= 1,
                "tile policy fields must be positive");
        HIPCHK(hipSetDevice(e->device));
        e->polA = *rows_policy;
        e->polT = *cols_policy;
        e->retile(e->exact_tiles);
    });
}

int mcp_set_block_budget(mcp_engine* e, int32_t rows_budget, int32_t cols_budget) {
    return guarded([&] {
        require(e && e->loaded, "mcp_set_block_budget needs a loaded problem");
        auto ok = [](int32_t b) { return b == 0 || b == 256 || b == 512 || b == 1024 || b == 2048; };
        require(ok(rows_budget) && ok(cols_budget) && ((rows_budget == 0) == (cols_budget == 0)),
                "block budgets must both be 0 or each one of 256, 512, 1024, 2048");
        HIPCHK(hipSetDevice(e->device));
        e->budA = rows_budget;
        e->budT = cols_budget;
        e->retile(e->exact_tiles);
    });
}

int mcp_get_block_tiling(mcp_engine* e, int32_t* rows_budget, int32_t* cols_budget, int32_t* row_tiles,
                         int32_t* col_tiles) {
    return guarded([&] {
        require(e != nullptr, "null engine");
        if (rows_budget) *rows_budget = e->budA;
        if (cols_budget) *cols_budget = e->budT;
        if (row_tiles) *row_tiles = (int32_t)e->bA.size();
        if (col_tiles) *col_tiles = (int32_t)e->bT.size();
    });
}

int mcp_get_tiling(mcp_engine* e, mcp_tile_policy* rows_policy, mcp_tile_policy* cols_policy,
                   int32_t* row_tiles, int32_t* col_tiles) {
    return guarded([&] {
        require(e != nullptr, "null engine");
        if (rows_policy) *rows_policy = e->polA;
        if (cols_policy) *cols_policy = e->polT;
        if (row_tiles) *row_tiles = (int32_t)e->tA.size();
        if (col_tiles) *col_tiles = (int32_t)e->tT.size();
    });
}

}  // extern "C"
