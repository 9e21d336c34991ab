// pdlp.hpp -- PDLP (PDHG) LP solve: the drop-in for markov-cero's
// lp::first_order::solve_pdlp (reference include/markov_cero/lp/first_order/pdlp.hpp).
//
// PdlpOptions / PdlpResult / PdlpStatus / Backend / RestartStrategy keep the
// reference's fields, meanings and defaults (pdlp.hpp:15-63).  Members marked
// "extension" drive the HIP engine and have no reference counterpart.
#pragma once

#include <cstddef>
#include <string>
#include <vector>

#include "markov_cero/model/model.hpp"

namespace markov_cero::lp::first_order {

enum class Backend { cpu, gpu };
enum class RestartStrategy { none, adaptive, fixed };
enum class Precision { fp64, mixed, packed };   // extension
enum class ReductionOrder { fast, reference };      // extension
// extension: `reference` keeps the reference's restart rule (pdlp.cpp:246-260);
// `kkt` uses the adaptive restarts of Applegate et al. (2021): candidate = the
// better of current iterate and average by the reference's score, restart on
// sufficient (0.2x) / necessary-without-progress (0.8x) decay or an epoch of
// 0.36 x all iterations, primal weight from the movement since the last
// restart (omega <- (|dy|/|dx|)^theta omega^(1-theta), theta = smoothing).
enum class RestartScheme { reference, kkt, halpern };

struct PdlpOptions {
    std::size_t max_iterations{100000};
    std::size_t restart_every{40};
    double primal_tolerance{1e-4};
    double dual_tolerance{1e-4};
    double gap_tolerance{1e-4};
    double step_size_reduction{0.9};
    std::size_t power_iterations{20};  // kept for source compatibility; never read
    Backend backend{Backend::cpu};
    RestartStrategy restart_strategy{RestartStrategy::adaptive};
    double restart_reduction_factor{0.368};
    bool adaptive_step_size{true};
    bool adaptive_primal_weight{true};
    double initial_primal_weight{0.0};
    double primal_weight_smoothing{0.5};
    bool ruiz_scaling{true};
    std::size_t ruiz_iterations{10};

    // --- extensions (Backend::gpu) ---
    RestartScheme restart_scheme{RestartScheme::reference};
    Precision precision{Precision::fp64};
    ReductionOrder reduction_order{ReductionOrder::fast};
    double polish_threshold{0.0};   // mixed: FP64 from this checked score (0 => only when FP32 stalls)
    int fp32_stall_checks{0};       // mixed: checks without a new best before FP64 (0 => 50)
    // Report infeasible_or_unbounded when the iterate's movement since the
    // last restart is a primal or dual ray certificate (off: the reference
    // never reports that status).
    bool detect_infeasibility{false};
    double infeasibility_tolerance{1e-8};
    // k in 0..8: after the Ruiz passes, scale each row and its bounds by
    // (1 + beta_i)^(-k/8), beta_i the largest magnitude of its finite bounds
    // (0: the reference's scaling).  The relative primal residual divides a
    // row's violation by 1 + |a_i x|, so rows with small bounds must be met
    // more tightly; the weight raises their dual step.
    int row_weighting{0};
    // Halpern scheme: a restart lowers the primal weight by at most this
    // factor (0: no limit; else a power of two in 2..2^20).
    int primal_weight_max_drop{0};
    int graph_chunk{0};             // iterations per captured HIP graph (0 => 64, <0 => eager launches)
    int device{0};
    bool require_gpu{false};        // true: fail instead of the cpu_fallback path
    bool autotune{false};           // time tile shapes on the device before solving
    std::vector<double> warm_primal;  // optional start x (original space), size n
    std::vector<double> warm_dual;    // optional start dual (reported convention), size m

    void set_tolerance(double tol) noexcept {
        primal_tolerance = tol;
        dual_tolerance = tol;
        gap_tolerance = tol;
    }
};

enum class PdlpStatus { optimal, iteration_limit, infeasible_or_unbounded, numerical_failure };

struct PdlpResult {
    PdlpStatus status{PdlpStatus::iteration_limit};
    std::vector<double> primal;
    std::vector<double> dual;
    double objective{0.0};
    double primal_infeasibility{0.0};
    double dual_infeasibility{0.0};
    double duality_gap{0.0};
    double tolerance{1e-4};
    std::size_t iterations{0};
    std::string message;
    double h2d_ms{0.0};
    double kernel_ms{0.0};
    double d2h_ms{0.0};
    double total_ms{0.0};
    // Path that actually ran: "cpu", "hip", or "cpu_fallback" (GPU requested,
    // no usable device).  Never claims a GPU that did not run.
    std::string backend_actually_used{"cpu"};
    // --- extensions ---
    std::size_t fp32_iterations{0};    // iterations with FP32 SpMVs (mixed)
    double setup_ms{0.0};
    double primal_weight{0.0};  // final omega (scaled space): seeds initial_primal_weight of a warm start
    double step_size{0.0};      // final eta
    std::size_t packed_iterations{0};  // iterations on the packed stream (packed, or mixed after FP32)
};

// Solve an LP with matrix-free PDHG (only A*x and A^T*y, no factorization).
[[nodiscard]] PdlpResult solve_pdlp(const model::Model& model, const PdlpOptions& options = {});

// Extension: solve independent LPs together on the GPU (one launch per phase
// for the whole batch, per-instance step/restart/weight state).  Results are
// in input order; with Backend::cpu the models are solved one after another.
[[nodiscard]] std::vector<PdlpResult> solve_pdlp_batch(const std::vector<const model::Model*>& models,
                                                       const PdlpOptions& options = {});

}  // namespace markov_cero::lp::first_order
