// verify.cpp -- zero-trust check of a returned primal against the original
// model.  The contract is that of reference src/verify/primal_verifier.cpp:
// each finite bound of a variable and of a row activity may be exceeded by at
// most abs + rel * max(1, |actual|, |bound|); row activities and the objective
// are recomputed in long double; integer-typed variables must lie within the
// integrality tolerance of an integer.  Those formulas are kept exactly, so
// verdicts agree with the reference's; the code is organised around one
// audit object instead.
#include "markov_cero/verify/primal_verifier.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <thread>
#include <vector>

namespace markov_cero::verify {
namespace {

// Collects violations of one candidate.
struct Audit {
    const Tolerance& feas;
    PrimalVerificationReport& rep;

    static double allowance(double actual, double bound, const Tolerance& t) {
        const double scale = std::max({1.0, std::abs(actual), std::abs(bound)});
        const double a = t.absolute + t.relative * scale;
        if (!std::isfinite(a)) throw std::invalid_argument("verification allowance overflows");
        return a;
    }

    // `actual` against the pair [lo, hi]; the excess over each finite side
    // feeds `worst` and, past its allowance, a violation of that side.
    void pair(double actual, const model::Bound& lo, const model::Bound& hi, const char* lo_name,
              const char* hi_name, std::size_t index, double& worst) {
        if (lo.is_finite()) side(actual, lo.value, std::max(0.0, lo.value - actual), lo_name, index, worst);
        if (hi.is_finite()) side(actual, hi.value, std::max(0.0, actual - hi.value), hi_name, index, worst);
    }

    void side(double actual, double bound, double excess, const char* name, std::size_t index, double& worst) {
        worst = std::max(worst, excess);
        const double allow = allowance(actual, bound, feas);
        if (excess > allow) rep.violations.push_back({name, index, actual, bound, excess, allow});
    }

    // Distance to the nearest integer (ties toward the lower one).
    void integral(double x, std::size_t index, double int_tol) {
        const double down = std::floor(x);
        const double nearest = (x - down <= 0.5) ? down : down + 1.0;
        const double frac = std::abs(x - nearest);
        rep.maximum_integrality_violation = std::max(rep.maximum_integrality_violation, frac);
        if (frac > int_tol) rep.violations.push_back({"integrality", index, x, nearest, frac, int_tol});
    }
};

// Row activities A x of a validated matrix, each row's terms accumulated in
// long double in increasing column order -- the order of
// SparseMatrixCSC::multiply, so the values are identical.  Large matrices are
// split by row ranges over threads: every thread walks all columns and adds
// only the entries of its rows (found by binary search in each sorted
// column), so no row's order of additions changes.
std::vector<double> row_activity(const model::SparseMatrixCSC& A, const std::vector<double>& x) {
    const std::size_t m = A.row_count, n = A.column_count, nnz = A.value.size();
    std::vector<double> out(m);
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const std::size_t parts = nnz < (std::size_t)1 << 20 ? 1 : std::min<std::size_t>({hw, 16, m});
    auto run = [&](std::size_t r0, std::size_t r1) {
        std::vector<long double> acc(r1 - r0, 0.0L);
        for (std::size_t j = 0; j < n; ++j) {
            const auto b = A.row_index.begin() + (std::ptrdiff_t)A.column_start[j];
            const auto e = A.row_index.begin() + (std::ptrdiff_t)A.column_start[j + 1];
            const long double xj = x[j];
            for (auto it = parts == 1 ? b : std::lower_bound(b, e, r0); it != e && *it < r1; ++it)
                acc[*it - r0] += static_cast<long double>(A.value[(std::size_t)(it - A.row_index.begin())]) * xj;
        }
        for (std::size_t i = r0; i < r1; ++i) out[i] = static_cast<double>(acc[i - r0]);
    };
    if (parts <= 1) {
        run(0, m);
    } else {
        std::vector<std::thread> pool;
        for (std::size_t k = 0; k < parts; ++k) pool.emplace_back(run, m * k / parts, m * (k + 1) / parts);
        for (auto& t : pool) t.join();
    }
    for (std::size_t i = 0; i < m; ++i)
        if (!std::isfinite(out[i])) throw std::overflow_error("row activity overflows in row " + std::to_string(i));
    return out;
}

}  // namespace

// Allowances above 1e-4 are refused (reference primal_verifier.cpp:39-43).
void Tolerance::validate() const {
    const bool ok = std::isfinite(absolute) && std::isfinite(relative) && absolute >= 0.0 && relative >= 0.0 &&
                    absolute <= 1e-4 && relative <= 1e-4;
    if (!ok) throw std::invalid_argument("verification tolerances must be finite, nonnegative and at most 1e-4");
}

PrimalVerificationReport verify_primal(const model::Model& md, const Candidate& cand, const Tolerance& feas,
                                       const Tolerance& objtol, double int_tol) {
    md.validate();
    feas.validate();
    objtol.validate();
    if (!(int_tol >= 0.0 && int_tol < 0.5)) throw std::invalid_argument("integrality tolerance must lie in [0, 0.5)");
    const std::size_t n = md.matrix.column_count;
    if (cand.primal.size() != n) throw std::invalid_argument("candidate has the wrong number of variables");
    if (!std::isfinite(cand.claimed_objective)) throw std::invalid_argument("claimed objective is not finite");
    for (const double x : cand.primal)
        if (!std::isfinite(x)) throw std::invalid_argument("candidate holds a non-finite value");

    PrimalVerificationReport rep;
    Audit audit{feas, rep};
    for (std::size_t j = 0; j < n; ++j) {
        audit.pair(cand.primal[j], md.variable_lower[j], md.variable_upper[j], "variable_lower", "variable_upper", j,
                   rep.maximum_variable_violation);
        if (md.variable_type[j] != model::VariableType::continuous) audit.integral(cand.primal[j], j, int_tol);
    }
    const std::vector<double> activity = row_activity(md.matrix, cand.primal);  // md validated above
    for (std::size_t i = 0; i < activity.size(); ++i)
        audit.pair(activity[i], md.row_lower[i], md.row_upper[i], "row_lower", "row_upper", i,
                   rep.maximum_row_violation);

    long double objective = md.objective_offset;
    for (std::size_t j = 0; j < n; ++j) objective += static_cast<long double>(md.objective[j]) * cand.primal[j];
    rep.recomputed_objective = static_cast<double>(objective);
    if (!std::isfinite(rep.recomputed_objective)) throw std::overflow_error("recomputed objective overflows");
    rep.objective_difference = std::abs(rep.recomputed_objective - cand.claimed_objective);
    const double allow = Audit::allowance(rep.recomputed_objective, cand.claimed_objective, objtol);
    if (rep.objective_difference > allow)
        rep.violations.push_back({"objective", 0, cand.claimed_objective, rep.recomputed_objective,
                                  rep.objective_difference, allow});
    rep.passed = rep.violations.empty();
    return rep;
}

}  // namespace markov_cero::verify
