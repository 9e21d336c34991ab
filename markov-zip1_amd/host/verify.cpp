// verify.cpp -- primal verification (reference src/verify/primal_verifier.cpp:9-104):
// bound and row checks with allowance abs + rel * max(1, |actual|, |bound|),
// long-double row activity and objective recomputation.
#include "markov_cero/verify.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace markov_cero::verify {
namespace {

double allowed(double actual, double bound, const Tolerance& t) {
    const double a = t.absolute + t.relative * std::max({1.0, std::abs(actual), std::abs(bound)});
    if (!std::isfinite(a)) throw std::invalid_argument("tolerance allowance overflow");
    return a;
}

// side = +1: actual must be >= bound; side = -1: actual must be <= bound.
void check(double actual, const model::Bound& bound, int side, const Tolerance& t, const char* category,
           std::size_t index, PrimalVerificationReport& rep, double& worst) {
    if (!bound.is_finite()) return;
    const double excess = std::max(0.0, side > 0 ? bound.value - actual : actual - bound.value);
    const double allow = allowed(actual, bound.value, t);
    worst = std::max(worst, excess);
    if (excess > allow) rep.violations.push_back({category, index, actual, bound.value, excess, allow});
}

}  // namespace

void Tolerance::validate() const {
    if (!std::isfinite(absolute) || !std::isfinite(relative) || absolute < 0.0 || relative < 0.0 ||
        absolute > 1e-4 || relative > 1e-4)
        throw std::invalid_argument("tolerances must be finite and nonnegative");
}

PrimalVerificationReport verify_primal(const model::Model& md, const Candidate& cand, const Tolerance& feas,
                                       const Tolerance& objtol, double int_tol) {
    md.validate();
    feas.validate();
    objtol.validate();
    if (!std::isfinite(int_tol) || int_tol < 0.0 || int_tol >= 0.5)
        throw std::invalid_argument("integrality tolerance must be finite and nonnegative");
    if (cand.primal.size() != md.matrix.column_count) throw std::invalid_argument("candidate dimension mismatch");
    if (!std::isfinite(cand.claimed_objective)) throw std::invalid_argument("claimed objective must be finite");
    PrimalVerificationReport rep;
    for (std::size_t j = 0; j < cand.primal.size(); ++j) {
        const double x = cand.primal[j];
        if (!std::isfinite(x)) throw std::invalid_argument("candidate contains non-finite value");
        check(x, md.variable_lower[j], +1, feas, "variable_lower", j, rep, rep.maximum_variable_violation);
        check(x, md.variable_upper[j], -1, feas, "variable_upper", j, rep, rep.maximum_variable_violation);
        if (md.variable_type[j] != model::VariableType::continuous) {
            const double lo = std::floor(x), hi = std::ceil(x);
            const double nearest = (x - lo <= hi - x) ? lo : hi;
            const double frac = std::abs(x - nearest);
            rep.maximum_integrality_violation = std::max(rep.maximum_integrality_violation, frac);
            if (frac > int_tol) rep.violations.push_back({"integrality", j, x, nearest, frac, int_tol});
        }
    }
    const auto act = md.matrix.multiply(cand.primal);
    for (std::size_t i = 0; i < act.size(); ++i) {
        check(act[i], md.row_lower[i], +1, feas, "row_lower", i, rep, rep.maximum_row_violation);
        check(act[i], md.row_upper[i], -1, feas, "row_upper", i, rep, rep.maximum_row_violation);
    }
    long double obj = md.objective_offset;
    for (std::size_t j = 0; j < cand.primal.size(); ++j)
        obj += static_cast<long double>(md.objective[j]) * cand.primal[j];
    rep.recomputed_objective = static_cast<double>(obj);
    if (!std::isfinite(rep.recomputed_objective)) throw std::overflow_error("recomputed objective is non-finite");
    rep.objective_difference = std::abs(rep.recomputed_objective - cand.claimed_objective);
    const double allow = allowed(rep.recomputed_objective, cand.claimed_objective, objtol);
    if (rep.objective_difference > allow)
        rep.violations.push_back({"objective", 0, cand.claimed_objective, rep.recomputed_objective,
                                  rep.objective_difference, allow});
    rep.passed = rep.violations.empty();
    return rep;
}

}  // namespace markov_cero::verify
