// markov-cero-solve -- CLI entry point of the PDLP path.
//
//   markov-cero-solve MODEL.(mps|mclp) [--engine pdlp] [--backend cpu|gpu]
//                     [--tolerance T] [--iteration-limit N] [--output f.json]
//                     [--precision fp64|mixed|packed] [--order fast|reference] [--autotune]
//                     [--restart-scheme reference|kkt|halpern] [--row-weighting 0..8]
//                     [--primal-weight-max-drop K]
//
// Mirrors the reference's pdlp branch (apps/markov_cero_solve.cpp:136-189):
// --backend/--tolerance/--iteration-limit map into PdlpOptions (an iteration
// limit of exactly 10000, the simplex default, means 100000), the returned
// primal is checked by verify_primal at the PDLP tolerance, and one JSON
// record with the reference's PDLP keys goes to stdout (and --output), plus a
// one-line verdict with the timing split on stderr.  Exit codes follow
// apps/json_output.hpp:19-32 (0 optimal, 3 invalid model, 6 iteration limit,
// 7 numerical failure).  Only the pdlp engine is provided.
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>

#include "markov_cero/io/mps.hpp"
#include "markov_cero/lp/first_order/pdlp.hpp"
#include "markov_cero/verify/primal_verifier.hpp"

namespace fo = markov_cero::lp::first_order;

namespace {

// %.17g (the reference's 17 significant digits, json_output.hpp:69-78) via
// std::to_chars: the same characters as an ostream at precision 17, without a
// stream per number (n = 4.8M primals take a fraction of a second).
std::string num(double v) {
    if (!std::isfinite(v)) return "null";
    char buf[40];
    const auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::general, 17);
    return std::string(buf, r.ptr);
}

std::string esc(const std::string& s) {
    std::string o;
    for (unsigned char c : s) {
        if (c == '"' || c == '\\') {
            o += '\\';
            o += (char)c;
        } else if (c == '\n') {
            o += "\\n";
        } else if (c < 0x20) {
            char b[8];
            std::snprintf(b, sizeof b, "\\u%04x", c);
            o += b;
        } else {
            o += (char)c;
        }
    }
    return o;
}

int usage() {
    std::cerr << "usage: markov-cero-solve MODEL.mps|MODEL.mclp [--engine pdlp] [--backend cpu|gpu]\n"
                 "       [--tolerance T] [--iteration-limit N] [--output FILE]\n"
                 "       [--precision fp64|mixed|packed] [--order fast|reference] [--autotune]\n"
                 "       [--restart-scheme reference|kkt|halpern] [--row-weighting 0..8]\n"
                 "       [--primal-weight-max-drop K]\n";
    return 4;
}

}  // namespace

int main(int argc, char** argv) {
    const auto started = std::chrono::steady_clock::now();
    std::string path, engine = "pdlp", backend = "cpu", output, precision = "fp64", order = "fast",
                scheme = "reference";
    double tol = 1e-4;
    std::size_t iteration_limit = 10000;
    bool autotune = false;
    int row_weighting = 0;
    int max_drop = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) {
                std::cerr << "missing value for " << a << "\n";
                std::exit(usage());
            }
            return argv[++i];
        };
        if (a == "--engine") engine = next();
        else if (a == "--backend") backend = next();
        else if (a == "--tolerance") tol = std::strtod(next().c_str(), nullptr);
        else if (a == "--iteration-limit") iteration_limit = std::strtoull(next().c_str(), nullptr, 10);
        else if (a == "--output") output = next();
        else if (a == "--precision") precision = next();
        else if (a == "--order") order = next();
        else if (a == "--autotune") autotune = true;
        else if (a == "--restart-scheme") scheme = next();
        else if (a == "--row-weighting") row_weighting = std::atoi(next().c_str());
        else if (a == "--primal-weight-max-drop") max_drop = std::atoi(next().c_str());
        else if (!a.empty() && a[0] == '-') return usage();
        else path = a;
    }
    if (path.empty()) return usage();
    if (engine != "pdlp" && engine != "auto") {
        std::cerr << "only --engine pdlp is provided by this build\n";
        return 4;
    }
    if (backend != "cpu" && backend != "gpu") {
        std::cerr << "invalid backend: " << backend << "\n";
        return 4;
    }
    if (scheme != "reference" && scheme != "kkt" && scheme != "halpern") {
        std::cerr << "invalid restart scheme: " << scheme << "\n";
        return 4;
    }
    if (row_weighting < 0 || row_weighting > 8) {
        std::cerr << "row weighting must be in 0..8: " << row_weighting << "\n";
        return 4;
    }
    if (max_drop != 0 && (max_drop < 2 || max_drop > (1 << 20) || (max_drop & (max_drop - 1)) != 0)) {
        std::cerr << "primal weight max drop must be 0 or a power of two in 2..2^20: " << max_drop << "\n";
        return 4;
    }
    if (!(tol > 0.0)) {
        std::cerr << "tolerance must be positive: " << tol << "\n";
        return 4;
    }
    std::string status = "NumericalFailure", message, error, original_message = "original primal not applicable";
    int code = 7;
    bool verified = false;
    std::size_t rows = 0, cols = 0, nnz = 0;
    fo::PdlpResult res;
    markov_cero::verify::PrimalVerificationReport rep;
    try {
        const bool binary = path.size() > 5 && path.substr(path.size() - 5) == ".mclp";
        const auto md = binary ? markov_cero::io::read_binary_lp(path) : markov_cero::io::read_mps_file(path);
        rows = md.matrix.row_count;
        cols = md.matrix.column_count;
        nnz = md.matrix.value.size();
        fo::PdlpOptions o;
        o.backend = backend == "gpu" ? fo::Backend::gpu : fo::Backend::cpu;
        o.max_iterations = (iteration_limit != 10000 && iteration_limit > 0) ? iteration_limit : 100000;
        o.set_tolerance(tol);
        o.precision = precision == "mixed"        ? fo::Precision::mixed
                      : precision == "packed" ? fo::Precision::packed
                                                  : fo::Precision::fp64;
        o.reduction_order = order == "reference" ? fo::ReductionOrder::reference : fo::ReductionOrder::fast;
        o.autotune = autotune;
        o.row_weighting = row_weighting;
        o.primal_weight_max_drop = max_drop;
        o.restart_scheme = scheme == "kkt"       ? fo::RestartScheme::kkt
                           : scheme == "halpern" ? fo::RestartScheme::halpern
                                                 : fo::RestartScheme::reference;
        res = fo::solve_pdlp(md, o);
        message = res.message;
        if (res.status == fo::PdlpStatus::optimal) {
            status = "Optimal";
            code = 0;
            markov_cero::verify::Candidate cand{res.primal, res.objective};
            rep = markov_cero::verify::verify_primal(md, cand, {tol, tol}, {tol, tol}, tol);
            verified = rep.passed;
            original_message = verified ? "original primal verified" : "original primal rejected";
            if (!verified) {
                status = "NumericalFailure";
                code = 7;
                message = "original-model verification failed";
            }
        } else if (res.status == fo::PdlpStatus::iteration_limit) {
            status = "IterationLimit";
            code = 6;
        }
    } catch (const markov_cero::io::MpsError& e) {
        status = "InvalidModel";
        code = 3;
        message = error = e.what();
    } catch (const std::invalid_argument& e) {
        status = "InvalidModel";
        code = 3;
        message = error = e.what();
    } catch (const std::exception& e) {
        status = "NumericalFailure";
        code = 7;
        message = error = e.what();
    }
    const double elapsed = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - started).count();
    std::ostringstream js;
    js << "{\"engine\":\"pdlp\",\"status\":\"" << status << "\",\"rows\":" << rows << ",\"cols\":" << cols
       << ",\"nonzeros\":" << nnz << ",\"verified\":" << (verified ? "true" : "false") << ",\"message\":\""
       << esc(message) << "\",\"objective\":" << num(res.objective) << ",\"primal\":[";
    for (std::size_t j = 0; j < res.primal.size(); ++j) js << (j ? "," : "") << num(res.primal[j]);
    js << "],\"original_verified\":" << (verified ? "true" : "false") << ",\"original_message\":\""
       << esc(original_message) << "\",\"maximum_primal_violation\":" << num(rep.maximum_row_violation)
       << ",\"maximum_variable_violation\":" << num(rep.maximum_variable_violation)
       << ",\"runtime_ms\":" << num(elapsed) << ",\"lp_iterations\":" << res.iterations
       << ",\"pdlp_tolerance\":" << num(tol) << ",\"relative_primal_residual\":" << num(res.primal_infeasibility)
       << ",\"relative_dual_residual\":" << num(res.dual_infeasibility)
       << ",\"relative_duality_gap\":" << num(res.duality_gap) << ",\"backend\":\"" << backend
       << "\",\"backend_actually_used\":\"" << esc(res.backend_actually_used) << "\",\"h2d_ms\":" << num(res.h2d_ms)
       << ",\"kernel_ms\":" << num(res.kernel_ms) << ",\"d2h_ms\":" << num(res.d2h_ms)
       << ",\"total_ms\":" << num(res.total_ms) << ",\"precision\":\"" << precision
       << "\",\"fp32_iterations\":" << res.fp32_iterations << ",\"packed_iterations\":" << res.packed_iterations
       << ",\"setup_ms\":" << num(res.setup_ms)
       << ",\"restart_scheme\":\"" << scheme << "\",\"dual\":[";
    // extension keys: the dual vector (reference sign convention, g = c + A^T y) and setup time
    for (std::size_t i = 0; i < res.dual.size(); ++i) js << (i ? "," : "") << num(res.dual[i]);
    js << "]";
    if (!error.empty()) js << ",\"error\":\"" << esc(error) << "\"";
    js << "}\n";
    std::cout << js.str();
    if (!output.empty()) {
        std::ofstream f(output);
        if (f)
            f << js.str();
        else
            std::cerr << "cannot write output\n";
    }
    const std::string timing = backend == "gpu" ? " [gpu H2D=" + num(res.h2d_ms) + "ms kernel=" + num(res.kernel_ms) +
                                                      "ms D2H=" + num(res.d2h_ms) + "ms total=" + num(res.total_ms) + "ms]"
                                                : " [cpu total=" + num(res.total_ms) + "ms]";
    std::cerr << "markov-cero " << status << " [tol=" << num(tol) << "]" << timing << " backend_actually_used="
              << res.backend_actually_used << (verified ? " VERIFIED\n" : " NOT VERIFIED\n");
    return code;
}
