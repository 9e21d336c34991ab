// mps.cpp -- free-format MPS reader/writer and the binary LP format.
#include "markov_cero/io/mps.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <unordered_map>
#include <vector>

namespace markov_cero::io {
namespace {

std::vector<std::string> split(const std::string& line) {
    std::vector<std::string> out;
    std::istringstream ss(line);
    std::string tok;
    while (ss >> tok) out.push_back(tok);
    return out;
}

double number(const std::string& s, std::size_t line) {
    char* end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || *end != '\0' || !std::isfinite(v)) throw MpsError(line, "bad number '" + s + "'");
    return v;
}

struct RowRec {
    std::string name;
    char type;
    double rhs = 0.0;
    double range = 0.0;
    bool has_range = false;
};

struct ColRec {
    std::string name;
    double obj = 0.0;
    model::Bound lo = model::Bound::finite(0.0);
    model::Bound hi = model::Bound::positive_infinity();
    model::VariableType type = model::VariableType::continuous;
};


// Reader with optional size caps (lim == nullptr: none).
model::Model parse_capped(std::istream& in, const MpsLimits* lim) {
    enum class Sec { none, name, objsense, rows, columns, rhs, ranges, bounds, end };
    Sec sec = Sec::none;
    std::string line, name = "MPS";
    std::size_t ln = 0, bytes = 0;
    model::ObjectiveSense sense = model::ObjectiveSense::minimize;
    std::vector<RowRec> rows;
    std::unordered_map<std::string, std::size_t> row_of, col_of;
    std::vector<ColRec> cols;
    struct Entry {
        std::size_t row, col;
        double v;
    };
    std::vector<Entry> entries;
    std::string objective_row;
    bool integer_block = false;
    auto col_index = [&](const std::string& c) {
        auto it = col_of.find(c);
        if (it != col_of.end()) return it->second;
        col_of.emplace(c, cols.size());
        if (lim && cols.size() >= lim->maximum_columns) throw MpsError(ln, "too many columns");
        cols.push_back(ColRec{c});
        return cols.size() - 1;
    };
    auto row_index = [&](const std::string& r) {
        auto it = row_of.find(r);
        if (it == row_of.end()) throw MpsError(ln, "unknown row '" + r + "'");
        return it->second;
    };
    while (std::getline(in, line)) {
        ++ln;
        if (lim) {
            bytes += line.size() + 1;
            if (ln > lim->maximum_lines) throw MpsError(ln, "too many lines");
            if (bytes > lim->maximum_bytes) throw MpsError(ln, "input exceeds the byte limit");
        }
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty() || line[0] == '*') continue;
        const auto f = split(line);
        if (f.empty()) continue;
        if (lim)
            for (const auto& tok : f)
                if (tok.size() > lim->maximum_name_bytes) throw MpsError(ln, "field exceeds the name limit");
        if (line[0] != ' ' && line[0] != '\t') {
            const std::string& h = f[0];
            if (h == "NAME") {
                sec = Sec::name;
                if (f.size() > 1) name = f[1];
            } else if (h == "OBJSENSE") {
                sec = Sec::objsense;
                if (f.size() > 1) sense = (f[1] == "MAX" || f[1] == "MAXIMIZE") ? model::ObjectiveSense::maximize
                                                                                : model::ObjectiveSense::minimize;
            } else if (h == "ROWS") {
                sec = Sec::rows;
            } else if (h == "COLUMNS") {
                sec = Sec::columns;
            } else if (h == "RHS") {
                sec = Sec::rhs;
            } else if (h == "RANGES") {
                sec = Sec::ranges;
            } else if (h == "BOUNDS") {
                sec = Sec::bounds;
            } else if (h == "ENDATA") {
                sec = Sec::end;
                break;
            } else {
                throw MpsError(ln, "unknown section '" + h + "'");
            }
            continue;
        }
        switch (sec) {
            case Sec::objsense:
                sense = (f[0] == "MAX" || f[0] == "MAXIMIZE") ? model::ObjectiveSense::maximize
                                                              : model::ObjectiveSense::minimize;
                break;
            case Sec::rows: {
                if (f.size() != 2 || f[0].size() != 1 || std::strchr("NELG", f[0][0]) == nullptr)
                    throw MpsError(ln, "bad ROWS record");
                if (row_of.count(f[1])) throw MpsError(ln, "duplicate row '" + f[1] + "'");
                if (f[0][0] == 'N' && objective_row.empty()) objective_row = f[1];
                row_of.emplace(f[1], rows.size());
                if (lim && rows.size() >= lim->maximum_rows) throw MpsError(ln, "too many rows");
                rows.push_back(RowRec{f[1], f[0][0]});
                break;
            }
            case Sec::columns: {
                if (f.size() == 3 && f[1].find("MARKER") != std::string::npos) {
                    integer_block = f[2].find("INTORG") != std::string::npos;
                    break;
                }
                if (f.size() != 3 && f.size() != 5) throw MpsError(ln, "bad COLUMNS record");
                const std::size_t c = col_index(f[0]);
                if (integer_block) cols[c].type = model::VariableType::integer;
                for (std::size_t p = 1; p + 1 < f.size(); p += 2) {
                    const std::size_t r = row_index(f[p]);
                    const double v = number(f[p + 1], ln);
                    if (rows[r].type == 'N') {
                        if (f[p] == objective_row) cols[c].obj += v;
                    } else {
                        if (lim && entries.size() >= lim->maximum_nonzeros) throw MpsError(ln, "too many nonzeros");
                        entries.push_back({r, c, v});
                    }
                }
                break;
            }
            case Sec::rhs:
            case Sec::ranges: {
                const std::size_t start = (f.size() % 2 == 1) ? 1 : 0;
                for (std::size_t p = start; p + 1 < f.size(); p += 2) {
                    const std::size_t r = row_index(f[p]);
                    if (rows[r].type == 'N') throw MpsError(ln, "objective-row RHS/RANGES values are unsupported");
                    const double v = number(f[p + 1], ln);
                    if (sec == Sec::rhs) {
                        rows[r].rhs = v;
                    } else {
                        rows[r].range = v;
                        rows[r].has_range = true;
                    }
                }
                break;
            }
            case Sec::bounds: {
                if (f.size() != 3 && f.size() != 4) throw MpsError(ln, "bad BOUNDS record");
                const std::string& t = f[0];
                ColRec& c = cols[col_index(f[2])];
                const double v = f.size() == 4 ? number(f[3], ln) : 0.0;
                if (t == "LO") c.lo = model::Bound::finite(v);
                else if (t == "UP") c.hi = model::Bound::finite(v);
                else if (t == "FX") c.lo = c.hi = model::Bound::finite(v);
                else if (t == "FR") { c.lo = model::Bound::negative_infinity(); c.hi = model::Bound::positive_infinity(); }
                else if (t == "MI") c.lo = model::Bound::negative_infinity();
                else if (t == "PL") c.hi = model::Bound::positive_infinity();
                else if (t == "BV") { c.type = model::VariableType::binary; c.lo = model::Bound::finite(0.0); c.hi = model::Bound::finite(1.0); }
                else if (t == "LI") { c.type = model::VariableType::integer; c.lo = model::Bound::finite(v); }
                else if (t == "UI") { c.type = model::VariableType::integer; c.hi = model::Bound::finite(v); }
                else throw MpsError(ln, "unsupported bound type: " + t);
                break;
            }
            default: throw MpsError(ln, "record outside a supported section");
        }
    }
    if (sec != Sec::end) throw MpsError(ln, "missing ENDATA");
    if (objective_row.empty()) throw MpsError(ln, "missing objective N row");
    std::vector<std::size_t> map(rows.size(), std::numeric_limits<std::size_t>::max());
    model::Model md;
    md.name = name;
    md.objective_sense = sense;
    for (std::size_t k = 0; k < rows.size(); ++k) {
        const RowRec& r = rows[k];
        if (r.type == 'N') continue;
        map[k] = md.row_name.size();
        md.row_name.push_back(r.name);
        model::Bound lo = model::Bound::negative_infinity(), hi = model::Bound::positive_infinity();
        if (r.type == 'E') lo = hi = model::Bound::finite(r.rhs);
        else if (r.type == 'L') hi = model::Bound::finite(r.rhs);
        else lo = model::Bound::finite(r.rhs);
        if (r.has_range) {
            const double w = std::abs(r.range);
            if (r.type == 'L') lo = model::Bound::finite(r.rhs - w);
            else if (r.type == 'G') hi = model::Bound::finite(r.rhs + w);
            else if (r.range >= 0.0) hi = model::Bound::finite(r.rhs + w);
            else lo = model::Bound::finite(r.rhs - w);
        }
        md.row_lower.push_back(lo);
        md.row_upper.push_back(hi);
    }
    model::SparseMatrixBuilder b(md.row_name.size(), cols.size());
    for (const auto& e : entries) b.add(map[e.row], e.col, e.v);
    md.matrix = b.build();
    for (const auto& c : cols) {
        md.variable_name.push_back(c.name);
        md.objective.push_back(c.obj);
        md.variable_lower.push_back(c.lo);
        md.variable_upper.push_back(c.hi);
        md.variable_type.push_back(c.type);
    }
    md.validate();
    return md;
}

}  // namespace

model::Model parse_mps(std::istream& in) { return parse_capped(in, nullptr); }

model::Model parse_mps(std::istream& in, const MpsLimits& limits) { return parse_capped(in, &limits); }

model::Model parse_mps_string(std::string_view text, const MpsLimits& limits) {
    std::istringstream in{std::string(text)};
    return parse_capped(in, &limits);
}

model::Model read_mps_file(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    return parse_mps(f);
}

void write_mps(const model::Model& md, std::ostream& out) {
    out << std::setprecision(17);
    out << "NAME          " << (md.name.empty() ? "LP" : md.name) << "\n";
    if (md.objective_sense == model::ObjectiveSense::maximize) out << "OBJSENSE\n    MAX\n";
    out << "ROWS\n N  COST\n";
    const std::size_t m = md.matrix.row_count, n = md.matrix.column_count;
    auto rname = [&](std::size_t i) { return i < md.row_name.size() ? md.row_name[i] : "R" + std::to_string(i); };
    auto cname = [&](std::size_t j) {
        return j < md.variable_name.size() ? md.variable_name[j] : "C" + std::to_string(j);
    };
    std::vector<char> type(m);
    for (std::size_t i = 0; i < m; ++i) {
        const auto& lo = md.row_lower[i];
        const auto& hi = md.row_upper[i];
        if (lo.is_finite() && hi.is_finite() && lo.value == hi.value) type[i] = 'E';
        else if (hi.is_finite()) type[i] = 'L';
        else if (lo.is_finite()) type[i] = 'G';
        else type[i] = 'N';
        out << ' ' << (type[i] == 'N' ? 'N' : type[i]) << "  " << rname(i) << "\n";
    }
    out << "COLUMNS\n";
    for (std::size_t j = 0; j < n; ++j) {
        if (md.objective[j] != 0.0) out << "    " << cname(j) << " COST " << md.objective[j] << "\n";
        for (std::size_t k = md.matrix.column_start[j]; k < md.matrix.column_start[j + 1]; ++k)
            out << "    " << cname(j) << ' ' << rname(md.matrix.row_index[k]) << ' ' << md.matrix.value[k] << "\n";
    }
    out << "RHS\n";
    for (std::size_t i = 0; i < m; ++i) {
        const double rhs = type[i] == 'G' ? md.row_lower[i].value : (type[i] == 'N' ? 0.0 : md.row_upper[i].value);
        if (rhs != 0.0) out << "    RHS " << rname(i) << ' ' << rhs << "\n";
    }
    bool ranges = false;
    for (std::size_t i = 0; i < m; ++i)
        if (type[i] == 'L' && md.row_lower[i].is_finite()) {
            if (!ranges) out << "RANGES\n";
            ranges = true;
            out << "    RNG " << rname(i) << ' ' << (md.row_upper[i].value - md.row_lower[i].value) << "\n";
        }
    out << "BOUNDS\n";
    for (std::size_t j = 0; j < n; ++j) {
        const auto& lo = md.variable_lower[j];
        const auto& hi = md.variable_upper[j];
        if (lo.is_finite() && hi.is_finite() && lo.value == hi.value) {
            out << " FX BND " << cname(j) << ' ' << lo.value << "\n";
            continue;
        }
        if (!lo.is_finite() && !hi.is_finite()) {
            out << " FR BND " << cname(j) << "\n";
            continue;
        }
        if (!lo.is_finite()) out << " MI BND " << cname(j) << "\n";
        else if (lo.value != 0.0) out << " LO BND " << cname(j) << ' ' << lo.value << "\n";
        if (hi.is_finite()) out << " UP BND " << cname(j) << ' ' << hi.value << "\n";
    }
    out << "ENDATA\n";
}

void write_binary_lp(const model::Model& md, const std::string& path) {
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    const int64_t m = (int64_t)md.matrix.row_count, n = (int64_t)md.matrix.column_count,
                  nnz = (int64_t)md.matrix.value.size();
    const int32_t sense = md.objective_sense == model::ObjectiveSense::maximize ? 1 : 0, pad = 0;
    auto put = [&](const void* p, std::size_t bytes) {
        if (bytes && std::fwrite(p, 1, bytes, f) != bytes) throw std::runtime_error("short write " + path);
    };
    put("MCLP0001", 8);
    put(&m, 8);
    put(&n, 8);
    put(&nnz, 8);
    put(&sense, 4);
    put(&pad, 4);
    put(&md.objective_offset, 8);
    put(md.matrix.column_start.data(), 8 * (n + 1));
    put(md.matrix.row_index.data(), 8 * nnz);
    put(md.matrix.value.data(), 8 * nnz);
    put(md.objective.data(), 8 * n);
    put(md.row_lower.data(), 16 * m);
    put(md.row_upper.data(), 16 * m);
    put(md.variable_lower.data(), 16 * n);
    put(md.variable_upper.data(), 16 * n);
    std::fclose(f);
}

model::Model read_binary_lp(const std::string& path) {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    auto get = [&](void* p, std::size_t bytes) {
        if (bytes && std::fread(p, 1, bytes, f) != bytes) {
            std::fclose(f);
            throw std::runtime_error("truncated binary LP " + path);
        }
    };
    char magic[8];
    get(magic, 8);
    if (std::memcmp(magic, "MCLP0001", 8) != 0) {
        std::fclose(f);
        throw std::runtime_error("not a binary LP file: " + path);
    }
    int64_t m, n, nnz;
    int32_t sense, pad;
    model::Model md;
    get(&m, 8);
    get(&n, 8);
    get(&nnz, 8);
    get(&sense, 4);
    get(&pad, 4);
    get(&md.objective_offset, 8);
    md.objective_sense = sense ? model::ObjectiveSense::maximize : model::ObjectiveSense::minimize;
    md.matrix.row_count = (std::size_t)m;
    md.matrix.column_count = (std::size_t)n;
    md.matrix.column_start.resize(n + 1);
    md.matrix.row_index.resize(nnz);
    md.matrix.value.resize(nnz);
    md.objective.resize(n);
    md.row_lower.resize(m);
    md.row_upper.resize(m);
    md.variable_lower.resize(n);
    md.variable_upper.resize(n);
    get(md.matrix.column_start.data(), 8 * (n + 1));
    get(md.matrix.row_index.data(), 8 * nnz);
    get(md.matrix.value.data(), 8 * nnz);
    get(md.objective.data(), 8 * n);
    get(md.row_lower.data(), 16 * m);
    get(md.row_upper.data(), 16 * m);
    get(md.variable_lower.data(), 16 * n);
    get(md.variable_upper.data(), 16 * n);
    std::fclose(f);
    md.name = path;
    md.variable_type.assign(n, model::VariableType::continuous);
    for (int64_t i = 0; i < m; ++i) md.row_name.push_back("R" + std::to_string(i));
    for (int64_t j = 0; j < n; ++j) md.variable_name.push_back("C" + std::to_string(j));
    return md;
}

}  // namespace markov_cero::io
