// MPS reader for the PDHG path. Behaviour follows the reference reader src/mps.cpp:1-399 rule for rule
// (free-format tokens first, fixed columns as the fallback; strict decimal numbers; duplicate entries summed
// and zero sums dropped; first RHS/RANGES/BOUNDS set only; RANGES semantics; +-1e20 sentinels read as
// infinite; names must be declared before use), with its own single-pass, in-memory implementation.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string_view>
#include <unordered_map>
#include <unordered_set>

#include "model.hpp"

namespace tpdhg {

bool Model::has_integers() const {
    return std::any_of(is_int.begin(), is_int.end(), [](char c) { return c != 0; });
}

size_t Model::nnz() const {
    size_t k = 0;
    for (const auto& c : cols) k += c.size();
    return k;
}

namespace {

constexpr double kSentinel = 1e20;  // |bound| >= 1e20 means infinite (src/mps.cpp:362-376)

// Strict decimal literal: [+-] digits [. digits] [(e|E|d|D) [+-] digits], at least one mantissa digit,
// finite value (src/mps.cpp:27-59). Fortran 'D' exponents are accepted.
double parse_number(std::string_view s) {
    size_t i = 0;
    const size_t n = s.size();
    auto run = [&] {
        size_t a = i;
        while (i < n && s[i] >= '0' && s[i] <= '9') ++i;
        return i - a;
    };
    if (i < n && (s[i] == '+' || s[i] == '-')) ++i;
    size_t mant = run();
    if (i < n && s[i] == '.') {
        ++i;
        mant += run();
    }
    bool ok = mant > 0;
    bool fortran = false;
    if (ok && i < n && (s[i] == 'e' || s[i] == 'E' || s[i] == 'd' || s[i] == 'D')) {
        fortran = s[i] == 'd' || s[i] == 'D';
        ++i;
        if (i < n && (s[i] == '+' || s[i] == '-')) ++i;
        ok = run() > 0;
    }
    if (!ok || i != n) throw ParseError("bad number '" + std::string(s) + "'");
    std::string buf(s);
    if (fortran || buf.find_first_of("Dd") != std::string::npos)
        for (char& ch : buf)
            if (ch == 'd' || ch == 'D') ch = 'E';
    double v = std::strtod(buf.c_str(), nullptr);
    if (!std::isfinite(v)) throw ParseError("number '" + std::string(s) + "' is not finite");
    return v;
}

bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }

// Fixed-column field [a, b) of a line, trimmed of blanks (and a trailing CR); "" when out of range.
std::string_view fixed_field(std::string_view line, size_t a, size_t b) {
    if (line.size() <= a) return {};
    std::string_view s = line.substr(a, b - a);
    size_t lo = s.find_first_not_of(" \t");
    if (lo == std::string_view::npos) return {};
    size_t hi = s.find_last_not_of(" \t\r");
    return s.substr(lo, hi - lo + 1);
}

enum class Sec { None, Name, Objsense, Rows, Columns, Rhs, Ranges, Bounds, Quad };

// Splits one data line into the fields of its section (src/mps.cpp:68-94).
void split(std::string_view line, Sec sec, bool fixed, std::vector<std::string_view>& t) {
    t.clear();
    if (!fixed) {
        size_t p = 0, e = line.size();
        while (p < e) {
            while (p < e && is_ws(line[p])) ++p;
            size_t s = p;
            while (p < e && !is_ws(line[p])) ++p;
            if (p > s) t.push_back(line.substr(s, p - s));
        }
        return;
    }
    if (sec == Sec::Rows) {
        t.push_back(fixed_field(line, 1, 3)), t.push_back(fixed_field(line, 4, 12));
    } else if (sec == Sec::Bounds) {
        t.push_back(fixed_field(line, 1, 3)), t.push_back(fixed_field(line, 4, 12));
        t.push_back(fixed_field(line, 14, 22));
        if (!fixed_field(line, 24, 36).empty()) t.push_back(fixed_field(line, 24, 36));
    } else {
        t.push_back(fixed_field(line, 4, 12)), t.push_back(fixed_field(line, 14, 22));
        t.push_back(fixed_field(line, 24, 36));
        if (!fixed_field(line, 39, 47).empty())
            t.push_back(fixed_field(line, 39, 47)), t.push_back(fixed_field(line, 49, 61));
    }
}

struct Reader {
    bool fixed;
    std::chrono::steady_clock::time_point deadline;
    Model md;
    Sec sec = Sec::None;
    std::string obj_row;
    std::unordered_map<std::string, int> row_id, col_id;
    std::unordered_set<std::string> extra_n_rows;  // N rows after the first are dropped
    std::vector<char> row_kind;
    std::vector<double> rhs, range;
    std::vector<char> ranged;
    bool in_int = false, saw_rows = false, saw_columns = false, saw_end = false;
    int cur = -1;                 // column currently being read
    std::vector<Entry> pending;   // its entries in file order
    std::string rhs_set, range_set, bound_set;
    std::map<std::pair<int, int>, double> quad;
    bool quad_full = false;

    int row_index(std::string_view name) {  // >= 0 row, -1 objective, -2 extra free row
        auto it = row_id.find(std::string(name));
        if (it != row_id.end()) return it->second;
        if (name == obj_row) return -1;
        if (extra_n_rows.count(std::string(name))) return -2;
        throw ParseError("unknown row '" + std::string(name) + "'");
    }
    int column_index(std::string_view name) {
        auto it = col_id.find(std::string(name));
        if (it == col_id.end()) throw ParseError("unknown column '" + std::string(name) + "'");
        return it->second;
    }
    static bool in_first_set(std::string& keep, std::string_view name) {
        if (keep.empty()) keep = std::string(name);
        return keep == name;
    }
    // Sorted, duplicate-summed (file order), zero-free entries of the column just finished.
    void close_column() {
        if (cur < 0 || pending.empty()) return;
        std::stable_sort(pending.begin(), pending.end(), [](const Entry& a, const Entry& b) { return a.index < b.index; });
        std::vector<Entry>& out = md.cols[size_t(cur)];
        for (const Entry& e : pending) {
            if (!out.empty() && out.back().index == e.index) out.back().value += e.value;
            else out.push_back(e);
        }
        out.erase(std::remove_if(out.begin(), out.end(), [](const Entry& e) { return e.value == 0; }), out.end());
        pending.clear();
    }

    void header(std::string_view line) {
        std::istringstream hs{std::string(line)};
        std::string word, arg;
        hs >> word >> arg;
        if (word == "NAME") {
            md.name = arg;
            sec = Sec::Name;
        } else if (word == "ENDATA") {
            saw_end = true;
        } else if (word == "OBJSENSE") {
            if (!arg.empty()) {
                if (arg == "MAX" || arg == "MAXIMIZE") md.maximize = true;
                else if (arg != "MIN" && arg != "MINIMIZE") throw ParseError("bad OBJSENSE " + arg);
            }
            sec = Sec::Objsense;
        } else if (word == "QUADOBJ" || word == "QMATRIX") {
            quad_full = word == "QMATRIX";
            sec = Sec::Quad;
        } else if (word == "QSECTION" || word == "QCMATRIX") {
            if (word == "QCMATRIX" || arg != obj_row)
                throw ParseError("quadratic constraints are not supported (" + word + " " + arg + ")");
            quad_full = true;
            sec = Sec::Quad;
        } else if (word == "ROWS") {
            if (saw_rows) throw ParseError("section ROWS appears twice");
            saw_rows = true, sec = Sec::Rows;
        } else if (word == "COLUMNS") {
            if (saw_columns) throw ParseError("section COLUMNS appears twice");
            saw_columns = true, sec = Sec::Columns;
        } else if (word == "RHS") {
            sec = Sec::Rhs;
        } else if (word == "RANGES") {
            sec = Sec::Ranges;
        } else if (word == "BOUNDS") {
            sec = Sec::Bounds;
        } else {
            throw ParseError("unsupported section " + word);
        }
    }

    void data(const std::vector<std::string_view>& t) {
        switch (sec) {
            case Sec::Objsense: {
                std::string_view v = t.empty() ? std::string_view() : t[0];
                if (v == "MAX" || v == "MAXIMIZE") md.maximize = true;
                else if (v != "MIN" && v != "MINIMIZE") throw ParseError("bad OBJSENSE " + std::string(v));
                return;
            }
            case Sec::Rows: {
                if (t.size() != 2 || t[1].empty()) throw ParseError("bad ROWS line");
                char kind = t[0].empty() ? '?' : char(std::toupper((unsigned char)t[0][0]));
                std::string name(t[1]);
                if (row_id.count(name) || name == obj_row || extra_n_rows.count(name))
                    throw ParseError("duplicate row name '" + name + "'");
                if (kind == 'N') {
                    if (obj_row.empty()) obj_row = name;
                    else extra_n_rows.insert(name);
                } else if (kind == 'E' || kind == 'L' || kind == 'G') {
                    row_id.emplace(name, int(md.row_names.size()));
                    md.row_names.push_back(name);
                    row_kind.push_back(kind);
                } else {
                    throw ParseError("bad row type " + std::string(t[0]));
                }
                return;
            }
            case Sec::Columns: {
                if (t.size() >= 2 && t[1] == "'MARKER'") {
                    bool open = false, close = false;
                    for (auto w : t) open |= w == "'INTORG'", close |= w == "'INTEND'";
                    if (open == close) throw ParseError("bad MARKER line");
                    in_int = open;
                    return;
                }
                if (t.size() != 3 && t.size() != 5) throw ParseError("bad COLUMNS line");
                int j;
                if (cur >= 0 && t[0] == md.col_names[size_t(cur)]) {
                    j = cur;
                } else {
                    std::string name(t[0]);
                    auto it = col_id.find(name);
                    if (it == col_id.end()) {
                        close_column();
                        j = int(md.col_names.size());
                        col_id.emplace(name, j);
                        md.col_names.push_back(name);
                        md.cost.push_back(0);
                        md.is_int.push_back(in_int);
                        md.cols.emplace_back();
                    } else {
                        j = it->second;
                        if (j != cur) throw ParseError("column '" + name + "' is listed in two separate blocks");
                    }
                }
                cur = j;
                for (size_t k = 1; k + 1 < t.size(); k += 2) {
                    int r = row_index(t[k]);
                    double v = parse_number(t[k + 1]);
                    if (r == -1) md.cost[size_t(j)] += v;
                    else if (r >= 0) pending.push_back({r, v});
                }
                return;
            }
            case Sec::Rhs:
            case Sec::Ranges: {
                size_t first = (fixed || t.size() % 2 == 1) ? 1 : 0;  // optional set name
                if (first == 1 && !t[0].empty() && !in_first_set(sec == Sec::Rhs ? rhs_set : range_set, t[0])) return;
                if (rhs.empty()) {
                    rhs.assign(md.row_names.size(), 0);
                    range.assign(md.row_names.size(), 0);
                    ranged.assign(md.row_names.size(), 0);
                }
                for (size_t k = first; k + 1 < t.size(); k += 2) {
                    int r = row_index(t[k]);
                    double v = parse_number(t[k + 1]);
                    if (sec == Sec::Rhs) {
                        if (r == -1) md.obj_const = -v;
                        else if (r >= 0) rhs[size_t(r)] = v;
                    } else if (r >= 0) {
                        range[size_t(r)] = v, ranged[size_t(r)] = 1;
                    }
                }
                return;
            }
            case Sec::Bounds: {
                if (md.col_lo.empty()) {
                    md.col_lo.assign(md.col_names.size(), 0);
                    md.col_up.assign(md.col_names.size(), kInf);
                }
                std::string_view type = t.empty() ? std::string_view() : t[0];
                const bool valueless = type == "FR" || type == "MI" || type == "PL" || type == "BV";
                std::string_view set, col, value;
                if (fixed) {
                    if (t.size() > 1) set = t[1];
                    if (t.size() > 2) col = t[2];
                    if (t.size() > 3) value = t[3];
                } else if (t.size() == 4) {
                    set = t[1], col = t[2], value = t[3];
                } else if (t.size() == 3) {
                    if (valueless) set = t[1], col = t[2];
                    else col = t[1], value = t[2];
                } else if (t.size() == 2) {
                    col = t[1];
                } else {
                    throw ParseError("bad BOUNDS line");
                }
                if (!set.empty() && !in_first_set(bound_set, set)) return;
                auto it = col_id.find(std::string(col));
                if (it == col_id.end()) throw ParseError("unknown column '" + std::string(col) + "' in BOUNDS");
                size_t j = size_t(it->second);
                double v = valueless || value.empty() ? 0 : parse_number(value);
                double &lo = md.col_lo[j], &up = md.col_up[j];
                if (type == "BV" || type == "LI" || type == "UI") md.is_int[j] = 1;
                if (type == "LO" || type == "LI") {
                    lo = v;
                } else if (type == "UP" || type == "UI") {
                    up = v;
                    if (v < 0 && lo == 0) lo = -kInf;
                } else if (type == "FX") {
                    lo = up = v;
                } else if (type == "FR") {
                    lo = -kInf, up = kInf;
                } else if (type == "MI") {
                    lo = -kInf;
                } else if (type == "PL") {
                    up = kInf;
                } else if (type == "BV") {
                    lo = 0, up = 1;
                } else {
                    throw ParseError("unsupported bound type " + std::string(type));
                }
                return;
            }
            case Sec::Quad: {
                if (t.size() != 3) throw ParseError("bad quadratic objective line");
                int a = column_index(t[0]), b = column_index(t[1]);
                double v = parse_number(t[2]);
                if (quad_full && a < b) return;  // both triangles listed: keep one
                quad[{std::max(a, b), std::min(a, b)}] += v;
                return;
            }
            default:
                return;  // data under NAME or before any section carries nothing
        }
    }

    Model run(std::string_view text) {
        std::vector<std::string_view> t;
        size_t pos = 0, lines = 0;
        while (pos < text.size() && !saw_end) {
            size_t nl = text.find('\n', pos);
            std::string_view line = text.substr(pos, nl == std::string_view::npos ? std::string_view::npos : nl - pos);
            pos = nl == std::string_view::npos ? text.size() : nl + 1;
            if ((++lines & 4095) == 0 && std::chrono::steady_clock::now() > deadline)
                throw ParseTimeLimit("parse exceeded the time limit");
            if (!line.empty() && line.back() == '\r') line.remove_suffix(1);
            if (line.find_first_not_of(" \t") == std::string_view::npos || line[0] == '*') continue;
            if (line[0] != ' ' && line[0] != '\t') {
                header(line);
                continue;
            }
            split(line, sec, fixed, t);
            data(t);
        }
        close_column();
        if (std::chrono::steady_clock::now() > deadline) throw ParseTimeLimit("parse exceeded the time limit");
        if (!saw_end) throw ParseError("missing ENDATA");
        for (const auto& [rc, v] : quad)
            if (v != 0) md.qobj.push_back({rc.first, rc.second, v});
        if (obj_row.empty()) throw ParseError("no objective row");
        const size_t m = md.row_names.size(), n = md.col_names.size();
        if (rhs.empty()) rhs.assign(m, 0), range.assign(m, 0), ranged.assign(m, 0);
        if (md.col_lo.empty()) md.col_lo.assign(n, 0), md.col_up.assign(n, kInf);
        md.row_lo.resize(m), md.row_up.resize(m);
        for (size_t i = 0; i < m; ++i) {  // row bounds from type, RHS and RANGES (src/mps.cpp:348-361)
            const double b = rhs[i], R = range[i];
            double &lo = md.row_lo[i], &up = md.row_up[i];
            if (row_kind[i] == 'E') {
                lo = up = b;
                if (ranged[i]) (R >= 0 ? up : lo) = b + R;
            } else if (row_kind[i] == 'L') {
                up = b, lo = ranged[i] ? b - std::fabs(R) : -kInf;
            } else {
                lo = b, up = ranged[i] ? b + std::fabs(R) : kInf;
            }
        }
        for (std::vector<double>* v : {&md.col_lo, &md.row_lo})
            for (double& x : *v)
                if (x <= -kSentinel) md.sentinel_bounds += std::isfinite(x) ? 1 : 0, x = -kInf;
        for (std::vector<double>* v : {&md.col_up, &md.row_up})
            for (double& x : *v)
                if (x >= kSentinel) md.sentinel_bounds += std::isfinite(x) ? 1 : 0, x = kInf;
        return std::move(md);
    }
};

Model parse_text(std::string_view text, bool fixed, std::chrono::steady_clock::time_point deadline) {
    Reader r;
    r.fixed = fixed, r.deadline = deadline;
    return r.run(text);
}

Model parse_both(std::string_view text, std::chrono::steady_clock::time_point deadline) {
    try {
        return parse_text(text, false, deadline);
    } catch (const ParseError& free_err) {
        try {
            return parse_text(text, true, deadline);
        } catch (const ParseError& fixed_err) {  // report both readings
            if (std::strcmp(free_err.what(), fixed_err.what()) == 0) throw;
            throw ParseError(std::string(free_err.what()) + " (free format); " + fixed_err.what() + " (fixed columns)");
        }
    }
}

}  // namespace

Model read_mps(const std::string& path, std::chrono::steady_clock::time_point deadline) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw ParseError("cannot open " + path);
    std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    return parse_both(text, deadline);
}

Model read_mps_text(const std::string& text, std::chrono::steady_clock::time_point deadline) {
    return parse_both(text, deadline);
}

const char* pdhg_input_error(const Model& model) {
    const bool quadratic = !model.qobj.empty(), integer = model.has_integers();
    if (quadratic && integer)
        return "PDHG supports continuous LPs only; quadratic objectives and integer variables are unsupported";
    if (quadratic) return "PDHG supports continuous LPs only; quadratic objectives are unsupported";
    if (integer)
        return "PDHG supports continuous LPs only; integer variables are unsupported (no LP relaxation is solved)";
    return nullptr;
}

bool pdhg_reject_unsupported_input(const Model& model) {
    const char* why = pdhg_input_error(model);
    if (!why) return false;
    std::printf("status unsupported_model\n");
    std::fflush(stdout);
    std::fprintf(stderr, "%s\n", why);
    return true;
}

}  // namespace tpdhg
