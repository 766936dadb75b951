// LP model and MPS input contract of the PDHG path.
//
// Same data model and semantics as the reference's Model / read_mps (src/taral.hpp:12-48, src/mps.cpp) and
// PDHG input gate (gpu/pdhg_input.hpp:7-26), written independently: a user's MPS file yields the same
// column-wise model, and the same files are rejected.
#pragma once
#include <chrono>
#include <cstddef>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

namespace tpdhg {

constexpr double kInf = std::numeric_limits<double>::infinity();

struct Entry {  // one nonzero of a column: row index and value (src/taral.hpp:12-15)
    int index;
    double value;
};

struct QEntry {  // one quadratic objective term, row >= col (src/taral.hpp:17-20)
    int row, col;
    double value;
};

// min (or max) cost'x + obj_const  s.t.  row_lo <= A x <= row_up,  col_lo <= x <= col_up (src/taral.hpp:22-35).
struct Model {
    std::string name;
    std::vector<std::string> row_names, col_names;
    std::vector<std::vector<Entry>> cols;  // column-wise A, rows ascending, no zeros
    std::vector<double> cost, col_lo, col_up, row_lo, row_up;
    double obj_const = 0;        // minus the RHS given on the objective row
    bool maximize = false;       // OBJSENSE MAX
    size_t sentinel_bounds = 0;  // finite |bound| >= 1e20 turned into infinity
    std::vector<char> is_int;    // INTORG..INTEND columns and BV/LI/UI bounds
    std::vector<QEntry> qobj;    // QUADOBJ / QMATRIX / QSECTION terms
    bool has_integers() const;
    size_t nnz() const;
};

struct ParseError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Raised when the deadline passes while reading; deliberately not a ParseError so that no fixed-column
// retry follows a timeout (src/taral.hpp:41-45).
struct ParseTimeLimit : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Reads an MPS file: whitespace-token (free) reading first, fixed columns when that fails
// (src/mps.cpp:388-399). Throws ParseError / ParseTimeLimit.
Model read_mps(const std::string& path,
               std::chrono::steady_clock::time_point deadline = std::chrono::steady_clock::time_point::max());

// Same as read_mps over text already in memory (used by the C ABI for in-memory models).
Model read_mps_text(const std::string& text,
                    std::chrono::steady_clock::time_point deadline = std::chrono::steady_clock::time_point::max());

// PDHG solves continuous LPs only: reason string for a model it must refuse, nullptr otherwise
// (gpu/pdhg_input.hpp:7-17, same messages).
const char* pdhg_input_error(const Model& model);

// Prints "status unsupported_model" to stdout and the reason to stderr when the model is refused; callers
// exit with status 2 before touching the device or any output file (gpu/pdhg_input.hpp:19-26).
bool pdhg_reject_unsupported_input(const Model& model);

}  // namespace tpdhg
