// Host-side problem preparation for the PDHG path: CSR storage, transpose, and the scaled problem
// (Ruiz + Pock-Chambolle), FP64. Semantics of gpu/pdhg.cu:48-133.
#pragma once
#include <cstdint>
#include <vector>

#include "model.hpp"

namespace tpdhg {

// Most nonzeros one matrix (or one batch pool) may hold: nonzeros are indexed with 32-bit integers on the host and
// the device, and the device's staging loops form indices up to 2^16 past a block's last nonzero before clamping.
constexpr int64_t kMaxNnz = int64_t(INT32_MAX) - (int64_t(1) << 16);

struct Csr {  // rows x cols, column indices ascending within each row
    int rows = 0;
    std::vector<int32_t> ptr, idx;
    std::vector<double> val;
    int64_t nnz() const { return int64_t(val.size()); }
};

// Column order of a's entries: column j holds positions [cptr[j], cptr[j+1]), position q is the CSR entry
// cperm[q] of row (*crow)[q], rows ascending within a column (a counting sort, done in row chunks on the host's
// cores with the sequential sort's result).
void column_order(const Csr& a, int cols, std::vector<int32_t>& cptr, std::vector<int32_t>& cperm,
                  std::vector<int32_t>* crow);

// Counting-sort transpose; `cols` is the column count of `a` (gpu/pdhg.cu:55-70).
Csr transpose(const Csr& a, int cols);

// Scaled problem: A_s = diag(Dr) A diag(Dc); x = Dc x_s; z = Dr z_s (gpu/pdhg.cu:72-78).
struct Scaled {
    int m = 0, n = 0;
    Csr A, AT;
    std::vector<double> c, l, u, rl, ru, Dr, Dc;
    double bnorm = 0, cnorm = 0;  // original-space norms for the relative stopping test
};

// Ruiz (10 passes, inf-norm) then Pock-Chambolle (alpha = 1) equilibration of the model's A, and the scaled
// cost, bounds and norms; sign = -1 turns a maximisation into a minimisation (gpu/pdhg.cu:80-133).
Scaled build_scaled(const Model& md, double sign);

// Same from a column-wise model given as arrays (CSC, rows ascending per column).
Scaled build_scaled_csc(int m, int n, const int64_t* colptr, const int32_t* rowidx, const double* val,
                        const double* cost, const double* col_lo, const double* col_up, const double* row_lo,
                        const double* row_up, double sign);

// Scaled cost, bounds and original-space norms from the model vectors with the scaling factors P.Dr, P.Dc
// already set (gpu/pdhg.cu:117-131); used again when a scenario changes only the vectors.
void scale_vectors(Scaled& P, const double* cost, const double* col_lo, const double* col_up, const double* row_lo,
                   const double* row_up, double sign);

// The reference's row measure b_i: the larger finite |row bound|, 0 when both are infinite.
double row_bound_measure(double lo, double up);

}  // namespace tpdhg
