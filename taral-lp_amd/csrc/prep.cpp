// Host problem preparation (gpu/pdhg.cu:48-133), written for large inputs: flat arrays, two-pass
// counting transposes, no per-column containers after the model is flattened.
#include "prep.hpp"

#include "parallel.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace tpdhg {

void column_order(const Csr& a, int cols, std::vector<int32_t>& cptr, std::vector<int32_t>& cperm,
                  std::vector<int32_t>* crow) {
    const int64_t nnz = a.nnz();
    cptr.assign(size_t(cols) + 1, 0), cperm.resize(size_t(nnz));
    if (crow) crow->resize(size_t(nnz));
    // row chunks of about equal nonzeros; chunk c places its entries of column j after those of chunks < c, so
    // every column lists its entries in row order, as one sequential counting sort would
    const int nc = nnz < (int64_t(1) << 20) || a.rows < 16 ? 1 : 16;
    std::vector<int> r0(size_t(nc) + 1, a.rows);
    r0[0] = 0;
    for (int c = 1; c < nc; ++c)
        r0[size_t(c)] = int(std::lower_bound(a.ptr.begin(), a.ptr.end() - 1, int32_t(nnz * c / nc)) - a.ptr.begin());
    std::vector<int32_t> at(size_t(nc) * size_t(cols), 0);  // per chunk: counts, then next positions
    parallel_for(nc, 1, [&](int64_t c) {
        int32_t* cnt = at.data() + size_t(c) * size_t(cols);
        for (int32_t k = a.ptr[size_t(r0[size_t(c)])]; k < a.ptr[size_t(r0[size_t(c) + 1])]; ++k) cnt[a.idx[size_t(k)]]++;
    });
    for (int j = 0; j < cols; ++j) {
        int32_t tot = 0;
        for (int c = 0; c < nc; ++c) tot += at[size_t(c) * size_t(cols) + size_t(j)];
        cptr[size_t(j) + 1] = cptr[size_t(j)] + tot;
    }
    parallel_for(cols, 4096, [&](int64_t j) {
        int32_t off = cptr[size_t(j)];
        for (int c = 0; c < nc; ++c) {
            int32_t& x = at[size_t(c) * size_t(cols) + size_t(j)];
            const int32_t n_c = x;
            x = off, off += n_c;
        }
    });
    parallel_for(nc, 1, [&](int64_t c) {
        int32_t* next = at.data() + size_t(c) * size_t(cols);
        for (int i = r0[size_t(c)]; i < r0[size_t(c) + 1]; ++i)
            for (int32_t k = a.ptr[size_t(i)]; k < a.ptr[size_t(i) + 1]; ++k) {
                const int32_t p = next[a.idx[size_t(k)]]++;
                cperm[size_t(p)] = k;
                if (crow) (*crow)[size_t(p)] = i;
            }
    });
}

Csr transpose(const Csr& a, int cols) {
    Csr t;
    t.rows = cols;
    std::vector<int32_t> perm;
    column_order(a, cols, t.ptr, perm, &t.idx);
    t.val.resize(perm.size());
    parallel_for(int64_t(perm.size()), 1 << 16, [&](int64_t q) { t.val[size_t(q)] = a.val[size_t(perm[size_t(q)])]; });
    return t;
}

double row_bound_measure(double lo, double up) {
    const bool fl = std::isfinite(lo), fu = std::isfinite(up);
    if (fl && fu) return std::max(std::fabs(lo), std::fabs(up));
    if (fl) return std::fabs(lo);
    if (fu) return std::fabs(up);
    return 0;
}

void scale_vectors(Scaled& P, const double* cost, const double* col_lo, const double* col_up, const double* row_lo,
                   const double* row_up, double sign) {
    const int m = P.m, n = P.n;
    P.c.resize(size_t(n)), P.l.resize(size_t(n)), P.u.resize(size_t(n));
    P.rl.resize(size_t(m)), P.ru.resize(size_t(m));
    double cn = 0, bn = 0;
    for (int j = 0; j < n; ++j) {
        const double c = sign * cost[j], d = P.Dc[size_t(j)];
        cn += c * c;
        P.c[size_t(j)] = c * d, P.l[size_t(j)] = col_lo[j] / d, P.u[size_t(j)] = col_up[j] / d;
    }
    for (int i = 0; i < m; ++i) {
        const double b = row_bound_measure(row_lo[i], row_up[i]);
        bn += b * b;
        P.rl[size_t(i)] = row_lo[i] * P.Dr[size_t(i)], P.ru[size_t(i)] = row_up[i] * P.Dr[size_t(i)];
    }
    P.bnorm = std::sqrt(bn), P.cnorm = std::sqrt(cn);
}

namespace {

// Divides A by sqrt of the row and column measures and accumulates the factors (gpu/pdhg.cu:95-102).
void divide_by_measures(Scaled& P, std::vector<double>& rs, std::vector<double>& cs) {
    for (double& v : rs) v = v > 0 ? std::sqrt(v) : 1;
    for (double& v : cs) v = v > 0 ? std::sqrt(v) : 1;
    Csr& A = P.A;
    parallel_for(P.m, 4096, [&](int64_t i) {
        const double r = rs[size_t(i)];
        for (int32_t k = A.ptr[size_t(i)]; k < A.ptr[size_t(i) + 1]; ++k)
            A.val[size_t(k)] /= r * cs[size_t(A.idx[size_t(k)])];
    });
    for (int i = 0; i < P.m; ++i) P.Dr[size_t(i)] /= rs[size_t(i)];
    for (int j = 0; j < P.n; ++j) P.Dc[size_t(j)] /= cs[size_t(j)];
}

void finish_scaling(Scaled& P, const double* cost, const double* col_lo, const double* col_up, const double* row_lo,
                    const double* row_up, double sign) {
    const int m = P.m, n = P.n;
    P.Dr.assign(size_t(m), 1.0), P.Dc.assign(size_t(n), 1.0);
    std::vector<double> rs(static_cast<size_t>(m)), cs(static_cast<size_t>(n));
    Csr& A = P.A;
    // Column view of A: position q of column j is the CSR entry cperm[q], rows ascending, so a per-column pass
    // meets a column's entries in the order a row-order scatter would (each row and each column is reduced by one
    // thread, and the measures equal the sequential ones bit for bit).
    std::vector<int32_t> cptr, cperm;
    column_order(A, n, cptr, cperm, nullptr);
    auto measures = [&](bool l1) {  // row and column infinity norms (Ruiz) or l1 norms (Pock-Chambolle)
        parallel_for(m, 4096, [&](int64_t i) {
            double r = 0;
            for (int32_t k = A.ptr[size_t(i)]; k < A.ptr[size_t(i) + 1]; ++k) {
                const double a = std::fabs(A.val[size_t(k)]);
                r = l1 ? r + a : std::max(r, a);
            }
            rs[size_t(i)] = r;
        });
        parallel_for(n, 4096, [&](int64_t j) {
            double c = 0;
            for (int32_t q = cptr[size_t(j)]; q < cptr[size_t(j) + 1]; ++q) {
                const double a = std::fabs(A.val[size_t(cperm[size_t(q)])]);
                c = l1 ? c + a : std::max(c, a);
            }
            cs[size_t(j)] = c;
        });
    };
    for (int pass = 0; pass < 10; ++pass) {  // Ruiz, infinity norm
        measures(false);
        divide_by_measures(P, rs, cs);
    }
    measures(true);  // Pock-Chambolle, alpha = 1: l1 norms, each summed in row order
    divide_by_measures(P, rs, cs);
    P.AT = transpose(A, n);
    scale_vectors(P, cost, col_lo, col_up, row_lo, row_up, sign);
}

}  // namespace

Scaled build_scaled_csc(int m, int n, const int64_t* colptr, const int32_t* rowidx, const double* val,
                        const double* cost, const double* col_lo, const double* col_up, const double* row_lo,
                        const double* row_up, double sign) {
    Scaled P;
    P.m = m, P.n = n;
    Csr T;  // A^T (columns of the model as rows), explicit zeros left out
    T.rows = n;
    T.ptr.assign(size_t(n) + 1, 0);
    int64_t keep = 0;
    for (int64_t k = 0; k < colptr[n]; ++k) keep += val[k] != 0;
    if (keep > kMaxNnz)
        throw std::length_error("the constraint matrix has " + std::to_string(keep) + " nonzeros; at most " +
                                std::to_string(kMaxNnz) + " are supported (32-bit nonzero indices)");
    T.idx.reserve(size_t(keep)), T.val.reserve(size_t(keep));
    for (int j = 0; j < n; ++j) {
        for (int64_t k = colptr[j]; k < colptr[j + 1]; ++k)
            if (val[k] != 0) T.idx.push_back(rowidx[k]), T.val.push_back(val[k]);
        T.ptr[size_t(j) + 1] = int32_t(T.idx.size());
    }
    P.A = transpose(T, m);
    finish_scaling(P, cost, col_lo, col_up, row_lo, row_up, sign);
    return P;
}

Scaled build_scaled(const Model& md, double sign) {
    const int n = int(md.cols.size()), m = int(md.row_lo.size());
    std::vector<int64_t> colptr(static_cast<size_t>(n) + 1, 0);
    for (int j = 0; j < n; ++j) colptr[size_t(j) + 1] = colptr[size_t(j)] + int64_t(md.cols[size_t(j)].size());
    const size_t nz = static_cast<size_t>(colptr[size_t(n)]);
    std::vector<int32_t> rowidx(nz);
    std::vector<double> val(nz);
    for (int j = 0; j < n; ++j) {
        size_t k = size_t(colptr[size_t(j)]);
        for (const Entry& e : md.cols[size_t(j)]) rowidx[k] = e.index, val[k] = e.value, ++k;
    }
    return build_scaled_csc(m, n, colptr.data(), rowidx.data(), val.data(), md.cost.data(), md.col_lo.data(),
                            md.col_up.data(), md.row_lo.data(), md.row_up.data(), sign);
}

}  // namespace tpdhg
