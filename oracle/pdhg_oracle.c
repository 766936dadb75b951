/* CPU oracle for the restarted PDHG LP path: TEST INFRASTRUCTURE ONLY (see pdhg_oracle.h).
 * Plain C99, single thread, built with -ffp-contract=off so every product and sum rounds on its own. */
#include "pdhg_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double or_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Progress trace to stderr every N evaluations (0 = off), from the ORACLE_TRACE environment variable. */
static int or_trace_every = -1;
static void or_trace_init(void) {
    if (or_trace_every < 0) {
        const char* e = getenv("ORACLE_TRACE");
        or_trace_every = e ? atoi(e) : 0;
    }
}

/* Deterministic power-iteration start, gpu/pdhg.cu:153 */
static double or_start_vec(int j) { return 0.5 + (double)(((uint32_t)j * 2654435761u) % 1000u) / 1000.0; }

/* 1/(1+|b_i|) in original row units, gpu/pdhg.cu:139-144 */
static double or_row_rel_weight(double lo, double up, double dr) {
    double b = 0;
    if (lo > -INFINITY) b = lo < 0 ? -lo : lo;
    if (up < INFINITY) {
        double a = up < 0 ? -up : up;
        b = a > b ? a : b;
    }
    return 1.0 / (1.0 + b / dr);
}

/* Counting-sort transpose of a CSR with `rows` rows into one with `cols` rows, gpu/pdhg.cu:55-70 */
static void or_transpose(int rows, const int32_t* ptr, const int32_t* idx, const double* val, int cols,
                         int32_t* tptr, int32_t* tidx, double* tval) {
    int64_t nnz = ptr[rows];
    memset(tptr, 0, sizeof(int32_t) * (size_t)(cols + 1));
    for (int64_t k = 0; k < nnz; ++k) tptr[idx[k] + 1]++;
    for (int j = 0; j < cols; ++j) tptr[j + 1] += tptr[j];
    int32_t* pos = (int32_t*)malloc(sizeof(int32_t) * (size_t)(cols > 0 ? cols : 1));
    memcpy(pos, tptr, sizeof(int32_t) * (size_t)cols);
    for (int i = 0; i < rows; ++i)
        for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) {
            int32_t p = pos[idx[k]]++;
            tidx[p] = i, tval[p] = val[k];
        }
    free(pos);
}

int or_build_scaled(int m, int n, const int64_t* colptr, const int32_t* rowidx, const double* val,
                    const double* cost, const double* col_lo, const double* col_up, const double* row_lo,
                    const double* row_up, double sign, or_scaled* P) {
    memset(P, 0, sizeof(*P));
    P->m = m, P->n = n;
    /* A^T straight from the column-wise model, zeros skipped (gpu/pdhg.cu:83-90) */
    int64_t nz = 0;
    for (int64_t k = 0; k < colptr[n]; ++k) nz += val[k] != 0;
    P->nnz = nz;
    size_t zb = (size_t)(nz > 0 ? nz : 1);
    int32_t* T_ptr = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
    int32_t* T_idx = (int32_t*)malloc(sizeof(int32_t) * zb);
    double* T_val = (double*)malloc(sizeof(double) * zb);
    T_ptr[0] = 0;
    int64_t q = 0;
    for (int j = 0; j < n; ++j) {
        for (int64_t k = colptr[j]; k < colptr[j + 1]; ++k)
            if (val[k] != 0) T_idx[q] = rowidx[k], T_val[q] = val[k], ++q;
        T_ptr[j + 1] = (int32_t)q;
    }
    P->aptr = (int32_t*)malloc(sizeof(int32_t) * (size_t)(m + 1));
    P->aidx = (int32_t*)malloc(sizeof(int32_t) * zb);
    P->aval = (double*)malloc(sizeof(double) * zb);
    or_transpose(n, T_ptr, T_idx, T_val, m, P->aptr, P->aidx, P->aval);
    free(T_ptr), free(T_idx), free(T_val);

    int32_t *ap = P->aptr, *ai = P->aidx;
    double* av = P->aval;
    P->Dr = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    P->Dc = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* rs = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    double* cs = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < m; ++i) P->Dr[i] = 1;
    for (int j = 0; j < n; ++j) P->Dc[j] = 1;
    /* apply(): divide by sqrt of the row and column measures (gpu/pdhg.cu:95-102) */
#define OR_APPLY()                                                                          \
    do {                                                                                    \
        for (int i = 0; i < m; ++i) rs[i] = rs[i] > 0 ? sqrt(rs[i]) : 1;                   \
        for (int j = 0; j < n; ++j) cs[j] = cs[j] > 0 ? sqrt(cs[j]) : 1;                   \
        for (int i = 0; i < m; ++i)                                                         \
            for (int32_t k = ap[i]; k < ap[i + 1]; ++k) av[k] /= rs[i] * cs[ai[k]];         \
        for (int i = 0; i < m; ++i) P->Dr[i] /= rs[i];                                      \
        for (int j = 0; j < n; ++j) P->Dc[j] /= cs[j];                                      \
    } while (0)
    for (int pass = 0; pass < 10; ++pass) { /* Ruiz, inf-norm (gpu/pdhg.cu:103-111) */
        memset(rs, 0, sizeof(double) * (size_t)m), memset(cs, 0, sizeof(double) * (size_t)n);
        for (int i = 0; i < m; ++i)
            for (int32_t k = ap[i]; k < ap[i + 1]; ++k) {
                double a = fabs(av[k]);
                if (a > rs[i]) rs[i] = a;
                if (a > cs[ai[k]]) cs[ai[k]] = a;
            }
        OR_APPLY();
    }
    memset(rs, 0, sizeof(double) * (size_t)m), memset(cs, 0, sizeof(double) * (size_t)n);
    for (int i = 0; i < m; ++i) /* Pock-Chambolle, alpha = 1 (gpu/pdhg.cu:112-115) */
        for (int32_t k = ap[i]; k < ap[i + 1]; ++k) rs[i] += fabs(av[k]), cs[ai[k]] += fabs(av[k]);
    OR_APPLY();
#undef OR_APPLY
    free(rs), free(cs);
    P->tptr = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
    P->tidx = (int32_t*)malloc(sizeof(int32_t) * zb);
    P->tval = (double*)malloc(sizeof(double) * zb);
    or_transpose(m, ap, ai, av, n, P->tptr, P->tidx, P->tval);

    P->c = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    P->l = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    P->u = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    P->rl = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    P->ru = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    for (int j = 0; j < n; ++j) { /* gpu/pdhg.cu:118-122 */
        double c = sign * cost[j];
        P->cnorm += c * c;
        P->c[j] = c * P->Dc[j], P->l[j] = col_lo[j] / P->Dc[j], P->u[j] = col_up[j] / P->Dc[j];
    }
    for (int i = 0; i < m; ++i) { /* gpu/pdhg.cu:123-130 */
        double lo = row_lo[i], up = row_up[i], b = 0;
        if (isfinite(lo) && isfinite(up)) b = fmax(fabs(lo), fabs(up));
        else if (isfinite(lo)) b = fabs(lo);
        else if (isfinite(up)) b = fabs(up);
        P->bnorm += b * b;
        P->rl[i] = lo * P->Dr[i], P->ru[i] = up * P->Dr[i];
    }
    P->bnorm = sqrt(P->bnorm), P->cnorm = sqrt(P->cnorm);
    return 0;
}

void or_free_scaled(or_scaled* s) {
    void* all[] = {s->aptr, s->aidx, s->aval, s->tptr, s->tidx, s->tval, s->c, s->l, s->u, s->rl, s->ru, s->Dr, s->Dc};
    for (size_t i = 0; i < sizeof(all) / sizeof(all[0]); ++i) free(all[i]);
    memset(s, 0, sizeof(*s));
}

void or_spmv(int rows, const int32_t* ptr, const int32_t* idx, const double* val, const double* v, double* y) {
    for (int r = 0; r < rows; ++r) {
        double s = 0;
        for (int32_t k = ptr[r]; k < ptr[r + 1]; ++k) s += val[k] * v[idx[k]];
        y[r] = s;
    }
}

/* Thread boundaries of the CPU backend's matrices (gpu/pdhg.cu:601-614): part[0] = 0, part[nt] = rows, and
 * part[t] is the first row whose prefix rows + nonzeros reaches t/nt of the total. */
static void or_partition(int rows, const int32_t* ptr, int nt, int* part) {
    long total = (long)ptr[rows] + rows;
    part[0] = 0;
    for (int t = 1; t < nt; ++t) {
        long target = total * t / nt;
        int lo = 0, hi = rows;
        while (lo < hi) {
            int mid = lo + (hi - lo) / 2;
            if ((long)ptr[mid] + mid < target)
                lo = mid + 1;
            else
                hi = mid;
        }
        part[t] = lo;
    }
    part[nt] = rows;
}

#define OT double
#define OS _d
#include "pdhg_oracle_t.inc"
#undef OT
#undef OS
#define OT float
#define OS _f
#include "pdhg_oracle_t.inc"
#undef OT
#undef OS

void or_eval(const or_scaled* P, int fp32, double f, const double* Ax, const double* z, const double* x,
             const double* ATz, const double* xlast, const double* zlast, double out[10]) {
    int m = P->m, n = P->n;
    if (fp32) {
        or_state_f S;
        memset(&S, 0, sizeof(S));
        S.m = m, S.n = n;
        S.c = or_cvt_f(P->c, n), S.l = or_cvt_f(P->l, n), S.u = or_cvt_f(P->u, n);
        S.rl = or_cvt_f(P->rl, m), S.ru = or_cvt_f(P->ru, m), S.Dr = or_cvt_f(P->Dr, m), S.Dc = or_cvt_f(P->Dc, n);
        S.xlast = or_cvt_f(xlast, n), S.zlast = or_cvt_f(zlast, m);
        float *a = or_cvt_f(Ax, m), *b = or_cvt_f(z, m), *c = or_cvt_f(x, n), *d = or_cvt_f(ATz, n);
        or_eval_t_f(&S, a, b, c, d, f, out);
        float* all[] = {S.c, S.l, S.u, S.rl, S.ru, S.Dr, S.Dc, S.xlast, S.zlast, a, b, c, d};
        for (size_t i = 0; i < sizeof(all) / sizeof(all[0]); ++i) free(all[i]);
    } else {
        or_state_d S;
        memset(&S, 0, sizeof(S));
        S.m = m, S.n = n;
        S.c = P->c, S.l = P->l, S.u = P->u, S.rl = P->rl, S.ru = P->ru, S.Dr = P->Dr, S.Dc = P->Dc;
        S.xlast = (double*)xlast, S.zlast = (double*)zlast;
        or_eval_t_d(&S, Ax, z, x, ATz, f, out);
    }
}

double or_power_norm(const or_scaled* P, int fp32, int iters) {
    double r;
    if (fp32) {
        or_state_f S;
        or_state_init_f(&S, P);
        r = or_power_t_f(&S, iters);
        or_state_free_f(&S);
    } else {
        or_state_d S;
        or_state_init_d(&S, P);
        r = or_power_t_d(&S, iters);
        or_state_free_d(&S);
    }
    return r;
}

int or_solve(const or_scaled* P, int fp32, double tol, long max_iter, int eval_freq, double row_rel,
             double time_limit, const double* x0, const double* z0, double* x_out, double* z_out,
             or_result* res) {
    if (eval_freq < 1) eval_freq = 1;
    or_trace_init();
    return fp32 ? or_run_f(P, tol, max_iter, eval_freq, row_rel, time_limit, x0, z0, x_out, z_out, res)
                : or_run_d(P, tol, max_iter, eval_freq, row_rel, time_limit, x0, z0, x_out, z_out, res);
}

#define OR_BLOCK_BODY(SUF)                                                                               \
    do {                                                                                                 \
        or_state##SUF S;                                                                                 \
        or_state_init##SUF(&S, P);                                                                       \
        S.tau = (OR_T)tau, S.sigma = (OR_T)sigma;                                                        \
        or_start##SUF(&S, P, x0, z0);                                                                    \
        if (steps > 0) or_iterate_mt##SUF(&S, steps, threads);                                           \
        const OR_T* src[8] = {S.x, S.z, S.xs, S.zs, S.Ax, S.ATz, S.Axs, S.ATzs};                         \
        double* dst[8] = {x, z, xs, zs, Ax, ATz, Axs, ATzs};                                             \
        for (int q = 0; q < 8; ++q) {                                                                    \
            int k = (q == 1 || q == 3 || q == 4 || q == 6) ? P->m : P->n;                                \
            for (int i = 0; dst[q] && i < k; ++i) dst[q][i] = (double)src[q][i];                         \
        }                                                                                                \
        or_state_free##SUF(&S);                                                                          \
    } while (0)

void or_block(const or_scaled* P, int fp32, double tau, double sigma, const double* x0, const double* z0, int steps,
              double* x, double* z, double* xs, double* zs, double* Ax, double* ATz, double* Axs, double* ATzs) {
    or_block_mt(P, fp32, tau, sigma, x0, z0, steps, 1, x, z, xs, zs, Ax, ATz, Axs, ATzs);
}

void or_block_mt(const or_scaled* P, int fp32, double tau, double sigma, const double* x0, const double* z0,
                 int steps, int threads, double* x, double* z, double* xs, double* zs, double* Ax, double* ATz,
                 double* Axs, double* ATzs) {
    if (threads < 1) threads = 1;
    if (fp32) {
#define OR_T float
        OR_BLOCK_BODY(_f);
#undef OR_T
    } else {
#define OR_T double
        OR_BLOCK_BODY(_d);
#undef OR_T
    }
}

void or_finish(const or_scaled* P, int n, const int64_t* colptr, const int32_t* rowidx, const double* val,
               const double* cost, double obj_const, const double* col_lo, const double* col_up,
               const double* row_lo, const double* row_up, const double* xs, double* objective,
               double* max_row_viol, double* max_bound_viol) {
    int m = P->m;
    double* x = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int j = 0; j < n; ++j) x[j] = xs[j] * P->Dc[j];
    long double obj = obj_const;
    for (int j = 0; j < n; ++j) obj += (long double)cost[j] * x[j];
    double* ax = (double*)calloc((size_t)(m > 0 ? m : 1), sizeof(double));
    double bviol = 0, rviol = 0;
    for (int j = 0; j < n; ++j) {
        for (int64_t k = colptr[j]; k < colptr[j + 1]; ++k) ax[rowidx[k]] += val[k] * x[j];
        bviol = fmax(bviol, fmax(col_lo[j] - x[j], x[j] - col_up[j]));
    }
    for (int i = 0; i < m; ++i) rviol = fmax(rviol, fmax(row_lo[i] - ax[i], ax[i] - row_up[i]));
    *objective = (double)obj, *max_row_viol = rviol, *max_bound_viol = bviol;
    free(x), free(ax);
}
