/* CPU oracle for the restarted PDHG LP path: TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's host preparation and CPU backend, used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the oracle. No product code links or
 * loads it. Each function names the reference lines it follows (paths relative to the reference root).
 *
 * Pinning (which check fixes which function) is listed at each declaration; a function with no such check
 * would say "parity unpinned" here, and none does.
 */
#ifndef TARAL_PDHG_ORACLE_H
#define TARAL_PDHG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* KKT scalar slots: gpu/pdhg.cu:136 (row part) and :145 (column part). out[0..4] row, out[5..9] column. */
enum { OR_PR2S, OR_PR2O, OR_PR2R, OR_DROW, OR_DZ2, OR_DR2S = 5, OR_DR2O, OR_POBJ, OR_DCOL, OR_DX2 };

typedef struct {
    int m, n;
    int64_t nnz;
    int32_t *aptr, *aidx; double *aval;   /* A_s, CSR over m rows */
    int32_t *tptr, *tidx; double *tval;   /* A_s^T, CSR over n rows */
    double *c, *l, *u, *rl, *ru, *Dr, *Dc;
    double bnorm, cnorm;
} or_scaled;

/* build_scaled, gpu/pdhg.cu:80-133 (with transpose :55-70). Model given column-wise (CSC, rows ascending per
 * column) as src/taral.hpp:27 stores it. Pinned by: closed-form identity A_s = diag(Dr) A diag(Dc) and the
 * Ruiz/Pock-Chambolle invariants on random matrices, plus the ledger trajectories through or_solve. */
int or_build_scaled(int m, int n, const int64_t* colptr, const int32_t* rowidx, const double* val,
                    const double* cost, const double* col_lo, const double* col_up, const double* row_lo,
                    const double* row_up, double sign, or_scaled* out);
void or_free_scaled(or_scaled* s);

/* y = M v over a CSR (double), the CPU dot of gpu/pdhg.cu:625-634. Pinned by brute-force dense products. */
void or_spmv(int rows, const int32_t* ptr, const int32_t* idx, const double* val, const double* v, double* y);

/* KKT scalars of one point, gpu/pdhg.cu:692-730 with one thread. fp32 != 0 rounds every vector entry to float
 * first, as Cpu<float> holds them. Pinned by brute-force sums on tiny cases and by the ledger trajectories. */
void or_eval(const or_scaled* P, int fp32, double f, const double* Ax, const double* z, const double* x,
             const double* ATz, const double* xlast, const double* zlast, double out[10]);

/* Power iteration estimate of ||A_s||_2, gpu/pdhg.cu:635-654. Pinned against numpy's dense SVD. */
double or_power_norm(const or_scaled* P, int fp32, int iters);

typedef struct {
    int status;          /* 0 near_optimal, 1 iteration_limit, 2 time_limit, 3 numerical_failure */
    long iterations;
    int restarts;
    int from_average;
    double rel_primal, rel_dual, rel_gap, norm_a, omega;
    double power_s, iterate_s;
} or_result;

/* The restarted averaged PDHG driver, gpu/pdhg.cu:761-849, over the single-thread CPU backend :583-741.
 * x_out (n) and z_out (m) receive the scaled-space point (may be NULL). x0/z0 (scaled space, may be NULL)
 * replace the cold start x0 = clamp(0, l, u), z0 = 0 of :660-663 (warm-start extension; NULL = reference).
 * Pinned by the reference ledgers: packing_10000 and transport_100x100 at 1e-4 and 1e-6 (iterations,
 * restarts, objective), gpu/gen_lp.py instances regenerated bit for bit. */
int or_solve(const or_scaled* P, int fp32, double tol, long max_iter, int eval_freq, double row_rel,
             double time_limit, const double* x0, const double* z0, double* x_out, double* z_out,
             or_result* res);

/* One iteration block of the CPU backend without the driver: set_steps(tau, sigma), start(x0, z0) (NULL = cold
 * start), iterate(steps) (gpu/pdhg.cu:655-691). Writes the state x, z, xs, zs, Ax, ATz, Axs, ATzs as doubles
 * (any may be NULL). Pinned through or_solve (same functions) and the ledger trajectories. */
void or_block(const or_scaled* P, int fp32, double tau, double sigma, const double* x0, const double* z0, int steps,
              double* x, double* z, double* xs, double* zs, double* Ax, double* ATz, double* Axs, double* ATzs);

/* or_block with the iterations on `threads` threads partitioned as the CPU backend's Pool (gpu/pdhg.cu:601-614,
 * 669-691); the result equals or_block's bit for bit (no phase reduces across rows). Used to time the CPU
 * baseline on all cores; pinned against or_block. */
void or_block_mt(const or_scaled* P, int fp32, double tau, double sigma, const double* x0, const double* z0,
                 int steps, int threads, double* x, double* z, double* xs, double* zs, double* Ax, double* ATz,
                 double* Axs, double* ATzs);

/* Objective on the ORIGINAL model in long double and max row/bound violation, gpu/pdhg.cu:866-876.
 * x is the scaled point; it is unscaled with Dc here. Pinned by the ledgers' objective and max_row_violation. */
void or_finish(const or_scaled* P, int n, const int64_t* colptr, const int32_t* rowidx, const double* val,
               const double* cost, double obj_const, const double* col_lo, const double* col_up,
               const double* row_lo, const double* row_up, const double* xs, double* objective,
               double* max_row_viol, double* max_bound_viol);

#ifdef __cplusplus
}
#endif
#endif
