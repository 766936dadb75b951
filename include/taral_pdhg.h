/* taral_pdhg.h — C ABI of the MI355X restarted-PDHG LP solver (libtaral_pdhg.so).
 *
 * Drop-in boundary for the reference's PDHG path (Hexraei/TARAL-LP, gpu/pdhg.cu + gpu/pdhg_input.hpp over the
 * Model/read_mps contract of src/taral.hpp and src/mps.cpp). Plain pointers and sizes only; every call
 * returns TPDHG_OK (0) or a negative code, with the message in tpdhg_last_error() (per thread). Nothing here
 * calls exit(): the reference's CHECK macro (gpu/pdhg.cu:36-43) exits the process, this boundary reports.
 * There is no CPU compute path: without a HIP device the device and solve calls fail with
 * TPDHG_ERR_NO_DEVICE.
 *
 * Handles are not thread-safe; one device handle owns one HIP stream and its device memory.
 */
#ifndef TARAL_PDHG_H
#define TARAL_PDHG_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TPDHG_OK = 0,
    TPDHG_ERR_ARG = -1,          /* bad argument or dimensions */
    TPDHG_ERR_PARSE = -2,        /* MPS parse error (reference: ParseError, pdhg.cu:977-980, exit 1) */
    TPDHG_ERR_UNSUPPORTED = -3,  /* QP or MILP input (reference: pdhg_reject_unsupported_input, exit 2) */
    TPDHG_ERR_DEVICE = -4,       /* HIP runtime failure (reference: CHECK -> exit 1) */
    TPDHG_ERR_NO_DEVICE = -5,    /* no HIP device visible */
    TPDHG_ERR_TIME = -6          /* parse deadline passed (reference: ParseTimeLimit) */
};

/* Solve status, the strings of gpu/pdhg.cu:753,814-826. */
enum { TPDHG_NEAR_OPTIMAL = 0, TPDHG_ITERATION_LIMIT = 1, TPDHG_TIME_LIMIT = 2, TPDHG_NUMERICAL_FAILURE = 3 };

const char* tpdhg_last_error(void);
const char* tpdhg_status_string(int status);
int tpdhg_device_count(void);

/* ------------------------------------------------------------------ model (input contract) */
typedef struct tpdhg_model tpdhg_model;

/* read_mps(path) (src/taral.hpp:48, src/mps.cpp:388-399): free format, then fixed columns. */
int tpdhg_model_read_mps(const char* path, tpdhg_model** out);
/* Same reader over MPS text in memory (len bytes). */
int tpdhg_model_read_mps_text(const char* text, int64_t len, tpdhg_model** out);
/* Column-wise LP from arrays (CSC; rows ascending within each column), the fields of Model
 * (src/taral.hpp:24-35) that the PDHG path reads. Bounds may be +-INFINITY. Arrays are copied. */
int tpdhg_model_from_csc(int32_t m, int32_t n, const int64_t* colptr, const int32_t* rowidx, const double* val,
                         const double* cost, const double* col_lo, const double* col_up, const double* row_lo,
                         const double* row_up, double obj_const, int32_t maximize, tpdhg_model** out);
void tpdhg_model_free(tpdhg_model* model);

typedef struct {
    int32_t m, n;
    int64_t nnz;
    int32_t maximize;
    int32_t has_integers;
    int64_t qobj_terms;
    int64_t sentinel_bounds;
    double obj_const;
} tpdhg_model_info;
int tpdhg_model_get_info(const tpdhg_model* model, tpdhg_model_info* info);
/* Copies the model out; any pointer may be NULL. Sizes: colptr n+1, rowidx/val nnz, cost/col_lo/col_up/is_int n,
 * row_lo/row_up m. */
int tpdhg_model_export(const tpdhg_model* model, int64_t* colptr, int32_t* rowidx, double* val, double* cost,
                       double* col_lo, double* col_up, double* row_lo, double* row_up, int8_t* is_int);
/* Name of row i (is_row != 0) or column i; "" for array-built models. */
const char* tpdhg_model_name_of(const tpdhg_model* model, int32_t is_row, int32_t i);
/* pdhg_input_error (gpu/pdhg_input.hpp:7-17): the reason PDHG refuses this model, or NULL. */
const char* tpdhg_model_input_error(const tpdhg_model* model);

/* ------------------------------------------------------------------ scaled problem (host, FP64) */
typedef struct tpdhg_scaled tpdhg_scaled;
typedef struct {
    int32_t m, n;
    int64_t nnz;
    const int32_t *a_ptr, *a_idx; /* A_s, CSR over m rows */
    const double* a_val;
    const int32_t *at_ptr, *at_idx; /* A_s^T, CSR over n rows */
    const double* at_val;
    const double *c, *l, *u, *rl, *ru, *Dr, *Dc;
    double bnorm, cnorm;
} tpdhg_scaled_arrays;

/* build_scaled(model, sign) (gpu/pdhg.cu:80-133): Ruiz x10 (inf-norm) then Pock-Chambolle (alpha 1). */
int tpdhg_scaled_build(const tpdhg_model* model, tpdhg_scaled** out);
/* Pointers into the handle's memory (valid until tpdhg_scaled_free). */
int tpdhg_scaled_view(const tpdhg_scaled* s, tpdhg_scaled_arrays* view);
void tpdhg_scaled_free(tpdhg_scaled* s);

/* ------------------------------------------------------------------ device backend (the Gpu<T> concept) */
typedef struct tpdhg_dev tpdhg_dev;

/* Gpu<T>::init (gpu/pdhg.cu:362-382): uploads A_s, A_s^T and the vectors; precision_bits 64 (FP64 iterates) or
 * 32 (FP32 iterates and SpMV, FP64 reductions; the reference's --fp32). */
int tpdhg_dev_create(const tpdhg_scaled_arrays* problem, int32_t precision_bits, int32_t eval_freq,
                     int32_t use_graph, tpdhg_dev** out);
/* Gpu<T>::power_norm (gpu/pdhg.cu:417-436). */
int tpdhg_dev_power_norm(tpdhg_dev* dev, int32_t iters, double* norm);
/* Gpu<T>::set_steps (gpu/pdhg.cu:442-445). */
int tpdhg_dev_set_steps(tpdhg_dev* dev, double tau, double sigma);
/* Gpu<T>::start (gpu/pdhg.cu:450-458); x0/z0 (scaled space) may be NULL for the reference cold start. */
int tpdhg_dev_start(tpdhg_dev* dev, const double* x0, const double* z0);
/* Gpu<T>::iterate (gpu/pdhg.cu:471-486). */
int tpdhg_dev_iterate(tpdhg_dev* dev, int32_t steps);
/* Gpu<T>::eval (gpu/pdhg.cu:487-499): out[0..4] = PR2S, PR2O, PR2R, DROW, DZ2; out[5..9] = DR2S, DR2O, POBJ,
 * DCOL, DX2. */
int tpdhg_dev_eval(tpdhg_dev* dev, int32_t avg, double f, double out[10]);
/* Current and f-scaled average KKT in one fused pass (no reference counterpart; same values as two evals). */
int tpdhg_dev_eval2(tpdhg_dev* dev, double f, double cur[10], double avg[10]);
/* Gpu<T>::restart (gpu/pdhg.cu:500-504). */
int tpdhg_dev_restart(tpdhg_dev* dev, int32_t avg, double f);
/* Gpu<T>::download (gpu/pdhg.cu:505-513): x (n) and z (m) in FP64, scaled space; either may be NULL. */
int tpdhg_dev_download(tpdhg_dev* dev, int32_t avg, double f, double* x, double* z);
/* y = A_s v (transpose 0; v n, y m) or A_s^T v (transpose 1; v m, y n) with the solver's SpMV kernel. */
int tpdhg_dev_spmv(tpdhg_dev* dev, int32_t transpose, const double* v, double* y);
/* Gpu<T>::mem_used and name (gpu/pdhg.cu:331-332); name_buf receives up to name_len-1 chars. */
int tpdhg_dev_info(tpdhg_dev* dev, int64_t* mem_used, char* name_buf, int32_t name_len);
/* Per-kernel HIP-event timing of the iteration kernels (on the solver stream) from now on. */
int tpdhg_dev_set_profile(tpdhg_dev* dev, int32_t on);
int tpdhg_dev_kernel_times(tpdhg_dev* dev, double* row_ms, double* col_ms, int64_t* row_launches,
                           int64_t* col_launches);
void tpdhg_dev_destroy(tpdhg_dev* dev);

/* ------------------------------------------------------------------ fast-path backend (no reference counterpart) */
/* Reflected Halpern PDHG on a problem recentred around an FP64 base point (x_b, z_b), scaled space: the inner LP
 * is  min (c + A^T z_b)' dx  s.t.  rl - A x_b <= A dx <= ru - A x_b,  l - x_b <= dx <= u - x_b. One iteration k
 * with anchors (x0, z0), lambda = (k+1)/(k+2) and reflection rho: x+ = proj(x - tau (c_in + A^T z)),
 * z+ = prox(z + sigma A (2 x+ - x)); x <- lambda (rho x+ - (rho-1) x) + (1-lambda) x0, and the same for z.
 * This is the engine of tpdhg_options.halpern; these calls expose it for tests and custom drivers. */
typedef struct tpdhg_fast tpdhg_fast;
int tpdhg_fast_create(const tpdhg_scaled_arrays* problem, int32_t precision_bits, int32_t eval_freq, int32_t use_graph,
                      double reflection, tpdhg_fast** out);
int tpdhg_fast_power_norm(tpdhg_fast* f, int32_t iters, double* norm);
/* Base point (n and m doubles, either may be NULL = 0); used from the next tpdhg_fast_begin_round. */
int tpdhg_fast_set_base(tpdhg_fast* f, const double* x_b, const double* z_b);
/* Inner problem from the base; inner iterates reset to dx = clamp(0, l_in, u_in), dz = 0. */
int tpdhg_fast_begin_round(tpdhg_fast* f);
int tpdhg_fast_set_steps(tpdhg_fast* f, double tau, double sigma);
/* Anchors = current inner point, iteration counter 0, first primal step. */
int tpdhg_fast_restart(tpdhg_fast* f);
int tpdhg_fast_iterate(tpdhg_fast* f, int32_t steps);
/* KKT scalars (slots as tpdhg_dev_eval) of base + inner on the original scaled problem, FP64; DX2/DZ2 measure
 * the distance from the anchors. */
int tpdhg_fast_eval(tpdhg_fast* f, double out[10]);
/* Fixed-point residual terms of the inner point u = (x, z) at the last tpdhg_fast_eval, FP64 sums: out[0] =
 * ||x - x+||^2, out[1] = ||z - z+||^2, out[2] = (z - z+)' A (x - x+), where (x+, z+) is one PDHG step from u:
 * x+ = proj(x - tau (c_in + A^T z)), z+ = prox(z + sigma A (2 x+ - x)). ||u - T u||_P^2 = out[0] / tau +
 * out[1] / sigma + 2 out[2] is what tpdhg_options.restart_metric = 1 measures. */
int tpdhg_fast_fixed_point(tpdhg_fast* f, double out[3]);
/* base += inner point. */
int tpdhg_fast_end_round(tpdhg_fast* f);
/* base + inner point (FP64, scaled space); either may be NULL. */
int tpdhg_fast_download(tpdhg_fast* f, double* x, double* z);
void tpdhg_fast_destroy(tpdhg_fast* f);

/* ------------------------------------------------------------------ full solve (the pdhg CLI path) */
typedef struct {
    double tol;        /* relative KKT tolerance (default 1e-4, gpu/pdhg.cu:745) */
    double time_limit; /* seconds (default 3600) */
    int64_t max_iter;  /* default 2e9 */
    int32_t eval_freq; /* iterations per evaluation (default 64) */
    double row_rel;    /* per-row relative primal test factor, 0 = off (gpu/pdhg.cu:748) */
    int32_t precision_bits; /* 64 (default) or 32 */
    int32_t refine;    /* 1: FP64 iterative refinement (needs halpern = 1): every restart recentres the FP32 inner
                          problem at the restart point with exact FP64 products */
    double inner_tol;  /* ignored (kept for the struct layout) */
    int32_t max_refinements; /* cap on the recentrings (default 1e9) */
    int32_t use_graph; /* capture each eval_freq-iteration block in a HIP graph (default 1) */
    int32_t verbose;
    int32_t profile_kernels; /* HIP-event timing of the row/column kernels */
    int32_t halpern;   /* 1: fast path, reflected Halpern PDHG on an FP64-recentred problem (no reference
                          counterpart; the reference algorithm is the default, halpern = 0) */
    double reflection; /* Halpern reflection coefficient in [1, 2) (default 1.9) */
    double primal_weight; /* initial primal weight of a warm-started solve, 0 = ||c||/||b|| as for a cold start
                             (no reference counterpart: the reference always cold-starts, gpu/pdhg.cu:450-458) */
    int32_t warm_fallback; /* 1: keep a warm start only when its relative KKT is below the cold start's */
    double progress_interval; /* > 0: one progress line on stderr every this many seconds of solving */
    int32_t omega_rule; /* fast path and batches: primal-weight guard after each restart. 0: the reference's
                           smoothed update only; 1: kept within two decades of the starting weight; 2: residual
                           balancing (no move against a 3x primal/gap imbalance); 3 (default): both; 4: both, the
                           two decades counted from the weight the first restart's update gives. The
                           reference scheme (halpern = 0) always uses the reference update */
    int32_t reuse_norm; /* 1: a session or batch runs the power iteration on its first solve only and reuses ||A_s||
                           afterwards (A_s never changes within one); 0 (default): every solve runs it, as each
                           reference CLI call does (gpu/pdhg.cu:763) */
    int32_t cpu_threads; /* -1 (default): the GPU. >= 0: the reference scheme on the CPU backend (the reference's
                            --device cpu --threads N, gpu/pdhg.cu:517-741, 913-934) with this many host threads,
                            0 = the reference's choice (1 below 200,000 nonzeros, else every hardware thread);
                            results do not depend on the thread count. halpern and refine must be 0 */
    int32_t restart_metric; /* fast path and batches: what the restart tests (sufficient decay 0.2, necessary decay
                               0.8 without progress, artificial at 36% of the iterations) measure. 1 (default): the
                               fixed-point residual ||u - T u||_P of the current point u = (x, z), T one PDHG step,
                               in the PDHG norm ||dx||^2 / tau + ||dz||^2 / sigma + 2 dz' A dx at the current weight;
                               0: the reference's weighted KKT error (gpu/pdhg.cu:783-786). The reference scheme
                               (halpern = 0) always uses the latter. No reference counterpart */
} tpdhg_options;
void tpdhg_options_default(tpdhg_options* o);

typedef struct {
    int32_t status; /* TPDHG_NEAR_OPTIMAL ... */
    int64_t iterations;
    int32_t restarts, refinements, from_average;
    double objective; /* original model, long double accumulation (gpu/pdhg.cu:868-869) */
    double rel_primal, rel_dual, rel_gap;
    double max_row_violation, max_bound_violation;
    double setup_s, parse_s, scale_s, device_init_s, transfer_s, solve_s, power_iter_s, total_s;
    double norm_a, primal_weight;
    int64_t device_mem_bytes;
    int32_t m, n;
    int64_t nnz;
    double row_kernel_ms, col_kernel_ms; /* HIP-event totals when profile_kernels */
    int64_t row_launches, col_launches;
    char device[64];
    int32_t warm_start_used;  /* the solve started from the caller's x0/y0 (0 after a fallback or a cold start) */
    double warm_kkt, cold_kkt; /* relative KKT of both starts when warm_fallback compared them, else 0 */
    int32_t threads;  /* host threads of the CPU backend; 0 for a GPU solve (the JSON's "backend" and "threads") */
} tpdhg_result;

/* Scales, uploads and solves (gpu/pdhg.cu:913-935 -> finish -> run). x (n) receives the original-space point
 * and y (m) the row multipliers y = Dr z (gpu/pdhg.cu:905-909); either may be NULL. x0/y0 (same spaces,
 * nullable) warm start the solve. Refuses QP/MILP models with TPDHG_ERR_UNSUPPORTED before any device work. */
int tpdhg_solve(const tpdhg_model* model, const tpdhg_options* opt, const double* x0, const double* y0,
                tpdhg_result* result, double* x, double* y);
/* A problem kept resident on the device: scaled and uploaded once by tpdhg_session_create, then solved from a
 * cold (x0 = y0 = NULL) or warm start by each tpdhg_session_solve, which repeats the power iteration and the
 * PDHG loop (no reference counterpart: the reference CLI re-reads and re-uploads per solve). The session keeps
 * a copy of the model. */
typedef struct tpdhg_session tpdhg_session;
int tpdhg_session_create(const tpdhg_model* model, const tpdhg_options* opt, tpdhg_session** out);
int tpdhg_session_solve(tpdhg_session* s, const double* x0, const double* y0, tpdhg_result* result, double* x,
                        double* y);
/* Replaces the cost and the row/column bounds (original space; n, n, n, m, m doubles) of a session's problem
 * while A stays resident: the vectors are rescaled with the session's Ruiz/Pock-Chambolle factors (which depend
 * on A only) and uploaded. For scenario families that share A (no reference counterpart). */
int tpdhg_session_set_vectors(tpdhg_session* s, const double* cost, const double* col_lo, const double* col_up,
                              const double* row_lo, const double* row_up);
/* Initial primal weight for the session's following warm-started solves (0 = ||c||/||b||); the learned warm
 * start predicts it with (x0, y0). No reference counterpart. */
int tpdhg_session_set_primal_weight(tpdhg_session* s, double primal_weight);
void tpdhg_session_free(tpdhg_session* s);

/* The JSON object gpu/pdhg.cu:881-897 writes (same keys and formats, plus the keys added here). */
int tpdhg_result_json(const tpdhg_model* model, const tpdhg_options* opt, const tpdhg_result* result, char* buf,
                      int32_t len);

/* ------------------------------------------------------------------ original-model recheck */
/* The checks of gpu/vector_check.py:46-74 and benchmarks/orig_check.py:63-88 on the model as read (host only):
 * objective c'x + const, the largest row violation relative to (1 + |violated bound|), the largest bound
 * violation, and, with row multipliers y, the weak-duality bound
 *   LB(y) = sum_j min_{x_j in [l_j,u_j]} (c - A'y)_j x_j + sum_i min_{r_i in [lo_i,up_i]} y_i r_i + const
 * for the sign of y (+ or -) with the smaller dual infeasibility (a maximisation is checked as min -c'x and its
 * bound reported in its own sense: an upper bound). */
typedef struct {
    double objective, row_viol_rel, bound_viol;
    int32_t dual_checked;      /* y was given */
    int32_t dual_bound_finite; /* no wrong-sign term on an infinite bound: dual_bound is a rigorous bound */
    int32_t dual_sign;         /* +1 or -1 */
    double dual_bound, dual_obj_loose, dual_infeas_rel, dual_gap_rel;
} tpdhg_recheck_result;
int tpdhg_recheck(const tpdhg_model* model, const double* x, const double* y, tpdhg_recheck_result* out);

/* ------------------------------------------------------------------ batches (no reference counterpart) */
/* Many LPs solved together by the fast path on one device: block-diagonal packing, one row launch and one column
 * launch per iteration for all live instances, per-instance step sizes, restarts and stopping, one HIP graph per
 * evaluation block. Instances whose scaled matrices are identical (scenario families) share one device copy.
 * Each instance follows exactly the decisions a single tpdhg_solve with the same options would take. */
typedef struct tpdhg_batch tpdhg_batch;
typedef struct {
    int32_t count, unique_matrices;
    int32_t cap;              /* nonzeros staged per row block (1024, 2048, 4096 or 8192) */
    int64_t nnz_total;        /* over all instances */
    int64_t nnz_stored;       /* held on the device (shared matrices once) */
    int64_t device_mem_bytes;
    double setup_s;           /* scaling, upload and stage-size autotuning */
    float tune_ms[4];         /* autotune timings (ms, 24 iterations) for 1024 / 2048 / 4096 / 8192; 0 when cap
                                 was given */
    int32_t persist_max;      /* live row/column blocks up to which an iteration block runs as one persistent
                                 launch (0: off, or switched off after a failed launch) */
    int64_t persistent_blocks, lockstep_blocks; /* iteration blocks run each way so far */
    int64_t stage_switches;   /* row-block stage changes so far: with an autotuned stage (cap 0) the batch moves to
                                 smaller stages as its live set shrinks below 256 blocks of the tuned stage */
} tpdhg_batch_info;
/* opt->halpern must be 1. cap: 0 = pick the row-block stage by timing (and follow the live set to smaller stages in
 * the batch's tail), else 1024, 2048, 4096 or 8192 throughout. The batch keeps copies of the models. */
int tpdhg_batch_create(const tpdhg_model* const* models, int32_t count, const tpdhg_options* opt, int32_t cap,
                       tpdhg_batch** out);
int tpdhg_batch_get_info(const tpdhg_batch* b, tpdhg_batch_info* info);
/* New cost and bounds (original space) for instance i; its matrix and scaling stay. */
int tpdhg_batch_set_vectors(tpdhg_batch* b, int32_t i, const double* cost, const double* col_lo, const double* col_up,
                            const double* row_lo, const double* row_up);
/* Solves every instance. x0/y0/primal_weight: NULL, or arrays of count entries (NULL entries: cold start);
 * results: count entries; x/y: NULL or count pointers (NULL entries skip that copy). */
int tpdhg_batch_solve(tpdhg_batch* b, const double* const* x0, const double* const* y0, const double* primal_weight,
                      tpdhg_result* results, double* const* x, double* const* y);
/* `iters` iterations of the whole batch with every instance live, each row and column launch timed by HIP events
 * on the batch stream: total ms and launch counts per side (the roofline of one batched iteration). The probe
 * launches are a separately named instantiation of the batched SpMV kernel, so kernel traces keep them apart from
 * a solve's launches, whose live sets shrink. The next solve starts over from its base points. */
int tpdhg_batch_probe(tpdhg_batch* b, int32_t iters, double* row_ms, double* col_ms, int64_t* row_launches,
                      int64_t* col_launches);
void tpdhg_batch_free(tpdhg_batch* b);

/* ------------------------------------------------------------------ learned warm start (no reference counterpart) */
/* MLP n_in -> h1 -> h2 -> n_out with ReLU, bf16 MFMA matrix products with FP32 accumulation, FP32 master
 * weights, Adam training on the GPU. n_in, h1, h2: positive multiples of 16. Parameters (FP32, row-major):
 * W1 [n_in x h1], b1 [h1], W2 [h1 x h2], b2 [h2], W3 [h2 x n_out], b3 [n_out]. bf16 rounding is applied to the
 * inputs, the weight copies, the hidden activations and the back-propagated errors before each product. */
typedef struct tpdhg_mlp tpdhg_mlp;
int tpdhg_mlp_create(int32_t n_in, int32_t h1, int32_t h2, int32_t n_out, uint64_t seed, tpdhg_mlp** out);
int64_t tpdhg_mlp_param_count(const tpdhg_mlp* net);
int tpdhg_mlp_set_params(tpdhg_mlp* net, const float* params);
int tpdhg_mlp_get_params(tpdhg_mlp* net, float* params);
/* y [batch x n_out] for x [batch x n_in]. */
int tpdhg_mlp_forward(tpdhg_mlp* net, const float* x, int32_t batch, float* y);
/* Mean squared error against t [batch x n_out] and its gradient (param_count floats, may be NULL). */
int tpdhg_mlp_loss_grad(tpdhg_mlp* net, const float* x, const float* t, int32_t batch, double* loss, float* grad);
/* Adam over shuffled mini-batches of the n samples (x [n x n_in], t [n x n_out]) for `epochs`; the data stays on
 * the device; *loss = mean loss of the last epoch. */
int tpdhg_mlp_train(tpdhg_mlp* net, const float* x, const float* t, int32_t n, int32_t batch, int32_t epochs, float lr,
                    uint64_t seed, double* loss);
void tpdhg_mlp_destroy(tpdhg_mlp* net);

/* ------------------------------------------------------------------ SpMV microbenchmark (gpu/csr_matvec.cu) */
/* y = M v for a CSR (rows x cols) on the device with the solver's SpMV kernel; `reps` timed launches after a
 * warm-up, mean and median kernel time from HIP events. */
int tpdhg_spmv_bench(int32_t precision_bits, int32_t rows, int32_t cols, const int32_t* ptr, const int32_t* idx,
                     const double* val, const double* v, double* y, int32_t reps, double* mean_ms,
                     double* median_ms);

/* ------------------------------------------------------------------ allocation-free SpMV */
/* y = M v on caller-owned DEVICE arrays, enqueued on `stream` (a hipStream_t; NULL = the default stream) and
 * returned without waiting. No reference counterpart: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops) and calls no hipMalloc, hipFree or allocating routine. ptr
 * (rows + 1) and idx (nnz) are int32 CSR with ptr[rows] == nnz and column indices in range (not checked: the
 * arrays stay on the device); val, v, y are float (precision_bits 32, FP32 accumulation) or double (64). The
 * partial sums of rows cut by the 2,048-nonzero work segments live in one static device array of 1 MiB per
 * device, so nnz is limited to tpdhg_spmv_noalloc_max_nnz() (134,217,728); larger inputs are refused with
 * TPDHG_ERR_ARG. Calls that overlap on one device share that array and race: order them on one stream. The same
 * product as tpdhg_spmv_bench, summed in another order. */
int tpdhg_spmv_noalloc(int32_t precision_bits, int32_t rows, int64_t nnz, const int32_t* ptr, const int32_t* idx,
                       const void* val, const void* v, void* y, void* stream);
int64_t tpdhg_spmv_noalloc_max_nnz(void);

/* The device's streaming ceiling, measured with the library's own kernels: median GB/s of a read-only pass and
 * of a copy (read + write bytes) over `bytes` bytes of HBM, `reps` timed launches each. The calibration kernel for
 * the FETCH_SIZE / WRITE_SIZE counters and the achievable bound the SpMV kernels are compared with. */
int tpdhg_stream_bench(int64_t bytes, int32_t reps, double* read_gbs, double* copy_gbs);

#ifdef __cplusplus
}
#endif
#endif
