// Restarted averaged PDHG driver and solve pipeline of the PDHG path: flat LP model, options, the host
// algorithm over the device backend (gpu/pdhg.cu:743-849), and the original-space report
// (gpu/pdhg.cu:858-911).
#pragma once
#include <chrono>
#include <cstdint>
#include <string>
#include <vector>

#include "batch.hpp"
#include "device.hpp"
#include "model.hpp"
#include "prep.hpp"

namespace tpdhg {

// Flat column-wise LP (CSC, rows ascending per column) with the Model fields the PDHG path reads.
struct LP {
    std::string name;
    int m = 0, n = 0;
    std::vector<int64_t> colptr{0};
    std::vector<int32_t> rowidx;
    std::vector<double> val, cost, col_lo, col_up, row_lo, row_up;
    std::vector<std::string> row_names, col_names;  // empty for array-built models
    std::vector<char> is_int;
    size_t qobj_terms = 0, sentinel_bounds = 0;
    double obj_const = 0;
    bool maximize = false;
    int64_t nnz() const { return colptr.back(); }
};

LP flatten(const Model& md);
// Reason string when PDHG must refuse the model (quadratic or integer), nullptr otherwise.
const char* input_error(const LP& lp);

struct Options {  // defaults of gpu/pdhg.cu:744-750 plus the options added here
    double tol = 1e-4, time_limit = 3600;
    int64_t max_iter = 2000000000LL;
    int eval_freq = 64;
    double row_rel = 0;
    bool verbose = false, graph = true;
    int precision_bits = 64;   // 32: FP32 iterates and SpMV, FP64 reductions (the reference's --fp32)
    bool refine = false;       // FP64 iterative refinement: recentre the FP32 inner problem at every restart
    int max_refinements = 1000000000;  // cap on the recentrings
    bool profile_kernels = false;
    bool halpern = false;      // fast path: reflected Halpern PDHG on the FP64-recentred problem
    double rho = 1.9;          // its reflection coefficient
    double omega0 = 0;         // initial primal weight of a warm-started solve (0: ||c||/||b|| as for a cold start)
    bool warm_fallback = false;  // cold-start instead when the warm point's relative KKT is not below the cold one's
    double progress_s = 0;       // > 0: a progress line on stderr every progress_s seconds of solving
    int omega_rule = 3;          // fast path and batches: primal-weight guard (driver.cpp:next_omega); the reference
                                 // scheme always uses the reference update
    bool reuse_norm = false;     // sessions and batches: power iteration on the first solve only
    int cpu_threads = -1;        // >= 0: the reference scheme on the CPU backend with this many threads (0: auto)
    int restart_metric = 1;      // fast path and batches: what the restart tests measure. 1: the fixed-point residual
                                 // ||u - T u||_P of the current point in the PDHG norm at the current weight; 0: the
                                 // weighted KKT error of the reference (gpu/pdhg.cu:783-786). The reference scheme
                                 // always uses the latter
};

struct Outcome {
    std::string status = "iteration_limit";
    int64_t iterations = 0;
    int restarts = 0, refinements = 0;
    bool from_average = false;
    double rel_primal = 0, rel_dual = 0, rel_gap = 0, norm_a = 0, omega = 1, power_s = 0, iterate_s = 0;
    std::vector<double> x, z;  // scaled-space point
    KernelTimes kernels;
    bool warm_used = false;            // the solve started from the caller's point
    double warm_kkt = 0, cold_kkt = 0;  // relative KKT of the warm and the cold start when the fallback compared them
};

// Restarted averaged PDHG on the device backend (gpu/pdhg.cu:761-849). x0/z0 (scaled space, nullable) warm start
// the iterates; omega0 > 0 replaces the initial primal weight. norm_a > 0 skips the power iteration.
Outcome run_pdhg(Device& dev, const Scaled& P, const Options& o, double time_budget, const double* x0 = nullptr,
                 const double* z0 = nullptr, double omega0 = 0, double norm_a = 0);

// The fast path (Options::halpern): reflected Halpern PDHG iterations on the FastDevice with the restart tests and
// primal-weight update of run_pdhg (the candidate is the current point), the tests applied to o.restart_metric's
// measure (by default the fixed-point residual of one PDHG step in its own norm). With o.refine (FP64 iterative
// refinement) every restart also recentres: the base point moves to the restart point in FP64 and the next inner
// problem is rebuilt with exact FP64 products, so the FP32 inner iterates only ever span one restart cycle (at most
// o.max_refinements recentrings). A point that meets the tolerance is confirmed by an FP64 evaluation after a
// recentring before the solve stops. x0/z0 (scaled space, nullable) become the first base point (warm start).
Outcome run_fast(FastDevice& dev, const Scaled& P, const Options& o, double time_budget, const double* x0 = nullptr,
                 const double* z0 = nullptr, double omega0 = 0, double norm_a = 0);

struct Timings {
    double parse = 0, scale = 0, init = 0, transfer = 0;
};

struct Report {  // what the CLI prints and the C ABI returns (gpu/pdhg.cu:866-897)
    Outcome out;
    double objective = 0, max_row_violation = 0, max_bound_violation = 0;
    double setup_s = 0, solve_s = 0, total_s = 0;
    Timings tm;
    std::string device_name, backend = "gpu";
    int threads = 0;  // host threads of the CPU backend (0 on the GPU)
    size_t device_mem_bytes = 0;
    int64_t nnz = 0;  // nonzeros of the scaled matrix (explicit zeros dropped)
    std::vector<double> x, y;  // original space: x = Dc x_s, y = Dr z_s
};

// A problem resident on the device: scaled once, uploaded once, then solved any number of times from cold or
// warm starts. Every solve runs the power iteration, unless Options::reuse_norm keeps the first solve's ||A_s||
// (A_s never changes).
class Session {
public:
    Session(const LP& lp, const Options& o, Timings tm = Timings());
    // Warm start from (x0, y0) (original space, nullable) with Options::omega0; with Options::warm_fallback the
    // start is kept only when its relative KKT is below the cold start's.
    Report solve(const double* x0_orig = nullptr, const double* y0_orig = nullptr);
    void set_omega0(double w) { o_.omega0 = w; }
    // New cost and bounds (original space) for the same A: rescaled with the session's factors and uploaded.
    // `lp` must be the model the session was created on; its vectors are overwritten.
    void set_vectors(LP& lp, const double* cost, const double* col_lo, const double* col_up, const double* row_lo,
                     const double* row_up);
    const Scaled& scaled() const { return P_; }
    Device* device() { return dev_.get(); }
    FastDevice* fast_device() { return fast_.get(); }
    const Options& options() const { return o_; }

private:
    const LP& lp_;
    Options o_;
    Timings tm_;
    Scaled P_;
    std::unique_ptr<Device> dev_;
    std::unique_ptr<FastDevice> fast_;
    double probe_kkt(const double* xs, const double* zs);  // relative KKT at a scaled-space start point
    double setup_s_ = 0, norm_a_ = 0;
    int threads_ = 0;  // CPU backend threads (0: GPU)
    std::chrono::steady_clock::time_point created_;
};

// Original-space point, objective and violations of rep.out (gpu/pdhg.cu:866-876).
void finish_report(const LP& lp, const Scaled& P, Report& rep);

// The fast path over a batch: every instance follows run_fast's decisions (restart tests, primal weight,
// recentring, FP64 confirmation) on its own scalars while all live instances iterate together; an instance
// leaves the lockstep when it stops. x0/z0/omega0 per instance (empty or null entries: cold start). norms, when
// it holds one value per instance, skips the power iteration; it receives the estimates otherwise.
std::vector<Outcome> run_batch(BatchDevice& b, const std::vector<const Scaled*>& P, const Options& o,
                               double time_budget, const std::vector<const double*>& x0,
                               const std::vector<const double*>& z0, const std::vector<double>& omega0,
                               std::vector<double>* norms);

// Many LPs resident on the device as one batch (the batch keeps its own copies of the models).
class BatchSession {
public:
    BatchSession(const std::vector<const LP*>& lps, const Options& o, int cap = 0);
    // x0/y0 per instance in original space (null entries: cold start), primal weight per instance (0: default);
    // with Options::warm_fallback each instance keeps its warm start only when its KKT beats its cold start's.
    std::vector<Report> solve(const std::vector<const double*>& x0 = {}, const std::vector<const double*>& y0 = {},
                              const std::vector<double>& omega0 = {});
    void set_vectors(int i, const double* cost, const double* col_lo, const double* col_up, const double* row_lo,
                     const double* row_up);
    int count() const { return int(lps_.size()); }
    const LP& lp(int i) const { return lps_[size_t(i)]; }
    BatchDevice& device() { return *dev_; }
    double setup_s() const { return setup_s_; }

private:
    std::vector<LP> lps_;
    std::vector<Scaled> P_;
    Options o_;
    std::unique_ptr<BatchDevice> dev_;
    std::vector<double> norms_;
    double setup_s_ = 0;
};

// Scales, uploads and solves on the current HIP device; throws std::runtime_error on device failure.
Report solve_lp(const LP& lp, const Options& o, Timings tm, std::chrono::steady_clock::time_point t_start,
                const double* x0_orig = nullptr, const double* y0_orig = nullptr);

std::string report_json(const Report& r, const LP& lp, const Options& o);

}  // namespace tpdhg
