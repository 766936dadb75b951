// Device backend of the PDHG path on one MI355X (gfx950): owns the scaled problem and the iterates in HBM and
// runs the hand-written HIP kernels. It offers the operations of the reference backend concept
// (gpu/pdhg.cu:313-514: init, power_norm, set_steps, start, iterate, eval, restart, download, name,
// mem_used), plus a fused current+average evaluation and per-kernel event timing. No HIP types appear here.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "prep.hpp"

namespace tpdhg {

// KKT scalar slots of one evaluation: row part out[0..4] (gpu/pdhg.cu:136), column part out[5..9] (:145).
enum KktSlot { PR2S = 0, PR2O, PR2R, DROW, DZ2, DR2S = 5, DR2O, POBJ, DCOL, DX2, NKKT = 10 };

struct KernelTimes {  // HIP-event totals of the two iteration kernels since the last reset
    double row_ms = 0, col_ms = 0;
    int64_t row_launches = 0, col_launches = 0;
};

class Device {
public:
    virtual ~Device() = default;
    // Power iteration estimate of ||A_s||_2 (gpu/pdhg.cu:416-436).
    virtual double power_norm(int iters) = 0;
    // Primal and dual step sizes, kept in device memory so a captured iteration block never changes.
    virtual void set_steps(double tau, double sigma) = 0;
    // x0 = clamp(x_init or 0, l, u), z0 = z_init or 0 (scaled space), then Ax, A^T z (gpu/pdhg.cu:450-458).
    virtual void start(const double* x_init, const double* z_init) = 0;
    // One block: first x-update, then `steps` x (dual row step, column step fused with the next x-update)
    // (gpu/pdhg.cu:459-486).
    virtual void iterate(int steps) = 0;
    // KKT scalars of current (cur) and average (avg, scaled by f) in one pass; either pointer may be null.
    virtual void eval2(double f, double* cur, double* avg) = 0;
    // Restart to the average (avg, scaled by f) or the current point (gpu/pdhg.cu:500-504).
    virtual void restart(bool avg, double f) = 0;
    // Current or f-scaled average iterate in FP64, scaled space (gpu/pdhg.cu:505-513).
    virtual void download(bool avg, double f, double* x, double* z) = 0;
    // y = A_s v (which = 0) or A_s^T v (which = 1) on host vectors, for checks.
    virtual void spmv(int which, const double* v, double* y) = 0;
    virtual void set_profile(bool on) = 0;
    virtual KernelTimes kernel_times() = 0;
    // Replaces c, l, u (n) and rl, ru (m), scaled space; A, A^T and the scaling stay (scenario families).
    virtual void set_vectors(const double* c, const double* l, const double* u, const double* rl, const double* ru) = 0;
    virtual void synchronize() = 0;
    virtual const std::string& name() const = 0;
    virtual size_t mem_used() const = 0;
    virtual int precision_bits() const = 0;
    virtual int m() const = 0;
    virtual int n() const = 0;
};

struct DeviceOptions {
    int precision_bits = 64;  // 64: FP64 iterates; 32: FP32 iterates and SpMV, FP64 reductions
    int eval_freq = 64;
    bool use_graph = true;
};

// Uploads the scaled problem (A, A^T, c, l, u, rl, ru, Dr, Dc) to the current HIP device. Throws
// std::runtime_error on any HIP failure; with no HIP device the error says so (there is no CPU path).
std::unique_ptr<Device> make_device(const Scaled& P, const DeviceOptions& opt);

// The reference scheme on host threads (the reference's CPU backend, `--device cpu`): threads > 0 threads, or 0
// for the reference's choice (one below 200,000 nonzeros, else every hardware thread). Results do not depend on
// the thread count.
std::unique_ptr<Device> make_cpu_device(const Scaled& P, const DeviceOptions& opt, int threads);
int cpu_threads_for(int64_t nnz, int threads);

// Same from raw scaled arrays (A and A^T as CSR, int32 indices).
std::unique_ptr<Device> make_device_raw(int m, int n, int64_t nnz, const int32_t* a_ptr, const int32_t* a_idx,
                                        const double* a_val, const int32_t* t_ptr, const int32_t* t_idx,
                                        const double* t_val, const double* c, const double* l, const double* u,
                                        const double* rl, const double* ru, const double* Dr, const double* Dc,
                                        const DeviceOptions& opt);

// Backend of the fast path: reflected Halpern PDHG iterations (FP32 or FP64 iterates) on a recentred problem
// around an FP64 base point (x_b, z_b), which is how FP64 iterative refinement and warm starts enter. In inner
// coordinates the LP is  min (c + A^T z_b)' dx  s.t.  rl - A x_b <= A dx <= ru - A x_b,  l - x_b <= dx <= u - x_b,
// with the row prox shifted by z_b so that inequality rows stay exact. One iteration k (anchor x0, z0,
// lambda = (k+1)/(k+2), reflection rho):
//   x+ = proj(x - tau (c_in + A^T z)),  z+ = prox(z + sigma A (2 x+ - x)),
//   x <- lambda (rho x+ - (rho-1) x) + (1 - lambda) x0,  z <- same with z+, z0.
// KKT scalars are evaluated in FP64 at the combined point base + inner on the original scaled problem, with the
// same ten slots as Device::eval2 (DX2/DZ2 measured from the anchors).
class FastDevice {
public:
    virtual ~FastDevice() = default;
    virtual double power_norm(int iters) = 0;
    // FP64 base point in scaled space (null = 0 for that part); takes effect at the next begin_round().
    virtual void set_base(const double* x_b, const double* z_b) = 0;
    // Inner problem from the current base; inner iterates reset to dx = clamp(0, l_in, u_in), dz = 0.
    virtual void begin_round() = 0;
    virtual void set_steps(double tau, double sigma) = 0;
    // Anchors = current point, Halpern counter = 0, first primal step from the current dual (after set_steps).
    virtual void restart() = 0;
    virtual void iterate(int steps) = 0;
    virtual void eval(double out[NKKT]) = 0;
    // Fixed-point residual of the last eval's inner point u = (x, z) under one iteration T of the (unreflected,
    // unanchored) PDHG step: out[0] = ||x - x+||^2, out[1] = ||z - z+||^2, out[2] = (z - z+)' A (x - x+), FP64 sums;
    // ||u - T u||_P^2 = out[0] / tau + out[1] / sigma + 2 out[2] in the norm of the PDHG metric.
    virtual void fixed_point(double out[3]) const = 0;
    // base += inner point (FP64).
    virtual void end_round() = 0;
    // base + inner point in FP64, scaled space; either may be null.
    virtual void download(double* x, double* z) = 0;
    virtual void set_profile(bool on) = 0;
    virtual KernelTimes kernel_times() = 0;
    // Replaces c, l, u (n) and rl, ru (m), scaled space, FP64; A, A^T and the scaling stay.
    virtual void set_vectors(const double* c, const double* l, const double* u, const double* rl, const double* ru) = 0;
    virtual const std::string& name() const = 0;
    virtual size_t mem_used() const = 0;
    virtual int precision_bits() const = 0;
};

struct FastOptions {
    int precision_bits = 32;  // iterate type: 32 (FP32 iterates and SpMV, FP64 evaluation and base) or 64
    int eval_freq = 64;
    bool use_graph = true;
    double rho = 1.9;  // reflection coefficient of the Halpern iteration, in [1, 2)
};

std::unique_ptr<FastDevice> make_fast_device(const Scaled& P, const FastOptions& opt);

// Number of visible HIP devices (0 when none); never throws.
int device_count();

// Times y = M v for a CSR on the device with the solver's SpMV kernel: `reps` launches after one warm-up,
// returns the mean and the median kernel time in ms (HIP events) and y (FP64).
// Median HBM rates (GB/s) of this library's streaming-read and copy kernels over `bytes` bytes (copy: read +
// write bytes), `reps` timed launches each.
void stream_bench(int64_t bytes, int reps, double* read_gbs, double* copy_gbs);
void spmv_bench(int precision_bits, int rows, int cols, const int32_t* ptr, const int32_t* idx, const double* val,
                const double* v, double* y, int reps, double* mean_ms, double* median_ms);

}  // namespace tpdhg
