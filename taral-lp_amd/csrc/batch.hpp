// Batched backend of the fast path: many LPs packed block-diagonally into one set of device buffers and advanced
// together, one row launch and one column launch per iteration for the whole batch (captured in a HIP graph per
// evaluation block). Each instance has its own scalars (step sizes, Halpern counter, activity) and its own KKT
// evaluation; instances with an identical scaled matrix share one copy of it on the device (scenario families).
//
// Arithmetic per instance is the FastDevice's exactly: the same row blocks, the same epilogues and the same
// reduction partition, so a batch reproduces the single-instance fast path bit for bit (non-slab layouts, rows of
// at most 2,048 nonzeros: a longer row is summed by one workgroup here and in 2,048-nonzero chunks there).
// No reference counterpart: the reference solves one LP per process (gpu/pdhg.cu:913-935).
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "device.hpp"
#include "prep.hpp"

namespace tpdhg {

class BatchDevice {
public:
    virtual ~BatchDevice() = default;
    virtual int count() const = 0;
    virtual int m(int i) const = 0;
    virtual int n(int i) const = 0;
    // Power iteration estimate of ||A_s|| for every instance (the single-instance rule, instances stop apart).
    virtual std::vector<double> power_norm(int iters) = 0;
    // FP64 base points (scaled space) per instance; null entries (or a null array) = 0.
    virtual void set_base(const std::vector<const double*>& xb, const std::vector<const double*>& zb) = 0;
    // The single-instance operations applied to the instances whose flag is set.
    virtual void begin_round(const std::vector<uint8_t>& on) = 0;
    virtual void set_steps(const std::vector<double>& tau, const std::vector<double>& sigma) = 0;
    virtual void restart(const std::vector<uint8_t>& on) = 0;
    virtual void end_round(const std::vector<uint8_t>& on) = 0;
    // `steps` iterations of the instances flagged in `on`; the flags stay fixed over the block.
    virtual void iterate(int steps, const std::vector<uint8_t>& on) = 0;
    // KKT slots (NKKT per instance, instance-major) of the flagged instances; other rows are left as they were.
    virtual void eval(const std::vector<uint8_t>& on, std::vector<double>& out) = 0;
    // Per instance (3 entries each, valid for the instances of the last eval): the fixed-point residual terms of
    // FastDevice::fixed_point.
    virtual void fixed_points(std::vector<double>& out) const = 0;
    // base + inner point of instance i in FP64, scaled space.
    virtual void download(int i, double* x, double* z) = 0;
    // New scaled c, l, u, rl, ru of instance i (its matrix and scaling stay).
    virtual void set_vectors(int i, const double* c, const double* l, const double* u, const double* rl,
                             const double* ru) = 0;
    virtual void set_profile(bool on) = 0;
    virtual KernelTimes kernel_times() = 0;
    virtual const std::string& name() const = 0;
    virtual size_t mem_used() const = 0;
    virtual int64_t nnz_stored() const = 0;   // matrix nonzeros held on the device (shared matrices once)
    virtual int64_t nnz_total() const = 0;    // sum over instances
    virtual int unique_matrices() const = 0;
    virtual int cap() const = 0;  // row-block stage size in use
    // Autotune timings (ms for 24 iterations, stage sizes 1024, 2048, 4096, 8192; empty when the size was given).
    virtual std::vector<float> tune_ms() const = 0;
    // `iters` iterations with every instance live, each row and column launch timed by HIP events on the batch
    // stream. The launches use their own instantiation of the batched SpMV kernel (a distinct kernel name, same
    // code), so a kernel trace separates these all-live launches from a solve's shrinking ones. Leaves the
    // iterates changed: the next solve starts over from its base points.
    virtual KernelTimes probe(int iters) = 0;
    // Iteration blocks run so far as one persistent launch / as lockstep launches, and the live-block limit of the
    // persistent path (0 when it is off).
    virtual int64_t persistent_blocks() const = 0;
    virtual int64_t lockstep_blocks() const = 0;
    virtual int persist_max() const = 0;
    // Stage changes made so far as the live set shrank (0 with a given stage).
    virtual int64_t stage_switches() const = 0;
};

struct BatchOptions {
    int precision_bits = 32;
    int eval_freq = 64;
    bool use_graph = true;
    double rho = 1.9;
    int cap = 0;  // nonzeros staged per row block: 1024, 2048, 4096 or 8192; 0 = pick by timing (autotune)
    // Largest live block count (per side) run as one persistent launch per iteration block; 0 = lockstep only
    // (the default: measured slower on the small live sets it serves, see the design notes; TPDHG_PERSIST_MAX
    // overrides).
    int persist_max = 0;
};

// Instances are the scaled problems; instances whose scaled A (and so A^T) is identical share the device copy.
std::unique_ptr<BatchDevice> make_batch_device(const std::vector<const Scaled*>& probs, const BatchOptions& opt);

}  // namespace tpdhg
