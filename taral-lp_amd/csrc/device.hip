// HIP kernels and device backend of the restarted PDHG path, written for gfx950 (CDNA4, wave64).
//
// SpMV design (shared by A x, A^T z, and the fused PDHG steps): rows are grouped on the host into row
// blocks of at most kCap nonzeros and kWG rows. One 256-thread workgroup takes one block: it streams the
// block's val/idx with unit-stride loads (each wave64 instruction reads 256 contiguous bytes), gathers the
// vector entries, stages the products in LDS, then reduces each row with a power-of-two group of lanes
// sized to the block (xor shuffles inside a wave) and applies the fused epilogue (dual step, column step
// plus next primal step, or a plain store). A row longer than kCap gets a block of its own and is reduced
// across the whole workgroup. Blocks are mapped XCD-aware: the hardware deals workgroups round-robin over
// the 8 XCDs, and block b is remapped so each XCD walks one contiguous range of rows (its L2 then sees a
// contiguous slice of the output and, for banded matrices, of the input vector).
//
// KKT evaluation is one fused pass over rows and columns for the current and the averaged point together,
// FP64 accumulation from the iterate type, per-workgroup partials and a fixed-order final sum, so results
// are deterministic run to run. The reference computes these in gpu/pdhg.cu:236-291 with one pass per point.
#include <hip/hip_runtime.h>

#include "hipcommon.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "device.hpp"

namespace tpdhg {
namespace {

// Dual (row) step, the semantics of gpu/pdhg.cu:179-190: s = (A x)_i, v = z + sigma (2 s - Ax_old),
// z+ = v - sigma proj_[rl, ru](v / sigma); Ax = s; running sums Axs += s, zs += z+.
template <class T>
struct EpiRow {
    T *Ax, *z, *Axs, *zs;
    const T *rl, *ru, *par;
    T sig;
    struct Pre {
        T zi, axi, lo, hi, axs, zs;
    };
    __device__ void load() { sig = par[1]; }
    __device__ Pre pre(int i) const { return Pre{z[i], Ax[i], rl[i], ru[i], Axs[i], zs[i]}; }
    __device__ void apply(int i, T s, const Pre& p) const {
        const T v = p.zi + sig * (2 * s - p.axi), q = v / sig;
        const T zn = q < p.lo ? v - sig * p.lo : (q > p.hi ? v - sig * p.hi : T(0));
        Ax[i] = s, z[i] = zn, Axs[i] = p.axs + s, zs[i] = p.zs + zn;
    }
    __device__ void operator()(int i, T s) const { apply(i, s, pre(i)); }
};

// Column step, gpu/pdhg.cu:193-206: s = (A^T z)_j; ATz = s, ATzs += s; when upd, the next primal step
// x+ = proj_[l, u](x - tau (c + s)), xs += x+.
template <class T>
struct EpiCol {
    T *ATz, *ATzs, *x, *xs;
    const T *c, *l, *u, *par;
    int upd;
    T tau;
    struct Pre {
        T atzs, xj, c, l, u, xs;
    };
    __device__ void load() { tau = par[0]; }
    __device__ Pre pre(int j) const {
        if (upd) return Pre{ATzs[j], x[j], c[j], l[j], u[j], xs[j]};
        return Pre{ATzs[j], T(0), T(0), T(0), T(0), T(0)};
    }
    __device__ void apply(int j, T s, const Pre& p) const {
        ATz[j] = s, ATzs[j] = p.atzs + s;
        if (upd) {
            const T xn = clampv(p.xj - tau * (p.c + s), p.l, p.u);
            x[j] = xn, xs[j] = p.xs + xn;
        }
    }
    __device__ void operator()(int j, T s) const { apply(j, s, pre(j)); }
};

// First primal step of a block from the stored A^T z (gpu/pdhg.cu:208-214).
template <class T>
__global__ __launch_bounds__(kWG) void k_xupd(int n, T* __restrict__ x, const T* __restrict__ ATz,
                                              const T* __restrict__ c, const T* __restrict__ l,
                                              const T* __restrict__ u, const T* __restrict__ par,
                                              T* __restrict__ xs) {
    const T tau = par[0];
    for (int j = int(blockIdx.x * blockDim.x + threadIdx.x); j < n; j += int(gridDim.x * blockDim.x)) {
        const T xn = clampv(x[j] - tau * (c[j] + ATz[j]), l[j], u[j]);
        x[j] = xn, xs[j] += xn;
    }
}

template <class T>
struct EvalArgs {
    const T *Ax, *z, *Axs, *zs, *rl, *ru, *Dr, *zlast;
    const T *x, *ATz, *xs, *ATzs, *c, *l, *u, *Dc, *xlast;
};

// KKT scalars of the current point and of the f-scaled average in one pass (gpu/pdhg.cu:236-271 twice).
// Blocks [0, nbr) cover rows, blocks [nbr, grid) cover columns; each writes 10 partials:
// rows: current PR2S..DZ2 then average PR2S..DZ2; columns: current DR2S..DX2 then average DR2S..DX2.
template <class T>
__global__ __launch_bounds__(kWG) void k_eval2(int m, int n, int nbr, double f, EvalArgs<T> a,
                                               double* __restrict__ part) {
    double v[10];
    for (int k = 0; k < 10; ++k) v[k] = 0;
    const int b = int(blockIdx.x);
    if (b < nbr) {
        for (int i = b * kWG + int(threadIdx.x); i < m; i += nbr * kWG) {
            const double lo = a.rl[i], up = a.ru[i], dr = a.Dr[i], zl = a.zlast[i];
            const double wrel = row_rel_weight(lo, up, dr);
            for (int h = 0; h < 2; ++h) {
                const double ax = h ? a.Axs[i] * f : double(a.Ax[i]), zz = h ? a.zs[i] * f : double(a.z[i]);
                const double r = ax < lo ? lo - ax : (ax > up ? ax - up : 0), ro = r / dr, rr = ro * wrel;
                double* o = v + 5 * h;
                o[0] += r * r, o[1] += ro * ro, o[2] += rr * rr;
                if (zz > 0 && up < INFINITY) o[3] -= zz * up;
                if (zz < 0 && lo > -INFINITY) o[3] -= zz * lo;
                const double dz = zz - zl;
                o[4] += dz * dz;
            }
        }
    } else {
        const int nbc = int(gridDim.x) - nbr;
        for (int j = (b - nbr) * kWG + int(threadIdx.x); j < n; j += nbc * kWG) {
            const double cj = a.c[j], lo = a.l[j], up = a.u[j], dc = a.Dc[j], xl = a.xlast[j];
            for (int h = 0; h < 2; ++h) {
                const double xx = h ? a.xs[j] * f : double(a.x[j]);
                const double d = cj + (h ? a.ATzs[j] * f : double(a.ATz[j]));
                double res = 0;
                double* o = v + 5 * h;
                o[2] += cj * xx;
                if (d > 0) {
                    if (lo > -INFINITY) o[3] += d * lo;
                    else res = d;
                } else if (d < 0) {
                    if (up < INFINITY) o[3] += d * up;
                    else res = -d;
                }
                const double ro = res / dc, dx = xx - xl;
                o[0] += res * res, o[1] += ro * ro, o[4] += dx * dx;
            }
        }
    }
    block_store<10>(v, part);
}


// Restart, gpu/pdhg.cu:293-299: x <- average (xs f) or current; xlast = x; clear both running sums.
template <class T>
__global__ __launch_bounds__(kWG) void k_restart(int n, int avg, double f, T* __restrict__ x, T* __restrict__ xs,
                                                 T* __restrict__ xlast, T* __restrict__ ys) {
    for (int j = int(blockIdx.x * blockDim.x + threadIdx.x); j < n; j += int(gridDim.x * blockDim.x)) {
        const T v = avg ? T(xs[j] * f) : x[j];
        x[j] = v, xlast[j] = v, xs[j] = 0, ys[j] = 0;
    }
}

// ------------------------------------------------------------------------------------------------ host side
template <class T>
class Dev final : public Device {
public:
    Dev(int m, int n, const int32_t* a_ptr, const int32_t* a_idx, const double* a_val, const int32_t* t_ptr,
        const int32_t* t_idx, const double* t_val, const double* c_, const double* l_, const double* u_,
        const double* rl_, const double* ru_, const double* Dr_, const double* Dc_, const DeviceOptions& opt)
        : m_(m), n_(n), eval_freq_(opt.eval_freq), use_graph_(opt.use_graph) {
        require_device();
        size_t free0 = 0, free1 = 0, total = 0;
        HIP_OK(hipMemGetInfo(&free0, &total));
        name_ = device_name();
        HIP_OK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
        A_ = upload_csr<T>(mem_, st_, m, n, a_ptr, a_idx, a_val);
        AT_ = upload_csr<T>(mem_, st_, n, m, t_ptr, t_idx, t_val);
        hl_.assign(l_, l_ + n), hu_.assign(u_, u_ + n);
        c = up(c_, n), l = up(l_, n), u = up(u_, n), rl = up(rl_, m), ru = up(ru_, m), Dr = up(Dr_, m), Dc = up(Dc_, n);
        par = zeros(2);
        x = zeros(n), ATz = zeros(n), xs = zeros(n), ATzs = zeros(n), xlast = zeros(n);
        z = zeros(m), Ax = zeros(m), zs = zeros(m), Axs = zeros(m), zlast = zeros(m);
        pv = zeros(n), pw = zeros(m), pu = zeros(n);
        nbr_ = red_blocks(m), nbc_ = red_blocks(n);
        part_ = mem_.alloc<double>(size_t(std::max(1, nbr_ + nbc_)) * 10 + size_t(kMaxRedBlocks));
        static_assert(kPowerGroup <= 32, "a power-iteration group's sums of squares go to out_");
        out_ = mem_.alloc<double>(32);
        f64buf_ = mem_.alloc<double>(size_t(std::max(m, n)));
        HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&hout_), 32 * sizeof(double), hipHostMallocDefault));
        HIP_OK(hipStreamSynchronize(st_));
        HIP_OK(hipMemGetInfo(&free1, &total));
        mem_used_ = free0 > free1 ? free0 - free1 : 0;
    }
    ~Dev() override {
        if (graph_) (void)hipGraphExecDestroy(graph_);
        for (hipEvent_t e : events_) (void)hipEventDestroy(e);
        if (hout_) (void)hipHostFree(hout_);
        if (st_) (void)hipStreamDestroy(st_);
    }

    double power_norm(int iters) override {
        std::vector<double> h(static_cast<size_t>(n_));
        double nn = 0;
        for (int j = 0; j < n_; ++j) {
            h[size_t(j)] = 0.5 + double((uint32_t(j) * 2654435761u) % 1000u) / 1000.0;  // gpu/pdhg.cu:153
            nn += h[size_t(j)] * h[size_t(j)];
        }
        std::vector<T> ht(static_cast<size_t>(n_));
        for (int j = 0; j < n_; ++j) ht[size_t(j)] = T(h[size_t(j)] / std::sqrt(nn));
        h2d(pv, ht.data(), size_t(n_));
        // The loop of gpu/pdhg.cu:417-436 with each step normalised on the device (x *= 1 / sqrt(s), s still on the
        // device), so kPowerGroup steps run between two host reads of their sums of squares; the estimate is the
        // one of the step the stopping rule picks, the group's later steps discarded.
        const int b = red_blocks(n_);
        if (!b) return 0;
        double est = 0;
        for (int it0 = 0; it0 < iters; it0 += kPowerGroup) {
            const int g = std::min(kPowerGroup, iters - it0);
            for (int q = 0; q < g; ++q) {
                spmv_dev(A_, pv, pw);
                spmv_dev(AT_, pw, pu);
                k_sumsq<T><<<b, kWG, 0, st_>>>(n_, pu, part_);
                k_finish<<<1, kWG, 0, st_>>>(0, b, 1, part_, out_ + q);
                k_scale_by_sum<T><<<grid_for(n_), kWG, 0, st_>>>(n_, pu, out_ + q);
                std::swap(pv, pu);
            }
            HIP_OK(hipMemcpyAsync(hout_, out_, size_t(g) * sizeof(double), hipMemcpyDeviceToHost, st_));
            HIP_OK(hipStreamSynchronize(st_));
            for (int q = 0; q < g; ++q) {
                const double s = hout_[q], prev = est;
                if (s <= 0) return 0;
                est = std::sqrt(std::sqrt(s));
                if (it0 + q >= 10 && std::fabs(est - prev) <= 1e-7 * est) return est;
            }
        }
        return est;
    }

    void set_steps(double tau, double sigma) override {
        T h[2] = {T(tau), T(sigma)};
        h2d(par, h, 2);
    }

    void start(const double* x_init, const double* z_init) override {
        // x0 = clamp(x_init or 0, l, u) in FP64 on the host with the FP64 bounds (gpu/pdhg.cu:452), z0 = z_init or 0;
        // the running sums are cleared so a handle can be started again.
        std::vector<T> hx(static_cast<size_t>(n_)), hz(static_cast<size_t>(m_), T(0));
        for (int j = 0; j < n_; ++j) {
            const double v = x_init ? x_init[j] : 0.0;
            hx[size_t(j)] = T(std::min(std::max(v, hl_[size_t(j)]), hu_[size_t(j)]));
        }
        if (z_init)
            for (int i = 0; i < m_; ++i) hz[size_t(i)] = T(z_init[i]);
        h2d(x, hx.data(), size_t(n_)), h2d(xlast, hx.data(), size_t(n_));
        h2d(z, hz.data(), size_t(m_)), h2d(zlast, hz.data(), size_t(m_));
        for (T* p : {xs, ATzs}) HIP_OK(hipMemsetAsync(p, 0, size_t(std::max(n_, 1)) * sizeof(T), st_));
        for (T* p : {zs, Axs}) HIP_OK(hipMemsetAsync(p, 0, size_t(std::max(m_, 1)) * sizeof(T), st_));
        products();
        HIP_OK(hipStreamSynchronize(st_));
        HIP_OK(hipGetLastError());
    }

    void iterate(int steps) override {
        if (steps <= 0) return;
        if (use_graph_ && !profile_ && steps == eval_freq_) {
            if (!graph_) {
                hipGraph_t g = nullptr;
                HIP_OK(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
                try {
                    launch_block(steps);
                } catch (...) {
                    (void)hipStreamEndCapture(st_, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                HIP_OK(hipStreamEndCapture(st_, &g));
                HIP_OK(hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0));
                HIP_OK(hipGraphDestroy(g));
            }
            HIP_OK(hipGraphLaunch(graph_, st_));
        } else {
            launch_block(steps);
        }
        HIP_OK(hipGetLastError());
    }

    void eval2(double f, double* cur, double* avg) override {
        if (nbr_ + nbc_ > 0) {
            EvalArgs<T> a{Ax, z, Axs, zs, rl, ru, Dr, zlast, x, ATz, xs, ATzs, c, l, u, Dc, xlast};
            k_eval2<T><<<nbr_ + nbc_, kWG, 0, st_>>>(m_, n_, nbr_, f, a, part_);
            k_finish<<<1, kWG, 0, st_>>>(0, nbr_, 10, part_, out_);
            k_finish<<<1, kWG, 0, st_>>>(nbr_, nbc_, 10, part_, out_ + 10);
            HIP_OK(hipMemcpyAsync(hout_, out_, 20 * sizeof(double), hipMemcpyDeviceToHost, st_));
            HIP_OK(hipStreamSynchronize(st_));
        } else {
            std::fill(hout_, hout_ + 20, 0.0);
        }
        HIP_OK(hipGetLastError());
        // row partials: current PR2S..DZ2 then average; column partials: current DR2S..DX2 then average
        if (cur) std::copy(hout_, hout_ + 5, cur), std::copy(hout_ + 10, hout_ + 15, cur + 5);
        if (avg) std::copy(hout_ + 5, hout_ + 10, avg), std::copy(hout_ + 15, hout_ + 20, avg + 5);
        collect_events();
    }

    void restart(bool avg, double f) override {
        k_restart<T><<<grid_for(n_), kWG, 0, st_>>>(n_, avg ? 1 : 0, f, x, xs, xlast, ATzs);
        k_restart<T><<<grid_for(m_), kWG, 0, st_>>>(m_, avg ? 1 : 0, f, z, zs, zlast, Axs);
        products();
        HIP_OK(hipGetLastError());
    }

    void download(bool avg, double f, double* hx, double* hz) override {
        if (hx && n_) {
            k_to_f64<T><<<grid_for(n_), kWG, 0, st_>>>(n_, avg ? f : 1.0, avg ? xs : x, f64buf_);
            HIP_OK(hipMemcpyAsync(hx, f64buf_, size_t(n_) * sizeof(double), hipMemcpyDeviceToHost, st_));
            HIP_OK(hipStreamSynchronize(st_));
        }
        if (hz && m_) {
            k_to_f64<T><<<grid_for(m_), kWG, 0, st_>>>(m_, avg ? f : 1.0, avg ? zs : z, f64buf_);
            HIP_OK(hipMemcpyAsync(hz, f64buf_, size_t(m_) * sizeof(double), hipMemcpyDeviceToHost, st_));
            HIP_OK(hipStreamSynchronize(st_));
        }
    }

    void spmv(int which, const double* v, double* y) override {
        const DevMat<T>& a = which ? AT_ : A_;
        const int cols = which ? m_ : n_, rows = which ? n_ : m_;
        T* in = which ? pw : pv;   // pv holds n entries, pw m, pu n
        T* res = which ? pu : pw;
        std::vector<T> h(v, v + cols);
        h2d(in, h.data(), size_t(cols));
        spmv_dev(a, in, res);
        if (rows) {
            k_to_f64<T><<<grid_for(rows), kWG, 0, st_>>>(rows, 1.0, res, f64buf_);
            HIP_OK(hipMemcpyAsync(y, f64buf_, size_t(rows) * sizeof(double), hipMemcpyDeviceToHost, st_));
        }
        HIP_OK(hipStreamSynchronize(st_));
        HIP_OK(hipGetLastError());
    }

    void set_profile(bool on) override { profile_ = on; }
    KernelTimes kernel_times() override {
        HIP_OK(hipStreamSynchronize(st_));
        collect_events();
        KernelTimes t = times_;
        times_ = KernelTimes();
        return t;
    }
    void set_vectors(const double* c_, const double* l_, const double* u_, const double* rl_, const double* ru_) override {
        auto put = [&](T* dst, const double* src, int k) {
            if (!k) return;
            std::vector<T> h(src, src + k);
            h2d(dst, h.data(), size_t(k));
        };
        put(c, c_, n_), put(l, l_, n_), put(u, u_, n_), put(rl, rl_, m_), put(ru, ru_, m_);
        hl_.assign(l_, l_ + n_), hu_.assign(u_, u_ + n_);
    }
    void synchronize() override { HIP_OK(hipStreamSynchronize(st_)); }
    const std::string& name() const override { return name_; }
    size_t mem_used() const override { return mem_used_; }
    int precision_bits() const override { return int(sizeof(T) * 8); }
    int m() const override { return m_; }
    int n() const override { return n_; }

private:
    T* up(const double* src, int k) {
        T* p = mem_.alloc<T>(size_t(k));
        if (k) {
            std::vector<T> h(src, src + k);
            HIP_OK(hipMemcpyAsync(p, h.data(), size_t(k) * sizeof(T), hipMemcpyHostToDevice, st_));
            HIP_OK(hipStreamSynchronize(st_));
        }
        return p;
    }
    T* zeros(int k) {
        T* p = mem_.alloc<T>(size_t(k));
        HIP_OK(hipMemsetAsync(p, 0, std::max<size_t>(size_t(k), 1) * sizeof(T), st_));
        return p;
    }
    void h2d(T* dst, const T* src, size_t k) {
        if (!k) return;
        HIP_OK(hipMemcpyAsync(dst, src, k * sizeof(T), hipMemcpyHostToDevice, st_));
        HIP_OK(hipStreamSynchronize(st_));
    }
    void spmv_dev(const DevMat<T>& a, const T* v, T* y) { launch_rows(a, v, EpiStore<T>{y}, st_); }
    void products() {  // Ax = A x, ATz = A^T z exactly
        spmv_dev(A_, x, Ax);
        spmv_dev(AT_, z, ATz);
    }
    hipEvent_t event(size_t i) {
        while (events_.size() <= i) {
            hipEvent_t e;
            HIP_OK(hipEventCreate(&e));
            events_.push_back(e);
        }
        return events_[i];
    }
    void launch_block(int steps) {
        k_xupd<T><<<grid_for(n_), kWG, 0, st_>>>(n_, x, ATz, c, l, u, par, xs);
        for (int s = 0; s < steps; ++s) {
            EpiRow<T> er{Ax, z, Axs, zs, rl, ru, par, T(0)};
            EpiCol<T> ec{ATz, ATzs, x, xs, c, l, u, par, s + 1 < steps ? 1 : 0, T(0)};
            if (profile_) {
                const size_t e = pending_.size();
                HIP_OK(hipEventRecord(event(e), st_));
                launch_rows(A_, x, er, st_);
                HIP_OK(hipEventRecord(event(e + 1), st_));
                launch_rows(AT_, z, ec, st_);
                HIP_OK(hipEventRecord(event(e + 2), st_));
                pending_.push_back(0), pending_.push_back(1), pending_.push_back(2);
            } else {
                launch_rows(A_, x, er, st_);
                launch_rows(AT_, z, ec, st_);
            }
        }
    }
    void collect_events() {
        if (pending_.empty()) return;
        HIP_OK(hipStreamSynchronize(st_));
        for (size_t i = 0; i + 2 < pending_.size(); i += 3) {
            float a = 0, b = 0;
            HIP_OK(hipEventElapsedTime(&a, events_[i], events_[i + 1]));
            HIP_OK(hipEventElapsedTime(&b, events_[i + 1], events_[i + 2]));
            times_.row_ms += a, times_.col_ms += b;
            times_.row_launches++, times_.col_launches++;
        }
        pending_.clear();
    }

    int m_, n_, eval_freq_;
    bool use_graph_, profile_ = false;
    Mem mem_;
    hipStream_t st_ = nullptr;
    hipGraphExec_t graph_ = nullptr;
    DevMat<T> A_, AT_;
    T *c, *l, *u, *rl, *ru, *Dr, *Dc, *par;
    T *x, *z, *Ax, *ATz, *xs, *zs, *Axs, *ATzs, *xlast, *zlast;
    T *pv, *pw, *pu;
    double *part_ = nullptr, *out_ = nullptr, *f64buf_ = nullptr, *hout_ = nullptr;
    int nbr_ = 0, nbc_ = 0;
    size_t mem_used_ = 0;
    std::string name_;
    std::vector<double> hl_, hu_;
    std::vector<hipEvent_t> events_;
    std::vector<int> pending_;
    KernelTimes times_;
};

}  // namespace

int device_count() {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) return 0;
    return count;
}

std::unique_ptr<Device> make_device_raw(int m, int n, int64_t nnz, const int32_t* a_ptr, const int32_t* a_idx,
                                        const double* a_val, const int32_t* t_ptr, const int32_t* t_idx,
                                        const double* t_val, const double* c, const double* l, const double* u,
                                        const double* rl, const double* ru, const double* Dr, const double* Dc,
                                        const DeviceOptions& opt) {
    if (m < 0 || n < 0 || nnz < 0 || nnz > INT32_MAX) throw std::runtime_error("problem dimensions out of range");
    if (a_ptr[m] != nnz || t_ptr[n] != nnz) throw std::runtime_error("CSR row pointers do not match nnz");
    if (opt.eval_freq < 1) throw std::runtime_error("eval_freq must be >= 1");
    if (opt.precision_bits == 32)
        return std::unique_ptr<Device>(
            new Dev<float>(m, n, a_ptr, a_idx, a_val, t_ptr, t_idx, t_val, c, l, u, rl, ru, Dr, Dc, opt));
    if (opt.precision_bits == 64)
        return std::unique_ptr<Device>(
            new Dev<double>(m, n, a_ptr, a_idx, a_val, t_ptr, t_idx, t_val, c, l, u, rl, ru, Dr, Dc, opt));
    throw std::runtime_error("precision must be 32 or 64 bits");
}

std::unique_ptr<Device> make_device(const Scaled& P, const DeviceOptions& opt) {
    return make_device_raw(P.m, P.n, P.A.nnz(), P.A.ptr.data(), P.A.idx.data(), P.A.val.data(), P.AT.ptr.data(),
                           P.AT.idx.data(), P.AT.val.data(), P.c.data(), P.l.data(), P.u.data(), P.rl.data(),
                           P.ru.data(), P.Dr.data(), P.Dc.data(), opt);
}

namespace {
template <class T>
void spmv_bench_t(int rows, int cols, const int32_t* ptr, const int32_t* idx, const double* val, const double* v,
                  double* y, int reps, double* mean_ms, double* median_ms) {
    require_device();
    Mem mem;
    hipStream_t st;
    HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{st};
    DevMat<T> a = upload_csr<T>(mem, st, rows, cols, ptr, idx, val);
    T* dv = mem.alloc<T>(size_t(cols));
    T* dy = mem.alloc<T>(size_t(rows));
    double* dy64 = mem.alloc<double>(size_t(rows));
    std::vector<T> hv(v, v + cols);
    if (cols) HIP_OK(hipMemcpyAsync(dv, hv.data(), size_t(cols) * sizeof(T), hipMemcpyHostToDevice, st));
    launch_rows(a, dv, EpiStore<T>{dy}, st);  // warm-up
    HIP_OK(hipStreamSynchronize(st));
    std::vector<hipEvent_t> ev(size_t(2 * std::max(reps, 1)));
    for (auto& e : ev) HIP_OK(hipEventCreate(&e));
    for (int r = 0; r < reps; ++r) {
        HIP_OK(hipEventRecord(ev[size_t(2 * r)], st));
        launch_rows(a, dv, EpiStore<T>{dy}, st);
        HIP_OK(hipEventRecord(ev[size_t(2 * r + 1)], st));
    }
    HIP_OK(hipStreamSynchronize(st));
    HIP_OK(hipGetLastError());
    std::vector<double> ts;
    for (int r = 0; r < reps; ++r) {
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, ev[size_t(2 * r)], ev[size_t(2 * r + 1)]));
        ts.push_back(ms);
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    double sum = 0;
    for (double t : ts) sum += t;
    std::sort(ts.begin(), ts.end());
    *mean_ms = reps ? sum / reps : 0;
    *median_ms = reps ? ts[ts.size() / 2] : 0;
    k_to_f64<T><<<grid_for(rows), kWG, 0, st>>>(rows, 1.0, dy, dy64);
    if (rows) HIP_OK(hipMemcpyAsync(y, dy64, size_t(rows) * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
}
}  // namespace

// Streaming ceiling of the device, measured with this library's own kernels: a read-only pass (each workgroup sums
// its grid-stride share of 16-byte words, one float per workgroup out) and a 16-byte-word copy. Four independent
// loads per thread per round keep enough requests in flight; the grid covers every CU several times.
__global__ __launch_bounds__(kWG) void k_stream_read(int64_t n4, const float4* __restrict__ a, float* __restrict__ out) {
    float s = 0;
    const int64_t stride = int64_t(gridDim.x) * kWG;
    int64_t i = int64_t(blockIdx.x) * kWG + threadIdx.x;
    for (; i + 3 * stride < n4; i += 4 * stride) {
        const float4 v0 = a[i], v1 = a[i + stride], v2 = a[i + 2 * stride], v3 = a[i + 3 * stride];
        s += (v0.x + v0.y + v0.z + v0.w) + (v1.x + v1.y + v1.z + v1.w) + (v2.x + v2.y + v2.z + v2.w) +
             (v3.x + v3.y + v3.z + v3.w);
    }
    for (; i < n4; i += stride) {
        const float4 v = a[i];
        s += v.x + v.y + v.z + v.w;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    __shared__ float red[kWG / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(kWG) void k_stream_copy(int64_t n4, const float4* __restrict__ a, float4* __restrict__ b) {
    const int64_t stride = int64_t(gridDim.x) * kWG;
    int64_t i = int64_t(blockIdx.x) * kWG + threadIdx.x;
    for (; i + 3 * stride < n4; i += 4 * stride) {
        const float4 v0 = a[i], v1 = a[i + stride], v2 = a[i + 2 * stride], v3 = a[i + 3 * stride];
        b[i] = v0, b[i + stride] = v1, b[i + 2 * stride] = v2, b[i + 3 * stride] = v3;
    }
    for (; i < n4; i += stride) b[i] = a[i];
}

void stream_bench(int64_t bytes, int reps, double* read_gbs, double* copy_gbs) {
    require_device();
    Mem mem;
    hipStream_t st;
    HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{st};
    const int64_t n4 = bytes / 16;
    const int grid = 8192;  // 32 workgroups per CU on 256 CUs
    float4* a = mem.alloc<float4>(size_t(n4));
    float4* b = mem.alloc<float4>(size_t(n4));
    float* out = mem.alloc<float>(size_t(grid));
    HIP_OK(hipMemsetAsync(a, 0, size_t(n4) * 16, st));
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    auto median_ms = [&](auto&& launch) {
        launch();  // warm-up
        std::vector<float> ts;
        for (int r = 0; r < reps; ++r) {
            HIP_OK(hipEventRecord(e0, st));
            launch();
            HIP_OK(hipEventRecord(e1, st));
            HIP_OK(hipEventSynchronize(e1));
            float ms = 0;
            HIP_OK(hipEventElapsedTime(&ms, e0, e1));
            ts.push_back(ms);
        }
        std::sort(ts.begin(), ts.end());
        return double(ts[ts.size() / 2]);
    };
    const double tr = median_ms([&] { k_stream_read<<<grid, kWG, 0, st>>>(n4, a, out); });
    const double tc = median_ms([&] { k_stream_copy<<<grid, kWG, 0, st>>>(n4, a, b); });
    HIP_OK(hipGetLastError());
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *read_gbs = double(n4) * 16 / (tr * 1e-3) / 1e9;
    *copy_gbs = 2.0 * double(n4) * 16 / (tc * 1e-3) / 1e9;
}

void spmv_bench(int precision_bits, int rows, int cols, const int32_t* ptr, const int32_t* idx, const double* val,
                const double* v, double* y, int reps, double* mean_ms, double* median_ms) {
    if (precision_bits == 32) spmv_bench_t<float>(rows, cols, ptr, idx, val, v, y, reps, mean_ms, median_ms);
    else if (precision_bits == 64) spmv_bench_t<double>(rows, cols, ptr, idx, val, v, y, reps, mean_ms, median_ms);
    else throw std::runtime_error("precision must be 32 or 64 bits");
}

}  // namespace tpdhg
