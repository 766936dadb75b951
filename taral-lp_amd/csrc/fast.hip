// Fast-path device backend: reflected Halpern PDHG on a problem recentred around an FP64 base point.
//
// Per iteration two fused kernels, both on the LDS-staged row-block SpMV of hipcommon.hpp:
//   row kernel (over A):    a = (A w)_i with w = 2 x+ - x; z+ = prox(z + sigma a); z <- Halpern blend of z+, z, z0
//   column kernel (over A^T): a = (A^T z)_j; x <- Halpern blend of x+, x, x0; next x+ = proj(x - tau (c + a));
//                             w = 2 x+ - x (the row kernel's gather vector)
// State traffic per iteration: 6 values per row (z, zh, rl, ru, z0 read; z written) and 9 per column (x, x+, x0,
// c, l, u read; x, x+, w written), fewer than the averaged scheme's 10 + 10, and no running sums. The step index
// inside a captured block and a device-side iteration counter give lambda, so one HIP graph serves every block. Evaluation adds one SpMV (A x) per block and an FP64 pass over base + inner.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "hipcommon.hpp"

namespace tpdhg {
namespace {

template <class T>
__device__ __forceinline__ T halpern_lambda(const int* kc, int s) {
    const int k = *kc + s;  // Halpern steps taken since the last restart
    return T(double(k + 1) / double(k + 2));
}

// Row kernel epilogue: dual prox step on the recentred rows (shift zh = base dual), then the Halpern blend.
template <class T>
struct EpiHRow {
    T* z;
    const T *z0, *zh, *rl, *ru, *par;
    const int* kc;
    int s;
    T rho, sig, lam;
    struct Pre {
        T zi, h, lo, hi, z0;
    };
    __device__ void load() { sig = par[1], lam = halpern_lambda<T>(kc, s); }
    __device__ Pre pre(int i) const { return Pre{z[i], zh[i], rl[i], ru[i], z0[i]}; }
    __device__ void apply(int i, T a, const Pre& p) const {
        const T v = p.zi + sig * a, q = (v + p.h) / sig;
        const T zp = q < p.lo ? v - sig * p.lo : (q > p.hi ? v - sig * p.hi : -p.h);
        z[i] = lam * (rho * zp - (rho - 1) * p.zi) + (1 - lam) * p.z0;
    }
    __device__ void operator()(int i, T a) const { apply(i, a, pre(i)); }
};

// Column kernel epilogue: Halpern blend of the primal (mode 0) or none (mode 1, first step after a restart), then
// the next primal step and the gather vector w = 2 x+ - x; A^T z is stored when ATz is given.
template <class T>
struct EpiHCol {
    T *x, *xp, *w, *ATz;
    const T *x0, *c, *l, *u, *par;
    const int* kc;
    int s, mode;
    T rho, tau, lam;
    struct Pre {
        T xj, xpj, x0j, c, l, u;
    };
    __device__ void load() { tau = par[0], lam = halpern_lambda<T>(kc, s); }
    __device__ Pre pre(int j) const {
        if (mode == 0) return Pre{x[j], xp[j], x0[j], c[j], l[j], u[j]};
        return Pre{x[j], T(0), T(0), c[j], l[j], u[j]};
    }
    __device__ void apply(int j, T a, const Pre& p) const {
        T xj = p.xj;
        if (mode == 0) {
            xj = lam * (rho * p.xpj - (rho - 1) * xj) + (1 - lam) * p.x0j;
            x[j] = xj;
        }
        const T xn = clampv(xj - tau * (p.c + a), p.l, p.u);
        xp[j] = xn, w[j] = 2 * xn - xj;
        if (ATz) ATz[j] = a;
    }
    __device__ void operator()(int j, T a) const { apply(j, a, pre(j)); }
};

__global__ void k_advance(int* kc, int steps) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *kc += steps;
}

__global__ void k_set_int(int* p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// Inner problem of a round from the FP64 base (rows): bounds shifted by A x_b, prox shift z_b, dz = 0.
template <class T>
__global__ __launch_bounds__(kWG) void k_begin_rows(int m, const double* __restrict__ rl, const double* __restrict__ ru,
                                                    const double* __restrict__ Axb, const double* __restrict__ zb,
                                                    T* __restrict__ rl_in, T* __restrict__ ru_in, T* __restrict__ zh,
                                                    T* __restrict__ z, T* __restrict__ z0) {
    for (int i = int(blockIdx.x * blockDim.x + threadIdx.x); i < m; i += int(gridDim.x * blockDim.x)) {
        rl_in[i] = T(rl[i] - Axb[i]), ru_in[i] = T(ru[i] - Axb[i]), zh[i] = T(zb[i]);
        z[i] = 0, z0[i] = 0;
    }
}

// Inner problem of a round (columns): reduced cost c + A^T z_b, bounds shifted by x_b, dx = clamp(0, l_in, u_in).
template <class T>
__global__ __launch_bounds__(kWG) void k_begin_cols(int n, const double* __restrict__ c, const double* __restrict__ l,
                                                    const double* __restrict__ u, const double* __restrict__ ATzb,
                                                    const double* __restrict__ xb, T* __restrict__ c_in,
                                                    T* __restrict__ l_in, T* __restrict__ u_in, T* __restrict__ x,
                                                    T* __restrict__ x0) {
    for (int j = int(blockIdx.x * blockDim.x + threadIdx.x); j < n; j += int(gridDim.x * blockDim.x)) {
        const T lo = T(l[j] - xb[j]), hi = T(u[j] - xb[j]);
        c_in[j] = T(c[j] + ATzb[j]), l_in[j] = lo, u_in[j] = hi;
        const T v = clampv(T(0), lo, hi);
        x[j] = v, x0[j] = v;
    }
}

template <class T>
__global__ __launch_bounds__(kWG) void k_accumulate(int k, double* __restrict__ base, const T* __restrict__ inner) {
    for (int i = int(blockIdx.x * blockDim.x + threadIdx.x); i < k; i += int(gridDim.x * blockDim.x))
        base[i] += double(inner[i]);
}

template <class T>
__global__ __launch_bounds__(kWG) void k_combine(int k, const double* __restrict__ base, const T* __restrict__ inner,
                                                 double* __restrict__ out) {
    for (int i = int(blockIdx.x * blockDim.x + threadIdx.x); i < k; i += int(gridDim.x * blockDim.x))
        out[i] = base[i] + double(inner[i]);
}

template <class T>
struct FastEvalArgs {
    const T *Ax, *z, *z0, *x, *ATz, *x0;             // inner point and anchors
    const double *Axb, *zb, *xb, *ATzb;              // FP64 base and its products
    const double *rl, *ru, *Dr, *c, *l, *u, *Dc;     // original scaled problem, FP64
    const T *Aw, *xp, *zh, *rl_in, *ru_in, *par;     // fixed-point residual: A w, x+, the inner row prox data
};

// KKT scalars of base + inner (the ten slots of the averaged scheme's evaluation, gpu/pdhg.cu:236-271), FP64,
// and the fixed-point residual terms of the inner point. Blocks [0, nbr) cover rows and write PR2S..DZ2, then
// ||z - z+||^2 and (z - z+)' A (x - x+), with A (x - x+) = (A x - A w) / 2 because w = 2 x+ - x; blocks
// [nbr, grid) cover columns and write DR2S..DX2, then ||x - x+||^2 and 0. x+ is the primal step the last column
// step took from the current point and z+ = prox(z + sigma A w) the dual step that follows it, so (x+, z+) = T(u).
template <class T>
__global__ __launch_bounds__(kWG) void k_eval_fast(int m, int n, int nbr, FastEvalArgs<T> a,
                                                   double* __restrict__ part) {
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    const int b = int(blockIdx.x);
    if (b < nbr) {
        const T sig = a.par[1];
        for (int i = b * kWG + int(threadIdx.x); i < m; i += nbr * kWG) {
            const T zp = row_prox(a.z[i], a.Aw[i], a.zh[i], a.rl_in[i], a.ru_in[i], sig);
            const double dzt = double(a.z[i]) - double(zp), adx = 0.5 * (double(a.Ax[i]) - double(a.Aw[i]));
            v[5] += dzt * dzt, v[6] += dzt * adx;
            const double ax = a.Axb[i] + double(a.Ax[i]), zz = a.zb[i] + double(a.z[i]);
            const double lo = a.rl[i], up = a.ru[i], dr = a.Dr[i];
            const double r = ax < lo ? lo - ax : (ax > up ? ax - up : 0), ro = r / dr;
            const double rr = ro * row_rel_weight(lo, up, dr);
            v[0] += r * r, v[1] += ro * ro, v[2] += rr * rr;
            if (zz > 0 && up < INFINITY) v[3] -= zz * up;
            if (zz < 0 && lo > -INFINITY) v[3] -= zz * lo;
            const double dz = double(a.z[i]) - double(a.z0[i]);
            v[4] += dz * dz;
        }
    } else {
        const int nbc = int(gridDim.x) - nbr;
        for (int j = (b - nbr) * kWG + int(threadIdx.x); j < n; j += nbc * kWG) {
            const double xx = a.xb[j] + double(a.x[j]), cj = a.c[j];
            const double d = cj + a.ATzb[j] + double(a.ATz[j]), lo = a.l[j], up = a.u[j];
            double res = 0;
            v[2] += cj * xx;
            if (d > 0) {
                if (lo > -INFINITY) v[3] += d * lo;
                else res = d;
            } else if (d < 0) {
                if (up < INFINITY) v[3] += d * up;
                else res = -d;
            }
            const double ro = res / a.Dc[j], dx = double(a.x[j]) - double(a.x0[j]);
            v[0] += res * res, v[1] += ro * ro, v[4] += dx * dx;
            const double dxt = double(a.x[j]) - double(a.xp[j]);
            v[5] += dxt * dxt;
        }
    }
    block_store<7>(v, part);
}

template <class T>
class FastDev final : public FastDevice {
public:
    FastDev(const Scaled& P, const FastOptions& o) : m_(P.m), n_(P.n), o_(o) {
        require_device();
        if (!(o.rho >= 1 && o.rho < 2)) throw std::runtime_error("reflection coefficient rho must be in [1, 2)");
        size_t free0 = 0, free1 = 0, total = 0;
        HIP_OK(hipMemGetInfo(&free0, &total));
        name_ = device_name();
        HIP_OK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
        {
            const HostLayout LA = choose_layout<T>(st_, m_, n_, P.A.ptr.data(), P.A.idx.data(), P.A.val.data());
            const HostLayout LT = choose_layout<T>(st_, n_, m_, P.AT.ptr.data(), P.AT.idx.data(), P.AT.val.data());
            A_ = upload_layout<T>(mem_, st_, m_, LA, P.A.val.data());
            AT_ = upload_layout<T>(mem_, st_, n_, LT, P.AT.val.data());
            A64_ = values64(A_, LA, P.A);
            AT64_ = values64(AT_, LT, P.AT);
        }
        c64 = up64(P.c), l64 = up64(P.l), u64 = up64(P.u), rl64 = up64(P.rl), ru64 = up64(P.ru);
        Dr64 = up64(P.Dr), Dc64 = up64(P.Dc);
        xb = zeros64(n_), zb = zeros64(m_), Axb = zeros64(m_), ATzb = zeros64(n_), buf64 = zeros64(std::max(m_, n_));
        c_in = alloc(n_), l_in = alloc(n_), u_in = alloc(n_), rl_in = alloc(m_), ru_in = alloc(m_), zh = alloc(m_);
        x = alloc(n_), xp = alloc(n_), w = alloc(n_), x0 = alloc(n_), ATz = alloc(n_);
        z = alloc(m_), z0 = alloc(m_), Ax = alloc(m_);
        pv = alloc(n_), pw = alloc(m_), pu = alloc(n_);
        for (auto [p, k] : {std::pair<T*, int>{xp, n_}, {w, n_}, {pw, m_}})  // read by eval before a first step
            HIP_OK(hipMemsetAsync(p, 0, size_t(std::max(k, 1)) * sizeof(T), st_));
        par_ = alloc(2);
        kc_ = mem_.alloc<int>(1);
        nbr_ = red_blocks(m_), nbc_ = red_blocks(n_);
        part_ = mem_.alloc<double>(size_t(std::max(1, nbr_ + nbc_)) * 7 + size_t(kMaxRedBlocks));
        static_assert(kPowerGroup <= 16, "a power-iteration group's sums of squares go to out_");
        out_ = mem_.alloc<double>(16);
        HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&hout_), 16 * sizeof(double), hipHostMallocDefault));
        k_set_int<<<1, 1, 0, st_>>>(kc_, 0);
        HIP_OK(hipStreamSynchronize(st_));
        HIP_OK(hipMemGetInfo(&free1, &total));
        mem_used_ = free0 > free1 ? free0 - free1 : 0;
    }
    ~FastDev() override {
        if (graph_) (void)hipGraphExecDestroy(graph_);
        for (hipEvent_t e : events_) (void)hipEventDestroy(e);
        if (hout_) (void)hipHostFree(hout_);
        if (st_) (void)hipStreamDestroy(st_);
    }

    // The iteration and stopping rule of gpu/pdhg.cu:417-436, with each step normalised on the device so that
    // kPowerGroup steps run between two host reads of their sums of squares; the estimate is the one of the step
    // the rule stops at (the group's later steps are discarded).
    double power_norm(int iters) override {
        std::vector<double> h(static_cast<size_t>(n_));
        double nn = 0;
        for (int j = 0; j < n_; ++j) {
            h[size_t(j)] = 0.5 + double((uint32_t(j) * 2654435761u) % 1000u) / 1000.0;
            nn += h[size_t(j)] * h[size_t(j)];
        }
        std::vector<T> ht(static_cast<size_t>(n_));
        for (int j = 0; j < n_; ++j) ht[size_t(j)] = T(h[size_t(j)] / std::sqrt(nn));
        h2d(pv, ht.data(), size_t(n_));
        const int b = red_blocks(n_);
        if (!b) return 0;
        double est = 0;
        for (int it0 = 0; it0 < iters; it0 += kPowerGroup) {
            const int g = std::min(kPowerGroup, iters - it0);
            for (int q = 0; q < g; ++q) {
                launch_rows(A_, pv, EpiStore<T>{pw}, st_);
                launch_rows(AT_, pw, EpiStore<T>{pu}, st_);
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

    void set_base(const double* x_b, const double* z_b) override {
        if (x_b && n_) HIP_OK(hipMemcpyAsync(xb, x_b, size_t(n_) * sizeof(double), hipMemcpyHostToDevice, st_));
        else if (n_) HIP_OK(hipMemsetAsync(xb, 0, size_t(n_) * sizeof(double), st_));
        if (z_b && m_) HIP_OK(hipMemcpyAsync(zb, z_b, size_t(m_) * sizeof(double), hipMemcpyHostToDevice, st_));
        else if (m_) HIP_OK(hipMemsetAsync(zb, 0, size_t(m_) * sizeof(double), st_));
        HIP_OK(hipStreamSynchronize(st_));
    }

    void begin_round() override {
        launch_rows(A64_, xb, EpiStore<double>{Axb}, st_);
        launch_rows(AT64_, zb, EpiStore<double>{ATzb}, st_);
        k_begin_rows<T><<<grid_for(m_), kWG, 0, st_>>>(m_, rl64, ru64, Axb, zb, rl_in, ru_in, zh, z, z0);
        k_begin_cols<T><<<grid_for(n_), kWG, 0, st_>>>(n_, c64, l64, u64, ATzb, xb, c_in, l_in, u_in, x, x0);
        HIP_OK(hipGetLastError());
    }

    void set_steps(double tau, double sigma) override {
        T h[2] = {T(tau), T(sigma)};
        HIP_OK(hipMemcpyAsync(par_, h, sizeof(h), hipMemcpyHostToDevice, st_));
        HIP_OK(hipStreamSynchronize(st_));
    }

    void restart() override {
        if (n_) HIP_OK(hipMemcpyAsync(x0, x, size_t(n_) * sizeof(T), hipMemcpyDeviceToDevice, st_));
        if (m_) HIP_OK(hipMemcpyAsync(z0, z, size_t(m_) * sizeof(T), hipMemcpyDeviceToDevice, st_));
        k_set_int<<<1, 1, 0, st_>>>(kc_, 0);
        launch_rows(AT_, z, col_epi(0, 1, true), st_);  // also stores A^T z of the new anchor
        HIP_OK(hipGetLastError());
    }

    void iterate(int steps) override {
        if (steps <= 0) return;
        if (o_.use_graph && !profile_ && steps == o_.eval_freq) {
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

    void eval(double out[NKKT]) override {
        launch_rows(A_, x, EpiStore<T>{Ax}, st_);  // A x of the inner point; A^T z was stored by the last step
        launch_rows(A_, w, EpiStore<T>{pw}, st_);  // A w, w = 2 x+ - x: the dual half of T(u)
        if (nbr_ + nbc_ > 0) {
            FastEvalArgs<T> a{Ax, z, z0, x, ATz, x0, Axb, zb, xb, ATzb, rl64, ru64, Dr64, c64, l64, u64, Dc64,
                              pw, xp, zh, rl_in, ru_in, par_};
            k_eval_fast<T><<<nbr_ + nbc_, kWG, 0, st_>>>(m_, n_, nbr_, a, part_);
            k_finish<<<1, kWG, 0, st_>>>(0, nbr_, 7, part_, out_);
            k_finish<<<1, kWG, 0, st_>>>(nbr_, nbc_, 7, part_, out_ + 7);
            HIP_OK(hipMemcpyAsync(hout_, out_, 14 * sizeof(double), hipMemcpyDeviceToHost, st_));
            HIP_OK(hipStreamSynchronize(st_));
        } else {
            std::fill(hout_, hout_ + 14, 0.0);
        }
        HIP_OK(hipGetLastError());
        std::copy(hout_, hout_ + 5, out), std::copy(hout_ + 7, hout_ + 12, out + 5);
        fp_[0] = hout_[12], fp_[1] = hout_[5], fp_[2] = hout_[6];
        collect_events();
    }

    void fixed_point(double out[3]) const override { std::copy(fp_, fp_ + 3, out); }

    void end_round() override {
        k_accumulate<T><<<grid_for(n_), kWG, 0, st_>>>(n_, xb, x);
        k_accumulate<T><<<grid_for(m_), kWG, 0, st_>>>(m_, zb, z);
        HIP_OK(hipStreamSynchronize(st_));
        HIP_OK(hipGetLastError());
    }

    void download(double* hx, double* hz) override {
        if (hx && n_) {
            k_combine<T><<<grid_for(n_), kWG, 0, st_>>>(n_, xb, x, buf64);
            HIP_OK(hipMemcpyAsync(hx, buf64, size_t(n_) * sizeof(double), hipMemcpyDeviceToHost, st_));
            HIP_OK(hipStreamSynchronize(st_));
        }
        if (hz && m_) {
            k_combine<T><<<grid_for(m_), kWG, 0, st_>>>(m_, zb, z, buf64);
            HIP_OK(hipMemcpyAsync(hz, buf64, size_t(m_) * sizeof(double), hipMemcpyDeviceToHost, st_));
            HIP_OK(hipStreamSynchronize(st_));
        }
        HIP_OK(hipGetLastError());
    }

    void set_vectors(const double* c_, const double* l_, const double* u_, const double* rl_, const double* ru_) override {
        auto put = [&](double* dst, const double* src, int k) {
            if (k) HIP_OK(hipMemcpyAsync(dst, src, size_t(k) * sizeof(double), hipMemcpyHostToDevice, st_));
        };
        put(c64, c_, n_), put(l64, l_, n_), put(u64, u_, n_), put(rl64, rl_, m_), put(ru64, ru_, m_);
        HIP_OK(hipStreamSynchronize(st_));
    }
    void set_profile(bool on) override { profile_ = on; }
    KernelTimes kernel_times() override {
        HIP_OK(hipStreamSynchronize(st_));
        collect_events();
        KernelTimes t = times_;
        times_ = KernelTimes();
        return t;
    }
    const std::string& name() const override { return name_; }
    size_t mem_used() const override { return mem_used_; }
    int precision_bits() const override { return int(sizeof(T) * 8); }

private:
    T* alloc(int k) { return mem_.alloc<T>(size_t(std::max(k, 1))); }
    double* zeros64(int k) {
        double* p = mem_.alloc<double>(size_t(std::max(k, 1)));
        HIP_OK(hipMemsetAsync(p, 0, size_t(std::max(k, 1)) * sizeof(double), st_));
        return p;
    }
    double* up64(const std::vector<double>& v) {
        double* p = mem_.alloc<double>(std::max<size_t>(v.size(), 1));
        if (!v.empty()) HIP_OK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, st_));
        HIP_OK(hipStreamSynchronize(st_));
        return p;
    }
    // FP64 copy of a matrix's values sharing its pointers and blocks (the same buffers when T is double).
    DevMat<double> values64(const DevMat<T>& a, const HostLayout& L, const Csr& h) {
        if constexpr (sizeof(T) == sizeof(double)) {
            DevMat<double> d;
            d.rows = a.rows, d.nblk = a.nblk, d.maxblk = a.maxblk, d.slab = a.slab;
            d.ptr = a.ptr, d.idx = a.idx, d.blk = a.blk, d.snblk = a.snblk, d.sboff = a.sboff;
            d.val = reinterpret_cast<double*>(a.val), d.part = reinterpret_cast<double*>(a.part);
            d.nch = a.nch, d.nlr = a.nlr, d.chunk = a.chunk, d.lrow = a.lrow;
            d.lpart = reinterpret_cast<double*>(a.lpart);
            d.sorted = a.sorted, d.cbits = a.cbits, d.sb = a.sb, d.sstart = a.sstart, d.nslab = a.nslab;
            d.bk = a.bk, d.roff = a.roff;
            return d;
        } else {
            return share_structure<double>(mem_, st_, a, L, h.val.data());
        }
    }
    void h2d(T* dst, const T* src, size_t k) {
        if (!k) return;
        HIP_OK(hipMemcpyAsync(dst, src, k * sizeof(T), hipMemcpyHostToDevice, st_));
        HIP_OK(hipStreamSynchronize(st_));
    }
    EpiHCol<T> col_epi(int s, int mode, bool store) {
        return EpiHCol<T>{x, xp, w, store ? ATz : nullptr, x0, c_in, l_in, u_in, par_, kc_, s, mode,
                          T(o_.rho), T(0), T(0)};
    }
    EpiHRow<T> row_epi(int s) { return EpiHRow<T>{z, z0, zh, rl_in, ru_in, par_, kc_, s, T(o_.rho), T(0), T(0)}; }
    hipEvent_t event(size_t i) {
        while (events_.size() <= i) {
            hipEvent_t e;
            HIP_OK(hipEventCreate(&e));
            events_.push_back(e);
        }
        return events_[i];
    }
    void launch_block(int steps) {
        for (int s = 0; s < steps; ++s) {
            const EpiHRow<T> er = row_epi(s);
            const EpiHCol<T> ec = col_epi(s, 0, s + 1 == steps);
            if (profile_) {
                const size_t e = pending_;
                HIP_OK(hipEventRecord(event(e), st_));
                launch_rows(A_, w, er, st_);
                HIP_OK(hipEventRecord(event(e + 1), st_));
                launch_rows(AT_, z, ec, st_);
                HIP_OK(hipEventRecord(event(e + 2), st_));
                pending_ += 3;
            } else {
                launch_rows(A_, w, er, st_);
                launch_rows(AT_, z, ec, st_);
            }
        }
        k_advance<<<1, 1, 0, st_>>>(kc_, steps);
    }
    void collect_events() {
        if (!pending_) return;
        HIP_OK(hipStreamSynchronize(st_));
        for (size_t i = 0; i + 2 < pending_; i += 3) {
            float a = 0, b = 0;
            HIP_OK(hipEventElapsedTime(&a, events_[i], events_[i + 1]));
            HIP_OK(hipEventElapsedTime(&b, events_[i + 1], events_[i + 2]));
            times_.row_ms += a, times_.col_ms += b;
            times_.row_launches++, times_.col_launches++;
        }
        pending_ = 0;
    }

    int m_, n_;
    FastOptions o_;
    bool profile_ = false;
    Mem mem_;
    hipStream_t st_ = nullptr;
    hipGraphExec_t graph_ = nullptr;
    DevMat<T> A_, AT_;
    DevMat<double> A64_, AT64_;
    double *c64, *l64, *u64, *rl64, *ru64, *Dr64, *Dc64;
    double *xb, *zb, *Axb, *ATzb, *buf64;
    T *c_in, *l_in, *u_in, *rl_in, *ru_in, *zh;
    T *x, *xp, *w, *x0, *ATz, *z, *z0, *Ax;
    T *pv, *pw, *pu, *par_;
    int* kc_ = nullptr;
    double *part_ = nullptr, *out_ = nullptr, *hout_ = nullptr;
    double fp_[3] = {0, 0, 0};  // fixed-point residual terms of the last eval
    int nbr_ = 0, nbc_ = 0;
    size_t mem_used_ = 0;
    std::string name_;
    std::vector<hipEvent_t> events_;
    size_t pending_ = 0;
    KernelTimes times_;
};

}  // namespace

std::unique_ptr<FastDevice> make_fast_device(const Scaled& P, const FastOptions& opt) {
    if (P.m < 0 || P.n < 0 || P.A.nnz() > INT32_MAX) throw std::runtime_error("problem dimensions out of range");
    if (opt.eval_freq < 1) throw std::runtime_error("eval_freq must be >= 1");
    if (opt.precision_bits == 32) return std::unique_ptr<FastDevice>(new FastDev<float>(P, opt));
    if (opt.precision_bits == 64) return std::unique_ptr<FastDevice>(new FastDev<double>(P, opt));
    throw std::runtime_error("precision must be 32 or 64 bits");
}

}  // namespace tpdhg
