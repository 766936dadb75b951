// CPU backend of the reference scheme: the reference's Cpu<T> (gpu/pdhg.cu:517-741, selected by `--device cpu
// --threads N`), the same iteration, evaluation and restart arithmetic as the HIP backend (device.hip) on host
// threads. Each row of A (and of A^T) is summed serially in CSR order by the thread owning it, so the iterates do
// not depend on the thread count; the KKT and power-iteration sums are taken over fixed chunks of rows, each chunk
// summed in order and the chunk sums added in chunk order, so they do not depend on it either. Rows are split
// between threads by rows plus nonzeros (a row costs one unit, a nonzero one more), contiguous ranges.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <exception>
#include <cstdio>
#include <functional>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "device.hpp"

namespace tpdhg {
namespace {

// A fixed team of host threads: run(f) calls f(t) for t in [0, size) (the caller is thread 0) and returns when every
// call has returned. Idle workers poll a generation counter for a short while and then sleep on a condition
// variable, so back-to-back phases of an iteration do not pay a wake-up while a long pause costs no CPU.
class Team {
public:
    explicit Team(int n) : n_(std::max(1, n)) {
        for (int t = 1; t < n_; ++t) threads_.emplace_back([this, t] { work(t); });
    }
    ~Team() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            gen_.fetch_add(1, std::memory_order_release);
        }
        cv_.notify_all();
        for (std::thread& th : threads_) th.join();
    }
    int size() const { return n_; }
    void run(const std::function<void(int)>& f) {
        if (n_ == 1) {
            f(0);
            return;
        }
        job_ = &f;
        left_.store(n_ - 1, std::memory_order_relaxed);
        {
            std::lock_guard<std::mutex> lk(mu_);
            gen_.fetch_add(1, std::memory_order_release);
        }
        cv_.notify_all();
        std::exception_ptr mine;
        try {
            f(0);
        } catch (...) {
            mine = std::current_exception();
        }
        while (left_.load(std::memory_order_acquire) != 0) std::this_thread::yield();  // f outlives every use
        if (mine) std::rethrow_exception(mine);
        if (err_) {  // a worker's exception, rethrown on the calling thread
            std::exception_ptr e = err_;
            err_ = nullptr;
            std::rethrow_exception(e);
        }
    }

private:
    void work(int t) {
        unsigned seen = 0;
        for (;;) {
            unsigned g = gen_.load(std::memory_order_acquire);
            for (int k = 0; g == seen && k < 20000; ++k) g = gen_.load(std::memory_order_acquire);
            if (g == seen) {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_.load(std::memory_order_acquire) != seen; });
                g = gen_.load(std::memory_order_acquire);
            }
            seen = g;
            if (stop_) return;
            try {
                (*job_)(t);
            } catch (...) {
                std::lock_guard<std::mutex> lk(mu_);
                if (!err_) err_ = std::current_exception();
            }
            left_.fetch_sub(1, std::memory_order_acq_rel);
        }
    }
    int n_;
    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::atomic<unsigned> gen_{0};
    std::atomic<int> left_{0};
    const std::function<void(int)>* job_ = nullptr;
    std::exception_ptr err_;  // first exception of a worker in the current run (written under mu_)
    bool stop_ = false;
};

constexpr int kChunk = 4096;  // rows per reduction chunk

template <class T>
T clampv(T v, T lo, T hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

template <class T>
struct HostCsr {
    int rows = 0;
    std::vector<int32_t> ptr, idx;
    std::vector<T> val;
    std::vector<int> part;  // thread t owns rows [part[t], part[t + 1])
    T dot(int i, const T* v) const {
        T s = 0;
        for (int32_t k = ptr[size_t(i)]; k < ptr[size_t(i) + 1]; ++k) s += val[size_t(k)] * v[idx[size_t(k)]];
        return s;
    }
};

// Contiguous row ranges of about equal rows + nonzeros.
std::vector<int> split_rows(int rows, const int32_t* ptr, int nt) {
    std::vector<int> part(size_t(nt) + 1, rows);
    part[0] = 0;
    const double total = double(rows) + double(ptr[rows]);
    int r = 0;
    for (int t = 1; t < nt; ++t) {
        const double target = total * t / nt;
        while (r < rows && double(r) + double(ptr[r]) < target) ++r;
        part[size_t(t)] = r;
    }
    return part;
}

template <class T>
class CpuDev final : public Device {
public:
    CpuDev(const Scaled& P, int threads) : m_(P.m), n_(P.n), team_(threads) {
        load(A_, P.A, threads);
        load(AT_, P.AT, threads);
        cT(c_, P.c), cT(l_, P.l), cT(u_, P.u), cT(rl_, P.rl), cT(ru_, P.ru), cT(Dr_, P.Dr), cT(Dc_, P.Dc);
        hl_ = P.l, hu_ = P.u;
        for (auto* v : {&x_, &ATz_, &xs_, &ATzs_, &xl_, &pv_, &pu_}) v->assign(size_t(n_), T(0));
        for (auto* v : {&z_, &Ax_, &zs_, &Axs_, &zl_, &pw_}) v->assign(size_t(m_), T(0));
        char nm[64];
        std::snprintf(nm, sizeof nm, "cpu x%d threads", team_.size());
        name_ = nm;
        mem_ = (A_.val.size() + AT_.val.size()) * (sizeof(T) + 4) + size_t(m_ + n_) * 16 * sizeof(T);
    }

    double power_norm(int iters) override {  // same start vector and stopping rule as the HIP backend
        double nn = 0;
        std::vector<double> h(static_cast<size_t>(n_));
        for (int j = 0; j < n_; ++j) {
            h[size_t(j)] = 0.5 + double((uint32_t(j) * 2654435761u) % 1000u) / 1000.0;
            nn += h[size_t(j)] * h[size_t(j)];
        }
        for (int j = 0; j < n_; ++j) pv_[size_t(j)] = T(h[size_t(j)] / std::sqrt(nn));
        double est = 0;
        for (int it = 0; it < iters; ++it) {
            product(A_, pv_.data(), pw_.data());
            product(AT_, pw_.data(), pu_.data());
            const double s = sumsq(pu_), prev = est;
            if (s <= 0) return 0;
            est = std::sqrt(std::sqrt(s));
            const double f = 1 / std::sqrt(s);
            team_.run([&](int t) {
                for (int j = AT_.part[size_t(t)]; j < AT_.part[size_t(t) + 1]; ++j) pu_[size_t(j)] = T(pu_[size_t(j)] * f);
            });
            std::swap(pv_, pu_);
            if (it >= 10 && std::fabs(est - prev) <= 1e-7 * est) break;
        }
        return est;
    }

    void set_steps(double tau, double sigma) override { tau_ = T(tau), sig_ = T(sigma); }

    void start(const double* x_init, const double* z_init) override {
        for (int j = 0; j < n_; ++j) {
            const double v = x_init ? x_init[j] : 0.0;
            x_[size_t(j)] = xl_[size_t(j)] = T(std::min(std::max(v, hl_[size_t(j)]), hu_[size_t(j)]));
        }
        for (int i = 0; i < m_; ++i) z_[size_t(i)] = zl_[size_t(i)] = z_init ? T(z_init[i]) : T(0);
        std::fill(xs_.begin(), xs_.end(), T(0)), std::fill(ATzs_.begin(), ATzs_.end(), T(0));
        std::fill(zs_.begin(), zs_.end(), T(0)), std::fill(Axs_.begin(), Axs_.end(), T(0));
        products();
    }

    // First primal step from the stored A^T z, then `steps` x (dual row step; column step fused with the next
    // primal step except after the last), as device.hip's launch_block.
    void iterate(int steps) override {
        if (steps <= 0) return;
        const T tau = tau_, sig = sig_;
        team_.run([&](int t) {
            for (int j = AT_.part[size_t(t)]; j < AT_.part[size_t(t) + 1]; ++j) {
                const size_t q = size_t(j);
                const T xn = clampv(x_[q] - tau * (c_[q] + ATz_[q]), l_[q], u_[q]);
                x_[q] = xn, xs_[q] += xn;
            }
        });
        for (int s = 0; s < steps; ++s) {
            team_.run([&](int t) {
                for (int i = A_.part[size_t(t)]; i < A_.part[size_t(t) + 1]; ++i) {
                    const size_t q = size_t(i);
                    const T a = A_.dot(i, x_.data());
                    const T v = z_[q] + sig * (2 * a - Ax_[q]), r = v / sig;
                    const T lo = rl_[q], hi = ru_[q];
                    const T zn = r < lo ? v - sig * lo : (r > hi ? v - sig * hi : T(0));
                    Ax_[q] = a, z_[q] = zn, Axs_[q] += a, zs_[q] += zn;
                }
            });
            const bool upd = s + 1 < steps;
            team_.run([&](int t) {
                for (int j = AT_.part[size_t(t)]; j < AT_.part[size_t(t) + 1]; ++j) {
                    const size_t q = size_t(j);
                    const T a = AT_.dot(j, z_.data());
                    ATz_[q] = a, ATzs_[q] += a;
                    if (upd) {
                        const T xn = clampv(x_[q] - tau * (c_[q] + a), l_[q], u_[q]);
                        x_[q] = xn, xs_[q] += xn;
                    }
                }
            });
        }
    }

    // KKT scalars of the current point and of the f-scaled average (device.hip k_eval2), FP64, chunked sums.
    void eval2(double f, double* cur, double* avg) override {
        const int rc = (m_ + kChunk - 1) / kChunk, cc = (n_ + kChunk - 1) / kChunk;
        std::vector<double> part(size_t(rc + cc) * 10, 0.0);
        team_.run([&](int t) {
            for (int k = t; k < rc + cc; k += team_.size()) {
                double* o = part.data() + size_t(k) * 10;
                if (k < rc) {
                    for (int i = k * kChunk; i < std::min(m_, (k + 1) * kChunk); ++i) row_terms(i, f, o);
                } else {
                    const int kk = k - rc;
                    for (int j = kk * kChunk; j < std::min(n_, (kk + 1) * kChunk); ++j) col_terms(j, f, o);
                }
            }
        });
        double out[20] = {0};
        for (int k = 0; k < rc; ++k)
            for (int q = 0; q < 10; ++q) out[q] += part[size_t(k) * 10 + size_t(q)];
        for (int k = rc; k < rc + cc; ++k)
            for (int q = 0; q < 10; ++q) out[10 + q] += part[size_t(k) * 10 + size_t(q)];
        // rows: current PR2S..DZ2 then average; columns: current DR2S..DX2 then average
        if (cur) std::copy(out, out + 5, cur), std::copy(out + 10, out + 15, cur + 5);
        if (avg) std::copy(out + 5, out + 10, avg), std::copy(out + 15, out + 20, avg + 5);
    }

    void restart(bool avg, double f) override {
        auto reset = [&](std::vector<T>& v, std::vector<T>& vs, std::vector<T>& vl, std::vector<T>& ps) {
            for (size_t q = 0; q < v.size(); ++q) {
                const T w = avg ? T(vs[q] * f) : v[q];
                v[q] = w, vl[q] = w, vs[q] = 0, ps[q] = 0;
            }
        };
        reset(x_, xs_, xl_, ATzs_);
        reset(z_, zs_, zl_, Axs_);
        products();
    }

    void download(bool avg, double f, double* x, double* z) override {
        if (x)
            for (int j = 0; j < n_; ++j) x[j] = double(avg ? xs_[size_t(j)] : x_[size_t(j)]) * (avg ? f : 1.0);
        if (z)
            for (int i = 0; i < m_; ++i) z[i] = double(avg ? zs_[size_t(i)] : z_[size_t(i)]) * (avg ? f : 1.0);
    }

    void spmv(int which, const double* v, double* y) override {
        const HostCsr<T>& a = which ? AT_ : A_;
        std::vector<T> in(v, v + (which ? m_ : n_)), out(static_cast<size_t>(a.rows));
        product(a, in.data(), out.data());
        for (int i = 0; i < a.rows; ++i) y[i] = double(out[size_t(i)]);
    }

    void set_profile(bool) override {}
    KernelTimes kernel_times() override { return KernelTimes(); }
    void set_vectors(const double* c, const double* l, const double* u, const double* rl, const double* ru) override {
        for (int j = 0; j < n_; ++j) c_[size_t(j)] = T(c[j]), l_[size_t(j)] = T(l[j]), u_[size_t(j)] = T(u[j]);
        for (int i = 0; i < m_; ++i) rl_[size_t(i)] = T(rl[i]), ru_[size_t(i)] = T(ru[i]);
        hl_.assign(l, l + n_), hu_.assign(u, u + n_);
    }
    void synchronize() override {}
    const std::string& name() const override { return name_; }
    size_t mem_used() const override { return mem_; }
    int precision_bits() const override { return int(sizeof(T) * 8); }
    int m() const override { return m_; }
    int n() const override { return n_; }

private:
    void load(HostCsr<T>& a, const Csr& src, int threads) {
        a.rows = src.rows, a.ptr = src.ptr, a.idx = src.idx;
        a.val.resize(src.val.size());
        for (size_t k = 0; k < src.val.size(); ++k) a.val[k] = T(src.val[k]);
        a.part = split_rows(src.rows, src.ptr.data(), std::max(1, threads));
    }
    static void cT(std::vector<T>& dst, const std::vector<double>& src) {
        dst.resize(src.size());
        for (size_t k = 0; k < src.size(); ++k) dst[k] = T(src[k]);
    }
    void product(const HostCsr<T>& a, const T* v, T* y) {
        team_.run([&](int t) {
            for (int i = a.part[size_t(t)]; i < a.part[size_t(t) + 1]; ++i) y[i] = a.dot(i, v);
        });
    }
    void products() {  // Ax = A x, ATz = A^T z exactly
        product(A_, x_.data(), Ax_.data());
        product(AT_, z_.data(), ATz_.data());
    }
    double sumsq(const std::vector<T>& v) {
        const int nc = (int(v.size()) + kChunk - 1) / kChunk;
        std::vector<double> part(static_cast<size_t>(nc), 0.0);
        team_.run([&](int t) {
            for (int k = t; k < nc; k += team_.size()) {
                double s = 0;
                for (int j = k * kChunk; j < std::min(int(v.size()), (k + 1) * kChunk); ++j)
                    s += double(v[size_t(j)]) * double(v[size_t(j)]);
                part[size_t(k)] = s;
            }
        });
        double s = 0;
        for (double p : part) s += p;
        return s;
    }
    static double row_rel_weight(double lo, double up, double dr) {
        double b = 0;
        if (lo > -INFINITY) b = lo < 0 ? -lo : lo;
        if (up < INFINITY) b = std::max(b, up < 0 ? -up : up);
        return 1.0 / (1.0 + b / dr);
    }
    void row_terms(int i, double f, double* v) const {
        const size_t q = size_t(i);
        const double lo = rl_[q], up = ru_[q], dr = Dr_[q], zl = zl_[q];
        const double wrel = row_rel_weight(lo, up, dr);
        for (int h = 0; h < 2; ++h) {
            const double ax = h ? Axs_[q] * f : double(Ax_[q]), zz = h ? zs_[q] * f : double(z_[q]);
            const double r = ax < lo ? lo - ax : (ax > up ? ax - up : 0), ro = r / dr, rr = ro * wrel;
            double* o = v + 5 * h;
            o[0] += r * r, o[1] += ro * ro, o[2] += rr * rr;
            if (zz > 0 && up < INFINITY) o[3] -= zz * up;
            if (zz < 0 && lo > -INFINITY) o[3] -= zz * lo;
            const double dz = zz - zl;
            o[4] += dz * dz;
        }
    }
    void col_terms(int j, double f, double* v) const {
        const size_t q = size_t(j);
        const double cj = c_[q], lo = l_[q], up = u_[q], dc = Dc_[q], xl = xl_[q];
        for (int h = 0; h < 2; ++h) {
            const double xx = h ? xs_[q] * f : double(x_[q]);
            const double d = cj + (h ? ATzs_[q] * f : double(ATz_[q]));
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

    int m_, n_;
    Team team_;
    HostCsr<T> A_, AT_;
    std::vector<T> c_, l_, u_, rl_, ru_, Dr_, Dc_;
    std::vector<double> hl_, hu_;
    std::vector<T> x_, ATz_, xs_, ATzs_, xl_, z_, Ax_, zs_, Axs_, zl_, pv_, pw_, pu_;
    T tau_ = 0, sig_ = 0;
    std::string name_;
    size_t mem_ = 0;
};

}  // namespace

int cpu_threads_for(int64_t nnz, int threads) {
    if (threads > 0) return threads;
    const int hw = int(std::max(1u, std::thread::hardware_concurrency()));
    return nnz < 200000 ? 1 : hw;  // the reference's automatic choice (gpu/pdhg.cu:986-989)
}

std::unique_ptr<Device> make_cpu_device(const Scaled& P, const DeviceOptions& opt, int threads) {
    if (P.m < 0 || P.n < 0) throw std::runtime_error("problem dimensions out of range");
    const int nt = cpu_threads_for(P.A.nnz(), threads);
    if (opt.precision_bits == 32) return std::unique_ptr<Device>(new CpuDev<float>(P, nt));
    if (opt.precision_bits == 64) return std::unique_ptr<Device>(new CpuDev<double>(P, nt));
    throw std::runtime_error("precision must be 32 or 64 bits");
}

}  // namespace tpdhg
