// Batched fast-path backend (batch.hpp): block-diagonal packing of many LPs.
//
// Layout in HBM. Row vectors of all instances are concatenated (instance i owns rows [roff_i, roff_i + m_i)),
// column vectors likewise ([coff_i, coff_i + n_i)). Matrices live in pools: one CSR (row pointers holding pool
// offsets, local column indices, values in the iterate type and in FP64) per distinct scaled matrix, for A and
// for A^T; instance i reads its rows through pbase_i (its first row pointer in the pool) and gathers from its own
// segment of the vector. A row-block table lists (instance, first row, end row) for every instance's row blocks,
// so one launch covers the whole batch; a workgroup whose instance is idle returns at once. Reductions use a
// second table with red_blocks(size_i) entries per instance, the single-instance partition, so sums come out
// bit-identical to the FastDevice's.
//
// Per iteration the batch costs one row launch and one column launch; a block of eval_freq iterations is one HIP
// graph whose kernels read the activity flags and the per-instance step sizes from device memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <numeric>

#include "batch.hpp"
#include "hipcommon.hpp"

namespace tpdhg {
namespace {

// Sorted pools pack (row-order position << kBatchColBits) | local column: instances need fewer than 2^19 columns
// (and rows), and a block holds at most 2^13 nonzeros.
constexpr int kBatchColBits = 19;
constexpr int kEvalK = 7;  // partial sums per evaluation block: 5 KKT slots and 2 fixed-point residual terms

template <class T>
__device__ __forceinline__ T batch_lambda(int k) {
    return T(double(k + 1) / double(k + 2));
}

// One row block of a batch: instance, local rows [r0, r1), absolute nonzero range [k0, k1) in the pool, the
// instance's first row pointer in the pool, the offset of its segment of the gathered vector and of the vector the
// epilogue writes. The tables and the live lists hold these records, so a workgroup finds everything about its
// block in one 32-byte read.
struct alignas(16) BMeta {
    int inst, r0, r1, k0, k1, pbase, voff, off;
};

// Epilogues: load(instance, offset) reads the per-instance scalars; pre(r) loads the operands of local row r and
// apply(r, sum, pre) finishes it (loaded early, before the gathers, so their latency overlaps them);
// operator()(r, sum) does both at once.

// y[off_i + r] = s
template <class T>
struct BStore {
    T* y;
    int base;
    struct Pre {};
    __device__ void load(int, int off) { base = off; }
    __device__ Pre pre(int) const { return {}; }
    __device__ void apply(int r, T s, const Pre&) const { y[base + r] = s; }
    __device__ void operator()(int r, T s) const { y[base + r] = s; }
};

// Dual prox step and Halpern blend of the fast path (fast.hip EpiHRow), per instance scalars.
template <class T>
struct BEpiHRow {
    T* z;
    const T *z0, *zh, *rl, *ru, *par;
    const int* kc;
    int s;
    T rho;
    int base;
    T sig, lam;
    struct Pre {
        T zi, h, lo, hi, z0;
    };
    __device__ void load(int i, int off) { base = off, sig = par[2 * i + 1], lam = batch_lambda<T>(kc[i] + s); }
    __device__ Pre pre(int r) const {
        const int g = base + r;
        return Pre{z[g], zh[g], rl[g], ru[g], z0[g]};
    }
    __device__ void apply(int r, T a, const Pre& p) const {
        const T v = p.zi + sig * a, q = (v + p.h) / sig;
        const T zp = q < p.lo ? v - sig * p.lo : (q > p.hi ? v - sig * p.hi : -p.h);
        z[base + r] = lam * (rho * zp - (rho - 1) * p.zi) + (1 - lam) * p.z0;
    }
    __device__ void operator()(int r, T a) const { apply(r, a, pre(r)); }
};

// Primal Halpern blend (mode 0) or none (mode 1), next primal step and gather vector (fast.hip EpiHCol).
template <class T>
struct BEpiHCol {
    T *x, *xp, *w, *ATz;
    const T *x0, *c, *l, *u, *par;
    const int* kc;
    int s, mode;
    T rho;
    int base;
    T tau, lam;
    struct Pre {
        T xj, xpj, x0j, c, l, u;
    };
    __device__ void load(int i, int off) { base = off, tau = par[2 * i], lam = batch_lambda<T>(kc[i] + s); }
    __device__ Pre pre(int r) const {
        const int j = base + r;
        if (mode == 0) return Pre{x[j], xp[j], x0[j], c[j], l[j], u[j]};
        return Pre{x[j], T(0), T(0), c[j], l[j], u[j]};
    }
    __device__ void apply(int r, T a, const Pre& p) const {
        const int j = base + r;
        T xj = p.xj;
        if (mode == 0) {
            xj = lam * (rho * p.xpj - (rho - 1) * xj) + (1 - lam) * p.x0j;
            x[j] = xj;
        }
        const T xn = clampv(xj - tau * (p.c + a), p.l, p.u);
        xp[j] = xn, w[j] = 2 * xn - xj;
        if (ATz) ATz[j] = a;
    }
    __device__ void operator()(int r, T a) const { apply(r, a, pre(r)); }
};

// Row-block SpMV over the batch with a CAP-entry LDS stage, one workgroup of WG threads per block (record m);
// rows are local to the instance. SORTED: the sorted layout of hipcommon.hpp (k_sorted) per instance: within a
// block the nonzeros are in column order with their row-order position packed above the column, each row summed
// in CSR order (block_row_sums' assignment of rows to threads and its summation order); the row pointers and the
// epilogue operands of a thread's first rows are loaded before the gathers. Otherwise the k_rowblocks scheme. A
// row longer than CAP is a block of its own, summed strided by the workgroup. sh (CAP values) and red (WG / 64)
// are the workgroup's LDS; the whole workgroup takes the same path, so the barriers inside are uniform.
template <class T, int CAP, bool SORTED, int WG, class Epi>
__device__ __forceinline__ void bblock(const BMeta& m, const int* __restrict__ ptr, const int* __restrict__ idx,
                                       const T* __restrict__ val, const T* __restrict__ v,
                                       const uint8_t* __restrict__ on, Epi epi, T* sh, T* red) {
    if (!on[m.inst]) return;  // the whole workgroup leaves together
    const int* P = ptr + m.pbase;
    const T* vi = v + m.voff;
    const int tid = int(threadIdx.x);
    const int r0 = m.r0, r1 = m.r1, k0 = m.k0, k1 = m.k1;
    constexpr uint32_t kMask = SORTED ? (1u << kBatchColBits) - 1u : 0xffffffffu;
    epi.load(m.inst, m.off);
    if (k1 - k0 > CAP) {
        T s = 0;
        gather_rounds<WG, 8>(k0, k1, idx, val, vi, [](int c) { return uint32_t(c) & kMask; },
                             [&](int, int, T x) { s += x; });
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if ((tid & 63) == 0) red[tid >> 6] = s;
        __syncthreads();
        if (tid == 0) epi(r0, tree_sum<WG / 64>(red));
        return;
    }
    if constexpr (SORTED) {
        const int nr = r1 - r0;
        constexpr int Q = 4;  // rows per thread loaded ahead (one per lane group when nr <= WG)
        int pa[Q], pe[Q];
        typename Epi::Pre pp[Q];
        const bool grouped = nr <= WG;
        int lg = 6;
        while (grouped && lg > 0 && (nr << lg) > WG) --lg;
        const int row = tid >> lg, lane = tid & ((1 << lg) - 1), lpr = 1 << lg;
        if (grouped) {
            if (row < nr) {
                pa[0] = P[r0 + row] - k0, pe[0] = P[r0 + row + 1] - k0;
                if (lane == 0) pp[0] = epi.pre(r0 + row);
            }
        } else {
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int r = r0 + q * WG + tid;
                if (r < r1) pa[q] = P[r] - k0, pe[q] = P[r + 1] - k0, pp[q] = epi.pre(r);
            }
        }
        // slots per thread per round: CAP / WG when that is under 8 (the 1,024-nonzero stage: 4, not 8 of which
        // half would load a clamped duplicate), else 8
        gather_rounds<WG, (CAP / WG < 8 ? CAP / WG : 8)>(k0, k1, idx, val, vi,
                                                         [](int c) { return uint32_t(c) & kMask; },
                                                         [&](int, int c, T x) { sh[uint32_t(c) >> kBatchColBits] = x; });
        __syncthreads();
        if (grouped) {
            T s = 0;
            if (row < nr)
                for (int k = pa[0] + lane; k < pe[0]; k += lpr) s += sh[k];
            for (int o = lpr >> 1; o > 0; o >>= 1) s += __shfl_xor(s, o);
            if (row < nr && lane == 0) epi.apply(r0 + row, s, pp[0]);
            return;
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int r = r0 + q * WG + tid;
            if (r < r1) {
                T s = 0;
                for (int j = pa[q]; j < pe[q]; ++j) s += sh[j];
                epi.apply(r, s, pp[q]);
            }
        }
        for (int r = r0 + Q * WG + tid; r < r1; r += WG) {  // rows beyond the first Q per thread
            T s = 0;
            for (int j = P[r] - k0; j < P[r + 1] - k0; ++j) s += sh[j];
            epi(r, s);
        }
        return;
    }
    gather_rounds<WG, CAP / WG>(k0, k1, idx, val, vi, [](int c) { return uint32_t(c) & kMask; },
                                [&](int k, int, T x) { sh[k - k0] = x; });
    __syncthreads();
    const int nr = r1 - r0;
    int lg = 6;
    while (lg > 0 && (nr << lg) > WG) --lg;
    const int row = tid >> lg, lane = tid & ((1 << lg) - 1), lpr = 1 << lg;
    T s = 0;
    if (row < nr) {
        const int a = P[r0 + row] - k0, e = P[r0 + row + 1] - k0;
        for (int k = a + lane; k < e; k += lpr) s += sh[k];
    }
    for (int o = lpr >> 1; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (row < nr && lane == 0) epi(r0 + row, s);
}

// One launch over block records: a side's whole table, or a live list (the live instances' blocks, so a launch
// covers only the instances still iterating). TAG only names a second instantiation (the all-live probe
// launches) for kernel traces.
template <class T, int CAP, bool SORTED, class Epi, int WG = stage_wg(CAP), int TAG = 0>
__global__ __launch_bounds__(WG) void k_bblocks(int nblk, const BMeta* __restrict__ meta, const int* __restrict__ ptr,
                                                 const int* __restrict__ idx, const T* __restrict__ val,
                                                 const T* __restrict__ v, const uint8_t* __restrict__ on, Epi epi) {
    __shared__ T sh[CAP];
    __shared__ T red[WG / 64];
    const BMeta m = meta[xcd_block(int(blockIdx.x), nblk)];
    bblock<T, CAP, SORTED, WG>(m, ptr, idx, val, v, on, epi, sh, red);
}

// ------------------------------------------------------------------ persistent blocks (small live sets)
// When few instances are live, one iteration is two launches of a few workgroups each, and each launch is bound by
// its chain of dependent loads and its launch gap rather than by bytes (the one-LP tail of a batch: about 7 us per
// launch). k_bpersist runs a whole block of iterations in one cooperative launch instead: workgroup g takes the
// live row blocks g, g + G, ... and the live column blocks likewise, and every step is row phase, grid barrier,
// column phase, grid barrier, with the same block code as k_bblocks (so results are bit-identical to the lockstep
// launches). Before its first phase each workgroup copies the vector entries its blocks will write (z of its row
// blocks; x, x+, w of its column blocks) to snapshot buffers, so a block that does not finish can be undone.
//
// The grid barrier counts arrivals on one counter (reset to 0 before each launch; arrival k of a workgroup waits
// for k * G). Its wait is bounded: after `spins` polls, or when another workgroup has already given up, the
// workgroup sets the error flag and returns, so every wave of the launch exits; the host then restores the
// snapshot and runs the block with the lockstep launches.
struct GridBar {
    unsigned* count;
    int* err;
    int spins;
    unsigned extra;  // added to every target; > 0 only to force the time-out path in tests (never reached)
};

__device__ __forceinline__ bool grid_barrier(const GridBar& g, unsigned target, int* flag) {
    __syncthreads();
    if (threadIdx.x == 0) {
        int fail = 0;
        __threadfence();  // release this workgroup's writes at agent scope
        atomicAdd(g.count, 1u);
        for (int k = 0; __hip_atomic_load(g.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target + g.extra;
             ++k) {
            if (k >= g.spins || __hip_atomic_load(g.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                atomicExch(g.err, 1);
                fail = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __threadfence();  // acquire the other workgroups' writes
        *flag = fail;
    }
    __syncthreads();
    return *flag == 0;
}

template <class T>
struct PersistArgs {
    int steps, na, nt;
    const BMeta *alist, *tlist;      // live blocks of A (row phase, gathers w) and of A^T (column phase, gathers z)
    const int *aptr, *aidx, *tptr, *tidx;
    const T *aval, *tval;
    const uint8_t* on;
    BEpiHRow<T> er;
    BEpiHCol<T> ec;  // ec.ATz is set on the block's last step only
    T* ATz;
    T *zs, *xs, *xps, *ws;  // snapshot of z (rows) and x, x+, w (columns) of the live blocks
    GridBar bar;
};

template <class T, int CAP, bool SORTED, int WG = stage_wg(CAP)>
__global__ __launch_bounds__(WG) void k_bpersist(PersistArgs<T> a) {
    __shared__ T sh[CAP];
    __shared__ T red[WG / 64];
    __shared__ int flag;
    const int g = int(blockIdx.x), G = int(gridDim.x), tid = int(threadIdx.x);
    for (int i = g; i < a.na; i += G) {
        const BMeta m = a.alist[i];
        if (!a.on[m.inst]) continue;
        const int base = m.off;
        for (int r = m.r0 + tid; r < m.r1; r += WG) a.zs[base + r] = a.er.z[base + r];
    }
    for (int i = g; i < a.nt; i += G) {
        const BMeta m = a.tlist[i];
        if (!a.on[m.inst]) continue;
        const int base = m.off;
        for (int r = m.r0 + tid; r < m.r1; r += WG) {
            const int j = base + r;
            a.xs[j] = a.ec.x[j], a.xps[j] = a.ec.xp[j], a.ws[j] = a.ec.w[j];
        }
    }
    __syncthreads();
    unsigned target = 0;
    for (int s = 0; s < a.steps; ++s) {
        BEpiHRow<T> er = a.er;
        er.s = s;
        for (int i = g; i < a.na; i += G) {
            bblock<T, CAP, SORTED, WG>(a.alist[i], a.aptr, a.aidx, a.aval, a.ec.w, a.on, er, sh, red);
            __syncthreads();
        }
        target += unsigned(G);
        if (!grid_barrier(a.bar, target, &flag)) return;
        BEpiHCol<T> ec = a.ec;
        ec.s = s;
        ec.ATz = s + 1 == a.steps ? a.ATz : nullptr;
        for (int i = g; i < a.nt; i += G) {
            bblock<T, CAP, SORTED, WG>(a.tlist[i], a.tptr, a.tidx, a.tval, a.er.z, a.on, ec, sh, red);
            __syncthreads();
        }
        if (s + 1 < a.steps) {
            target += unsigned(G);
            if (!grid_barrier(a.bar, target, &flag)) return;
        }
    }
}

// Undoes a persistent block that did not finish: workgroup i < na restores z of live row block alist[i], the
// others x, x+, w of live column block tlist[i - na].
template <class T>
__global__ __launch_bounds__(kWG) void k_brestore(PersistArgs<T> a) {
    const int i = int(blockIdx.x), tid = int(threadIdx.x);
    const bool row = i < a.na;
    const BMeta m = row ? a.alist[i] : a.tlist[i - a.na];
    if (!a.on[m.inst]) return;
    const int base = m.off;
    for (int r = m.r0 + tid; r < m.r1; r += kWG) {
        const int j = base + r;
        if (row) {
            a.er.z[j] = a.zs[j];
        } else {
            a.ec.x[j] = a.xs[j], a.ec.xp[j] = a.xps[j], a.ec.w[j] = a.ws[j];
        }
    }
}

__global__ void k_reset_bar(unsigned* count, int* err) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *count = 0u, *err = 0;
}

// Reduction / elementwise table: (instance, q, nb, slot) entries; entry q of nb covers local indices
// q * kWG + t, stepping nb * kWG (the single-instance grid-stride partition with nb = red_blocks(size)); slot is the
// entry's index in the whole table, where its partial sums go, so a launch over a compacted list of the live
// instances' entries fills the same slots as one over the whole table.
struct Seg {
    const int* tab;
    const int* off;   // first global index of each instance
    const int* size;  // element count of each instance
};

#define SEG_LOOP(S, on, ...)                                                          \
    const int* e_ = (S).tab + 4 * blockIdx.x;                                          \
    const int inst_ = e_[0], q_ = e_[1], nb_ = e_[2], slot_ = e_[3];                   \
    (void)slot_;                                                                       \
    const bool act_ = (on)[inst_] != 0;                                                \
    const int base_ = (S).off[inst_], size_ = (S).size[inst_];                         \
    if (act_)                                                                          \
        for (int r_ = q_ * kWG + int(threadIdx.x); r_ < size_; r_ += nb_ * kWG) {      \
            const int g = base_ + r_;                                                  \
            __VA_ARGS__                                                                \
        }

template <class T>
__global__ __launch_bounds__(kWG) void k_bbegin_rows(Seg S, const uint8_t* __restrict__ on, const double* __restrict__ rl,
                                                     const double* __restrict__ ru, const double* __restrict__ Axb,
                                                     const double* __restrict__ zb, T* __restrict__ rl_in,
                                                     T* __restrict__ ru_in, T* __restrict__ zh, T* __restrict__ z,
                                                     T* __restrict__ z0) {
    SEG_LOOP(S, on, {
        rl_in[g] = T(rl[g] - Axb[g]), ru_in[g] = T(ru[g] - Axb[g]), zh[g] = T(zb[g]);
        z[g] = 0, z0[g] = 0;
    })
}

template <class T>
__global__ __launch_bounds__(kWG) void k_bbegin_cols(Seg S, const uint8_t* __restrict__ on, const double* __restrict__ c,
                                                     const double* __restrict__ l, const double* __restrict__ u,
                                                     const double* __restrict__ ATzb, const double* __restrict__ xb,
                                                     T* __restrict__ c_in, T* __restrict__ l_in, T* __restrict__ u_in,
                                                     T* __restrict__ x, T* __restrict__ x0) {
    SEG_LOOP(S, on, {
        const T lo = T(l[g] - xb[g]), hi = T(u[g] - xb[g]);
        c_in[g] = T(c[g] + ATzb[g]), l_in[g] = lo, u_in[g] = hi;
        const T v = clampv(T(0), lo, hi);
        x[g] = v, x0[g] = v;
    })
}

template <class T>
__global__ __launch_bounds__(kWG) void k_baccumulate(Seg S, const uint8_t* __restrict__ on, double* __restrict__ base,
                                                     const T* __restrict__ inner) {
    SEG_LOOP(S, on, { base[g] += double(inner[g]); })
}

template <class T>
__global__ __launch_bounds__(kWG) void k_bcopy(Seg S, const uint8_t* __restrict__ on, T* __restrict__ dst,
                                               const T* __restrict__ src) {
    SEG_LOOP(S, on, { dst[g] = src[g]; })
}

template <class T>
__global__ __launch_bounds__(kWG) void k_bscale(Seg S, const uint8_t* __restrict__ on, T* __restrict__ x,
                                                const double* __restrict__ fac) {
    SEG_LOOP(S, on, { x[g] = T(x[g] * fac[inst_]); })
}

template <class T>
__global__ __launch_bounds__(kWG) void k_bsumsq(Seg S, const uint8_t* __restrict__ on, const T* __restrict__ x,
                                                double* __restrict__ part) {
    double v[1] = {0};
    SEG_LOOP(S, on, { v[0] += double(x[g]) * double(x[g]); })
    block_store<1>(v, part, slot_);
}

// fac[i] = 1 / sqrt(s[i]) for the flagged instances with s[i] > 0 (the device-side normalisation of a power step;
// the same arithmetic as k_scale_by_sum), 1 otherwise.
__global__ void k_inv_sqrt_pos(int B, const uint8_t* __restrict__ on, const double* __restrict__ s,
                               double* __restrict__ fac) {
    for (int i = int(blockIdx.x * blockDim.x + threadIdx.x); i < B; i += int(gridDim.x * blockDim.x))
        if (on[i]) {
            const double sv = s[i];
            fac[i] = sv > 0 ? 1 / sqrt(sv) : 1.0;
        }
}

__global__ void k_bset_counters(int B, const uint8_t* __restrict__ on, int* __restrict__ kc, int add, int reset) {
    for (int i = int(blockIdx.x * blockDim.x + threadIdx.x); i < B; i += int(gridDim.x * blockDim.x))
        if (on[i]) kc[i] = reset ? 0 : kc[i] + add;
}

template <class T>
struct BEvalArgs {
    const T *Ax, *z, *z0, *x, *ATz, *x0;
    const double *Axb, *zb, *xb, *ATzb;
    const double *rl, *ru, *Dr, *c, *l, *u, *Dc;
    const T *Aw, *xp, *zh, *rl_in, *ru_in, *par;  // fixed-point residual: A w, x+, the inner row prox data
};

// KKT slots of base + inner per instance and the fixed-point residual terms of the inner point (fast.hip
// k_eval_fast, per segment, the same seven partials in the same order).
template <class T>
__global__ __launch_bounds__(kWG) void k_beval_rows(Seg S, const uint8_t* __restrict__ on, BEvalArgs<T> a,
                                                    double* __restrict__ part) {
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    SEG_LOOP(S, on, {
        const T zp = row_prox(a.z[g], a.Aw[g], a.zh[g], a.rl_in[g], a.ru_in[g], a.par[2 * inst_ + 1]);
        const double dzt = double(a.z[g]) - double(zp), adx = 0.5 * (double(a.Ax[g]) - double(a.Aw[g]));
        v[5] += dzt * dzt, v[6] += dzt * adx;
        const double ax = a.Axb[g] + double(a.Ax[g]), zz = a.zb[g] + double(a.z[g]);
        const double lo = a.rl[g], up = a.ru[g], dr = a.Dr[g];
        const double r = ax < lo ? lo - ax : (ax > up ? ax - up : 0), ro = r / dr;
        const double rr = ro * row_rel_weight(lo, up, dr);
        v[0] += r * r, v[1] += ro * ro, v[2] += rr * rr;
        if (zz > 0 && up < INFINITY) v[3] -= zz * up;
        if (zz < 0 && lo > -INFINITY) v[3] -= zz * lo;
        const double dz = double(a.z[g]) - double(a.z0[g]);
        v[4] += dz * dz;
    })
    block_store<7>(v, part, slot_);
}

template <class T>
__global__ __launch_bounds__(kWG) void k_beval_cols(Seg S, const uint8_t* __restrict__ on, BEvalArgs<T> a,
                                                    double* __restrict__ part) {
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    SEG_LOOP(S, on, {
        const double xx = a.xb[g] + double(a.x[g]), cj = a.c[g];
        const double d = cj + a.ATzb[g] + double(a.ATz[g]), lo = a.l[g], up = a.u[g];
        double res = 0;
        v[2] += cj * xx;
        if (d > 0) {
            if (lo > -INFINITY) v[3] += d * lo;
            else res = d;
        } else if (d < 0) {
            if (up < INFINITY) v[3] += d * up;
            else res = -d;
        }
        const double ro = res / a.Dc[g], dx = double(a.x[g]) - double(a.x0[g]);
        v[0] += res * res, v[1] += ro * ro, v[4] += dx * dx;
        const double dxt = double(a.x[g]) - double(a.xp[g]);
        v[5] += dxt * dxt;
    })
    block_store<7>(v, part, slot_);
}

// Per instance: out[inst * stride + ooff + k] = sum of its K partials (k_finish per segment, same order).
__global__ __launch_bounds__(kWG) void k_bfinish(const int* __restrict__ first, const int* __restrict__ count,
                                                 const uint8_t* __restrict__ on, int K, const double* __restrict__ part,
                                                 int stride, int ooff, double* __restrict__ out) {
    __shared__ double sh[kWG];
    const int inst = int(blockIdx.x), t = int(threadIdx.x);
    if (!on[inst]) return;
    const int b0 = first[inst], blocks = count[inst];
    for (int k = 0; k < K; ++k) {
        double s = 0;
        for (int q = t; q < blocks; q += kWG) s += part[size_t(b0 + q) * size_t(K) + size_t(k)];
        sh[t] = s;
        __syncthreads();
        for (int o = kWG / 2; o > 0; o >>= 1) {
            if (t < o) sh[t] += sh[t + o];
            __syncthreads();
        }
        if (t == 0) out[size_t(inst) * size_t(stride) + size_t(ooff + k)] = sh[0];
        __syncthreads();
    }
}

template <class T>
__global__ __launch_bounds__(kWG) void k_bcombine(int k, const double* __restrict__ base, const T* __restrict__ inner,
                                                  double* __restrict__ out) {
    for (int i = int(blockIdx.x * blockDim.x + threadIdx.x); i < k; i += int(gridDim.x * blockDim.x))
        out[i] = base[i] + double(inner[i]);
}

// Row blocks with a stage of `cap` nonzeros (row_blocks of hipcommon.hpp for cap = kCap), the row-sum chains
// bounded for the stage's workgroup (chain_fits).
std::vector<int32_t> row_blocks_cap(int rows, const int32_t* ptr, int cap) {
    const int wg = stage_wg(cap);
    std::vector<int32_t> b;
    b.push_back(0);
    int r0 = 0;
    while (r0 < rows) {
        int r = r0;
        int64_t lmax = 0;
        while (r < rows && r - r0 < kWG && int64_t(ptr[r + 1]) - ptr[r0] <= cap) {
            const int64_t lm = std::max<int64_t>(lmax, int64_t(ptr[r + 1]) - ptr[r]);
            if (!chain_fits(r - r0 + 1, lm, wg)) break;
            lmax = lm, ++r;
        }
        if (r == r0) r = r0 + 1;
        b.push_back(r);
        r0 = r;
    }
    return b;
}

struct Pool {  // distinct matrices of one side, host copy
    std::vector<int32_t> ptr, idx;
    std::vector<double> val;
    std::vector<int32_t> pbase;   // per distinct matrix: first row pointer in ptr
    std::vector<int32_t> rows;    // per distinct matrix
};

uint64_t mix_hash(uint64_t h, uint64_t v) {
    h ^= v + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2);
    return h;
}

uint64_t csr_hash(const Csr& a) {
    uint64_t h = mix_hash(uint64_t(a.rows), uint64_t(a.nnz()));
    const size_t step = std::max<size_t>(1, a.val.size() / 4096);
    for (size_t k = 0; k < a.val.size(); k += step) {
        uint64_t bits;
        std::memcpy(&bits, &a.val[k], sizeof bits);
        h = mix_hash(h, bits ^ (uint64_t(uint32_t(a.idx[k])) << 32));
    }
    return h;
}

bool csr_equal(const Csr& a, const Csr& b) {
    return a.rows == b.rows && a.ptr == b.ptr && a.idx == b.idx && a.val == b.val;
}

template <class T>
class BatchDev final : public BatchDevice {
public:
    BatchDev(const std::vector<const Scaled*>& probs, const BatchOptions& o) : B_(int(probs.size())), o_(o) {
        require_device();
        if (B_ <= 0) throw std::runtime_error("empty batch");
        if (!(o.rho >= 1 && o.rho < 2)) throw std::runtime_error("reflection coefficient rho must be in [1, 2)");
        if (o.eval_freq < 1) throw std::runtime_error("eval_freq must be >= 1");
        if (o.cap != 0 && o.cap != 1024 && o.cap != 2048 && o.cap != 4096 && o.cap != 8192)
            throw std::runtime_error("batch row-block stage must be 0 (autotune), 1024, 2048, 4096 or 8192 nonzeros");
        size_t free0 = 0, free1 = 0, total = 0;
        HIP_OK(hipMemGetInfo(&free0, &total));
        name_ = device_name();
        persist_max_ = o.persist_max;
        if (const char* e = std::getenv("TPDHG_PERSIST_MAX")) persist_max_ = std::atoi(e);
        if (const char* e = std::getenv("TPDHG_PERSIST_SPINS")) persist_spins_ = std::max(1, std::atoi(e));
        // failure injection for tests: every grid barrier waits for one arrival too many and gives up after
        // persist_spins_ polls, so the restore-and-replay path runs while every wave still exits
        if (const char* e = std::getenv("TPDHG_PERSIST_TEST_UNREACHABLE")) persist_unreachable_ = std::atoi(e) > 0 ? 1u : 0u;
        HIP_OK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
        build_host(probs);
        upload_all(probs);
        const char* pick = o.cap ? nullptr : std::getenv("TPDHG_TUNED_STAGE");
        if (o.cap) {
            set_cap(o.cap);
            tuned_cap_ = o.cap;
        } else if (pick) {  // the autotuner's pick given (tests of the stage following the live set)
            const int c = std::atoi(pick);
            if (c != 1024 && c != 2048 && c != 4096 && c != 8192)
                throw std::runtime_error("TPDHG_TUNED_STAGE must be 1024, 2048, 4096 or 8192");
            set_cap(c);
            tuned_cap_ = c;
        } else {
            autotune();
        }
        HIP_OK(hipMemGetInfo(&free1, &total));
        mem_used_ = free0 > free1 ? free0 - free1 : 0;
        // the stage follows the live set only after tuning (TPDHG_FIXED_STAGE keeps the tuned one throughout)
        auto_stage_ = o.cap == 0 && std::getenv("TPDHG_FIXED_STAGE") == nullptr;
    }
    ~BatchDev() override {
        if (graph_) (void)hipGraphExecDestroy(graph_);
        for (hipEvent_t e : events_) (void)hipEventDestroy(e);
        if (hflags_) (void)hipHostFree(hflags_);
        if (hpar_) (void)hipHostFree(hpar_);
        if (hout_) (void)hipHostFree(hout_);
        if (hlist_) (void)hipHostFree(hlist_);
        if (herr_) (void)hipHostFree(herr_);
        if (st_) (void)hipStreamDestroy(st_);
    }

    int count() const override { return B_; }
    int m(int i) const override { return m_[size_t(i)]; }
    int n(int i) const override { return n_[size_t(i)]; }
    int64_t nnz_stored() const override { return int64_t(PA_.val.size()); }
    int64_t nnz_total() const override { return nnz_total_; }
    int unique_matrices() const override { return int(PA_.rows.size()); }
    int cap() const override { return cap_; }
    std::vector<float> tune_ms() const override { return tune_ms_; }

    // A solve starts with every instance live: on the tuned stage (the previous solve may have ended on a smaller
    // one), so every solve of a batch sums in the same order from its first product on.
    void tuned_stage() {
        if (auto_stage_ && cap_ != tuned_cap_) set_cap(tuned_cap_);
    }

    std::vector<double> power_norm(int iters) override {
        tuned_stage();
        static_assert(kPowerGroup <= 2 * kEvalK, "a power-iteration group's sums of squares go to out_");
        std::vector<double> est(static_cast<size_t>(B_), 0.0), prev(static_cast<size_t>(B_), 0.0);
        {
            std::vector<T> hv(static_cast<size_t>(ntot_));
            for (int i = 0; i < B_; ++i) {
                double nn = 0;
                std::vector<double> h(static_cast<size_t>(n_[size_t(i)]));
                for (int j = 0; j < n_[size_t(i)]; ++j) {
                    h[size_t(j)] = 0.5 + double((uint32_t(j) * 2654435761u) % 1000u) / 1000.0;
                    nn += h[size_t(j)] * h[size_t(j)];
                }
                for (int j = 0; j < n_[size_t(i)]; ++j)
                    hv[size_t(coff_[size_t(i)]) + size_t(j)] = T(h[size_t(j)] / std::sqrt(nn));
            }
            HIP_OK(hipMemcpyAsync(pv_, hv.data(), hv.size() * sizeof(T), hipMemcpyHostToDevice, st_));
            HIP_OK(hipStreamSynchronize(st_));
        }
        // Each instance follows the single solve's power iteration (FastDev::power_norm: steps normalised on the
        // device, kPowerGroup steps between host reads, the stopping rule of gpu/pdhg.cu:417-436); an instance that
        // stops inside a group keeps iterating to the group's end, its later steps discarded.
        std::vector<uint8_t> on(static_cast<size_t>(B_), 1);
        T *pv = pv_, *pu = pu_;
        bool any = true;
        for (int it0 = 0; it0 < iters && any; it0 += kPowerGroup) {
            const int g = std::min(kPowerGroup, iters - it0);
            upload_flags(on);
            for (int q = 0; q < g; ++q) {
                double* sq = out_ + size_t(q) * size_t(B_);
                spmv_rows(A_, pv, BStore<T>{pw_, 0}, dflags_);
                spmv_rows(AT_, pw_, BStore<T>{pu, 0}, dflags_);
                k_bsumsq<T><<<ncseg_, kWG, 0, st_>>>(cseg(), dflags_, pu, part_);
                k_bfinish<<<B_, kWG, 0, st_>>>(d_cfirst_, d_ccount_, dflags_, 1, part_, 1, 0, sq);
                k_inv_sqrt_pos<<<(B_ + kWG - 1) / kWG, kWG, 0, st_>>>(B_, dflags_, sq, fac_);
                k_bscale<T><<<ncseg_, kWG, 0, st_>>>(cseg(), dflags_, pu, fac_);
                std::swap(pv, pu);
            }
            HIP_OK(hipMemcpyAsync(hout_, out_, size_t(g) * size_t(B_) * sizeof(double), hipMemcpyDeviceToHost, st_));
            HIP_OK(hipStreamSynchronize(st_));
            any = false;
            for (int i = 0; i < B_; ++i) {
                for (int q = 0; q < g && on[size_t(i)]; ++q) {
                    const double s = hout_[size_t(q) * size_t(B_) + size_t(i)];
                    prev[size_t(i)] = est[size_t(i)];
                    if (s <= 0) {
                        est[size_t(i)] = 0, on[size_t(i)] = 0;
                        break;
                    }
                    est[size_t(i)] = std::sqrt(std::sqrt(s));
                    if (it0 + q >= 10 && std::fabs(est[size_t(i)] - prev[size_t(i)]) <= 1e-7 * est[size_t(i)])
                        on[size_t(i)] = 0;
                }
                any = any || on[size_t(i)];
            }
        }
        HIP_OK(hipGetLastError());
        return est;
    }

    void set_base(const std::vector<const double*>& xb, const std::vector<const double*>& zb) override {
        tuned_stage();
        std::vector<double> hx(static_cast<size_t>(ntot_), 0.0), hz(static_cast<size_t>(mtot_), 0.0);
        for (int i = 0; i < B_; ++i) {
            if (size_t(i) < xb.size() && xb[size_t(i)])
                std::copy(xb[size_t(i)], xb[size_t(i)] + n_[size_t(i)], hx.begin() + coff_[size_t(i)]);
            if (size_t(i) < zb.size() && zb[size_t(i)])
                std::copy(zb[size_t(i)], zb[size_t(i)] + m_[size_t(i)], hz.begin() + roff_[size_t(i)]);
        }
        HIP_OK(hipMemcpyAsync(xb_, hx.data(), hx.size() * sizeof(double), hipMemcpyHostToDevice, st_));
        HIP_OK(hipMemcpyAsync(zb_, hz.data(), hz.size() * sizeof(double), hipMemcpyHostToDevice, st_));
        HIP_OK(hipStreamSynchronize(st_));
    }

    void begin_round(const std::vector<uint8_t>& on) override {
        upload_flags(on);
        spmv_flagged(A64_, xb_, BStore<double>{Axb_, 0}, on, true);
        spmv_flagged(AT64_, zb_, BStore<double>{ATzb_, 0}, on, false);
        const auto rs = rsegs(on), cs = csegs(on);
        if (rs.second)
            k_bbegin_rows<T><<<rs.second, kWG, 0, st_>>>(rs.first, dflags_, rl64_, ru64_, Axb_, zb_, rl_in_, ru_in_, zh_,
                                                         z_, z0_);
        if (cs.second)
            k_bbegin_cols<T><<<cs.second, kWG, 0, st_>>>(cs.first, dflags_, c64_, l64_, u64_, ATzb_, xb_, c_in_, l_in_, u_in_,
                                                  x_, x0_);
        HIP_OK(hipStreamSynchronize(st_));
        HIP_OK(hipGetLastError());
    }

    void set_steps(const std::vector<double>& tau, const std::vector<double>& sigma) override {
        for (int i = 0; i < B_; ++i) hpar_[2 * i] = T(tau[size_t(i)]), hpar_[2 * i + 1] = T(sigma[size_t(i)]);
        HIP_OK(hipMemcpyAsync(par_, hpar_, size_t(2 * B_) * sizeof(T), hipMemcpyHostToDevice, st_));
        HIP_OK(hipStreamSynchronize(st_));
    }

    void restart(const std::vector<uint8_t>& on) override {
        upload_flags(on);
        const auto rs = rsegs(on), cs = csegs(on);
        if (cs.second) k_bcopy<T><<<cs.second, kWG, 0, st_>>>(cs.first, dflags_, x0_, x_);
        if (rs.second) k_bcopy<T><<<rs.second, kWG, 0, st_>>>(rs.first, dflags_, z0_, z_);
        k_bset_counters<<<(B_ + kWG - 1) / kWG, kWG, 0, st_>>>(B_, dflags_, kc_, 0, 1);
        spmv_flagged(AT_, z_, col_epi(0, 1, true), on, false);
        HIP_OK(hipStreamSynchronize(st_));
        HIP_OK(hipGetLastError());
    }

    void end_round(const std::vector<uint8_t>& on) override {
        upload_flags(on);
        const auto rs = rsegs(on), cs = csegs(on);
        if (cs.second) k_baccumulate<T><<<cs.second, kWG, 0, st_>>>(cs.first, dflags_, xb_, x_);
        if (rs.second) k_baccumulate<T><<<rs.second, kWG, 0, st_>>>(rs.first, dflags_, zb_, z_);
        HIP_OK(hipStreamSynchronize(st_));
        HIP_OK(hipGetLastError());
    }

    void iterate(int steps, const std::vector<uint8_t>& on) override {
        if (steps <= 0) return;
        upload_act(on);
        ensure_lists(on);
        if (persistent_fits()) {
            launch_persistent(steps);
            return;
        }
        ++lockstep_blocks_;
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

    void eval(const std::vector<uint8_t>& on, std::vector<double>& out) override {
        upload_flags(on);
        spmv_flagged(A_, x_, BStore<T>{Ax_, 0}, on, true);
        spmv_flagged(A_, w_, BStore<T>{pw_, 0}, on, true);  // A w, w = 2 x+ - x: the dual half of T(u)
        BEvalArgs<T> a{Ax_, z_, z0_, x_, ATz_, x0_, Axb_, zb_, xb_, ATzb_, rl64_, ru64_, Dr64_, c64_, l64_, u64_, Dc64_,
                       pw_, xp_, zh_, rl_in_, ru_in_, par_};
        const auto rs = rsegs(on), cs = csegs(on);
        if (rs.second) k_beval_rows<T><<<rs.second, kWG, 0, st_>>>(rs.first, dflags_, a, part_);
        if (cs.second)
            k_beval_cols<T><<<cs.second, kWG, 0, st_>>>(cs.first, dflags_, a, part_ + size_t(nrseg_) * kEvalK);
        k_bfinish<<<B_, kWG, 0, st_>>>(d_rfirst_, d_rcount_, dflags_, kEvalK, part_, 2 * kEvalK, 0, out_);
        k_bfinish<<<B_, kWG, 0, st_>>>(d_cfirst_after_rows_, d_ccount_, dflags_, kEvalK, part_, 2 * kEvalK, kEvalK, out_);
        HIP_OK(hipMemcpyAsync(hout_, out_, size_t(B_) * 2 * kEvalK * sizeof(double), hipMemcpyDeviceToHost, st_));
        HIP_OK(hipStreamSynchronize(st_));
        HIP_OK(hipGetLastError());
        out.resize(size_t(B_) * NKKT);
        fp_.resize(size_t(B_) * 3);
        for (int i = 0; i < B_; ++i) {
            if (!on[size_t(i)]) continue;
            const double* h = hout_ + size_t(i) * 2 * kEvalK;  // rows: 5 KKT slots + 2, columns: 5 + 2
            std::copy(h, h + 5, out.begin() + size_t(i) * NKKT);
            std::copy(h + kEvalK, h + kEvalK + 5, out.begin() + size_t(i) * NKKT + 5);
            fp_[size_t(i) * 3] = h[kEvalK + 5], fp_[size_t(i) * 3 + 1] = h[5], fp_[size_t(i) * 3 + 2] = h[6];
        }
        collect_events();
    }

    void fixed_points(std::vector<double>& out) const override { out = fp_; }

    void download(int i, double* hx, double* hz) override {
        const int n = n_[size_t(i)], m = m_[size_t(i)];
        if (hx && n) {
            k_bcombine<T><<<grid_for(n), kWG, 0, st_>>>(n, xb_ + coff_[size_t(i)], x_ + coff_[size_t(i)], buf64_);
            HIP_OK(hipMemcpyAsync(hx, buf64_, size_t(n) * sizeof(double), hipMemcpyDeviceToHost, st_));
            HIP_OK(hipStreamSynchronize(st_));
        }
        if (hz && m) {
            k_bcombine<T><<<grid_for(m), kWG, 0, st_>>>(m, zb_ + roff_[size_t(i)], z_ + roff_[size_t(i)], buf64_);
            HIP_OK(hipMemcpyAsync(hz, buf64_, size_t(m) * sizeof(double), hipMemcpyDeviceToHost, st_));
            HIP_OK(hipStreamSynchronize(st_));
        }
        HIP_OK(hipGetLastError());
    }

    void set_vectors(int i, const double* c, const double* l, const double* u, const double* rl,
                     const double* ru) override {
        if (i < 0 || i >= B_) throw std::invalid_argument("instance index out of range");
        const size_t n = size_t(n_[size_t(i)]), m = size_t(m_[size_t(i)]);
        const size_t co = size_t(coff_[size_t(i)]), ro = size_t(roff_[size_t(i)]);
        auto put = [&](double* dst, const double* src, size_t k) {
            if (k) HIP_OK(hipMemcpyAsync(dst, src, k * sizeof(double), hipMemcpyHostToDevice, st_));
        };
        put(c64_ + co, c, n), put(l64_ + co, l, n), put(u64_ + co, u, n), put(rl64_ + ro, rl, m), put(ru64_ + ro, ru, m);
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

    int64_t persistent_blocks() const override { return persistent_blocks_; }
    int64_t lockstep_blocks() const override { return lockstep_blocks_; }
    int persist_max() const override { return persist_ok_ ? persist_max_ : 0; }
    int64_t stage_switches() const override { return stage_switches_; }

    KernelTimes probe(int iters) override {
        if (iters < 1) throw std::invalid_argument("probe needs at least one iteration");
        std::vector<uint8_t> all(static_cast<size_t>(B_), 1);
        HIP_OK(hipStreamSynchronize(st_));
        collect_events();
        const KernelTimes saved = times_;
        upload_act(all);
        ensure_lists(all);
        launch_block(2, true);  // warm-up
        HIP_OK(hipStreamSynchronize(st_));
        collect_events();
        times_ = KernelTimes();
        launch_block(iters, true);
        HIP_OK(hipStreamSynchronize(st_));
        HIP_OK(hipGetLastError());
        collect_events();
        const KernelTimes t = times_;
        times_ = saved;
        return t;
    }

private:
    struct Side {  // device copy of one side's pool and the block records for the current stage size
        int rows_total = 0, nblk = 0;
        int *ptr = nullptr, *idx = nullptr;
        BMeta* meta = nullptr;
        T* val = nullptr;
    };
    struct Side64 {
        int nblk = 0;
        int *ptr = nullptr, *idx = nullptr;
        BMeta* meta = nullptr;
        double* val = nullptr;
    };

    void build_host(const std::vector<const Scaled*>& probs) {
        m_.resize(size_t(B_)), n_.resize(size_t(B_)), roff_.resize(size_t(B_)), coff_.resize(size_t(B_));
        mat_of_.resize(size_t(B_));
        std::multimap<uint64_t, int> seen;  // hash -> distinct matrix index
        std::vector<const Scaled*> uniq;
        int64_t mt = 0, nt = 0;
        for (int i = 0; i < B_; ++i) {
            const Scaled& P = *probs[size_t(i)];
            if (P.m <= 0 || P.n <= 0) throw std::runtime_error("batch instances need at least one row and one column");
            m_[size_t(i)] = P.m, n_[size_t(i)] = P.n, roff_[size_t(i)] = int(mt), coff_[size_t(i)] = int(nt);
            mt += P.m, nt += P.n;
            nnz_total_ += P.A.nnz();
            const uint64_t h = csr_hash(P.A);
            int found = -1;
            for (auto it = seen.lower_bound(h); it != seen.end() && it->first == h; ++it)
                if (csr_equal(uniq[size_t(it->second)]->A, P.A)) {
                    found = it->second;
                    break;
                }
            if (found < 0) {
                found = int(uniq.size());
                uniq.push_back(&P);
                seen.emplace(h, found);
            }
            mat_of_[size_t(i)] = found;
        }
        if (mt > INT32_MAX || nt > INT32_MAX) throw std::runtime_error("batch too large for 32-bit vector offsets");
        sorted_ = sorted_layout_enabled();
        for (int i = 0; i < B_; ++i)
            if (m_[size_t(i)] >= (1 << kBatchColBits) || n_[size_t(i)] >= (1 << kBatchColBits)) sorted_ = false;
        mtot_ = int(mt), ntot_ = int(nt);
        auto add = [](Pool& pool, const Csr& a) {
            const int64_t at = int64_t(pool.val.size());
            if (at + a.nnz() > kMaxNnz)
                throw std::length_error("the batch's distinct matrices hold more than " + std::to_string(kMaxNnz) +
                                        " nonzeros (32-bit nonzero indices)");
            pool.pbase.push_back(int32_t(pool.ptr.size()));
            pool.rows.push_back(a.rows);
            for (int32_t p : a.ptr) pool.ptr.push_back(int32_t(at + p));
            pool.idx.insert(pool.idx.end(), a.idx.begin(), a.idx.end());
            pool.val.insert(pool.val.end(), a.val.begin(), a.val.end());
        };
        for (const Scaled* P : uniq) add(PA_, P->A), add(PT_, P->AT);
        // reduction / elementwise tables (single-instance partition)
        auto seg = [&](const std::vector<int>& size, std::vector<int32_t>& tab, std::vector<int32_t>& first,
                       std::vector<int32_t>& count) {
            for (int i = 0; i < B_; ++i) {
                const int nb = red_blocks(size[size_t(i)]);
                first.push_back(int32_t(tab.size() / 4)), count.push_back(nb);
                for (int q = 0; q < nb; ++q) {
                    const int32_t slot = int32_t(tab.size() / 4);
                    tab.push_back(i), tab.push_back(q), tab.push_back(nb), tab.push_back(slot);
                }
            }
        };
        seg(m_, rtab_, rfirst_, rcount_);
        seg(n_, ctab_, cfirst_, ccount_);
        nrseg_ = int(rfirst_.empty() ? 0 : rtab_.size() / 4), ncseg_ = int(ctab_.size() / 4);
    }

    BMeta* up_meta(const std::vector<BMeta>& v) {
        BMeta* p = mem_.alloc<BMeta>(std::max<size_t>(v.size(), 1));
        if (!v.empty()) HIP_OK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(BMeta), hipMemcpyHostToDevice, st_));
        HIP_OK(hipStreamSynchronize(st_));
        return p;
    }
    int* up_int(const std::vector<int32_t>& v) {
        int* p = mem_.alloc<int>(std::max<size_t>(v.size(), 1));
        if (!v.empty()) HIP_OK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice, st_));
        HIP_OK(hipStreamSynchronize(st_));
        return p;
    }
    template <class U>
    U* up_vals(const std::vector<double>& v) {
        std::vector<U> h(v.size());
        for (size_t k = 0; k < v.size(); ++k) h[k] = U(v[k]);
        U* p = mem_.alloc<U>(std::max<size_t>(v.size(), 1));
        if (!h.empty()) HIP_OK(hipMemcpyAsync(p, h.data(), h.size() * sizeof(U), hipMemcpyHostToDevice, st_));
        HIP_OK(hipStreamSynchronize(st_));
        return p;
    }
    double* cat64(const std::vector<const Scaled*>& probs, int rowside, std::vector<double> Scaled::*f) {
        std::vector<double> h(static_cast<size_t>(rowside ? mtot_ : ntot_));
        for (int i = 0; i < B_; ++i) {
            const std::vector<double>& v = probs[size_t(i)]->*f;
            std::copy(v.begin(), v.end(), h.begin() + (rowside ? roff_[size_t(i)] : coff_[size_t(i)]));
        }
        double* p = mem_.alloc<double>(std::max<size_t>(h.size(), 1));
        HIP_OK(hipMemcpyAsync(p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, st_));
        HIP_OK(hipStreamSynchronize(st_));
        return p;
    }
    T* zerosT(int k) {
        T* p = mem_.alloc<T>(size_t(std::max(k, 1)));
        HIP_OK(hipMemsetAsync(p, 0, size_t(std::max(k, 1)) * sizeof(T), st_));
        return p;
    }
    double* zeros64(int k) {
        double* p = mem_.alloc<double>(size_t(std::max(k, 1)));
        HIP_OK(hipMemsetAsync(p, 0, size_t(std::max(k, 1)) * sizeof(double), st_));
        return p;
    }

    void upload_all(const std::vector<const Scaled*>& probs) {
        // matrices: pointers and indices once; values in the iterate type and in FP64 (shared when T is double)
        // (sorted pools get their index and value arrays per stage, build_stage)
        A_.ptr = up_int(PA_.ptr), AT_.ptr = up_int(PT_.ptr);
        A_.rows_total = mtot_, AT_.rows_total = ntot_;
        A64_.ptr = A_.ptr, AT64_.ptr = AT_.ptr;
        if (!sorted_) {
            base_idx_a_ = up_int(PA_.idx), base_idx_t_ = up_int(PT_.idx);
            base_val_a_ = up_vals<T>(PA_.val), base_val_t_ = up_vals<T>(PT_.val);
            if constexpr (sizeof(T) == sizeof(double)) {
                base_val64_a_ = reinterpret_cast<double*>(base_val_a_), base_val64_t_ = reinterpret_cast<double*>(base_val_t_);
            } else {
                base_val64_a_ = up_vals<double>(PA_.val), base_val64_t_ = up_vals<double>(PT_.val);
            }
        }
        nnz_of_.resize(size_t(B_));
        for (int i = 0; i < B_; ++i) nnz_of_[size_t(i)] = probs[size_t(i)]->A.nnz(), nnz_all_ += nnz_of_[size_t(i)];
        d_roff_ = up_int(roff_), d_coff_ = up_int(coff_);
        d_m_ = up_int(m_), d_n_ = up_int(n_);
        d_rtab_ = up_int(rtab_), d_ctab_ = up_int(ctab_);
        d_rlive_ = mem_.alloc<int>(std::max<size_t>(rtab_.size(), 4)), d_clive_ = mem_.alloc<int>(std::max<size_t>(ctab_.size(), 4));
        d_rfirst_ = up_int(rfirst_), d_rcount_ = up_int(rcount_), d_cfirst_ = up_int(cfirst_), d_ccount_ = up_int(ccount_);
        {
            std::vector<int32_t> cf(cfirst_);
            for (int32_t& v : cf) v += nrseg_;  // column partials follow the row partials in part_
            d_cfirst_after_rows_ = up_int(cf);
        }
        c64_ = cat64(probs, 0, &Scaled::c), l64_ = cat64(probs, 0, &Scaled::l), u64_ = cat64(probs, 0, &Scaled::u);
        Dc64_ = cat64(probs, 0, &Scaled::Dc);
        rl64_ = cat64(probs, 1, &Scaled::rl), ru64_ = cat64(probs, 1, &Scaled::ru), Dr64_ = cat64(probs, 1, &Scaled::Dr);
        xb_ = zeros64(ntot_), zb_ = zeros64(mtot_), Axb_ = zeros64(mtot_), ATzb_ = zeros64(ntot_);
        int maxlen = 1;
        for (int i = 0; i < B_; ++i) maxlen = std::max(maxlen, std::max(m_[size_t(i)], n_[size_t(i)]));
        buf64_ = zeros64(maxlen);
        c_in_ = zerosT(ntot_), l_in_ = zerosT(ntot_), u_in_ = zerosT(ntot_);
        rl_in_ = zerosT(mtot_), ru_in_ = zerosT(mtot_), zh_ = zerosT(mtot_);
        x_ = zerosT(ntot_), xp_ = zerosT(ntot_), w_ = zerosT(ntot_), x0_ = zerosT(ntot_), ATz_ = zerosT(ntot_);
        z_ = zerosT(mtot_), z0_ = zerosT(mtot_), Ax_ = zerosT(mtot_);
        pv_ = zerosT(ntot_), pw_ = zerosT(mtot_), pu_ = zerosT(ntot_);
        par_ = zerosT(2 * B_);
        kc_ = mem_.alloc<int>(size_t(B_));
        HIP_OK(hipMemsetAsync(kc_, 0, size_t(B_) * sizeof(int), st_));
        fac_ = zeros64(B_);
        part_ = mem_.alloc<double>(size_t(std::max(1, nrseg_ + ncseg_)) * kEvalK);
        out_ = zeros64(B_ * 2 * kEvalK);
        dflags_ = mem_.alloc<uint8_t>(size_t(B_)), dact_ = mem_.alloc<uint8_t>(size_t(B_));
        HIP_OK(hipMemsetAsync(dflags_, 0, size_t(B_), st_));
        HIP_OK(hipMemsetAsync(dact_, 0, size_t(B_), st_));
        HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&hflags_), size_t(B_), hipHostMallocDefault));
        HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&hpar_), size_t(2 * B_) * sizeof(T), hipHostMallocDefault));
        HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&hout_), size_t(B_) * 2 * kEvalK * sizeof(double),
                             hipHostMallocDefault));
        HIP_OK(hipStreamSynchronize(st_));
    }

    // One row-block stage: the block records of both sides (device and host copies), each instance's range of
    // records, and, for sorted pools, the index and value arrays permuted for this stage's blocks (every block's
    // nonzeros in column order with packed row positions). Every stage built stays resident, so the batch can
    // change stage between iteration blocks without touching the host.
    struct Stage {
        BMeta *meta_a = nullptr, *meta_t = nullptr;
        std::vector<BMeta> hmeta_a, hmeta_t;
        std::vector<int32_t> afirst, acount, tfirst, tcount;
        int *idx_a = nullptr, *idx_t = nullptr;
        T *val_a = nullptr, *val_t = nullptr;
        double *val64_a = nullptr, *val64_t = nullptr;
    };

    // Sorted pool of one side for a stage: indices packed with row positions and values in the iterate type and
    // in FP64, permuted by the sort, in new device arrays.
    void repack(const Pool& pool, const std::vector<std::vector<int32_t>>& per, int cap, int*& idx, T*& val,
                double*& val64) {
        const size_t nnz = pool.val.size();
        std::vector<int32_t> pk(nnz);
        std::vector<int64_t> perm(nnz);
        for (size_t u = 0; u < pool.rows.size(); ++u) {
            // blocks hold local rows; pool pointers are absolute, so sort_blocks sees absolute nonzero ranges
            sort_blocks(per[u], pool.ptr.data() + pool.pbase[u], pool.idx.data(), 0, nullptr, kBatchColBits, pk, perm,
                        cap);
        }
        std::vector<double> pv(nnz);
        parallel_for(int64_t(nnz), 1 << 16, [&](int64_t k) { pv[size_t(k)] = pool.val[size_t(perm[size_t(k)])]; });
        idx = up_int(pk);
        std::vector<T> tv(nnz);
        parallel_for(int64_t(nnz), 1 << 16, [&](int64_t k) { tv[size_t(k)] = T(pv[size_t(k)]); });
        val = mem_.alloc<T>(std::max<size_t>(nnz, 1));
        if (nnz) HIP_OK(hipMemcpyAsync(val, tv.data(), nnz * sizeof(T), hipMemcpyHostToDevice, st_));
        if constexpr (sizeof(T) != sizeof(double)) {
            val64 = mem_.alloc<double>(std::max<size_t>(nnz, 1));
            if (nnz) HIP_OK(hipMemcpyAsync(val64, pv.data(), nnz * sizeof(double), hipMemcpyHostToDevice, st_));
        } else {
            val64 = reinterpret_cast<double*>(val);
        }
        HIP_OK(hipStreamSynchronize(st_));
    }

    Stage& build_stage(int cap) {
        Stage sg;
        auto table = [&](const Pool& pool, bool a_side) {
            std::vector<std::vector<int32_t>> per(pool.rows.size());
            for (size_t u = 0; u < pool.rows.size(); ++u) {
                // local row pointers of distinct matrix u
                std::vector<int32_t> lp(static_cast<size_t>(pool.rows[u]) + 1);
                const int32_t b0 = pool.ptr[size_t(pool.pbase[u])];
                for (int r = 0; r <= pool.rows[u]; ++r) lp[size_t(r)] = pool.ptr[size_t(pool.pbase[u]) + size_t(r)] - b0;
                per[u] = sorted_ ? sorted_blocks(pool.rows[u], lp.data(), cap, cap)
                                 : row_blocks_cap(pool.rows[u], lp.data(), cap);
            }
            if (sorted_) {
                if (a_side) repack(pool, per, cap, sg.idx_a, sg.val_a, sg.val64_a);
                else repack(pool, per, cap, sg.idx_t, sg.val_t, sg.val64_t);
            }
            std::vector<int32_t> t;
            for (int i = 0; i < B_; ++i) {
                const std::vector<int32_t>& b = per[size_t(mat_of_[size_t(i)])];
                for (size_t k = 0; k + 1 < b.size(); ++k) t.push_back(i), t.push_back(b[k]), t.push_back(b[k + 1]);
            }
            return t;
        };
        const std::vector<int32_t> ta = table(PA_, true), tt = table(PT_, false);
        if (!sorted_) {  // row-block pools are not permuted: every stage reads the base arrays
            sg.idx_a = base_idx_a_, sg.idx_t = base_idx_t_, sg.val_a = base_val_a_, sg.val_t = base_val_t_;
            sg.val64_a = base_val64_a_, sg.val64_t = base_val64_t_;
        }
        // block records: A's rows gather the column-side vector and write the row side, A^T's the reverse
        auto records = [&](const std::vector<int32_t>& t, const Pool& pool, const std::vector<int32_t>& voff,
                           const std::vector<int32_t>& off) {
            std::vector<BMeta> m(t.size() / 3);
            for (size_t k = 0; k < m.size(); ++k) {
                const int i = t[3 * k], r0 = t[3 * k + 1], r1 = t[3 * k + 2];
                const int pb = pool.pbase[size_t(mat_of_[size_t(i)])];
                m[k] = BMeta{i, r0, r1, pool.ptr[size_t(pb + r0)], pool.ptr[size_t(pb + r1)], pb, voff[size_t(i)],
                             off[size_t(i)]};
            }
            return m;
        };
        sg.hmeta_a = records(ta, PA_, coff_, roff_), sg.hmeta_t = records(tt, PT_, roff_, coff_);
        auto ranges = [&](const std::vector<BMeta>& m, std::vector<int32_t>& first, std::vector<int32_t>& count) {
            first.assign(size_t(B_), 0), count.assign(size_t(B_), 0);
            for (size_t k = 0; k < m.size(); ++k) {
                const int i = m[k].inst;
                if (count[size_t(i)] == 0) first[size_t(i)] = int32_t(k);
                count[size_t(i)]++;
            }
        };
        ranges(sg.hmeta_a, sg.afirst, sg.acount), ranges(sg.hmeta_t, sg.tfirst, sg.tcount);
        sg.meta_a = up_meta(sg.hmeta_a), sg.meta_t = up_meta(sg.hmeta_t);
        {  // live-block list buffers, grown to the largest table seen
            const size_t na = std::max<size_t>(sg.hmeta_a.size(), 1), nt = std::max<size_t>(sg.hmeta_t.size(), 1);
            if (na > list_cap_a_) alist_ = mem_.alloc<BMeta>(na), list_cap_a_ = na;
            if (nt > list_cap_t_) tlist_ = mem_.alloc<BMeta>(nt), list_cap_t_ = nt;
            if (na + nt > hlist_cap_) {
                if (hlist_) HIP_OK(hipHostFree(hlist_));
                hlist_ = nullptr;
                HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&hlist_), (na + nt) * sizeof(BMeta), hipHostMallocDefault));
                hlist_cap_ = na + nt;
            }
        }
        return stages_.emplace(cap, std::move(sg)).first->second;
    }

    // Makes stage `cap` current (building it first if needed); the live lists and the captured graph start over.
    void set_cap(int cap) {
        auto it = stages_.find(cap);
        Stage& sg = it == stages_.end() ? build_stage(cap) : it->second;
        A_.meta = A64_.meta = sg.meta_a, AT_.meta = AT64_.meta = sg.meta_t;
        A_.idx = A64_.idx = sg.idx_a, AT_.idx = AT64_.idx = sg.idx_t;
        A_.val = sg.val_a, AT_.val = sg.val_t, A64_.val = sg.val64_a, AT64_.val = sg.val64_t;
        A_.nblk = A64_.nblk = int(sg.hmeta_a.size()), AT_.nblk = AT64_.nblk = int(sg.hmeta_t.size());
        cur_ = &sg;
        cap_ = cap;
        list_inst_.clear(), nalist_ = ntlist_ = 0;
        if (graph_) {
            (void)hipGraphExecDestroy(graph_);
            graph_ = nullptr;
        }
    }

    // Stage for a live set of `nnz` nonzeros (per side) when the stage was autotuned: the tuned stage while all
    // instances are live or while the live nonzeros still make kStageFill blocks of it; below that the largest
    // smaller stage that still makes kStageFill blocks, else the smallest. A nearly idle GPU runs the few blocks of
    // a batch's tail faster when they are smaller (one live 13k x 28k LP: 10.5 us per iteration with 1,024-nonzero
    // blocks against 15.8 us with 8,192; eight 8k-14k-row LPs: 18.5 us with 2,048 against 20.0 us).
    int stage_for(int64_t nnz, bool all_live) const {
        if (!auto_stage_) return cap_;  // a given stage, or the autotuner timing each stage in turn
        if (all_live || nnz >= int64_t(kStageFill) * tuned_cap_) return tuned_cap_;
        for (int c = tuned_cap_ / 2; c >= 1024; c /= 2)
            if (nnz >= int64_t(kStageFill) * c) return c;
        return 1024;
    }

    // Times a few iteration blocks for each stage size and keeps the fastest (all instances active).
    void autotune() {
        std::vector<uint8_t> all(static_cast<size_t>(B_), 1);
        const bool saved_graph = o_.use_graph;
        o_.use_graph = false;
        hipEvent_t e0, e1;
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        float best = 0;
        int best_cap = kCap;
        for (int cap : {1024, 2048, 4096, 8192}) {
            set_cap(cap);
            upload_act(all);
            ensure_lists(all);
            launch_block(4);  // warm-up
            HIP_OK(hipEventRecord(e0, st_));
            for (int r = 0; r < 3; ++r) launch_block(8);
            HIP_OK(hipEventRecord(e1, st_));
            HIP_OK(hipEventSynchronize(e1));
            float ms = 0;
            HIP_OK(hipEventElapsedTime(&ms, e0, e1));
            tune_ms_.push_back(ms);
            if (cap == 1024 || ms < best) best = ms, best_cap = cap;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        o_.use_graph = saved_graph;
        set_cap(best_cap);
        tuned_cap_ = best_cap;
        HIP_OK(hipMemsetAsync(kc_, 0, size_t(B_) * sizeof(int), st_));
        HIP_OK(hipStreamSynchronize(st_));
    }

    // Persistent blocks: used while both live block lists fit one co-resident grid of at most persist_max_
    // workgroups (TPDHG_PERSIST_MAX; 0 turns the path off), never under per-launch profiling.
    template <int CAP, bool SRT>
    const void* persist_kernel() {
        return reinterpret_cast<const void*>(&k_bpersist<T, CAP, SRT>);
    }
    const void* persist_kernel_for_cap() {
        switch (cap_) {
            case 1024: return sorted_ ? persist_kernel<1024, true>() : persist_kernel<1024, false>();
            case 4096: return sorted_ ? persist_kernel<4096, true>() : persist_kernel<4096, false>();
            case 8192: return sorted_ ? persist_kernel<8192, true>() : persist_kernel<8192, false>();
            default: return sorted_ ? persist_kernel<2048, true>() : persist_kernel<2048, false>();
        }
    }
    int persist_grid_max() {  // co-resident workgroups of the persistent kernel at the current stage
        auto it = persist_cap_.find(cap_);
        if (it != persist_cap_.end()) return it->second;
        int dev = 0, cus = 0, coop = 0, per_cu = 0;
        HIP_OK(hipGetDevice(&dev));
        HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        HIP_OK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
        if (coop) HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persist_kernel_for_cap(), stage_wg(cap_), 0));
        const int g = coop ? per_cu * cus : 0;
        persist_cap_[cap_] = g;
        return g;
    }
    bool persistent_fits() {
        if (!persist_ok_ || persist_max_ <= 0 || profile_) return false;
        const int need = std::max(nalist_, ntlist_);
        return need > 0 && need <= persist_max_ && need <= persist_grid_max();
    }
    void launch_persistent(int steps) {
        if (!zs_) {
            zs_ = mem_.alloc<T>(size_t(std::max(mtot_, 1)));
            xs_ = mem_.alloc<T>(size_t(std::max(ntot_, 1))), xps_ = mem_.alloc<T>(size_t(std::max(ntot_, 1)));
            ws_ = mem_.alloc<T>(size_t(std::max(ntot_, 1)));
            bar_ = mem_.alloc<unsigned>(1), err_ = mem_.alloc<int>(1);
            HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&herr_), sizeof(int), hipHostMallocDefault));
        }
        const Side& a = A_;
        const Side& t = AT_;
        PersistArgs<T> pa{steps, nalist_, ntlist_, alist_, tlist_, a.ptr, a.idx, t.ptr, t.idx, a.val, t.val,
                          dact_, row_epi(0), col_epi(0, 0, false), ATz_, zs_, xs_, xps_, ws_,
                          GridBar{bar_, err_, persist_spins_, persist_unreachable_}};
        k_reset_bar<<<1, 1, 0, st_>>>(bar_, err_);
        const int G = std::max(nalist_, ntlist_);
        void* args[] = {&pa};
        const hipError_t e = hipLaunchCooperativeKernel(persist_kernel_for_cap(), dim3(unsigned(G)),
                                                        dim3(unsigned(stage_wg(cap_))), args, 0, st_);
        if (e != hipSuccess) {  // not launchable as one co-resident grid here: lockstep from now on
            (void)hipGetLastError();
            persist_ok_ = false;
            std::fprintf(stderr, "batch: persistent blocks off (cooperative launch of %d workgroups: %s)\n", G,
                         hipGetErrorString(e));
            ++lockstep_blocks_;
            launch_block(steps);
            return;
        }
        HIP_OK(hipMemcpyAsync(herr_, err_, sizeof(int), hipMemcpyDeviceToHost, st_));
        HIP_OK(hipStreamSynchronize(st_));
        if (*herr_) {  // a grid barrier gave up: undo the partial block and run it in lockstep
            persist_ok_ = false;
            std::fprintf(stderr, "batch: a persistent block's grid barrier timed out; restored and replayed in "
                                 "lockstep (persistent blocks off for this batch)\n");
            k_brestore<T><<<nalist_ + ntlist_, kWG, 0, st_>>>(pa);
            HIP_OK(hipGetLastError());
            ++lockstep_blocks_;
            launch_block(steps);
            return;
        }
        k_bset_counters<<<(B_ + kWG - 1) / kWG, kWG, 0, st_>>>(B_, dact_, kc_, steps, 0);
        HIP_OK(hipGetLastError());
        ++persistent_blocks_;
    }

    Seg rseg() const { return Seg{d_rtab_, d_roff_, d_m_}; }
    Seg cseg() const { return Seg{d_ctab_, d_coff_, d_n_}; }
    // Whether the live lists hold every instance flagged in `on`.
    bool covers(const std::vector<uint8_t>& on) const {
        bool c = list_inst_.size() == size_t(B_);
        for (int i = 0; i < B_ && c; ++i) c = !on[size_t(i)] || list_inst_[size_t(i)];
        return c;
    }
    // Row / column entries for a launch over the instances flagged in `on`: the live list when it covers them
    // (a batch's tail then launches a few workgroups instead of the whole table), else the whole table.
    std::pair<Seg, int> rsegs(const std::vector<uint8_t>& on) const {
        return covers(on) ? std::make_pair(Seg{d_rlive_, d_roff_, d_m_}, nrlive_) : std::make_pair(rseg(), nrseg_);
    }
    std::pair<Seg, int> csegs(const std::vector<uint8_t>& on) const {
        return covers(on) ? std::make_pair(Seg{d_clive_, d_coff_, d_n_}, nclive_) : std::make_pair(cseg(), ncseg_);
    }

    void upload_flags(const std::vector<uint8_t>& on) {
        std::copy(on.begin(), on.end(), hflags_);
        HIP_OK(hipMemcpyAsync(dflags_, hflags_, size_t(B_), hipMemcpyHostToDevice, st_));
        HIP_OK(hipStreamSynchronize(st_));
    }
    void upload_act(const std::vector<uint8_t>& on) {
        std::copy(on.begin(), on.end(), hflags_);
        HIP_OK(hipMemcpyAsync(dact_, hflags_, size_t(B_), hipMemcpyHostToDevice, st_));
        HIP_OK(hipStreamSynchronize(st_));
    }

    template <class S, class U, class Epi, int TAG = 0>
    void spmv_rows(const S& side, const U* v, const Epi& epi, const uint8_t* on, const BMeta* list = nullptr,
                   int nlist = 0) {
        const int nb = list ? nlist : side.nblk;
        if (nb == 0) return;
#define TPDHG_BB(CAPV, SRT)                                                                                      \
    k_bblocks<U, CAPV, SRT, Epi, stage_wg(CAPV), TAG><<<nb, stage_wg(CAPV), 0, st_>>>(                          \
        nb, list ? list : side.meta, side.ptr, side.idx, side.val, v, on, epi)
        if (sorted_) {
            switch (cap_) {
                case 1024: TPDHG_BB(1024, true); break;
                case 4096: TPDHG_BB(4096, true); break;
                case 8192: TPDHG_BB(8192, true); break;
                default: TPDHG_BB(2048, true);
            }
        } else {
            switch (cap_) {
                case 1024: TPDHG_BB(1024, false); break;
                case 4096: TPDHG_BB(4096, false); break;
                case 8192: TPDHG_BB(8192, false); break;
                default: TPDHG_BB(2048, false);
            }
        }
#undef TPDHG_BB
    }
    // A product over the instances flagged in `on` (dflags_ holds the same flags): over the live lists when they
    // hold every flagged instance (a batch's tail then launches a few blocks instead of the whole table), else over
    // the side's whole table. a_side: the list of A (else of A^T).
    template <class S, class U, class Epi>
    void spmv_flagged(const S& side, const U* v, const Epi& epi, const std::vector<uint8_t>& on, bool a_side) {
        if (covers(on)) spmv_rows(side, v, epi, dflags_, a_side ? alist_ : tlist_, a_side ? nalist_ : ntlist_);
        else spmv_rows(side, v, epi, dflags_);
    }

    // Live-block lists of both sides for the instances in `on`: rebuilt when an instance in `on` is missing from
    // the current lists or when the live blocks have dropped to half the lists' length (so at most ~log2 of the
    // block count rebuilds per solve); a rebuild drops the captured graph, whose launches have the old lengths.
    // With an autotuned stage the stage follows the live set first (stage_for), which also starts the lists over.
    void ensure_lists(const std::vector<uint8_t>& on) {
        int64_t live_nnz = 0;
        bool all_live = true;
        for (int i = 0; i < B_; ++i) {
            if (on[size_t(i)]) live_nnz += nnz_of_[size_t(i)];
            else all_live = false;
        }
        const int want = stage_for(live_nnz, all_live);
        if (want != cap_) {
            set_cap(want);
            ++stage_switches_;
        }
        const Stage& sg = *cur_;
        int64_t live_a = 0;
        bool missing = list_inst_.size() != size_t(B_);
        for (int i = 0; i < B_ && !missing; ++i) missing = on[size_t(i)] && !list_inst_[size_t(i)];
        for (int i = 0; i < B_; ++i)
            if (on[size_t(i)]) live_a += sg.acount[size_t(i)];
        if (!missing && 2 * live_a > int64_t(nalist_)) return;
        std::vector<BMeta> la, lt;
        list_inst_.assign(size_t(B_), 0);
        for (int i = 0; i < B_; ++i) {
            if (!on[size_t(i)]) continue;
            list_inst_[size_t(i)] = 1;
            for (int k = 0; k < sg.acount[size_t(i)]; ++k) la.push_back(sg.hmeta_a[size_t(sg.afirst[size_t(i)] + k)]);
            for (int k = 0; k < sg.tcount[size_t(i)]; ++k) lt.push_back(sg.hmeta_t[size_t(sg.tfirst[size_t(i)] + k)]);
        }
        std::copy(la.begin(), la.end(), hlist_);
        std::copy(lt.begin(), lt.end(), hlist_ + la.size());
        if (!la.empty()) HIP_OK(hipMemcpyAsync(alist_, hlist_, la.size() * sizeof(BMeta), hipMemcpyHostToDevice, st_));
        if (!lt.empty())
            HIP_OK(hipMemcpyAsync(tlist_, hlist_ + la.size(), lt.size() * sizeof(BMeta), hipMemcpyHostToDevice, st_));
        HIP_OK(hipStreamSynchronize(st_));
        nalist_ = int(la.size()), ntlist_ = int(lt.size());
        {  // the live instances' reduction / elementwise entries (their slots unchanged)
            std::vector<int32_t> rl, cl;
            for (int i = 0; i < B_; ++i) {
                if (!list_inst_[size_t(i)]) continue;
                rl.insert(rl.end(), rtab_.begin() + 4 * rfirst_[size_t(i)], rtab_.begin() + 4 * (rfirst_[size_t(i)] + rcount_[size_t(i)]));
                cl.insert(cl.end(), ctab_.begin() + 4 * cfirst_[size_t(i)], ctab_.begin() + 4 * (cfirst_[size_t(i)] + ccount_[size_t(i)]));
            }
            if (!rl.empty()) HIP_OK(hipMemcpyAsync(d_rlive_, rl.data(), rl.size() * sizeof(int), hipMemcpyHostToDevice, st_));
            if (!cl.empty()) HIP_OK(hipMemcpyAsync(d_clive_, cl.data(), cl.size() * sizeof(int), hipMemcpyHostToDevice, st_));
            HIP_OK(hipStreamSynchronize(st_));
            nrlive_ = int(rl.size() / 4), nclive_ = int(cl.size() / 4);
        }
        ++list_builds_;
        if (graph_) {
            (void)hipGraphExecDestroy(graph_);
            graph_ = nullptr;
        }
    }
    BEpiHCol<T> col_epi(int s, int mode, bool store) {
        return BEpiHCol<T>{x_, xp_, w_, store ? ATz_ : nullptr, x0_, c_in_, l_in_, u_in_, par_, kc_, s, mode,
                           T(o_.rho), 0, T(0), T(0)};
    }
    BEpiHRow<T> row_epi(int s) {
        return BEpiHRow<T>{z_, z0_, zh_, rl_in_, ru_in_, par_, kc_, s, T(o_.rho), 0, T(0), T(0)};
    }
    hipEvent_t event(size_t i) {
        while (events_.size() <= i) {
            hipEvent_t e;
            HIP_OK(hipEventCreate(&e));
            events_.push_back(e);
        }
        return events_[i];
    }
    void launch_block(int steps, bool probe = false) {
        for (int s = 0; s < steps; ++s) {
            const BEpiHRow<T> er = row_epi(s);
            const BEpiHCol<T> ec = col_epi(s, 0, s + 1 == steps);
            if (probe) {
                const size_t e = pending_;
                HIP_OK(hipEventRecord(event(e), st_));
                spmv_rows<Side, T, BEpiHRow<T>, 1>(A_, w_, er, dact_, alist_, nalist_);
                HIP_OK(hipEventRecord(event(e + 1), st_));
                spmv_rows<Side, T, BEpiHCol<T>, 1>(AT_, z_, ec, dact_, tlist_, ntlist_);
                HIP_OK(hipEventRecord(event(e + 2), st_));
                pending_ += 3;
            } else if (profile_) {
                const size_t e = pending_;
                HIP_OK(hipEventRecord(event(e), st_));
                spmv_rows(A_, w_, er, dact_, alist_, nalist_);
                HIP_OK(hipEventRecord(event(e + 1), st_));
                spmv_rows(AT_, z_, ec, dact_, tlist_, ntlist_);
                HIP_OK(hipEventRecord(event(e + 2), st_));
                pending_ += 3;
            } else {
                spmv_rows(A_, w_, er, dact_, alist_, nalist_);
                spmv_rows(AT_, z_, ec, dact_, tlist_, ntlist_);
            }
        }
        k_bset_counters<<<(B_ + kWG - 1) / kWG, kWG, 0, st_>>>(B_, dact_, kc_, steps, 0);
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

    int B_;
    BatchOptions o_;
    bool persist_ok_ = true;
    int persist_max_ = 256, persist_spins_ = 1 << 20;
    unsigned persist_unreachable_ = 0;
    std::map<int, int> persist_cap_;
    int64_t persistent_blocks_ = 0, lockstep_blocks_ = 0;
    T *zs_ = nullptr, *xs_ = nullptr, *xps_ = nullptr, *ws_ = nullptr;
    unsigned* bar_ = nullptr;
    int *err_ = nullptr, *herr_ = nullptr;
    bool profile_ = false;
    Mem mem_;
    hipStream_t st_ = nullptr;
    hipGraphExec_t graph_ = nullptr;
    std::vector<int32_t> m_, n_, roff_, coff_, mat_of_;
    int mtot_ = 0, ntot_ = 0;
    int64_t nnz_total_ = 0;
    Pool PA_, PT_;
    std::vector<int32_t> rtab_, ctab_, rfirst_, rcount_, cfirst_, ccount_;
    int nrseg_ = 0, ncseg_ = 0, cap_ = kCap, tuned_cap_ = kCap;
    static constexpr int kStageFill = 256;  // blocks that keep every CU busy
    bool auto_stage_ = false;               // the stage was autotuned, so it may change with the live set
    std::map<int, Stage> stages_;
    Stage* cur_ = nullptr;
    std::vector<int64_t> nnz_of_;            // nonzeros of each instance's A
    int64_t nnz_all_ = 0;
    int *base_idx_a_ = nullptr, *base_idx_t_ = nullptr;
    T *base_val_a_ = nullptr, *base_val_t_ = nullptr;
    double *base_val64_a_ = nullptr, *base_val64_t_ = nullptr;
    bool sorted_ = false;  // sorted pools (every instance below 2^19 rows and columns)
    std::vector<uint8_t> list_inst_;
    BMeta *alist_ = nullptr, *tlist_ = nullptr, *hlist_ = nullptr;
    size_t list_cap_a_ = 0, list_cap_t_ = 0, hlist_cap_ = 0;
    int nalist_ = 0, ntlist_ = 0;
    int64_t list_builds_ = 0, stage_switches_ = 0;
    std::vector<float> tune_ms_;
    Side A_, AT_;
    Side64 A64_, AT64_;
    int *d_roff_ = nullptr, *d_coff_ = nullptr, *d_m_ = nullptr, *d_n_ = nullptr;
    int *d_rtab_ = nullptr, *d_ctab_ = nullptr, *d_rfirst_ = nullptr, *d_rcount_ = nullptr;
    int *d_rlive_ = nullptr, *d_clive_ = nullptr;
    int nrlive_ = 0, nclive_ = 0;
    int *d_cfirst_ = nullptr, *d_cfirst_after_rows_ = nullptr, *d_ccount_ = nullptr;
    double *c64_ = nullptr, *l64_ = nullptr, *u64_ = nullptr, *Dc64_ = nullptr, *rl64_ = nullptr, *ru64_ = nullptr,
           *Dr64_ = nullptr;
    double *xb_ = nullptr, *zb_ = nullptr, *Axb_ = nullptr, *ATzb_ = nullptr, *buf64_ = nullptr, *fac_ = nullptr;
    T *c_in_ = nullptr, *l_in_ = nullptr, *u_in_ = nullptr, *rl_in_ = nullptr, *ru_in_ = nullptr, *zh_ = nullptr;
    T *x_ = nullptr, *xp_ = nullptr, *w_ = nullptr, *x0_ = nullptr, *ATz_ = nullptr, *z_ = nullptr, *z0_ = nullptr,
      *Ax_ = nullptr;
    T *pv_ = nullptr, *pw_ = nullptr, *pu_ = nullptr, *par_ = nullptr, *hpar_ = nullptr;
    int* kc_ = nullptr;
    double *part_ = nullptr, *out_ = nullptr, *hout_ = nullptr;
    std::vector<double> fp_;  // fixed-point residual terms of the last eval, 3 per instance
    uint8_t *dflags_ = nullptr, *dact_ = nullptr, *hflags_ = nullptr;
    size_t mem_used_ = 0;
    std::string name_;
    std::vector<hipEvent_t> events_;
    size_t pending_ = 0;
    KernelTimes times_;
};

}  // namespace

std::unique_ptr<BatchDevice> make_batch_device(const std::vector<const Scaled*>& probs, const BatchOptions& opt) {
    for (const Scaled* p : probs)
        if (!p) throw std::invalid_argument("null batch instance");
    if (opt.precision_bits == 32) return std::unique_ptr<BatchDevice>(new BatchDev<float>(probs, opt));
    if (opt.precision_bits == 64) return std::unique_ptr<BatchDevice>(new BatchDev<double>(probs, opt));
    throw std::runtime_error("precision must be 32 or 64 bits");
}

}  // namespace tpdhg
