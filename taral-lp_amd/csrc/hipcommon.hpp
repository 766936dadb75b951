// Shared HIP building blocks of the PDHG device backends (gfx950, wave64): error checking, the LDS-staged
// row-block SpMV kernel with pluggable epilogues, deterministic block reductions, device buffers and CSR upload.
// Included by each backend translation unit; everything has internal linkage.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "parallel.hpp"
#include "prep.hpp"

namespace tpdhg {
namespace {

constexpr int kWG = 256;    // threads per workgroup: 4 waves of 64
constexpr int kCap = 2048;  // nonzeros staged per row block (8 per thread)
constexpr int kMaxRedBlocks = 1024;
constexpr int kSlabsK = 8;  // XCDs: the slab layout uses a multiple of this many column slabs

#define HIP_OK(call)                                                                                      \
    do {                                                                                                  \
        hipError_t e_ = (call);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " (" #call ")"); \
    } while (0)

// ------------------------------------------------------------------------------------------------ kernels
__device__ __forceinline__ int xcd_block(int b, int nblk) {
    const int q = nblk >> 3, r = nblk & 7, x = b & 7, i = b >> 3;
    return x * q + (x < r ? x : r) + i;
}

template <class T>
__device__ __forceinline__ T clampv(T v, T lo, T hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

// Dual prox step of a recentred row (the fast path's row step before its Halpern blend): for dual z, row product
// a, base dual h, shifted bounds [lo, hi] and step sig, z+ = v - sig * clamp((v + h) / sig, lo, hi) with
// v = z + sig * a, written casewise (the middle case is exactly -h).
template <class T>
__device__ __forceinline__ T row_prox(T zi, T a, T h, T lo, T hi, T sig) {
    const T v = zi + sig * a, q = (v + h) / sig;
    return q < lo ? v - sig * lo : (q > hi ? v - sig * hi : -h);
}

// Epilogues of the row-block kernels: load() reads the launch's scalars, pre(r) the operands of row r and
// apply(r, sum, pre) finishes the row; the block kernels call pre before their gathers so that the operands'
// latency overlaps them, and operator()(r, sum) does both at once.
template <class T>
struct EpiStore {  // y = M v
    T* y;
    struct Pre {};
    __device__ void load() {}
    __device__ Pre pre(int) const { return {}; }
    __device__ void apply(int r, T s, const Pre&) const { y[r] = s; }
    __device__ void operator()(int r, T s) const { y[r] = s; }
};

// Gathers of nonzeros [k0, k1), U per thread per round (nonzero base + u * WG + tid): each lane first loads its U
// index words and values, then its U gathered entries, all at indices clamped into [k0, k1) so that no load sits
// behind a bounds test, and only then hands the in-range products to use(k, word, product) in the order of k.
// With a bounds test around each load the compiler waits for every gather before issuing the next one (one load
// chain in flight per wave); this way a wave keeps 2U and then U loads in flight.
template <int WG, int U, class T, class W, class Col, class Use>
__device__ __forceinline__ void gather_rounds(int k0, int k1, const W* __restrict__ w, const T* __restrict__ val,
                                              const T* __restrict__ v, Col col, Use use) {
    if (k1 <= k0) return;
    const int tid = int(threadIdx.x);
    for (int base = k0; base < k1; base += U * WG) {
        W p[U];
        T a[U], g[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = min(base + u * WG + tid, k1 - 1);
            p[u] = w[k];
            a[u] = val[k];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) g[u] = v[col(p[u])];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (base + u * WG + tid < k1) use(base + u * WG + tid, p[u], a[u] * g[u]);
    }
}

struct ColPlain {
    __device__ int operator()(int c) const { return c; }
};
struct ColPacked {
    uint32_t mask;
    __device__ uint32_t operator()(uint32_t p) const { return p & mask; }
};

// Row sums of a staged block (products of rows [r0, r1) in CSR order at sh[off(r) ...], off(r1) = the block's
// length) by a workgroup of WG threads: blocks of at most WG rows give each row a power-of-two lane group (xor
// shuffles inside a wave), larger blocks one thread per row.
template <class T, int WG = kWG, class Off, class Out>
__device__ __forceinline__ void block_row_sums_at(Off&& off, int r0, int r1, const T* sh, Out&& out) {
    const int tid = int(threadIdx.x), nr = r1 - r0;
    if (nr <= WG) {
        int lg = 6;
        while (lg > 0 && (nr << lg) > WG) --lg;
        const int row = tid >> lg, lane = tid & ((1 << lg) - 1), lpr = 1 << lg;
        T s = 0;
        if (row < nr) {
            const int a = off(r0 + row), e = off(r0 + row + 1);
            for (int k = a + lane; k < e; k += lpr) s += sh[k];
        }
        for (int o = lpr >> 1; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (row < nr && lane == 0) out(r0 + row, s);
        return;
    }
    for (int rb = r0; rb < r1; rb += 4 * WG) {  // rows rb + q * WG + tid; the row pointers loaded up front
        int a[4], e[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int r = min(rb + q * WG + tid, r1 - 1);
            a[q] = off(r), e[q] = off(r + 1);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int r = rb + q * WG + tid;
            if (r < r1) {
                T s = 0;
                for (int j = a[q]; j < e[q]; ++j) s += sh[j];
                out(r, s);
            }
        }
    }
}

// The same with the block's rows located by CSR row pointers (row r's products at sh[ptr[r] - k0 ...]).
template <class T, int WG = kWG, class Out>
__device__ __forceinline__ void block_row_sums(const int* __restrict__ ptr, int r0, int r1, int k0, const T* sh,
                                               Out&& out) {
    block_row_sums_at<T, WG>([&](int r) { return ptr[r] - k0; }, r0, r1, sh, out);
}

// A staged block from start to finish: stage() fills sh, then the rows are summed and finished exactly as
// block_row_sums_at sums them (same rows per thread, same order). When the block has at most WG rows (one lane
// group per row), each group's row offsets and epilogue operands are loaded before stage(), so their latency
// overlaps the gathers; larger blocks load them after (four rows per thread ahead would cost the registers of
// four operand sets: 50 -> 84 VGPRs for the 8,192-nonzero blocks, and 4% on config 2, whose blocks are larger).
template <class T, int WG, class Off, class Epi, class Stage>
__device__ __forceinline__ void staged_rows(Off&& off, int r0, int r1, const T* sh, const Epi& epi, Stage&& stage) {
    const int tid = int(threadIdx.x), nr = r1 - r0;
    if (nr > WG) {
        stage();
        __syncthreads();
        block_row_sums_at<T, WG>(off, r0, r1, sh, [&](int r, T s) { epi(r, s); });
        return;
    }
    int lg = 6;
    while (lg > 0 && (nr << lg) > WG) --lg;
    const int row = tid >> lg, lane = tid & ((1 << lg) - 1), lpr = 1 << lg;
    int a = 0, e = 0;
    typename Epi::Pre pp{};
    if (row < nr) {
        a = off(r0 + row), e = off(r0 + row + 1);
        if (lane == 0) pp = epi.pre(r0 + row);
    }
    stage();
    __syncthreads();
    T s = 0;
    for (int k = a + lane; k < e; k += lpr) s += sh[k];
    for (int o = lpr >> 1; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (row < nr && lane == 0) epi.apply(r0 + row, s, pp);
}

// Row-block SpMV. Workgroups [0, nblk) take the row blocks; a row longer than kCap nonzeros forms a block of its
// own, which is skipped here: workgroups [nblk, nblk + nch) each take one kCap-nonzero chunk of such a row
// (chunk c: row, first and end nonzero in chunk[3c..3c+2]) and write its sum to lpart[c]; k_long_rows then adds
// each long row's chunk sums in chunk order and applies the epilogue (deterministic, no atomics).
template <class T, class Epi>
__global__ __launch_bounds__(kWG) void k_rowblocks(int nblk, const int* __restrict__ blk, const int* __restrict__ ptr,
                                                   const int* __restrict__ idx, const T* __restrict__ val,
                                                   const T* __restrict__ v, Epi epi, const int* __restrict__ chunk,
                                                   T* __restrict__ lpart) {
    __shared__ T sh[kCap];
    __shared__ T red[kWG / 64];
    const int tid = int(threadIdx.x);
    if (int(blockIdx.x) >= nblk) {  // a chunk of a long row
        const int c = int(blockIdx.x) - nblk;
        const int k0 = chunk[3 * c + 1], k1 = chunk[3 * c + 2];
        T s = 0;
        gather_rounds<kWG, 8>(k0, k1, idx, val, v, ColPlain{}, [&](int, int, T x) { s += x; });
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if ((tid & 63) == 0) red[tid >> 6] = s;
        __syncthreads();
        if (tid == 0) lpart[c] = (red[0] + red[1]) + (red[2] + red[3]);
        return;
    }
    const int b = xcd_block(int(blockIdx.x), nblk);
    const int r0 = blk[b], r1 = blk[b + 1];
    const int k0 = ptr[r0], k1 = ptr[r1];
    if (k1 - k0 > kCap) return;  // a long row: its chunks and k_long_rows handle it
    epi.load();
    // a block holds at most kWG rows, so each row gets a power-of-two lane group (the widest that fits)
    staged_rows<T, kWG>([&](int r) { return ptr[r] - k0; }, r0, r1, sh, epi, [&] {
        gather_rounds<kWG, kCap / kWG>(k0, k1, idx, val, v, ColPlain{}, [&](int k, int, T x) { sh[k - k0] = x; });
    });
}

// Long rows: row lrow[3i], chunks [lrow[3i+1], lrow[3i+1] + lrow[3i+2]) of lpart, summed in order.
template <class T, class Epi>
__global__ __launch_bounds__(kWG) void k_long_rows(int nlr, const int* __restrict__ lrow, const T* __restrict__ lpart,
                                                   Epi epi) {
    epi.load();
    for (int i = int(blockIdx.x * blockDim.x + threadIdx.x); i < nlr; i += int(gridDim.x * blockDim.x)) {
        const int c0 = lrow[3 * i + 1], nc = lrow[3 * i + 2];
        T s = lpart[c0];
        for (int c = 1; c < nc; ++c) s += lpart[c0 + c];
        epi(lrow[3 * i], s);
    }
}

// Pairwise sum of N values (N a power of two): ((r0 + r1) + (r2 + r3)) + ...
template <int N, class T>
__device__ __forceinline__ T tree_sum(const T* r) {
    if constexpr (N == 1)
        return r[0];
    else
        return tree_sum<N / 2>(r) + tree_sum<N / 2>(r + N / 2);
}

// Threads per workgroup of the kernels that stage SB nonzeros in LDS (sorted, sorted slab, batched blocks): one
// per 16 staged nonzeros, between kWG and 1,024. Measured at config 1 (FP32 slabs, 16,384-nonzero blocks): A x
// 209 -> 174 us from 256 to 1,024 threads; two such workgroups fill a CU's 160 KB of LDS and its 32 wave slots, and
// the row-sum phase after the gathers is split over 4x the threads. Equal SB gives equal WG in every kernel, so
// equal blocks are summed in the same order everywhere.
constexpr int stage_wg(int sb) { return sb / 16 < kWG ? kWG : (sb / 16 > 1024 ? 1024 : sb / 16); }

// Sorted layout. Inside each row block (at most SB nonzeros, rows of at most kCap) the nonzeros are stored in
// ascending column order, each with its row-order position in the block packed above the column:
// pk = (pos << cbits) | col. A wave's 64 lanes then gather from nearby columns, so several lanes share a cache
// line and the L2 sees fewer requests than nonzeros; each product goes to LDS at its row-order position and the
// rows are summed there (block_row_sums). Rows longer than kCap are chunked as in k_rowblocks.
template <class T, int SB, int WG, class Epi>
__global__ __launch_bounds__(WG) void k_sorted(int nblk, const int* __restrict__ blk, const int* __restrict__ ptr,
                                               const uint32_t* __restrict__ pk, const T* __restrict__ val,
                                               const T* __restrict__ v, int cbits, Epi epi,
                                               const int* __restrict__ chunk, T* __restrict__ lpart) {
    __shared__ T sh[SB];
    __shared__ T red[WG / 64];
    const int tid = int(threadIdx.x);
    const uint32_t mask = (1u << cbits) - 1u;
    if (int(blockIdx.x) >= nblk) {  // a chunk of a long row
        const int c = int(blockIdx.x) - nblk;
        const int k0 = chunk[3 * c + 1], k1 = chunk[3 * c + 2];
        T s = 0;
        gather_rounds<WG, kCap / WG>(k0, k1, pk, val, v, ColPacked{mask}, [&](int, uint32_t, T x) { s += x; });
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if ((tid & 63) == 0) red[tid >> 6] = s;
        __syncthreads();
        if (tid == 0) lpart[c] = tree_sum<WG / 64>(red);
        return;
    }
    const int b = xcd_block(int(blockIdx.x), nblk);
    const int r0 = blk[b], r1 = blk[b + 1];
    const int k0 = ptr[r0], k1 = ptr[r1];
    if (k1 - k0 > kCap && r1 - r0 == 1) return;  // a long row: its chunks and k_long_rows handle it
    epi.load();
    staged_rows<T, WG>([&](int r) { return ptr[r] - k0; }, r0, r1, sh, epi, [&] {
        gather_rounds<WG, 8>(k0, k1, pk, val, v, ColPacked{mask}, [&](int, uint32_t p, T x) { sh[p >> cbits] = x; });
    });
}

// Sorted slab layout: workgroup b takes row block b / nslab of slab b % nslab (columns local to the slab,
// gathered from v + sstart[slab]) and writes the slab's partial row sums part[slab * rows + row]; same staging as
// k_sorted, with WG threads. A block's nonzeros are [bk[b], bk[b + 1]) and its rows are located by their 16-bit
// offsets inside the block (roff), half the bytes of per-slab int32 row pointers.
template <class T, int SB, int WG>
__global__ __launch_bounds__(WG) void k_sorted_slab(int rows, int nslab, const int* __restrict__ snblk,
                                                    const int* __restrict__ sboff, const int* __restrict__ blk_all,
                                                    const int* __restrict__ bk_all,
                                                    const uint16_t* __restrict__ roff_all,
                                                    const uint32_t* __restrict__ pk,
                                                    const T* __restrict__ val, const T* __restrict__ v,
                                                    const int* __restrict__ sstart, int cbits, T* __restrict__ part) {
    __shared__ T sh[SB];
    __shared__ T red[WG / 64];
    const int q = int(blockIdx.x) % nslab, lb = int(blockIdx.x) / nslab;
    if (lb >= snblk[q]) return;
    const int* blk = blk_all + sboff[q];
    const int* bk = bk_all + sboff[q];
    const uint16_t* roff = roff_all + size_t(q) * size_t(rows);
    const T* vq = v + sstart[q];
    T* y = part + size_t(q) * size_t(rows);
    const int tid = int(threadIdx.x);
    const uint32_t mask = (1u << cbits) - 1u;
    const int r0 = blk[lb], r1 = blk[lb + 1];
    const int k0 = bk[lb], k1 = bk[lb + 1];
    if (k1 - k0 > kCap && r1 - r0 == 1) {  // a long row of this slab: strided sums and a workgroup reduction
        T s = 0;
        gather_rounds<WG, 8>(k0, k1, pk, val, vq, ColPacked{mask}, [&](int, uint32_t, T x) { s += x; });
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if ((tid & 63) == 0) red[tid >> 6] = s;
        __syncthreads();
        if (tid == 0) y[r0] = tree_sum<WG / 64>(red);
        return;
    }
    gather_rounds<WG, 8>(k0, k1, pk, val, vq, ColPacked{mask}, [&](int, uint32_t p, T x) { sh[p >> cbits] = x; });
    __syncthreads();
    const int len = k1 - k0;
    block_row_sums_at<T, WG>([&](int r) { return r < r1 ? int(roff[r]) : len; }, r0, r1, sh,
                             [&](int r, T s) { y[r] = s; });
}

// Slab layout, first kernel: workgroup b handles row block b / nslab of slab b % nslab (on XCD b % 8) and writes
// the slab's partial row sums part[slab * rows + row]; same staging and reduction as k_rowblocks.
template <class T>
__global__ __launch_bounds__(kWG) void k_slab_partial(int rows, int nslab, const int* __restrict__ snblk,
                                                      const int* __restrict__ sboff, const int* __restrict__ blk_all,
                                                      const int* __restrict__ ptr_all, const int* __restrict__ idx,
                                                      const T* __restrict__ val, const T* __restrict__ v,
                                                      T* __restrict__ part) {
    __shared__ T sh[kCap];
    __shared__ T red[kWG / 64];
    const int q = int(blockIdx.x) % nslab, lb = int(blockIdx.x) / nslab;
    if (lb >= snblk[q]) return;
    const int* blk = blk_all + sboff[q];
    const int* ptr = ptr_all + size_t(q) * (size_t(rows) + 1);
    T* y = part + size_t(q) * size_t(rows);
    const int tid = int(threadIdx.x);
    const int r0 = blk[lb], r1 = blk[lb + 1];
    const int k0 = ptr[r0], k1 = ptr[r1];
    if (k1 - k0 > kCap) {
        T s = 0;
        gather_rounds<kWG, 8>(k0, k1, idx, val, v, ColPlain{}, [&](int, int, T x) { s += x; });
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if ((tid & 63) == 0) red[tid >> 6] = s;
        __syncthreads();
        if (tid == 0) y[r0] = (red[0] + red[1]) + (red[2] + red[3]);
        return;
    }
    gather_rounds<kWG, kCap / kWG>(k0, k1, idx, val, v, ColPlain{}, [&](int k, int, T x) { sh[k - k0] = x; });
    __syncthreads();
    const int nr = r1 - r0;
    int lg = 6;
    while (lg > 0 && (nr << lg) > kWG) --lg;
    const int row = tid >> lg, lane = tid & ((1 << lg) - 1), lpr = 1 << lg;
    T s = 0;
    if (row < nr) {
        const int a = ptr[r0 + row] - k0, e = ptr[r0 + row + 1] - k0;
        for (int k = a + lane; k < e; k += lpr) s += sh[k];
    }
    for (int o = lpr >> 1; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (row < nr && lane == 0) y[r0 + row] = s;
}

// Slab layout, second kernel: the nslab partial sums of each row in slab order, then the epilogue.
template <class T, class Epi>
__global__ __launch_bounds__(kWG) void k_combine_epi(int rows, int nslab, const T* __restrict__ part, Epi epi) {
    epi.load();
    for (int i = int(blockIdx.x * blockDim.x + threadIdx.x); i < rows; i += int(gridDim.x * blockDim.x)) {
        T s = part[i];
        for (int q = 1; q < nslab; ++q) s += part[size_t(q) * size_t(rows) + size_t(i)];
        epi(i, s);
    }
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    return v;
}

// Workgroup sum of K doubles per thread into part[slot * K + k], slot = blockIdx unless given (lane order fixed:
// deterministic).
template <int K>
__device__ __forceinline__ void block_store(double (&v)[K], double* __restrict__ part, int slot = -1) {
    __shared__ double sh[K][kWG / 64];
    const int lane = int(threadIdx.x) & 63, w = int(threadIdx.x) >> 6;
    for (int k = 0; k < K; ++k) {
        const double s = wave_sum(v[k]);
        if (lane == 0) sh[k][w] = s;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        const int k = int(threadIdx.x);
        double s = 0;
        for (int q = 0; q < kWG / 64; ++q) s += sh[k][q];
        part[size_t(slot < 0 ? int(blockIdx.x) : slot) * K + k] = s;
    }
}

__device__ __forceinline__ double row_rel_weight(double lo, double up, double dr) {
    double b = 0;
    if (lo > -INFINITY) b = lo < 0 ? -lo : lo;
    if (up < INFINITY) {
        const double a = up < 0 ? -up : up;
        b = a > b ? a : b;
    }
    return 1.0 / (1.0 + b / dr);
}

template <class T>
__global__ __launch_bounds__(kWG) void k_sumsq(int n, const T* __restrict__ x, double* __restrict__ part) {
    double v[1] = {0};
    for (int j = int(blockIdx.x * blockDim.x + threadIdx.x); j < n; j += int(gridDim.x * blockDim.x))
        v[0] += double(x[j]) * double(x[j]);
    block_store<1>(v, part);
}

// out[k] = sum over b < blocks of part[(b0 + b) * K + k] for k < K: each thread sums a fixed strided subset of
// the partials, then a fixed-order tree in LDS, so the result is the same on every run.
[[maybe_unused]] __global__ __launch_bounds__(kWG) void k_finish(int b0, int blocks, int K, const double* __restrict__ part,
                                                double* __restrict__ out) {
    __shared__ double sh[kWG];
    const int t = int(threadIdx.x);
    for (int k = 0; k < K; ++k) {
        double s = 0;
        for (int q = t; q < blocks; q += kWG) s += part[size_t(b0 + q) * size_t(K) + size_t(k)];
        sh[t] = s;
        __syncthreads();
        for (int o = kWG / 2; o > 0; o >>= 1) {
            if (t < o) sh[t] += sh[t + o];
            __syncthreads();
        }
        if (t == 0) out[k] = sh[0];
        __syncthreads();
    }
}

template <class T>
__global__ __launch_bounds__(kWG) void k_scale(int n, T* __restrict__ x, double f) {
    for (int j = int(blockIdx.x * blockDim.x + threadIdx.x); j < n; j += int(gridDim.x * blockDim.x))
        x[j] = T(x[j] * f);
}

// x *= 1 / sqrt(s[0]) when s[0] > 0 (a power-iteration step normalised on the device: s is the step's sum of
// squares, still on the device).
template <class T>
__global__ __launch_bounds__(kWG) void k_scale_by_sum(int n, T* __restrict__ x, const double* __restrict__ s) {
    const double sv = *s;
    if (!(sv > 0)) return;
    const double f = 1 / sqrt(sv);
    for (int j = int(blockIdx.x * blockDim.x + threadIdx.x); j < n; j += int(gridDim.x * blockDim.x))
        x[j] = T(x[j] * f);
}

// Power-iteration steps run between two reads of their sums of squares by the host.
constexpr int kPowerGroup = 14;

template <class T>
__global__ __launch_bounds__(kWG) void k_to_f64(int n, double f, const T* __restrict__ src, double* __restrict__ dst) {
    for (int j = int(blockIdx.x * blockDim.x + threadIdx.x); j < n; j += int(gridDim.x * blockDim.x))
        dst[j] = double(src[j]) * f;
}

template <class T>
__global__ __launch_bounds__(kWG) void k_from_f64(int n, const double* __restrict__ src, T* __restrict__ dst) {
    for (int j = int(blockIdx.x * blockDim.x + threadIdx.x); j < n; j += int(gridDim.x * blockDim.x))
        dst[j] = T(src[j]);
}

// Row blocks: maximal runs of rows holding at most kCap nonzeros and at most kWG rows; a row longer than
// kCap forms a block of its own.
// Lanes the row-sum phase (block_row_sums_at) gives each row of a block of nr rows on wg threads: the widest
// power-of-two group, at most 64, with nr groups fitting the workgroup; one thread per row beyond wg rows.
constexpr int row_lanes(int nr, int wg) {
    int lg = 6;
    while (lg > 0 && (nr << lg) > wg) --lg;
    return nr > wg ? 1 : 1 << lg;
}

// Blocks are cut so that no lane of the row-sum phase adds more than kRowChain products in sequence. The lanes per
// row follow the block's row count, so a long row among many short or empty rows would otherwise be summed by one
// or two lanes: a 1,000-nonzero row cut into a block with 255 empty rows took one lane through 1,000 dependent LDS
// reads, and a product of 125 such rows followed by 875 empty ones 35.8 us instead of 3.6 (the transport LPs' A
// in column slabs, where each supply row fills one slab and is empty in the others, ran 3.8x slower than in one
// piece for this reason).
constexpr int kRowChain = 64;

// Whether a block of nr rows whose longest row has lmax nonzeros keeps every lane within kRowChain additions.
constexpr bool chain_fits(int nr, int64_t lmax, int wg) {
    return nr <= 1 || (lmax + row_lanes(nr, wg) - 1) / row_lanes(nr, wg) <= kRowChain;
}

std::vector<int32_t> row_blocks(int rows, const int32_t* ptr) {
    std::vector<int32_t> b;
    b.push_back(0);
    int r0 = 0;
    while (r0 < rows) {
        int r = r0;
        int64_t lmax = 0;
        while (r < rows && r - r0 < kWG && int64_t(ptr[r + 1]) - ptr[r0] <= kCap) {
            const int64_t lm = std::max<int64_t>(lmax, int64_t(ptr[r + 1]) - ptr[r]);
            if (!chain_fits(r - r0 + 1, lm, kWG)) break;
            lmax = lm, ++r;
        }
        if (r == r0) r = r0 + 1;  // long row
        b.push_back(r);
        r0 = r;
    }
    return b;
}

// Row blocks of the sorted layout: maximal runs of rows with at most sb nonzeros and at most 8 rows per thread of
// the block's workgroup (stage_wg(sb)), no row longer than long_above inside a run and every lane's row-sum chain
// within kRowChain; a row longer than long_above forms a block of its own (long_above <= sb). The row cap bounds the row-sum phase where many rows are empty (a
// block-structured matrix cut into column slabs leaves most rows of a slab empty; without it one workgroup once
// wrote 280,000 empty rows' sums and a product took 316 instead of 23 us).
std::vector<int32_t> sorted_blocks(int rows, const int32_t* ptr, int sb, int long_above = kCap) {
    const int wg = stage_wg(sb), max_rows = 8 * wg;
    std::vector<int32_t> b;
    b.push_back(0);
    int r0 = 0;
    while (r0 < rows) {
        int r = r0;
        if (int64_t(ptr[r0 + 1]) - ptr[r0] > long_above) {
            r = r0 + 1;
        } else {
            int64_t lmax = 0;
            while (r < rows && r - r0 < max_rows && int64_t(ptr[r + 1]) - ptr[r] <= long_above &&
                   int64_t(ptr[r + 1]) - ptr[r0] <= sb) {
                const int64_t lm = std::max<int64_t>(lmax, int64_t(ptr[r + 1]) - ptr[r]);
                if (!chain_fits(r - r0 + 1, lm, wg)) break;
                lmax = lm, ++r;
            }
            if (r == r0) r = r0 + 1;
        }
        b.push_back(r);
        r0 = r;
    }
    return b;
}

[[maybe_unused]] int grid_for(int k) { return std::max(1, std::min(2048, (k + kWG - 1) / kWG)); }
[[maybe_unused]] int red_blocks(int k) { return k > 0 ? std::max(1, std::min(kMaxRedBlocks, (k + kWG - 1) / kWG)) : 0; }

struct Mem {  // device allocations owned by one backend
    std::vector<void*> ptrs;
    template <class U>
    U* alloc(size_t count) {
        void* p = nullptr;
        HIP_OK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(U)));
        ptrs.push_back(p);
        return static_cast<U*>(p);
    }
    ~Mem() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

// One CSR on the device in one of two layouts.
//  plain: row blocks over the whole matrix (k_rowblocks).
//  slab:  for a gathered vector of at least kSlabMinCols entries (half an XCD's L2 in FP32) whose (row, slab)
//         pairs are mostly occupied (slabs_pay), the columns are split into nslab equal ranges (slab_count: 8 or
//         16) and the matrix stored slab-major as nslab CSR pieces, with absolute row pointers (row blocks) or
//         block starts and 16-bit in-block row offsets (sorted blocks). Workgroup b works on slab b % nslab, i.e.
//         on XCD b % 8, so each XCD gathers
//         from slices of the vector that stay in its own L2; each slab writes partial row sums and a combine
//         kernel adds them in slab order and applies the epilogue.
constexpr int kSlabs = kSlabsK;
constexpr int kSlabMinCols = 1 << 19;  // 2 MB of FP32 gather target and up

template <class T>
struct DevMat {
    int rows = 0, nblk = 0, maxblk = 0;
    bool slab = false;
    int *ptr = nullptr, *idx = nullptr, *blk = nullptr;  // unsorted slab: ptr holds nslab * (rows + 1) entries
    int *snblk = nullptr, *sboff = nullptr;              // slab: block count and block-list offset per slab
    T* val = nullptr;
    T* part = nullptr;  // slab: nslab * rows partial sums
    int nch = 0, nlr = 0;                                // plain: chunks of long rows, long rows
    int *chunk = nullptr, *lrow = nullptr;
    T* lpart = nullptr;                                  // plain: one partial sum per chunk
    bool sorted = false;                                 // idx holds packed (position, column) words
    int cbits = 0, sb = 0;
    int* sstart = nullptr;                               // sorted slab: first column of each slab
    int nslab = 0;                                       // slab: number of column slabs (2, 4, 8 or 16)
    int* bk = nullptr;           // sorted slab: first nonzero of each block (indexed like blk; ptr is not uploaded)
    uint16_t* roff = nullptr;    // sorted slab: nslab * rows offsets of each row's first nonzero inside its block
};

struct HostLayout {
    bool slab = false;
    int nblk = 0, maxblk = 0;
    std::vector<int32_t> ptr, idx, blk, snblk, sboff;
    std::vector<int32_t> chunk, lrow;  // plain: long-row chunks (row, k0, k1) and long rows (row, first chunk, count)
    std::vector<int64_t> perm;  // slab or sorted layout: position k holds the CSR entry perm[k]
    bool sorted = false;
    int cbits = 0, sb = 0;  // sorted: column bits of the packed words, nonzeros per block (power of two)
    std::vector<int32_t> sstart;
    int nslab = 0;
    // sorted slab (ptr left empty): first nonzero of each block, indexed like blk, and each row's offset inside its
    // block per slab (nslab * rows; a block holds at most 2^14 nonzeros, so 16 bits suffice)
    std::vector<int32_t> bk;
    std::vector<uint16_t> roff;
};

// Column slabs of the slab layout: the fewest of 8 and 16 that keep a slab's slice of the gathered vector within
// 2^17 entries. Measured at config 1 (FP32, sorted slab blocks with batched gathers): A x gathering 2M entries
// 155 us with 16 slabs, 166 with 8, 207 with 4; A^T z gathering 1M entries 154 us with 8, 186 with 16, 166 with
// 4 (narrower slices pack a block's gathers into fewer cache lines; more slabs cost partial sums and row
// pointers). Workgroup b works on slab b mod nslab, i.e. on XCD b mod 8: with 16 slabs each XCD serves two
// slices, with fewer than 8 XCDs x and x + nslab share one. TPDHG_SLABS=2|4|8|16 forces a count.
int slab_count(int cols) {
    if (const char* e = std::getenv("TPDHG_SLABS")) {
        const int k = std::atoi(e);
        if (k == 2 || k == 4 || k == 8 || k == 16) return k;
    }
    return int64_t(cols) > int64_t(kSlabs) << 17 ? 2 * kSlabs : kSlabs;
}

// Largest sorted blocks. Plain layout (256 threads): 8,192 (32 KB of FP32 products in LDS; measured at config 1
// against 4,096 and 2,048), 13 position bits above 19 column bits. Slabs: 16,384 when a slab spans at most 2^18
// columns (14 position bits; 64 KB of FP32 products, 1,024 threads: 174 us against 209 us for 8,192 blocks at
// config 1), else 8,192.
constexpr int kSortPlain = 8192;
constexpr int kSortSlab = 16384;
// The block size shrinks (down to kCap) until there is at least one block per CU (256). Larger blocks amortise each
// workgroup's fixed work (block record, LDS staging, row sums): a 100k x 200k planted LP with 4M nonzeros runs
// 34.2 us per FP32 fast-path iteration on 8,192-nonzero blocks (488 per product) against 43.5 us on the 2,048-
// nonzero blocks a minimum of 1,024 blocks gave; the block-angular LP of 6.1M nonzeros 41.6 against 43.4 us.
constexpr int kSortMinBlocks = 256;

int sorted_block_size(int64_t nnz, int max_sb, int value_bytes) {
    int sb = value_bytes > 4 ? max_sb / 2 : max_sb;  // same LDS bytes for FP64 values
    if (const char* e = std::getenv("TPDHG_SORT_MAX_BLOCK")) {  // tuning: cap on the block size
        const int cap = std::atoi(e);
        while (sb > kCap && sb > cap) sb /= 2;
    }
    int64_t min_blocks = kSortMinBlocks;
    if (const char* e = std::getenv("TPDHG_SORT_MIN_BLOCKS")) min_blocks = std::atoll(e);  // tuning
    while (sb > kCap && nnz / sb < min_blocks) sb /= 2;
    return sb;
}

// Layout choice: TPDHG_LAYOUT=sorted or =unsorted forces one; TPDHG_LAYOUT=slab forces the sorted slab layout on
// every matrix of at least 2 * kSlabsK columns (whatever its size or slab occupancy: the production layout of the
// largest matrices, made reachable at sizes a test solves in seconds); otherwise ("auto") single matrices use the
// sorted layout from kSortMinNnz nonzeros up (below that the solve is launch-bound and the row-block kernels
// measured faster per iteration: 11.7 against 14.8 us on 10k-row LPs) and batches use it always.
enum class LayoutMode { unsorted, sorted, slab, automatic };
constexpr int64_t kSortMinNnz = int64_t(1) << 20;

LayoutMode layout_mode() {
    const char* e = std::getenv("TPDHG_LAYOUT");
    if (e && std::string(e) == "unsorted") return LayoutMode::unsorted;
    if (e && std::string(e) == "sorted") return LayoutMode::sorted;
    if (e && std::string(e) == "slab") return LayoutMode::slab;
    return LayoutMode::automatic;
}

[[maybe_unused]] bool sorted_layout_enabled() { return layout_mode() != LayoutMode::unsorted; }

// Sorts the entries of every block of a row range by column, writing packed words and the permutation (entries
// [base + k0, base + k1) of idx_in map to the same range of pk / perm; positions are relative to the block).
void sort_blocks(const std::vector<int32_t>& blk, const int32_t* ptr, const int32_t* cols_local, int64_t perm_base_in,
                 const int64_t* perm_in, int cbits, std::vector<int32_t>& pk, std::vector<int64_t>& perm,
                 int long_above = kCap) {
    // blocks are sorted independently (each writes only its own nonzero range), on the host's cores
    parallel_for(int64_t(blk.size()) - 1, 8, [&](int64_t b) {
        const int32_t k0 = ptr[blk[size_t(b)]], k1 = ptr[blk[size_t(b) + 1]];
        const int32_t len = k1 - k0;
        std::vector<int32_t> order(static_cast<size_t>(len));
        for (int32_t i = 0; i < len; ++i) order[size_t(i)] = i;
        const bool single_long = blk[size_t(b) + 1] - blk[size_t(b)] == 1 && len > long_above;
        if (!single_long)
            std::stable_sort(order.begin(), order.end(),
                             [&](int32_t a, int32_t c) { return cols_local[k0 + a] < cols_local[k0 + c]; });
        for (int32_t i = 0; i < len; ++i) {
            const int32_t src = k0 + order[size_t(i)];
            const uint32_t pos = single_long ? 0u : uint32_t(order[size_t(i)]);
            pk[size_t(k0 + i)] = int32_t((pos << cbits) | uint32_t(cols_local[src]));
            perm[size_t(k0 + i)] = perm_in ? perm_in[src] : perm_base_in + src;
        }
    });
}

// Whether column slabs pay: at least half of the (nonempty row, slab) pairs hold a nonzero. A random sparse matrix
// fills most pairs (config 1: 92% with 2.5 nonzeros per row and slab); a block-structured one leaves most of a
// slab's rows empty, and the slab layout then mostly stores, writes and adds empty partial sums (config 2's
// matrices in 8 slabs: 76 and 37 us per product against 23 and 17 in one piece).
bool slabs_pay(int rows, int cols, const int32_t* ptr, const int32_t* idx, int nslab) {
    std::vector<uint8_t> hits(static_cast<size_t>(rows));
    parallel_for(rows, 8192, [&](int64_t i) {
        uint32_t hit = 0;
        for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) hit |= 1u << int(int64_t(idx[k]) * nslab / cols);
        hits[size_t(i)] = uint8_t(__builtin_popcount(hit));
    });
    int64_t pairs = 0, nonempty = 0;
    for (int i = 0; i < rows; ++i) pairs += hits[size_t(i)], nonempty += ptr[i + 1] > ptr[i];
    return 2 * pairs >= nonempty * nslab;
}

// one_piece: no column slabs whatever the rule says (the alternative choose_layout times against slabs).
[[maybe_unused]] HostLayout make_layout(int rows, int cols, const int32_t* ptr, const int32_t* idx,
                                         int value_bytes = 4, bool one_piece = false) {
    HostLayout L;
    const int64_t nnz = ptr[rows];
    if (nnz > kMaxNnz)
        throw std::length_error("a matrix of " + std::to_string(nnz) + " nonzeros exceeds the " +
                                std::to_string(kMaxNnz) + " the kernels index with 32-bit integers");
    const LayoutMode mode = layout_mode();
    if (one_piece)
        L.slab = false;
    else if (mode == LayoutMode::slab)
        L.slab = rows > 0 && cols >= 2 * kSlabs;
    else
        L.slab = cols >= kSlabMinCols && rows > 0 && slabs_pay(rows, cols, ptr, idx, slab_count(cols));
    const bool want_sorted = mode == LayoutMode::sorted || mode == LayoutMode::slab ||
                             (mode == LayoutMode::automatic && nnz >= kSortMinNnz);
    if (!L.slab) {
        L.ptr.assign(ptr, ptr + rows + 1);
        if (want_sorted && cols <= (1 << 19)) {
            L.sorted = true, L.cbits = 19, L.sb = sorted_block_size(nnz, kSortPlain, value_bytes);
            L.blk = sorted_blocks(rows, ptr, L.sb);
            L.idx.resize(size_t(nnz)), L.perm.resize(size_t(nnz));
            sort_blocks(L.blk, ptr, idx, 0, nullptr, L.cbits, L.idx, L.perm);
        } else {
            L.idx.assign(idx, idx + nnz);
            L.blk = row_blocks(rows, ptr);
        }
        L.nblk = int(L.blk.size()) - 1;
        for (int i = 0; i < rows; ++i) {
            if (ptr[i + 1] - ptr[i] <= kCap) continue;
            L.lrow.push_back(i), L.lrow.push_back(int32_t(L.chunk.size() / 3));
            int nc = 0;
            for (int32_t k = ptr[i]; k < ptr[i + 1]; k += kCap, ++nc)
                L.chunk.push_back(i), L.chunk.push_back(k), L.chunk.push_back(std::min(k + kCap, ptr[i + 1]));
            L.lrow.push_back(nc);
        }
        return L;
    }
    const int NS = slab_count(cols);
    L.nslab = NS;
    auto slab_of = [&](int32_t c) { return int(int64_t(c) * NS / cols); };
    // per slab: row lengths, then their prefix; row i's nonzeros of slab q go, in row order, to
    // soff[q] + sp[q][i] ... (rows and slabs are independent, so these passes run on the host's cores)
    std::vector<std::vector<int32_t>> sp(NS, std::vector<int32_t>(size_t(rows) + 1, 0));
    parallel_for(rows, 8192, [&](int64_t i) {
        for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) sp[size_t(slab_of(idx[k]))][size_t(i) + 1]++;
    });
    parallel_for(NS, 1, [&](int64_t q) {
        for (int i = 0; i < rows; ++i) sp[size_t(q)][size_t(i) + 1] += sp[size_t(q)][size_t(i)];
    });
    std::vector<int64_t> soff(NS + 1, 0);
    for (int q = 0; q < NS; ++q) soff[size_t(q) + 1] = soff[size_t(q)] + sp[size_t(q)][size_t(rows)];
    L.idx.resize(size_t(nnz)), L.perm.resize(size_t(nnz));
    parallel_for(rows, 8192, [&](int64_t i) {
        int64_t at[2 * kSlabs];
        for (int q = 0; q < NS; ++q) at[q] = soff[size_t(q)] + sp[size_t(q)][size_t(i)];
        for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) {
            const int64_t p = at[slab_of(idx[k])]++;
            L.idx[size_t(p)] = idx[k], L.perm[size_t(p)] = k;
        }
    });
    L.ptr.resize(size_t(NS) * (size_t(rows) + 1));
    L.snblk.resize(NS), L.sboff.resize(NS);
    // slab q holds the columns c with c * NS / cols == q, i.e. [ceil(q * cols / NS), ...)
    L.sstart.resize(NS + 1);
    for (int q = 0; q <= NS; ++q) L.sstart[size_t(q)] = int32_t((int64_t(q) * cols + NS - 1) / NS);
    int64_t width = 0;
    for (int q = 0; q < NS; ++q) width = std::max<int64_t>(width, L.sstart[size_t(q) + 1] - L.sstart[size_t(q)]);
    L.sorted = want_sorted && width <= (1 << 19);
    if (L.sorted)
        L.cbits = width <= (1 << 18) ? 18 : 19,
        L.sb = sorted_block_size(nnz, L.cbits == 18 ? kSortSlab : kSortSlab / 2, value_bytes);
    std::vector<int32_t> local, pk;
    std::vector<int64_t> perm2;
    if (L.sorted) {
        local.resize(size_t(nnz)), pk.resize(size_t(nnz)), perm2.resize(size_t(nnz));
        parallel_for(nnz, 1 << 16, [&](int64_t k) {
            local[size_t(k)] = L.idx[size_t(k)] - L.sstart[size_t(slab_of(L.idx[size_t(k)]))];
        });
    }
    parallel_for(NS, 1, [&](int64_t q) {
        for (int i = 0; i <= rows; ++i)
            L.ptr[size_t(q) * (size_t(rows) + 1) + size_t(i)] = int32_t(soff[size_t(q)] + sp[size_t(q)][size_t(i)]);
    });
    std::vector<std::vector<int32_t>> bq(static_cast<size_t>(NS));  // each slab's blocks, cut and sorted on its own
    parallel_for(NS, 1, [&](int64_t q) {
        const int32_t* qptr = L.ptr.data() + size_t(q) * (size_t(rows) + 1);
        bq[size_t(q)] = L.sorted ? sorted_blocks(rows, qptr, L.sb) : row_blocks(rows, sp[size_t(q)].data());
        if (L.sorted) sort_blocks(bq[size_t(q)], qptr, local.data(), 0, L.perm.data(), L.cbits, pk, perm2);
    });
    for (int q = 0; q < NS; ++q) {
        const std::vector<int32_t>& b = bq[size_t(q)];
        L.snblk[size_t(q)] = int(b.size()) - 1;
        L.sboff[size_t(q)] = int(L.blk.size());
        L.maxblk = std::max(L.maxblk, L.snblk[size_t(q)]);
        L.blk.insert(L.blk.end(), b.begin(), b.end());
    }
    if (L.sorted) {
        L.idx.swap(pk), L.perm.swap(perm2);
        // the kernels locate blocks and rows by bk and roff; the per-slab row pointers stay on the host
        L.bk.resize(L.blk.size()), L.roff.assign(size_t(NS) * size_t(rows), 0);
        parallel_for(NS, 1, [&](int64_t q) {
            const int32_t* qptr = L.ptr.data() + size_t(q) * (size_t(rows) + 1);
            const std::vector<int32_t>& b = bq[size_t(q)];
            const size_t base = size_t(L.sboff[size_t(q)]);
            for (size_t lb = 0; lb < b.size(); ++lb) L.bk[base + lb] = qptr[b[lb]];
            for (size_t lb = 0; lb + 1 < b.size(); ++lb)
                for (int32_t r = b[lb]; r < b[lb + 1]; ++r) {  // a one-row block (a long row) gives offset 0
                    const int32_t o = qptr[r] - qptr[b[lb]];
                    if (o > 65535) throw std::logic_error("sorted slab block beyond 16-bit row offsets");
                    L.roff[size_t(q) * size_t(rows) + size_t(r)] = uint16_t(o);
                }
        });
        L.ptr.clear(), L.ptr.shrink_to_fit();
    }
    return L;
}

template <class T>
T* upload_values(Mem& mem, hipStream_t st, const HostLayout& L, const double* val, int64_t nnz) {
    T* d = mem.alloc<T>(size_t(std::max<int64_t>(nnz, 1)));
    if (nnz) {
        std::vector<T> h(static_cast<size_t>(nnz));
        if (!L.perm.empty())
            parallel_for(nnz, 1 << 16, [&](int64_t k) { h[size_t(k)] = T(val[L.perm[size_t(k)]]); });
        else
            parallel_for(nnz, 1 << 16, [&](int64_t k) { h[size_t(k)] = T(val[k]); });
        HIP_OK(hipMemcpyAsync(d, h.data(), size_t(nnz) * sizeof(T), hipMemcpyHostToDevice, st));
        HIP_OK(hipStreamSynchronize(st));
    }
    return d;
}

template <class T>
DevMat<T> upload_layout(Mem& mem, hipStream_t st, int rows, const HostLayout& L, const double* val) {
    DevMat<T> d;
    d.rows = rows, d.slab = L.slab, d.nblk = L.nblk, d.maxblk = L.maxblk;
    auto up = [&](const std::vector<int32_t>& v) {
        int* p = mem.alloc<int>(std::max<size_t>(v.size(), 1));
        if (!v.empty()) HIP_OK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice, st));
        return p;
    };
    d.ptr = up(L.ptr), d.idx = up(L.idx), d.blk = up(L.blk);
    d.sorted = L.sorted, d.cbits = L.cbits, d.sb = L.sb, d.nslab = L.nslab;
    if (L.sorted && L.slab) {
        d.sstart = up(L.sstart), d.bk = up(L.bk);
        d.roff = mem.alloc<uint16_t>(std::max<size_t>(L.roff.size(), 1));
        if (!L.roff.empty())
            HIP_OK(hipMemcpyAsync(d.roff, L.roff.data(), L.roff.size() * sizeof(uint16_t), hipMemcpyHostToDevice, st));
    }
    d.nch = int(L.chunk.size() / 3), d.nlr = int(L.lrow.size() / 3);
    if (d.nch) d.chunk = up(L.chunk), d.lrow = up(L.lrow), d.lpart = mem.alloc<T>(size_t(d.nch));
    if (L.slab) {
        d.snblk = up(L.snblk), d.sboff = up(L.sboff);
        d.part = mem.alloc<T>(size_t(L.nslab) * size_t(std::max(rows, 1)));
    }
    d.val = upload_values<T>(mem, st, L, val, int64_t(L.idx.size()));
    HIP_OK(hipStreamSynchronize(st));
    return d;
}

// The same structure with values of another type (U), sharing pointers, indices and blocks.
template <class U, class T>
DevMat<U> share_structure(Mem& mem, hipStream_t st, const DevMat<T>& a, const HostLayout& L, const double* val) {
    DevMat<U> d;
    d.rows = a.rows, d.nblk = a.nblk, d.maxblk = a.maxblk, d.slab = a.slab;
    d.ptr = a.ptr, d.idx = a.idx, d.blk = a.blk, d.snblk = a.snblk, d.sboff = a.sboff;
    d.nch = a.nch, d.nlr = a.nlr, d.chunk = a.chunk, d.lrow = a.lrow;
    d.sorted = a.sorted, d.cbits = a.cbits, d.sb = a.sb, d.sstart = a.sstart, d.nslab = a.nslab;
    d.bk = a.bk, d.roff = a.roff;
    if (a.nch) d.lpart = mem.alloc<U>(size_t(a.nch));
    d.val = upload_values<U>(mem, st, L, val, int64_t(L.idx.size()));
    if (a.slab) d.part = mem.alloc<U>(size_t(a.nslab) * size_t(std::max(a.rows, 1)));
    return d;
}


template <class T, int SB, class Epi>
void launch_sorted(const DevMat<T>& a, const T* v, const Epi& epi, hipStream_t st) {
    if (a.slab) {
        constexpr int WG = stage_wg(SB);
        if (a.maxblk > 0)
            k_sorted_slab<T, SB, WG><<<a.nslab * a.maxblk, WG, 0, st>>>(
                a.rows, a.nslab, a.snblk, a.sboff, a.blk, a.bk, a.roff, reinterpret_cast<const uint32_t*>(a.idx),
                a.val, v, a.sstart, a.cbits, a.part);
        k_combine_epi<T, Epi><<<grid_for(a.rows), kWG, 0, st>>>(a.rows, a.nslab, a.part, epi);
        return;
    }
    if constexpr (SB > kSortPlain) {
        throw std::logic_error("sorted plain layout with blocks above 8,192 nonzeros");
    } else {
        if (a.nblk == 0) return;
        constexpr int WG = stage_wg(SB);
        k_sorted<T, SB, WG, Epi><<<a.nblk + a.nch, WG, 0, st>>>(a.nblk, a.blk, a.ptr,
                                                               reinterpret_cast<const uint32_t*>(a.idx), a.val, v,
                                                               a.cbits, epi, a.chunk, a.lpart);
        if (a.nlr) k_long_rows<T, Epi><<<(a.nlr + kWG - 1) / kWG, kWG, 0, st>>>(a.nlr, a.lrow, a.lpart, epi);
    }
}

template <class T, class Epi>
void launch_rows(const DevMat<T>& a, const T* v, const Epi& epi, hipStream_t st) {
    if (a.rows == 0) return;
    if (a.sorted) {
        switch (a.sb) {
            case 2048: launch_sorted<T, 2048>(a, v, epi, st); return;
            case 4096: launch_sorted<T, 4096>(a, v, epi, st); return;
            case 8192: launch_sorted<T, 8192>(a, v, epi, st); return;
            case 16384: launch_sorted<T, 16384>(a, v, epi, st); return;
            default: throw std::logic_error("unsupported sorted block size");
        }
    }
    if (a.slab) {
        if (a.maxblk > 0)
            k_slab_partial<T><<<a.nslab * a.maxblk, kWG, 0, st>>>(a.rows, a.nslab, a.snblk, a.sboff, a.blk, a.ptr,
                                                                 a.idx, a.val, v, a.part);
        k_combine_epi<T, Epi><<<grid_for(a.rows), kWG, 0, st>>>(a.rows, a.nslab, a.part, epi);
        return;
    }
    if (a.nblk == 0) return;
    k_rowblocks<T, Epi><<<a.nblk + a.nch, kWG, 0, st>>>(a.nblk, a.blk, a.ptr, a.idx, a.val, v, epi, a.chunk, a.lpart);
    if (a.nlr) k_long_rows<T, Epi><<<(a.nlr + kWG - 1) / kWG, kWG, 0, st>>>(a.nlr, a.lrow, a.lpart, epi);
}

// Median time (ms) of `reps` products y = M v with layout L and values in T, on scratch buffers freed on return
// (v is zero: the time depends on the indices, not on the values).
template <class T>
double time_product(hipStream_t st, int rows, int cols, const HostLayout& L, const double* val, int reps = 7) {
    Mem tmp;
    const DevMat<T> d = upload_layout<T>(tmp, st, rows, L, val);
    T* v = tmp.alloc<T>(size_t(std::max(cols, 1)));
    T* y = tmp.alloc<T>(size_t(std::max(rows, 1)));
    HIP_OK(hipMemsetAsync(v, 0, size_t(std::max(cols, 1)) * sizeof(T), st));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    std::vector<float> ms;
    try {
        for (int r = 0; r < reps + 2; ++r) {  // two warm-up products
            HIP_OK(hipEventRecord(e0, st));
            launch_rows(d, v, EpiStore<T>{y}, st);
            HIP_OK(hipEventRecord(e1, st));
            HIP_OK(hipEventSynchronize(e1));
            float t = 0;
            HIP_OK(hipEventElapsedTime(&t, e0, e1));
            if (r >= 2) ms.push_back(t);
        }
        HIP_OK(hipGetLastError());
    } catch (...) {
        (void)hipEventDestroy(e0), (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0), (void)hipEventDestroy(e1);
    std::sort(ms.begin(), ms.end());
    return ms[ms.size() / 2];
}

// Layout of one matrix for the single-LP backends. make_layout's occupancy rule decides between column slabs and
// one piece; when it picks slabs (automatic mode), the one-piece layout is timed against them (median of 7 products
// each, the solver's own kernels in T) and taken when it is at least kOnePieceGain times faster, a margin that keeps
// the choice, and so the summation order, the same from run to run. A guard against the rule's misjudgements: it
// first caught the transport LP's A (2,000 rows of 1,000 nonzeros over 1M columns: 52.3 us per product in 8 slabs
// against 13.6 in one piece), whose cause, unbounded row-sum chains, kRowChain now removes (11.8 us in slabs, which
// the timing keeps); config 1's matrices keep their slabs.
constexpr double kOnePieceGain = 1.25;

template <class T>
HostLayout choose_layout(hipStream_t st, int rows, int cols, const int32_t* ptr, const int32_t* idx,
                         const double* val) {
    HostLayout L = make_layout(rows, cols, ptr, idx, int(sizeof(T)));
    if (!L.slab || layout_mode() != LayoutMode::automatic) return L;
    HostLayout one = make_layout(rows, cols, ptr, idx, int(sizeof(T)), true);
    const double t_slab = time_product<T>(st, rows, cols, L, val), t_one = time_product<T>(st, rows, cols, one, val);
    return t_one * kOnePieceGain <= t_slab ? std::move(one) : std::move(L);
}

// A CSR uploaded in the layout choose_layout picks.
template <class T>
DevMat<T> upload_csr(Mem& mem, hipStream_t st, int rows, int cols, const int32_t* ptr, const int32_t* idx,
                     const double* val) {
    return upload_layout<T>(mem, st, rows, choose_layout<T>(st, rows, cols, ptr, idx, val), val);
}

[[maybe_unused]] std::string device_name() {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, dev));
    std::string nm = prop.name;
    if (nm.empty()) nm = prop.gcnArchName;
    return nm;
}

[[maybe_unused]] void require_device() {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        throw std::runtime_error("no HIP device visible: the PDHG solver runs on an AMD GPU only (there is no CPU path)");
}

}  // namespace
}  // namespace tpdhg
