// Shared HIP building blocks of the PDHG device backends (gfx950, wave64): error checking, the LDS-staged
// row-block SpMV kernel with pluggable epilogues, deterministic block reductions, device buffers and CSR upload.
// Included by each backend translation unit; everything has internal linkage.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace tpdhg {
namespace {

constexpr int kWG = 256;    // threads per workgroup: 4 waves of 64
constexpr int kCap = 2048;  // nonzeros staged per row block (8 per thread)
constexpr int kMaxRedBlocks = 1024;

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

template <class T>
struct EpiStore {  // y = M v
    T* y;
    __device__ void load() {}
    __device__ void operator()(int r, T s) const { y[r] = s; }
};

template <class T, class Epi>
__global__ __launch_bounds__(kWG) void k_rowblocks(int nblk, const int* __restrict__ blk, const int* __restrict__ ptr,
                                                   const int* __restrict__ idx, const T* __restrict__ val,
                                                   const T* __restrict__ v, Epi epi) {
    __shared__ T sh[kCap];
    __shared__ T red[kWG / 64];
    const int b = xcd_block(int(blockIdx.x), nblk);
    const int tid = int(threadIdx.x);
    const int r0 = blk[b], r1 = blk[b + 1];
    const int k0 = ptr[r0], k1 = ptr[r1];
    epi.load();
    if (k1 - k0 > kCap) {  // a single long row: strided partial sums and a workgroup reduction
        T s = 0;
        for (int k = k0 + tid; k < k1; k += kWG) s += val[k] * v[idx[k]];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if ((tid & 63) == 0) red[tid >> 6] = s;
        __syncthreads();
        if (tid == 0) epi(r0, (red[0] + red[1]) + (red[2] + red[3]));
        return;
    }
#pragma unroll
    for (int u = 0; u < kCap / kWG; ++u) {
        const int k = k0 + u * kWG + tid;
        if (k < k1) sh[k - k0] = val[k] * v[idx[k]];
    }
    __syncthreads();
    const int nr = r1 - r0;
    int lg = 6;  // lanes per row = 2^lg: the widest group with nr groups fitting the workgroup
    while (lg > 0 && (nr << lg) > kWG) --lg;
    const int row = tid >> lg, lane = tid & ((1 << lg) - 1), lpr = 1 << lg;
    T s = 0;
    if (row < nr) {
        const int a = ptr[r0 + row] - k0, e = ptr[r0 + row + 1] - k0;
        for (int k = a + lane; k < e; k += lpr) s += sh[k];
    }
    for (int o = lpr >> 1; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (row < nr && lane == 0) epi(r0 + row, s);
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    return v;
}

// Workgroup sum of K doubles per thread into part[blockIdx * K + k] (lane order fixed: deterministic).
template <int K>
__device__ __forceinline__ void block_store(double (&v)[K], double* __restrict__ part) {
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
        part[size_t(blockIdx.x) * K + k] = s;
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
__global__ __launch_bounds__(kWG) void k_finish(int b0, int blocks, int K, const double* __restrict__ part,
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
std::vector<int32_t> row_blocks(int rows, const int32_t* ptr) {
    std::vector<int32_t> b;
    b.push_back(0);
    int r0 = 0;
    while (r0 < rows) {
        int r = r0;
        while (r < rows && r - r0 < kWG && int64_t(ptr[r + 1]) - ptr[r0] <= kCap) ++r;
        if (r == r0) r = r0 + 1;  // long row
        b.push_back(r);
        r0 = r;
    }
    return b;
}

int grid_for(int k) { return std::max(1, std::min(2048, (k + kWG - 1) / kWG)); }
int red_blocks(int k) { return k > 0 ? std::max(1, std::min(kMaxRedBlocks, (k + kWG - 1) / kWG)) : 0; }

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

template <class T>
struct DevMat {
    int rows = 0, nblk = 0;
    int *ptr = nullptr, *idx = nullptr, *blk = nullptr;
    T* val = nullptr;
};

template <class T>
DevMat<T> upload_csr(Mem& mem, hipStream_t st, int rows, const int32_t* ptr, const int32_t* idx, const double* val) {
    DevMat<T> d;
    d.rows = rows;
    const int64_t nnz = ptr[rows];
    std::vector<int32_t> blk = row_blocks(rows, ptr);
    d.nblk = int(blk.size()) - 1;
    d.ptr = mem.alloc<int>(size_t(rows) + 1), d.idx = mem.alloc<int>(size_t(nnz)), d.blk = mem.alloc<int>(blk.size());
    d.val = mem.alloc<T>(size_t(nnz));
    HIP_OK(hipMemcpyAsync(d.ptr, ptr, (size_t(rows) + 1) * sizeof(int), hipMemcpyHostToDevice, st));
    if (nnz) HIP_OK(hipMemcpyAsync(d.idx, idx, size_t(nnz) * sizeof(int), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(d.blk, blk.data(), blk.size() * sizeof(int), hipMemcpyHostToDevice, st));
    if (nnz) {
        std::vector<T> h(val, val + nnz);
        HIP_OK(hipMemcpyAsync(d.val, h.data(), size_t(nnz) * sizeof(T), hipMemcpyHostToDevice, st));
        HIP_OK(hipStreamSynchronize(st));
    }
    HIP_OK(hipStreamSynchronize(st));
    return d;
}

template <class T, class Epi>
void launch_rows(const DevMat<T>& a, const T* v, const Epi& epi, hipStream_t st) {
    if (a.nblk == 0) return;
    k_rowblocks<T, Epi><<<a.nblk, kWG, 0, st>>>(a.nblk, a.blk, a.ptr, a.idx, a.val, v, epi);
}

std::string device_name() {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, dev));
    std::string nm = prop.name;
    if (nm.empty()) nm = prop.gcnArchName;
    return nm;
}

void require_device() {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        throw std::runtime_error("no HIP device visible: the PDHG solver runs on an AMD GPU only (there is no CPU path)");
}

}  // namespace
}  // namespace tpdhg
