// y = M v on caller-owned device arrays, allocating nothing: the entry point for callers that cannot allocate
// device memory (code under HIP graph capture, serving loops). No reference counterpart.
//
// The nonzeros are cut into equal segments of kSegNnz; workgroup b stages the products of segment b in LDS and
// sums every row that intersects it. A row lying wholly inside the segment is written to y directly; a row cut by
// a segment boundary leaves the segment's partial sum in `scratch`: its first row's in slot 2b when that row began
// in an earlier segment, else its last row's in slot 2b + 1 when that row continues into a later one. The second
// kernel (one thread per row) writes empty rows as 0 and adds each cut row's partials in segment order: slot
// 2 b0 + 1 of its first segment b0, then slot 2b of every later segment up to its last (deterministic, no atomics).
// Every slot the second kernel reads was written by the first kernel of the same call, so nothing needs clearing.
//
// `scratch` is a file-scope device array of 2 * kMaxSegs doubles (1 MiB per device, held until the process ends);
// host code reaches it only through the address hipGetSymbolAddress returns for the current device. Calls that
// overlap on one device share it and race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "hipcommon.hpp"
#include "noalloc.hpp"

namespace tpdhg {
namespace {

constexpr int kSegNnz = 2048;       // nonzeros per segment (one workgroup, 8 per thread)
constexpr int kMaxSegs = 1 << 16;   // segments the scratch array holds: at most 134,217,728 nonzeros per call
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[2 * kMaxSegs];

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// The scratch array's address on the current device, looked up once per device.
void* scratch_ptr() {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("hipGetSymbolAddress returned no address for the scratch array");
        g_scratch[dev] = p;
    }
    return g_scratch[dev];
}

// Row containing nonzero k: the last r with ptr[r] <= k (ptr[rows] > k).
__device__ __forceinline__ int row_of(const int* __restrict__ ptr, int rows, int64_t k) {
    int lo = 0, hi = rows;  // invariant: ptr[lo] <= k < ptr[hi]
    while (hi - lo > 1) {
        const int mid = lo + (hi - lo) / 2;
        if (int64_t(ptr[mid]) <= k) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <class T>
__global__ __launch_bounds__(kWG) void k_seg_rows(int rows, int64_t nnz, const int* __restrict__ ptr,
                                                  const int* __restrict__ idx, const T* __restrict__ val,
                                                  const T* __restrict__ v, T* __restrict__ y, T* __restrict__ carry) {
    __shared__ T sh[kSegNnz];
    __shared__ int rr[2];
    const int b = int(blockIdx.x), tid = int(threadIdx.x);
    const int64_t k0 = int64_t(b) * kSegNnz, k1 = k0 + kSegNnz < nnz ? k0 + kSegNnz : nnz;
    if (tid == 0) rr[0] = row_of(ptr, rows, k0), rr[1] = row_of(ptr, rows, k1 - 1);
    const int n = int(k1 - k0);
    gather_rounds<kWG, kSegNnz / kWG>(0, n, idx + k0, val + k0, v, ColPlain{}, [&](int k, int, T x) { sh[k] = x; });
    __syncthreads();
    const int rf = rr[0], rl = rr[1];
    for (int r = rf + tid; r <= rl; r += kWG) {
        const int64_t e0 = ptr[r], e1 = ptr[r + 1];
        if (e0 == e1) continue;  // empty rows: the second kernel writes them
        const int a = int((e0 > k0 ? e0 : k0) - k0), e = int((e1 < k1 ? e1 : k1) - k0);
        T s = 0;
        for (int k = a; k < e; ++k) s += sh[k];
        if (r == rf && e0 < k0)
            carry[2 * b] = s;  // began in an earlier segment
        else if (r == rl && e1 > k1)
            carry[2 * b + 1] = s;  // continues into a later segment
        else
            y[r] = s;
    }
}

template <class T>
__global__ __launch_bounds__(kWG) void k_seg_fix(int rows, const int* __restrict__ ptr, const T* __restrict__ carry,
                                                 T* __restrict__ y) {
    for (int r = int(blockIdx.x * blockDim.x + threadIdx.x); r < rows; r += int(gridDim.x * blockDim.x)) {
        const int64_t e0 = ptr[r], e1 = ptr[r + 1];
        if (e0 == e1) {
            y[r] = 0;
            continue;
        }
        const int64_t b0 = e0 / kSegNnz, b1 = (e1 - 1) / kSegNnz;
        if (b0 == b1) continue;  // written whole by the first kernel
        T s = carry[2 * b0 + 1];
        for (int64_t q = b0 + 1; q <= b1; ++q) s += carry[2 * q];
        y[r] = s;
    }
}

template <class T>
void launch(int rows, int64_t nnz, const int* ptr, const int* idx, const void* val, const void* v, void* y,
            hipStream_t st) {
    T* carry = static_cast<T*>(scratch_ptr());  // never the symbol's host-side name
    const int nseg = int((nnz + kSegNnz - 1) / kSegNnz);
    if (nseg > 0)
        k_seg_rows<T><<<nseg, kWG, 0, st>>>(rows, nnz, ptr, idx, static_cast<const T*>(val), static_cast<const T*>(v),
                                            static_cast<T*>(y), carry);
    k_seg_fix<T><<<grid_for(rows), kWG, 0, st>>>(rows, ptr, carry, static_cast<T*>(y));
    HIP_OK(hipGetLastError());
}

}  // namespace

int64_t spmv_noalloc_max_nnz() { return int64_t(kMaxSegs) * kSegNnz; }

void spmv_noalloc(int precision_bits, int rows, int64_t nnz, const int* ptr, const int* idx, const void* val,
                  const void* v, void* y, hipStream_t st) {
    if (rows < 0 || nnz < 0) throw std::invalid_argument("rows and nnz must be >= 0");
    if (nnz > spmv_noalloc_max_nnz())
        throw std::length_error("nnz " + std::to_string(nnz) + " exceeds the " + std::to_string(spmv_noalloc_max_nnz()) +
                                " nonzeros the static scratch array covers");
    if (precision_bits != 32 && precision_bits != 64) throw std::invalid_argument("precision must be 32 or 64 bits");
    if (rows == 0) return;
    if (precision_bits == 32) launch<float>(rows, nnz, ptr, idx, val, v, y, st);
    else launch<double>(rows, nnz, ptr, idx, val, v, y, st);
}

}  // namespace tpdhg
