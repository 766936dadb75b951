// Warm-start network of the learned path: a 3-layer MLP (n_in -> h1 -> h2 -> n_out, ReLU) trained and run on the
// GPU with bf16 MFMA matrix products (v_mfma_f32_16x16x32_bf16: bf16 operands, FP32 accumulation), FP32 master
// weights and Adam. Every matrix product of the forward and backward passes is one launch of k_gemm_bf16: 64 x 64
// output tiles staged through LDS, four waves per workgroup. Shapes are padded to multiples of 16 (batch, widths).
//
// Where bf16 rounding happens (the CPU oracle rounds at the same places): inputs X, weight copies, hidden
// activations H1 and H2, and the back-propagated dY, dZ2, dZ1 before each product; biases, pre-activations,
// outputs, losses, gradients and Adam state stay FP32.
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <random>
#include <stdexcept>

#include "hipcommon.hpp"
#include "mlp.hpp"

namespace tpdhg {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float f) {  // round to nearest even (finite inputs)
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kTM = 64, kTN = 64, kTK = 32;  // workgroup tile (M x N) and K step
constexpr int kLdsK = kTK + 8;                 // LDS row length in bf16: 80-byte rows spread a wave's 16-byte reads

// The 8 bf16 of operand X at rows/cols [p0, p0 + 8) of one K step, as one thread's share of an LDS tile load.
// X(p, k) = X[p * s_p + k * s_k]; with s_k == 1 the thread takes 8 consecutive k of one row p (contiguous in
// memory), else 8 consecutive p at one k (contiguous when s_p == 1), so a wave's loads are coalesced either way.
struct TileLoad {
    uint16_t v[8];
    int p, k;       // first element within the tile
    bool along_k;   // the 8 elements run along k (else along p)
};
__device__ __forceinline__ void tile_fetch(TileLoad& t, const uint16_t* __restrict__ X, int s_p, int s_k, int p0,
                                           int k0, int P, int K) {
    const int tid = int(threadIdx.x);
    t.along_k = s_k == 1;
    if (t.along_k) t.p = tid >> 2, t.k = (tid & 3) * 8;  // 64 rows x 4 chunks of 8 k
    else t.k = tid >> 3, t.p = (tid & 7) * 8;            // 32 k x 8 chunks of 8 rows
    const int gp = p0 + t.p, gk = k0 + t.k;
    // P and K are multiples of 16 and chunks start at multiples of 8: a chunk is wholly inside or outside
    const bool in = gp < P && gk < K;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int dp = t.along_k ? 0 : j, dk = t.along_k ? j : 0;
        t.v[j] = in ? X[size_t(gp + dp) * size_t(s_p) + size_t(gk + dk) * size_t(s_k)] : uint16_t(0);
    }
}
// Stores the fetched chunk into the LDS tile sh[p][k] (k-contiguous rows of kLdsK).
__device__ __forceinline__ void tile_store(const TileLoad& t, uint16_t (*sh)[kLdsK]) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        if (t.along_k) sh[t.p][t.k + j] = t.v[j];
        else sh[t.p + j][t.k] = t.v[j];
    }
}

// C[M x N] = alpha * op(A)[M x K] op(B)[K x N] (+ bias[n]) (ReLU), A(i, k) = A[i * sa_m + k * sa_k],
// B(k, j) = B[k * sb_k + j * sb_n]; FP32 result to C (row-major, ldc) and, when Cb is given, a bf16 copy.
// M, N, K multiples of 16; grid (ceil(N/64), ceil(M/64)), 256 threads. Each workgroup stages 64 x 32 tiles of A
// and 32 x 64 tiles of B (both stored k-contiguous) through LDS per K step, the next step's global loads issued
// before the current step's products; each of its four waves owns a 32 x 32 quarter of the 64 x 64 output tile as
// 2 x 2 v_mfma_f32_16x16x32_bf16 tiles (bf16 operands, FP32 accumulation, K = 32 per instruction).
// MFMA 16x16x32 bf16 lane layout: lane l holds A(row l%16, k 8(l/16)..+7), B(k 8(l/16)..+7, col l%16) and
// D(rows 4(l/16)..+3, col l%16).
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const uint16_t* __restrict__ A, int sa_m,
                                                   int sa_k, const uint16_t* __restrict__ B, int sb_k, int sb_n,
                                                   float alpha, const float* __restrict__ bias, int relu,
                                                   float* __restrict__ C, int ldc, uint16_t* __restrict__ Cb) {
    __shared__ __align__(16) uint16_t As[kTM][kLdsK];  // As[i][k]
    __shared__ __align__(16) uint16_t Bs[kTN][kLdsK];  // Bs[j][k]
    const int wave = int(threadIdx.x) >> 6, lane = int(threadIdx.x) & 63;
    const int m0 = int(blockIdx.y) * kTM, n0 = int(blockIdx.x) * kTN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;  // the wave's quarter of the tile
    const int li = lane & 15, lk = 8 * (lane >> 4);
    floatx4 acc[2][2] = {};
    TileLoad ta, tb;
    tile_fetch(ta, A, sa_m, sa_k, m0, 0, M, K);
    tile_fetch(tb, B, sb_n, sb_k, n0, 0, N, K);  // B^T(j, k) = B[j * sb_n + k * sb_k]
    for (int k0 = 0; k0 < K; k0 += kTK) {
        __syncthreads();  // the previous step's LDS reads are done
        tile_store(ta, As), tile_store(tb, Bs);
        __syncthreads();
        if (k0 + kTK < K) {
            tile_fetch(ta, A, sa_m, sa_k, m0, k0 + kTK, M, K);
            tile_fetch(tb, B, sb_n, sb_k, n0, k0 + kTK, N, K);
        }
        bf16x8 a[2], b[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            a[t] = *reinterpret_cast<const bf16x8*>(&As[wm + 16 * t + li][lk]);
            b[t] = *reinterpret_cast<const bf16x8*>(&Bs[wn + 16 * t + li][lk]);
        }
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
                acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ti], b[tj], acc[ti][tj], 0, 0, 0);
    }
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) {
            const int c = n0 + wn + 16 * tj + li;
            if (c >= N) continue;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = m0 + wm + 16 * ti + 4 * (lane >> 4) + j;
                if (row >= M) continue;
                float v = alpha * acc[ti][tj][j];
                if (bias) v += bias[c];
                if (relu) v = v > 0.f ? v : 0.f;
                C[size_t(row) * size_t(ldc) + size_t(c)] = v;
                if (Cb) Cb[size_t(row) * size_t(ldc) + size_t(c)] = to_bf16(v);
            }
        }
}

__global__ void k_to_bf16(int64_t k, const float* __restrict__ src, uint16_t* __restrict__ dst) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < k; i += int64_t(gridDim.x) * blockDim.x)
        dst[i] = to_bf16(src[i]);
}

// Rows idx[0..valid) of src (rows x cols, FP32) into the first rows of dst (B x ld) as FP32 and bf16; rows from
// `valid` on and columns from `cols` on are zero.
__global__ void k_gather_rows(int B, int valid, int cols, int ld, const int* __restrict__ idx,
                              const float* __restrict__ src, float* __restrict__ dst, uint16_t* __restrict__ dstb) {
    const int total = B * ld;
    for (int t = int(blockIdx.x * blockDim.x + threadIdx.x); t < total; t += int(gridDim.x * blockDim.x)) {
        const int i = t / ld, j = t % ld;
        const float v = (i < valid && j < cols) ? src[size_t(idx[i]) * size_t(cols) + size_t(j)] : 0.f;
        if (dst) dst[t] = v;
        if (dstb) dstb[t] = to_bf16(v);
    }
}

// dY = 2 (Y - T) / (B * n_out) over the real outputs (zero in the padding) and per-block partial sums of the
// squared error for the loss (Y, T: B x ld).
__global__ __launch_bounds__(256) void k_mse(int B, int n_out, int ld, const float* __restrict__ Y,
                                             const float* __restrict__ T, float* __restrict__ dY,
                                             uint16_t* __restrict__ dYb, double* __restrict__ part) {
    double v[1] = {0};
    const float sc = 2.f / float(B * n_out);
    for (int t = int(blockIdx.x * blockDim.x + threadIdx.x); t < B * ld; t += int(gridDim.x * blockDim.x)) {
        const int j = t % ld;
        float g = 0;
        if (j < n_out) {
            const float d = Y[t] - T[t];
            v[0] += double(d) * double(d);
            g = sc * d;
        }
        dY[t] = g, dYb[t] = to_bf16(g);
    }
    block_store<1>(v, part);
}

// dZ = dH * (H > 0) as FP32 and bf16.
__global__ void k_relu_back(int64_t k, const float* __restrict__ dH, const float* __restrict__ H,
                            float* __restrict__ dZ, uint16_t* __restrict__ dZb) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < k; i += int64_t(gridDim.x) * blockDim.x) {
        const float g = H[i] > 0.f ? dH[i] : 0.f;
        dZ[i] = g, dZb[i] = to_bf16(g);
    }
}

// Column sums of a B x ld FP32 matrix (bias gradients), fixed order.
__global__ void k_colsum(int B, int ncols, int ld, const float* __restrict__ X, float* __restrict__ out) {
    for (int j = int(blockIdx.x * blockDim.x + threadIdx.x); j < ncols; j += int(gridDim.x * blockDim.x)) {
        float s = 0;
        for (int i = 0; i < B; ++i) s += X[size_t(i) * size_t(ld) + size_t(j)];
        out[j] = s;
    }
}

__global__ void k_adam(int64_t k, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, float lr, float b1, float b2, float c1, float c2, float eps) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < k; i += int64_t(gridDim.x) * blockDim.x) {
        const float gi = g[i];
        const float a = b1 * m1[i] + (1 - b1) * gi, b = b2 * m2[i] + (1 - b2) * gi * gi;
        m1[i] = a, m2[i] = b;
        w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
    }
}

int pad16(int k) { return (k + 15) / 16 * 16; }

class MlpImpl final : public Mlp {
public:
    MlpImpl(int n_in, int h1, int h2, int n_out, uint64_t seed) : in_(n_in), h1_(h1), h2_(h2), out_(n_out) {
        if (n_in % 16 || h1 % 16 || h2 % 16 || n_in <= 0 || h1 <= 0 || h2 <= 0 || n_out <= 0)
            throw std::invalid_argument("MLP widths: n_in, h1, h2 must be positive multiples of 16, n_out positive");
        require_device();
        op_ = pad16(n_out);
        HIP_OK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
        // parameter block (FP32): W1 [in x h1], b1, W2 [h1 x h2], b2, W3 [h2 x op], b3 [op]; padding columns of W3, b3
        // stay zero and get zero gradient
        off_[0] = 0, off_[1] = off_[0] + size_t(in_) * h1_, off_[2] = off_[1] + size_t(h1_);
        off_[3] = off_[2] + size_t(h1_) * h2_, off_[4] = off_[3] + size_t(h2_);
        off_[5] = off_[4] + size_t(h2_) * op_, off_[6] = off_[5] + size_t(op_);
        np_ = off_[6];
        w_ = mem_.alloc<float>(np_), g_ = mem_.alloc<float>(np_), m1_ = mem_.alloc<float>(np_), m2_ = mem_.alloc<float>(np_);
        wb_ = mem_.alloc<uint16_t>(np_);
        std::mt19937_64 rng(seed);
        std::vector<float> h(np_, 0.f);
        auto init = [&](size_t off, int rows, int cols, int real_cols) {  // He-uniform
            std::uniform_real_distribution<float> U(-std::sqrt(6.f / float(rows)), std::sqrt(6.f / float(rows)));
            for (int i = 0; i < rows; ++i)
                for (int j = 0; j < cols; ++j) h[off + size_t(i) * size_t(cols) + size_t(j)] = j < real_cols ? U(rng) : 0.f;
        };
        init(off_[0], in_, h1_, h1_), init(off_[2], h1_, h2_, h2_), init(off_[4], h2_, op_, out_);
        HIP_OK(hipMemcpyAsync(w_, h.data(), np_ * sizeof(float), hipMemcpyHostToDevice, st_));
        HIP_OK(hipMemsetAsync(m1_, 0, np_ * sizeof(float), st_));
        HIP_OK(hipMemsetAsync(m2_, 0, np_ * sizeof(float), st_));
        part_ = mem_.alloc<double>(1024), out_d_ = mem_.alloc<double>(4);
        HIP_OK(hipStreamSynchronize(st_));
        refresh_bf16();
    }
    ~MlpImpl() override {
        if (st_) (void)hipStreamDestroy(st_);
    }

    int64_t param_count() const override { return int64_t(in_) * h1_ + h1_ + int64_t(h1_) * h2_ + h2_ + int64_t(h2_) * out_ + out_; }

    void set_params(const float* p) override {
        std::vector<float> h(np_, 0.f);
        unpack(p, h.data());
        HIP_OK(hipMemcpyAsync(w_, h.data(), np_ * sizeof(float), hipMemcpyHostToDevice, st_));
        HIP_OK(hipStreamSynchronize(st_));
        refresh_bf16();
    }
    void get_params(float* p) override {
        std::vector<float> h(np_);
        HIP_OK(hipMemcpyAsync(h.data(), w_, np_ * sizeof(float), hipMemcpyDeviceToHost, st_));
        HIP_OK(hipStreamSynchronize(st_));
        pack(h.data(), p);
    }

    void forward(const float* x, int batch, float* y) override {
        ensure_batch(batch);
        upload_inputs(x, batch);
        run_forward(bp_);
        std::vector<float> h(size_t(bp_) * op_);
        HIP_OK(hipMemcpyAsync(h.data(), Y_, h.size() * sizeof(float), hipMemcpyDeviceToHost, st_));
        HIP_OK(hipStreamSynchronize(st_));
        for (int i = 0; i < batch; ++i) std::memcpy(y + size_t(i) * out_, h.data() + size_t(i) * op_, size_t(out_) * sizeof(float));
    }

    double loss_grad(const float* x, const float* t, int batch, float* grad) override {
        ensure_batch(batch);
        upload_inputs(x, batch);
        std::vector<float> ht(size_t(bp_) * op_, 0.f);
        for (int i = 0; i < batch; ++i) std::memcpy(ht.data() + size_t(i) * op_, t + size_t(i) * out_, size_t(out_) * sizeof(float));
        HIP_OK(hipMemcpyAsync(T_, ht.data(), ht.size() * sizeof(float), hipMemcpyHostToDevice, st_));
        const double loss = run_forward_backward(batch);
        if (grad) {
            std::vector<float> h(np_);
            HIP_OK(hipMemcpyAsync(h.data(), g_, np_ * sizeof(float), hipMemcpyDeviceToHost, st_));
            HIP_OK(hipStreamSynchronize(st_));
            pack(h.data(), grad);
        }
        return loss;
    }

    double train(const float* x, const float* t, int n, int batch, int epochs, float lr, uint64_t seed) override {
        if (n <= 0 || batch <= 0 || epochs < 0) throw std::invalid_argument("train: empty data or negative epochs");
        batch = std::min(batch, n);
        ensure_batch(batch);
        if (n > data_cap_) {  // training set resident on the device, kept for later calls
            data_cap_ = n;
            dx_ = mem_.alloc<float>(size_t(n) * in_), dt_ = mem_.alloc<float>(size_t(n) * out_);
            didx_ = mem_.alloc<int>(size_t(n));
        }
        float* dx = dx_;
        float* dt = dt_;
        int* didx = didx_;
        HIP_OK(hipMemcpyAsync(dx, x, size_t(n) * in_ * sizeof(float), hipMemcpyHostToDevice, st_));
        HIP_OK(hipMemcpyAsync(dt, t, size_t(n) * out_ * sizeof(float), hipMemcpyHostToDevice, st_));
        std::mt19937_64 rng(seed);
        std::vector<int> perm(static_cast<size_t>(n));
        double last = 0;
        for (int e = 0; e < epochs; ++e) {
            for (int i = 0; i < n; ++i) perm[size_t(i)] = i;
            std::shuffle(perm.begin(), perm.end(), rng);
            HIP_OK(hipMemcpyAsync(didx, perm.data(), size_t(n) * sizeof(int), hipMemcpyHostToDevice, st_));
            double sum = 0;
            int batches = 0;
            for (int b0 = 0; b0 + batch <= n; b0 += batch) {
                k_gather_rows<<<64, 256, 0, st_>>>(bp_, batch, in_, in_, didx + b0, dx, nullptr, Xb_);
                k_gather_rows<<<64, 256, 0, st_>>>(bp_, batch, out_, op_, didx + b0, dt, T_, nullptr);
                sum += run_forward_backward(batch);
                ++batches;
                adam_step(lr);
            }
            last = batches ? sum / batches : 0;
        }
        HIP_OK(hipStreamSynchronize(st_));
        return last;
    }

private:
    void ensure_batch(int batch) {
        if (batch <= 0) throw std::invalid_argument("batch must be positive");
        const int bp = pad16(batch);
        if (bp == bp_) return;
        if (bp > cap_) {
            cap_ = bp;
            Xb_ = mem_.alloc<uint16_t>(size_t(cap_) * in_);
            Z1_ = mem_.alloc<float>(size_t(cap_) * h1_), H1b_ = mem_.alloc<uint16_t>(size_t(cap_) * h1_);
            Z2_ = mem_.alloc<float>(size_t(cap_) * h2_), H2b_ = mem_.alloc<uint16_t>(size_t(cap_) * h2_);
            Y_ = mem_.alloc<float>(size_t(cap_) * op_), T_ = mem_.alloc<float>(size_t(cap_) * op_);
            dY_ = mem_.alloc<float>(size_t(cap_) * op_), dYb_ = mem_.alloc<uint16_t>(size_t(cap_) * op_);
            dH_ = mem_.alloc<float>(size_t(cap_) * std::max(h1_, h2_));
            dZ_ = mem_.alloc<float>(size_t(cap_) * std::max(h1_, h2_));
            dZb_ = mem_.alloc<uint16_t>(size_t(cap_) * std::max(h1_, h2_));
            HIP_OK(hipMemsetAsync(Xb_, 0, size_t(cap_) * in_ * sizeof(uint16_t), st_));
            HIP_OK(hipMemsetAsync(T_, 0, size_t(cap_) * op_ * sizeof(float), st_));
        }
        bp_ = bp;
    }
    void upload_inputs(const float* x, int batch) {
        std::vector<float> h(size_t(bp_) * in_, 0.f);
        std::memcpy(h.data(), x, size_t(batch) * in_ * sizeof(float));
        if (!xin_ || xin_cap_ < bp_) {
            xin_ = mem_.alloc<float>(size_t(bp_) * in_);
            xin_cap_ = bp_;
        }
        HIP_OK(hipMemcpyAsync(xin_, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, st_));
        k_to_bf16<<<64, 256, 0, st_>>>(int64_t(bp_) * in_, xin_, Xb_);
        rows_ = batch;
    }
    void gemm(int M, int N, int K, const uint16_t* A, int sa_m, int sa_k, const uint16_t* B, int sb_k, int sb_n,
              const float* bias, int relu, float* C, int ldc, uint16_t* Cb) {
        dim3 grid((N + kTN - 1) / kTN, (M + kTM - 1) / kTM);
        k_gemm_bf16<<<grid, 256, 0, st_>>>(M, N, K, A, sa_m, sa_k, B, sb_k, sb_n, 1.f, bias, relu, C, ldc, Cb);
    }
    void run_forward(int bp) {
        const uint16_t *W1 = wb_ + off_[0], *W2 = wb_ + off_[2], *W3 = wb_ + off_[4];
        gemm(bp, h1_, in_, Xb_, in_, 1, W1, h1_, 1, w_ + off_[1], 1, Z1_, h1_, H1b_);  // H1 = relu(X W1 + b1)
        gemm(bp, h2_, h1_, H1b_, h1_, 1, W2, h2_, 1, w_ + off_[3], 1, Z2_, h2_, H2b_);  // H2 = relu(H1 W2 + b2)
        gemm(bp, op_, h2_, H2b_, h2_, 1, W3, op_, 1, w_ + off_[5], 0, Y_, op_, nullptr);  // Y = H2 W3 + b3
        HIP_OK(hipGetLastError());
    }
    double run_forward_backward(int batch) {
        run_forward(bp_);
        const int bp = bp_;
        // rows beyond the batch carry zero inputs; their loss terms are masked by using B = batch rows only
        k_mse<<<1, 256, 0, st_>>>(batch, out_, op_, Y_, T_, dY_, dYb_, part_);
        k_finish<<<1, 256, 0, st_>>>(0, 1, 1, part_, out_d_);
        if (bp > batch) {
            HIP_OK(hipMemsetAsync(dY_ + size_t(batch) * op_, 0, size_t(bp - batch) * op_ * sizeof(float), st_));
            HIP_OK(hipMemsetAsync(dYb_ + size_t(batch) * op_, 0, size_t(bp - batch) * op_ * sizeof(uint16_t), st_));
        }
        const uint16_t *W2 = wb_ + off_[2], *W3 = wb_ + off_[4];
        float* g = g_;
        // layer 3: dW3 = H2^T dY, db3 = colsum dY, dH2 = dY W3^T
        gemm(h2_, op_, bp, H2b_, 1, h2_, dYb_, op_, 1, nullptr, 0, g + off_[4], op_, nullptr);
        k_colsum<<<(op_ + 255) / 256, 256, 0, st_>>>(bp, op_, op_, dY_, g + off_[5]);
        gemm(bp, h2_, op_, dYb_, op_, 1, W3, 1, op_, nullptr, 0, dH_, h2_, nullptr);
        k_relu_back<<<256, 256, 0, st_>>>(int64_t(bp) * h2_, dH_, Z2_, dZ_, dZb_);
        // layer 2: dW2 = H1^T dZ2, db2 = colsum dZ2, dH1 = dZ2 W2^T
        gemm(h1_, h2_, bp, H1b_, 1, h1_, dZb_, h2_, 1, nullptr, 0, g + off_[2], h2_, nullptr);
        k_colsum<<<(h2_ + 255) / 256, 256, 0, st_>>>(bp, h2_, h2_, dZ_, g + off_[3]);
        gemm(bp, h1_, h2_, dZb_, h2_, 1, W2, 1, h2_, nullptr, 0, dH_, h1_, nullptr);
        k_relu_back<<<256, 256, 0, st_>>>(int64_t(bp) * h1_, dH_, Z1_, dZ_, dZb_);
        // layer 1: dW1 = X^T dZ1, db1 = colsum dZ1
        gemm(in_, h1_, bp, Xb_, 1, in_, dZb_, h1_, 1, nullptr, 0, g + off_[0], h1_, nullptr);
        k_colsum<<<(h1_ + 255) / 256, 256, 0, st_>>>(bp, h1_, h1_, dZ_, g + off_[1]);
        double hl = 0;
        HIP_OK(hipMemcpyAsync(&hl, out_d_, sizeof(double), hipMemcpyDeviceToHost, st_));
        HIP_OK(hipStreamSynchronize(st_));
        HIP_OK(hipGetLastError());
        return hl / double(batch * out_);
    }
    void adam_step(float lr) {
        ++step_;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, float(step_)), c2 = 1.f - std::pow(b2, float(step_));
        k_adam<<<512, 256, 0, st_>>>(int64_t(np_), w_, g_, m1_, m2_, lr, b1, b2, c1, c2, 1e-8f);
        refresh_bf16();
    }
    void refresh_bf16() {
        k_to_bf16<<<512, 256, 0, st_>>>(int64_t(np_), w_, wb_);
        HIP_OK(hipGetLastError());
    }
    // external layout: W1 [in x h1], b1, W2 [h1 x h2], b2, W3 [h2 x out], b3 [out]
    void unpack(const float* p, float* h) const {
        size_t q = 0;
        for (size_t k = 0; k < off_[4]; ++k) h[k] = p[q++];
        for (int i = 0; i < h2_; ++i)
            for (int j = 0; j < out_; ++j) h[off_[4] + size_t(i) * op_ + size_t(j)] = p[q++];
        for (int j = 0; j < out_; ++j) h[off_[5] + size_t(j)] = p[q++];
    }
    void pack(const float* h, float* p) const {
        size_t q = 0;
        for (size_t k = 0; k < off_[4]; ++k) p[q++] = h[k];
        for (int i = 0; i < h2_; ++i)
            for (int j = 0; j < out_; ++j) p[q++] = h[off_[4] + size_t(i) * op_ + size_t(j)];
        for (int j = 0; j < out_; ++j) p[q++] = h[off_[5] + size_t(j)];
    }

    int in_, h1_, h2_, out_, op_ = 0;
    size_t off_[7] = {};
    size_t np_ = 0;
    Mem mem_;
    hipStream_t st_ = nullptr;
    float *w_ = nullptr, *g_ = nullptr, *m1_ = nullptr, *m2_ = nullptr;
    uint16_t* wb_ = nullptr;
    int bp_ = 0, cap_ = 0, rows_ = 0, xin_cap_ = 0;
    int64_t step_ = 0;
    uint16_t *Xb_ = nullptr, *H1b_ = nullptr, *H2b_ = nullptr, *dYb_ = nullptr, *dZb_ = nullptr;
    float *Z1_ = nullptr, *Z2_ = nullptr, *Y_ = nullptr, *T_ = nullptr, *dY_ = nullptr, *dH_ = nullptr, *dZ_ = nullptr;
    float* xin_ = nullptr;
    float *dx_ = nullptr, *dt_ = nullptr;
    int* didx_ = nullptr;
    int data_cap_ = 0;
    double *part_ = nullptr, *out_d_ = nullptr;
};

}  // namespace

std::unique_ptr<Mlp> make_mlp(int n_in, int h1, int h2, int n_out, uint64_t seed) {
    return std::unique_ptr<Mlp>(new MlpImpl(n_in, h1, h2, n_out, seed));
}

}  // namespace tpdhg
