// Warm-start network of the learned path (no reference counterpart): n_in -> h1 -> h2 -> n_out MLP with ReLU,
// bf16 MFMA products with FP32 accumulation, FP32 master weights, trained with Adam on the GPU (mlp.hip).
#pragma once
#include <cstdint>
#include <memory>

namespace tpdhg {

class Mlp {
public:
    virtual ~Mlp() = default;
    // Parameters in the external layout W1 [n_in x h1], b1 [h1], W2 [h1 x h2], b2 [h2], W3 [h2 x n_out],
    // b3 [n_out], row-major FP32.
    virtual int64_t param_count() const = 0;
    virtual void set_params(const float* p) = 0;
    virtual void get_params(float* p) = 0;
    // y [batch x n_out] for x [batch x n_in].
    virtual void forward(const float* x, int batch, float* y) = 0;
    // Mean squared error over batch x n_out and its gradient (external layout; may be null).
    virtual double loss_grad(const float* x, const float* t, int batch, float* grad) = 0;
    // Adam (beta 0.9 / 0.999) over shuffled mini-batches, data resident on the device; returns the mean loss of
    // the last epoch.
    virtual double train(const float* x, const float* t, int n, int batch, int epochs, float lr, uint64_t seed) = 0;
};

// Throws std::runtime_error without a HIP device or for widths that are not positive multiples of 16 (n_in,
// h1, h2; n_out is padded internally).
std::unique_ptr<Mlp> make_mlp(int n_in, int h1, int h2, int n_out, uint64_t seed);

}  // namespace tpdhg
