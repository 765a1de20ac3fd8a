// mlp.hpp -- the warm-start network: 6 -> 64 -> 64 -> 1 with SiLU. Layer 1
// (K = 6) in FP32 on the vector units; layers 2 and 3 on BF16 MFMA
// (v_mfma_f32_32x32x8bf16_1k) with FP32 accumulation.
//
// Parameters are one flat float vector: W1 (6x64, row-major [in][out]), b1
// (64), W2 (64x64, [in][out]), b2 (64), W3 (64), b3 (1). Inputs are rows of
// 8 floats (6 features, 2 zero pads). W2 and W3 are rounded to BF16 when
// they are loaded into the MFMA operands, as are the layer-1 activations;
// W1, the biases and all accumulation stay FP32.
#pragma once
#include "common.hpp"

namespace sph {

constexpr int MLP_NIN = 6, MLP_NH = 64, MLP_XSTRIDE = 8;
constexpr int MLP_OFF_W1 = 0, MLP_OFF_B1 = 384, MLP_OFF_W2 = 448, MLP_OFF_B2 = 4544, MLP_OFF_W3 = 4608,
              MLP_OFF_B3 = 4672, MLP_NPARAM = 4673;

// y[i] = net(x[i]) for i < n (device pointers).
void mlp_forward(hipStream_t s, const float *theta, const float *x, unsigned n, float *y);

// Gradient of the mean squared error sum_i (net(x_i) - t_i)^2 / n with
// respect to theta, and the loss (device pointers; scratch grows as needed).
class MLPTrainer {
public:
    DBuf<float> partial, grad, loss_part, adam_m, adam_v, loss;
    unsigned waves = 0;
    void loss_grad(hipStream_t s, const float *theta, const float *x, const float *t, unsigned n);
    // One Adam step on theta with the gradient of the last loss_grad.
    void adam(hipStream_t s, float *theta, int step, float lr);
};

}  // namespace sph
