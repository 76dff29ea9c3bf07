// The per-molecule energy surrogate: an MLP 32 -> 64 -> 64 -> 1 with SiLU,
// BF16 weights on the matrix cores (v_mfma_f32_16x16x16_bf16), FP32
// accumulation, FP32 master weights for the optimizer.
//
// Parameter vector (FP32, G6_MLP_NPARAM entries), in this order:
//   W1[64][32] (output-major), b1[64], W2[64][64], b2[64], w3[64], b3
// Normalisation (G6_MLP_NNORM entries): mean[32], inv_std[32] of the
// features, then y_shift, y_scale; the network sees (x - mean) * inv_std and
// its output o gives the energy y_scale * o + y_shift.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace g6 {

constexpr int kMlpIn = 32, kMlpHid = 64;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kMlpHid * kMlpIn, kOffW2 = kOffB1 + kMlpHid,
              kOffB2 = kOffW2 + kMlpHid * kMlpHid, kOffW3 = kOffB2 + kMlpHid,
              kOffB3 = kOffW3 + kMlpHid, kMlpParams = kOffB3 + 1;
constexpr int kMlpNorm = 2 * kMlpIn + 2;

struct Mlp {
  float *p = nullptr;       // [kMlpParams] FP32 master weights
  uint16_t *wb = nullptr;   // [64*32 + 64*64] BF16 copies of W1 and W2
  float *norm = nullptr;    // [kMlpNorm]
  float *adam_m = nullptr, *adam_v = nullptr;  // [kMlpParams]
  float *part = nullptr;    // [max blocks][kMlpParams + 1] gradient and loss partials
  float *grad = nullptr;    // [kMlpParams + 1] (last: loss)
  int part_blocks = 0;
  long long adam_t = 0;
  hipStream_t stream = nullptr;
};

// y[r] = energy of feature row x[r*32..] (r < m), device pointers; rows
// listed by idx when idx != nullptr.  With segcnt, the rows form segments of
// seg rows of which the first segcnt[k] are live; 16-row tiles wholly past
// them are not computed.  Stream-ordered, no host sync.
void mlp_launch_forward(const Mlp &net, long long m, const float *x, const int *idx, float *y,
                        hipStream_t s, int seg = 0, const int *segcnt = nullptr);
// gradient of the mean squared error in normalised units over the rows
// (idx or 0..m-1), into net.grad (loss in net.grad[kMlpParams])
int mlp_launch_grad(Mlp &net, long long m, const float *x, const float *t, const int *idx,
                    hipStream_t s);
// one Adam step with the gradient in net.grad; refreshes the BF16 copies
void mlp_launch_adam(Mlp &net, float lr, hipStream_t s, float *loss_out = nullptr);
Mlp *mlp_of(int64_t h);  // nullptr for an unknown g6_mlp handle

}  // namespace g6
