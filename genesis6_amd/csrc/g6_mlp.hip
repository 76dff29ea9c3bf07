// The energy surrogate's network on the matrix cores: inference, the
// gradient of the training loss, and Adam, all hand-written.
//
// A wave owns a tile of 16 feature rows.  The forward layer products run as
// gfx950's v_mfma_f32_16x16x32_bf16 (A: lane l holds row l%16, k =
// 8(l/16)..+7; B: k = 8(l/16)..+7 of column l%16; D: rows 4(l/16)..+3 of
// column l%16).  Their activation operands are split into two BF16 terms,
// hi = bf16(v) and lo = bf16(v - hi) (16 significant bits), each multiplied
// by the BF16 weights: a single BF16 rounding (8 bits) would make the
// network's output step by the rounding of its inputs, and a volume move's
// energy change (every feature moved by ~0.2 %) would be lost in those
// steps.  Activations cross from the D layout to the A layout through LDS.
// The backward pass runs v_mfma_f32_16x16x16_bf16 on the hi terms and reuses
// the D layout directly as the transposed A operand of the weight gradients
// (dW = dZ^T H contracts over the tile's 16 rows).  Gradient partials are
// summed in a fixed order (waves of a block, then blocks), so a gradient is
// bitwise reproducible.
#include <math.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "g6_internal.h"
#include "g6_mlp.h"

namespace g6 {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8b __attribute__((ext_vector_type(8)));
typedef unsigned short v8u __attribute__((ext_vector_type(8)));

// FP32 -> BF16, round to nearest even (finite inputs)
__device__ inline uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ inline float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ inline float dsilu(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

__device__ inline v4s pack4(uint16_t a, uint16_t b, uint16_t c, uint16_t d) {
  v4s r;
  r.x = (short)a;
  r.y = (short)b;
  r.z = (short)c;
  r.w = (short)d;
  return r;
}
__device__ inline v4s lds4(const uint16_t *p) { return pack4(p[0], p[1], p[2], p[3]); }
__device__ inline v4f mfma(v4s a, v4s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// eight consecutive BF16 values (16-byte aligned) as an MFMA operand
__device__ inline v8b lds8(const uint16_t *p) {
  return __builtin_bit_cast(v8b, *reinterpret_cast<const v8u *>(p));
}
__device__ inline v4f mfma32(v8b a, v8b b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ inline void wsync_m() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// sum over the 16 lanes of a row group (butterfly; every lane gets it)
__device__ inline float sum16(float v) {
#pragma unroll
  for (int d = 8; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

template <bool GRAD>
__global__ __launch_bounds__(256) void k_mlp(const float *__restrict__ P,
                                             const uint16_t *__restrict__ WB,
                                             const float *__restrict__ NRM, long long m,
                                             const float *__restrict__ x,
                                             const float *__restrict__ t,
                                             const int *__restrict__ idx, float *__restrict__ y,
                                             float *__restrict__ part, int seg,
                                             const int *__restrict__ segcnt) {
  __shared__ __align__(16) uint16_t s_w1[kMlpHid * kMlpIn], s_w2[kMlpHid * kMlpHid];
  __shared__ float s_b1[kMlpHid], s_b2[kMlpHid], s_w3[kMlpHid], s_mean[kMlpIn], s_istd[kMlpIn];
  __shared__ __align__(16) uint16_t s_xn[4][16 * kMlpIn], s_h1[4][16 * kMlpHid];
  __shared__ __align__(16) uint16_t s_h1lo[4][16 * kMlpHid];
  __shared__ uint16_t s_dz[GRAD ? 4 : 1][GRAD ? 16 * kMlpHid : 1];
  __shared__ float s_acc[GRAD ? kMlpParams + 1 : 1];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int r16 = lane & 15, q = lane >> 4;
  for (int k = tid; k < kMlpHid * kMlpIn; k += 256) s_w1[k] = WB[k];
  for (int k = tid; k < kMlpHid * kMlpHid; k += 256) s_w2[k] = WB[kMlpHid * kMlpIn + k];
  if (tid < kMlpHid) {
    s_b1[tid] = P[kOffB1 + tid];
    s_b2[tid] = P[kOffB2 + tid];
    s_w3[tid] = P[kOffW3 + tid];
  }
  if (tid < kMlpIn) {
    s_mean[tid] = NRM[tid];
    s_istd[tid] = NRM[kMlpIn + tid];
  }
  if (GRAD)
    for (int k = tid; k < kMlpParams + 1; k += 256) s_acc[k] = 0.0f;
  __syncthreads();
  const float b3 = P[kOffB3], yshift = NRM[2 * kMlpIn], yscale = NRM[2 * kMlpIn + 1];
  v4f gW2[4][4], gW1[4][2];
  float gb1[4], gb2[4], gw3[4], gb3 = 0.0f, loss = 0.0f;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    gb1[c] = gb2[c] = gw3[c] = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; k++) gW2[c][k] = v4f{0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 2; k++) gW1[c][k] = v4f{0, 0, 0, 0};
  }
  const long long ntile = (m + 15) / 16;
  uint16_t *xn = s_xn[wid], *h1 = s_h1[wid], *h1lo = s_h1lo[wid];
  for (long long tile = (long long)blockIdx.x * 4 + wid; tile < ntile;
       tile += (long long)gridDim.x * 4) {
    const long long row0 = tile * 16;
    // rows in segments of seg with segcnt[k] live rows each: a tile wholly
    // past its segment's live rows is skipped
    if (segcnt && row0 / seg == (row0 + 15) / seg && row0 % seg >= segcnt[row0 / seg]) continue;
    wsync_m();  // the previous tile's LDS reads are done
    // layer 1: z1 = W1 xn + b1; the lane's A operand is its own row r16,
    // features 8q..8q+7, normalised in FP32 and split into hi + lo
    v4f z1[4];
    {
      const long long r = row0 + r16;
      const bool ok = r < m;
      const long long src = ok ? (idx ? (long long)idx[r] : r) : 0;
      v8u ahi, alo;
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int f = 8 * q + j;
        const float v = ok ? (x[src * kMlpIn + f] - s_mean[f]) * s_istd[f] : 0.0f;
        const uint16_t h = f2bf(v);
        ahi[j] = h;
        alo[j] = f2bf(v - bf2f(h));
        if (GRAD) xn[r16 * kMlpIn + f] = h;  // the weight gradient's operand
      }
      const v8b a_hi = __builtin_bit_cast(v8b, ahi), a_lo = __builtin_bit_cast(v8b, alo);
#pragma unroll
      for (int ct = 0; ct < 4; ct++) {
        const v8b w = lds8(s_w1 + (16 * ct + r16) * kMlpIn + 8 * q);
        v4f acc = {0, 0, 0, 0};
        acc = mfma32(a_hi, w, acc);
        acc = mfma32(a_lo, w, acc);
        const float bb = s_b1[16 * ct + r16];
        z1[ct] = v4f{acc.x + bb, acc.y + bb, acc.z + bb, acc.w + bb};
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const float hv = silu(z1[ct][i]);
          const uint16_t h = f2bf(hv);
          h1[(4 * q + i) * kMlpHid + 16 * ct + r16] = h;
          h1lo[(4 * q + i) * kMlpHid + 16 * ct + r16] = f2bf(hv - bf2f(h));
        }
      }
    }
    wsync_m();
    // layer 2: z2 = W2 (h1_hi + h1_lo) + b2, k = 32 kb + 8q..+7
    v4f z2[4];
    {
      v8b ah[2], al[2];
#pragma unroll
      for (int kb = 0; kb < 2; kb++) {
        ah[kb] = lds8(h1 + r16 * kMlpHid + 32 * kb + 8 * q);
        al[kb] = lds8(h1lo + r16 * kMlpHid + 32 * kb + 8 * q);
      }
#pragma unroll
      for (int ct = 0; ct < 4; ct++) {
        const uint16_t *w = s_w2 + (16 * ct + r16) * kMlpHid + 8 * q;
        v4f acc = {0, 0, 0, 0};
#pragma unroll
        for (int kb = 0; kb < 2; kb++) {
          const v8b wb = lds8(w + 32 * kb);
          acc = mfma32(ah[kb], wb, acc);
          acc = mfma32(al[kb], wb, acc);
        }
        const float bb = s_b2[16 * ct + r16];
        z2[ct] = v4f{acc.x + bb, acc.y + bb, acc.z + bb, acc.w + bb};
      }
    }
    // output o = w3 . silu(z2) + b3 for rows 4q..4q+3
    float o[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      float s = 0.0f;
#pragma unroll
      for (int ct = 0; ct < 4; ct++) s += silu(z2[ct][i]) * s_w3[16 * ct + r16];
      o[i] = sum16(s) + b3;
    }
    if (!GRAD) {
      if (r16 == 0)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const long long r = row0 + 4 * q + i;
          if (r < m) y[r] = yscale * o[i] + yshift;
        }
      continue;
    }
    // ---- backward: loss = mean over rows of (o - (t - shift)/scale)^2 ----
    float g[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const long long r = row0 + 4 * q + i;
      float d = 0.0f;
      if (r < m) {
        const long long src = idx ? (long long)idx[r] : r;
        d = o[i] - (t[src] - yshift) / yscale;
      }
      g[i] = 2.0f * d / (float)m;
      if (r16 == 0) {
        loss += d * d;
        gb3 += g[i];
      }
    }
    v4f dz2[4];
#pragma unroll
    for (int ct = 0; ct < 4; ct++) {
      const float w3 = s_w3[16 * ct + r16];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        dz2[ct][i] = g[i] * w3 * dsilu(z2[ct][i]);
        gw3[ct] += g[i] * silu(z2[ct][i]);
        gb2[ct] += dz2[ct][i];
      }
    }
    // dW2[c][k] += sum_rows dz2[row][c] h1[row][k]
    {
      v4s bh[4];
#pragma unroll
      for (int kt = 0; kt < 4; kt++) {
        const uint16_t *p = h1 + (4 * q) * kMlpHid + 16 * kt + r16;
        bh[kt] = pack4(p[0], p[kMlpHid], p[2 * kMlpHid], p[3 * kMlpHid]);
      }
      uint16_t *dz = s_dz[GRAD ? wid : 0];
#pragma unroll
      for (int ct = 0; ct < 4; ct++) {
        const v4s a = pack4(f2bf(dz2[ct][0]), f2bf(dz2[ct][1]), f2bf(dz2[ct][2]), f2bf(dz2[ct][3]));
#pragma unroll
        for (int kt = 0; kt < 4; kt++) gW2[ct][kt] = mfma(a, bh[kt], gW2[ct][kt]);
#pragma unroll
        for (int i = 0; i < 4; i++) dz[(4 * q + i) * kMlpHid + 16 * ct + r16] = f2bf(dz2[ct][i]);
      }
      wsync_m();
      // dh1[row][k] = sum_c dz2[row][c] W2[c][k]; dz1 = dh1 silu'(z1)
      v4s ad[4];
#pragma unroll
      for (int cb = 0; cb < 4; cb++) ad[cb] = lds4(dz + r16 * kMlpHid + 16 * cb + 4 * q);
      v4f dz1[4];
#pragma unroll
      for (int kt = 0; kt < 4; kt++) {
        v4f acc = {0, 0, 0, 0};
#pragma unroll
        for (int cb = 0; cb < 4; cb++) {
          const uint16_t *w = s_w2 + (16 * cb + 4 * q) * kMlpHid + 16 * kt + r16;
          acc = mfma(ad[cb], pack4(w[0], w[kMlpHid], w[2 * kMlpHid], w[3 * kMlpHid]), acc);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
          dz1[kt][i] = acc[i] * dsilu(z1[kt][i]);
          gb1[kt] += dz1[kt][i];
        }
      }
      // dW1[c][k] += sum_rows dz1[row][c] xn[row][k]
      v4s bx[2];
#pragma unroll
      for (int kt = 0; kt < 2; kt++) {
        const uint16_t *p = xn + (4 * q) * kMlpIn + 16 * kt + r16;
        bx[kt] = pack4(p[0], p[kMlpIn], p[2 * kMlpIn], p[3 * kMlpIn]);
      }
#pragma unroll
      for (int ct = 0; ct < 4; ct++) {
        const v4s a = pack4(f2bf(dz1[ct][0]), f2bf(dz1[ct][1]), f2bf(dz1[ct][2]), f2bf(dz1[ct][3]));
#pragma unroll
        for (int kt = 0; kt < 2; kt++) gW1[ct][kt] = mfma(a, bx[kt], gW1[ct][kt]);
      }
    }
  }
  if (!GRAD) return;
  // block partial: waves in order, within a wave the four row groups in order
  for (int w = 0; w < 4; w++) {
    __syncthreads();
    if (wid != w) continue;
#pragma unroll
    for (int ct = 0; ct < 4; ct++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int c = 16 * ct + 4 * q + i;
#pragma unroll
        for (int kt = 0; kt < 4; kt++) s_acc[kOffW2 + c * kMlpHid + 16 * kt + r16] += gW2[ct][kt][i];
#pragma unroll
        for (int kt = 0; kt < 2; kt++) s_acc[kOffW1 + c * kMlpIn + 16 * kt + r16] += gW1[ct][kt][i];
      }
    for (int qq = 0; qq < 4; qq++) {
      if (q == qq) {
#pragma unroll
        for (int ct = 0; ct < 4; ct++) {
          s_acc[kOffB1 + 16 * ct + r16] += gb1[ct];
          s_acc[kOffB2 + 16 * ct + r16] += gb2[ct];
          s_acc[kOffW3 + 16 * ct + r16] += gw3[ct];
        }
        if (r16 == 0) {
          s_acc[kOffB3] += gb3;
          s_acc[kMlpParams] += loss;
        }
      }
      wsync_m();
    }
  }
  __syncthreads();
  for (int k = tid; k < kMlpParams + 1; k += 256)
    part[(size_t)blockIdx.x * (kMlpParams + 1) + k] = s_acc[k];
}

// grad[k] = sum over blocks b = 0.. of part[b][k]; the loss entry is
// divided by the row count
__global__ void k_mlp_reduce(const float *__restrict__ part, int nb, float *__restrict__ grad,
                             long long m) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > kMlpParams) return;
  // block order; sixteen loads in flight at a time
  float s = 0.0f;
  int b = 0;
  for (; b + 16 <= nb; b += 16) {
    float v[16];
#pragma unroll
    for (int u = 0; u < 16; u++) v[u] = part[(size_t)(b + u) * (kMlpParams + 1) + k];
#pragma unroll
    for (int u = 0; u < 16; u++) s += v[u];
  }
  for (; b < nb; b++) s += part[(size_t)b * (kMlpParams + 1) + k];
  grad[k] = k == kMlpParams ? s / (float)m : s;
}

// Adam step on every parameter; loss_out (if set) receives the batch loss
// grad[kMlpParams], so a training epoch keeps its losses on the device
__global__ void k_mlp_adam(float *__restrict__ p, float *__restrict__ mo, float *__restrict__ ve,
                           const float *__restrict__ grad, uint16_t *__restrict__ wb, float lr,
                           float c1, float c2, float *__restrict__ loss_out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k == 0 && loss_out) *loss_out = grad[kMlpParams];
  if (k >= kMlpParams) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gk = grad[k];
  const float m1 = b1 * mo[k] + (1.0f - b1) * gk;
  const float v1 = b2 * ve[k] + (1.0f - b2) * gk * gk;
  mo[k] = m1;
  ve[k] = v1;
  const float pk = p[k] - lr * (m1 * c1) / (sqrtf(v1 * c2) + eps);
  p[k] = pk;
  if (k < kOffB1) wb[k] = f2bf(pk);
  else if (k >= kOffW2 && k < kOffB2) wb[kMlpHid * kMlpIn + (k - kOffW2)] = f2bf(pk);
}

// One training step after k_mlp<true>: each thread sums its parameter's
// block partials (k_mlp_reduce's order) and applies k_mlp_adam's update to
// it; the loss entry goes to loss_out.  Equal to k_mlp_reduce followed by
// k_mlp_adam, with one launch fewer.
__global__ void k_mlp_reduce_adam(const float *__restrict__ part, int nb, long long m,
                                  float *__restrict__ p, float *__restrict__ mo,
                                  float *__restrict__ ve, float *__restrict__ grad,
                                  uint16_t *__restrict__ wb, float lr, float c1, float c2,
                                  float *__restrict__ loss_out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > kMlpParams) return;
  float s = 0.0f;
  int b = 0;
  for (; b + 16 <= nb; b += 16) {
    float v[16];
#pragma unroll
    for (int u = 0; u < 16; u++) v[u] = part[(size_t)(b + u) * (kMlpParams + 1) + k];
#pragma unroll
    for (int u = 0; u < 16; u++) s += v[u];
  }
  for (; b < nb; b++) s += part[(size_t)b * (kMlpParams + 1) + k];
  if (k == kMlpParams) {
    const float loss = s / (float)m;
    grad[k] = loss;
    if (loss_out) *loss_out = loss;
    return;
  }
  grad[k] = s;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gk = s;
  const float m1 = b1 * mo[k] + (1.0f - b1) * gk;
  const float v1 = b2 * ve[k] + (1.0f - b2) * gk * gk;
  mo[k] = m1;
  ve[k] = v1;
  const float pk = p[k] - lr * (m1 * c1) / (sqrtf(v1 * c2) + eps);
  p[k] = pk;
  if (k < kOffB1) wb[k] = f2bf(pk);
  else if (k >= kOffW2 && k < kOffB2) wb[kMlpHid * kMlpIn + (k - kOffW2)] = f2bf(pk);
}

// BF16 copies of W1 and W2 from the master weights
__global__ void k_mlp_cast(const float *__restrict__ p, uint16_t *__restrict__ wb) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < kMlpHid * kMlpIn) wb[k] = f2bf(p[kOffW1 + k]);
  else if (k < kMlpHid * kMlpIn + kMlpHid * kMlpHid) wb[k] = f2bf(p[kOffW2 + k - kMlpHid * kMlpIn]);
}

constexpr int kMlpGradBlocks = 240;

void mlp_launch_forward(const Mlp &net, long long m, const float *x, const int *idx, float *y,
                        hipStream_t s, int seg, const int *segcnt) {
  if (m <= 0) return;
  const long long blocks = ((m + 15) / 16 + 3) / 4;
  k_mlp<false><<<(unsigned)blocks, 256, 0, s>>>(net.p, net.wb, net.norm, m, x, nullptr, idx, y,
                                                nullptr, seg, segcnt);
}

int mlp_launch_grad(Mlp &net, long long m, const float *x, const float *t, const int *idx,
                    hipStream_t s) {
  if (m <= 0) return G6_EARG;
  long long blocks = ((m + 15) / 16 + 3) / 4;
  if (blocks > net.part_blocks) blocks = net.part_blocks;
  k_mlp<true><<<(unsigned)blocks, 256, 0, s>>>(net.p, net.wb, net.norm, m, x, t, idx, nullptr,
                                               net.part, 0, nullptr);
  k_mlp_reduce<<<(kMlpParams + 1 + 255) / 256, 256, 0, s>>>(net.part, (int)blocks, net.grad, m);
  return 0;
}

// gradient of one batch and the Adam step, three launches in two
static int mlp_launch_train_step(Mlp &net, long long m, const float *x, const float *t,
                                 const int *idx, float lr, float *loss_out, hipStream_t s) {
  if (m <= 0) return G6_EARG;
  long long blocks = ((m + 15) / 16 + 3) / 4;
  if (blocks > net.part_blocks) blocks = net.part_blocks;
  k_mlp<true><<<(unsigned)blocks, 256, 0, s>>>(net.p, net.wb, net.norm, m, x, t, idx, nullptr,
                                               net.part, 0, nullptr);
  net.adam_t++;
  const double c1 = 1.0 / (1.0 - pow(0.9, (double)net.adam_t));
  const double c2 = 1.0 / (1.0 - pow(0.999, (double)net.adam_t));
  k_mlp_reduce_adam<<<(kMlpParams + 1 + 255) / 256, 256, 0, s>>>(
      net.part, (int)blocks, m, net.p, net.adam_m, net.adam_v, net.grad, net.wb, lr, (float)c1,
      (float)c2, loss_out);
  return 0;
}

void mlp_launch_adam(Mlp &net, float lr, hipStream_t s, float *loss_out) {
  net.adam_t++;
  const double c1 = 1.0 / (1.0 - pow(0.9, (double)net.adam_t));
  const double c2 = 1.0 / (1.0 - pow(0.999, (double)net.adam_t));
  k_mlp_adam<<<(kMlpParams + 255) / 256, 256, 0, s>>>(net.p, net.adam_m, net.adam_v, net.grad,
                                                      net.wb, lr, (float)c1, (float)c2, loss_out);
}

static std::mutex g_mmu;
static std::vector<Mlp *> g_mlp;

static Mlp *mget(g6_mlp h) {
  std::lock_guard<std::mutex> lk(g_mmu);
  if (h <= 0 || h > (g6_mlp)g_mlp.size()) return nullptr;
  return g_mlp[h - 1];
}

static void mlp_release(Mlp &n) {
  if (n.p) (void)hipFree(n.p);
  if (n.wb) (void)hipFree(n.wb);
  if (n.norm) (void)hipFree(n.norm);
  if (n.adam_m) (void)hipFree(n.adam_m);
  if (n.adam_v) (void)hipFree(n.adam_v);
  if (n.part) (void)hipFree(n.part);
  if (n.grad) (void)hipFree(n.grad);
  if (n.stream) (void)hipStreamDestroy(n.stream);
}

Mlp *mlp_of(g6_mlp h) { return mget(h); }

// device scratch of a host-array call
struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace g6

using namespace g6;

extern "C" {

int g6_mlp_create(g6_mlp *hm) {
  if (!hm) return G6_EARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return G6_ENODEVICE;
  Mlp *n = new Mlp();
  n->part_blocks = kMlpGradBlocks;
  bool ok = hipMalloc(&n->p, sizeof(float) * kMlpParams) == hipSuccess &&
            hipMalloc(&n->wb, sizeof(uint16_t) * (kMlpHid * kMlpIn + kMlpHid * kMlpHid)) ==
                hipSuccess &&
            hipMalloc(&n->norm, sizeof(float) * kMlpNorm) == hipSuccess &&
            hipMalloc(&n->adam_m, sizeof(float) * kMlpParams) == hipSuccess &&
            hipMalloc(&n->adam_v, sizeof(float) * kMlpParams) == hipSuccess &&
            hipMalloc(&n->part, sizeof(float) * (kMlpParams + 1) * kMlpGradBlocks) ==
                hipSuccess &&
            hipMalloc(&n->grad, sizeof(float) * (kMlpParams + 1)) == hipSuccess &&
            hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) == hipSuccess;
  if (ok) {
    // zero weights, identity normalisation, cleared optimizer state
    std::vector<float> z(kMlpParams, 0.0f), nm(kMlpNorm, 0.0f);
    for (int f = 0; f < kMlpIn; f++) nm[kMlpIn + f] = 1.0f;
    nm[2 * kMlpIn + 1] = 1.0f;
    ok = hipMemcpy(n->p, z.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice) ==
             hipSuccess &&
         hipMemcpy(n->adam_m, z.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice) ==
             hipSuccess &&
         hipMemcpy(n->adam_v, z.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice) ==
             hipSuccess &&
         hipMemcpy(n->norm, nm.data(), sizeof(float) * kMlpNorm, hipMemcpyHostToDevice) ==
             hipSuccess;
    if (ok) {
      k_mlp_cast<<<(kMlpHid * (kMlpIn + kMlpHid) + 255) / 256, 256, 0, n->stream>>>(n->p, n->wb);
      ok = hipStreamSynchronize(n->stream) == hipSuccess;
    }
  }
  if (!ok) {
    mlp_release(*n);
    delete n;
    return G6_EHIP;
  }
  std::lock_guard<std::mutex> lk(g_mmu);
  g_mlp.push_back(n);
  *hm = (g6_mlp)g_mlp.size();
  return 0;
}

int g6_mlp_free(g6_mlp hm) {
  Mlp *n = mget(hm);
  if (!n) return G6_EARG;
  (void)hipStreamSynchronize(n->stream);
  mlp_release(*n);
  delete n;
  std::lock_guard<std::mutex> lk(g_mmu);
  g_mlp[hm - 1] = nullptr;
  return 0;
}

int g6_mlp_set_params(g6_mlp hm, const float *params, const float *norm) {
  Mlp *n = mget(hm);
  if (!n) return G6_EARG;
  G6_CHECK(hipStreamSynchronize(n->stream));
  if (params) {
    G6_CHECK(hipMemcpy(n->p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    k_mlp_cast<<<(kMlpHid * (kMlpIn + kMlpHid) + 255) / 256, 256, 0, n->stream>>>(n->p, n->wb);
    // a new starting point: the optimizer state restarts
    G6_CHECK(hipMemsetAsync(n->adam_m, 0, sizeof(float) * kMlpParams, n->stream));
    G6_CHECK(hipMemsetAsync(n->adam_v, 0, sizeof(float) * kMlpParams, n->stream));
    n->adam_t = 0;
  }
  if (norm) G6_CHECK(hipMemcpy(n->norm, norm, sizeof(float) * kMlpNorm, hipMemcpyHostToDevice));
  G6_CHECK(hipStreamSynchronize(n->stream));
  return 0;
}

int g6_mlp_get_params(g6_mlp hm, float *params, float *norm) {
  Mlp *n = mget(hm);
  if (!n) return G6_EARG;
  G6_CHECK(hipStreamSynchronize(n->stream));
  if (params) G6_CHECK(hipMemcpy(params, n->p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
  if (norm) G6_CHECK(hipMemcpy(norm, n->norm, sizeof(float) * kMlpNorm, hipMemcpyDeviceToHost));
  return 0;
}

int g6_mlp_forward(g6_mlp hm, int64_t m, const float *x, float *y) {
  Mlp *n = mget(hm);
  if (!n || m < 1 || !x || !y) return G6_EARG;
  DevBuf dx, dy;
  G6_CHECK(hipMalloc(&dx.p, sizeof(float) * kMlpIn * m));
  G6_CHECK(hipMalloc(&dy.p, sizeof(float) * m));
  G6_CHECK(hipMemcpy(dx.p, x, sizeof(float) * kMlpIn * m, hipMemcpyHostToDevice));
  mlp_launch_forward(*n, m, (const float *)dx.p, nullptr, (float *)dy.p, n->stream);
  G6_CHECK(hipGetLastError());
  G6_CHECK(hipMemcpyAsync(y, dy.p, sizeof(float) * m, hipMemcpyDeviceToHost, n->stream));
  G6_CHECK(hipStreamSynchronize(n->stream));
  return 0;
}

int g6_mlp_loss_grad(g6_mlp hm, int64_t m, const float *x, const float *t, float *loss,
                     float *grad) {
  Mlp *n = mget(hm);
  if (!n || m < 1 || !x || !t) return G6_EARG;
  DevBuf dx, dt;
  G6_CHECK(hipMalloc(&dx.p, sizeof(float) * kMlpIn * m));
  G6_CHECK(hipMalloc(&dt.p, sizeof(float) * m));
  G6_CHECK(hipMemcpy(dx.p, x, sizeof(float) * kMlpIn * m, hipMemcpyHostToDevice));
  G6_CHECK(hipMemcpy(dt.p, t, sizeof(float) * m, hipMemcpyHostToDevice));
  const int rc = mlp_launch_grad(*n, m, (const float *)dx.p, (const float *)dt.p, nullptr,
                                 n->stream);
  if (rc != 0) return rc;
  G6_CHECK(hipGetLastError());
  std::vector<float> g(kMlpParams + 1);
  G6_CHECK(hipMemcpyAsync(g.data(), n->grad, sizeof(float) * (kMlpParams + 1),
                          hipMemcpyDeviceToHost, n->stream));
  G6_CHECK(hipStreamSynchronize(n->stream));
  if (grad) memcpy(grad, g.data(), sizeof(float) * kMlpParams);
  if (loss) *loss = g[kMlpParams];
  return 0;
}

int g6_mlp_train(g6_mlp hm, int64_t m, const float *x, const float *t, int epochs, int batch,
                 float lr, uint32_t seed, float *loss_per_epoch) {
  Mlp *n = mget(hm);
  if (!n || m < 1 || !x || !t || epochs < 1 || batch < 16 || !(lr > 0)) return G6_EARG;
  const int64_t nb = m / batch;  // full batches only
  DevBuf dx, dt, di[2], dl;
  G6_CHECK(hipMalloc(&dx.p, sizeof(float) * kMlpIn * m));
  G6_CHECK(hipMalloc(&dt.p, sizeof(float) * m));
  G6_CHECK(hipMalloc(&di[0].p, sizeof(int) * m));
  G6_CHECK(hipMalloc(&di[1].p, sizeof(int) * m));
  G6_CHECK(hipMalloc(&dl.p, sizeof(float) * (size_t)epochs * (nb > 0 ? nb : 1)));
  G6_CHECK(hipMemcpy(dx.p, x, sizeof(float) * kMlpIn * m, hipMemcpyHostToDevice));
  G6_CHECK(hipMemcpy(dt.p, t, sizeof(float) * m, hipMemcpyHostToDevice));
  McHost rng;
  rng.seed(seed);
  std::vector<int> perm(m);
  for (int64_t r = 0; r < m; r++) perm[r] = (int)r;
  // Fisher-Yates shuffle from the MT19937 stream, one per epoch in turn.
  // The draws do not depend on the array, so they are made first and the
  // swap loop prefetches the element it will swap 32 steps ahead (the array
  // is far larger than the caches).
  std::vector<uint32_t> ks(m > 1 ? m - 1 : 1);
  auto shuffle = [&]() {
    const int64_t cnt = m - 1;
    for (int64_t q = 0; q < cnt; q++) ks[q] = rng.next() % (uint32_t)(m - q);
    for (int64_t q = 0; q < cnt; q++) {
      if (q + 32 < cnt) __builtin_prefetch(&perm[ks[q + 32]], 1);
      std::swap(perm[m - 1 - q], perm[ks[q]]);
    }
  };
  shuffle();
  G6_CHECK(hipMemcpy(di[0].p, perm.data(), sizeof(int) * m, hipMemcpyHostToDevice));
  for (int e = 0; e < epochs; e++) {
    // the epoch's batches run back to back on the device; each batch loss
    // stays there (written by the Adam step).  Meanwhile the host shuffles
    // the next epoch's order into the other index buffer.
    const int *idx = (const int *)di[e & 1].p;
    for (int64_t b = 0; b < nb; b++) {
      const int rc = mlp_launch_train_step(*n, batch, (const float *)dx.p, (const float *)dt.p,
                                           idx + b * batch, lr, (float *)dl.p + e * nb + b,
                                           n->stream);
      if (rc != 0) return rc;
    }
    G6_CHECK(hipGetLastError());
    if (e + 1 < epochs) {
      shuffle();
      G6_CHECK(hipMemcpyAsync(di[(e + 1) & 1].p, perm.data(), sizeof(int) * m,
                              hipMemcpyHostToDevice, n->stream));
      G6_CHECK(hipStreamSynchronize(n->stream));  // perm is the next shuffle's input
    }
  }
  std::vector<float> lossb((size_t)epochs * (nb > 0 ? nb : 1), 0.0f);
  if (nb > 0)
    G6_CHECK(hipMemcpyAsync(lossb.data(), dl.p, sizeof(float) * epochs * nb,
                            hipMemcpyDeviceToHost, n->stream));
  G6_CHECK(hipStreamSynchronize(n->stream));
  const int64_t per = nb > 0 ? nb : 1;
  for (int e = 0; e < epochs; e++) {
    double s = 0.0;
    for (int64_t b = 0; b < per; b++) s += lossb[(size_t)e * per + b];
    if (loss_per_epoch) loss_per_epoch[e] = (float)(s / (double)per);
  }
  return 0;
}

}  // extern "C"
