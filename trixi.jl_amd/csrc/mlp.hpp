// mlp.hpp -- learned troubled-cell indicator for 2D p3 elements: a 19 -> 64 ->
// 64 -> 1 perceptron (ReLU, ReLU, sigmoid) whose output times alpha_max
// replaces the Hennemann-Gassner blending factor. There is no counterpart in
// Trixi.jl itself (its earlier neural-network indicators moved to a separate
// package, NEWS.md:495); the inputs are the HG indicator's own quantities
// (indicators_2d.jl:30-102): the 16 nodal values of density*pressure,
// normalised, and three functions of the modal energies.
//
// Matrix products run on MFMA: v_mfma_f32_16x16x16_bf16 (BF16 operands, FP32
// accumulation). One wavefront handles 16 elements at a time. Activations are
// kept transposed (hidden x element), so the accumulator layout of one layer,
//   lane l: column (element) l % 16, rows 4 (l / 16) + r, r = 0..3,
// is exactly the B-operand layout of the next layer,
//   lane l: column l % 16, k = 4 (l / 16) + j, j = 0..3,
// and the A operand (weights) is read as
//   lane l: row l % 16, k = 4 (l / 16) + j.
// Weight gradients sum over elements (K = element) after a transpose of the
// per-wave activations through LDS. Everything after the products (biases,
// ReLU, sigmoid, loss) is FP32; Adam keeps FP32 master weights and refreshes
// the BF16 copies.
#pragma once
#include <stdint.h>

#include "dg_common.hpp"

namespace trixi_hip {
namespace mlp {

constexpr int NF = 19;  // features: 16 normalised rho*p values + 3 energy features
constexpr int KF = 32;  // padded feature count (two K blocks of 16)
constexpr int NH = 64;  // hidden width
// FP32 parameter vector, in this order:
constexpr int OFF_W1 = 0;                  // W1[NH][NF] (hidden, feature)
constexpr int OFF_B1 = OFF_W1 + NH * NF;   // b1[NH]
constexpr int OFF_W2 = OFF_B1 + NH;        // W2[NH][NH] (out, in)
constexpr int OFF_B2 = OFF_W2 + NH * NH;   // b2[NH]
constexpr int OFF_W3 = OFF_B2 + NH;        // w3[NH]
constexpr int OFF_B3 = OFF_W3 + NH;        // b3
constexpr int NP = OFF_B3 + 1;             // 5505 parameters
constexpr int NPL = NP + 1;                // per-wave partial row: gradient + loss

constexpr int WAVES = 4;  // wavefronts per workgroup
constexpr int NTHR = 64 * WAVES;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round-to-nearest-even FP32 -> BF16 (finite inputs)
__device__ __forceinline__ uint16_t to_bf16(float x) {
  uint32_t b = __float_as_uint(x);
  b += 0x7fffu + ((b >> 16) & 1u);
  return (uint16_t)(b >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
  bf16x4 r;
  r[0] = (short)to_bf16(a);
  r[1] = (short)to_bf16(b);
  r[2] = (short)to_bf16(c);
  r[3] = (short)to_bf16(d);
  return r;
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// BF16 weight images in LDS: W1 [NH][KF] (features padded with zeros), W2
// [NH][NH] and its transpose.
struct WeightsLDS {
  uint16_t w1[NH * KF];
  uint16_t w2[NH * NH];
  uint16_t w2t[NH * NH];
  float b1[NH], b2[NH], w3[NH], b3;
};

__device__ __forceinline__ void load_weights(const float *__restrict__ w, WeightsLDS &s) {
  for (int i = threadIdx.x; i < NH * KF; i += blockDim.x) {
    const int h = i / KF, f = i - h * KF;
    s.w1[i] = f < NF ? to_bf16(w[OFF_W1 + h * NF + f]) : 0;
  }
  for (int i = threadIdx.x; i < NH * NH; i += blockDim.x) {
    const int o = i / NH, in = i - o * NH;
    const uint16_t v = to_bf16(w[OFF_W2 + i]);
    s.w2[i] = v;
    s.w2t[in * NH + o] = v;
  }
  for (int i = threadIdx.x; i < NH; i += blockDim.x) {
    s.b1[i] = w[OFF_B1 + i];
    s.b2[i] = w[OFF_B2 + i];
    s.w3[i] = w[OFF_W3 + i];
  }
  if (threadIdx.x == 0) s.b3 = w[OFF_B3];
}

__device__ __forceinline__ bf16x4 lds4(const uint16_t *p) {
  // four consecutive BF16 values (8-byte aligned by construction)
  const uint2 v = *reinterpret_cast<const uint2 *>(p);
  bf16x4 r;
  r[0] = (short)(v.x & 0xffffu);
  r[1] = (short)(v.x >> 16);
  r[2] = (short)(v.y & 0xffffu);
  r[3] = (short)(v.y >> 16);
  return r;
}

// Forward pass of one 16-element tile. x: [n][KF] FP32 features (zero
// padded). Leaves per lane: z1/z2 pre-activations (C layout), the BF16
// operands xb/h1b and the output y of element l % 16.
struct Tile {
  bf16x4 xb[2];
  f32x4 z1[4], z2[4];
  bf16x4 h1b[4];
  float y;
};

// The layers of the forward pass with X^T already in t.xb (B-operand layout:
// lane l holds element l % 16, features kb * 16 + 4 (l / 16) + j). W is any
// struct with the BF16 images w1 [NH][KF], w2 [NH][NH] and FP32 b1, b2, w3, b3.
template <typename W> __device__ __forceinline__ void forward_layers(const W &s, Tile &t) {
  const int l = threadIdx.x & 63;
  const int g = l >> 4;
  // layer 1: Z1^T = W1 X^T + b1
#pragma unroll
  for (int hb = 0; hb < 4; ++hb) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) acc = mfma(lds4(&s.w1[(hb * 16 + (l & 15)) * KF + kb * 16 + 4 * g]), t.xb[kb], acc);
    float h[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      acc[r] += s.b1[hb * 16 + 4 * g + r];
      h[r] = acc[r] > 0.f ? acc[r] : 0.f;
    }
    t.z1[hb] = acc;
    t.h1b[hb] = pack4(h[0], h[1], h[2], h[3]);
  }
  // layer 2: Z2^T = W2 H1^T + b2; layer 3 partial dot with w3
  float part = 0.f;
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc = mfma(lds4(&s.w2[(ob * 16 + (l & 15)) * NH + kb * 16 + 4 * g]), t.h1b[kb], acc);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int h = ob * 16 + 4 * g + r;
      acc[r] += s.b2[h];
      part += s.w3[h] * (acc[r] > 0.f ? acc[r] : 0.f);
    }
    t.z2[ob] = acc;
  }
  // sum over the four lanes holding the same element
  part += __shfl_xor(part, 16);
  part += __shfl_xor(part, 32);
  const float o = part + s.b3;
  t.y = 1.f / (1.f + __expf(-o));
}

// The operands one lane needs for forward_layers, held in registers: the A
// fragments of W1 (hb, kb) and W2 (ob, kb), the biases and output weights of
// the rows the lane's accumulators hold (h = 16 hb + 4 (l / 16) + r). Loading
// them all at kernel entry makes one memory round trip of the whole network
// instead of one per layer.
struct WeightRegs {
  bf16x4 w1[4][2];
  bf16x4 w2[4][4];
  float b1[4][4], b2[4][4], w3[4][4];
  float b3;
};

template <typename W> __device__ __forceinline__ void load_weight_regs(const W &s, WeightRegs &r) {
  const int l = threadIdx.x & 63;
  const int g = l >> 4;
#pragma unroll
  for (int hb = 0; hb < 4; ++hb) {
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) r.w1[hb][kb] = lds4(&s.w1[(hb * 16 + (l & 15)) * KF + kb * 16 + 4 * g]);
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) r.w2[hb][kb] = lds4(&s.w2[(hb * 16 + (l & 15)) * NH + kb * 16 + 4 * g]);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      r.b1[hb][q] = s.b1[hb * 16 + 4 * g + q];
      r.b2[hb][q] = s.b2[hb * 16 + 4 * g + q];
      r.w3[hb][q] = s.w3[hb * 16 + 4 * g + q];
    }
  }
  r.b3 = s.b3;
}

// forward_layers with the operands from registers: the same products, bias
// additions and activations in the same order, so the same result bit for bit.
__device__ __forceinline__ void forward_layers_regs(const WeightRegs &w, Tile &t) {
#pragma unroll
  for (int hb = 0; hb < 4; ++hb) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) acc = mfma(w.w1[hb][kb], t.xb[kb], acc);
    float h[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      acc[r] += w.b1[hb][r];
      h[r] = acc[r] > 0.f ? acc[r] : 0.f;
    }
    t.z1[hb] = acc;
    t.h1b[hb] = pack4(h[0], h[1], h[2], h[3]);
  }
  float part = 0.f;
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc = mfma(w.w2[ob][kb], t.h1b[kb], acc);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      acc[r] += w.b2[ob][r];
      part += w.w3[ob][r] * (acc[r] > 0.f ? acc[r] : 0.f);
    }
    t.z2[ob] = acc;
  }
  part += __shfl_xor(part, 16);
  part += __shfl_xor(part, 32);
  const float o = part + w.b3;
  t.y = 1.f / (1.f + __expf(-o));
}

__device__ __forceinline__ void forward_tile(const float *__restrict__ x, int n, int tile, const WeightsLDS &s,
                                             Tile &t) {
  const int l = threadIdx.x & 63;
  const int e = tile * 16 + (l & 15);
  const int g = l >> 4;
  // X^T as the B operand: element e, features kb*16 + 4g + j
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) {
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (e < n) v = *reinterpret_cast<const float4 *>(x + (size_t)e * KF + kb * 16 + 4 * g);
    t.xb[kb] = pack4(v.x, v.y, v.z, v.w);
  }
  forward_layers(s, t);
}

// The network as inference reads it: BF16 images of W1 (features padded with
// zeros) and W2, FP32 biases and output weights. Kept in device memory
// (refreshed whenever the parameters change) and copied into LDS with
// 16-byte loads, so that a launch converts nothing.
struct InferImage {
  uint16_t w1[NH * KF];
  uint16_t w2[NH * NH];
  float b1[NH], b2[NH], w3[NH], b3, pad[3];
};
static_assert(sizeof(InferImage) % 16 == 0, "16-byte copies");

// FP32 parameters -> InferImage (one workgroup).
__global__ __launch_bounds__(256) void mlp_pack_kernel(const float *__restrict__ w, InferImage *__restrict__ img) {
  for (int i = threadIdx.x; i < NH * KF; i += blockDim.x) {
    const int h = i / KF, f = i - h * KF;
    img->w1[i] = f < NF ? to_bf16(w[OFF_W1 + h * NF + f]) : 0;
  }
  for (int i = threadIdx.x; i < NH * NH; i += blockDim.x) img->w2[i] = to_bf16(w[OFF_W2 + i]);
  for (int i = threadIdx.x; i < NH; i += blockDim.x) {
    img->b1[i] = w[OFF_B1 + i];
    img->b2[i] = w[OFF_B2 + i];
    img->w3[i] = w[OFF_W3 + i];
  }
  if (threadIdx.x == 0) {
    img->b3 = w[OFF_B3];
    img->pad[0] = img->pad[1] = img->pad[2] = 0.f;
  }
}

// The clipping of calc_indicator_hennemann_gassner! (indicators_2d.jl:86-99)
// applied to a blending factor: below alpha_min it is 0 (pure DG, the element
// kernel then skips the FV subcells), above 1 - alpha_min 1 (pure FV), and it
// never exceeds alpha_max.
__device__ __forceinline__ double hg_clip(double al, double alpha_min, double alpha_max) {
  if (al < alpha_min) al = 0.0;
  if (al > 1 - alpha_min) al = 1.0;
  return al < alpha_max ? al : alpha_max;
}

// Inference over E elements: alpha_raw[e] = hg_clip(alpha_max * y(features[e])),
// or the HG value (alpha_hg, clipped by the HG kernel) when a feature lies
// outside [lo, hi] or y is not finite (fallback != 0); with fallback == 0 the
// network output is always used.
__global__ __launch_bounds__(NTHR) void mlp_infer_kernel(const float *__restrict__ w, const float *__restrict__ x,
                                                         const float *__restrict__ lo, const float *__restrict__ hi,
                                                         const double *__restrict__ alpha_hg, double alpha_max,
                                                         double alpha_min, int fallback, int E,
                                                         double *__restrict__ alpha_raw, float *__restrict__ y_out) {
  __shared__ __attribute__((aligned(16))) WeightsLDS s;
  load_weights(w, s);
  __syncthreads();
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int ntiles = (E + 15) / 16;
  for (int tile = blockIdx.x * WAVES + wave; tile < ntiles; tile += gridDim.x * WAVES) {
    Tile t;
    forward_tile(x, E, tile, s, t);
    const int e = tile * 16 + l;
    if (l < 16 && e < E) {
      if (y_out) y_out[e] = t.y;
      if (alpha_raw) {
        bool use_nn = true;
        if (fallback) {
          use_nn = isfinite(t.y);
          for (int f = 0; f < NF; ++f) {
            const float v = x[(size_t)e * KF + f];
            use_nn = use_nn && v >= lo[f] && v <= hi[f];
          }
        }
        alpha_raw[e] = use_nn ? hg_clip(alpha_max * (double)t.y, alpha_min, alpha_max) : alpha_hg[e];
      }
    }
  }
}

// Loss L = sum_e (y_e - target_e)^2 / n and its gradient. Each wavefront
// handles tiles w, w + W, ... and writes one partial row [NPL] (gradient,
// then loss) to partial[global wave]; mlp_reduce_kernel sums the rows.
__global__ __launch_bounds__(NTHR) void mlp_grad_kernel(const float *__restrict__ w, const float *__restrict__ x,
                                                        const float *__restrict__ target, int n,
                                                        float *__restrict__ partial) {
  __shared__ __attribute__((aligned(16))) WeightsLDS s;
  // per-wave transposes [hidden or feature][element], BF16
  __shared__ __attribute__((aligned(16))) uint16_t s_x[WAVES][KF * 16];
  __shared__ __attribute__((aligned(16))) uint16_t s_h1[WAVES][NH * 16];
  __shared__ __attribute__((aligned(16))) uint16_t s_d2[WAVES][NH * 16];
  __shared__ __attribute__((aligned(16))) uint16_t s_d1[WAVES][NH * 16];
  load_weights(w, s);
  __syncthreads();
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int g = l >> 4, c = l & 15;
  const int gw = blockIdx.x * WAVES + wave, nw = gridDim.x * WAVES;
  const int ntiles = (n + 15) / 16;
  const float inv_n = 1.f / (float)n;

  f32x4 gW1[4][2], gW2[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int b = 0; b < 2; ++b) gW1[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < 4; ++b) gW2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
  float gb1[16], gb2[16], gw3[16];  // per lane: hidden blk*16 + 4g + r (blk = idx / 4, r = idx % 4)
#pragma unroll
  for (int i = 0; i < 16; ++i) gb1[i] = gb2[i] = gw3[i] = 0.f;
  float gb3 = 0.f, loss = 0.f;

  for (int tile = gw; tile < ntiles; tile += nw) {
    Tile t;
    forward_tile(x, n, tile, s, t);
    const int e = tile * 16 + c;
    const bool ok = e < n;
    const float tgt = ok ? target[e] : 0.f;
    const float res = ok ? t.y - tgt : 0.f;
    if (g == 0) {
      loss += res * res * inv_n;
      gb3 += 2.f * res * t.y * (1.f - t.y) * inv_n;
    }
    const float dout = 2.f * res * t.y * (1.f - t.y) * inv_n;  // dL/do for element c
    // layer 3 and layer 2 backward (C layout)
    f32x4 d2[4];
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int h = ob * 16 + 4 * g + r;
        const float z = t.z2[ob][r];
        gw3[ob * 4 + r] += dout * (z > 0.f ? z : 0.f);
        const float dz = z > 0.f ? s.w3[h] * dout : 0.f;
        d2[ob][r] = dz;
        gb2[ob * 4 + r] += dz;
      }
    }
    bf16x4 d2b[4];
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) d2b[ob] = pack4(d2[ob][0], d2[ob][1], d2[ob][2], d2[ob][3]);
    // dH1^T = W2^T dZ2^T, then through the first ReLU
    f32x4 d1[4];
#pragma unroll
    for (int ib = 0; ib < 4; ++ib) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ob = 0; ob < 4; ++ob) acc = mfma(lds4(&s.w2t[(ib * 16 + c) * NH + ob * 16 + 4 * g]), d2b[ob], acc);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float dz = t.z1[ib][r] > 0.f ? acc[r] : 0.f;
        d1[ib][r] = dz;
        gb1[ib * 4 + r] += dz;
      }
    }
    // transposes: [hidden][element] images of X, H1, dZ2, dZ1 (BF16)
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
      for (int j = 0; j < 4; ++j) s_x[wave][(kb * 16 + 4 * g + j) * 16 + c] = (uint16_t)t.xb[kb][j];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = (b * 16 + 4 * g + j) * 16 + c;
        s_h1[wave][row] = (uint16_t)t.h1b[b][j];
        s_d2[wave][row] = (uint16_t)d2b[b][j];
        s_d1[wave][row] = to_bf16(d1[b][j]);
      }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // dW2 (out x in) += dZ2^T H1: A = dZ2^T rows, B = H1 columns, K = element
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
      const bf16x4 a = lds4(&s_d2[wave][(ob * 16 + c) * 16 + 4 * g]);
#pragma unroll
      for (int ib = 0; ib < 4; ++ib) gW2[ob][ib] = mfma(a, lds4(&s_h1[wave][(ib * 16 + c) * 16 + 4 * g]), gW2[ob][ib]);
    }
    // dW1 (hidden x feature) += dZ1^T X
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) {
      const bf16x4 a = lds4(&s_d1[wave][(hb * 16 + c) * 16 + 4 * g]);
#pragma unroll
      for (int fb = 0; fb < 2; ++fb) gW1[hb][fb] = mfma(a, lds4(&s_x[wave][(fb * 16 + c) * 16 + 4 * g]), gW1[hb][fb]);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }

  // per-lane vectors summed over the 16 lanes of a row group (elements)
#pragma unroll
  for (int i = 0; i < 16; ++i) {
#pragma unroll
    for (int m = 1; m < 16; m <<= 1) {
      gb1[i] += __shfl_xor(gb1[i], m);
      gb2[i] += __shfl_xor(gb2[i], m);
      gw3[i] += __shfl_xor(gw3[i], m);
    }
  }
#pragma unroll
  for (int m = 1; m < 16; m <<= 1) {
    gb3 += __shfl_xor(gb3, m);
    loss += __shfl_xor(loss, m);
  }
  float *row = partial + (size_t)gw * NPL;
  // accumulator layout: row 4g + r, column c of each 16 x 16 tile
#pragma unroll
  for (int hb = 0; hb < 4; ++hb)
#pragma unroll
    for (int fb = 0; fb < 2; ++fb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int h = hb * 16 + 4 * g + r, f = fb * 16 + c;
        if (f < NF) row[OFF_W1 + h * NF + f] = gW1[hb][fb][r];
      }
#pragma unroll
  for (int ob = 0; ob < 4; ++ob)
#pragma unroll
    for (int ib = 0; ib < 4; ++ib)
#pragma unroll
      for (int r = 0; r < 4; ++r) row[OFF_W2 + (ob * 16 + 4 * g + r) * NH + ib * 16 + c] = gW2[ob][ib][r];
  if (c == 0) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int h = (i / 4) * 16 + 4 * g + (i % 4);
      row[OFF_B1 + h] = gb1[i];
      row[OFF_B2 + h] = gb2[i];
      row[OFF_W3 + h] = gw3[i];
    }
  }
  if (l == 0) {
    row[OFF_B3] = gb3;
    row[NP] = loss;
  }
}

// out[p] = sum over rows of partial[row][p], p < NPL, rows in order.
__global__ __launch_bounds__(256) void mlp_reduce_kernel(const float *__restrict__ partial, int rows,
                                                         float *__restrict__ out) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= NPL) return;
  float acc = 0.f;
  for (int r = 0; r < rows; ++r) acc += partial[(size_t)r * NPL + p];
  out[p] = acc;
}

// Adam (Kingma & Ba 2015) on the FP32 master weights; step = 1, 2, ...
__global__ __launch_bounds__(256) void mlp_adam_kernel(float *__restrict__ w, float *__restrict__ m,
                                                       float *__restrict__ v, const float *__restrict__ grad, float lr,
                                                       float beta1, float beta2, float eps, int step) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= NP) return;
  const float gr = grad[p];
  const float mi = beta1 * m[p] + (1.f - beta1) * gr;
  const float vi = beta2 * v[p] + (1.f - beta2) * gr * gr;
  m[p] = mi;
  v[p] = vi;
  const float mh = mi / (1.f - powf(beta1, (float)step));
  const float vh = vi / (1.f - powf(beta2, (float)step));
  w[p] -= lr * mh / (sqrtf(vh) + eps);
}

}  // namespace mlp
}  // namespace trixi_hip
