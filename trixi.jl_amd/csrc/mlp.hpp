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

// The A fragment of W1 for (hb, kb) and of W2 for (ob, kb) held by this lane
// (row 16 block + l % 16, k = 16 kb + 4 (l / 16) + 0..3), from the row-major
// LDS images; the inference image's overloads follow its definition.
__device__ __forceinline__ bf16x4 frag_w1(const WeightsLDS &s, int hb, int kb) {
  const int l = threadIdx.x & 63;
  return lds4(&s.w1[(hb * 16 + (l & 15)) * KF + kb * 16 + 4 * (l >> 4)]);
}
__device__ __forceinline__ bf16x4 frag_w2(const WeightsLDS &s, int ob, int kb) {
  const int l = threadIdx.x & 63;
  return lds4(&s.w2[(ob * 16 + (l & 15)) * NH + kb * 16 + 4 * (l >> 4)]);
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
    for (int kb = 0; kb < 2; ++kb) acc = mfma(frag_w1(s, hb, kb), t.xb[kb], acc);
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
    for (int kb = 0; kb < 4; ++kb) acc = mfma(frag_w2(s, ob, kb), t.h1b[kb], acc);
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
// instead of one per layer (load_weight_regs, after InferImage).
struct WeightRegs {
  bf16x4 w1[4][2];
  bf16x4 w2[4][4];
  float b1[4][4], b2[4][4], w3[4][4];
  float b3;
};

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
// zeros) and W2 in MFMA fragment order, [block][kb][lane][4] (lane l: row
// 16 block + l % 16, k = 16 kb + 4 (l / 16) + 0..3), so that the 64 lanes of a
// wavefront read a fragment as one contiguous 512-byte block; FP32 biases and
// output weights. Kept in device memory, refreshed whenever the parameters
// change, so that a launch converts nothing.
struct InferImage {
  uint16_t w1f[4 * 2 * 64 * 4];
  uint16_t w2f[4 * 4 * 64 * 4];
  float b1[NH], b2[NH], w3[NH], b3, pad[3];
};
static_assert(sizeof(InferImage) % 16 == 0, "16-byte copies");

__device__ __forceinline__ bf16x4 frag_w1(const InferImage &s, int hb, int kb) {
  return lds4(&s.w1f[((hb * 2 + kb) * 64 + (threadIdx.x & 63)) * 4]);
}
__device__ __forceinline__ bf16x4 frag_w2(const InferImage &s, int ob, int kb) {
  return lds4(&s.w2f[((ob * 4 + kb) * 64 + (threadIdx.x & 63)) * 4]);
}

__device__ __forceinline__ void load_weight_regs(const InferImage &s, WeightRegs &r) {
  const int l = threadIdx.x & 63;
  const int g = l >> 4;
#pragma unroll
  for (int hb = 0; hb < 4; ++hb) {
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) r.w1[hb][kb] = frag_w1(s, hb, kb);
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) r.w2[hb][kb] = frag_w2(s, hb, kb);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      r.b1[hb][q] = s.b1[hb * 16 + 4 * g + q];
      r.b2[hb][q] = s.b2[hb * 16 + 4 * g + q];
      r.w3[hb][q] = s.w3[hb * 16 + 4 * g + q];
    }
  }
  r.b3 = s.b3;
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

}  // namespace mlp
}  // namespace trixi_hip
