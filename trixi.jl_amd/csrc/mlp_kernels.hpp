// mlp_kernels.hpp -- the launches of the learned troubled-cell indicator's
// network (mlp.hpp): parameter packing, inference over a feature matrix,
// loss gradient, the fixed-order reduction of per-wave gradient rows and
// Adam. Included by the one translation unit that launches them (the element
// kernels include only mlp.hpp's device functions).
#pragma once
#include "mlp.hpp"

namespace trixi_hip {
namespace mlp {

// FP32 parameters -> InferImage (one workgroup).
__global__ __launch_bounds__(256) void mlp_pack_kernel(const float *__restrict__ w, InferImage *__restrict__ img) {
  for (int i = threadIdx.x; i < NH; i += blockDim.x) {
    img->b1[i] = w[OFF_B1 + i];
    img->b2[i] = w[OFF_B2 + i];
    img->w3[i] = w[OFF_W3 + i];
  }
  if (threadIdx.x == 0) {
    img->b3 = w[OFF_B3];
    img->pad[0] = img->pad[1] = img->pad[2] = 0.f;
  }
  // W1 and W2 in fragment order: [block][kb][lane][4]
  for (int i = threadIdx.x; i < 4 * 2 * 64 * 4; i += blockDim.x) {
    const int jj = i & 3, l = (i >> 2) & 63, kb = (i >> 8) & 1, hb = i >> 9;
    const int h = hb * 16 + (l & 15), f = kb * 16 + 4 * (l >> 4) + jj;
    img->w1f[i] = f < NF ? to_bf16(w[OFF_W1 + h * NF + f]) : 0;
  }
  for (int i = threadIdx.x; i < 4 * 4 * 64 * 4; i += blockDim.x) {
    const int jj = i & 3, l = (i >> 2) & 63, kb = (i >> 8) & 3, ob = i >> 10;
    const int o = ob * 16 + (l & 15), k = kb * 16 + 4 * (l >> 4) + jj;
    img->w2f[i] = to_bf16(w[OFF_W2 + o * NH + k]);
  }
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

// out[p] = sum over rows of partial[row][p], p < NPL, in a fixed order: a
// workgroup takes RED_COLS consecutive columns; its RED_SPLIT row groups sum
// rows g, g + RED_SPLIT, ... in row order (each load instruction of a
// wavefront reads two rows of 32 consecutive columns), then the group sums are
// added in group order. Deterministic, and RED_SPLIT independent load streams
// per column instead of one serial chain over all rows.
constexpr int RED_COLS = 32, RED_SPLIT = 8;
__global__ __launch_bounds__(RED_COLS *RED_SPLIT) void mlp_reduce_kernel(const float *__restrict__ partial, int rows,
                                                                         float *__restrict__ out) {
  __shared__ float s_part[RED_SPLIT][RED_COLS];
  const int c = threadIdx.x % RED_COLS, g = threadIdx.x / RED_COLS;
  const int p = blockIdx.x * RED_COLS + c;
  float acc = 0.f;
  if (p < NPL)
    for (int r = g; r < rows; r += RED_SPLIT) acc += partial[(size_t)r * NPL + p];
  s_part[g][c] = acc;
  __syncthreads();
  if (g == 0 && p < NPL) {
    float t = s_part[0][c];
#pragma unroll
    for (int k = 1; k < RED_SPLIT; ++k) t += s_part[k][c];
    out[p] = t;
  }
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
