// indicator2d.hpp -- the shock-capturing indicator of 2D polydeg-3 elements
// in one launch: IndicatorHennemannGassner's blending factor and, with FEAT,
// the learned indicator's features, the BF16-MFMA network and its clipped
// output, with the HG factor as fallback.
//
//   calc_indicator_hennemann_gassner!  indicators_2d.jl:30-102 (indicators.jl:119-156)
//   multiply_scalar_dimensionwise!     interpolation.jl:207-235
//   density_pressure / density / pressure  compressible_euler_2d.jl:2205-2210
//
// Work decomposition: one wavefront of 16 elements per workgroup. Lane l owns
// element l % 16 and the node row j = l / 16 (nodes i + 4 j, i = 0..3), so
// the modal transform along x runs in registers, the transform along y reads
// the other rows of the element through LDS, and the modal energies, mean
// and deviation are sums over the four lanes of the element (xor shuffles).
// The lane's four normalised nodal values are exactly its part of the
// network's B operand (features 4 j .. 4 j + 3), and the lane with j = 0 holds
// the three energy features: the features never leave registers. Each lane
// loads its operands of the network (MFMA fragments, biases, the training-box
// bounds) from the device image at kernel entry, next to the state loads, so
// the kernel makes one memory round trip and the workgroup has no barrier.
// Arithmetic up to the features is FP64 and in the operation order of
// hg_indicator_kernel (the same factors).
#pragma once
#include "dg_common.hpp"
#include "mlp.hpp"

namespace trixi_hip {

struct Vinv4 {
  double v[16];  // inverse_vandermonde_legendre, row-major [mode][node]
};

// Everything the 2D p3 indicator needs besides the state, shared by the
// stand-alone kernel and the epilogue of the fused stage. FEAT = false:
// alpha_raw[e] = the HG factor. FEAT = true: the features as well (into
// feat_out [E][32] when not null), and with net != 0 alpha_raw[e] = the
// network's clipped factor, or the HG factor where an input leaves [lo, hi]
// or the output is not finite (fallback != 0); with net == 0 alpha_raw
// receives the HG factor. With FEAT, img, lo and hi must point to device
// memory of their sizes even when net == 0 (they are read unconditionally, so
// that no branch sits between the loads and their use). alpha_hg_out (may be
// null) receives the HG factor in every case.
struct Ind2DArgs {
  HGParams hp;
  Vinv4 vi;
  const mlp::InferImage *img;
  const float *lo, *hi;
  int net, fallback;
  double *alpha_raw;
  double *alpha_hg_out;
  float *feat_out;
};

// One wavefront computes the indicator of the 16 elements e0 .. e0 + 15 (those
// below E): lane l owns element l % 16 and node row j = l / 16; un holds the
// lane's 16 state values (nodes 4 j .. 4 j + 3, four variables each). s_mx,
// s_sq: 256 doubles each of LDS owned by the wavefront. PRELOAD: the network
// operands are loaded into registers at entry (the stand-alone kernel, one
// wavefront per SIMD, where latency is exposed), from simg (the workgroup's
// LDS copy of the image) with SHARED_IMG, else from ia.img; otherwise they are
// read where the products use them (inside the element kernel, whose register
// budget they would otherwise set). All give the same result bit for bit.
template <typename T, bool FEAT, bool PRELOAD = true, bool SHARED_IMG = false>
__device__ __forceinline__ void indicator_2d_p3_wave(const T *un, int e0, int E, const Ind2DArgs &ia, double *s_mx,
                                                     double *s_sq, const mlp::InferImage *simg = nullptr) {
  constexpr int n = 4, V = 4;
  const HGParams &hp = ia.hp;
  const Vinv4 &vi = ia.vi;
  const mlp::InferImage *__restrict__ img = ia.img;
  const float *__restrict__ lo = ia.lo;
  const float *__restrict__ hi = ia.hi;
  const int net = ia.net, fallback = ia.fallback;
  double *__restrict__ alpha_raw = ia.alpha_raw;
  double *__restrict__ alpha_hg_out = ia.alpha_hg_out;
  float *__restrict__ feat_out = ia.feat_out;
  const int l = threadIdx.x & 63;
  const int el = l & 15, j = l >> 4;
  const int e = e0 + el;
  const bool valid = e < E;
  // with the network: its operands and the lane's training-box bounds, issued
  // behind the state loads, so that their latency overlaps the modal
  // transform (one memory round trip for the whole kernel)
  mlp::WeightRegs wr;
  float lo_r[7], hi_r[7];
  if constexpr (FEAT && PRELOAD) {
    if constexpr (SHARED_IMG) mlp::load_weight_regs(*simg, wr);
    else mlp::load_weight_regs(*img, wr);
#pragma unroll
    for (int i = 0; i < n; ++i) {
      lo_r[i] = lo[4 * j + i];
      hi_r[i] = hi[4 * j + i];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      lo_r[4 + i] = lo[16 + i];
      hi_r[4 + i] = hi[16 + i];
    }
  }
  double q[n];
#pragma unroll
  for (int i = 0; i < n; ++i) {
    const double rho = un[i * V], m2 = double(un[i * V + 1]) * double(un[i * V + 1]) +
                                       double(un[i * V + 2]) * double(un[i * V + 2]);
    q[i] = hp.variable == 1   ? rho
           : hp.variable == 2 ? (hp.gamma - 1) * (double(un[i * V + 3]) - 0.5 * m2 / rho)
                              : (hp.gamma - 1) * (rho * double(un[i * V + 3]) - 0.5 * m2);
  }
  // modal transform along x (own row), then along y (all rows of the element)
  double *mx = &s_mx[el * 16];
#pragma unroll
  for (int a = 0; a < n; ++a) {
    double r = 0;
#pragma unroll
    for (int i = 0; i < n; ++i) r += vi.v[a * n + i] * q[i];
    mx[4 * j + a] = r;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  // row j of the inverse Vandermonde matrix, selected from registers: a
  // lane-indexed read of the kernel argument would be a memory load (one more
  // round trip); the empty asm keeps the selects from being folded back
  double vj[n];
#pragma unroll
  for (int b = 0; b < n; ++b) {
    double x = 0.0;
#pragma unroll
    for (int jj = 0; jj < n; ++jj) {
      double c = vi.v[jj * n + b];
      asm volatile("" : "+v"(c));
      x = j == jj ? c : x;
    }
    vj[b] = x;
  }
  double m[n];  // modes (a, j)
#pragma unroll
  for (int a = 0; a < n; ++a) {
    double r = 0;
#pragma unroll
    for (int b = 0; b < n; ++b) r += vj[b] * mx[4 * b + a];
    m[a] = r;
  }
  // squared modes (a, j) of the element to LDS; every lane of the element
  // then sums them in the loop order of indicators_2d.jl:59-75 (te2 over
  // j < 2, i < 2; te1 adds (i < 3, j = 2), then (i = 2, j < 2); te adds
  // (i, j = 3), then (i = 3, j < 3)), so that the energies and the factor are
  // those of hg_indicator_kernel bit for bit
  double *sq = &s_sq[el * 16];
#pragma unroll
  for (int a = 0; a < n; ++a) sq[a + 4 * j] = m[a] * m[a];
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  double te2 = 0;
#pragma unroll
  for (int jj = 0; jj < 2; ++jj)
#pragma unroll
    for (int i = 0; i < 2; ++i) te2 += sq[i + 4 * jj];
  double te1 = te2;
#pragma unroll
  for (int i = 0; i < 3; ++i) te1 += sq[i + 4 * 2];
#pragma unroll
  for (int jj = 0; jj < 2; ++jj) te1 += sq[2 + 4 * jj];
  double te = te1;
#pragma unroll
  for (int i = 0; i < 4; ++i) te += sq[i + 4 * 3];
#pragma unroll
  for (int jj = 0; jj < 3; ++jj) te += sq[3 + 4 * jj];
  const double ef1 = te != 0 ? (te - te1) / te : 0.0;
  const double ef2 = te1 != 0 ? (te1 - te2) / te1 : 0.0;
  const double energy = ef1 > ef2 ? ef1 : ef2;
  auto hg_alpha = [&]() {
    return mlp::hg_clip(1.0 / (1.0 + exp(-hp.parameter_s / hp.threshold * (energy - hp.threshold))), hp.alpha_min,
                        hp.alpha_max);
  };
  if constexpr (!FEAT) {
    const double a_hg = hg_alpha();
    if (valid && j == 0) {
      alpha_raw[e] = a_hg;
      if (alpha_hg_out) alpha_hg_out[e] = a_hg;
    }
  } else {
    // features: nodal values centred by the element mean and divided by the
    // largest deviation (+ 1e-7 |mean|); the HG logits of the two energy
    // ratios / 8 in [-2, 2]; log10 of the smoothest-modes share / 8 in [-2, 0]
    double s = q[0] + q[1] + q[2] + q[3];
    s += __shfl_xor(s, 16);
    s += __shfl_xor(s, 32);
    const double mean = s / 16;
    double dev = 0;
#pragma unroll
    for (int i = 0; i < n; ++i) dev = fmax(dev, fabs(q[i] - mean));
    dev = fmax(dev, __shfl_xor(dev, 16));
    dev = fmax(dev, __shfl_xor(dev, 32));
    const double scale = dev + 1e-7 * fabs(mean) + 1e-300;
    float f[8];
#pragma unroll
    for (int i = 0; i < n; ++i) f[i] = float((q[i] - mean) / scale);
    const double k = hp.parameter_s / hp.threshold;
    const double smooth = te > 0 ? te2 / te : 1.0;
    f[4] = float(fmin(2.0, fmax(-2.0, k * (ef1 - hp.threshold) / 8)));
    f[5] = float(fmin(2.0, fmax(-2.0, k * (ef2 - hp.threshold) / 8)));
    f[6] = fminf(0.f, fmaxf(-2.f, log10f(float(smooth) + 1e-16f) / 8));
    f[7] = 0.f;
    double a = 0.0;
    bool need_hg = true;  // the HG factor is the result (no network) or the fallback
    if (net) {
      mlp::Tile t;
      t.xb[0] = mlp::pack4(f[0], f[1], f[2], f[3]);
      t.xb[1] = j == 0 ? mlp::pack4(f[4], f[5], f[6], 0.f) : mlp::pack4(0.f, 0.f, 0.f, 0.f);
      if constexpr (PRELOAD) {
        mlp::forward_layers_regs(wr, t);
      } else {
        mlp::forward_layers(*img, t);
#pragma unroll
        for (int i = 0; i < n; ++i) {
          lo_r[i] = lo[4 * j + i];
          hi_r[i] = hi[4 * j + i];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          lo_r[4 + i] = lo[16 + i];
          hi_r[4 + i] = hi[16 + i];
        }
      }
      // the training box: every feature of the element inside [lo, hi]
      bool in_box = true;
#pragma unroll
      for (int i = 0; i < n; ++i) in_box = in_box && f[i] >= lo_r[i] && f[i] <= hi_r[i];
      if (j == 0)
#pragma unroll
        for (int i = 0; i < 3; ++i) in_box = in_box && f[4 + i] >= lo_r[4 + i] && f[4 + i] <= hi_r[4 + i];
      const unsigned long long out = __builtin_amdgcn_ballot_w64(!in_box);
      const bool element_in_box = ((out >> el) & 0x0001000100010001ull) == 0;
      const bool use_nn = !fallback || (element_in_box && isfinite(t.y));
      if (use_nn) a = mlp::hg_clip(hp.alpha_max * (double)t.y, hp.alpha_min, hp.alpha_max);
      need_hg = !use_nn;
    }
    const double a_hg = (need_hg || alpha_hg_out) ? hg_alpha() : 0.0;
    if (need_hg) a = a_hg;
    if (valid) {
      if (j == 0) {
        if (alpha_raw) alpha_raw[e] = a;
        if (alpha_hg_out) alpha_hg_out[e] = a_hg;
      }
      if (feat_out) {
        float *fo = feat_out + (size_t)e * kMlpFeatureStride;
        *reinterpret_cast<float4 *>(fo + 4 * j) = make_float4(f[0], f[1], f[2], f[3]);
        *reinterpret_cast<float4 *>(fo + 16 + 4 * j) =
            j == 0 ? make_float4(f[4], f[5], f[6], 0.f) : make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
  }
}

// Stand-alone launch: one wavefront per 16 elements. The HG factor alone
// (FEAT = false) runs one wavefront per workgroup. With the features and the
// network, a workgroup is kIndWaves wavefronts that share one LDS copy of the
// network image, read from global memory once per workgroup instead of once
// per wavefront (the image is 13 KB; at 128^2 elements that is 3.3 instead of
// 13 MB of L2 reads per launch).
constexpr int kIndWaves = 4;
template <typename T, bool FEAT>
__global__ __launch_bounds__(FEAT ? 64 * kIndWaves : 64) void indicator_2d_p3_kernel(const T *__restrict__ u, int E,
                                                                                    Ind2DArgs ia) {
  constexpr int V = 4, W = FEAT ? kIndWaves : 1;
  __shared__ double s_mx[W][16 * 16];  // x-transformed rows [element][row j][mode i]
  __shared__ double s_sq[W][16 * 16];  // squared modes [element][a + 4 j]
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int el = l & 15, j = l >> 4;
  const int e0 = (blockIdx.x * W + w) * 16;
  const int e = e0 + el;
  const int ec = e < E ? e : E - 1;  // lanes past the end compute on the last element, store nothing
  // the lane's four nodes (i + 4 j): 16 contiguous values, loaded at once in
  // 16-byte pieces before the (wave-uniform) choice of the indicator variable
  constexpr int NCH = 4 * V * (int)sizeof(T) / 16;
  float4 raw[NCH];
  const float4 *src = reinterpret_cast<const float4 *>(u + ((size_t)ec * 16 + 4 * j) * V);
#pragma unroll
  for (int c = 0; c < NCH; ++c) raw[c] = src[c];
  if constexpr (FEAT) {
    __shared__ __attribute__((aligned(16))) mlp::InferImage s_img;
    if (ia.net) {  // uniform: copy the image in 16-byte pieces
      constexpr int NC = (int)(sizeof(mlp::InferImage) / 16);
      const float4 *gi = reinterpret_cast<const float4 *>(ia.img);
      float4 *si = reinterpret_cast<float4 *>(&s_img);
      for (int c = threadIdx.x; c < NC; c += 64 * W) si[c] = gi[c];
    }
    __syncthreads();
    indicator_2d_p3_wave<T, FEAT, true, true>(reinterpret_cast<const T *>(raw), e0, E, ia, s_mx[w], s_sq[w], &s_img);
  } else {
    indicator_2d_p3_wave<T, FEAT>(reinterpret_cast<const T *>(raw), e0, E, ia, s_mx[w], s_sq[w]);
  }
}

}  // namespace trixi_hip
