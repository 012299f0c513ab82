// dg_common.hpp -- device-side physics for the DGSEM compressible-Euler RHS.
//
// Pointwise functions of Trixi.jl restated for CDNA4 (gfx950), templated on
// the arithmetic type T (float for the FP32 bulk path, double for the plain
// FP64 path) and the number of space dimensions D (V = D + 2 variables).
//
//   cons2prim              src/equations/compressible_euler_3d.jl:1783-1794 (2D :2038-2046)
//   flux                   compressible_euler_3d.jl:420-447 (2D :415-434)
//   flux_ranocha           compressible_euler_3d.jl:746-793 (2D :686-722), in the
//                          per-node-logarithm form of the reference's own
//                          flux_ranocha_turbo kernel (dg_3d_compressible_euler.jl:265-420)
//   ln_mean / inv_ln_mean  src/auxiliary/math.jl:196-263
//   flux_lax_friedrichs    src/equations/numerical_fluxes.jl:32-45, 166-178, 229-253
//   max_abs_speed(_naive)  compressible_euler_3d.jl:1112-1177 (2D :1423-1482)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace trixi_hip {

enum VolumeKind { VOL_WEAK = 0, VOL_FD_RANOCHA = 1, VOL_FD_CENTRAL = 2, VOL_HG_RANOCHA = 3,
                  VOL_HG_CENTRAL = 4, VOL_PURE_FV = 5 };
enum FluxKind { FLUX_LLF = 0, FLUX_LLF_NAIVE = 1, FLUX_RANOCHA = 2, FLUX_CENTRAL = 3, FLUX_HLLC = 4 };

constexpr int ipow_c(int b, int e) { return e == 0 ? 1 : b * ipow_c(b, e - 1); }

// ---------------------------------------------------------------------------
// Precision traits. The FP32 path uses the hardware reciprocal, base-2
// logarithm and square root (v_rcp_f32, v_log_f32, v_sqrt_f32, 1 ulp); the
// FP64 path refines v_rcp_f64 with two Newton steps and calls the correctly
// rounded double log and sqrt.
// ---------------------------------------------------------------------------
template <typename T> struct Prec;

template <> struct Prec<float> {
  static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
  static __device__ __forceinline__ float ln(float x) {
    return __builtin_amdgcn_logf(x) * 0.69314718055994530942f;
  }
  static __device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
  // Series branch of the logarithmic mean is taken for f^2 < 1e-2 in FP32:
  // its truncation error (first omitted term ~ f^8 / 9 <= 1.2e-9) is below FP32
  // round-off, and the cancellation-prone far branch is only used for
  // |y - x| / (y + x) >= 0.1.  sqrt(1e-2) = 0.1.
  static constexpr float sqrt_f2_threshold = 0.1f;
};

template <> struct Prec<double> {
  static __device__ __forceinline__ double rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    return r;
  }
  static __device__ __forceinline__ double ln(double x) { return ::log(x); }
  static __device__ __forceinline__ double sqrt_(double x) { return ::sqrt(x); }
  // Reference threshold f^2 < 1e-4 (math.jl:199).
  static constexpr double sqrt_f2_threshold = 0.01;
};

// fused multiply-add in the operand precision (__builtin_fma is the double
// builtin and would promote float operands)
__device__ __forceinline__ float tfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double tfma(double a, double b, double c) { return __builtin_fma(a, b, c); }

// v_max / v_min (one instruction; a compare-and-select takes two plus a
// hazard wait). Differs from a > b ? a : b only for NaN operands, where the
// non-NaN one is returned (the select returned b).
__device__ __forceinline__ float tmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ float tmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ double tmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ double tmin(double a, double b) { return __builtin_fmin(a, b); }
// |a| as an operand modifier of the consuming instruction
__device__ __forceinline__ float tabs(float a) { return __builtin_fabsf(a); }
__device__ __forceinline__ double tabs(double a) { return __builtin_fabs(a); }

// ---------------------------------------------------------------------------
// Logarithmic means with one reciprocal each.
//
// With s = x + y, d = y - x, f = d / s:  f^2 < threshold  <=>  |d| < sqrt(thr) * s.
//   ln_mean near:  s / (2 + f2 (2/3 + f2 (2/5 + 2/7 f2)))
//               =  s * 0.5 (1 - f2/3 - 4 f2^2/45 - 44 f2^3/945 - 428 f2^4/14175)
//                  (series of f / atanh f; truncation < 1e-18 at f2 < 1e-4 and
//                   < 4e-11 at f2 < 1e-2)
//   ln_mean far:   d / (log y - log x)
//   inv_ln_mean near: (2 + f2 (2/3 + f2 (2/5 + 2/7 f2))) / s   (math.jl:242-247)
//   inv_ln_mean far:  (log y - log x) / d
// The near/far selection happens before the single reciprocal.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T ln_mean_l(T x, T y, T lx, T ly) {
  T s = x + y, d = y - x;
  bool near = tabs(d) < Prec<T>::sqrt_f2_threshold * s;
  T r = Prec<T>::rcp(near ? s : (ly - lx));
  T t = d * r;
  T z = t * t;
  T q = tfma(tfma(tfma(tfma(T(-428.0 / 28350.0), z, T(-44.0 / 1890.0)),
                                                  z, T(-4.0 / 90.0)), z, T(-1.0 / 6.0)), z, T(0.5));
  return near ? s * q : t;
}

template <typename T>
__device__ __forceinline__ T inv_ln_mean_l(T x, T y, T lx, T ly) {
  T s = x + y, d = y - x;
  bool near = tabs(d) < Prec<T>::sqrt_f2_threshold * s;
  T r = Prec<T>::rcp(near ? s : d);
  T t = d * r;
  T z = t * t;
  T p = tfma(tfma(tfma(T(2.0 / 7.0), z, T(2.0 / 5.0)), z, T(2.0 / 3.0)), z, T(2.0));
  return near ? p * r : (ly - lx) * r;
}

// Per-node data of the flux_ranocha volume kernel: density, its log, the
// inverse temperature beta = rho/p, its log, half velocities, half pressure.
// flux_ranocha's inv_ln_mean(rho_ll p_rr, rho_rr p_ll) p_ll p_rr equals
// inv_ln_mean(beta_ll, beta_rr) (compressible_euler_3d.jl:752-757).
// The "log" fields hold ln(x) in FP64 and log2(x)/2 in FP32 (see
// flux_ranocha_fast); lnv() returns the natural log in both cases. r and b
// hold rho and beta, or, in the square-root representation SQ (FP32 with
// polydeg <= 3), r = sqrt(rho) and b = sqrt(p/rho) = beta^(-1/2): the fast
// FP32 flux then forms 2^(a_x + a_y) = r_x r_y by a product. SQ costs two
// registers per node in the line loop, which polydeg 4 lines cannot afford
// without losing occupancy. rho_of / beta_of recover rho and beta.
template <typename T, int D> struct NodeR {
  T r, lrho, b, lbeta, h[D], hp;
};
constexpr int nodeR_count(int D) { return 5 + D; }

template <typename T, int N> constexpr bool node_sqrt_rep() { return sizeof(T) == 4 && N <= 3; }

template <bool SQ, typename T, int D> __device__ __forceinline__ T rho_of(const NodeR<T, D> &w) {
  if constexpr (SQ) return w.r * w.r;
  else return w.r;
}
template <bool SQ, typename T, int D> __device__ __forceinline__ T beta_of(const NodeR<T, D> &w) {
  if constexpr (SQ) return Prec<T>::rcp(w.b * w.b);
  else return w.b;
}

template <typename T> struct LogRep;
template <> struct LogRep<double> {
  static __device__ __forceinline__ double store(double x) { return ::log(x); }
  static __device__ __forceinline__ double lnv(double a) { return a; }
};
template <> struct LogRep<float> {
  static __device__ __forceinline__ float store(float x) { return 0.5f * __builtin_amdgcn_logf(x); }
  static __device__ __forceinline__ float lnv(float a) { return a * 1.38629436111989061883f; }  // 2 ln 2
};

template <typename T, int D, bool SQ = false>
__device__ __forceinline__ void nodeR_from_cons(const T *u, T gamma_m1, NodeR<T, D> &w) {
  T rho = u[0];
  T irho = Prec<T>::rcp(rho);
  T ke2 = 0;
  T v[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    v[d] = u[1 + d] * irho;
    ke2 = tfma(u[1 + d], v[d], ke2);
  }
  T p = gamma_m1 * (u[D + 1] - T(0.5) * ke2);
  w.lrho = LogRep<T>::store(rho);
  w.lbeta = w.lrho - LogRep<T>::store(p);
  if constexpr (SQ) {
    w.r = __builtin_amdgcn_sqrtf(rho);  // v_sqrt_f32, 1 ulp
    w.b = __builtin_amdgcn_sqrtf(p * irho);
  } else {
    w.r = rho;
    w.b = rho * Prec<T>::rcp(p);
  }
#pragma unroll
  for (int d = 0; d < D; ++d) w.h[d] = T(0.5) * v[d];
  w.hp = T(0.5) * p;
}

// flux_ranocha from per-node data, orientation o (compile-time).
template <typename T, int D, int O, bool SQ = false>
__device__ __forceinline__ void flux_ranocha_r(const NodeR<T, D> &l, const NodeR<T, D> &r,
                                               T inv_gamma_m1, T *f) {
  T rho_mean = ln_mean_l(rho_of<SQ>(l), rho_of<SQ>(r), LogRep<T>::lnv(l.lrho), LogRep<T>::lnv(r.lrho));
  T inv_rho_p_mean =
      inv_ln_mean_l(beta_of<SQ>(l), beta_of<SQ>(r), LogRep<T>::lnv(l.lbeta), LogRep<T>::lnv(r.lbeta));
  T vavg[D];
  T hh = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    vavg[d] = l.h[d] + r.h[d];
    hh = tfma(l.h[d], r.h[d], hh);  // = velocity_square_avg / 2
  }
  T p_avg = l.hp + r.hp;
  T cross = tfma(l.hp, r.h[O], r.hp * l.h[O]);  // = 0.5 (p_ll v_rr + p_rr v_ll) / 2
  T f1 = rho_mean * vavg[O];
  f[0] = f1;
#pragma unroll
  for (int d = 0; d < D; ++d) f[1 + d] = f1 * vavg[d];
  f[1 + O] += p_avg;
  T inner = tfma(inv_rho_p_mean, inv_gamma_m1, hh + hh);
  f[D + 1] = tfma(f1, inner, cross + cross);
}

// FP32 flux_ranocha without divisions. With a = log2(x)/2 stored per node and
// w = ln(y/x)/2 = ln2 (a_y - a_x), the logarithmic mean satisfies exactly
//   ln_mean(x, y) = sqrt(x y) sinh(w)/w = 2^(a_x + a_y) * P(w^2),
//   1/ln_mean(x, y) = 2^-(a_x + a_y) * Q(w^2),
// P(w^2) = sum_k w^(2k)/(2k+1)! = 1 + w^2/6 + w^4/120 + w^6/5040 + ...,
// Q = w/sinh(w) = 1 - w^2/6 + 7 w^4/360 - 31 w^6/15120 + ...
// Truncated after w^4 the relative error is <= 7.6e-10 (P) and 7.8e-9 (Q) for
// |w| <= 1/8 (y/x <= e^(1/4)), below FP32 round-off; callers use this only when
// every pair of the line is within that range (flux_ranocha_r otherwise).
// Coefficients below are pre-multiplied by ln2^(2k) so that the series variable
// is (a_y - a_x)^2. In the square-root representation (SQ) the factors
// 2^(a_x + a_y) = sqrt(x y) are products of the per-node roots, otherwise one
// exp2 each. The energy component is returned halved, f[D+1] = f_energy / 2,
// from half_inv_gamma_m1 = 1/(2 (gamma - 1)): callers accumulate it with twice
// the derivative weights, which is exact (power-of-two scalings) and saves the
// two doublings of flux_ranocha_r.
template <int D, int O, bool SQ>
__device__ __forceinline__ void flux_ranocha_fast(const NodeR<float, D> &l, const NodeR<float, D> &r,
                                                  float half_inv_gamma_m1, float *f) {
  const float d1 = r.lrho - l.lrho, z1 = d1 * d1;
  const float p1 = tfma(tfma(0.00192362582153f, z1, 0.0800755023197f), z1, 1.0f);
  const float d2 = r.lbeta - l.lbeta, z2 = d2 * d2;
  const float q2 = tfma(tfma(0.00448846025023f, z2, -0.0800755023197f), z2, 1.0f);
  float rho_mean, inv_rho_p_mean;
  if constexpr (SQ) {
    rho_mean = (l.r * r.r) * p1;
    inv_rho_p_mean = (l.b * r.b) * q2;
  } else {
    rho_mean = __builtin_amdgcn_exp2f(l.lrho + r.lrho) * p1;
    inv_rho_p_mean = __builtin_amdgcn_exp2f(-(l.lbeta + r.lbeta)) * q2;
  }
  float vavg[D];
  float hh = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    vavg[d] = l.h[d] + r.h[d];
    hh = tfma(l.h[d], r.h[d], hh);
  }
  const float p_avg = l.hp + r.hp;
  const float cross = tfma(l.hp, r.h[O], r.hp * l.h[O]);
  const float f1 = rho_mean * vavg[O];
  f[0] = f1;
#pragma unroll
  for (int d = 0; d < D; ++d) f[1 + d] = f1 * vavg[d];
  f[1 + O] += p_avg;
  const float inner_h = tfma(inv_rho_p_mean, half_inv_gamma_m1, hh);
  f[D + 1] = tfma(f1, inner_h, cross);  // half the energy flux
}

// |a_y - a_x| bound of the FP32 fast path: |w| <= 1/8  <=>  |a_y - a_x| <= 1/(8 ln2)
constexpr float kFastLogSpan = 0.1803368801f;

// Physical flux of a conservative state in orientation o (runtime), plus the
// primitive data LLF needs.
template <typename T, int D> struct ConsP {
  T u[D + 2];
  T v[D];
  T p;
  T irho;
};

template <typename T, int D>
__device__ __forceinline__ void consp_from_cons(const T *u, T gamma_m1, ConsP<T, D> &c) {
  T irho = Prec<T>::rcp(u[0]);
  T ke2 = 0;
#pragma unroll
  for (int k = 0; k < D + 2; ++k) c.u[k] = u[k];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    c.v[d] = u[1 + d] * irho;
    ke2 = tfma(u[1 + d], c.v[d], ke2);
  }
  c.p = gamma_m1 * (u[D + 1] - T(0.5) * ke2);
  c.irho = irho;
}

template <typename T, int D, int O>
__device__ __forceinline__ void phys_flux(const ConsP<T, D> &c, T *f) {
  f[0] = c.u[1 + O];
#pragma unroll
  for (int d = 0; d < D; ++d) f[1 + d] = c.u[1 + O] * c.v[d];
  f[1 + O] += c.p;
  f[D + 1] = (c.u[D + 1] + c.p) * c.v[O];
}

// flux_central (numerical_fluxes.jl:17-25)
template <typename T, int D, int O>
__device__ __forceinline__ void flux_central_c(const ConsP<T, D> &l, const ConsP<T, D> &r, T *f) {
  T fl[D + 2], fr[D + 2];
  phys_flux<T, D, O>(l, fl);
  phys_flux<T, D, O>(r, fr);
#pragma unroll
  for (int k = 0; k < D + 2; ++k) f[k] = T(0.5) * (fl[k] + fr[k]);
}

// FluxLaxFriedrichs(max_abs_speed | max_abs_speed_naive) with the sound speeds given
template <typename T, int D, int O>
__device__ __forceinline__ void flux_llf_pc(const ConsP<T, D> &l, const ConsP<T, D> &r, T c_ll, T c_rr, bool naive,
                                            T *f) {
  T fl[D + 2], fr[D + 2];
  phys_flux<T, D, O>(l, fl);
  phys_flux<T, D, O>(r, fr);
  T al = tabs(l.v[O]), ar = tabs(r.v[O]);
  T lambda = naive ? tmax(al, ar) + tmax(c_ll, c_rr) : tmax(al + c_ll, ar + c_rr);
  T hl = T(-0.5) * lambda;
#pragma unroll
  for (int k = 0; k < D + 2; ++k) f[k] = T(0.5) * (fl[k] + fr[k]) + hl * (r.u[k] - l.u[k]);
}

// Twice the same flux, (fL + fR) - lambda (uR - uL): one operation less per
// component; callers scale by half the weight they apply. Scaling by 2 is
// exact, so the scaled result equals 2 * flux_llf_pc bit for bit.
template <typename T, int D, int O>
__device__ __forceinline__ void flux_llf_pc2(const ConsP<T, D> &l, const ConsP<T, D> &r, T c_ll, T c_rr, bool naive,
                                             T *f) {
  T fl[D + 2], fr[D + 2];
  phys_flux<T, D, O>(l, fl);
  phys_flux<T, D, O>(r, fr);
  T al = tabs(l.v[O]), ar = tabs(r.v[O]);
  T lambda = naive ? tmax(al, ar) + tmax(c_ll, c_rr) : tmax(al + c_ll, ar + c_rr);
#pragma unroll
  for (int k = 0; k < D + 2; ++k) f[k] = tfma(-lambda, r.u[k] - l.u[k], fl[k] + fr[k]);
}

// Own face state of an FP32 flux_ranocha line from its node record in the
// square-root representation: the same velocity and pressure as
// consp_from_cons (v = 2 h, p = 2 hp exactly), and the sound speed
// sqrt(gamma p / rho) = sqrt(gamma) b.
template <int D>
__device__ __forceinline__ void consp_from_node(const float *u, const NodeR<float, D> &w, float sqrt_gamma,
                                                ConsP<float, D> &c, float &cs) {
#pragma unroll
  for (int k = 0; k < D + 2; ++k) c.u[k] = u[k];
#pragma unroll
  for (int d = 0; d < D; ++d) c.v[d] = 2.0f * w.h[d];
  c.p = 2.0f * w.hp;
  c.irho = 0.0f;  // not needed by phys_flux / flux_llf_pc
  cs = sqrt_gamma * w.b;
}

// FluxLaxFriedrichs(max_abs_speed | max_abs_speed_naive)
template <typename T, int D, int O>
__device__ __forceinline__ void flux_llf_c(const ConsP<T, D> &l, const ConsP<T, D> &r, bool naive,
                                           T gamma, T *f) {
  const T c_ll = Prec<T>::sqrt_(gamma * l.p * l.irho);
  const T c_rr = Prec<T>::sqrt_(gamma * r.p * r.irho);
  flux_llf_pc<T, D, O>(l, r, c_ll, c_rr, naive, f);
}

// flux_hllc (compressible_euler_3d.jl:1417-1541, 2D compressible_euler_2d.jl:
// 1714-1816, orientation form): Toro's HLL with contact, signal speeds from
// the Roe averages of the normal velocity and the enthalpy (Einfeldt), the
// star-state speed from the jump conditions; the upwind side's flux, or that
// side's flux plus its signal speed times the jump to its star state.
// Quotients as products with refined reciprocals (FP64 to round-off of the
// reference's divisions).
template <typename T, int D, int O>
__device__ __forceinline__ void flux_hllc_c(const ConsP<T, D> &l, const ConsP<T, D> &r, T gamma, T *f) {
  constexpr int V = D + 2;
  const T rho_ll = l.u[0], rho_rr = r.u[0];
  const T c_ll = Prec<T>::sqrt_(gamma * l.p * l.irho), c_rr = Prec<T>::sqrt_(gamma * r.p * r.irho);
  T f_ll[V], f_rr[V];
  phys_flux<T, D, O>(l, f_ll);
  phys_flux<T, D, O>(r, f_rr);
  const T sr_ll = Prec<T>::sqrt_(rho_ll), sr_rr = Prec<T>::sqrt_(rho_rr);
  const T sum_sr = sr_ll + sr_rr, isum = Prec<T>::rcp(sum_sr);
  const T vel_L = l.v[O], vel_R = r.v[O];
  const T vel_roe = (sr_ll * vel_L + sr_rr * vel_R) * isum;
  T mag = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const T vr = sr_ll * l.v[d] + sr_rr * r.v[d];
    mag += vr * vr;
  }
  mag = mag * (isum * isum);
  const T H_ll = (l.u[V - 1] + l.p) * l.irho, H_rr = (r.u[V - 1] + r.p) * r.irho;
  const T H_roe = (sr_ll * H_ll + sr_rr * H_rr) * isum;
  const T c_roe = Prec<T>::sqrt_((gamma - T(1)) * (H_roe - T(0.5) * mag));
  const T Ssl = tmin(vel_L - c_ll, vel_roe - c_roe);
  const T Ssr = tmax(vel_R + c_rr, vel_roe + c_roe);
  if (Ssl >= T(0)) {
#pragma unroll
    for (int v = 0; v < V; ++v) f[v] = f_ll[v];
  } else if (Ssr <= T(0)) {
#pragma unroll
    for (int v = 0; v < V; ++v) f[v] = f_rr[v];
  } else {
    const T sMu_L = Ssl - vel_L, sMu_R = Ssr - vel_R;
    const T SStar = (r.p - l.p + rho_ll * vel_L * sMu_L - rho_rr * vel_R * sMu_R) *
                    Prec<T>::rcp(rho_ll * sMu_L - rho_rr * sMu_R);
    const bool left = SStar >= T(0);  // Ssl <= 0 holds here
    const ConsP<T, D> &s = left ? l : r;
    const T *fs = left ? f_ll : f_rr;
    const T sMu = left ? sMu_L : sMu_R, S = left ? Ssl : Ssr;
    const T rho_s = s.u[0];
    const T densStar = rho_s * sMu * Prec<T>::rcp(S - SStar);
    const T enerStar = s.u[V - 1] * s.irho + (SStar - s.v[O]) * (SStar + s.p * Prec<T>::rcp(rho_s * sMu));
    T ustar[V];
    ustar[0] = densStar;
#pragma unroll
    for (int d = 0; d < D; ++d) ustar[1 + d] = densStar * (d == O ? SStar : s.v[d]);
    ustar[V - 1] = densStar * enerStar;
#pragma unroll
    for (int v = 0; v < V; ++v) f[v] = fs[v] + S * (ustar[v] - s.u[v]);
  }
}

// flux_ranocha from conservative states (surface / FV use).
template <typename T, int D, int O>
__device__ __forceinline__ void flux_ranocha_c(const T *ul, const T *ur, T gamma_m1, T *f) {
  NodeR<T, D> a, b;
  nodeR_from_cons<T, D>(ul, gamma_m1, a);
  nodeR_from_cons<T, D>(ur, gamma_m1, b);
  flux_ranocha_r<T, D, O>(a, b, T(1) / gamma_m1, f);
}

// Runtime-selected two-point flux of conservative states (surface and FV
// subcells). flux_hllc is compiled in only where HLLC is true (the pure-FV
// element kernels and the mortar fluxes): as one more run-time branch it would
// raise the register count, and lower the occupancy, of every other variant.
template <typename T, int D, int O, bool HLLC = false>
__device__ __forceinline__ void num_flux_c(int kind, const T *ul, const T *ur, T gamma, T *f) {
  if (kind == FLUX_RANOCHA) {
    flux_ranocha_c<T, D, O>(ul, ur, gamma - T(1), f);
  } else {
    ConsP<T, D> l, r;
    consp_from_cons<T, D>(ul, gamma - T(1), l);
    consp_from_cons<T, D>(ur, gamma - T(1), r);
    if (kind == FLUX_CENTRAL) {
      flux_central_c<T, D, O>(l, r, f);
      return;
    }
    if constexpr (HLLC) {
      if (kind == FLUX_HLLC) {
        flux_hllc_c<T, D, O>(l, r, gamma, f);
        return;
      }
    }
    flux_llf_c<T, D, O>(l, r, kind == FLUX_LLF_NAIVE, gamma, f);
  }
}

// source_terms_convergence_test (compressible_euler_3d.jl:127-153, 2D :120-146),
// given si, co = sincos(omega (x_1 + ... + x_D - t)), omega = pi.
template <typename T, int D>
__device__ __forceinline__ void source_convergence_sc(T si, T co, T gamma, T *s) {
  const T c = 2, A = T(0.1), omega = T(3.14159265358979323846);
  T rho = c + A * si;
  T rho_x = omega * A * co;
  if constexpr (D == 3) {
    T tmp = (2 * rho - T(1.5)) * (gamma - 1);
    s[0] = 2 * rho_x;
    s[1] = rho_x * (2 + tmp);
    s[2] = s[1];
    s[3] = s[1];
    s[4] = rho_x * (4 * rho + 3 * tmp);
  } else {
    T tmp = (2 * rho - 1) * (gamma - 1);
    s[0] = rho_x;
    s[1] = rho_x * (1 + tmp);
    s[2] = s[1];
    s[3] = 2 * rho_x * (rho + tmp);
  }
}

// NaN-propagating maximum of the wave-speed reductions (max_dt).
__device__ __forceinline__ double nanmax(double a, double b) {
  return (a != a || b != b) ? __builtin_nan("") : (a > b ? a : b);
}

// Per-direction wave speeds |v_d| + c of one node (max_abs_speeds,
// compressible_euler_3d.jl:1770-1775) in FP64 from its conservative values
// (load(v) returns value v as double). Shared by max_speed_nodes_kernel and the
// fused last stage of a step, so that both give the same bits.
template <int D, typename L> __device__ __forceinline__ void node_wave_speeds(L load, double gamma, double *lam) {
  const double rho = load(0);
  double ke2 = 0, v[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    v[d] = load(1 + d) / rho;
    ke2 += load(1 + d) * v[d];
  }
  const double p = (gamma - 1) * (load(D + 1) - 0.5 * ke2);
  const double c = sqrt(gamma * p / rho);
#pragma unroll
  for (int d = 0; d < D; ++d) lam[d] = fabs(v[d]) + c;
}

__device__ __forceinline__ void tsincos(float x, float *s, float *c) { sincosf(x, s, c); }
__device__ __forceinline__ void tsincos(double x, double *s, double *c) { sincos(x, s, c); }

// Parameters of IndicatorHennemannGassner (indicators.jl:50, 119-131): the
// threshold 0.5 10^(-1.8 nnodes^0.25) and the sharpness log((1 - 1e-4) / 1e-4)
// are set by the host from the number of nodes per direction.
struct HGParams {
  double gamma, alpha_max, alpha_min, threshold, parameter_s;
  int variable;  // indicator variable: 0 density_pressure, 1 density, 2 pressure
};

// Features of the learned indicator (2D p3 only; feat [E][32], zero padded;
// computed by indicator_2d_p3_kernel):
//   0..15  nodal rho*p, centred by its element mean and divided by the largest
//          deviation (+ 1e-7 |mean|), so that they lie in [-1, 1];
//   16, 17 the HG logits s/T (r - T) of the two energy ratios r, / 8, clipped
//          to [-2, 2];
//   18     log10(E_clip2 / E_total + 1e-16) / 8, clipped to [-2, 0].
constexpr int kMlpFeatureStride = 32;

}  // namespace trixi_hip
