// dg_aux.hpp -- the kernels around the element kernel: Hennemann-Gassner
// indicator, 2N Runge-Kutta stage update, precision conversion, max_dt,
// error norms and integrals. All reductions run in FP64 with a fixed
// two-pass order (per-block partials, then one block), so results are
// deterministic and the kernels can live inside a captured graph.
#pragma once
#include "dg_common.hpp"

namespace trixi_hip {

// ---------------------------------------------------------------------------
// IndicatorHennemannGassner, raw (before smoothing):
//   indicators.jl:119-156, indicators_2d.jl:30-102, indicators_3d.jl:39-122,
//   multiply_scalar_dimensionwise! interpolation.jl:207-235 / 348-386,
//   density_pressure compressible_euler_3d.jl:1951-1956.
// FP64 arithmetic in both precision modes (the clipping at alpha_min is
// threshold-sensitive). One workgroup handles IPB elements, one thread per
// node for the modal transform, one thread per element for the energies
// (reference loop order).
// ---------------------------------------------------------------------------
template <int N> struct VinvConst {
  double Vinv[(N + 1) * (N + 1)];
};

template <typename T, int D, int N, int IPB, int NT>
__global__ __launch_bounds__(NT) void hg_indicator_kernel(const T *__restrict__ u,
                                                          double *__restrict__ alpha_raw, int E,
                                                          HGParams hp, VinvConst<N> vc) {
  constexpr int n = N + 1, NN = ipow_c(n, D), V = D + 2;
  __shared__ double s_a[IPB * NN], s_b[IPB * NN];
  const int tid = threadIdx.x;
  const int e0 = blockIdx.x * IPB;
  if (e0 >= E) return;
  const int ne = min(IPB, E - e0);
  for (int q = tid; q < ne * NN; q += NT) {
    const T *un = u + (size_t)(e0 * NN + q) * V;
    double rho = un[0], m2 = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) m2 += double(un[1 + d]) * double(un[1 + d]);
    // density_pressure (euler_2d.jl:2205-2210), density, pressure
    s_a[q] = hp.variable == 1   ? rho
             : hp.variable == 2 ? (hp.gamma - 1) * (double(un[D + 1]) - 0.5 * m2 / rho)
                                : (hp.gamma - 1) * (rho * double(un[D + 1]) - 0.5 * m2);
  }
  __syncthreads();
  double *src = s_a, *dst = s_b;
#pragma unroll
  for (int o = 0; o < D; ++o) {
    const int st = o == 0 ? 1 : (o == 1 ? n : n * n);
    for (int q = tid; q < ne * NN; q += NT) {
      const int el = q / NN, node = q - el * NN;
      const int io = (node / st) % n;
      const int b0 = el * NN + node - io * st;
      double res = 0;
#pragma unroll
      for (int ii = 0; ii < n; ++ii) res += vc.Vinv[io * n + ii] * src[b0 + ii * st];
      dst[q] = res;
    }
    __syncthreads();
    double *t = src;
    src = dst;
    dst = t;
  }
  if (tid < ne) {
    const double *m = src + tid * NN;
    auto m2 = [&](int i, int j, int kk) {
      double x = m[i + n * j + n * n * kk];
      return x * x;
    };
    double te2 = 0, te1, te;
    if constexpr (D == 2) {
      for (int j = 0; j < n - 2; ++j)
        for (int i = 0; i < n - 2; ++i) te2 += m2(i, j, 0);
      te1 = te2;
      for (int i = 0; i < n - 1; ++i) te1 += m2(i, n - 2, 0);
      for (int j = 0; j < n - 2; ++j) te1 += m2(n - 2, j, 0);
      te = te1;
      for (int i = 0; i < n; ++i) te += m2(i, n - 1, 0);
      for (int j = 0; j < n - 1; ++j) te += m2(n - 1, j, 0);
    } else {
      for (int kk = 0; kk < n - 2; ++kk)
        for (int j = 0; j < n - 2; ++j)
          for (int i = 0; i < n - 2; ++i) te2 += m2(i, j, kk);
      te1 = te2;
      for (int j = 0; j < n - 1; ++j)
        for (int i = 0; i < n - 1; ++i) te1 += m2(i, j, n - 2);
      for (int kk = 0; kk < n - 2; ++kk)
        for (int i = 0; i < n - 1; ++i) te1 += m2(i, n - 2, kk);
      for (int kk = 0; kk < n - 2; ++kk)
        for (int j = 0; j < n - 2; ++j) te1 += m2(n - 2, j, kk);
      te = te1;
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) te += m2(i, j, n - 1);
      for (int kk = 0; kk < n - 1; ++kk)
        for (int i = 0; i < n; ++i) te += m2(i, n - 1, kk);
      for (int kk = 0; kk < n - 1; ++kk)
        for (int j = 0; j < n - 1; ++j) te += m2(n - 1, j, kk);
    }
    double ef1 = te != 0 ? (te - te1) / te : 0.0;
    double ef2 = te1 != 0 ? (te1 - te2) / te1 : 0.0;
    double energy = ef1 > ef2 ? ef1 : ef2;
    double al = 1.0 / (1.0 + exp(-hp.parameter_s / hp.threshold * (energy - hp.threshold)));
    if (al < hp.alpha_min) al = 0.0;
    if (al > 1 - hp.alpha_min) al = 1.0;
    alpha_raw[e0 + tid] = al < hp.alpha_max ? al : hp.alpha_max;
  }
}

// Stand-alone smoothing (for trixi_hip_indicator; the element kernel smooths inline).
template <int D>
__global__ void hg_smooth_kernel(const double *__restrict__ alpha_raw,
                                 const int32_t *__restrict__ nbr, double *__restrict__ alpha,
                                 int E, int smooth) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double al = alpha_raw[e];
  if (smooth) {
    double m = al;
    for (int dir = 0; dir < 2 * D; ++dir) {
      const int nb = nbr[dir + 2 * D * e];  // < 0: boundary, no interface
      double an = nb >= 0 ? 0.5 * alpha_raw[nb] : 0.0;
      m = an > m ? an : m;
    }
    al = m;
  }
  alpha[e] = al;
}

// ---------------------------------------------------------------------------
// 2N low-storage stage update (methods_2N.jl:153-157):
//   u_tmp = du - u_tmp * a_stage;  u += u_tmp * b_stage_dt
// u FP64 register; du and u_tmp stored in the RHS precision TD, arithmetic in
// FP64; optional FP32 shadow of u.
// b_stage_dt = b_stage * dt with dt taken from tdev[1] when tdev != null.
// ---------------------------------------------------------------------------
template <typename TD>
__global__ void rk2n_update_kernel(const TD *__restrict__ du, TD *__restrict__ u_tmp,
                                   double *__restrict__ u, float *__restrict__ u32, size_t count,
                                   double a_stage, double b_stage, const double *__restrict__ tdev,
                                   int first_stage) {
  const double b_dt = tdev ? b_stage * tdev[1] : b_stage;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) {
    double ut = first_stage ? 0.0 : double(u_tmp[i]);
    ut = double(du[i]) - ut * a_stage;
    double un = tfma(ut, b_dt, u[i]);
    u_tmp[i] = TD(ut);
    u[i] = un;
    if (u32) u32[i] = float(un);
  }
}

template <typename TO, typename TI>
__global__ void convert_kernel(TO *__restrict__ out, const TI *__restrict__ in, size_t count) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) out[i] = TO(in[i]);
}

// ---------------------------------------------------------------------------
// max_dt (stepsize_dg3d.jl:8-32, stepsize_dg2d.jl:8-30), max_abs_speeds
// (compressible_euler_3d.jl:1770-1775). Pass 1: one thread per element, block
// max into partials. Pass 2: one block reduces and writes
//   out[0] = cfl * 2 / (nnodes * max_scaled_speed)
// NaN propagates (fmax would drop it; the reference uses Base.max).
// ---------------------------------------------------------------------------

// max_dt (stepsize_dg3d.jl:8-32, stepsize_dg2d.jl:8-30) in two passes: per
// block, max over its elements of inv_jac * sum_d max_nodes lambda_d with
// lambda_d = |v_d| + c (max_abs_speeds, compressible_euler_3d.jl:1770-1775);
// one thread per node (groups of EB = 256 / NN elements, coalesced loads; a
// block walks groups with a grid stride and keeps its running maximum). The
// node maxima are reduced through LDS in parallel: S threads per (element,
// direction) each take every S-th node, then one thread per pair combines
// the S partials. Maxima are exact, so the result does not depend on the
// order; the direction sum runs in d order. NaN speeds propagate (nanmax).
template <typename T, int D, int N>
__global__ __launch_bounds__(256) void max_speed_nodes_kernel(const T *__restrict__ u,
                                                              const double *__restrict__ inv_jac, int E,
                                                              double gamma, double *__restrict__ partial) {
  constexpr int n = N + 1, NN = ipow_c(n, D), V = D + 2;
  constexpr int EB = NN >= 256 ? 1 : 256 / NN;
  constexpr int S = 8, NPAIR = EB * D;
  static_assert(NN <= 256, "one thread per node");
  __shared__ double s_l[D][EB * NN];
  __shared__ double s_p[NPAIR * S];
  __shared__ double s_e[EB];
  const int tid = threadIdx.x;
  const int ngroups = (E + EB - 1) / EB;
  double bmax = 4.9406564584124654e-324;  // nextfloat(0.0), as the reference starts
  // groups of EB elements, grid-stride (block-uniform loop: the barriers are safe)
  for (int grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const int e0 = grp * EB;
    const int ne = min(EB, E - e0);
    if (tid < ne * NN) {
      const T *un = u + ((size_t)e0 * NN + tid) * V;
      double lam[D];
      node_wave_speeds<D>([&](int v) { return double(un[v]); }, gamma, lam);
#pragma unroll
      for (int d = 0; d < D; ++d) s_l[d][tid] = lam[d];
    }
    __syncthreads();
    for (int k = tid; k < NPAIR * S; k += 256) {
      const int pair = k / S, sub = k - pair * S;
      const int el = pair / D, d = pair - el * D;
      double m = 0.0;
      if (el < ne)
        for (int q = sub; q < NN; q += S) m = nanmax(m, s_l[d][el * NN + q]);
      s_p[k] = m;
    }
    __syncthreads();
    if (tid < ne) {
      double sum = 0;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        double m = 0.0;
#pragma unroll
        for (int j = 0; j < S; ++j) m = nanmax(m, s_p[(tid * D + d) * S + j]);
        sum += m;
      }
      s_e[tid] = inv_jac[e0 + tid] * sum;
    }
    __syncthreads();
    if (tid == 0)
      for (int i = 0; i < ne; ++i) bmax = nanmax(bmax, s_e[i]);
    __syncthreads();  // s_l, s_p and s_e are rewritten by the next group
  }
  if (tid == 0) partial[blockIdx.x] = bmax;
}

// Max (NaN-propagating) of np partials by one workgroup of kRedThreads
// threads: independent strided loads, then a tree in LDS; s_red[0] holds the
// result for every thread after the call.
constexpr int kRedThreads = 1024;
__device__ __forceinline__ void block_nanmax(const double *__restrict__ partial, int np, double *s_red) {
  double m = 4.9406564584124654e-324;
#pragma unroll 8
  for (int i = threadIdx.x; i < np; i += kRedThreads) m = nanmax(m, partial[i]);
  s_red[threadIdx.x] = m;
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) s_red[threadIdx.x] = nanmax(s_red[threadIdx.x], s_red[threadIdx.x + w]);
    __syncthreads();
  }
}

// Device step control of a 2N run with StepsizeCallback(cfl) and limit_dt
// (stepsize.jl:93-124, 146-155; time_integration.jl:46-55; the step loop of
// methods_2N.jl:131-168), after the stages of one step:
//   t += dt; steps += 1; the step was the last one -> done
//   dt = cfl * 2 / (nnodes * max_e inv_jac_e sum_d max lambda_d)
//   if t + dt > t_end or isapprox(t + dt, t_end): dt = t_end - t (next step is the last)
//   NaN dt -> error, done; steps == max_steps -> done; done -> dt = 0.
// tdt = {t, dt}; ctl = {t_end, cfl, last, done, steps, error, max_steps}. A
// step replayed after done leaves u and t unchanged (dt = 0).
static __global__ __launch_bounds__(kRedThreads) void step_control_kernel(const double *__restrict__ partial, int np,
                                                                   int nnodes, double *__restrict__ tdt,
                                                                   double *__restrict__ ctl) {
  __shared__ double s_red[kRedThreads];
  block_nanmax(partial, np, s_red);
  if (threadIdx.x != 0) return;
  if (ctl[3] != 0.0) {  // done: this replay changed nothing
    tdt[1] = 0.0;
    return;
  }
  const double t = tdt[0] + tdt[1];
  const double steps = ctl[4] + 1.0;
  bool done = ctl[2] != 0.0 || steps >= ctl[6];
  double dt = ctl[1] * (2.0 / (nnodes * s_red[0]));
  if (!done && dt != dt) {  // the next step would start with a NaN dt
    ctl[5] = 1.0;
    done = true;
  }
  double last = 0.0;
  if (!done) {
    const double t_end = ctl[0], s = t + dt;
    const double tol = 1.4901161193847656e-08 * fmax(fabs(s), fabs(t_end));  // isapprox rtol sqrt(eps)
    if (s > t_end || s == t_end || fabs(s - t_end) <= tol) {
      dt = t_end - t;
      last = 1.0;
    }
  }
  tdt[0] = t;
  tdt[1] = done ? 0.0 : dt;
  ctl[2] = last;
  ctl[3] = done ? 1.0 : 0.0;
  ctl[4] = steps;
}

static __global__ __launch_bounds__(kRedThreads) void max_dt_finalize_kernel(const double *__restrict__ partial,
                                                                      int np, int nnodes, double cfl,
                                                                      double *__restrict__ out) {
  __shared__ double s_red[kRedThreads];
  block_nanmax(partial, np, s_red);
  if (threadIdx.x == 0) out[0] = cfl * (2.0 / (nnodes * s_red[0]));
}

// ---------------------------------------------------------------------------
// Exact solutions for calc_error_norms (initial conditions, host-side setup
// in the reference; the device needs them to compare at analysis nodes).
// ---------------------------------------------------------------------------
struct ExactParams {
  int kind;
  double gamma;
  double params[4];
  double modes[4 * 8 * 4];
};

template <int D>
__device__ void exact_solution(const ExactParams &ep, const double *x, double t, double *u) {
  const int V = D + 2;
  const double g = ep.gamma, igm1 = 1.0 / (g - 1);
  auto prim2cons = [&](const double *w) {
    u[0] = w[0];
    double ke2 = 0;
    for (int d = 0; d < D; ++d) {
      u[1 + d] = w[0] * w[1 + d];
      ke2 += u[1 + d] * w[1 + d];
    }
    u[V - 1] = w[V - 1] * igm1 + 0.5 * ke2;
  };
  switch (ep.kind) {
    case 0: {  // initial_condition_constant
      u[0] = 1;
      u[1] = 0.1;
      u[2] = -0.2;
      if (D == 3) u[3] = 0.7;
      u[V - 1] = 10;
      break;
    }
    case 1: {  // initial_condition_convergence_test
      double s = 0;
      for (int d = 0; d < D; ++d) s += x[d];
      double ini = 2 + 0.1 * sin(M_PI * (s - t));
      for (int v = 0; v < V - 1; ++v) u[v] = ini;
      u[V - 1] = ini * ini;
      break;
    }
    case 2: {  // initial_condition_weak_blast_wave
      if (D == 2) {  // compressible_euler_2d.jl:181-200: polar coordinates
        const double r = sqrt(x[0] * x[0] + x[1] * x[1]);
        const double phi = atan2(x[1], x[0]);
        const double w[4] = {r > 0.5 ? 1.0 : 1.1691, r > 0.5 ? 0.0 : 0.1882 * cos(phi),
                             r > 0.5 ? 0.0 : 0.1882 * sin(phi), r > 0.5 ? 1.0 : 1.245};
        prim2cons(w);
        break;
      }
      // compressible_euler_3d.jl:163-183: spherical coordinates
      double xn = x[0], yn = x[1], zn = D == 3 ? x[2] : 0.0;
      double r = sqrt(xn * xn + yn * yn + zn * zn);
      double phi = atan2(yn, xn);
      double theta = r == 0 ? 0.0 : acos(zn / r);
      double w[5];
      w[0] = r > 0.5 ? 1.0 : 1.1691;
      w[1] = r > 0.5 ? 0.0 : 0.1882 * cos(phi) * sin(theta);
      w[2] = r > 0.5 ? 0.0 : 0.1882 * sin(phi) * sin(theta);
      w[3] = r > 0.5 ? 0.0 : 0.1882 * cos(theta);
      w[4] = r > 0.5 ? 1.0 : 1.245;
      if (D == 3) prim2cons(w);
      break;
    }
    case 3: {  // initial_condition_isentropic_vortex (elixir_euler_vortex.jl)
      double rho = 1, v1 = 1, v2 = 1, p = ep.params[0];
      double rt = p / rho;
      double c0 = -x[1], c1 = x[0];
      double r2 = c0 * c0 + c1 * c1;
      double du = 5.0 / (2 * M_PI) * exp(0.5 * (1 - r2));
      double dtemp = -(g - 1) / (2 * g * rt) * du * du;
      rho = rho * pow(1 + dtemp, 1 / (g - 1));
      v1 = v1 + du * c0;
      v2 = v2 + du * c1;
      p = p * pow(1 + dtemp, g / (g - 1));
      double w[4] = {rho, v1, v2, p};
      if (D == 2) prim2cons(w);
      break;
    }
    case 4: {  // initial_condition_kelvin_helmholtz_instability (noise-free)
      double B = tanh(15 * x[1] + 7.5) - tanh(15 * x[1] - 7.5);
      double w[4] = {0.5 + 0.75 * B, 0.5 * (B - 1), 0.1 * sin(M_PI * 2 * x[0]), 1.0};
      if (D == 2) prim2cons(w);
      break;
    }
    case 6: {  // initial_condition_density_pulse (elixir_euler_amr.jl:15-26)
      double r2 = 0;
      for (int d = 0; d < D; ++d) r2 += x[d] * x[d];
      const double rho = 1 + exp(-r2) / 2;
      u[0] = rho;
      for (int d = 0; d < D; ++d) u[1 + d] = rho;
      u[V - 1] = igm1 + 0.5 * rho * D;
      break;
    }
    case 7: {  // medium blast wave, 2D (elixir_euler_blast_wave_amr.jl:14-33)
      const double r = sqrt(x[0] * x[0] + x[1] * x[1]);
      const double phi = atan2(x[1], x[0]);
      double w[4] = {r > 0.5 ? 1.0 : 1.1691, r > 0.5 ? 0.0 : 0.1882 * cos(phi), r > 0.5 ? 0.0 : 0.1882 * sin(phi),
                     r > 0.5 ? 1.0e-3 : 1.245};
      if (D == 2) prim2cons(w);
      break;
    }
    case 8: {  // blob in a supersonic stream, 2D (elixir_euler_blob_amr.jl:13-58)
      const double R = 1, dens0 = 1, Chi = 10, tau_cr = 1 / 1.6;
      const double velx0 = 2 * R * sqrt(Chi) / tau_cr, cs = velx0 / 2.7, p0 = cs * cs * dens0 / g;
      const double xr = x[0] + 15, yr = x[1];
      const double r = sqrt(xr * xr + yr * yr);
      const double shape = 0.5 * (1 + (tanh(2 * (r + R)) - (tanh(2 * (r - R)) + 1)));
      double w[4] = {dens0 + (Chi - 1) * shape, velx0 - velx0 * shape, 0.0, p0};
      if (D == 2) prim2cons(w);
      break;
    }
    default: {  // seeded Fourier modes
      double w[5];
      for (int fld = 0; fld < 1 + D; ++fld) {
        double s = 0;
        for (int m = 0; m < 8; ++m) {
          const double *md = &ep.modes[(fld * 8 + m) * 4];
          double arg = 0;
          for (int d = 0; d < D; ++d) arg += md[d] * (x[d] + 1.0);
          s += sin(M_PI * arg + md[3]);
        }
        w[fld] = fld == 0 ? 1.0 + ep.params[0] * s : ep.params[1] * s;
      }
      w[D + 1] = 1.0;
      prim2cons(w);
      break;
    }
  }
}

template <int N, int NA> struct AnalysisConst {
  double vdm[NA * (N + 1)];  // [a*n + j]
  double aw[NA];
  double nodes[N + 1];
};

// calc_error_norms (analysis_dg3d.jl:123-161): one block per element. The
// element's solution and node coordinates are staged in LDS; each thread
// interpolates to its analysis points with the tensor-product Vandermonde
// (the reference applies it dimension by dimension; same sums, other order).
// partial[e] = {l2 contributions [V], linf [V]}. The time is tdev[0] when
// tdev is given (the device clock of a running solve), else t.
template <typename T, int D, int N, int NA>
__global__ __launch_bounds__(256) void error_norms_kernel(const T *__restrict__ u,
                                                          const double *__restrict__ inv_jac,
                                                          const double *__restrict__ centers,
                                                          int E, double t_host, const double *__restrict__ tdev,
                                                          ExactParams ep, AnalysisConst<N, NA> ac,
                                                          double *__restrict__ partial) {
  constexpr int n = N + 1, NN = ipow_c(n, D), V = D + 2, NQ = ipow_c(NA, D);
  constexpr int W = V + D;  // interpolate u and x together
  __shared__ double s0[NN * W];
  __shared__ double s_red[256];
  const int e = blockIdx.x;
  if (e >= E) return;
  const int tid = threadIdx.x;
  const double t = tdev ? tdev[0] : t_host;
  const double jac = 1.0 / inv_jac[e];
  for (int q = tid; q < NN; q += 256) {
    int idx[3] = {q % n, (q / n) % n, q / (n * n)};
    for (int v = 0; v < V; ++v) s0[q * W + v] = double(u[((size_t)e * NN + q) * V + v]);
    for (int d = 0; d < D; ++d) s0[q * W + V + d] = centers[d + D * e] + jac * ac.nodes[idx[d]];
  }
  __syncthreads();
  const double vol = D == 3 ? jac * jac * jac : jac * jac;
  double l2[V], li[V];
  for (int v = 0; v < V; ++v) l2[v] = li[v] = 0;
  for (int q = tid; q < NQ; q += 256) {
    int id[3] = {q % NA, (q / NA) % NA, q / (NA * NA)};
    double w = 1;
    for (int d = 0; d < D; ++d) w *= ac.aw[id[d]];
    double val[W];
    for (int c = 0; c < W; ++c) val[c] = 0;
    for (int node = 0; node < NN; ++node) {
      int jd[3] = {node % n, (node / n) % n, node / (n * n)};
      double c = 1;
      for (int d = 0; d < D; ++d) c *= ac.vdm[id[d] * n + jd[d]];
      for (int cc = 0; cc < W; ++cc) val[cc] += c * s0[node * W + cc];
    }
    double ue[V];
    exact_solution<D>(ep, &val[V], t, ue);
    for (int v = 0; v < V; ++v) {
      double diff = ue[v] - val[v];
      l2[v] += diff * diff * (w * vol);
      double ad = fabs(diff);
      li[v] = (ad > li[v] || ad != ad) ? ad : li[v];
    }
  }
  for (int v = 0; v < V; ++v) {
    s_red[tid] = l2[v];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (tid < w) s_red[tid] += s_red[tid + w];
      __syncthreads();
    }
    if (tid == 0) partial[(size_t)e * 2 * V + v] = s_red[0];
    __syncthreads();
    s_red[tid] = li[v];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (tid < w) s_red[tid] = nanmax(s_red[tid], s_red[tid + w]);
      __syncthreads();
    }
    if (tid == 0) partial[(size_t)e * 2 * V + V + v] = s_red[0];
    __syncthreads();
  }
}

// integrate_via_indices (analysis_dg3d.jl:364-386) of u and entropy_math
// (compressible_euler_3d.jl:1970-1976): partial[e] = {int u [V], int S}.
template <typename T, int D, int N>
__global__ __launch_bounds__(256) void integrals_kernel(const T *__restrict__ u,
                                                        const double *__restrict__ inv_jac,
                                                        int E, double gamma,
                                                        const double *__restrict__ weights,
                                                        double *__restrict__ partial) {
  constexpr int n = N + 1, NN = ipow_c(n, D), V = D + 2;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const double jac = 1.0 / inv_jac[e];
  const double vol = D == 3 ? jac * jac * jac : jac * jac;
  double acc[V + 1];
  for (int v = 0; v <= V; ++v) acc[v] = 0;
  for (int node = 0; node < NN; ++node) {
    int idx[3] = {node % n, (node / n) % n, node / (n * n)};
    double w = vol;
    for (int d = 0; d < D; ++d) w *= weights[idx[d]];
    const T *un = u + ((size_t)e * NN + node) * V;
    double m2 = 0;
    for (int d = 0; d < D; ++d) m2 += double(un[1 + d]) * double(un[1 + d]);
    double rho = un[0];
    double p = (gamma - 1) * (double(un[V - 1]) - 0.5 * m2 / rho);
    double s = log(p) - gamma * log(rho);
    double S = -s * rho / (gamma - 1);
    for (int v = 0; v < V; ++v) acc[v] += w * double(un[v]);
    acc[V] += w * S;
  }
  for (int v = 0; v <= V; ++v) partial[(size_t)e * (V + 1) + v] = acc[v];
}

// Column sums / maxima of a partial table [rows, cols] -> out[cols]; one block.
static __global__ __launch_bounds__(256) void reduce_columns_kernel(const double *__restrict__ partial,
                                                             int rows, int cols, int nsum,
                                                             double *__restrict__ out) {
  __shared__ double s_red[256];
  for (int c = 0; c < cols; ++c) {
    const bool is_sum = c < nsum;
    double acc = is_sum ? 0.0 : 0.0;
    for (int r = threadIdx.x; r < rows; r += 256) {
      double x = partial[(size_t)r * cols + c];
      acc = is_sum ? acc + x : nanmax(acc, x);
    }
    s_red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (threadIdx.x < w)
        s_red[threadIdx.x] = is_sum ? s_red[threadIdx.x] + s_red[threadIdx.x + w]
                                    : nanmax(s_red[threadIdx.x], s_red[threadIdx.x + w]);
      __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = s_red[0];
    __syncthreads();
  }
}

// Outer states of the boundaries for BoundaryConditionDirichlet(initial_condition)
// (equations.jl:159-182; calc_boundary_flux_by_direction!, dg_2d.jl:826-860):
// ghost[b][l] = u_exact(x, t) at face node l of boundary b. x = bx[b][l] are the
// element's face node coordinates (calc_boundary_node_coordinates!,
// containers_2d.jl:394-413), computed once at set-up; l orders the face's free
// indices with the lower dimension fastest. The time is tdev[0] + t * tdev[1]
// (a stage of a running step) or t.
template <typename T, int D, int N>
__global__ __launch_bounds__(256) void boundary_state_kernel(const double *__restrict__ bx, int nodes_total,
                                                             double t, const double *__restrict__ tdev,
                                                             ExactParams ep, T *__restrict__ ghost) {
  constexpr int V = D + 2;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nodes_total) return;
  const double tt = tdev ? (tdev[0] + t * tdev[1]) : t;
  double x[3], ub[V];
#pragma unroll
  for (int d = 0; d < D; ++d) x[d] = bx[(size_t)k * D + d];
  exact_solution<D>(ep, x, tt, ub);
#pragma unroll
  for (int v = 0; v < V; ++v) ghost[(size_t)k * V + v] = T(ub[v]);
}

// compute_coefficients! on the device (dg.jl:1293-1324, the reference's own
// GPU kernel, and init_elements! for the node coordinates,
// containers_3d.jl:87-128): u at every LGL node is the exact solution at
// x_d = center_d + jacobian * node_{i_d}, with the product and the sum rounded
// separately as the host computes them; jacobian = 1 / inverse_jacobian. One
// thread per node, state layout [E][n^D][V] in FP64.
template <int D, int N> struct NodesConst {
  double x[N + 1];
};
template <int D, int N>
__global__ __launch_bounds__(256) void initial_state_kernel(const double *__restrict__ centers,
                                                            const double *__restrict__ inv_jac, int E,
                                                            NodesConst<D, N> nodes, ExactParams ep, double t,
                                                            double *__restrict__ u) {
  constexpr int n = N + 1, NN = ipow_c(n, D), V = D + 2;
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)E * NN) return;
  const int e = (int)(q / NN), l = (int)(q - (long)e * NN);
  const double jac = 1.0 / inv_jac[e];
  double x[3], uu[V];
  int r = l;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const int i = r % n;
    r /= n;
    x[d] = __dadd_rn(centers[(size_t)D * e + d], __dmul_rn(jac, nodes.x[i]));
  }
  exact_solution<D>(ep, x, t, uu);
#pragma unroll
  for (int v = 0; v < V; ++v) u[(size_t)q * V + v] = uu[v];
}

// boundary_condition_slip_wall, TreeMesh methods (compressible_euler_2d.jl:322-420,
// compressible_euler_3d.jl:319-415): the flux at a wall face node from the
// inner state alone. The state is rotated to the outward normal n = +-e_o;
// the tangents of rotate_to_x fix the order of the kinetic-energy terms (3D
// o = x: (z, y), y: (x, z), z: (y, x); 2D the other axis). The wall pressure p*
// solves the pressure Riemann problem (Toro 2009, sec. 6.3.3): a rarefaction
// for an outward normal velocity v_n <= 0 (clipped at 0 when the base of the
// power is negative), a shock otherwise. The flux is p* in the normal momentum
// component and zero elsewhere, for either side of the element. FP64
// arithmetic in both precisions (a surface set), stored in T. One thread per
// boundary face node; bnd[b] = {element, 0-based direction}; the flux goes to
// face slot b of mflux, face node = the element kernel's line index (free
// indices, lower dimension fastest).
template <typename T, int D, int N>
__global__ __launch_bounds__(256) void slip_wall_flux_kernel(const T *__restrict__ u, const int32_t *__restrict__ bnd,
                                                             int B, double gamma, T *__restrict__ mflux) {
  constexpr int n = N + 1, V = D + 2, L = D == 3 ? n * n : n, NN = ipow_c(n, D);
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= B * L) return;
  const int b = k / L, fn = k - b * L;
  const int e = bnd[2 * b], dir = bnd[2 * b + 1], o = dir >> 1;
  const bool plus = (dir & 1) != 0;
  int idx[3] = {0, 0, 0};
  {
    int r = fn;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      if (d == o) {
        idx[d] = plus ? n - 1 : 0;
      } else {
        idx[d] = r % n;
        r /= n;
      }
    }
  }
  const int node = D == 3 ? idx[0] + n * (idx[1] + n * idx[2]) : idx[0] + n * idx[1];
  const T *q = u + ((size_t)e * NN + node) * V;
  const double rho = double(q[0]);
  const double mo = double(q[1 + o]);
  const double mn = plus ? mo : -mo;
  double mt0, mt1 = 0.0;
  if constexpr (D == 3) {
    mt0 = double(q[1 + (o == 0 ? 2 : (o == 1 ? 0 : 1))]);
    mt1 = double(q[1 + (o == 0 ? 1 : (o == 1 ? 2 : 0))]);
  } else {
    mt0 = double(q[1 + (1 - o)]);
  }
  const double vn = mn / rho;
  double ke2 = mn * vn;
  ke2 += mt0 * (mt0 / rho);
  if constexpr (D == 3) ke2 += mt1 * (mt1 / rho);
  const double p = (gamma - 1) * (double(q[V - 1]) - 0.5 * ke2);
  double p_star;
  if (vn <= 0) {
    const double c = sqrt(gamma * p / rho);
    const double base = 1 + 0.5 * (gamma - 1) * vn / c;
    p_star = base >= 0 ? p * pow(base, 2 * gamma * (1.0 / (gamma - 1))) : 0.0;
  } else {
    const double A = 2 / ((gamma + 1) * rho);
    const double Bc = p * (gamma - 1) / (gamma + 1);
    p_star = p + 0.5 * vn / A * (vn + sqrt(vn * vn + 4 * A * (p + Bc)));
  }
  T *f = mflux + (size_t)k * V;
#pragma unroll
  for (int v = 0; v < V; ++v) f[v] = T(0);
  f[1 + o] = T(p_star);
}

// L2 mortars (prolong2mortars! + calc_mortar_flux! + mortar_fluxes_to_elements!,
// dg_3d.jl:770-1335, dg_2d.jl:897-1243). Per mortar: the large element's face
// trace is interpolated to each small face (forward matrices, first free index
// first), the numerical flux is evaluated at the small-face nodes (the large
// side is the left state when large_sides == 1) and stored for the small
// element's face, and the large element's face flux is the sum of the reverse
// L2 projections of the small-face fluxes (3D order: upper left, upper right,
// lower left, lower right; 2D: upper, lower). One thread per small-face node
// (S = 2^(D-1) faces of n^(D-1) nodes), MPB mortars per 256-thread block.
// mort[m] = {small ids [S], large id, orientation, large_side};
// mflux[(slot * L + node) * V + v], slot = slot0 + m (S + 1) + face (S = large).
template <typename T, int N> struct MortarConst {
  T fu[(N + 1) * (N + 1)], fl[(N + 1) * (N + 1)], ru[(N + 1) * (N + 1)], rl[(N + 1) * (N + 1)];
};

template <typename T, int D, int N>
__global__ __launch_bounds__(256) void mortar_flux_kernel(const T *__restrict__ u, const int32_t *__restrict__ mort,
                                                          int M, int slot0, int surface_flux, T gamma,
                                                          MortarConst<T, N> mc, T *__restrict__ mflux) {
  constexpr int n = N + 1, V = D + 2, S = 1 << (D - 1), L = D == 3 ? n * n : n, NN = ipow_c(n, D);
  constexpr int TPM = S * L, MPB = 256 / TPM, W = S + 3;
  static_assert(MPB >= 1, "one mortar per block at least");
  __shared__ T s_ul[MPB][L][V];
  __shared__ T s_fs[MPB][S][L][V];
  const int tid = threadIdx.x, mm = tid / TPM, r = tid - mm * TPM;
  const int m = blockIdx.x * MPB + mm;
  const bool active = mm < MPB && m < M;
  const int32_t *md = mort + (size_t)(active ? m : 0) * W;
  const int large = md[S], o = md[S + 1];
  const bool large_left = md[S + 2] == 1;
  const int st[3] = {1, n, n * n};
  int fd0 = o == 0 ? 1 : 0, fd1 = o == 2 ? 1 : 2;  // free dimensions (3D); 2D: fd0 only
  if (D == 2) fd0 = 1 - o;
  // face node (i + n j) of an element's face on side idx_o
  auto node_of = [&](int fn, int idx_o) {
    const int i = fn % n, j = fn / n;
    int node = idx_o * st[o] + i * st[fd0];
    if (D == 3) node += j * st[fd1];
    return node;
  };
  if (active && r < L) {
    const T *src = u + ((size_t)large * NN + node_of(r, large_left ? n - 1 : 0)) * V;
#pragma unroll
    for (int v = 0; v < V; ++v) s_ul[mm][r][v] = src[v];
  }
  __syncthreads();
  if (active) {
    const int q = r / L, fn = r - q * L, i = fn % n, j = fn / n;
    const T *F0 = (q & 1) ? mc.fu : mc.fl;
    const T *F1 = (q >> 1) ? mc.fu : mc.fl;
    T up[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      if constexpr (D == 3) {
        T acc = 0;
        for (int b = 0; b < n; ++b) {
          T t = 0;
          for (int a2 = 0; a2 < n; ++a2) t += F0[i * n + a2] * s_ul[mm][a2 + n * b][v];
          acc += F1[j * n + b] * t;
        }
        up[v] = acc;
      } else {
        T acc = 0;
        for (int a2 = 0; a2 < n; ++a2) acc += F0[i * n + a2] * s_ul[mm][a2][v];
        up[v] = acc;
      }
    }
    const T *us = u + ((size_t)md[q] * NN + node_of(fn, large_left ? 0 : n - 1)) * V;
    T usv[V], f[V];
#pragma unroll
    for (int v = 0; v < V; ++v) usv[v] = us[v];
    const T *ul = large_left ? up : usv;
    const T *ur = large_left ? usv : up;
    if (o == 0) num_flux_c<T, D, 0, true>(surface_flux, ul, ur, gamma, f);
    else if (o == 1) num_flux_c<T, D, 1, true>(surface_flux, ul, ur, gamma, f);
    else if constexpr (D == 3) num_flux_c<T, D, 2, true>(surface_flux, ul, ur, gamma, f);
    T *dst = mflux + ((size_t)(slot0 + m * (S + 1) + q) * L + fn) * V;
#pragma unroll
    for (int v = 0; v < V; ++v) {
      dst[v] = f[v];
      s_fs[mm][q][fn][v] = f[v];
    }
  }
  __syncthreads();
  if (active && r < L) {
    const int i = r % n, j = r / n;
    T fl[V];
#pragma unroll
    for (int v = 0; v < V; ++v) fl[v] = 0;
    for (int k = 0; k < S; ++k) {
      const int q = D == 3 ? (k < 2 ? k + 2 : k - 2) : 1 - k;
      const T *R0 = (q & 1) ? mc.ru : mc.rl;
      const T *R1 = (q >> 1) ? mc.ru : mc.rl;
#pragma unroll
      for (int v = 0; v < V; ++v) {
        if constexpr (D == 3) {
          T acc = 0;
          for (int b = 0; b < n; ++b) {
            T t = 0;
            for (int a2 = 0; a2 < n; ++a2) t += R0[i * n + a2] * s_fs[mm][q][a2 + n * b][v];
            acc += R1[j * n + b] * t;
          }
          fl[v] += acc;
        } else {
          T acc = 0;
          for (int a2 = 0; a2 < n; ++a2) acc += R0[i * n + a2] * s_fs[mm][q][a2][v];
          fl[v] += acc;
        }
      }
    }
    T *dst = mflux + ((size_t)(slot0 + m * (S + 1) + S) * L + r) * V;
#pragma unroll
    for (int v = 0; v < V; ++v) dst[v] = fl[v];
  }
}

// HG smoothing over an adjacency list (interfaces and mortars; apply_smoothing!,
// indicators_3d.jl:125-173): alpha_e = max(alpha_raw_e, 0.5 max_nbr alpha_raw_nbr).
static __global__ void hg_smooth_adj_kernel(const double *__restrict__ alpha_raw, const int32_t *__restrict__ adj_ptr,
                                     const int32_t *__restrict__ adj, double *__restrict__ alpha, int E) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double m = alpha_raw[e];
  for (int k = adj_ptr[e]; k < adj_ptr[e + 1]; ++k) {
    const double an = 0.5 * alpha_raw[adj[k]];
    m = an > m ? an : m;
  }
  alpha[e] = m;
}

// IndicatorMax (indicators_3d.jl:239-259, indicators_2d.jl): alpha[e] = max over
// the nodes of element e of conserved variable var. One thread per element.
template <int D, int N>
__global__ __launch_bounds__(256) void indicator_max_kernel(const double *__restrict__ u, int E, int var,
                                                            double *__restrict__ alpha) {
  constexpr int n = N + 1, NN = ipow_c(n, D), V = D + 2;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const double *ue = u + (size_t)e * NN * V + var;
  double m = ue[0];
  for (int node = 1; node < NN; ++node) m = fmax(m, ue[(size_t)node * V]);
  alpha[e] = m;
}

// AMR solution transfer with the L2 adaptor (refine_element! / coarsen_elements!,
// amr_dg3d.jl:138-230, 337-430; amr_dg2d.jl): new element e is
//   kind 0      a copy of old element src,
//   kind 1 + q  child q of old element src: per dimension d the interpolation
//               forward_upper (bit d of q set) or forward_lower,
//   kind -1     the sum over the 2^D old elements src + q of the projections
//               reverse_upper / reverse_lower per dimension (bit d of q).
// mats = [forward_upper, forward_lower, reverse_upper, reverse_lower], each
// n x n row-major with rows = output nodes. One thread per new node; the
// tensor contraction runs x innermost, then y, then z.
template <int D, int N>
__global__ __launch_bounds__(256) void amr_transfer_kernel(const double *__restrict__ u_old, double *__restrict__ u_new,
                                                           const int32_t *__restrict__ kind,
                                                           const int32_t *__restrict__ src, int E_new,
                                                           const double *__restrict__ mats) {
  constexpr int n = N + 1, NN = ipow_c(n, D), V = D + 2;
  __shared__ double s_m[4 * n * n];
  for (int k = threadIdx.x; k < 4 * n * n; k += blockDim.x) s_m[k] = mats[k];
  __syncthreads();
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)E_new * NN) return;
  const int e = (int)(idx / NN), node = (int)(idx % NN);
  const int i = node % n, j = (node / n) % n, k = D == 3 ? node / (n * n) : 0;
  const int kd = kind[e], s0 = src[e];
  double out[V];
  for (int v = 0; v < V; ++v) out[v] = 0;
  if (kd == 0) {
    const double *ue = u_old + ((size_t)s0 * NN + node) * V;
    for (int v = 0; v < V; ++v) out[v] = ue[v];
  } else {
    const bool coarsen = kd < 0;
    const int nq = coarsen ? (1 << D) : 1;
    for (int qq = 0; qq < nq; ++qq) {
      const int q = coarsen ? qq : kd - 1;
      const double *ue = u_old + (size_t)(coarsen ? s0 + qq : s0) * NN * V;
      const int base = coarsen ? 2 : 0;
      const double *mx = &s_m[(base + ((q & 1) ? 0 : 1)) * n * n + i * n];
      const double *my = &s_m[(base + ((q & 2) ? 0 : 1)) * n * n + j * n];
      const double *mz = &s_m[(base + ((q & 4) ? 0 : 1)) * n * n + k * n];
      for (int c = 0; c < (D == 3 ? n : 1); ++c) {
        double t2[V];
        for (int v = 0; v < V; ++v) t2[v] = 0;
        for (int b = 0; b < n; ++b) {
          double t1[V];
          for (int v = 0; v < V; ++v) t1[v] = 0;
          for (int a = 0; a < n; ++a) {
            const double *un = ue + (size_t)(a + n * b + n * n * c) * V;
            for (int v = 0; v < V; ++v) t1[v] += mx[a] * un[v];
          }
          for (int v = 0; v < V; ++v) t2[v] += my[b] * t1[v];
        }
        if (D == 3) {
          for (int v = 0; v < V; ++v) out[v] += mz[c] * t2[v];
        } else {
          for (int v = 0; v < V; ++v) out[v] += t2[v];
        }
      }
    }
  }
  double *o = u_new + ((size_t)e * NN + node) * V;
  for (int v = 0; v < V; ++v) o[v] = out[v];
}

}  // namespace trixi_hip
