// dg_element.hpp -- the fused per-element DGSEM right-hand-side kernel.
//
// One launch computes, for a tile of EPB consecutive elements per workgroup,
// everything rhs_hyperbolic! (src/solvers/dgsem_tree/dg_2d.jl:113-186) does
// for a conforming periodic TreeMesh:
//   set_zero!                    src/solvers/solvers.jl:8-29  (du is written, not accumulated)
//   calc_volume_integral!        calc_volume_integral.jl:180-272
//     weak_form_kernel!          dg_3d.jl:133-164, dg_2d.jl:195-221
//     flux_differencing_kernel!  dg_3d.jl:166-214, dg_2d.jl:223-259
//     fv_kernel!/calcflux_fv!    dg_3d.jl:268-306, 359-389; dg_2d.jl:407-464
//   prolong2interfaces! +
//   calc_interface_flux!         dg_3d.jl:530-602, dg_2d.jl:515-632
//     (each element evaluates the numerical flux on its own 2D faces from its
//      own face trace and the neighbour's, read directly from u)
//   calc_surface_integral!       dg_3d.jl:1337-1394, dg_2d.jl:1245-1287
//   apply_jacobian!              dg_3d.jl:1396-1415
//   calc_sources!                dg_3d.jl:1417-1438
//
// Work decomposition: after the tile is staged in LDS, each thread owns one
// node LINE (fixed direction o, the other indices fixed). For that line it
// evaluates every two-point volume flux between line nodes exactly once
// (n(n-1)/2 fluxes, the symmetric update of flux_differencing_kernel!), plus
// the two surface fluxes at the line ends, keeps the line's du in registers and
// adds it to the LDS accumulator. Directions are processed one after another
// with a barrier between them, so every du entry has exactly one writer per
// phase: no atomics, bitwise reproducible.
#pragma once
#include "dg_common.hpp"

namespace trixi_hip {

template <typename T, int N> struct DGConst {
  static constexpr int n = N + 1;
  T Dsplit[n * n];  // derivative_split, row-major [a*n + b] = Dsplit[a, b]
  T Dhat[n * n];    // derivative_hat
  T inv_w[n];       // inverse_weights
  double nodes[n];  // LGL nodes (coordinates of sources)
};

template <typename T> struct ElemArgs {
  const T *__restrict__ u;               // [V, NN, E]
  T *__restrict__ du;                    // [V, NN, E]
  const int32_t *__restrict__ nbr;       // [2D, E] neighbour element per direction
  const double *__restrict__ inv_jac;    // [E]
  const double *__restrict__ centers;    // [D, E]
  const double *__restrict__ alpha_raw;  // [E] HG indicator before smoothing (HG only)
  double *__restrict__ alpha_out;        // [E] smoothed blending factor (HG only, may be null)
  const double *__restrict__ tdev;       // device {t, dt} (may be null: use t)
  double t;                              // time, or c_stage when tdev != null
  int E;
  int surface_flux;
  int fv_flux;
  int source;
  int alpha_smooth;
  int xcd_remap;
  T gamma;
};

// Workgroups are dispatched round-robin over the 8 XCDs. Remap so that each
// XCD walks a contiguous range of the (space-filling, depth-first) element
// order: neighbours of an element then mostly sit in the same XCD's L2.
__device__ __forceinline__ int xcd_remap_block(int b, int nb) {
  const int X = 8;
  int q = nb / X, r = nb % X;
  int x = b % X, i = b / X;
  return (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// Cooperative contiguous copy global <-> LDS with 16-byte accesses when the
// tile size allows it.
template <typename T, int NT>
__device__ __forceinline__ void tile_copy(T *__restrict__ dst, const T *__restrict__ src, int cnt,
                                          bool vec_ok) {
  const int tid = threadIdx.x;
  if (vec_ok) {
    const int c4 = cnt * (int)sizeof(T) / 16;
    const float4 *s4 = reinterpret_cast<const float4 *>(src);
    float4 *d4 = reinterpret_cast<float4 *>(dst);
    for (int i = tid; i < c4; i += NT) d4[i] = s4[i];
  } else {
    for (int i = tid; i < cnt; i += NT) dst[i] = src[i];
  }
}

// One direction O of the line phase.
template <typename T, int D, int N, int VOL, int EPB, int NT, int O>
__device__ __forceinline__ void line_phase(const ElemArgs<T> &a, const DGConst<T, N> &k, int e0,
                                           int ne, const T *__restrict__ s_u,
                                           const T *__restrict__ s_w, T *__restrict__ s_du,
                                           const double *__restrict__ s_alpha, int SP) {
  constexpr int n = N + 1, NN = ipow_c(n, D), L = ipow_c(n, D - 1), V = D + 2;
  constexpr int st = ipow_c(n, O);
  constexpr bool RAN = (VOL == VOL_FD_RANOCHA || VOL == VOL_HG_RANOCHA);
  constexpr bool CEN = (VOL == VOL_FD_CENTRAL || VOL == VOL_HG_CENTRAL);
  constexpr bool HG = (VOL == VOL_HG_RANOCHA || VOL == VOL_HG_CENTRAL);
  const T gamma = a.gamma;
  const T gm1 = gamma - T(1);
  const T igm1 = T(1) / gm1;

  for (int l = threadIdx.x; l < EPB * L; l += NT) {
    const int el = l / L;
    if (el >= ne) continue;
    const int line = l - el * L;
    const int base = (line % st) + (line / st) * st * n;
    const int q0 = el * NN + base;  // tile-local node index of line position 0
    const int e = e0 + el;

    T acc[n][V];
#pragma unroll
    for (int i = 0; i < n; ++i)
#pragma unroll
      for (int v = 0; v < V; ++v) acc[i][v] = T(0);

    // ---------------- volume integral ----------------
    bool do_fd = (VOL != VOL_PURE_FV);
    double alpha = 0.0;
    bool blend = false;
    if constexpr (HG) {
      alpha = s_alpha[el];
      const double atol = sizeof(T) == 8 ? 1.8189894035458565e-12 : 1.1920928955078125e-5;
      blend = !(alpha <= atol && alpha >= -atol);
    }
    if constexpr (VOL == VOL_WEAK) {
#pragma unroll
      for (int b = 0; b < n; ++b) {
        ConsP<T, D> cb;
        consp_from_cons<T, D>(&s_u[(q0 + b * st) * V], gm1, cb);
        T fb[V];
        phys_flux<T, D, O>(cb, fb);
#pragma unroll
        for (int i = 0; i < n; ++i) {
          const T dh = k.Dhat[i * n + b];
#pragma unroll
          for (int v = 0; v < V; ++v) acc[i][v] = __builtin_fma(dh, fb[v], acc[i][v]);
        }
      }
    }
    if constexpr (RAN) {
      NodeR<T, D> w[n];
#pragma unroll
      for (int i = 0; i < n; ++i) {
        const int q = q0 + i * st;
        w[i].rho = s_w[0 * SP + q];
        w[i].lrho = s_w[1 * SP + q];
        w[i].beta = s_w[2 * SP + q];
        w[i].lbeta = s_w[3 * SP + q];
#pragma unroll
        for (int d = 0; d < D; ++d) w[i].h[d] = s_w[(4 + d) * SP + q];
        w[i].hp = s_w[(4 + D) * SP + q];
      }
#pragma unroll
      for (int i = 0; i < n; ++i)
#pragma unroll
        for (int j = i + 1; j < n; ++j) {
          T f[V];
          flux_ranocha_r<T, D, O>(w[i], w[j], igm1, f);
          const T di = k.Dsplit[i * n + j], dj = k.Dsplit[j * n + i];
#pragma unroll
          for (int v = 0; v < V; ++v) {
            acc[i][v] = __builtin_fma(di, f[v], acc[i][v]);
            acc[j][v] = __builtin_fma(dj, f[v], acc[j][v]);
          }
        }
    }
    if constexpr (CEN) {
      T fph[n][V];
#pragma unroll
      for (int i = 0; i < n; ++i) {
        ConsP<T, D> ci;
        consp_from_cons<T, D>(&s_u[(q0 + i * st) * V], gm1, ci);
        phys_flux<T, D, O>(ci, fph[i]);
      }
#pragma unroll
      for (int i = 0; i < n; ++i)
#pragma unroll
        for (int j = i + 1; j < n; ++j) {
          const T di = k.Dsplit[i * n + j], dj = k.Dsplit[j * n + i];
#pragma unroll
          for (int v = 0; v < V; ++v) {
            T f = T(0.5) * (fph[i][v] + fph[j][v]);
            acc[i][v] = __builtin_fma(di, f, acc[i][v]);
            acc[j][v] = __builtin_fma(dj, f, acc[j][v]);
          }
        }
    }
    (void)do_fd;
    if constexpr (HG || VOL == VOL_PURE_FV) {
      T afv = T(1);
      bool run_fv = true;
      if constexpr (HG) {
        run_fv = blend;
        if (blend) {
          const T s = T(1.0 - alpha);
#pragma unroll
          for (int i = 0; i < n; ++i)
#pragma unroll
            for (int v = 0; v < V; ++v) acc[i][v] *= s;
        }
        afv = T(alpha);
      }
      if (run_fv) {
        // first-order FV on the LGL subcells along this line; the subcell
        // fluxes at the element faces are zero (create_f_threaded, dg_3d.jl:11-54)
        T fprev[V];
#pragma unroll
        for (int v = 0; v < V; ++v) fprev[v] = T(0);
#pragma unroll
        for (int i = 0; i < n; ++i) {
          T fnext[V];
          if (i < n - 1) {
            num_flux_c<T, D, O>(a.fv_flux, &s_u[(q0 + i * st) * V], &s_u[(q0 + (i + 1) * st) * V],
                                gamma, fnext);
          } else {
#pragma unroll
            for (int v = 0; v < V; ++v) fnext[v] = T(0);
          }
          const T c = afv * k.inv_w[i];
#pragma unroll
          for (int v = 0; v < V; ++v) {
            acc[i][v] = __builtin_fma(c, fnext[v] - fprev[v], acc[i][v]);
            fprev[v] = fnext[v];
          }
        }
      }
    }

    // ---------------- interface fluxes + surface integral ----------------
    const T factor = k.inv_w[0];  // inverse_weights[1] of the reference
    {
      // -O face: this element is the right state of the interface
      const int nb = a.nbr[(2 * O) + 2 * D * e];
      const T *un = a.u + ((size_t)nb * NN + base + (n - 1) * st) * V;
      T ul[V], ur[V], F[V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        ul[v] = un[v];
        ur[v] = s_u[q0 * V + v];
      }
      num_flux_c<T, D, O>(a.surface_flux, ul, ur, gamma, F);
#pragma unroll
      for (int v = 0; v < V; ++v) acc[0][v] = __builtin_fma(-F[v], factor, acc[0][v]);
    }
    {
      // +O face: this element is the left state
      const int nb = a.nbr[(2 * O + 1) + 2 * D * e];
      const T *un = a.u + ((size_t)nb * NN + base) * V;
      T ul[V], ur[V], F[V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        ul[v] = s_u[(q0 + (n - 1) * st) * V + v];
        ur[v] = un[v];
      }
      num_flux_c<T, D, O>(a.surface_flux, ul, ur, gamma, F);
#pragma unroll
      for (int v = 0; v < V; ++v) acc[n - 1][v] = __builtin_fma(F[v], factor, acc[n - 1][v]);
    }

    // ---------------- accumulate into LDS ----------------
#pragma unroll
    for (int i = 0; i < n; ++i) {
      T *dst = &s_du[(q0 + i * st) * V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        if constexpr (O == 0) dst[v] = acc[i][v];
        else dst[v] += acc[i][v];
      }
    }
  }
}

template <typename T, int D, int N, int VOL, int EPB, int NT>
__global__ __launch_bounds__(NT) void dg_element_kernel(ElemArgs<T> a, DGConst<T, N> k) {
  constexpr int n = N + 1, NN = ipow_c(n, D), V = D + 2;
  constexpr bool RAN = (VOL == VOL_FD_RANOCHA || VOL == VOL_HG_RANOCHA);
  constexpr bool HG = (VOL == VOL_HG_RANOCHA || VOL == VOL_HG_CENTRAL);
  constexpr int NW = RAN ? nodeR_count(D) : 1;
  // pad the SoA stride so successive variables start in different LDS banks
  constexpr int SP = EPB * NN + 1;
  __shared__ __attribute__((aligned(16))) T s_u[EPB * NN * V];
  __shared__ __attribute__((aligned(16))) T s_du[EPB * NN * V];
  __shared__ T s_w[NW * SP];
  __shared__ double s_alpha[HG ? EPB : 1];

  const int tid = threadIdx.x;
  const int blk = a.xcd_remap ? xcd_remap_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int e0 = blk * EPB;
  if (e0 >= a.E) return;  // uniform per block
  const int ne = min(EPB, a.E - e0);
  constexpr bool vec_ok_c = ((NN * V * (int)sizeof(T)) % 16) == 0;

  // phase A: stage the tile
  tile_copy<T, NT>(s_u, a.u + (size_t)e0 * NN * V, ne * NN * V, vec_ok_c);
  if constexpr (HG) {
    if (tid < ne) {
      const int e = e0 + tid;
      double al = a.alpha_raw[e];
      if (a.alpha_smooth) {
        // apply_smoothing! (indicators_2d.jl:104-135, indicators_3d.jl:125-173):
        // alpha_e = max(alpha_e, 0.5 alpha_nbr) over the element's interfaces
        double m = al;
#pragma unroll
        for (int dir = 0; dir < 2 * D; ++dir) {
          double an = 0.5 * a.alpha_raw[a.nbr[dir + 2 * D * e]];
          m = an > m ? an : m;
        }
        al = m;
      }
      s_alpha[tid] = al;
      if (a.alpha_out) a.alpha_out[e] = al;
    }
  }
  __syncthreads();

  // phase A2: per-node primitive data for flux_ranocha
  if constexpr (RAN) {
    const T gm1 = a.gamma - T(1);
    for (int q = tid; q < ne * NN; q += NT) {
      NodeR<T, D> w;
      nodeR_from_cons<T, D>(&s_u[q * V], gm1, w);
      s_w[0 * SP + q] = w.rho;
      s_w[1 * SP + q] = w.lrho;
      s_w[2 * SP + q] = w.beta;
      s_w[3 * SP + q] = w.lbeta;
#pragma unroll
      for (int d = 0; d < D; ++d) s_w[(4 + d) * SP + q] = w.h[d];
      s_w[(4 + D) * SP + q] = w.hp;
    }
    __syncthreads();
  }

  // phase B: lines, one direction at a time
  line_phase<T, D, N, VOL, EPB, NT, 0>(a, k, e0, ne, s_u, s_w, s_du, s_alpha, SP);
  __syncthreads();
  line_phase<T, D, N, VOL, EPB, NT, 1>(a, k, e0, ne, s_u, s_w, s_du, s_alpha, SP);
  __syncthreads();
  if constexpr (D == 3) {
    line_phase<T, D, N, VOL, EPB, NT, 2>(a, k, e0, ne, s_u, s_w, s_du, s_alpha, SP);
    __syncthreads();
  }

  // phase C: Jacobian and sources, node-wise in LDS
  const double tt = a.tdev ? (a.tdev[0] + a.t * a.tdev[1]) : a.t;
  for (int q = tid; q < ne * NN; q += NT) {
    const int el = q / NN, node = q - el * NN;
    const int e = e0 + el;
    const double ij = a.inv_jac[e];
    const T fac = T(-ij);
    T *d = &s_du[q * V];
#pragma unroll
    for (int v = 0; v < V; ++v) d[v] *= fac;
    if (a.source) {
      const double jac = 1.0 / ij;
      T x[D];
      int rem = node;
#pragma unroll
      for (int dd = 0; dd < D; ++dd) {
        const int idx = rem % n;
        rem /= n;
        x[dd] = T(a.centers[dd + D * e] + jac * k.nodes[idx]);
      }
      T s[V];
      source_convergence<T, D>(x, T(tt), a.gamma, s);
#pragma unroll
      for (int v = 0; v < V; ++v) d[v] += s[v];
    }
  }
  __syncthreads();
  tile_copy<T, NT>(a.du + (size_t)e0 * NN * V, s_du, ne * NN * V, vec_ok_c);
}

}  // namespace trixi_hip
