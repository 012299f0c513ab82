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
// Work decomposition. The workgroup has D groups of WPD wavefronts, one group
// per coordinate direction. Phase A stages the tile (16-byte loads) and
// computes per-node primitive data into LDS. Phase B: every lane of group O
// owns one node LINE in direction O; it evaluates each two-point volume flux
// of the line exactly once (n(n-1)/2 fluxes, the symmetric update of
// flux_differencing_kernel!) plus the two surface fluxes at the line ends, and
// writes the line's du into the LDS buffer of direction O. Phase C sums the D
// buffers per node, applies the Jacobian and sources, and stores with 16-byte
// writes. Every LDS word has one writer per phase: no atomics, bitwise
// reproducible results.
//
// LDS layout: structure-of-arrays per variable, node index q = el*NN + node.
// For the 3D polydeg-3 tile of 4 elements the node index is XOR-swizzled
// (i, j, k) -> (i^el, j^el, k^el) and each direction group uses its own lane
// order, which makes every line access of every direction, and the sequential
// accesses of phases A and C, free of bank conflicts for 32 and for 64 banks.
#pragma once
#include "dg_common.hpp"
#include "indicator2d.hpp"

namespace trixi_hip {

template <typename T, int N> struct DGConst {
  static constexpr int n = N + 1;
  T Dsplit[n * n];  // derivative_split, row-major [a*n + b] = Dsplit[a, b]
  T Dsplit2[n * n];  // 2 * derivative_split (halved energy flux of flux_ranocha_fast)
  T Dhat[n * n];    // derivative_hat
  T inv_w[n];       // inverse_weights
  double nodes[n];  // LGL nodes (coordinates of sources)
};

template <typename T> struct ElemArgs {
  const T *__restrict__ u;               // [V, NN, E]
  T *__restrict__ du;                    // [V, NN, E]
  const int32_t *__restrict__ nbr;       // [2D, E] neighbour element per direction; -(b + 1) at boundary b
  const T *__restrict__ ghost;           // [B + F, n^(D-1), V] outer states of the boundaries (slots < nbnd)
  const T *__restrict__ mflux;           // [B + F, n^(D-1), V] given face fluxes of mortar faces (slots >= nbnd)
  int nbnd;                              // B: face slots below are boundaries, the others mortar faces
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
  // fused 2N stage (STAGE variants, methods_2N.jl:153-157): instead of storing
  // du, u_tmp = du - a_stage u_tmp and u_reg_out = u_reg_in + b_dt u_tmp, with
  // the arithmetic in FP64 and u_tmp stored in the RHS precision T. In FP32
  // that rounds the register once per stage (relative 2^-24 of a value that
  // is itself a combination of FP32 RHS values, whose own error is larger);
  // the equal-L2 run of the 3D MMS case bounds the total effect (density L2
  // errors of the FP32 and FP64 runs differ by 3.6e-7 after 100 steps).
  // b_dt = b_stage * tdev[1] (or b_stage without tdev); u32_out receives the
  // FP32 shadow of u_reg_out (FP32 mode, may be null). u_reg_in == u_reg_out
  // is allowed when the stage reads its state from u. With u64 != null (FP32
  // stage on a mesh without mortars and slip walls) the stage reads its tile
  // and the neighbour face states from the FP64 register u64 (= u_reg_in of
  // every element, so u_reg_out must be another buffer), rounded to FP32 on
  // load: the values the shadow would hold, bit for bit, without its 4 bytes
  // of writes and reads per value.
  T *u_tmp;
  const double *u_reg_in;
  double *u_reg_out;
  float *u32_out;
  const double *u64;
  double a_stage, b_stage;
  int first_stage;
  T gamma;
  T inv_gamma_m1;  // 1 / (gamma - 1), rounded once in T
  T sqrt_gamma;
  // fused last stage of a step (STAGE variants): with speed_partial != null
  // the epilogue also reduces the wave speed of the new state (the max_dt input
  // of the next step, max_speed_nodes_kernel's arithmetic) to one FP64 partial
  // per workgroup, speed_partial[workgroup]; gamma64 is gamma in FP64
  double *speed_partial;
  double gamma64;
  // fused stage of a 2D polydeg-3 shock-capturing variant: with ind_next != 0
  // the epilogue also computes the raw blending factors of the new state (the
  // next stage's indicator: HG, or the learned one with ind.net) into
  // ind.alpha_raw, with indicator_2d_p3_kernel's arithmetic
  int ind_next;
  Ind2DArgs ind;
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

template <int BYTES> struct ChunkOf;
template <> struct ChunkOf<16> { using type = float4; };
template <> struct ChunkOf<8> { using type = float2; };
template <> struct ChunkOf<4> { using type = float; };

// A full tile of EPB elements moves between global memory and LDS in chunks of
// CB bytes, the largest of 16/8/4 that divides the tile size (tile offsets are
// multiples of the tile size and the arrays are 16-byte aligned).
template <typename T, int D, int N, int EPB> struct TileIO {
  static constexpr int NN = ipow_c(N + 1, D), V = D + 2;
  static constexpr int TB = EPB * NN * V * (int)sizeof(T);
  static constexpr int CB = TB % 16 == 0 ? 16 : TB % 8 == 0 ? 8 : 4;
  using C = typename ChunkOf<CB>::type;
  static constexpr int NCH = TB / CB;
};

// LDS index map and lane -> line assignment (see the header comment).
template <int D, int N, int EPB> struct Layout {
  static constexpr int n = N + 1, NN = ipow_c(n, D), L = ipow_c(n, D - 1);
  static constexpr bool SWZ = (D == 3 && N == 3 && EPB == 4);
  static constexpr int SIZE = SWZ ? EPB * NN : EPB * NN + (EPB * NN) / 32 + 1;
  static __device__ __forceinline__ int phys(int q) {
    if constexpr (SWZ) {
      const int el = q >> 6, i = q & 3, j = (q >> 2) & 3, k = (q >> 4) & 3;
      return (el << 6) | (i ^ el) | ((j ^ el) << 2) | ((k ^ el) << 4);
    } else {
      return q + (q >> 5);
    }
  }
  // slot of node q0 + i n^O of a line that starts at q0 (index 0 along O),
  // given p0 = phys(q0): in the swizzled layout only the O-th index field
  // changes, from 0 ^ el to i ^ el
  template <int O> static __device__ __forceinline__ int phys_line(int p0, int q0, int i) {
    if constexpr (SWZ) return p0 ^ (i << (2 * O));
    else return phys(q0 + i * ipow_c(n, O));
  }
  // line = r0 + n*r1 with (r0, r1) the two free indices in increasing dimension order
  template <int O> static __device__ __forceinline__ void lane_line(int ll, int &el, int &line) {
    if constexpr (SWZ) {
      if constexpr (O == 0) {
        el = ll & 3;
        line = ((ll >> 2) & 3) + 4 * (ll >> 4);  // (j, k)
      } else if constexpr (O == 1) {
        el = ll & 3;
        line = ((ll >> 2) & 3) + 4 * (ll >> 4);  // (i, k)
      } else {
        el = ll >> 4;
        line = ll & 15;  // (i, j)
      }
    } else {
      el = ll / L;
      line = ll - el * L;
    }
  }
};

template <int VOL> struct VolTraits {
  static constexpr bool RAN = (VOL == VOL_FD_RANOCHA || VOL == VOL_HG_RANOCHA);
  static constexpr bool CEN = (VOL == VOL_FD_CENTRAL || VOL == VOL_HG_CENTRAL);
  static constexpr bool HG = (VOL == VOL_HG_RANOCHA || VOL == VOL_HG_CENTRAL);
  static constexpr bool FV = HG || VOL == VOL_PURE_FV;
  // conservative state needed in LDS (weak form, central flux, FV subcells)
  static constexpr bool CONS = !RAN || HG;
};

// Per-lane line context: which line the lane owns and the four face states
// (neighbour states loaded from global memory at kernel entry, own states read
// from the staged tile) so that no load latency sits inside the line phase.
template <typename T, int V> struct LineCtx {
  int el, line, base;
  bool valid;
  int fm, fp;  // BND variants: offsets of the given fluxes of mortar faces at the -O / +O face, or -1
  T ul_m[V], ur_p[V], own0[V], own1[V];
};

// Line of this lane in direction o (wave-uniform, runtime) and the ids of the
// two neighbours across its end faces. Branch-free over o, so that these loads
// and the tile loads are issued back to back. Lanes without a line (or on a
// missing element of a partial tile) use the tile's last element, so every
// lane issues the same in-bounds loads.
template <typename T, int D, int N, int EPB>
__device__ __forceinline__ void line_ids(const ElemArgs<T> &a, int e0, int ne, int ll, int o, LineCtx<T, D + 2> &c,
                                         int &nbm, int &nbp) {
  using LY = Layout<D, N, EPB>;
  constexpr int n = N + 1;
  int el0, line0, el2, line2;
  LY::template lane_line<0>(ll, el0, line0);  // directions 0 and 1 share the lane map
  LY::template lane_line<D - 1>(ll, el2, line2);
  c.el = o == D - 1 ? el2 : el0;
  c.line = o == D - 1 ? line2 : line0;
  c.valid = (ll < EPB * LY::L) && (c.el < ne);
  // line = r0 + n r1 over the two free indices; base = node of index 0 along o
  const int b0 = c.line * n;                                // o = 0: (j, k) -> j n + k n^2
  const int b1 = (c.line % n) + (c.line / n) * n * n;       // o = 1: (i, k) -> i + k n^2
  const int b2 = c.line;                                    // o = 2: (i, j) -> i + j n
  c.base = o == 0 ? b0 : (o == 1 ? b1 : b2);
  const int e = e0 + min(c.el, ne - 1);
  nbm = a.nbr[(2 * o) + 2 * D * e];
  nbp = a.nbr[(2 * o + 1) + 2 * D * e];
}

// Neighbour face states of the line in direction o (consumed in phase B). With
// BND (a mesh with boundaries), the outer state at a boundary face comes from
// the ghost record of that boundary (face node = the lane's line index);
// periodic meshes use the instantiation without that select.
template <typename T, int D, int N, int EPB, bool BND>
__device__ __forceinline__ void line_loads(const ElemArgs<T> &a, LineCtx<T, D + 2> &c, int o, int nbm, int nbp,
                                           const double *u64) {
  using LY = Layout<D, N, EPB>;
  constexpr int n = N + 1, NN = LY::NN, V = D + 2, L = LY::L;
  const int st = o == 0 ? 1 : (o == 1 ? n : n * n);
  if constexpr (sizeof(T) == 4) {
    if (u64) {  // FP32 stage reading the FP64 register (wave-uniform)
      // With BND only boundary faces (no mortar or slip-wall slots) occur
      // here: their outer state is the ghost record, loaded in FP32 beside the
      // (clamped, in-bounds) register load and selected per lane.
      const bool im = !BND || nbm >= 0, ip = !BND || nbp >= 0;
      const double *vm = u64 + ((size_t)(im ? nbm : 0) * NN + c.base + (n - 1) * st) * V;
      const double *vp = u64 + ((size_t)(ip ? nbp : 0) * NN + c.base) * V;
      if constexpr (BND) {
        const T *gm = a.ghost + ((size_t)(im ? 0 : -nbm - 1) * L + c.line) * V;
        const T *gp = a.ghost + ((size_t)(ip ? 0 : -nbp - 1) * L + c.line) * V;
        c.fm = -1;
        c.fp = -1;
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const T xm = T(vm[v]), gmv = gm[v], xp = T(vp[v]), gpv = gp[v];
          c.ul_m[v] = im ? xm : gmv;
          c.ur_p[v] = ip ? xp : gpv;
        }
      } else {
#pragma unroll
        for (int v = 0; v < V; ++v) {
          c.ul_m[v] = T(vm[v]);
          c.ur_p[v] = T(vp[v]);
        }
      }
      return;
    }
  }
  const T *un_m, *un_p;
  if constexpr (BND) {
    // a face slot (-nb - 1) below nbnd is a boundary (its ghost state), above a
    // mortar face (its flux is given; the ghost slot only keeps the load valid)
    un_m = nbm >= 0 ? a.u + ((size_t)nbm * NN + c.base + (n - 1) * st) * V
                    : a.ghost + ((size_t)(-nbm - 1) * L + c.line) * V;
    un_p = nbp >= 0 ? a.u + ((size_t)nbp * NN + c.base) * V : a.ghost + ((size_t)(-nbp - 1) * L + c.line) * V;
    c.fm = (nbm < 0 && -nbm - 1 >= a.nbnd) ? ((-nbm - 1) * L + c.line) * V : -1;
    c.fp = (nbp < 0 && -nbp - 1 >= a.nbnd) ? ((-nbp - 1) * L + c.line) * V : -1;
  } else {
    un_m = a.u + ((size_t)nbm * NN + c.base + (n - 1) * st) * V;
    un_p = a.u + ((size_t)nbp * NN + c.base) * V;
  }
#pragma unroll
  for (int v = 0; v < V; ++v) {
    c.ul_m[v] = un_m[v];
    c.ur_p[v] = un_p[v];
  }
}

// own face states from the staged tile (AoS [q][v])
template <typename T, int D, int N, int EPB, int O>
__device__ __forceinline__ void line_own(const T *__restrict__ staged, LineCtx<T, D + 2> &c) {
  using LY = Layout<D, N, EPB>;
  constexpr int n = N + 1, NN = LY::NN, V = D + 2;
  constexpr int st = ipow_c(n, O);
  if (!c.valid) return;
  const int q0 = c.el * NN + c.base;
#pragma unroll
  for (int v = 0; v < V; ++v) {
    c.own0[v] = staged[q0 * V + v];
    c.own1[v] = staged[(q0 + (n - 1) * st) * V + v];
  }
}

// One line of direction O: volume terms, the two surface terms, LDS write.
template <typename T, int D, int N, int VOL, int EPB, int O, bool BND>
__device__ __forceinline__ void line_work(const ElemArgs<T> &a, const DGConst<T, N> &k, const LineCtx<T, D + 2> &c,
                                          const T *__restrict__ s_w, const T *__restrict__ s_c,
                                          T *__restrict__ s_d, const double *__restrict__ s_alpha) {
  using LY = Layout<D, N, EPB>;
  using VT = VolTraits<VOL>;
  constexpr int n = N + 1, NN = LY::NN, V = D + 2, SZ = LY::SIZE;
  constexpr int st = ipow_c(n, O);
  if (!c.valid) return;
  const int el = c.el;
  const int q0 = el * NN + c.base;
  const T gamma = a.gamma;
  const T gm1 = gamma - T(1);
  const T igm1 = a.inv_gamma_m1;
  // node records in the square-root representation (FP32, polydeg <= 3); the
  // own face states' primitives and sound speeds then come from the records of
  // the two end nodes
  constexpr bool SQ = node_sqrt_rep<T, N>();
  constexpr bool OWN_FROM_W = VT::RAN && SQ;
  NodeR<T, D> wend[2];

  int pq[n];
  pq[0] = LY::phys(q0);
#pragma unroll
  for (int i = 1; i < n; ++i) pq[i] = LY::template phys_line<O>(pq[0], q0, i);

  T acc[n][V];
#pragma unroll
  for (int i = 0; i < n; ++i)
#pragma unroll
    for (int v = 0; v < V; ++v) acc[i][v] = T(0);

  double alpha = 0.0;
  bool blend = false;
  if constexpr (VT::HG) {
    alpha = s_alpha[el];
    const double atol = sizeof(T) == 8 ? 1.8189894035458565e-12 : 1.1920928955078125e-5;
    blend = !(alpha <= atol && alpha >= -atol);
  }

  // ---------------- interface fluxes + surface integral ----------------
  // (evaluated after the volume integral, see the end of this function)
  const T factor = k.inv_w[0];  // inverse_weights[1] of the reference
  // mortar faces (BND variants): the flux computed by mortar_flux_kernel
  auto given_flux = [&](int off, T *F) {
    if constexpr (BND) {
      if (off >= 0) {
#pragma unroll
        for (int v = 0; v < V; ++v) F[v] = a.mflux[off + v];
      }
    }
  };
  auto faces = [&]() {
    T F[V];
    bool done = false;
    if constexpr (OWN_FROM_W) {
      if (a.surface_flux == FLUX_LLF || a.surface_flux == FLUX_LLF_NAIVE) {
        const bool naive = a.surface_flux == FLUX_LLF_NAIVE;
        ConsP<T, D> own_face[2];
        T own_c[2];
        consp_from_node<D>(c.own0, wend[0], a.sqrt_gamma, own_face[0], own_c[0]);
        consp_from_node<D>(c.own1, wend[1], a.sqrt_gamma, own_face[1], own_c[1]);
        // twice the LLF flux with half the weight (bitwise the same terms);
        // a given mortar flux is doubled to match
        const T hfactor = T(0.5) * factor;
        auto given_flux2 = [&](int off, T *F2) {
          if constexpr (BND) {
            if (off >= 0) {
#pragma unroll
              for (int v = 0; v < V; ++v) F2[v] = T(2) * a.mflux[off + v];
            }
          }
        };
        ConsP<T, D> nb;
        consp_from_cons<T, D>(c.ul_m, gm1, nb);  // -O face: element is the right state
        flux_llf_pc2<T, D, O>(nb, own_face[0], Prec<T>::sqrt_(gamma * nb.p * nb.irho), own_c[0], naive, F);
        given_flux2(c.fm, F);
#pragma unroll
        for (int v = 0; v < V; ++v) acc[0][v] = tfma(-F[v], hfactor, acc[0][v]);
        consp_from_cons<T, D>(c.ur_p, gm1, nb);  // +O face: element is the left state
        flux_llf_pc2<T, D, O>(own_face[1], nb, own_c[1], Prec<T>::sqrt_(gamma * nb.p * nb.irho), naive, F);
        given_flux2(c.fp, F);
#pragma unroll
        for (int v = 0; v < V; ++v) acc[n - 1][v] = tfma(F[v], hfactor, acc[n - 1][v]);
        done = true;
      }
    }
    if (!done) {
      num_flux_c<T, D, O, VOL == VOL_PURE_FV>(a.surface_flux, c.ul_m, c.own0, gamma, F);  // -O face: element is the right state
      given_flux(c.fm, F);
#pragma unroll
      for (int v = 0; v < V; ++v) acc[0][v] = tfma(-F[v], factor, acc[0][v]);
      num_flux_c<T, D, O, VOL == VOL_PURE_FV>(a.surface_flux, c.own1, c.ur_p, gamma, F);  // +O face: element is the left state
      given_flux(c.fp, F);
#pragma unroll
      for (int v = 0; v < V; ++v) acc[n - 1][v] = tfma(F[v], factor, acc[n - 1][v]);
    }
  };

  // ---------------- volume integral ----------------
  if constexpr (VOL == VOL_WEAK) {
#pragma unroll
    for (int b = 0; b < n; ++b) {
      T cb_u[V];
#pragma unroll
      for (int v = 0; v < V; ++v) cb_u[v] = s_c[v * SZ + pq[b]];
      ConsP<T, D> cb;
      consp_from_cons<T, D>(cb_u, gm1, cb);
      T fb[V];
      phys_flux<T, D, O>(cb, fb);
#pragma unroll
      for (int i = 0; i < n; ++i) {
        const T dh = k.Dhat[i * n + b];
#pragma unroll
        for (int v = 0; v < V; ++v) acc[i][v] = tfma(dh, fb[v], acc[i][v]);
      }
    }
  }
  if constexpr (VT::RAN) {
    NodeR<T, D> w[n];
    auto load_w = [&](int i) {
      w[i].r = s_w[0 * SZ + pq[i]];
      w[i].lrho = s_w[1 * SZ + pq[i]];
      w[i].b = s_w[2 * SZ + pq[i]];
      w[i].lbeta = s_w[3 * SZ + pq[i]];
#pragma unroll
      for (int d = 0; d < D; ++d) w[i].h[d] = s_w[(4 + d) * SZ + pq[i]];
      w[i].hp = s_w[(4 + D) * SZ + pq[i]];
    };
#pragma unroll
    for (int i = 0; i < n; ++i) load_w(i);
    if constexpr (OWN_FROM_W) {  // the end records, for the faces after the volume terms
      wend[0] = w[0];
      wend[1] = w[n - 1];
    }
    bool fast = false;
    if constexpr (sizeof(T) == 4) {
      // the division-free series applies when every log ratio on the line is in range
      T lo_r = w[0].lrho, hi_r = w[0].lrho, lo_b = w[0].lbeta, hi_b = w[0].lbeta;
#pragma unroll
      for (int i = 1; i < n; ++i) {
        lo_r = tmin(lo_r, w[i].lrho);
        hi_r = tmax(hi_r, w[i].lrho);
        lo_b = tmin(lo_b, w[i].lbeta);
        hi_b = tmax(hi_b, w[i].lbeta);
      }
      const bool in_range = (hi_r - lo_r <= T(kFastLogSpan)) && (hi_b - lo_b <= T(kFastLogSpan));
      fast = __builtin_amdgcn_ballot_w64(!in_range) == 0;  // wave-uniform
    }
    if (fast) {
      if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int i = 0; i < n; ++i)
#pragma unroll
          for (int j = i + 1; j < n; ++j) {
            T f[V];
            flux_ranocha_fast<D, O, SQ>(w[i], w[j], T(0.5) * igm1, f);
            const T di = k.Dsplit[i * n + j], dj = k.Dsplit[j * n + i];
#pragma unroll
            for (int v = 0; v < V - 1; ++v) {
              acc[i][v] = tfma(di, f[v], acc[i][v]);
              acc[j][v] = tfma(dj, f[v], acc[j][v]);
            }
            acc[i][V - 1] = tfma(k.Dsplit2[i * n + j], f[V - 1], acc[i][V - 1]);
            acc[j][V - 1] = tfma(k.Dsplit2[j * n + i], f[V - 1], acc[j][V - 1]);
          }
      }
    } else {
#pragma unroll
      for (int i = 0; i < n; ++i)
#pragma unroll
        for (int j = i + 1; j < n; ++j) {
          T f[V];
          if constexpr (sizeof(T) == 4) {
            // Rare path (a log ratio of the line outside the series range):
            // fresh copies of the two records from LDS. The opaque zero offset
            // keeps them apart from the register copies used by the fast path,
            // which then die at the branch; reusing them here would set the
            // register allocation of the whole kernel (144 instead of 106
            // VGPRs for 3D p3, one wave per SIMD less).
            int z0 = 0;
            asm volatile("" : "+v"(z0));
            const T *sw = s_w + z0;
            NodeR<T, D> wi, wj;
            auto read = [&](NodeR<T, D> &x, int q) {
              x.r = sw[0 * SZ + q];
              x.lrho = sw[1 * SZ + q];
              x.b = sw[2 * SZ + q];
              x.lbeta = sw[3 * SZ + q];
#pragma unroll
              for (int d = 0; d < D; ++d) x.h[d] = sw[(4 + d) * SZ + q];
              x.hp = sw[(4 + D) * SZ + q];
            };
            read(wi, pq[i]);
            read(wj, pq[j]);
            flux_ranocha_r<T, D, O, SQ>(wi, wj, igm1, f);
          } else {
            flux_ranocha_r<T, D, O, SQ>(w[i], w[j], igm1, f);
          }
          const T di = k.Dsplit[i * n + j], dj = k.Dsplit[j * n + i];
#pragma unroll
          for (int v = 0; v < V; ++v) {
            acc[i][v] = tfma(di, f[v], acc[i][v]);
            acc[j][v] = tfma(dj, f[v], acc[j][v]);
          }
        }
    }
  }
  if constexpr (VT::CEN) {
    T fph[n][V];
#pragma unroll
    for (int i = 0; i < n; ++i) {
      T ci_u[V];
#pragma unroll
      for (int v = 0; v < V; ++v) ci_u[v] = s_c[v * SZ + pq[i]];
      ConsP<T, D> ci;
      consp_from_cons<T, D>(ci_u, gm1, ci);
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
          acc[i][v] = tfma(di, f, acc[i][v]);
          acc[j][v] = tfma(dj, f, acc[j][v]);
        }
      }
  }
  if constexpr (VT::FV) {
    T afv = T(1);
    bool run_fv = true;
    if constexpr (VT::HG) {
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
      // first-order FV on the LGL subcells of this line; subcell fluxes at the
      // element faces are zero (create_f_threaded, dg_3d.jl:11-54)
      T fprev[V], cur[V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        fprev[v] = T(0);
        cur[v] = s_c[v * SZ + pq[0]];
      }
#pragma unroll
      for (int i = 0; i < n; ++i) {
        T fnext[V];
        if (i < n - 1) {
          T nxt[V];
#pragma unroll
          for (int v = 0; v < V; ++v) nxt[v] = s_c[v * SZ + pq[i + 1]];
          num_flux_c<T, D, O, VOL == VOL_PURE_FV>(a.fv_flux, cur, nxt, gamma, fnext);
#pragma unroll
          for (int v = 0; v < V; ++v) cur[v] = nxt[v];
        } else {
#pragma unroll
          for (int v = 0; v < V; ++v) fnext[v] = T(0);
        }
        const T cw = afv * k.inv_w[i];
#pragma unroll
        for (int v = 0; v < V; ++v) {
          acc[i][v] = tfma(cw, fnext[v] - fprev[v], acc[i][v]);
          fprev[v] = fnext[v];
        }
      }
    }
  }

  // keep the face work (which waits on the neighbour loads of phase A) behind
  // the volume arithmetic
  __builtin_amdgcn_sched_barrier(0);
  faces();

#pragma unroll
  for (int i = 0; i < n; ++i)
#pragma unroll
    for (int v = 0; v < V; ++v) s_d[(O * V + v) * SZ + pq[i]] = acc[i][v];
}

// Source-term phase tables of one element: with x_d = c_d + J xi_{i_d} the
// phase theta = pi (x_1 + ... + x_D - t) splits into
//   theta = alpha_{i_1} + gamma_{i_2} (+ gamma_{i_3}),
//   alpha_i = pi (c_1 + ... + c_D - t) + pi J xi_i,  gamma_j = pi J xi_j,
// so each element needs 2n sincos evaluations instead of one per node; the
// node values follow by angle addition.
// Entry idx of an element's table from its inverse Jacobian and the sum of its
// centre coordinates.
template <typename T, int N>
__device__ __forceinline__ void source_table_entry(const DGConst<T, N> &k, int idx, double inv_jac, double csum,
                                                   double tt, T *__restrict__ tab) {
  constexpr int n = N + 1;
  const double jac = 1.0 / inv_jac;
  const double pi = 3.14159265358979323846;
  const int which = idx / n, i = idx - which * n;  // which: 0 = alpha, 1 = gamma
  // Select the node from registers: a dynamically indexed kernel-argument load
  // would be a vector-memory load whose wait also drains the neighbour loads
  // in flight. The empty asm makes each value opaque, so the selects are not
  // folded back into an indexed load.
  double xi = 0.0;
#pragma unroll
  for (int m = 0; m < n; ++m) {
    double nm = k.nodes[m];
    asm volatile("" : "+v"(nm));
    xi = i == m ? nm : xi;
  }
  const double ph = (which == 0 ? pi * (csum - tt) : 0.0) + pi * jac * xi;
  T si, co;
  tsincos(T(ph), &si, &co);
  tab[2 * idx] = si;
  tab[2 * idx + 1] = co;
}

// LDS region A of a variant (node records and/or conservative state, later the
// output staging), in elements of T
template <typename T, int D, int N, int VOL, int EPB> constexpr int region_a_size() {
  constexpr int NW = VolTraits<VOL>::RAN ? nodeR_count(D) : 0;
  constexpr int NC = VolTraits<VOL>::CONS ? D + 2 : 0;
  constexpr int ALIGN = 16 / (int)sizeof(T);
  return ((NW + NC) * Layout<D, N, EPB>::SIZE + ALIGN - 1) / ALIGN * ALIGN;
}
// The fused last stage can reduce the wave speed when region A holds the D
// per-node FP64 speeds of the tile (not so for the weak form, central flux and
// pure FV in 3D FP32, whose region A is the conservative state only).
template <typename T, int D, int N, int VOL, int EPB> constexpr bool stage_speed_fits() {
  return (size_t)region_a_size<T, D, N, VOL, EPB>() * sizeof(T) >=
         (size_t)D * EPB * ipow_c(N + 1, D) * sizeof(double);
}
// The fused stage of a 2D polydeg-3 shock-capturing variant can compute the
// next stage's indicator when the tile is whole wavefronts of 16 elements, one
// wavefront each, and region A holds their 512 doubles of scratch apiece.
template <typename T, int D, int N, int VOL, int EPB, int NT> constexpr bool stage_ind2d_fits() {
  return D == 2 && N == 3 && VolTraits<VOL>::HG && EPB % 16 == 0 && EPB / 16 <= NT / 64 &&
         (size_t)region_a_size<T, D, N, VOL, EPB>() * sizeof(T) >= (size_t)(EPB / 16) * 512 * sizeof(double);
}

// Wave-speed reduction of the fused last stage (max_speed_nodes_kernel,
// dg_aux.hpp, for the tile's new state): per node the FP64 wave speeds of
// node_wave_speeds into s_lam [D][EPB NN] (LDS region A, free after the
// update), per (element, direction) the maximum over the nodes (S threads per
// pair, then one), per element inv_jac * sum over d, and the NaN-propagating
// maximum over the tile to speed_partial[e0 / EPB]. Maxima are exact, so the
// partials combine to the same maximum as the stand-alone kernel's.
template <typename T, int D, int N, int EPB, int NT>
__device__ __forceinline__ void stage_wave_speed(const ElemArgs<T> &a, const double *s_un, double *s_lam, int ne,
                                                 int e0) {
  constexpr int NN = ipow_c(N + 1, D), V = D + 2, S = 8, NPAIR = EPB * D;
  static_assert(EPB <= NT, "one thread per element");
  __shared__ double s_pm[NPAIR * S];
  __shared__ double s_em[EPB];
  const int tid = threadIdx.x;
  __syncthreads();  // s_un complete; region A (s_o) no longer read
  for (int q = tid; q < ne * NN; q += NT) {
    double lam[D];
    node_wave_speeds<D>([&](int v) { return s_un[q * V + v]; }, a.gamma64, lam);
#pragma unroll
    for (int d = 0; d < D; ++d) s_lam[d * EPB * NN + q] = lam[d];
  }
  __syncthreads();
  for (int k = tid; k < NPAIR * S; k += NT) {
    const int pair = k / S, sub = k - pair * S;
    const int el = pair / D, d = pair - el * D;
    double m = 0.0;
    if (el < ne)
      for (int q = sub; q < NN; q += S) m = nanmax(m, s_lam[d * EPB * NN + el * NN + q]);
    s_pm[k] = m;
  }
  __syncthreads();
  if (tid < ne) {
    double sum = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double m = 0.0;
#pragma unroll
      for (int j = 0; j < S; ++j) m = nanmax(m, s_pm[(tid * D + d) * S + j]);
      sum += m;
    }
    s_em[tid] = a.inv_jac[e0 + tid] * sum;
  }
  __syncthreads();
  if (tid == 0) {
    double bmax = 4.9406564584124654e-324;  // nextfloat(0.0), as max_speed_nodes_kernel starts
    for (int i = 0; i < ne; ++i) bmax = nanmax(bmax, s_em[i]);
    a.speed_partial[e0 / EPB] = bmax;
  }
}

// The next stage's indicator of a 2D polydeg-3 tile, from its new state in
// LDS (s_un, FP64 [q][v]; in FP32 mode rounded to the FP32 state the next
// stage reads): one wavefront per 16 elements runs indicator_2d_p3_wave with
// the network operands read where used, s_tmp (region A) holding its 512
// doubles of scratch per wavefront.
template <typename T, int EPB>
__device__ __forceinline__ void stage_indicator_2d(const ElemArgs<T> &a, const double *s_un, double *s_tmp, int ne,
                                                   int e0) {
  __syncthreads();  // s_un complete; region A no longer read
  const int w = threadIdx.x >> 6;
  if (w < EPB / 16 && w * 16 < ne) {
    const int l = threadIdx.x & 63, el = l & 15, j = l >> 4;
    const int elc = min(w * 16 + el, ne - 1);  // lanes past the tile's end repeat its last element
    T uv[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) uv[k] = T(s_un[(elc * 16 + 4 * j) * 4 + k]);
    double *s_mx = s_tmp + w * 512, *s_sq = s_mx + 256;
    if (a.ind.net) indicator_2d_p3_wave<T, true, false>(uv, e0 + w * 16, a.E, a.ind, s_mx, s_sq);
    else indicator_2d_p3_wave<T, false, false>(uv, e0 + w * 16, a.E, a.ind, s_mx, s_sq);
  }
}

// One tile of EPB elements per workgroup. Every global load of phase A is
// issued before the first one is consumed: neighbour ids, the tile (in chunks,
// into registers), the per-element scalars, then the neighbour face states,
// which stay in flight until phase B.
// Minimum wavefronts per SIMD the register allocation must allow. The FP64 3D
// polydeg-3 flux_ranocha variants are held to 168 VGPRs (from 184, some values
// spilled): 3 instead of 2 waves per SIMD, RHS 108.9 -> 102.4 us and CK2N54
// step 613.5 -> 574.7 us on C2 (A/B on one box); no constraint elsewhere.
template <typename T, int D, int N, int VOL> constexpr int min_waves_per_simd() {
  return sizeof(T) == 8 && D == 3 && N == 3 && VOL == VOL_FD_RANOCHA ? 3 : 1;
}

template <typename T, int D, int N, int VOL, int EPB, int WPD, bool BND, bool STAGE>
__global__ __launch_bounds__(D *WPD * 64) __attribute__((amdgpu_waves_per_eu(min_waves_per_simd<T, D, N, VOL>())))
void dg_element_kernel(ElemArgs<T> a, DGConst<T, N> k) {
  using LY = Layout<D, N, EPB>;
  using VT = VolTraits<VOL>;
  using TIO = TileIO<T, D, N, EPB>;
  using C = typename TIO::C;
  constexpr int NT = D * WPD * 64;
  constexpr int n = N + 1, NN = LY::NN, V = D + 2, SZ = LY::SIZE;
  constexpr int NW = VT::RAN ? nodeR_count(D) : 0;
  constexpr int NC = VT::CONS ? V : 0;
  constexpr int ALIGN = 16 / (int)sizeof(T);
  // primitive / conservative data, later output staging; a multiple of 16 bytes
  constexpr int SIZE_A = region_a_size<T, D, N, VOL, EPB>();
  static_assert(SIZE_A == ((NW + NC) * SZ + ALIGN - 1) / ALIGN * ALIGN, "region A");
  constexpr int SIZE_B = D * V * SZ;  // du per direction, earlier input staging
  static_assert(SIZE_A >= EPB * NN * V && SIZE_B >= EPB * NN * V, "staging does not fit");
  static_assert(EPB * LY::L <= WPD * 64, "one line per lane");
  static_assert(EPB * 2 * n <= NT, "one source-table entry per thread");
  constexpr int RPT = (TIO::NCH + NT - 1) / NT;  // tile chunks per thread
  __shared__ __attribute__((aligned(16))) T smem[SIZE_A + SIZE_B];
  __shared__ double s_alpha[VT::HG ? EPB : 1];
  __shared__ T s_src[EPB * 4 * n];  // per element: (sin, cos) of alpha_i, gamma_j
  __shared__ T s_fac[EPB];          // per element: -inverse_jacobian
  T *s_w = smem;            // [NW][SZ]
  T *s_c = smem + NW * SZ;  // [NC][SZ]
  T *s_d = smem + SIZE_A;   // [D][V][SZ]
  T *s_o = smem;            // [EPB*NN][V] output staging (region A, free after phase B)

  const int tid = threadIdx.x;
  // direction of this wave group (wave-uniform, kept in a scalar register)
  const int grp = __builtin_amdgcn_readfirstlane(tid / (WPD * 64));
  const int ll = tid - grp * (WPD * 64);
  const int e0 = (a.xcd_remap ? xcd_remap_block(blockIdx.x, gridDim.x) : (int)blockIdx.x) * EPB;
  if (e0 >= a.E) return;  // uniform per block
  const int ne = min(EPB, a.E - e0);
  const bool full = ne == EPB;
  const double tt = a.tdev ? (a.tdev[0] + a.t * a.tdev[1]) : a.t;

  // ---- phase A: issue the loads ----
  LineCtx<T, V> lc;
  int nbm, nbp;
  line_ids<T, D, N, EPB>(a, e0, ne, ll, grp, lc, nbm, nbp);
  // chunk indices are clamped to the tile's extent, so all lanes load (a
  // partial last tile is staged by the scalar loop below instead)
  const int nch = full ? TIO::NCH : max(1, ne * NN * V * (int)sizeof(T) / TIO::CB);
  const double *u64 = STAGE && sizeof(T) == 4 ? a.u64 : nullptr;  // FP64 source of the FP32 stage, or null
  const C *csrc = reinterpret_cast<const C *>(a.u + (size_t)e0 * NN * V);
  C rc[RPT];
  if (STAGE && sizeof(T) == 4 && u64) {
    constexpr int WC = TIO::CB / (int)sizeof(T);
    const double *src64 = u64 + (size_t)e0 * NN * V;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int ic = min(tid + r * NT, nch - 1);
      T x[WC];
#pragma unroll
      for (int w = 0; w < WC; ++w) x[w] = T(src64[(size_t)ic * WC + w]);
      __builtin_memcpy(&rc[r], x, sizeof(C));
    }
  } else {
#pragma unroll
    for (int r = 0; r < RPT; ++r) rc[r] = csrc[min(tid + r * NT, nch - 1)];
  }
  // element of this thread's source-table entry: the table is built by the
  // last EPB * 2n threads, away from wave 0, which has the most node work of
  // phases A and C (threads outside the table repeat the nearest entry)
  const int ts = tid - (NT - EPB * 2 * n);
  const int sel = min(max(ts, 0) / (2 * n), ne - 1);
  const double ij = a.inv_jac[e0 + sel];
  double cen[D];
#pragma unroll
  for (int d = 0; d < D; ++d) cen[d] = a.centers[d + D * (e0 + sel)];
  line_loads<T, D, N, EPB, BND>(a, lc, grp, nbm, nbp, u64);

  // ---- phase A: stage the tile and the per-element data into LDS ----
  {
    // lanes past the last chunk rewrite that chunk with the value it holds
    // (an unconditional store keeps the loads above where they are issued)
    C *dst = reinterpret_cast<C *>(s_d);
#pragma unroll
    for (int r = 0; r < RPT; ++r) dst[min(tid + r * NT, nch - 1)] = rc[r];
    if (!full) {  // the bytes of a partial tile past its last whole chunk
      const T *src = a.u + (size_t)e0 * NN * V;
      const double *src64 = u64 ? u64 + (size_t)e0 * NN * V : nullptr;
      for (int i = nch * (TIO::CB / (int)sizeof(T)) + tid; i < ne * NN * V; i += NT)
        s_d[i] = src64 ? T(src64[i]) : src[i];
    }
  }
  if constexpr (VT::HG) {
    if (tid < ne) {
      const int e = e0 + tid;
      double al = a.alpha_raw[e];
      if (a.alpha_smooth) {
        // apply_smoothing! (indicators_2d.jl:104-135, indicators_3d.jl:125-173):
        // alpha_e = max(alpha_e, 0.5 alpha_nbr) over the element's interfaces
        double m = al;
#pragma unroll
        for (int dir = 0; dir < 2 * D; ++dir) {
          const int nb = a.nbr[dir + 2 * D * e];  // < 0: boundary, no interface
          double an = nb >= 0 ? 0.5 * a.alpha_raw[nb] : 0.0;
          m = an > m ? an : m;
        }
        al = m;
      }
      s_alpha[tid] = al;
      if (a.alpha_out) a.alpha_out[e] = al;
    }
  }
  if (ts >= 0 && ts < ne * 2 * n) {
    const int idx = ts - sel * 2 * n;
    if (idx == 0) s_fac[sel] = T(-ij);
    if (a.source) {
      double csum = 0;
#pragma unroll
      for (int d = 0; d < D; ++d) csum += cen[d];
      source_table_entry<T, N>(k, idx, ij, csum, tt, &s_src[sel * 4 * n]);
    }
  }
  __syncthreads();
  if (grp == 0) line_own<T, D, N, EPB, 0>(s_d, lc);
  else if (grp == 1) line_own<T, D, N, EPB, 1>(s_d, lc);
  else if constexpr (D == 3) line_own<T, D, N, EPB, 2>(s_d, lc);
  {
    const T gm1 = a.gamma - T(1);
    for (int q = tid; q < ne * NN; q += NT) {
      const int p = LY::phys(q);
      T uq[V];
#pragma unroll
      for (int v = 0; v < V; ++v) uq[v] = s_d[q * V + v];
      if constexpr (VT::CONS) {
#pragma unroll
        for (int v = 0; v < V; ++v) s_c[v * SZ + p] = uq[v];
      }
      if constexpr (VT::RAN) {
        NodeR<T, D> w;
        nodeR_from_cons<T, D, node_sqrt_rep<T, N>()>(uq, gm1, w);
        s_w[0 * SZ + p] = w.r;
        s_w[1 * SZ + p] = w.lrho;
        s_w[2 * SZ + p] = w.b;
        s_w[3 * SZ + p] = w.lbeta;
#pragma unroll
        for (int d = 0; d < D; ++d) s_w[(4 + d) * SZ + p] = w.h[d];
        s_w[(4 + D) * SZ + p] = w.hp;
      }
    }
  }
  __syncthreads();

  // ---- phase B: one direction per wave group, one line per lane ----
  if (grp == 0) line_work<T, D, N, VOL, EPB, 0, BND>(a, k, lc, s_w, s_c, s_d, s_alpha);
  else if (grp == 1) line_work<T, D, N, VOL, EPB, 1, BND>(a, k, lc, s_w, s_c, s_d, s_alpha);
  else if constexpr (D == 3) line_work<T, D, N, VOL, EPB, 2, BND>(a, k, lc, s_w, s_c, s_d, s_alpha);
  __syncthreads();

  // ---- phase C: sum directions, Jacobian, sources; stage the output ----
  // The fused stage's register loads (u_tmp, u) of a full tile are issued
  // here, all at once, so that their latency overlaps the rest of phase C.
  constexpr int W = TIO::CB / (int)sizeof(T);  // values per chunk
  constexpr int NPRE = STAGE ? RPT * W : 1;
  double pre_t[NPRE], pre_u[NPRE];
  const size_t base = (size_t)e0 * NN * V;
  if constexpr (STAGE) {
    if (full) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const int ic = min(tid + r * NT, TIO::NCH - 1);  // clamped: every lane loads in bounds
#pragma unroll
        for (int w = 0; w < W; ++w) {
          const size_t g = base + (size_t)ic * W + w;
          pre_t[r * W + w] = a.first_stage ? 0.0 : double(a.u_tmp[g]);
          pre_u[r * W + w] = a.u_reg_in[g];
        }
      }
    }
  }
  for (int q = tid; q < ne * NN; q += NT) {
    const int el = q / NN, node = q - el * NN;
    const int p = LY::phys(q);
    const T fac = s_fac[el];
    T d[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      T sum = s_d[(0 * V + v) * SZ + p];
#pragma unroll
      for (int o = 1; o < D; ++o) sum += s_d[(o * V + v) * SZ + p];
      d[v] = sum * fac;
    }
    if (a.source) {
      // theta = alpha_i + gamma_j (+ gamma_k) by angle addition
      const T *tb = &s_src[el * 4 * n];
      const int i0 = node % n, i1 = (node / n) % n;
      T si = tb[2 * i0], co = tb[2 * i0 + 1];
      T sg = tb[2 * (n + i1)], cg = tb[2 * (n + i1) + 1];
      T s2 = tfma(si, cg, co * sg), c2 = tfma(co, cg, -si * sg);
      if constexpr (D == 3) {
        const int i2 = node / (n * n);
        sg = tb[2 * (n + i2)];
        cg = tb[2 * (n + i2) + 1];
        si = tfma(s2, cg, c2 * sg);
        co = tfma(c2, cg, -s2 * sg);
      } else {
        si = s2;
        co = c2;
      }
      T s[V];
      source_convergence_sc<T, D>(si, co, a.gamma, s);
#pragma unroll
      for (int v = 0; v < V; ++v) d[v] += s[v];
    }
#pragma unroll
    for (int v = 0; v < V; ++v) s_o[q * V + v] = d[v];
  }
  __syncthreads();
  if constexpr (STAGE) {
    // the 2N update of the tile's values, in the same arithmetic as
    // rk2n_update_kernel (bitwise equal to the unfused stage); with the speed
    // reduction the new values are kept in LDS too (region B, [q][v], free
    // after phase C)
    const double b_dt = a.tdev ? a.b_stage * a.tdev[1] : a.b_stage;
    constexpr bool SPEED_FITS = stage_speed_fits<T, D, N, VOL, EPB>();
    const bool speed = SPEED_FITS && a.speed_partial != nullptr;
    constexpr bool IND2D = stage_ind2d_fits<T, D, N, VOL, EPB, NT>();
    const bool ind = IND2D && a.ind_next != 0;
    double *s_un = reinterpret_cast<double *>(s_d);
    static_assert(SIZE_B * sizeof(T) >= (size_t)EPB * NN * V * sizeof(double), "new state staging does not fit");
    auto update = [&](size_t g, T du_v, double ut, double ur) {
      ut = double(du_v) - ut * a.a_stage;
      const double un = tfma(ut, b_dt, ur);
      a.u_tmp[g] = T(ut);
      a.u_reg_out[g] = un;
      if (a.u32_out) a.u32_out[g] = float(un);
      if (speed || ind) s_un[g - base] = un;
    };
    if (full) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const int i = tid + r * NT;
        if (i < TIO::NCH) {
#pragma unroll
          for (int w = 0; w < W; ++w)
            update(base + (size_t)i * W + w, s_o[i * W + w], pre_t[r * W + w], pre_u[r * W + w]);
        }
      }
    } else {
      for (int i = tid; i < ne * NN * V; i += NT) {
        const size_t g = base + i;
        update(g, s_o[i], a.first_stage ? 0.0 : double(a.u_tmp[g]), a.u_reg_in[g]);
      }
    }
    if constexpr (SPEED_FITS) {
      if (speed) stage_wave_speed<T, D, N, EPB, NT>(a, s_un, reinterpret_cast<double *>(smem), ne, e0);
    }
    if constexpr (IND2D) {
      if (ind) stage_indicator_2d<T, EPB>(a, s_un, reinterpret_cast<double *>(smem), ne, e0);
    }
  } else if (full) {
    const C *src = reinterpret_cast<const C *>(s_o);
    C *dst = reinterpret_cast<C *>(a.du + (size_t)e0 * NN * V);
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int i = tid + r * NT;
      if (i < TIO::NCH) dst[i] = src[i];
    }
  } else {
    T *dst = a.du + (size_t)e0 * NN * V;
    for (int i = tid; i < ne * NN * V; i += NT) dst[i] = s_o[i];
  }
}

}  // namespace trixi_hip
