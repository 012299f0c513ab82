// trixi_hip.hip -- C ABI of the MI355X DGSEM compressible-Euler RHS (see
// include/trixi_hip.h): context management, kernel dispatch, the 2N RK stage
// and step (hipGraph-captured), reductions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/trixi_hip.h"
#include "dg_aux.hpp"
#include "elem_registry.hpp"
#include "indicator2d.hpp"
#include "mlp_kernels.hpp"

using namespace trixi_hip;

namespace trixi_hip {
std::map<ElemKey, ElemLaunch> &elem_registry() {
  static std::map<ElemKey, ElemLaunch> r;
  return r;
}
std::map<ElemKey, ElemCaps> &elem_caps() {
  static std::map<ElemKey, ElemCaps> r;
  return r;
}
}  // namespace trixi_hip

namespace {

// ---------------------------------------------------------------------------
// Element kernels: instantiated in elem_*.hip and looked up here by key.
// ---------------------------------------------------------------------------
// Tile when the configuration asks for the default (elements_per_block = 0) and
// neither the process's own autotuning (trixi_hip_autotune) nor the measured
// table below has an entry for the kernel variant: one wavefront per direction.
struct DefaultTile {
  int D, N, epb, nt;
};
const DefaultTile kDefaultTiles[] = {{2, 2, 21, 128}, {2, 3, 16, 128}, {2, 4, 12, 128},
                                     {3, 2, 7, 192},  {3, 3, 4, 192},  {3, 4, 2, 192}};

// Tiles chosen by on-device sweeps of every instantiated tile per kernel
// variant (python -m trixi_hip.autotune on MI355X writes tile_table.inc).
struct TunedTile {
  int fp32, D, N, vol, bnd, stage, epb, nt;
};
const TunedTile kTunedTiles[] = {
#include "tile_table.inc"
    {-1, 0, 0, 0, 0, 0, 0, 0},  // end marker (matches no variant)
};

// (precision, ndims, polydeg, volume integral, boundaries, stage) -> tile
// measured by trixi_hip_autotune in this process; consulted before kTunedTiles.
struct TuneKey {
  int fp32, D, N, vol, bnd, stage;
  bool operator<(const TuneKey &o) const {
    return std::tie(fp32, D, N, vol, bnd, stage) < std::tie(o.fp32, o.D, o.N, o.vol, o.bnd, o.stage);
  }
};
std::map<TuneKey, std::pair<int, int>> g_tuned;
std::mutex g_tuned_mutex;

int vol_kind(const trixi_hip_config &c) {
  switch (c.volume_integral) {
    case TRIXI_HIP_VI_WEAK: return VOL_WEAK;
    case TRIXI_HIP_VI_FLUX_DIFFERENCING:
      return c.volume_flux == TRIXI_HIP_FLUX_RANOCHA ? VOL_FD_RANOCHA
             : c.volume_flux == TRIXI_HIP_FLUX_CENTRAL ? VOL_FD_CENTRAL : -1;
    case TRIXI_HIP_VI_SHOCK_CAPTURING_HG:
      return c.volume_flux == TRIXI_HIP_FLUX_RANOCHA ? VOL_HG_RANOCHA
             : c.volume_flux == TRIXI_HIP_FLUX_CENTRAL ? VOL_HG_CENTRAL : -1;
    case TRIXI_HIP_VI_PURE_FV: return VOL_PURE_FV;
    default: return -1;
  }
}

// ---------------------------------------------------------------------------
// Aux-kernel dispatch on (T, D, N) via a compile-time list.
// ---------------------------------------------------------------------------
template <typename F> bool dispatch_dn(int D, int N, F &&f) {
#define TRIXI_DN(DD, NN_) \
  if (D == DD && N == NN_) { f(std::integral_constant<int, DD>{}, std::integral_constant<int, NN_>{}); return true; }
  TRIXI_DN(2, 2) TRIXI_DN(2, 3) TRIXI_DN(2, 4) TRIXI_DN(3, 2) TRIXI_DN(3, 3) TRIXI_DN(3, 4)
#undef TRIXI_DN
  return false;
}

constexpr int kMaxPartials = 1 << 16;
constexpr int kMaxSpeedBlocks = 1 << 16;  // workgroups (and partials) of the wave-speed reduction
constexpr int kMlpMaxWaves = 1024;  // partial gradient rows of one training step
constexpr int kEagerSolveSteps = 8;  // trixi_hip_solve with at most this many steps launches without graphs

}  // namespace

struct trixi_hip_ctx {
  trixi_hip_config cfg{};
  int D = 0, N = 0, n = 0, NN = 0, V = 0, E = 0, I = 0;
  size_t nvals = 0;  // V * NN * E
  bool fp32 = false;
  int vol = 0;
  hipStream_t own_stream = nullptr, stream = nullptr;
  std::string err;
  int device = 0;

  // mesh / basis on the device
  int32_t *d_nbr = nullptr;
  int B = 0;                     // boundaries (non-periodic mesh)
  double *d_bx = nullptr;        // [B, n^(D-1), D] coordinates of the boundary face nodes
  bool slip = false;             // boundary_condition_slip_wall: boundary slots carry given fluxes
  int32_t *d_bnd = nullptr;      // [B][2] {element, 0-based direction} (slip walls)
  // L2 mortars (refined meshes): M mortars, each with S = 2^(D-1) small faces
  // and one large face; face slots B + m (S + 1) + face in ghost / mflux
  int M = 0;
  int32_t *d_mort = nullptr;     // [M][S + 3] {small ids, large id, orientation, large_side}
  void *d_mflux = nullptr;       // [B + M (S + 1), n^(D-1), V] given fluxes of mortar faces (and slip walls)
  std::vector<double> mfu, mfl, mru, mrl;  // forward/reverse upper/lower, row-major n x n
  int32_t *d_adj_ptr = nullptr, *d_adj = nullptr;  // element adjacency (interfaces + mortars) for HG smoothing
  bool ghost_ready = false;      // ghost holds the states of a time-independent boundary function
  void *d_ghost = nullptr;       // [B, n^(D-1), V] outer boundary states, native precision
  double *d_invj = nullptr, *d_centers = nullptr;
  // resident state: the FP64 RK register u (a pair in FP64 mode, one buffer in
  // FP32 mode), the 2N register u_tmp in the RHS precision, the FP32 shadow of
  // u (a pair in FP32 mode). The fused
  // stage kernel reads the native state of buffer par and writes the other one
  // (neighbour faces are read while other workgroups write), then par flips.
  double *d_ub[2] = {nullptr, nullptr};
  void *d_utmp = nullptr;
  float *d_u32b[2] = {nullptr, nullptr};
  int par = 0;
  // FP32 mode: the shadow d_u32b[par] lags the register d_ub[par] (the fused
  // stages of a mesh without boundaries and mortars do not write it); readers
  // of the shadow call sync_u32 first
  bool u32_stale = false;
  void *d_du = nullptr;  // native precision
  // staging for trixi_hip_rhs (FP64 interface)
  double *d_hu = nullptr, *d_hdu = nullptr;
  float *d_su32 = nullptr, *d_sdu32 = nullptr;
  // HG: raw factors (scratch of the RHS and indicator entry points) and the
  // smoothed ones
  double *d_alpha_raw = nullptr, *d_alpha = nullptr;
  // raw factors of state buffer par for the stages: a fused 2D p3 stage writes
  // those of its new state into d_alpha_st[par ^ 1] (other workgroups still
  // read d_alpha_st[par]); alpha_fresh: d_alpha_st[par] holds those of the
  // resident state, so the next stage launches no indicator
  double *d_alpha_st[2] = {nullptr, nullptr};
  bool alpha_fresh = false;
  // the fused 2D p3 stage computes the next stage's indicator in its epilogue
  // instead of the stand-alone indicator_2d_p3_kernel reading the state again;
  // TRIXI_HIP_FUSED_INDICATOR=0 turns it off
  bool fused_ind = true;
  // reductions
  double *d_partial = nullptr, *d_scalars = nullptr;  // scalars: [0..1] {t, dt}, [2] max_dt, [3..] reductions
  size_t partial_cap = 0;

  // host copies of basis
  std::vector<double> nodes, weights, Dsplit, Dhat, Vinv, a_weights, a_vdm;
  int na = 0;
  HGParams hg{};
  ExactParams exact{};

  // element-kernel launch (elem_stage: the variant with the fused 2N update,
  // null when not instantiated for the tile or switched off); each with its
  // own tile (elements per workgroup, threads)
  ElemLaunch elem = nullptr, elem_stage = nullptr;
  int epb = 0, nt = 0, epb_st = 0, nt_st = 0;
  std::vector<unsigned char> kconst;  // DGConst<T, N> blob
  // workgroups of the element kernel walk contiguous element ranges per XCD
  // (neighbour faces then mostly hit the same L2), separately for the RHS and
  // the fused-stage variant. FP32: both (3D p3 RHS 36.7 -> 35.8 us, CK2N54 step
  // 395.6 -> 382.6 us). FP64: the stage only (step 565.2 -> 555.6 us; the RHS is
  // slower with it, 101.9 vs 102.8 us). A/B on one box; TRIXI_HIP_XCD_REMAP=0/1
  // overrides both.
  int xcd_remap = 0, xcd_stage = 0;
  // the last stage of a step reduces the wave speed in its epilogue (see
  // stage_wave_speed) instead of max_speed_nodes_kernel reading the state again.
  // FP32 default: 3D p3 CK2N54 step 382.6 -> 370.4 us. FP64 is slower with it
  // (611.6 vs 623.2 us: at two workgroups per CU the epilogue's barriers are not
  // hidden). A/B on one box; TRIXI_HIP_FUSED_SPEED=0/1 overrides.
  bool fused_speed = false;

  // captured RK steps, keyed by scheme, stream, entry point and start parity
  // (two slots: consecutive steps of a scheme with an odd stage count alternate)
  struct GraphSlot {
    std::vector<double> key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  } graphs[2];
  int graph_next = 0;

  // blending factor source for shock capturing (TRIXI_HIP_IND_*)
  int ind_mode = TRIXI_HIP_IND_HG;
  // learned indicator (2D p3): d_mlp = weights [NP] | Adam m [NP] | Adam v [NP]
  // | feature box lo [NF] | hi [NF] | gradient + loss [NPL]
  float *d_mlp = nullptr, *d_feat = nullptr, *d_mlp_partial = nullptr;
  mlp::InferImage *d_mlp_img = nullptr;  // BF16 inference image of the parameters
  double *d_alpha_hg = nullptr;
  bool mlp_ready = false;
  int mlp_step = 0;

  // analysis records of trixi_hip_solve (ring of rec_cap records)
  double *d_records = nullptr;
  int rec_cap = 0, rec_interval = 0;
  double tv_err = -1, tv_int = -1;  // total volume as the error norms / integrals compute it

  // device allocations and host synchronisations made for this context
  // (trixi_hip_get_counters)
  long long n_alloc = 0, n_sync = 0;
  // size class of every device block this context holds (block pool)
  std::map<void *, size_t> blocks;
};

namespace {

// Device blocks and streams of destroyed contexts are kept per device and
// handed to later contexts (an AMR run creates a context per adaptation, of
// slightly different sizes): a block size class is 256-byte granules rounded
// up to the next of 4/8, 5/8, 6/8, 7/8 or 8/8 of a power of two (at most 25 %
// unused); above 64 MiB, 2 MiB granules (at most 3 % unused). At most
// kPoolBytes per device and kPoolStreams streams are kept; the rest is freed.
// A block or stream enters the pool only after the work of its context has
// completed (destroy synchronises the stream first).
constexpr int kPoolDevices = 64;
constexpr size_t kPoolBytes = (size_t)1 << 30;
constexpr size_t kPoolStreams = 8;
struct DevicePool {
  std::mutex m;
  std::multimap<size_t, void *> blocks[kPoolDevices];
  size_t cached[kPoolDevices] = {};
  std::vector<hipStream_t> streams[kPoolDevices];
};
DevicePool &device_pool() {
  static DevicePool *p = new DevicePool;  // never destroyed: blocks outlive static destruction
  return *p;
}
size_t size_class(size_t bytes) {
  const size_t g = (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
  constexpr size_t kBig = (size_t)64 << 20, kBigGranule = (size_t)2 << 20;
  if (g > kBig) return (g + kBigGranule - 1) / kBigGranule * kBigGranule;
  size_t p = 256;
  while (p < g) p <<= 1;
  if (p <= 1024) return p;
  size_t c = p / 2;
  while (c < g) c += p / 8;
  return c;
}
bool pool_device_ok(int dev) { return dev >= 0 && dev < kPoolDevices; }

// Every device allocation and every host synchronisation of the library goes
// through these, so that trixi_hip_get_counters can report them.
template <typename P> hipError_t dev_alloc(trixi_hip_ctx *c, P **p, size_t bytes) {
  ++c->n_alloc;
  const size_t cls = size_class(bytes);
  void *ptr = nullptr;
  if (pool_device_ok(c->device)) {
    DevicePool &pool = device_pool();
    std::lock_guard<std::mutex> lock(pool.m);
    auto it = pool.blocks[c->device].find(cls);
    if (it != pool.blocks[c->device].end()) {
      ptr = it->second;
      pool.blocks[c->device].erase(it);
      pool.cached[c->device] -= cls;
    }
  }
  if (!ptr) {
    hipError_t e = hipMalloc(&ptr, cls);
    if (e != hipSuccess) return e;
  }
  c->blocks[ptr] = cls;
  *p = static_cast<P *>(ptr);
  return hipSuccess;
}
// Returns a block of this context to the pool (its context's work on it must
// have completed).
void dev_release(trixi_hip_ctx *c, void *ptr) {
  if (!ptr) return;
  auto it = c->blocks.find(ptr);
  const size_t cls = it == c->blocks.end() ? 0 : it->second;
  if (it != c->blocks.end()) c->blocks.erase(it);
  if (cls && pool_device_ok(c->device)) {
    DevicePool &pool = device_pool();
    std::lock_guard<std::mutex> lock(pool.m);
    if (pool.cached[c->device] + cls <= kPoolBytes) {
      pool.blocks[c->device].emplace(cls, ptr);
      pool.cached[c->device] += cls;
      return;
    }
  }
  (void)hipFree(ptr);
}
hipError_t stream_acquire(int dev, hipStream_t *s) {
  if (pool_device_ok(dev)) {
    DevicePool &pool = device_pool();
    std::lock_guard<std::mutex> lock(pool.m);
    if (!pool.streams[dev].empty()) {
      *s = pool.streams[dev].back();
      pool.streams[dev].pop_back();
      return hipSuccess;
    }
  }
  return hipStreamCreate(s);
}
void stream_release(int dev, hipStream_t s) {
  if (pool_device_ok(dev)) {
    DevicePool &pool = device_pool();
    std::lock_guard<std::mutex> lock(pool.m);
    if (pool.streams[dev].size() < kPoolStreams) {
      pool.streams[dev].push_back(s);
      return;
    }
  }
  (void)hipStreamDestroy(s);
}
hipError_t host_sync(trixi_hip_ctx *c) {
  ++c->n_sync;
  return hipStreamSynchronize(c->stream);
}
hipError_t sync_copy(trixi_hip_ctx *c, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
  ++c->n_sync;
  return hipMemcpy(dst, src, bytes, kind);
}

int set_err(trixi_hip_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

// message of the last failed context-free call of this thread (fixed buffer:
// those calls allocate nothing, host memory included)
thread_local char g_free_err[256] = "";

int free_err(int code, const char *msg, const char *detail = nullptr) {
  std::snprintf(g_free_err, sizeof g_free_err, "%s%s%s", msg, detail ? ": " : "", detail ? detail : "");
  return code;
}

// the current resident state: FP64 register u, FP32 shadow (FP32 mode), native
double *cur_u64(trixi_hip_ctx *c) { return c->d_ub[c->par]; }
float *cur_u32(trixi_hip_ctx *c) { return c->d_u32b[c->par]; }
const void *native_u(trixi_hip_ctx *c) { return c->fp32 ? (const void *)cur_u32(c) : (const void *)cur_u64(c); }

// outputs of the fused stage kernel (ElemArgs fields of the STAGE variants)
struct StageOut {
  void *u_tmp = nullptr;
  const double *u_reg_in = nullptr;
  double *u_reg_out = nullptr;
  float *u32_out = nullptr;
  const double *u64 = nullptr;  // FP32 stage: read the state from this FP64 register
  double a_stage = 0, b_stage = 0;
  int first_stage = 0;
  double *speed_partial = nullptr;  // the last stage of a step: wave-speed partials of the new state
  bool ind_next = false;            // 2D p3 shock capturing: the next stage's raw factors too
};

#define HIP_TRY(call)                                                                            \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      return set_err(ctx, TRIXI_HIP_EHIP, std::string(#call " failed: ") + hipGetErrorString(e_)); \
  } while (0)

template <typename T> void build_kconst(trixi_hip_ctx *ctx) {
  dispatch_dn(ctx->D, ctx->N, [&](auto Dc, auto Nc) {
    constexpr int N = decltype(Nc)::value;
    constexpr int n = N + 1;
    DGConst<T, N> k{};
    for (int i = 0; i < n * n; ++i) {
      k.Dsplit[i] = T(ctx->Dsplit[i]);
      k.Dsplit2[i] = T(2.0 * ctx->Dsplit[i]);
      k.Dhat[i] = T(ctx->Dhat[i]);
    }
    for (int i = 0; i < n; ++i) {
      k.inv_w[i] = T(1.0 / ctx->weights[i]);
      k.nodes[i] = ctx->nodes[i];
    }
    ctx->kconst.resize(sizeof(k));
    std::memcpy(ctx->kconst.data(), &k, sizeof(k));
    (void)Dc;
  });
}

TuneKey tune_key(const trixi_hip_ctx *ctx, int stage) {
  return TuneKey{ctx->fp32 ? 1 : 0, ctx->D, ctx->N, ctx->vol, ctx->B + ctx->M > 0 ? 1 : 0, stage};
}

ElemLaunch find_launch(const trixi_hip_ctx *ctx, int stage, int epb, int nt) {
  auto it = elem_registry().find(
      ElemKey{ctx->fp32 ? 1 : 0, ctx->D, ctx->N, ctx->vol, epb, nt, ctx->B + ctx->M > 0 ? 1 : 0, stage});
  return it == elem_registry().end() ? nullptr : it->second;
}

// What the context's fused-stage kernel can add in its epilogue: the wave
// speed of the new state (the last stage of a step then replaces the
// stand-alone max_speed_nodes_kernel) and the next stage's 2D p3 indicator.
ElemCaps stage_caps(const trixi_hip_ctx *ctx) {
  if (!ctx->elem_stage) return ElemCaps{};
  auto it = elem_caps().find(ElemKey{ctx->fp32 ? 1 : 0, ctx->D, ctx->N, ctx->vol, ctx->epb_st, ctx->nt_st,
                                     ctx->B + ctx->M > 0 ? 1 : 0, 1});
  return it != elem_caps().end() ? it->second : ElemCaps{};
}
bool stage_speed_ok(const trixi_hip_ctx *ctx) { return stage_caps(ctx).speed; }

// Default tile of a kernel variant: this process's autotuning, else the
// measured table, else one wavefront per direction.
void default_tile(const trixi_hip_ctx *ctx, int stage, int &epb, int &nt) {
  const TuneKey k = tune_key(ctx, stage);
  {
    std::lock_guard<std::mutex> lock(g_tuned_mutex);
    auto it = g_tuned.find(k);
    if (it != g_tuned.end() && find_launch(ctx, stage, it->second.first, it->second.second)) {
      epb = it->second.first;
      nt = it->second.second;
      return;
    }
  }
  for (const auto &t : kTunedTiles)
    if (t.fp32 == k.fp32 && t.D == k.D && t.N == k.N && t.vol == k.vol && t.bnd == k.bnd && t.stage == k.stage &&
        find_launch(ctx, stage, t.epb, t.nt)) {
      epb = t.epb;
      nt = t.nt;
      return;
    }
  for (const auto &t : kDefaultTiles)
    if (t.D == ctx->D && t.N == ctx->N) {
      epb = t.epb;
      nt = t.nt;
    }
}

// The RHS and the fused-stage kernels of the context: the given tile for both,
// or (epb, nt = 0) each variant's default tile.
int select_tile(trixi_hip_ctx *ctx, int epb, int nt) {
  int e_r = epb, n_r = nt, e_s = epb, n_s = nt;
  if (epb <= 0 || nt <= 0) {
    default_tile(ctx, 0, e_r, n_r);
    default_tile(ctx, 1, e_s, n_s);
  }
  ctx->elem = find_launch(ctx, 0, e_r, n_r);
  if (!ctx->elem)
    return set_err(ctx, TRIXI_HIP_EINVAL,
                   "no element kernel instantiated for this (precision, ndims, polydeg, volume "
                   "integral, tile, boundaries)");
  const char *unfused = std::getenv("TRIXI_HIP_UNFUSED_STAGE");
  ctx->elem_stage = (unfused && std::atoi(unfused) != 0) ? nullptr : find_launch(ctx, 1, e_s, n_s);
  ctx->epb = e_r;
  ctx->nt = n_r;
  ctx->epb_st = ctx->elem_stage ? e_s : 0;
  ctx->nt_st = ctx->elem_stage ? n_s : 0;
  return TRIXI_HIP_OK;
}

// Arguments of the 2D p3 indicator (stand-alone kernel or fused-stage
// epilogue): HG parameters, the inverse Vandermonde matrix, and with the
// network (or features) its device image, box bounds and modes.
Ind2DArgs ind2d_args(const trixi_hip_ctx *ctx, const HGParams &params, bool net, double *alpha_out, double *alpha_hg,
                     float *feat) {
  Ind2DArgs ia{};
  ia.hp = params;
  for (int i = 0; i < 16; ++i) ia.vi.v[i] = ctx->Vinv[i];
  ia.img = ctx->d_mlp_img;
  ia.lo = ctx->d_mlp ? ctx->d_mlp + 3 * mlp::NP : nullptr;
  ia.hi = ctx->d_mlp ? ia.lo + mlp::NF : nullptr;
  ia.net = net ? 1 : 0;
  ia.fallback = ctx->ind_mode == TRIXI_HIP_IND_MLP_WITH_FALLBACK ? 1 : 0;
  ia.alpha_raw = alpha_out;
  ia.alpha_hg_out = alpha_hg;
  ia.feat_out = feat;
  return ia;
}

// Raw (unsmoothed) blending factors of state u into alpha_out: the HG factor,
// or (2D p3, net = true) the learned indicator's, with the HG fallback. With
// feat != null (2D p3) the learned indicator's features [E][32] too, and with
// alpha_hg != null the HG factor. 2D p3 runs indicator_2d_p3_kernel, the other
// (ndims, polydeg) hg_indicator_kernel.
template <typename T>
int launch_hg(trixi_hip_ctx *ctx, const T *u, double *alpha_out, float *feat, const HGParams *hp = nullptr,
              bool net = false, double *alpha_hg = nullptr) {
  const HGParams params = hp ? *hp : ctx->hg;
  if (ctx->D == 2 && ctx->N == 3) {
    const int nb = (ctx->E + 15) / 16;  // one wavefront per 16 elements
    Ind2DArgs ia = ind2d_args(ctx, params, net, alpha_out, alpha_hg, feat);
    if (feat || net) {
      // the network buffers exist whenever features or the network are asked
      // for (trixi_hip_mlp_features and trixi_hip_set_indicator allocate them)
      if (!ctx->d_mlp || !ctx->d_mlp_img) return set_err(ctx, TRIXI_HIP_ESTATE, "learned-indicator buffers missing");
      hipLaunchKernelGGL((indicator_2d_p3_kernel<T, true>), dim3((nb + kIndWaves - 1) / kIndWaves),
                         dim3(64 * kIndWaves), 0, ctx->stream, u, ctx->E, ia);
    } else {
      hipLaunchKernelGGL((indicator_2d_p3_kernel<T, false>), dim3(nb), dim3(64), 0, ctx->stream, u, ctx->E, ia);
    }
    HIP_TRY(hipGetLastError());
    return TRIXI_HIP_OK;
  }
  if (feat || net) return set_err(ctx, TRIXI_HIP_EINVAL, "the learned indicator is defined for 2D polydeg 3");
  bool ok = dispatch_dn(ctx->D, ctx->N, [&](auto Dc, auto Nc) {
    constexpr int D = decltype(Dc)::value, N = decltype(Nc)::value;
    constexpr int NN = ipow_c(N + 1, D);
    constexpr int IPB = NN >= 256 ? 1 : 256 / NN;
    VinvConst<N> vc{};
    for (int i = 0; i < (N + 1) * (N + 1); ++i) vc.Vinv[i] = ctx->Vinv[i];
    int nb = (ctx->E + IPB - 1) / IPB;
    hipLaunchKernelGGL((hg_indicator_kernel<T, D, N, IPB, 256>), dim3(nb), dim3(256), 0, ctx->stream, u,
                       alpha_out, ctx->E, params, vc);
  });
  if (!ok) return set_err(ctx, TRIXI_HIP_EINVAL, "unsupported (ndims, polydeg) for the indicator");
  HIP_TRY(hipGetLastError());
  if (alpha_hg) HIP_TRY(hipMemcpyAsync(alpha_hg, alpha_out, (size_t)ctx->E * 8, hipMemcpyDeviceToDevice, ctx->stream));
  return TRIXI_HIP_OK;
}

int mlp_blocks(int n) {
  const int ntiles = (n + 15) / 16;
  const int waves = std::min(std::max(ntiles, 1), kMlpMaxWaves);
  return (waves + mlp::WAVES - 1) / mlp::WAVES;
}

// Raw (unsmoothed) blending factors of state u into alpha_out, from the
// indicator selected by ind_mode.
// (features, network and fallback in one launch: indicator_2d_p3_kernel).
template <typename T> int launch_indicator_raw(trixi_hip_ctx *ctx, const T *u, double *alpha_out) {
  return launch_hg<T>(ctx, u, alpha_out, nullptr, nullptr, ctx->ind_mode != TRIXI_HIP_IND_HG);
}

bool is_hg(const trixi_hip_ctx *ctx) { return ctx->vol == VOL_HG_RANOCHA || ctx->vol == VOL_HG_CENTRAL; }

// Whether the context's stages compute the next stage's indicator in their
// epilogue.
bool stage_fuses_indicator(const trixi_hip_ctx *ctx) {
  return ctx->fused_ind && is_hg(ctx) && stage_caps(ctx).ind2d;
}

// FP32 fused stages read the FP64 register directly (no shadow traffic) when
// nothing reads the native state between stages: no mortar or slip-wall flux
// kernel (Dirichlet outer states come from the boundary function), and no
// stand-alone indicator before each stage
bool stage_reads_u64(const trixi_hip_ctx *c) {
  return c->fp32 && c->M == 0 && !c->slip && (!is_hg(c) || stage_fuses_indicator(c));
}

template <typename T>
int launch_rhs_T(trixi_hip_ctx *ctx, T *du, const T *u, double t, const double *tdev,
                 const StageOut *so = nullptr) {
  // shock capturing: raw indicator first (needs the whole state), smoothing
  // happens inline. A stage reads the factors of state buffer par, computed by
  // the previous fused stage when fresh.
  const bool hg = is_hg(ctx);
  double *alpha_in = so ? ctx->d_alpha_st[ctx->par] : ctx->d_alpha_raw;
  if (hg) {
    if (!(so && ctx->alpha_fresh)) {
      int r = launch_indicator_raw<T>(ctx, u, alpha_in);
      if (r) return r;
    }
    if (ctx->M > 0 && ctx->cfg.alpha_smooth) {  // smoothing across interfaces and mortars, before the element kernel
      hipLaunchKernelGGL(hg_smooth_adj_kernel, dim3((ctx->E + 255) / 256), dim3(256), 0, ctx->stream,
                         alpha_in, ctx->d_adj_ptr, ctx->d_adj, ctx->d_alpha, ctx->E);
      HIP_TRY(hipGetLastError());
    }
  }
  // outer states of the Dirichlet boundaries at this stage's time; a
  // time-independent boundary function is evaluated once
  const bool t_dep = ctx->exact.kind == TRIXI_HIP_IC_CONVERGENCE_TEST;
  if (ctx->B > 0 && !ctx->slip && (t_dep || !ctx->ghost_ready)) {
    ctx->ghost_ready = !t_dep;
    bool ok = dispatch_dn(ctx->D, ctx->N, [&](auto Dc, auto Nc) {
      constexpr int D = decltype(Dc)::value, N = decltype(Nc)::value;
      const int total = ctx->B * ipow_c(N + 1, D - 1);
      hipLaunchKernelGGL((boundary_state_kernel<T, D, N>), dim3((total + 255) / 256), dim3(256), 0, ctx->stream,
                         ctx->d_bx, total, t, tdev, ctx->exact, (T *)ctx->d_ghost);
    });
    if (!ok) return TRIXI_HIP_EINVAL;
    HIP_TRY(hipGetLastError());
  }
  if (ctx->slip) {  // wall fluxes from the current inner states (given fluxes of the boundary slots)
    bool ok = dispatch_dn(ctx->D, ctx->N, [&](auto Dc, auto Nc) {
      constexpr int D = decltype(Dc)::value, N = decltype(Nc)::value;
      const int total = ctx->B * ipow_c(N + 1, D - 1);
      hipLaunchKernelGGL((slip_wall_flux_kernel<T, D, N>), dim3((total + 255) / 256), dim3(256), 0, ctx->stream, u,
                         ctx->d_bnd, ctx->B, ctx->cfg.gamma, (T *)ctx->d_mflux);
    });
    if (!ok) return TRIXI_HIP_EINVAL;
    HIP_TRY(hipGetLastError());
  }
  if (ctx->M > 0) {  // fluxes of the mortar faces
    bool ok = dispatch_dn(ctx->D, ctx->N, [&](auto Dc, auto Nc) {
      constexpr int D = decltype(Dc)::value, N = decltype(Nc)::value;
      constexpr int n = N + 1, S = 1 << (D - 1), TPM = S * (D == 3 ? n * n : n), MPB = 256 / TPM;
      MortarConst<T, N> mc{};
      for (int i = 0; i < n * n; ++i) {
        mc.fu[i] = T(ctx->mfu[i]);
        mc.fl[i] = T(ctx->mfl[i]);
        mc.ru[i] = T(ctx->mru[i]);
        mc.rl[i] = T(ctx->mrl[i]);
      }
      hipLaunchKernelGGL((mortar_flux_kernel<T, D, N>), dim3((ctx->M + MPB - 1) / MPB), dim3(256), 0, ctx->stream,
                         u, ctx->d_mort, ctx->M, ctx->B, ctx->cfg.surface_flux, T(ctx->cfg.gamma), mc,
                         (T *)ctx->d_mflux);
    });
    if (!ok) return TRIXI_HIP_EINVAL;
    HIP_TRY(hipGetLastError());
  }
  ElemArgs<T> a{};
  a.u = u;
  a.du = du;
  a.nbr = ctx->d_nbr;
  a.ghost = (const T *)ctx->d_ghost;
  a.mflux = (const T *)ctx->d_mflux;
  a.nbnd = ctx->slip ? 0 : ctx->B;  // slip walls: every face slot carries a given flux
  a.inv_jac = ctx->d_invj;
  a.centers = ctx->d_centers;
  const bool pre_smoothed = hg && ctx->M > 0 && ctx->cfg.alpha_smooth;  // smoothed across mortars above
  a.alpha_raw = pre_smoothed ? ctx->d_alpha : alpha_in;
  a.alpha_out = pre_smoothed ? nullptr : ctx->d_alpha;
  a.tdev = tdev;
  a.t = t;
  a.E = ctx->E;
  a.surface_flux = ctx->cfg.surface_flux;
  a.fv_flux = ctx->cfg.fv_flux;
  a.source = ctx->cfg.source_terms == TRIXI_HIP_SRC_CONVERGENCE_TEST;
  a.alpha_smooth = ctx->M > 0 ? 0 : ctx->cfg.alpha_smooth;
  a.xcd_remap = so ? ctx->xcd_stage : ctx->xcd_remap;
  a.gamma = T(ctx->cfg.gamma);
  a.inv_gamma_m1 = T(1) / (a.gamma - T(1));  // the device expression, evaluated once in T
  a.sqrt_gamma = T(std::sqrt(ctx->cfg.gamma));
  if (so) {
    a.u_tmp = (T *)so->u_tmp;
    a.u_reg_in = so->u_reg_in;
    a.u_reg_out = so->u_reg_out;
    a.u32_out = so->u32_out;
    a.u64 = so->u64;
    a.a_stage = so->a_stage;
    a.b_stage = so->b_stage;
    a.first_stage = so->first_stage;
    a.speed_partial = so->speed_partial;
    if (so->ind_next) {  // raw factors of the new state into the other buffer of the pair
      a.ind_next = 1;
      a.ind = ind2d_args(ctx, ctx->hg, ctx->ind_mode != TRIXI_HIP_IND_HG, ctx->d_alpha_st[ctx->par ^ 1], nullptr,
                         nullptr);
    }
  }
  a.gamma64 = ctx->cfg.gamma;
  const int epb = so ? ctx->epb_st : ctx->epb;
  int nblocks = (ctx->E + epb - 1) / epb;
  hipError_t e = (so ? ctx->elem_stage : ctx->elem)(&a, ctx->kconst.data(), nblocks, ctx->stream);
  if (e != hipSuccess) return set_err(ctx, TRIXI_HIP_EHIP, std::string("element kernel launch: ") + hipGetErrorString(e));
  return TRIXI_HIP_OK;
}

int launch_rhs_native(trixi_hip_ctx *ctx, void *du, const void *u, double t, const double *tdev) {
  if (ctx->fp32) return launch_rhs_T<float>(ctx, (float *)du, (const float *)u, t, tdev);
  return launch_rhs_T<double>(ctx, (double *)du, (const double *)u, t, tdev);
}

int grid_for(size_t count) {
  size_t b = (count + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (int)b;
}

template <typename TO, typename TI> int convert(trixi_hip_ctx *ctx, TO *out, const TI *in, size_t count) {
  hipLaunchKernelGGL((convert_kernel<TO, TI>), dim3(grid_for(count)), dim3(256), 0, ctx->stream, out, in, count);
  HIP_TRY(hipGetLastError());
  return TRIXI_HIP_OK;
}

// the shadow from the register, where fused stages left it behind
int sync_u32(trixi_hip_ctx *ctx) {
  if (!ctx->fp32 || !ctx->u32_stale) return TRIXI_HIP_OK;
  int r = convert<float, double>(ctx, cur_u32(ctx), cur_u64(ctx), ctx->nvals);
  if (!r) ctx->u32_stale = false;
  return r;
}

int launch_update(trixi_hip_ctx *ctx, double a_stage, double b_stage, const double *tdev, int first) {
  size_t cnt = ctx->nvals;
  if (ctx->fp32)
    hipLaunchKernelGGL((rk2n_update_kernel<float>), dim3(grid_for(cnt)), dim3(256), 0, ctx->stream,
                       (const float *)ctx->d_du, (float *)ctx->d_utmp, cur_u64(ctx), cur_u32(ctx), cnt, a_stage, b_stage,
                       tdev, first);
  else
    hipLaunchKernelGGL((rk2n_update_kernel<double>), dim3(grid_for(cnt)), dim3(256), 0, ctx->stream,
                       (const double *)ctx->d_du, (double *)ctx->d_utmp, cur_u64(ctx), (float *)nullptr, cnt, a_stage,
                       b_stage, tdev, first);
  HIP_TRY(hipGetLastError());
  return TRIXI_HIP_OK;
}

// One 2N stage (methods_2N.jl:146-159) of the resident state: the RHS at
// t (tdev == null) or tdev[0] + t tdev[1], and the update with b_dt = b_stage
// (tdev == null) or b_stage tdev[1]. With the fused kernel the update runs in
// the element kernel's epilogue and the state moves to the other buffer of
// its pair (par flips); otherwise the RHS writes du and rk2n_update_kernel
// updates in place.
int launch_stage(trixi_hip_ctx *ctx, double a_stage, double b_stage, double t, const double *tdev, int first,
                 double *speed_partial = nullptr) {
  if (!ctx->elem_stage) {
    int r = sync_u32(ctx);
    if (!r) r = launch_rhs_native(ctx, ctx->d_du, native_u(ctx), t, tdev);
    ctx->alpha_fresh = false;
    return r ? r : launch_update(ctx, a_stage, b_stage, tdev, first);
  }
  if (is_hg(ctx) && !ctx->alpha_fresh) {  // the stage's indicator reads the native state
    int r = sync_u32(ctx);
    if (r) return r;
  }
  StageOut so;
  so.u_tmp = ctx->d_utmp;
  so.u_reg_in = cur_u64(ctx);
  const bool from64 = stage_reads_u64(ctx);
  so.u_reg_out = ctx->d_ub[ctx->par ^ 1];
  so.u32_out = ctx->fp32 && !from64 ? ctx->d_u32b[ctx->par ^ 1] : nullptr;
  so.u64 = from64 ? so.u_reg_in : nullptr;
  so.a_stage = a_stage;
  so.b_stage = b_stage;
  so.first_stage = first;
  so.speed_partial = speed_partial;
  so.ind_next = stage_fuses_indicator(ctx);
  int r = ctx->fp32 ? launch_rhs_T<float>(ctx, nullptr, (const float *)native_u(ctx), t, tdev, &so)
                    : launch_rhs_T<double>(ctx, nullptr, (const double *)native_u(ctx), t, tdev, &so);
  if (r) return r;
  ctx->par ^= 1;
  if (from64) ctx->u32_stale = true;
  ctx->alpha_fresh = so.ind_next;
  return TRIXI_HIP_OK;
}

// Before a captured step is recorded or replayed: with the fused indicator the
// graph's first stage reads the factors of the state it starts from, so they
// are computed here when no fused stage left them.
int refresh_alpha(trixi_hip_ctx *ctx) {
  if (!stage_fuses_indicator(ctx) || ctx->alpha_fresh) return TRIXI_HIP_OK;
  int r = sync_u32(ctx);
  if (r) return r;
  r = ctx->fp32 ? launch_indicator_raw<float>(ctx, (const float *)native_u(ctx), ctx->d_alpha_st[ctx->par])
                    : launch_indicator_raw<double>(ctx, (const double *)native_u(ctx), ctx->d_alpha_st[ctx->par]);
  if (r) return r;
  ctx->alpha_fresh = true;
  return TRIXI_HIP_OK;
}

__global__ void set_time_kernel(double *tdt, double t, double dt) {
  if (threadIdx.x == 0) {
    tdt[0] = t;
    tdt[1] = dt;
  }
}

void drop_graphs(trixi_hip_ctx *ctx) {
  for (auto &g : ctx->graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g = trixi_hip_ctx::GraphSlot{};
  }
}

hipGraphExec_t find_graph(trixi_hip_ctx *ctx, const std::vector<double> &key) {
  for (auto &g : ctx->graphs)
    if (g.exec && g.key == key) return g.exec;
  return nullptr;
}

// Captures what record() launches on the context's stream into a slot; the
// host-side parity is restored afterwards (capture executes nothing).
template <typename F> int capture_graph(trixi_hip_ctx *ctx, const std::vector<double> &key, F &&record,
                                        hipGraphExec_t *out) {
  auto &g = ctx->graphs[ctx->graph_next];
  ctx->graph_next ^= 1;
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (g.graph) (void)hipGraphDestroy(g.graph);
  g = trixi_hip_ctx::GraphSlot{};
  const int par0 = ctx->par;
  const bool ghost0 = ctx->ghost_ready;  // a graph that is never replayed computes no boundary states
  HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  int r = record();
  hipGraph_t graph = nullptr;
  hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
  ctx->par = par0;
  if (r || ec != hipSuccess) ctx->ghost_ready = ghost0;
  if (r) {
    if (graph) (void)hipGraphDestroy(graph);
    return r;
  }
  if (ec != hipSuccess) return set_err(ctx, TRIXI_HIP_EHIP, std::string("end capture: ") + hipGetErrorString(ec));
  g.graph = graph;
  HIP_TRY(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
  g.key = key;
  *out = g.exec;
  return TRIXI_HIP_OK;
}

}  // namespace

extern "C" {

int trixi_hip_create(const trixi_hip_config *cfg, const trixi_hip_mesh *mesh,
                     const trixi_hip_basis *basis, trixi_hip_ctx **out) {
  if (!cfg || !mesh || !basis || !out) return TRIXI_HIP_EINVAL;
  *out = nullptr;
  trixi_hip_ctx *ctx = new trixi_hip_ctx();
  ctx->cfg = *cfg;
  ctx->D = cfg->ndims;
  ctx->N = cfg->polydeg;
  ctx->n = cfg->polydeg + 1;
  ctx->V = ctx->D + 2;
  ctx->E = cfg->nelements;
  ctx->I = cfg->ninterfaces;
  auto fail = [&](int code, const char *msg) {
    fprintf(stderr, "trixi_hip_create: %s\n", msg);
    delete ctx;
    return code;
  };
  if (ctx->D != 2 && ctx->D != 3) return fail(TRIXI_HIP_EINVAL, "ndims must be 2 or 3");
  ctx->B = cfg->nboundaries;
  ctx->M = cfg->nmortars;
  if (cfg->hg_variable < TRIXI_HIP_VAR_DENSITY_PRESSURE || cfg->hg_variable > TRIXI_HIP_VAR_PRESSURE)
    return fail(TRIXI_HIP_EINVAL, "hg_variable must be a TRIXI_HIP_VAR_* code");
  const int S = 1 << (ctx->D - 1);
  if (ctx->E <= 0 || ctx->B < 0 || ctx->M < 0 ||
      2 * ctx->I + ctx->B + ctx->M * (S + 1) != 2 * ctx->D * ctx->E)
    return fail(TRIXI_HIP_EINVAL,
                "every element face must be in one interface, boundary or mortar: "
                "2 ninterfaces + nboundaries + nmortars (2^(ndims-1) + 1) == 2 ndims nelements");
  if (ctx->M > 0 && (!mesh->mortar_neighbor_ids || !mesh->mortar_large_sides || !mesh->mortar_orientations ||
                     !basis->mortar_forward_upper || !basis->mortar_forward_lower || !basis->mortar_reverse_upper ||
                     !basis->mortar_reverse_lower))
    return fail(TRIXI_HIP_EINVAL, "mortar arrays or mortar matrices missing");
  if (ctx->B > 0 && cfg->boundary_condition != TRIXI_HIP_BC_DIRICHLET &&
      cfg->boundary_condition != TRIXI_HIP_BC_SLIP_WALL)
    return fail(TRIXI_HIP_EINVAL,
                "a mesh with boundaries needs boundary_condition = TRIXI_HIP_BC_DIRICHLET or TRIXI_HIP_BC_SLIP_WALL");
  ctx->slip = ctx->B > 0 && cfg->boundary_condition == TRIXI_HIP_BC_SLIP_WALL;
  if (ctx->B > 0 && (!mesh->boundary_neighbor_ids || !mesh->boundary_orientations || !mesh->boundary_neighbor_sides))
    return fail(TRIXI_HIP_EINVAL, "boundary arrays missing");
  bool dn_ok = dispatch_dn(ctx->D, ctx->N, [](auto, auto) {});
  if (!dn_ok) return fail(TRIXI_HIP_EINVAL, "unsupported (ndims, polydeg): supported are 2D/3D with polydeg 2..4");
  ctx->NN = ctx->D == 3 ? ctx->n * ctx->n * ctx->n : ctx->n * ctx->n;
  ctx->nvals = (size_t)ctx->V * ctx->NN * ctx->E;
  ctx->fp32 = cfg->precision == TRIXI_HIP_FP32;
  ctx->vol = vol_kind(*cfg);
  if (ctx->vol < 0) return fail(TRIXI_HIP_EINVAL, "unsupported volume integral / volume flux");
  if (cfg->surface_flux < 0 || cfg->surface_flux > TRIXI_HIP_FLUX_HLLC || cfg->fv_flux < 0 ||
      cfg->fv_flux > TRIXI_HIP_FLUX_HLLC)
    return fail(TRIXI_HIP_EINVAL, "unsupported surface or FV flux");
  if ((cfg->surface_flux == TRIXI_HIP_FLUX_HLLC || cfg->fv_flux == TRIXI_HIP_FLUX_HLLC) && ctx->vol != VOL_PURE_FV)
    return fail(TRIXI_HIP_EINVAL, "flux_hllc is supported with VolumeIntegralPureLGLFiniteVolume only");
  const int n = ctx->n, D = ctx->D, E = ctx->E, I = ctx->I;
  // element neighbour table from the interface container (containers_3d.jl:229-279):
  // interface i with orientation o joins left (direction 2o+1 of left) and right (2o of right)
  constexpr int32_t kUnset = INT32_MIN;
  std::vector<int32_t> nbr((size_t)2 * D * E, kUnset);
  for (int i = 0; i < I; ++i) {
    int l = mesh->neighbor_ids[2 * i], r = mesh->neighbor_ids[2 * i + 1], o = mesh->orientations[i];
    if (l < 0 || l >= E || r < 0 || r >= E || o < 0 || o >= D) return fail(TRIXI_HIP_EINVAL, "bad interface data");
    nbr[(2 * o + 1) + 2 * D * l] = r;
    nbr[(2 * o) + 2 * D * r] = l;
  }
  // boundaries (init_boundaries!, containers_2d.jl:475-543): boundary b at the
  // face of its element in direction 2o+1 (neighbor_side 1) or 2o (side 2),
  // stored as -(b + 1) in the neighbour table
  std::vector<int32_t> bnd((size_t)2 * ctx->B);
  for (int b = 0; b < ctx->B; ++b) {
    const int e = mesh->boundary_neighbor_ids[b], o = mesh->boundary_orientations[b];
    const int side = mesh->boundary_neighbor_sides[b];
    if (e < 0 || e >= E || o < 0 || o >= D || (side != 1 && side != 2))
      return fail(TRIXI_HIP_EINVAL, "bad boundary data");
    const int dir = side == 1 ? 2 * o + 1 : 2 * o;
    if (nbr[dir + 2 * D * e] != kUnset) return fail(TRIXI_HIP_EINVAL, "a face is both an interface and a boundary");
    nbr[dir + 2 * D * e] = -(b + 1);
    bnd[2 * b] = e;
    bnd[2 * b + 1] = dir;
  }
  // mortars (init_mortars!, containers_3d.jl:652-788): small element q at its face
  // towards the large one, the large element at its face towards the smalls
  std::vector<int32_t> mort((size_t)ctx->M * (S + 3));
  for (int m = 0; m < ctx->M; ++m) {
    const int o = mesh->mortar_orientations[m], side = mesh->mortar_large_sides[m];
    if (o < 0 || o >= D || (side != 1 && side != 2)) return fail(TRIXI_HIP_EINVAL, "bad mortar data");
    for (int k = 0; k <= S; ++k) {
      const int e = mesh->mortar_neighbor_ids[k + (S + 1) * m];
      if (e < 0 || e >= E) return fail(TRIXI_HIP_EINVAL, "bad mortar data");
      const bool large = k == S;
      const int dir = (side == 1) == large ? 2 * o + 1 : 2 * o;
      if (nbr[dir + 2 * D * e] != kUnset) return fail(TRIXI_HIP_EINVAL, "a mortar face is also another face");
      nbr[dir + 2 * D * e] = -(ctx->B + m * (S + 1) + k + 1);
      mort[(size_t)m * (S + 3) + k] = e;
    }
    mort[(size_t)m * (S + 3) + S + 1] = o;
    mort[(size_t)m * (S + 3) + S + 2] = side;
  }
  // element adjacency for the smoothing of alpha on refined meshes (CSR): per
  // element its mortar partners (mortar order), then its interface neighbours
  // (interface order)
  std::vector<int32_t> adj_ptr, adj;
  if (ctx->M > 0) {
    adj_ptr.assign((size_t)E + 1, 0);
    for (int m = 0; m < ctx->M; ++m) {
      const int32_t *mm = &mort[(size_t)m * (S + 3)];
      for (int q = 0; q < S; ++q) {
        ++adj_ptr[mm[q] + 1];
        ++adj_ptr[mm[S] + 1];
      }
    }
    for (int i = 0; i < I; ++i) {
      ++adj_ptr[mesh->neighbor_ids[2 * i] + 1];
      ++adj_ptr[mesh->neighbor_ids[2 * i + 1] + 1];
    }
    for (int e = 0; e < E; ++e) adj_ptr[e + 1] += adj_ptr[e];
    adj.resize(adj_ptr[E]);
    std::vector<int32_t> fill(adj_ptr.begin(), adj_ptr.end() - 1);
    for (int m = 0; m < ctx->M; ++m) {
      const int32_t *mm = &mort[(size_t)m * (S + 3)];
      for (int q = 0; q < S; ++q) {
        adj[fill[mm[q]]++] = mm[S];
        adj[fill[mm[S]]++] = mm[q];
      }
    }
    for (int i = 0; i < I; ++i) {
      const int l = mesh->neighbor_ids[2 * i], r = mesh->neighbor_ids[2 * i + 1];
      adj[fill[l]++] = r;
      adj[fill[r]++] = l;
    }
  }
  // the L2 mortar matrices serve the mortars and the AMR solution transfer
  if (basis->mortar_forward_upper && basis->mortar_forward_lower && basis->mortar_reverse_upper &&
      basis->mortar_reverse_lower) {
    ctx->mfu.assign(basis->mortar_forward_upper, basis->mortar_forward_upper + n * n);
    ctx->mfl.assign(basis->mortar_forward_lower, basis->mortar_forward_lower + n * n);
    ctx->mru.assign(basis->mortar_reverse_upper, basis->mortar_reverse_upper + n * n);
    ctx->mrl.assign(basis->mortar_reverse_lower, basis->mortar_reverse_lower + n * n);
  }
  // coordinates of the boundary face nodes: x_d = center_d + J node_{i_d}, with
  // i_o at the face and the free indices in face-node order (lower dimension fastest)
  const int Lf = D == 3 ? n * n : n;
  std::vector<double> bx((size_t)ctx->B * Lf * D);
  for (int b = 0; b < ctx->B; ++b) {
    const int e = bnd[2 * b], dir = bnd[2 * b + 1], o = dir / 2;
    const double jac = 1.0 / mesh->inverse_jacobian[e];
    for (int l = 0; l < Lf; ++l) {
      int idx[3], r = l;
      idx[o] = (dir & 1) ? n - 1 : 0;
      for (int d = 0; d < D; ++d)
        if (d != o) {
          idx[d] = r % n;
          r /= n;
        }
      for (int d = 0; d < D; ++d)
        bx[((size_t)b * Lf + l) * D + d] = std::fma(jac, basis->nodes[idx[d]], mesh->element_centers[d + D * e]);
    }
  }
  for (size_t k = 0; k < nbr.size(); ++k)
    if (nbr[k] == kUnset) return fail(TRIXI_HIP_EINVAL, "element face without interface, boundary or mortar");
  // basis
  ctx->nodes.assign(basis->nodes, basis->nodes + n);
  ctx->weights.assign(basis->weights, basis->weights + n);
  ctx->Dsplit.assign(basis->derivative_split, basis->derivative_split + n * n);
  ctx->Dhat.assign(basis->derivative_hat, basis->derivative_hat + n * n);
  ctx->Vinv.assign(basis->inverse_vandermonde_legendre, basis->inverse_vandermonde_legendre + n * n);
  ctx->na = basis->analysis_nnodes;
  if (ctx->na != 2 * ctx->N + 1) return fail(TRIXI_HIP_EINVAL, "analysis basis must have 2*polydeg+1 nodes");
  ctx->a_weights.assign(basis->analysis_weights, basis->analysis_weights + ctx->na);
  ctx->a_vdm.assign(basis->analysis_vandermonde, basis->analysis_vandermonde + ctx->na * n);
  // HG parameters (indicators.jl:130-135)
  ctx->hg.gamma = cfg->gamma;
  ctx->hg.alpha_max = cfg->alpha_max;
  ctx->hg.alpha_min = cfg->alpha_min;
  ctx->hg.threshold = 0.5 * std::pow(10.0, -1.8 * std::pow((double)n, 0.25));
  ctx->hg.parameter_s = std::log((1 - 0.0001) / 0.0001);
  ctx->hg.variable = cfg->hg_variable;
  ctx->xcd_remap = cfg->precision == TRIXI_HIP_FP32 ? 1 : 0;
  ctx->xcd_stage = 1;
  ctx->fused_speed = cfg->precision == TRIXI_HIP_FP32;
  if (const char *env = std::getenv("TRIXI_HIP_FUSED_INDICATOR")) ctx->fused_ind = std::atoi(env) != 0;
  if (const char *env = std::getenv("TRIXI_HIP_XCD_REMAP")) ctx->xcd_remap = ctx->xcd_stage = std::atoi(env) != 0;
  if (const char *env = std::getenv("TRIXI_HIP_FUSED_SPEED")) ctx->fused_speed = std::atoi(env) != 0;
  ctx->exact.kind = TRIXI_HIP_IC_CONSTANT;
  ctx->exact.gamma = cfg->gamma;

  if (ctx->fp32) build_kconst<float>(ctx);
  else build_kconst<double>(ctx);
  if (hipGetDevice(&ctx->device) != hipSuccess) return fail(TRIXI_HIP_EHIP, "hipGetDevice failed");
  if (select_tile(ctx, cfg->elements_per_block, cfg->threads_per_block) != TRIXI_HIP_OK) {
    fprintf(stderr, "trixi_hip_create: %s\n", ctx->err.c_str());
    delete ctx;
    return TRIXI_HIP_EINVAL;
  }

  if (stream_acquire(ctx->device, &ctx->own_stream) != hipSuccess)
    return fail(TRIXI_HIP_EHIP, "hipStreamCreate failed");
  ctx->stream = ctx->own_stream;
  const size_t nv = ctx->nvals;
  const size_t tsz = ctx->fp32 ? 4 : 8;
  ctx->partial_cap = std::max<size_t>((size_t)E * 2 * ctx->V, kMaxPartials);
  bool ok = dev_alloc(ctx, &ctx->d_nbr, nbr.size() * 4) == hipSuccess &&
            dev_alloc(ctx, &ctx->d_invj, E * 8) == hipSuccess &&
            dev_alloc(ctx, &ctx->d_centers, (size_t)D * E * 8) == hipSuccess &&
            dev_alloc(ctx, &ctx->d_ub[0], nv * 8) == hipSuccess && dev_alloc(ctx, &ctx->d_utmp, nv * tsz) == hipSuccess &&
            dev_alloc(ctx, &ctx->d_ub[1], nv * 8) == hipSuccess &&
            dev_alloc(ctx, &ctx->d_du, nv * tsz) == hipSuccess &&
            dev_alloc(ctx, &ctx->d_alpha_raw, E * 8) == hipSuccess && dev_alloc(ctx, &ctx->d_alpha, E * 8) == hipSuccess &&
            dev_alloc(ctx, &ctx->d_alpha_st[0], E * 8) == hipSuccess &&
            dev_alloc(ctx, &ctx->d_alpha_st[1], E * 8) == hipSuccess &&
            dev_alloc(ctx, &ctx->d_partial, ctx->partial_cap * 8) == hipSuccess &&
            dev_alloc(ctx, &ctx->d_scalars, 64 * 8) == hipSuccess;
  // face slots: boundaries, then mortar faces; ghost states of the mortar slots
  // are a valid placeholder state (the element kernel replaces their fluxes)
  const size_t nslots = (size_t)ctx->B + (size_t)ctx->M * (S + 1);
  const size_t slot_vals = nslots * Lf * ctx->V;
  if (ok && ctx->B > 0) ok = dev_alloc(ctx, &ctx->d_bx, bx.size() * 8) == hipSuccess;
  if (ok && nslots > 0) ok = dev_alloc(ctx, &ctx->d_ghost, slot_vals * tsz) == hipSuccess;
  if (ok && (ctx->M > 0 || ctx->slip)) ok = dev_alloc(ctx, &ctx->d_mflux, slot_vals * tsz) == hipSuccess;
  if (ok && ctx->slip) ok = dev_alloc(ctx, &ctx->d_bnd, bnd.size() * 4) == hipSuccess;
  if (ok && ctx->M > 0)
    ok = dev_alloc(ctx, &ctx->d_mort, mort.size() * 4) == hipSuccess &&
         dev_alloc(ctx, &ctx->d_adj_ptr, adj_ptr.size() * 4) == hipSuccess &&
         dev_alloc(ctx, &ctx->d_adj, std::max<size_t>(adj.size(), 1) * 4) == hipSuccess;
  if (ok && ctx->fp32)
    ok = dev_alloc(ctx, &ctx->d_u32b[0], nv * 4) == hipSuccess && dev_alloc(ctx, &ctx->d_u32b[1], nv * 4) == hipSuccess &&
         dev_alloc(ctx, &ctx->d_su32, nv * 4) == hipSuccess &&
         dev_alloc(ctx, &ctx->d_sdu32, nv * 4) == hipSuccess;
  if (!ok) {
    trixi_hip_destroy(ctx);
    fprintf(stderr, "trixi_hip_create: device allocation failed\n");
    return TRIXI_HIP_ENOMEM;
  }
  // uploads on the context's stream, one synchronisation at the end (the host
  // arrays live until then)
  auto up = [&](void *dst, const void *src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream) == hipSuccess;
  };
  ok = up(ctx->d_nbr, nbr.data(), nbr.size() * 4) && up(ctx->d_invj, mesh->inverse_jacobian, E * 8) &&
       up(ctx->d_centers, mesh->element_centers, (size_t)D * E * 8) &&
       hipMemsetAsync(ctx->d_alpha_raw, 0, E * 8, ctx->stream) == hipSuccess &&
       hipMemsetAsync(ctx->d_alpha, 0, E * 8, ctx->stream) == hipSuccess &&
       up(ctx->d_scalars + 32, ctx->weights.data(), n * 8) && (ctx->B == 0 || up(ctx->d_bx, bx.data(), bx.size() * 8));
  std::vector<double> g;
  std::vector<float> gf;
  if (ok && nslots > 0) {
    g.resize(slot_vals);
    for (size_t k = 0; k < slot_vals; ++k) {
      const int v = (int)(k % ctx->V);
      g[k] = v == 0 ? 1.0 : (v == ctx->V - 1 ? 2.5 : 0.0);  // rho = 1, v = 0, p = 1 (gamma 1.4)
    }
    if (ctx->fp32) {
      gf.assign(g.begin(), g.end());
      ok = up(ctx->d_ghost, gf.data(), slot_vals * 4);
    } else {
      ok = up(ctx->d_ghost, g.data(), slot_vals * 8);
    }
  }
  if (ok && ctx->slip) ok = up(ctx->d_bnd, bnd.data(), bnd.size() * 4);
  if (ok && ctx->M > 0)
    ok = up(ctx->d_mort, mort.data(), mort.size() * 4) && up(ctx->d_adj_ptr, adj_ptr.data(), adj_ptr.size() * 4) &&
         (adj.empty() || up(ctx->d_adj, adj.data(), adj.size() * 4));
  ok = ok && host_sync(ctx) == hipSuccess;
  if (!ok) {
    trixi_hip_destroy(ctx);
    fprintf(stderr, "trixi_hip_create: upload failed\n");
    return TRIXI_HIP_EHIP;
  }
  *out = ctx;
  return TRIXI_HIP_OK;
}

int trixi_hip_destroy(trixi_hip_ctx *ctx) {
  if (!ctx) return TRIXI_HIP_OK;
  if (ctx->stream) (void)host_sync(ctx);
  if (ctx->own_stream && ctx->own_stream != ctx->stream) (void)hipStreamSynchronize(ctx->own_stream);
  drop_graphs(ctx);
  void *ptrs[] = {ctx->d_nbr, ctx->d_invj, ctx->d_centers, ctx->d_ub[0], ctx->d_ub[1], ctx->d_utmp,
                  ctx->d_u32b[0], ctx->d_u32b[1], ctx->d_du,
                  ctx->d_hu, ctx->d_hdu, ctx->d_su32, ctx->d_sdu32, ctx->d_alpha_raw, ctx->d_alpha,
                  ctx->d_alpha_st[0], ctx->d_alpha_st[1],
                  ctx->d_partial, ctx->d_scalars, ctx->d_mlp, ctx->d_feat, ctx->d_mlp_partial,
                  ctx->d_alpha_hg, ctx->d_mlp_img, ctx->d_records, ctx->d_bx, ctx->d_ghost, ctx->d_mort,
                  ctx->d_mflux, ctx->d_bnd,
                  ctx->d_adj_ptr, ctx->d_adj};
  for (void *p : ptrs) dev_release(ctx, p);
  while (!ctx->blocks.empty()) dev_release(ctx, ctx->blocks.begin()->first);
  if (ctx->own_stream) stream_release(ctx->device, ctx->own_stream);
  delete ctx;
  return TRIXI_HIP_OK;
}

const char *trixi_hip_last_error(const trixi_hip_ctx *ctx) {
  if (ctx) return ctx->err.c_str();
  return g_free_err[0] ? g_free_err : "null context";
}

int trixi_hip_set_stream(trixi_hip_ctx *ctx, void *hip_stream) {
  if (!ctx) return TRIXI_HIP_EINVAL;
  ctx->stream = hip_stream ? (hipStream_t)hip_stream : ctx->own_stream;
  return TRIXI_HIP_OK;
}

int trixi_hip_synchronize(trixi_hip_ctx *ctx) {
  if (!ctx) return TRIXI_HIP_EINVAL;
  HIP_TRY(host_sync(ctx));
  return TRIXI_HIP_OK;
}

int trixi_hip_upload_u(trixi_hip_ctx *ctx, const double *u_host) {
  if (!ctx || !u_host) return TRIXI_HIP_EINVAL;
  ctx->alpha_fresh = false;
  HIP_TRY(hipMemcpyAsync(cur_u64(ctx), u_host, ctx->nvals * 8, hipMemcpyHostToDevice, ctx->stream));
  if (ctx->fp32) {
    int r = convert<float, double>(ctx, cur_u32(ctx), cur_u64(ctx), ctx->nvals);
    if (r) return r;
    ctx->u32_stale = false;
  }
  HIP_TRY(host_sync(ctx));
  return TRIXI_HIP_OK;
}

int trixi_hip_download_u(trixi_hip_ctx *ctx, double *u_host) {
  if (!ctx || !u_host) return TRIXI_HIP_EINVAL;
  HIP_TRY(hipMemcpyAsync(u_host, cur_u64(ctx), ctx->nvals * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(host_sync(ctx));
  return TRIXI_HIP_OK;
}

int trixi_hip_set_u_device(trixi_hip_ctx *ctx, const double *u_dev) {
  if (!ctx || !u_dev) return TRIXI_HIP_EINVAL;
  ctx->alpha_fresh = false;
  HIP_TRY(hipMemcpyAsync(cur_u64(ctx), u_dev, ctx->nvals * 8, hipMemcpyDeviceToDevice, ctx->stream));
  if (ctx->fp32) {
    int r = convert<float, double>(ctx, cur_u32(ctx), cur_u64(ctx), ctx->nvals);
    if (r) return r;
    ctx->u32_stale = false;
  }
  return TRIXI_HIP_OK;
}

int trixi_hip_get_u_device(trixi_hip_ctx *ctx, double *u_dev) {
  if (!ctx || !u_dev) return TRIXI_HIP_EINVAL;
  HIP_TRY(hipMemcpyAsync(u_dev, cur_u64(ctx), ctx->nvals * 8, hipMemcpyDeviceToDevice, ctx->stream));
  return TRIXI_HIP_OK;
}

int trixi_hip_rhs(trixi_hip_ctx *ctx, double *du, const double *u, double t, int where) {
  if (!ctx || !du || !u) return TRIXI_HIP_EINVAL;
  const size_t nv = ctx->nvals;
  const double *du_in = u;
  double *du_out = du;
  if (where == TRIXI_HIP_HOST) {
    if (!ctx->d_hu) {
      HIP_TRY(dev_alloc(ctx, &ctx->d_hu, nv * 8));
      HIP_TRY(dev_alloc(ctx, &ctx->d_hdu, nv * 8));
    }
    HIP_TRY(hipMemcpyAsync(ctx->d_hu, u, nv * 8, hipMemcpyHostToDevice, ctx->stream));
    du_in = ctx->d_hu;
    du_out = ctx->d_hdu;
  } else if (where != TRIXI_HIP_DEVICE) {
    return set_err(ctx, TRIXI_HIP_EINVAL, "where must be TRIXI_HIP_HOST or TRIXI_HIP_DEVICE");
  } else if (((uintptr_t)u % 16) || ((uintptr_t)du % 16)) {
    return set_err(ctx, TRIXI_HIP_EINVAL, "device arrays must be 16-byte aligned");
  }
  int r;
  if (ctx->fp32) {
    if ((r = convert<float, double>(ctx, ctx->d_su32, du_in, nv))) return r;
    if ((r = launch_rhs_T<float>(ctx, ctx->d_sdu32, ctx->d_su32, t, nullptr))) return r;
    if ((r = convert<double, float>(ctx, du_out, ctx->d_sdu32, nv))) return r;
  } else {
    if ((r = launch_rhs_T<double>(ctx, du_out, du_in, t, nullptr))) return r;
  }
  if (where == TRIXI_HIP_HOST) {
    HIP_TRY(hipMemcpyAsync(du, du_out, nv * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(host_sync(ctx));
  }
  return TRIXI_HIP_OK;
}

int trixi_hip_rhs_native(trixi_hip_ctx *ctx, void *du_dev, const void *u_dev, double t) {
  if (!ctx || !du_dev || !u_dev) return TRIXI_HIP_EINVAL;
  if (((uintptr_t)u_dev % 16) || ((uintptr_t)du_dev % 16))
    return set_err(ctx, TRIXI_HIP_EINVAL, "device arrays must be 16-byte aligned");
  return launch_rhs_native(ctx, du_dev, u_dev, t, nullptr);
}

int trixi_hip_rk_stage(trixi_hip_ctx *ctx, double a_stage, double b_stage_dt, double t_stage,
                       int first_stage) {
  if (!ctx) return TRIXI_HIP_EINVAL;
  return launch_stage(ctx, a_stage, b_stage_dt, t_stage, nullptr, first_stage);
}

int trixi_hip_rk_step(trixi_hip_ctx *ctx, const double *a, const double *b, const double *c,
                      int nstages, double t, double dt, int use_graph) {
  if (!ctx || !a || !b || !c || nstages <= 0 || nstages > 16) return TRIXI_HIP_EINVAL;
  double *tdt = ctx->d_scalars;
  hipLaunchKernelGGL(set_time_kernel, dim3(1), dim3(1), 0, ctx->stream, tdt, t, dt);
  HIP_TRY(hipGetLastError());
  auto record = [&]() -> int {
    for (int s = 0; s < nstages; ++s) {
      // stage kernels read t and dt from device memory; ElemArgs.t carries c_stage
      int r = launch_stage(ctx, a[s], b[s], c[s], tdt, s == 0);
      if (r) return r;
    }
    return TRIXI_HIP_OK;
  };
  if (!use_graph) return record();
  int r0 = refresh_alpha(ctx);  // the graph's first stage reads the current factors
  if (r0) return r0;
  std::vector<double> key{-2.0, (double)nstages};
  for (int s = 0; s < nstages; ++s) {
    key.push_back(a[s]);
    key.push_back(b[s]);
    key.push_back(c[s]);
  }
  key.push_back((double)(uintptr_t)ctx->stream);
  key.push_back((double)ctx->par);  // buffers the captured stages read and write
  hipGraphExec_t x = find_graph(ctx, key);
  if (!x) {
    int r = capture_graph(ctx, key, record, &x);
    if (r) return r;
  }
  HIP_TRY(hipGraphLaunch(x, ctx->stream));
  if (ctx->elem_stage) {
    ctx->par ^= (nstages & 1);
    if (stage_reads_u64(ctx)) ctx->u32_stale = true;
  }
  return TRIXI_HIP_OK;
}

}  // extern "C"

namespace {

// Partial maxima of the scaled wave speed of the resident FP64 state into
// d_partial (one per block); returns the number of partials, or < 0.
int launch_max_speed(trixi_hip_ctx *ctx) {
  int nb = -1;
  dispatch_dn(ctx->D, ctx->N, [&](auto Dc, auto Nc) {
    constexpr int D = decltype(Dc)::value, N = decltype(Nc)::value;
    constexpr int NN = ipow_c(N + 1, D);
    constexpr int EB = NN >= 256 ? 1 : 256 / NN;
    nb = std::min((ctx->E + EB - 1) / EB, kMaxSpeedBlocks);  // beyond that, several groups per block
    if ((size_t)nb > ctx->partial_cap) {
      nb = -1;
      return;
    }
    hipLaunchKernelGGL((max_speed_nodes_kernel<double, D, N>), dim3(nb), dim3(256), 0, ctx->stream, cur_u64(ctx),
                       ctx->d_invj, ctx->E, ctx->cfg.gamma, ctx->d_partial);
  });
  if (nb < 0) return set_err(ctx, TRIXI_HIP_EINVAL, "too many elements for the reduction buffer");
  HIP_TRY(hipGetLastError());
  return nb;
}

// d_scalars layout: [0] t, [1] dt, [2] max_dt result, [8, 8 + 2V) error-norm
// sums, [20, 21 + V) integral sums, [32, 32 + n) LGL weights, [48, 56) step
// control {t_end, cfl, last, done, steps, error, max_steps, analyses made}
constexpr int kErrSums = 8, kIntSums = 20, kWeights = 32, kCtl = 48;

// one analysis record: {t, iter, l2 sums [V], linf [V], integral sums [V], entropy sum}
int rec_width(const trixi_hip_ctx *ctx) { return 3 + 3 * ctx->V; }

// total_volume (analysis_dg3d.jl:123-161 normalisation, integrate with
// normalize=true), evaluated once in the two forms the reductions use.
int total_volumes(trixi_hip_ctx *ctx) {
  if (ctx->tv_err > 0) return TRIXI_HIP_OK;
  std::vector<double> ij(ctx->E);
  HIP_TRY(sync_copy(ctx, ij.data(), ctx->d_invj, ctx->E * 8, hipMemcpyDeviceToHost));
  double a = 0, b = 0;
  for (int e = 0; e < ctx->E; ++e) {
    a += std::pow(1.0 / ij[e], ctx->D) * std::pow(2.0, ctx->D);
    b += std::pow(2.0 / ij[e], ctx->D);
  }
  ctx->tv_err = a;
  ctx->tv_int = b;
  return TRIXI_HIP_OK;
}

// calc_error_norms sums of the resident FP64 state at time t (tdev[0] if tdev)
// -> out[0, 2V): {sum w J (u - u_exact)^2 [V], max |u - u_exact| [V]}.
int enqueue_error_sums(trixi_hip_ctx *ctx, double t, const double *tdev, double *out) {
  const int V = ctx->V, E = ctx->E;
  if ((size_t)E * 2 * V > ctx->partial_cap) return set_err(ctx, TRIXI_HIP_EINVAL, "partials too small");
  bool ok = dispatch_dn(ctx->D, ctx->N, [&](auto Dc, auto Nc) {
    constexpr int D = decltype(Dc)::value, N = decltype(Nc)::value;
    constexpr int NA = 2 * N + 1;
    AnalysisConst<N, NA> ac{};
    for (int i = 0; i < NA * (N + 1); ++i) ac.vdm[i] = ctx->a_vdm[i];
    for (int i = 0; i < NA; ++i) ac.aw[i] = ctx->a_weights[i];
    for (int i = 0; i <= N; ++i) ac.nodes[i] = ctx->nodes[i];
    hipLaunchKernelGGL((error_norms_kernel<double, D, N, NA>), dim3(E), dim3(256), 0, ctx->stream, cur_u64(ctx),
                       ctx->d_invj, ctx->d_centers, E, t, tdev, ctx->exact, ac, ctx->d_partial);
  });
  if (!ok) return TRIXI_HIP_EINVAL;
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(reduce_columns_kernel, dim3(1), dim3(256), 0, ctx->stream, ctx->d_partial, E, 2 * V, V, out);
  HIP_TRY(hipGetLastError());
  return TRIXI_HIP_OK;
}

// integrate_via_indices sums of u and the mathematical entropy -> out[0, V + 1).
int enqueue_integral_sums(trixi_hip_ctx *ctx, double *out) {
  const int V = ctx->V, E = ctx->E;
  if ((size_t)E * (V + 1) > ctx->partial_cap) return set_err(ctx, TRIXI_HIP_EINVAL, "partials too small");
  bool ok = dispatch_dn(ctx->D, ctx->N, [&](auto Dc, auto Nc) {
    constexpr int D = decltype(Dc)::value, N = decltype(Nc)::value;
    hipLaunchKernelGGL((integrals_kernel<double, D, N>), dim3((E + 255) / 256), dim3(256), 0, ctx->stream,
                       cur_u64(ctx), ctx->d_invj, E, ctx->cfg.gamma, ctx->d_scalars + kWeights, ctx->d_partial);
  });
  if (!ok) return TRIXI_HIP_EINVAL;
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(reduce_columns_kernel, dim3(1), dim3(256), 0, ctx->stream, ctx->d_partial, E, V + 1, V + 1,
                     out);
  HIP_TRY(hipGetLastError());
  return TRIXI_HIP_OK;
}

// Appends {t, steps, sums} of the current device state to the record ring.
__global__ void record_kernel(const double *__restrict__ scalars, double *__restrict__ ctl,
                              double *__restrict__ records, int cap, int V) {
  if (threadIdx.x != 0) return;
  const long long k = (long long)ctl[7];
  double *rec = records + (size_t)(k % cap) * (3 + 3 * V);
  rec[0] = scalars[0];
  rec[1] = ctl[4];
  for (int i = 0; i < 2 * V; ++i) rec[2 + i] = scalars[kErrSums + i];
  for (int i = 0; i <= V; ++i) rec[2 + 2 * V + i] = scalars[kIntSums + i];
  ctl[7] = double(k + 1);
}

// One analysis of the running solve, at the device time, without host sync.
int enqueue_record(trixi_hip_ctx *ctx) {
  int r = enqueue_error_sums(ctx, 0.0, ctx->d_scalars, ctx->d_scalars + kErrSums);
  if (!r) r = enqueue_integral_sums(ctx, ctx->d_scalars + kIntSums);
  if (r) return r;
  hipLaunchKernelGGL(record_kernel, dim3(1), dim3(64), 0, ctx->stream, ctx->d_scalars, ctx->d_scalars + kCtl,
                     ctx->d_records, ctx->rec_cap, ctx->V);
  HIP_TRY(hipGetLastError());
  return TRIXI_HIP_OK;
}

__global__ void solve_init_kernel(double *tdt, double *ctl, double t, double dt, double t_end, double cfl,
                                  double last, double max_steps) {
  if (threadIdx.x == 0) {
    tdt[0] = t;
    tdt[1] = dt;
    ctl[0] = t_end;
    ctl[1] = cfl;
    ctl[2] = last;
    ctl[3] = 0.0;
    ctl[4] = 0.0;
    ctl[5] = 0.0;
    ctl[6] = max_steps;
    ctl[7] = 0.0;
  }
}

}  // namespace

extern "C" {

int trixi_hip_max_dt(trixi_hip_ctx *ctx, double cfl, double *dt) {
  if (!ctx || !dt) return TRIXI_HIP_EINVAL;
  const int nb = launch_max_speed(ctx);
  if (nb < 0) return nb;
  hipLaunchKernelGGL(max_dt_finalize_kernel, dim3(1), dim3(kRedThreads), 0, ctx->stream, ctx->d_partial, nb, ctx->n,
                     cfl, ctx->d_scalars + 2);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(dt, ctx->d_scalars + 2, 8, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(host_sync(ctx));
  return TRIXI_HIP_OK;
}

}  // extern "C"

namespace {

// Scratch of trixi_hip_max_dt_noalloc: the partial maxima of the wave-speed
// reduction, one per workgroup (64 KiB). Every device the library runs on
// holds its own copy until the process ends. In host code `scratch` names a
// host-side stand-in, never the device memory: kernels get the per-device
// address that hipGetSymbolAddress returns (scratch_on_current_device).
constexpr int kNoAllocPartials = 8192;
static __device__ __align__(256) double scratch[kNoAllocPartials];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};
// Device address of `scratch` on the current device, looked up once per device.
int scratch_on_current_device(double **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return free_err(TRIXI_HIP_EHIP, "hipGetDevice failed", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return free_err(TRIXI_HIP_EINVAL, "device ordinal beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return free_err(TRIXI_HIP_EHIP, "hipGetSymbolAddress(scratch) failed", hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return TRIXI_HIP_OK;
}

}  // namespace

extern "C" {

int trixi_hip_max_dt_noalloc(int ndims, int polydeg, int nelements, double gamma, const double *u_dev,
                             const double *inverse_jacobian_dev, double cfl, double *dt_dev, void *hip_stream) {
  if (!u_dev || !inverse_jacobian_dev || !dt_dev)
    return free_err(TRIXI_HIP_EINVAL, "max_dt_noalloc: null device pointer");
  if (nelements <= 0 || nelements > INT_MAX - 256)
    return free_err(TRIXI_HIP_EINVAL, "max_dt_noalloc: nelements must be in [1, 2^31 - 257]");
  if (!dispatch_dn(ndims, polydeg, [](auto, auto) {}))
    return free_err(TRIXI_HIP_EINVAL, "max_dt_noalloc: ndims must be 2 or 3 and polydeg 2, 3 or 4");
  double *partial = nullptr;
  int r = scratch_on_current_device(&partial);
  if (r) return r;
  hipStream_t s = static_cast<hipStream_t>(hip_stream);
  int nb = -1, n = polydeg + 1;
  dispatch_dn(ndims, polydeg, [&](auto Dc, auto Nc) {
    constexpr int D = decltype(Dc)::value, N = decltype(Nc)::value;
    constexpr int NN = ipow_c(N + 1, D);
    constexpr int EB = NN >= 256 ? 1 : 256 / NN;
    // more element groups than partials: each workgroup walks several groups
    nb = std::min((nelements + EB - 1) / EB, kNoAllocPartials);
    hipLaunchKernelGGL((max_speed_nodes_kernel<double, D, N>), dim3(nb), dim3(256), 0, s, u_dev,
                       inverse_jacobian_dev, nelements, gamma, partial);
  });
  if (nb <= 0 || nb > kNoAllocPartials)
    return free_err(TRIXI_HIP_EINVAL, "max_dt_noalloc: more partials than the 8192 of the scratch buffer");
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return free_err(TRIXI_HIP_EHIP, "max_dt_noalloc: launch failed", hipGetErrorString(e));
  // reads exactly the nb partials written above: nothing of an earlier call is used
  hipLaunchKernelGGL(max_dt_finalize_kernel, dim3(1), dim3(kRedThreads), 0, s, partial, nb, n, cfl, dt_dev);
  e = hipGetLastError();
  if (e != hipSuccess) return free_err(TRIXI_HIP_EHIP, "max_dt_noalloc: launch failed", hipGetErrorString(e));
  g_free_err[0] = '\0';
  return TRIXI_HIP_OK;
}

int trixi_hip_solve(trixi_hip_ctx *ctx, const double *a, const double *b, const double *c, int nstages,
                    double t0, double t_end, double cfl, int max_steps, int poll_every, double *t_out,
                    int *steps_out) {
  if (!ctx || !a || !b || !c || nstages <= 0 || nstages > 16 || !(cfl > 0) || max_steps <= 0 ||
      poll_every <= 0 || !(t_end >= t0))
    return TRIXI_HIP_EINVAL;
  if (t_out) *t_out = t0;
  if (steps_out) *steps_out = 0;
  double *tdt = ctx->d_scalars, *ctl = ctx->d_scalars + kCtl;
  // first step size on the host side of the stream (dt = cfl max_dt(u0)) and its limit_dt
  double dt = 0;
  int r = trixi_hip_max_dt(ctx, cfl, &dt);
  if (r) return r;
  if (dt != dt) return set_err(ctx, TRIXI_HIP_ESTATE, "time step size `dt` is NaN");
  double last = 0;
  {
    const double s = t0 + dt;
    if (s > t_end || s == t_end || std::fabs(s - t_end) <= 1.4901161193847656e-08 * std::fmax(std::fabs(s), std::fabs(t_end))) {
      dt = t_end - t0;
      last = 1;
    }
  }
  hipLaunchKernelGGL(solve_init_kernel, dim3(1), dim3(1), 0, ctx->stream, tdt, ctl, t0, dt, t_end, cfl, last,
                     (double)max_steps);
  HIP_TRY(hipGetLastError());
  // one step = the stages + the wave-speed reduction + the step control, one
  // graph per (scheme, stream, start parity of the state buffers)
  std::vector<double> key{-1.0, (double)nstages};
  for (int s = 0; s < nstages; ++s) {
    key.push_back(a[s]);
    key.push_back(b[s]);
    key.push_back(c[s]);
  }
  key.push_back((double)(uintptr_t)ctx->stream);
  // the last stage reduces the wave speed of the new state in its epilogue
  // when its variant can (else max_speed_nodes_kernel after it); both give
  // the same partial maxima element by element
  const bool fused_speed = ctx->fused_speed && stage_speed_ok(ctx) &&
                           (size_t)((ctx->E + ctx->epb_st - 1) / ctx->epb_st) <= ctx->partial_cap;
  auto record = [&]() -> int {
    for (int s = 0; s < nstages; ++s) {
      int rr = launch_stage(ctx, a[s], b[s], c[s], tdt, s == 0,
                            fused_speed && s == nstages - 1 ? ctx->d_partial : nullptr);
      if (rr) return rr;
    }
    const int nb = fused_speed ? (ctx->E + ctx->epb_st - 1) / ctx->epb_st : launch_max_speed(ctx);
    if (nb < 0) return nb;
    hipLaunchKernelGGL(step_control_kernel, dim3(1), dim3(kRedThreads), 0, ctx->stream, ctx->d_partial, nb, ctx->n, tdt,
                       ctl);
    HIP_TRY(hipGetLastError());
    return TRIXI_HIP_OK;
  };
  // with the fused stage an odd stage count ends a step in the other buffer:
  // consecutive replays alternate between the graphs of the two parities
  const int flip = ctx->elem_stage ? (nstages & 1) : 0;
  auto step_graph = [&](hipGraphExec_t *x) -> int {
    std::vector<double> k = key;
    k.push_back((double)ctx->par);
    *x = find_graph(ctx, k);
    return *x ? TRIXI_HIP_OK : capture_graph(ctx, k, record, x);
  };
  // A solve of a few steps (an AMR run solves up to the next adaptation, each
  // time on a new context) launches the steps directly: capturing and
  // instantiating the step graphs would cost more than the replays save. The
  // kernels and their arguments are the same; record() leaves the parity
  // flipped as a replay does.
  const bool eager = max_steps <= kEagerSolveSteps;
  if ((r = refresh_alpha(ctx))) return r;  // the step graphs' first stage reads the current factors
  if (!eager) {
    hipGraphExec_t x = nullptr;
    if ((r = step_graph(&x))) return r;
    if (flip) {
      ctx->par ^= 1;
      r = step_graph(&x);
      ctx->par ^= 1;
      if (r) return r;
    }
  }
  // Replay without host synchronisation, reading the control block between
  // batches. A batch is at most poll_every steps and, near the end, no more than
  // the steps left at the current dt (floor), so few replays run after the end.
  // With analysis on (trixi_hip_set_analysis), batches end at multiples of the
  // interval, where the analysis is enqueued behind them; the first record is
  // the initial state and the last one the final state (analysis.jl:121-122,
  // 143-230).
  const int every = ctx->d_records ? ctx->rec_interval : 0;
  if (every > 0 && (r = enqueue_record(ctx))) return r;
  double h[8] = {0}, cur[2] = {t0, dt};
  int launched = 0;
  bool recorded = false;  // a record was enqueued behind the last batch
  for (;;) {
    const double left = (t_end - cur[0]) / cur[1];  // > 0: not done, dt > 0
    int batch = left >= poll_every ? poll_every : (left >= 1.0 ? (int)left : 1);
    batch = std::min(batch, std::max(1, max_steps - (int)h[4]));
    if (every > 0) batch = std::min(batch, every - (int)h[4] % every);
    for (int i = 0; i < batch; ++i) {
      if (eager) {
        if ((r = record())) return r;
        continue;
      }
      hipGraphExec_t x = nullptr;
      if ((r = step_graph(&x))) return r;
      HIP_TRY(hipGraphLaunch(x, ctx->stream));
      ctx->par ^= flip;
      if (ctx->elem_stage && stage_reads_u64(ctx)) ctx->u32_stale = true;
    }
    launched += batch;
    recorded = every > 0 && ((int)h[4] + batch) % every == 0;
    if (recorded && (r = enqueue_record(ctx))) return r;
    HIP_TRY(hipMemcpyAsync(h, ctl, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(cur, tdt, sizeof(cur), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(host_sync(ctx));
    if (t_out) *t_out = cur[0];
    if (steps_out) *steps_out = (int)h[4];
    if (h[5] != 0.0) return set_err(ctx, TRIXI_HIP_ESTATE, "time step size `dt` is NaN");
    if (h[3] != 0.0) {
      if (every > 0 && !recorded) {  // the final state, if the last batch did not record it
        if ((r = enqueue_record(ctx))) return r;
        HIP_TRY(host_sync(ctx));
      }
      break;
    }
    if (launched > max_steps) return set_err(ctx, TRIXI_HIP_ESTATE, "step control did not finish");
  }
  return TRIXI_HIP_OK;
}

int trixi_hip_indicator(trixi_hip_ctx *ctx, int mode, double *alpha_host) {
  if (!ctx || mode < TRIXI_HIP_IND_HG || mode > TRIXI_HIP_IND_MLP_WITH_FALLBACK) return TRIXI_HIP_EINVAL;
  if (mode != TRIXI_HIP_IND_HG && !ctx->mlp_ready)
    return set_err(ctx, TRIXI_HIP_ESTATE, "the learned indicator has no parameters (trixi_hip_mlp_set_params)");
  int r = sync_u32(ctx);
  if (r) return r;
  const int saved = ctx->ind_mode;
  ctx->ind_mode = mode;
  r = ctx->fp32 ? launch_indicator_raw<float>(ctx, cur_u32(ctx), ctx->d_alpha_raw)
                    : launch_indicator_raw<double>(ctx, cur_u64(ctx), ctx->d_alpha_raw);
  ctx->ind_mode = saved;
  if (r) return r;
  bool ok = true;
  if (ctx->M > 0 && ctx->cfg.alpha_smooth) {  // neighbours across mortars too
    hipLaunchKernelGGL(hg_smooth_adj_kernel, dim3((ctx->E + 255) / 256), dim3(256), 0, ctx->stream,
                       ctx->d_alpha_raw, ctx->d_adj_ptr, ctx->d_adj, ctx->d_alpha, ctx->E);
  } else {
    ok = dispatch_dn(ctx->D, ctx->N, [&](auto Dc, auto) {
      constexpr int D = decltype(Dc)::value;
      hipLaunchKernelGGL((hg_smooth_kernel<D>), dim3((ctx->E + 255) / 256), dim3(256), 0, ctx->stream,
                         ctx->d_alpha_raw, ctx->d_nbr, ctx->d_alpha, ctx->E, ctx->cfg.alpha_smooth);
    });
  }
  if (!ok) return TRIXI_HIP_EINVAL;
  HIP_TRY(hipGetLastError());
  if (alpha_host) {
    HIP_TRY(hipMemcpyAsync(alpha_host, ctx->d_alpha, ctx->E * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(host_sync(ctx));
  }
  return TRIXI_HIP_OK;
}

// ---------------------------------------------------------------------------
// Learned indicator
// ---------------------------------------------------------------------------
static int mlp_check(trixi_hip_ctx *ctx) {
  if (ctx->D != 2 || ctx->N != 3)
    return set_err(ctx, TRIXI_HIP_EINVAL, "the learned indicator is defined for 2D elements of polydeg 3");
  if (ctx->d_mlp) return TRIXI_HIP_OK;
  const size_t nmlp = 3 * (size_t)mlp::NP + 2 * mlp::NF + mlp::NPL;
  HIP_TRY(dev_alloc(ctx, &ctx->d_mlp, nmlp * 4));
  HIP_TRY(dev_alloc(ctx, &ctx->d_feat, (size_t)ctx->E * kMlpFeatureStride * 4));
  HIP_TRY(dev_alloc(ctx, &ctx->d_alpha_hg, (size_t)ctx->E * 8));
  HIP_TRY(dev_alloc(ctx, &ctx->d_mlp_partial, (size_t)kMlpMaxWaves * mlp::NPL * 4));
  HIP_TRY(dev_alloc(ctx, &ctx->d_mlp_img, sizeof(mlp::InferImage)));
  return TRIXI_HIP_OK;
}

// Refresh the BF16 inference image after the parameters changed.
static int mlp_pack(trixi_hip_ctx *ctx) {
  ctx->alpha_fresh = false;  // the learned factors change with the parameters
  hipLaunchKernelGGL(mlp::mlp_pack_kernel, dim3(1), dim3(256), 0, ctx->stream, ctx->d_mlp, ctx->d_mlp_img);
  HIP_TRY(hipGetLastError());
  return TRIXI_HIP_OK;
}

static float *mlp_grad_buffer(trixi_hip_ctx *ctx) { return ctx->d_mlp + 3 * mlp::NP + 2 * mlp::NF; }

int trixi_hip_mlp_set_params(trixi_hip_ctx *ctx, const float *params, const float *feature_lo,
                             const float *feature_hi) {
  if (!ctx || !params || !feature_lo || !feature_hi) return TRIXI_HIP_EINVAL;
  int r = mlp_check(ctx);
  if (r) return r;
  std::vector<float> h(3 * (size_t)mlp::NP + 2 * mlp::NF, 0.f);
  std::memcpy(h.data(), params, mlp::NP * 4);
  std::memcpy(h.data() + 3 * mlp::NP, feature_lo, mlp::NF * 4);
  std::memcpy(h.data() + 3 * mlp::NP + mlp::NF, feature_hi, mlp::NF * 4);
  HIP_TRY(hipMemcpyAsync(ctx->d_mlp, h.data(), h.size() * 4, hipMemcpyHostToDevice, ctx->stream));
  if ((r = mlp_pack(ctx))) return r;
  HIP_TRY(host_sync(ctx));
  ctx->mlp_ready = true;
  ctx->mlp_step = 0;
  return TRIXI_HIP_OK;
}

int trixi_hip_mlp_get_params(trixi_hip_ctx *ctx, float *params) {
  if (!ctx || !params) return TRIXI_HIP_EINVAL;
  if (!ctx->mlp_ready) return set_err(ctx, TRIXI_HIP_ESTATE, "no parameters set");
  HIP_TRY(hipMemcpyAsync(params, ctx->d_mlp, mlp::NP * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(host_sync(ctx));
  return TRIXI_HIP_OK;
}

int trixi_hip_set_indicator(trixi_hip_ctx *ctx, int mode) {
  if (!ctx || mode < TRIXI_HIP_IND_HG || mode > TRIXI_HIP_IND_MLP_WITH_FALLBACK) return TRIXI_HIP_EINVAL;
  if (mode != TRIXI_HIP_IND_HG && ctx->hg.variable != TRIXI_HIP_VAR_DENSITY_PRESSURE)
    return set_err(ctx, TRIXI_HIP_EINVAL, "the learned indicator reads density_pressure");
  if (mode != TRIXI_HIP_IND_HG && !ctx->mlp_ready)
    return set_err(ctx, TRIXI_HIP_ESTATE, "the learned indicator has no parameters (trixi_hip_mlp_set_params)");
  if (mode != ctx->ind_mode) {  // the captured steps launch the indicator kernels
    HIP_TRY(host_sync(ctx));
    drop_graphs(ctx);
  }
  if (mode != ctx->ind_mode) ctx->alpha_fresh = false;
  ctx->ind_mode = mode;
  return TRIXI_HIP_OK;
}

int trixi_hip_mlp_features(trixi_hip_ctx *ctx, float *features_host, double *alpha_hg_host) {
  if (!ctx) return TRIXI_HIP_EINVAL;
  int r = mlp_check(ctx);
  if (r) return r;
  if ((r = sync_u32(ctx))) return r;
  r = ctx->fp32 ? launch_hg<float>(ctx, cur_u32(ctx), ctx->d_alpha_hg, ctx->d_feat)
                : launch_hg<double>(ctx, cur_u64(ctx), ctx->d_alpha_hg, ctx->d_feat);  // no network: HG factors
  if (r) return r;
  if (features_host)
    HIP_TRY(hipMemcpyAsync(features_host, ctx->d_feat, (size_t)ctx->E * kMlpFeatureStride * 4,
                           hipMemcpyDeviceToHost, ctx->stream));
  if (alpha_hg_host)
    HIP_TRY(hipMemcpyAsync(alpha_hg_host, ctx->d_alpha_hg, (size_t)ctx->E * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(host_sync(ctx));
  return TRIXI_HIP_OK;
}

int trixi_hip_mlp_forward(trixi_hip_ctx *ctx, const float *x_dev, int n, float *y_dev) {
  if (!ctx || !x_dev || !y_dev || n <= 0) return TRIXI_HIP_EINVAL;
  if (!ctx->mlp_ready) return set_err(ctx, TRIXI_HIP_ESTATE, "no parameters set");
  if (((uintptr_t)x_dev & 15) != 0) return set_err(ctx, TRIXI_HIP_EINVAL, "features must be 16-byte aligned");
  const float *lo = ctx->d_mlp + 3 * mlp::NP, *hi = lo + mlp::NF;
  hipLaunchKernelGGL(mlp::mlp_infer_kernel, dim3(mlp_blocks(n)), dim3(mlp::NTHR), 0, ctx->stream, ctx->d_mlp, x_dev,
                     lo, hi, (const double *)nullptr, 0.0, 0.0, 0, n, (double *)nullptr, y_dev);
  HIP_TRY(hipGetLastError());
  return TRIXI_HIP_OK;
}

static int mlp_gradient(trixi_hip_ctx *ctx, const float *x_dev, const float *target_dev, int n) {
  const int nb = mlp_blocks(n);
  hipLaunchKernelGGL(mlp::mlp_grad_kernel, dim3(nb), dim3(mlp::NTHR), 0, ctx->stream, ctx->d_mlp, x_dev, target_dev,
                     n, ctx->d_mlp_partial);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(mlp::mlp_reduce_kernel, dim3((mlp::NPL + mlp::RED_COLS - 1) / mlp::RED_COLS),
                     dim3(mlp::RED_COLS * mlp::RED_SPLIT), 0, ctx->stream, ctx->d_mlp_partial, nb * mlp::WAVES,
                     mlp_grad_buffer(ctx));
  HIP_TRY(hipGetLastError());
  return TRIXI_HIP_OK;
}

int trixi_hip_mlp_loss_grad(trixi_hip_ctx *ctx, const float *x_dev, const float *target_dev, int n, float *loss,
                            float *grad_host) {
  if (!ctx || !x_dev || !target_dev || n <= 0) return TRIXI_HIP_EINVAL;
  if (!ctx->mlp_ready) return set_err(ctx, TRIXI_HIP_ESTATE, "no parameters set");
  if (((uintptr_t)x_dev & 15) != 0) return set_err(ctx, TRIXI_HIP_EINVAL, "features must be 16-byte aligned");
  int r = mlp_gradient(ctx, x_dev, target_dev, n);
  if (r) return r;
  if (grad_host)
    HIP_TRY(hipMemcpyAsync(grad_host, mlp_grad_buffer(ctx), mlp::NP * 4, hipMemcpyDeviceToHost, ctx->stream));
  if (loss) HIP_TRY(hipMemcpyAsync(loss, mlp_grad_buffer(ctx) + mlp::NP, 4, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(host_sync(ctx));
  return TRIXI_HIP_OK;
}

int trixi_hip_mlp_train_step(trixi_hip_ctx *ctx, const float *x_dev, const float *target_dev, int n,
                             float learning_rate, float *loss) {
  if (!ctx || !x_dev || !target_dev || n <= 0 || !(learning_rate > 0.f)) return TRIXI_HIP_EINVAL;
  if (!ctx->mlp_ready) return set_err(ctx, TRIXI_HIP_ESTATE, "no parameters set");
  if (((uintptr_t)x_dev & 15) != 0) return set_err(ctx, TRIXI_HIP_EINVAL, "features must be 16-byte aligned");
  int r = mlp_gradient(ctx, x_dev, target_dev, n);
  if (r) return r;
  ctx->mlp_step += 1;
  hipLaunchKernelGGL(mlp::mlp_adam_kernel, dim3((mlp::NP + 255) / 256), dim3(256), 0, ctx->stream, ctx->d_mlp,
                     ctx->d_mlp + mlp::NP, ctx->d_mlp + 2 * mlp::NP, mlp_grad_buffer(ctx), learning_rate, 0.9f,
                     0.999f, 1e-8f, ctx->mlp_step);
  HIP_TRY(hipGetLastError());
  if ((r = mlp_pack(ctx))) return r;
  if (loss) {
    HIP_TRY(hipMemcpyAsync(loss, mlp_grad_buffer(ctx) + mlp::NP, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(host_sync(ctx));
  }
  return TRIXI_HIP_OK;
}

int trixi_hip_indicator_hg(trixi_hip_ctx *ctx, double alpha_max, double alpha_min, int alpha_smooth, int variable,
                           double *alpha_host) {
  if (!ctx || !alpha_host || variable < TRIXI_HIP_VAR_DENSITY_PRESSURE || variable > TRIXI_HIP_VAR_PRESSURE ||
      !(alpha_max >= 0) || !(alpha_min >= 0))
    return TRIXI_HIP_EINVAL;
  HGParams hp = ctx->hg;
  hp.alpha_max = alpha_max;
  hp.alpha_min = alpha_min;
  hp.variable = variable;
  int r = launch_hg<double>(ctx, cur_u64(ctx), ctx->d_alpha_raw, nullptr, &hp);
  if (r) return r;
  const double *out = ctx->d_alpha_raw;
  if (alpha_smooth) {
    if (ctx->M > 0) {  // neighbours across mortars too
      hipLaunchKernelGGL(hg_smooth_adj_kernel, dim3((ctx->E + 255) / 256), dim3(256), 0, ctx->stream,
                         ctx->d_alpha_raw, ctx->d_adj_ptr, ctx->d_adj, ctx->d_alpha, ctx->E);
    } else {
      dispatch_dn(ctx->D, ctx->N, [&](auto Dc, auto) {
        constexpr int D = decltype(Dc)::value;
        hipLaunchKernelGGL((hg_smooth_kernel<D>), dim3((ctx->E + 255) / 256), dim3(256), 0, ctx->stream,
                           ctx->d_alpha_raw, ctx->d_nbr, ctx->d_alpha, ctx->E, 1);
      });
    }
    HIP_TRY(hipGetLastError());
    out = ctx->d_alpha;
  }
  HIP_TRY(hipMemcpyAsync(alpha_host, out, (size_t)ctx->E * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(host_sync(ctx));
  return TRIXI_HIP_OK;
}

int trixi_hip_indicator_max(trixi_hip_ctx *ctx, int variable, double *alpha_host) {
  if (!ctx || !alpha_host || variable < 0 || variable >= ctx->V) return TRIXI_HIP_EINVAL;
  bool ok = dispatch_dn(ctx->D, ctx->N, [&](auto Dc, auto Nc) {
    constexpr int D = decltype(Dc)::value, N = decltype(Nc)::value;
    hipLaunchKernelGGL((indicator_max_kernel<D, N>), dim3((ctx->E + 255) / 256), dim3(256), 0, ctx->stream,
                       cur_u64(ctx), ctx->E, variable, ctx->d_alpha_raw);
  });
  if (!ok) return set_err(ctx, TRIXI_HIP_EINVAL, "unsupported ndims/polydeg");
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(alpha_host, ctx->d_alpha_raw, (size_t)ctx->E * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(host_sync(ctx));
  return TRIXI_HIP_OK;
}

int trixi_hip_transfer(trixi_hip_ctx *to, trixi_hip_ctx *from, const int32_t *kind, const int32_t *src) {
  if (!to || !from || !kind || !src || to == from) return TRIXI_HIP_EINVAL;
  if (to->D != from->D || to->N != from->N || to->device != from->device)
    return set_err(to, TRIXI_HIP_EINVAL, "transfer needs contexts of the same ndims, polydeg and device");
  const int nch = 1 << to->D;
  for (int e = 0; e < to->E; ++e) {  // every element reads inside `from`
    const int k = kind[e], s = src[e];
    const bool ok = (k == 0 || (k >= 1 && k <= nch)) ? (s >= 0 && s < from->E)
                    : (k == -1 ? (s >= 0 && s + nch <= from->E) : false);
    if (!ok) return set_err(to, TRIXI_HIP_EINVAL, "transfer map entry out of range");
  }
  const int n = to->n, E = to->E;
  to->alpha_fresh = false;
  std::vector<double> mats;
  for (const std::vector<double> *m : {&to->mfu, &to->mfl, &to->mru, &to->mrl}) {
    if ((int)m->size() != n * n) return set_err(to, TRIXI_HIP_EINVAL, "mortar matrices missing");
    mats.insert(mats.end(), m->begin(), m->end());
  }
  // staging: kind [E] | src [E] | matrices [4 n n] (8-byte aligned after the ints)
  const size_t ints = ((size_t)2 * E + 1) & ~(size_t)1;
  char *d_tmp = nullptr;
  if (dev_alloc(to, &d_tmp, ints * 4 + mats.size() * 8) != hipSuccess)
    return set_err(to, TRIXI_HIP_ENOMEM, "transfer staging allocation failed");
  int32_t *d_kind = (int32_t *)d_tmp, *d_src = d_kind + E;
  double *d_mats = (double *)(d_tmp + ints * 4);
  int r = TRIXI_HIP_OK;
  if (host_sync(from) != hipSuccess ||
      hipMemcpyAsync(d_kind, kind, (size_t)E * 4, hipMemcpyHostToDevice, to->stream) != hipSuccess ||
      hipMemcpyAsync(d_src, src, (size_t)E * 4, hipMemcpyHostToDevice, to->stream) != hipSuccess ||
      hipMemcpyAsync(d_mats, mats.data(), mats.size() * 8, hipMemcpyHostToDevice, to->stream) != hipSuccess) {
    r = set_err(to, TRIXI_HIP_EHIP, "transfer staging failed");
  } else {
    const long total = (long)E * to->NN;
    bool ok = dispatch_dn(to->D, to->N, [&](auto Dc, auto Nc) {
      constexpr int D = decltype(Dc)::value, N = decltype(Nc)::value;
      hipLaunchKernelGGL((amr_transfer_kernel<D, N>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                         to->stream, cur_u64(from), cur_u64(to), d_kind, d_src, E, d_mats);
    });
    if (!ok || hipGetLastError() != hipSuccess) r = set_err(to, TRIXI_HIP_EHIP, "transfer launch failed");
    if (!r && to->fp32) r = convert<float, double>(to, cur_u32(to), cur_u64(to), to->nvals);
    if (!r) to->u32_stale = false;
  }
  if (host_sync(to) != hipSuccess && !r) r = set_err(to, TRIXI_HIP_EHIP, "transfer failed");
  dev_release(to, d_tmp);
  return r;
}

int trixi_hip_set_exact_solution(trixi_hip_ctx *ctx, int ic_kind, const double *params, int nparams,
                                 const double *modes) {
  if (!ctx || ic_kind < 0 || ic_kind > TRIXI_HIP_IC_BLOB) return TRIXI_HIP_EINVAL;
  // captured steps hold the boundary-state kernel with the old function
  HIP_TRY(host_sync(ctx));
  drop_graphs(ctx);
  ctx->ghost_ready = false;
  ctx->exact.kind = ic_kind;
  ctx->exact.gamma = ctx->cfg.gamma;
  for (int i = 0; i < 4; ++i) ctx->exact.params[i] = (params && i < nparams) ? params[i] : 0.0;
  if (ic_kind == TRIXI_HIP_IC_FOURIER_MODES) {
    if (!modes) return set_err(ctx, TRIXI_HIP_EINVAL, "Fourier-mode exact solution needs modes");
    std::memcpy(ctx->exact.modes, modes, sizeof(ctx->exact.modes));
  }
  return TRIXI_HIP_OK;
}

int trixi_hip_compute_coefficients(trixi_hip_ctx *ctx, double t) {
  if (!ctx) return TRIXI_HIP_EINVAL;
  const long total = (long)ctx->E * ctx->NN;
  bool ok = dispatch_dn(ctx->D, ctx->N, [&](auto Dc, auto Nc) {
    constexpr int D = decltype(Dc)::value, N = decltype(Nc)::value;
    NodesConst<D, N> nc{};
    for (int i = 0; i <= N; ++i) nc.x[i] = ctx->nodes[i];
    hipLaunchKernelGGL((initial_state_kernel<D, N>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       ctx->stream, ctx->d_centers, ctx->d_invj, ctx->E, nc, ctx->exact, t, cur_u64(ctx));
  });
  if (!ok) return TRIXI_HIP_EINVAL;
  HIP_TRY(hipGetLastError());
  if (ctx->fp32) {
    int r = convert<float, double>(ctx, cur_u32(ctx), cur_u64(ctx), ctx->nvals);
    if (r) return r;
    ctx->u32_stale = false;
  }
  ctx->alpha_fresh = false;
  return TRIXI_HIP_OK;
}

int trixi_hip_error_norms(trixi_hip_ctx *ctx, double t, double *l2, double *linf) {
  if (!ctx || !l2 || !linf) return TRIXI_HIP_EINVAL;
  const int V = ctx->V;
  int r = enqueue_error_sums(ctx, t, nullptr, ctx->d_scalars + kErrSums);
  if (r) return r;
  double res[16];
  HIP_TRY(hipMemcpyAsync(res, ctx->d_scalars + kErrSums, 2 * V * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(host_sync(ctx));
  if ((r = total_volumes(ctx))) return r;
  for (int v = 0; v < V; ++v) {
    l2[v] = std::sqrt(res[v] / ctx->tv_err);
    linf[v] = res[V + v];
  }
  return TRIXI_HIP_OK;
}

int trixi_hip_integrals(trixi_hip_ctx *ctx, double *cons_integrals, double *entropy_integral) {
  if (!ctx || !cons_integrals || !entropy_integral) return TRIXI_HIP_EINVAL;
  const int V = ctx->V;
  int r = enqueue_integral_sums(ctx, ctx->d_scalars + kIntSums);
  if (r) return r;
  double res[8];
  HIP_TRY(hipMemcpyAsync(res, ctx->d_scalars + kIntSums, (V + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(host_sync(ctx));
  if ((r = total_volumes(ctx))) return r;
  for (int v = 0; v < V; ++v) cons_integrals[v] = res[v] / ctx->tv_int;
  *entropy_integral = res[V] / ctx->tv_int;
  return TRIXI_HIP_OK;
}

int trixi_hip_set_analysis(trixi_hip_ctx *ctx, int interval, int capacity) {
  if (!ctx || interval < 0 || (interval > 0 && capacity <= 0)) return TRIXI_HIP_EINVAL;
  if (ctx->d_records) {
    HIP_TRY(host_sync(ctx));
    dev_release(ctx, ctx->d_records);
    ctx->d_records = nullptr;
  }
  ctx->rec_interval = interval;
  ctx->rec_cap = interval > 0 ? capacity : 0;
  if (interval > 0) HIP_TRY(dev_alloc(ctx, &ctx->d_records, (size_t)capacity * rec_width(ctx) * 8));
  return TRIXI_HIP_OK;
}

int trixi_hip_get_analysis(trixi_hip_ctx *ctx, double *records, int max_records, int *count) {
  if (!ctx || !count || max_records < 0 || (max_records > 0 && !records)) return TRIXI_HIP_EINVAL;
  *count = 0;
  if (!ctx->d_records) return TRIXI_HIP_OK;
  double total = 0;
  HIP_TRY(hipMemcpyAsync(&total, ctx->d_scalars + kCtl + 7, 8, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(host_sync(ctx));
  const int W = rec_width(ctx), V = ctx->V;
  const long long made = (long long)total;
  const int kept = (int)std::min<long long>(made, ctx->rec_cap);
  std::vector<double> raw((size_t)ctx->rec_cap * W);
  HIP_TRY(sync_copy(ctx, raw.data(), ctx->d_records, raw.size() * 8, hipMemcpyDeviceToHost));
  int r = total_volumes(ctx);
  if (r) return r;
  const int nout = std::min(kept, max_records);
  for (int i = 0; i < nout; ++i) {  // the newest nout records, oldest first
    const long long k = made - nout + i;
    const double *src = raw.data() + (size_t)(k % ctx->rec_cap) * W;
    double *dst = records + (size_t)i * W;
    dst[0] = src[0];
    dst[1] = src[1];
    for (int v = 0; v < V; ++v) {
      dst[2 + v] = std::sqrt(src[2 + v] / ctx->tv_err);
      dst[2 + V + v] = src[2 + V + v];
      dst[2 + 2 * V + v] = src[2 + 2 * V + v] / ctx->tv_int;
    }
    dst[2 + 3 * V] = src[2 + 3 * V] / ctx->tv_int;
  }
  *count = nout;
  return TRIXI_HIP_OK;
}

int trixi_hip_get_counters(const trixi_hip_ctx *ctx, int64_t *device_allocations, int64_t *host_synchronizations) {
  if (!ctx) return TRIXI_HIP_EINVAL;
  if (device_allocations) *device_allocations = ctx->n_alloc;
  if (host_synchronizations) *host_synchronizations = ctx->n_sync;
  return TRIXI_HIP_OK;
}

int trixi_hip_autotune(trixi_hip_ctx *ctx, int repetitions, trixi_hip_tile_timing *timings, int max_timings,
                       int *ntimings) {
  if (!ctx || repetitions <= 0 || max_timings < 0 || (max_timings > 0 && !timings)) return TRIXI_HIP_EINVAL;
  if (ntimings) *ntimings = 0;
  const int reps = repetitions + (repetitions & 1);  // even: the fused stage returns to its buffer parity
  struct Cand {
    int stage, epb, nt;
    ElemLaunch fn;
    double best_us;
  };
  std::vector<Cand> cand;
  const bool fused = ctx->elem_stage != nullptr;  // the stage variant is in use (not switched off)
  for (const auto &kv : elem_registry()) {
    const ElemKey &k = kv.first;
    if (k.fp32 == (ctx->fp32 ? 1 : 0) && k.D == ctx->D && k.N == ctx->N && k.vol == ctx->vol &&
        k.bnd == (ctx->B + ctx->M > 0 ? 1 : 0) && (k.stage == 0 || fused))
      cand.push_back(Cand{k.stage, k.epb, k.nt, kv.second, 1e300});
  }
  // captured steps hold the kernels of the old tiles
  HIP_TRY(host_sync(ctx));
  drop_graphs(ctx);
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  HIP_TRY(hipEventCreate(&ev0));
  if (hipEventCreate(&ev1) != hipSuccess) {
    (void)hipEventDestroy(ev0);
    return set_err(ctx, TRIXI_HIP_EHIP, "hipEventCreate failed");
  }
  const ElemLaunch keep_elem = ctx->elem, keep_stage = ctx->elem_stage;
  const int keep[4] = {ctx->epb, ctx->nt, ctx->epb_st, ctx->nt_st};
  // The RHS kernels write d_du from the resident state. The stage kernels run
  // as first stages with a = b = 0, which leaves the state unchanged (u_tmp =
  // du, u = u + 0 u_tmp) and, over an even number of launches, in its buffer.
  auto launch = [&](const Cand &c) -> int {
    if (c.stage == 0) {
      int r0 = sync_u32(ctx);
      return r0 ? r0 : launch_rhs_native(ctx, ctx->d_du, native_u(ctx), 0.0, nullptr);
    }
    return launch_stage(ctx, 0.0, 0.0, 0.0, nullptr, 1);
  };
  int r = TRIXI_HIP_OK;
  for (int round = 0; round < 3 && !r; ++round) {
    for (auto &c : cand) {
      if (c.stage == 0) {
        ctx->elem = c.fn;
        ctx->epb = c.epb;
        ctx->nt = c.nt;
      } else {
        ctx->elem_stage = c.fn;
        ctx->epb_st = c.epb;
        ctx->nt_st = c.nt;
      }
      for (int i = 0; i < 2 && !r; ++i) r = launch(c);
      if (!r && hipEventRecord(ev0, ctx->stream) != hipSuccess) r = TRIXI_HIP_EHIP;
      for (int i = 0; i < reps && !r; ++i) r = launch(c);
      if (!r && hipEventRecord(ev1, ctx->stream) != hipSuccess) r = TRIXI_HIP_EHIP;
      ++ctx->n_sync;
      float ms = 0.f;
      if (!r && (hipEventSynchronize(ev1) != hipSuccess || hipEventElapsedTime(&ms, ev0, ev1) != hipSuccess))
        r = TRIXI_HIP_EHIP;
      if (!r) c.best_us = std::min(c.best_us, 1e3 * ms / reps);
      if (r) break;
    }
  }
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  ctx->elem = keep_elem;
  ctx->elem_stage = keep_stage;
  ctx->epb = keep[0];
  ctx->nt = keep[1];
  ctx->epb_st = keep[2];
  ctx->nt_st = keep[3];
  if (r) return r == TRIXI_HIP_EHIP ? set_err(ctx, r, "autotune: a timed launch failed") : r;
  // the fastest tile of each variant becomes the context's and this process's default
  for (int stage = 0; stage < 2; ++stage) {
    const Cand *best = nullptr;
    for (const auto &c : cand)
      if (c.stage == stage && (!best || c.best_us < best->best_us)) best = &c;
    if (!best) continue;
    if (stage == 0) {
      ctx->elem = best->fn;
      ctx->epb = best->epb;
      ctx->nt = best->nt;
    } else {
      ctx->elem_stage = best->fn;
      ctx->epb_st = best->epb;
      ctx->nt_st = best->nt;
    }
    std::lock_guard<std::mutex> lock(g_tuned_mutex);
    g_tuned[tune_key(ctx, stage)] = {best->epb, best->nt};
  }
  const int nout = std::min<int>((int)cand.size(), max_timings);
  for (int i = 0; i < nout; ++i)
    timings[i] = trixi_hip_tile_timing{cand[i].stage, cand[i].epb, cand[i].nt, cand[i].best_us};
  if (ntimings) *ntimings = (int)cand.size();
  return TRIXI_HIP_OK;
}

int trixi_hip_get_stage_tile(const trixi_hip_ctx *ctx, int *elements_per_block, int *threads_per_block) {
  if (!ctx) return TRIXI_HIP_EINVAL;
  if (elements_per_block) *elements_per_block = ctx->epb_st;
  if (threads_per_block) *threads_per_block = ctx->nt_st;
  return TRIXI_HIP_OK;
}

int trixi_hip_get_tile(const trixi_hip_ctx *ctx, int *elements_per_block, int *threads_per_block) {
  if (!ctx) return TRIXI_HIP_EINVAL;
  if (elements_per_block) *elements_per_block = ctx->epb;
  if (threads_per_block) *threads_per_block = ctx->nt;
  return TRIXI_HIP_OK;
}

}  // extern "C"
