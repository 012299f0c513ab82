// elem_registry.hpp -- registry of dg_element_kernel instantiations. The
// instantiations live in their own translation units (elem_*.hip, built in
// parallel); each registers a type-erased launcher under its key at load time,
// and the dispatcher in trixi_hip.hip looks the key up for a context.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <tuple>

#include "dg_element.hpp"

namespace trixi_hip {

// (precision, ndims, polydeg, volume integral, elements per block, threads per
// block, boundaries, stage) -- the workgroup has D * WPD wavefronts, one group
// of WPD waves per direction; bnd selects the variant that reads boundary
// states, stage the variant that applies the 2N update instead of storing du.
struct ElemKey {
  int fp32, D, N, vol, epb, nt, bnd, stage;
  bool operator<(const ElemKey &o) const {
    return std::tie(fp32, D, N, vol, epb, nt, bnd, stage) <
           std::tie(o.fp32, o.D, o.N, o.vol, o.epb, o.nt, o.bnd, o.stage);
  }
};

// Type-erased launcher: args and constants are passed as raw blobs that the
// instantiation reinterprets as ElemArgs<T> / DGConst<T, N>.
using ElemLaunch = hipError_t (*)(const void *args, const void *kconst, int nblocks, hipStream_t s);

std::map<ElemKey, ElemLaunch> &elem_registry();  // defined in trixi_hip.hip
// what a fused-stage variant's epilogue can add: the wave-speed reduction
// (stage_speed_fits) and the next stage's 2D p3 indicator (stage_ind2d_fits)
struct ElemCaps {
  bool speed = false, ind2d = false;
};
std::map<ElemKey, ElemCaps> &elem_caps();

template <typename T, int D, int N, int VOL, int EPB, int WPD, bool BND, bool STAGE>
hipError_t launch_elem(const void *args, const void *kconst, int nblocks, hipStream_t s) {
  const ElemArgs<T> &a = *static_cast<const ElemArgs<T> *>(args);
  const DGConst<T, N> &k = *static_cast<const DGConst<T, N> *>(kconst);
  hipLaunchKernelGGL((dg_element_kernel<T, D, N, VOL, EPB, WPD, BND, STAGE>), dim3(nblocks), dim3(D * WPD * 64), 0,
                     s, a, k);
  return hipGetLastError();
}

template <typename T, int D, int N, int VOL, int EPB, int WPD, bool BND, bool STAGE> struct Reg {
  Reg() {
    const ElemKey key{sizeof(T) == 4, D, N, VOL, EPB, D * WPD * 64, BND ? 1 : 0, STAGE ? 1 : 0};
    elem_registry()[key] = &launch_elem<T, D, N, VOL, EPB, WPD, BND, STAGE>;
    elem_caps()[key] = ElemCaps{STAGE && stage_speed_fits<T, D, N, VOL, EPB>(),
                                STAGE && stage_ind2d_fits<T, D, N, VOL, EPB, D * WPD * 64>()};
  }
};

}  // namespace trixi_hip

// Tiles: EPB = elements per workgroup, WPD = wavefronts per direction (lines
// per direction EPB * n^(D-1) <= 64 * WPD).
#define TRIXI_CAT2(a, b) a##b
#define TRIXI_CAT(a, b) TRIXI_CAT2(a, b)
#define TRIXI_REG(T, D, N, VOL, EPB, WPD, BND, STAGE) \
  static trixi_hip::Reg<T, D, N, VOL, EPB, WPD, BND, STAGE> TRIXI_CAT(trixi_reg_, __COUNTER__);

#define TRIXI_REG_ALLVOL(T, D, N, EPB, WPD, BND, STAGE)          \
  TRIXI_REG(T, D, N, trixi_hip::VOL_WEAK, EPB, WPD, BND, STAGE)       \
  TRIXI_REG(T, D, N, trixi_hip::VOL_FD_RANOCHA, EPB, WPD, BND, STAGE) \
  TRIXI_REG(T, D, N, trixi_hip::VOL_FD_CENTRAL, EPB, WPD, BND, STAGE) \
  TRIXI_REG(T, D, N, trixi_hip::VOL_HG_RANOCHA, EPB, WPD, BND, STAGE) \
  TRIXI_REG(T, D, N, trixi_hip::VOL_HG_CENTRAL, EPB, WPD, BND, STAGE) \
  TRIXI_REG(T, D, N, trixi_hip::VOL_PURE_FV, EPB, WPD, BND, STAGE)

// the default tiles (one wave per direction), also listed in kDefaultTiles
#define TRIXI_REG_DEFAULTS(T, BND, STAGE)          \
  TRIXI_REG_ALLVOL(T, 2, 2, 21, 1, BND, STAGE)     \
  TRIXI_REG_ALLVOL(T, 2, 3, 16, 1, BND, STAGE)     \
  TRIXI_REG_ALLVOL(T, 2, 4, 12, 1, BND, STAGE)     \
  TRIXI_REG_ALLVOL(T, 3, 2, 7, 1, BND, STAGE)      \
  TRIXI_REG_ALLVOL(T, 3, 3, 4, 1, BND, STAGE)      \
  TRIXI_REG_ALLVOL(T, 3, 4, 2, 1, BND, STAGE)

// alternative tiles (two wavefronts per direction) of the headline volume
// integrals, RHS and fused-stage variants, periodic meshes
#define TRIXI_REG_ALT_TILES(T)                                        \
  TRIXI_REG(T, 3, 3, trixi_hip::VOL_FD_RANOCHA, 8, 2, false, false)  \
  TRIXI_REG(T, 3, 3, trixi_hip::VOL_FD_RANOCHA, 8, 2, false, true)   \
  TRIXI_REG(T, 3, 4, trixi_hip::VOL_FD_RANOCHA, 5, 2, false, false)  \
  TRIXI_REG(T, 3, 4, trixi_hip::VOL_FD_RANOCHA, 5, 2, false, true)   \
  TRIXI_REG(T, 2, 3, trixi_hip::VOL_FD_RANOCHA, 32, 2, false, false) \
  TRIXI_REG(T, 2, 3, trixi_hip::VOL_FD_RANOCHA, 32, 2, false, true)  \
  TRIXI_REG(T, 2, 3, trixi_hip::VOL_HG_RANOCHA, 32, 2, false, false) \
  TRIXI_REG(T, 2, 3, trixi_hip::VOL_HG_RANOCHA, 32, 2, false, true)
