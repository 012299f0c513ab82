// FP32 element kernels with the alternative tiles of the headline workloads
// (two wavefronts per direction), RHS and fused-stage variants, periodic
// meshes: candidates of trixi_hip_autotune besides the default tiles.
#include "elem_registry.hpp"

TRIXI_REG_ALT_TILES(float)
