// FP32 element kernels for periodic meshes (no boundary states).
#include "elem_registry.hpp"

TRIXI_REG_DEFAULTS(float, false, false)

// alternative tiles for the headline volume integrals, selectable through
// elements_per_block / threads_per_block (on MI355X they measured slower:
// 3D p3 8x384 47.3 us vs 41.9 us, 3D p4 5x384 155 us vs 92 us)
TRIXI_REG(float, 3, 3, trixi_hip::VOL_FD_RANOCHA, 8, 2, false, false)
TRIXI_REG(float, 3, 4, trixi_hip::VOL_FD_RANOCHA, 5, 2, false, false)
TRIXI_REG(float, 2, 3, trixi_hip::VOL_HG_RANOCHA, 32, 2, false, false)
TRIXI_REG(float, 2, 3, trixi_hip::VOL_FD_RANOCHA, 32, 2, false, false)
