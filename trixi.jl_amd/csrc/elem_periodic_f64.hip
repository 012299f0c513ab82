// FP64 element kernels for periodic meshes (no boundary states).
#include "elem_registry.hpp"

TRIXI_REG_DEFAULTS(double, false, false)

// alternative tiles (see elem_periodic_f32.hip)
TRIXI_REG(double, 3, 3, trixi_hip::VOL_FD_RANOCHA, 8, 2, false, false)
TRIXI_REG(double, 3, 4, trixi_hip::VOL_FD_RANOCHA, 5, 2, false, false)
TRIXI_REG(double, 2, 3, trixi_hip::VOL_HG_RANOCHA, 32, 2, false, false)
