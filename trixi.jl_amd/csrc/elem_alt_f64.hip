// FP64 element kernels with the alternative tiles of the headline workloads
// (see elem_alt_f32.hip).
#include "elem_registry.hpp"

TRIXI_REG_ALT_TILES(double)
