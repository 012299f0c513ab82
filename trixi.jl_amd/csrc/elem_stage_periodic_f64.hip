// F64 element kernels that apply the 2N stage update (default tiles, periodic meshes).
#include "elem_registry.hpp"

TRIXI_REG_DEFAULTS(double, false, true)
