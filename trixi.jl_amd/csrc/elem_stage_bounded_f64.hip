// F64 element kernels that apply the 2N stage update (default tiles, bounded meshes).
#include "elem_registry.hpp"

TRIXI_REG_DEFAULTS(double, true, true)
