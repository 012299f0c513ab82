// FP64 element kernels for meshes with boundaries (default tiles only).
#include "elem_registry.hpp"

TRIXI_REG_DEFAULTS(double, true, false)
