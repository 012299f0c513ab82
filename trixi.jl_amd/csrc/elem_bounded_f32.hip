// FP32 element kernels for meshes with boundaries (default tiles only).
#include "elem_registry.hpp"

TRIXI_REG_DEFAULTS(float, true, false)
