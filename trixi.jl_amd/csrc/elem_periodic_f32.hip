// FP32 element kernels for periodic meshes (no boundary states).
#include "elem_registry.hpp"

TRIXI_REG_DEFAULTS(float, false, false)
