"""FP64 CPU oracle for the DGSEM compressible-Euler RHS (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
package, and only as the checker; the HIP product path never imports it.
"""
from .oracle import Oracle, ControllerThreeLevel, ControllerThreeLevelCombined, build, LIB_PATH  # noqa: F401
