"""ctypes binding of the C ABI declared in include/trixi_hip.h.

The shared library ``libtrixi_hip.so`` is built in-tree (``trixi_hip.build``)
and loaded from the package directory. There is no fallback: if the library
is missing, or no GPU is visible, every entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libtrixi_hip.so")

OK = 0
EINVAL = -1
EHIP = -2
ENOMEM = -3
ESTATE = -4

FP64 = 0
FP32 = 1

VI_WEAK = 0
VI_FLUX_DIFFERENCING = 1
VI_SHOCK_CAPTURING_HG = 2
VI_PURE_FV = 3

FLUX_LLF = 0
FLUX_LLF_NAIVE = 1
FLUX_RANOCHA = 2
FLUX_CENTRAL = 3
FLUX_HLLC = 4

BC_NONE = 0
BC_DIRICHLET = 1
BC_SLIP_WALL = 2

SRC_NONE = 0
SRC_CONVERGENCE_TEST = 1

IC_CONSTANT = 0
IC_CONVERGENCE_TEST = 1
IC_WEAK_BLAST_WAVE = 2
IC_ISENTROPIC_VORTEX = 3
IC_KELVIN_HELMHOLTZ = 4
IC_FOURIER_MODES = 5
IC_DENSITY_PULSE = 6
IC_BLAST_WAVE = 7
IC_BLOB = 8

VAR_DENSITY_PRESSURE = 0
VAR_DENSITY = 1
VAR_PRESSURE = 2

IND_HG = 0
IND_MLP = 1
IND_MLP_WITH_FALLBACK = 2

MLP_NPARAMS = 5505
MLP_NFEATURES = 19
MLP_FEATURE_STRIDE = 32

HOST = 0
DEVICE = 1


class Config(ctypes.Structure):
    _fields_ = [
        ("ndims", ctypes.c_int32),
        ("polydeg", ctypes.c_int32),
        ("nelements", ctypes.c_int32),
        ("ninterfaces", ctypes.c_int32),
        ("gamma", ctypes.c_double),
        ("precision", ctypes.c_int32),
        ("volume_integral", ctypes.c_int32),
        ("volume_flux", ctypes.c_int32),
        ("surface_flux", ctypes.c_int32),
        ("fv_flux", ctypes.c_int32),
        ("alpha_max", ctypes.c_double),
        ("alpha_min", ctypes.c_double),
        ("alpha_smooth", ctypes.c_int32),
        ("source_terms", ctypes.c_int32),
        ("elements_per_block", ctypes.c_int32),
        ("threads_per_block", ctypes.c_int32),
        ("nboundaries", ctypes.c_int32),
        ("boundary_condition", ctypes.c_int32),
        ("nmortars", ctypes.c_int32),
        ("hg_variable", ctypes.c_int32),
    ]


class Mesh(ctypes.Structure):
    _fields_ = [
        ("neighbor_ids", ctypes.POINTER(ctypes.c_int32)),
        ("orientations", ctypes.POINTER(ctypes.c_int32)),
        ("inverse_jacobian", ctypes.POINTER(ctypes.c_double)),
        ("element_centers", ctypes.POINTER(ctypes.c_double)),
        ("boundary_neighbor_ids", ctypes.POINTER(ctypes.c_int32)),
        ("boundary_orientations", ctypes.POINTER(ctypes.c_int32)),
        ("boundary_neighbor_sides", ctypes.POINTER(ctypes.c_int32)),
        ("mortar_neighbor_ids", ctypes.POINTER(ctypes.c_int32)),
        ("mortar_large_sides", ctypes.POINTER(ctypes.c_int32)),
        ("mortar_orientations", ctypes.POINTER(ctypes.c_int32)),
    ]


class Basis(ctypes.Structure):
    _fields_ = [
        ("nodes", ctypes.POINTER(ctypes.c_double)),
        ("weights", ctypes.POINTER(ctypes.c_double)),
        ("derivative_split", ctypes.POINTER(ctypes.c_double)),
        ("derivative_hat", ctypes.POINTER(ctypes.c_double)),
        ("inverse_vandermonde_legendre", ctypes.POINTER(ctypes.c_double)),
        ("analysis_nnodes", ctypes.c_int32),
        ("analysis_weights", ctypes.POINTER(ctypes.c_double)),
        ("analysis_vandermonde", ctypes.POINTER(ctypes.c_double)),
        ("mortar_forward_upper", ctypes.POINTER(ctypes.c_double)),
        ("mortar_forward_lower", ctypes.POINTER(ctypes.c_double)),
        ("mortar_reverse_upper", ctypes.POINTER(ctypes.c_double)),
        ("mortar_reverse_lower", ctypes.POINTER(ctypes.c_double)),
    ]


# (name, restype, argtypes) for every exported entry point of include/trixi_hip.h
_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int)
_F = ctypes.c_float
SIGNATURES = [
    ("trixi_hip_create", _I, [ctypes.POINTER(Config), ctypes.POINTER(Mesh), ctypes.POINTER(Basis), ctypes.POINTER(_P)]),
    ("trixi_hip_destroy", _I, [_P]),
    ("trixi_hip_last_error", ctypes.c_char_p, [_P]),
    ("trixi_hip_set_stream", _I, [_P, _P]),
    ("trixi_hip_synchronize", _I, [_P]),
    ("trixi_hip_upload_u", _I, [_P, _P]),
    ("trixi_hip_download_u", _I, [_P, _P]),
    ("trixi_hip_set_u_device", _I, [_P, _P]),
    ("trixi_hip_get_u_device", _I, [_P, _P]),
    ("trixi_hip_rhs", _I, [_P, _P, _P, _D, _I]),
    ("trixi_hip_rhs_native", _I, [_P, _P, _P, _D]),
    ("trixi_hip_rk_stage", _I, [_P, _D, _D, _D, _I]),
    ("trixi_hip_rk_step", _I, [_P, _PD, _PD, _PD, _I, _D, _D, _I]),
    ("trixi_hip_solve", _I, [_P, _PD, _PD, _PD, _I, _D, _D, _D, _I, _I, _PD, _PI]),
    ("trixi_hip_max_dt", _I, [_P, _D, _PD]),
    ("trixi_hip_max_dt_noalloc", _I, [_I, _I, _I, _D, _P, _P, _D, _P, _P]),
    ("trixi_hip_indicator", _I, [_P, _I, _P]),
    ("trixi_hip_indicator_hg", _I, [_P, _D, _D, _I, _I, _P]),
    ("trixi_hip_indicator_max", _I, [_P, _I, _P]),
    ("trixi_hip_transfer", _I, [_P, _P, _P, _P]),
    ("trixi_hip_set_exact_solution", _I, [_P, _I, _PD, _I, _PD]),
    ("trixi_hip_compute_coefficients", _I, [_P, _D]),
    ("trixi_hip_error_norms", _I, [_P, _D, _PD, _PD]),
    ("trixi_hip_integrals", _I, [_P, _PD, _PD]),
    ("trixi_hip_set_analysis", _I, [_P, _I, _I]),
    ("trixi_hip_get_analysis", _I, [_P, _PD, _I, _PI]),
    ("trixi_hip_get_tile", _I, [_P, _PI, _PI]),
    ("trixi_hip_get_stage_tile", _I, [_P, _PI, _PI]),
    ("trixi_hip_autotune", _I, [_P, _I, _P, _I, _PI]),
    ("trixi_hip_get_counters", _I, [_P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    ("trixi_hip_mlp_set_params", _I, [_P, _P, _P, _P]),
    ("trixi_hip_mlp_get_params", _I, [_P, _P]),
    ("trixi_hip_set_indicator", _I, [_P, _I]),
    ("trixi_hip_mlp_features", _I, [_P, _P, _P]),
    ("trixi_hip_mlp_forward", _I, [_P, _P, _I, _P]),
    ("trixi_hip_mlp_loss_grad", _I, [_P, _P, _P, _I, _P, _P]),
    ("trixi_hip_mlp_train_step", _I, [_P, _P, _P, _I, _F, _P]),
    ("trixi_hip_tree_create", _P, [_I, _P, _D, _I, _P, _P, _P]),
    ("trixi_hip_tree_destroy", None, [_P]),
    ("trixi_hip_tree_last_error", ctypes.c_char_p, [_P]),
    ("trixi_hip_tree_refine", _I, [_P, _P, _P]),
    ("trixi_hip_tree_coarsen", _I, [_P, _I, _P, _P, _P]),
    ("trixi_hip_tree_sizes", _I, [_P, _P]),
    ("trixi_hip_tree_leaves", _I, [_P, _P, _P, _P, _P, _P]),
    ("trixi_hip_tree_connectivity", _I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
]

class TileTiming(ctypes.Structure):
    """trixi_hip_tile_timing"""
    _fields_ = [("stage", ctypes.c_int32), ("elements_per_block", ctypes.c_int32),
                ("threads_per_block", ctypes.c_int32), ("microseconds", ctypes.c_double)]


_lib = None


def load(path=None):
    """Load libtrixi_hip.so and declare its signatures (raises if absent)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"libtrixi_hip.so not found at {p}: build it with __graft_entry__.build() "
            "(there is no CPU fallback for the HIP path)")
    # PyTorch ships its own libamdhip64.so.7 with the same soname as the system
    # one; whichever is loaded first serves the whole process, and torch's
    # device arrays only work on its own copy. Load torch first when present.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(p)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class TrixiHipError(RuntimeError):
    pass


def check(rc, ctx=None, what=""):
    if rc != OK:
        msg = ""
        if ctx is not None and _lib is not None:
            m = _lib.trixi_hip_last_error(ctx)
            msg = m.decode() if m else ""
        raise TrixiHipError(f"{what} failed with code {rc}: {msg}")
