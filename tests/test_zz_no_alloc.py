"""trixi_hip_max_dt_noalloc: the wave-speed reduction with its partial maxima in
the library's fixed device array instead of a context's buffers. Called
through the C ABI with device pointers; results must equal trixi_hip_max_dt
of a context holding the same state bit for bit (same kernels, and a maximum
does not depend on the reduction order)."""
import ctypes

import numpy as np
import pytest

import trixi_hip as th
from gpu_helpers import build

torch = pytest.importorskip("torch")

MMS3 = dict(ndims=3, polydeg=3, level=3, cmin=(0, 0, 0), cmax=(2, 2, 2), ic="convergence_test",
            source="convergence_test", surface_flux="llf")
VORTEX2 = dict(ndims=2, polydeg=4, level=3, cmin=(-10, -10), cmax=(10, 10), ic="isentropic_vortex",
               ic_params=(25.0,))
# 262144 elements in 16384 groups of 16: twice the 8192 partials of the scratch
# array, so each workgroup walks two element groups
VORTEX2_L9 = dict(VORTEX2, polydeg=3, level=9)


def _ctx_max_dt(semi, u, cfl):
    assert semi._lib.trixi_hip_upload_u(semi.ctx, u.ctypes.data_as(ctypes.c_void_p)) == 0
    dt = ctypes.c_double()
    assert semi._lib.trixi_hip_max_dt(semi.ctx, cfl, ctypes.byref(dt)) == 0
    return dt.value


def _noalloc_max_dt(lib, spec, semi, u, cfl, dt_dev):
    ud = torch.from_numpy(u).cuda()
    ij = torch.from_numpy(np.ascontiguousarray(semi.inverse_jacobian, dtype=np.float64)).cuda()
    stream = torch.cuda.current_stream()
    print(f"noalloc max_dt: ndims {spec['ndims']} polydeg {spec['polydeg']} E {semi.mesh.nelements}", flush=True)
    rc = lib.trixi_hip_max_dt_noalloc(spec["ndims"], spec["polydeg"], semi.mesh.nelements, 1.4,
                                      ctypes.c_void_p(ud.data_ptr()), ctypes.c_void_p(ij.data_ptr()), cfl,
                                      ctypes.c_void_p(dt_dev.data_ptr()), ctypes.c_void_p(stream.cuda_stream))
    assert rc == 0, lib.trixi_hip_last_error(None)
    stream.synchronize()
    return float(dt_dev.cpu()[0])


@pytest.mark.gpu
def test_noalloc_max_dt_equals_context_max_dt_on_successive_calls():
    """Three calls in a row, each on a smaller mesh than the one before (and
    another dimension or polydeg), each against trixi_hip_max_dt of a context."""
    lib = th._lib.load()
    dt_dev = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    rng = np.random.default_rng(7)
    for spec, cfl in ((VORTEX2_L9, 0.5), (MMS3, 0.7), (VORTEX2, 0.9)):
        semi, _ = build(spec, "fp64")
        u = semi.compute_coefficients(0.0)
        u = u * (1 + 0.01 * rng.standard_normal(u.shape))
        ref = _ctx_max_dt(semi, u, cfl)
        got = _noalloc_max_dt(lib, spec, semi, u, cfl, dt_dev)
        assert got == ref, (spec["ndims"], got, ref)
        semi.close()


@pytest.mark.gpu
def test_noalloc_max_dt_rejects_bad_arguments():
    lib = th._lib.load()
    x = torch.zeros(8, dtype=torch.float64, device="cuda")
    p = ctypes.c_void_p(x.data_ptr())
    assert lib.trixi_hip_max_dt_noalloc(3, 3, 0, 1.4, p, p, 1.0, p, None) == th._lib.EINVAL
    assert b"nelements" in lib.trixi_hip_last_error(None)
    assert lib.trixi_hip_max_dt_noalloc(3, 7, 1, 1.4, p, p, 1.0, p, None) == th._lib.EINVAL
    assert b"polydeg" in lib.trixi_hip_last_error(None)
    assert lib.trixi_hip_max_dt_noalloc(3, 3, 1, 1.4, None, p, 1.0, p, None) == th._lib.EINVAL
