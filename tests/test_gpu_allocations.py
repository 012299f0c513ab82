"""The counterpart of the reference's allocation test of the RHS,
@test_allocations(Trixi.rhs_hyperbolic!, semi, sol, 1000)
(test/test_tree_3d_euler.jl:23): after set-up, the RHS, the 2N stage and the
graph-replayed step make no device allocation and no host synchronisation
(the library's own counters, trixi_hip_get_counters), and the device's free
memory does not move; a device-controlled solve synchronises once per poll of
its control block."""
import ctypes

import numpy as np
import pytest

from gpu_helpers import build

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

SPECS = [
    dict(ndims=3, polydeg=3, level=3, cmin=(0, 0, 0), cmax=(2, 2, 2), ic="convergence_test",
         source="convergence_test", surface_flux="llf"),
    dict(ndims=2, polydeg=3, level=4, cmin=(-1, -1), cmax=(1, 1), ic="kelvin_helmholtz", ic_params=(1e-3, 0),
         volume_integral="hg", volume_flux="ranocha", surface_flux="llf_naive", fv_flux="llf_naive",
         alpha_max=0.002, alpha_min=1e-4),
]


def counters(semi):
    a, s = ctypes.c_int64(), ctypes.c_int64()
    assert semi._lib.trixi_hip_get_counters(semi.ctx, ctypes.byref(a), ctypes.byref(s)) == 0
    return a.value, s.value


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("spec", SPECS, ids=["euler3d_mms", "khi2d_hg"])
def test_rhs_stage_step_allocate_and_synchronise_nothing(spec, precision):
    import trixi_hip as th
    semi, _ = build(spec, precision)
    lib, ctx = semi._lib, semi.ctx
    u0 = semi.compute_coefficients(0.0)
    assert lib.trixi_hip_upload_u(ctx, u0.ctypes.data_as(ctypes.c_void_p)) == 0
    dtype = torch.float32 if precision == "fp32" else torch.float64
    u = torch.from_numpy(u0).to(device="cuda", dtype=dtype)
    du = torch.empty_like(u)
    rk = th.CarpenterKennedy2N54()
    a, b, c = (np.ascontiguousarray(x, dtype=np.float64) for x in (rk.a, rk.b, rk.c))
    pd = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    stream = torch.cuda.Stream()
    assert lib.trixi_hip_set_stream(ctx, ctypes.c_void_p(stream.cuda_stream)) == 0

    def run(k):
        for i in range(k):
            assert lib.trixi_hip_rhs_native(ctx, ctypes.c_void_p(du.data_ptr()), ctypes.c_void_p(u.data_ptr()),
                                            0.1) == 0
            assert lib.trixi_hip_rk_stage(ctx, a[1], 1e-4 * b[1], 0.1, int(i == 0)) == 0
            assert lib.trixi_hip_rk_step(ctx, pd(a), pd(b), pd(c), len(a), 0.1, 1e-4, 1) == 0

    run(1)  # set-up: the first step captures its graphs
    stream.synchronize()
    before = counters(semi)
    free0 = torch.cuda.mem_get_info()[0]
    run(1000)
    after = counters(semi)
    stream.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert after == before, (before, after)
    assert free1 == free0, (free0, free1)
    lib.trixi_hip_set_stream(ctx, None)
    semi.close()


def test_device_solve_synchronises_once_per_poll():
    import trixi_hip as th
    semi, _ = build(SPECS[0], "fp32")
    lib, ctx = semi._lib, semi.ctx
    u0 = semi.compute_coefficients(0.0)
    assert lib.trixi_hip_upload_u(ctx, u0.ctypes.data_as(ctypes.c_void_p)) == 0
    rk = th.CarpenterKennedy2N54()
    a, b, c = (np.ascontiguousarray(x, dtype=np.float64) for x in (rk.a, rk.b, rk.c))
    pd = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    t_out, steps = ctypes.c_double(), ctypes.c_int()
    # the first solve captures the step graphs
    assert lib.trixi_hip_solve(ctx, pd(a), pd(b), pd(c), len(a), 0.0, 0.01, 0.5, 100000, 50, ctypes.byref(t_out),
                               ctypes.byref(steps)) == 0
    a0, s0 = counters(semi)
    poll = 10
    assert lib.trixi_hip_solve(ctx, pd(a), pd(b), pd(c), len(a), 0.01, 0.3, 0.5, 100000, poll, ctypes.byref(t_out),
                               ctypes.byref(steps)) == 0
    a1, s1 = counters(semi)
    assert steps.value > 4 * poll and t_out.value == pytest.approx(0.3, abs=1e-14)
    assert a1 == a0
    # one synchronisation for the initial max_dt, then one per batch of at most
    # `poll` steps (plus the batches shortened near t_end)
    assert s1 - s0 <= 1 + steps.value // poll + 3, (s1 - s0, steps.value)
    semi.close()
