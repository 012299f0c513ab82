"""The element-kernel tile autotuner: the tiles it chooses from give bitwise-
equal RHS and fused-stage results (FP64 always; FP32 on smooth states, where
every line takes the log-mean series: that choice is made per wavefront, so on
states with large jumps FP32 tiles may differ at round-off level), autotuning
leaves the resident state as it was, and the tuned tile becomes the default of
later contexts."""
import ctypes

import numpy as np
import pytest

from gpu_helpers import build

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

MMS3 = dict(ndims=3, polydeg=3, level=3, cmin=(0, 0, 0), cmax=(2, 2, 2), ic="convergence_test",
            source="convergence_test", surface_flux="llf")
FOURIER3_P4 = dict(ndims=3, polydeg=4, level=2, cmin=(-1, -1, -1), cmax=(1, 1, 1), ic="fourier_modes",
                   ic_params=(0,), surface_flux="llf")
KHI2 = dict(ndims=2, polydeg=3, level=5, cmin=(-1, -1), cmax=(1, 1), ic="kelvin_helmholtz", ic_params=(1e-3, 0),
            volume_integral="hg", volume_flux="ranocha", surface_flux="llf_naive", fv_flux="llf_naive",
            alpha_max=0.002, alpha_min=1e-4)
FD2 = dict(ndims=2, polydeg=3, level=5, cmin=(-10, -10), cmax=(10, 10), ic="isentropic_vortex", ic_params=(25.0,),
           surface_flux="llf")


def _state(semi):
    u = np.empty(semi.nvalues())
    assert semi._lib.trixi_hip_download_u(semi.ctx, u.ctypes.data_as(ctypes.c_void_p)) == 0
    return u


def _rhs_and_steps(semi, u0, nsteps=3):
    """du of u0 (native precision, widened) and the state after nsteps graph-replayed CK2N54 steps."""
    import trixi_hip as th
    lib, ctx = semi._lib, semi.ctx
    assert lib.trixi_hip_upload_u(ctx, u0.ctypes.data_as(ctypes.c_void_p)) == 0
    dtype = torch.float32 if semi.precision == "fp32" else torch.float64
    u = torch.from_numpy(u0).to(device="cuda", dtype=dtype)
    du = torch.empty_like(u)
    assert lib.trixi_hip_rhs_native(ctx, ctypes.c_void_p(du.data_ptr()), ctypes.c_void_p(u.data_ptr()), 0.1) == 0
    torch.cuda.synchronize()
    rk = th.CarpenterKennedy2N54()
    a, b, c = (np.ascontiguousarray(x, dtype=np.float64) for x in (rk.a, rk.b, rk.c))
    pd = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    for k in range(nsteps):
        assert lib.trixi_hip_rk_step(ctx, pd(a), pd(b), pd(c), 5, 0.1 + 1e-3 * k, 1e-3, 1) == 0
    return du.double().cpu().numpy(), _state(semi)


@pytest.mark.parametrize("spec,precision,one,alt", [
    (MMS3, "fp64", (4, 192), (8, 384)), (MMS3, "fp32", (4, 192), (8, 384)),
    (FOURIER3_P4, "fp64", (2, 192), (5, 384)), (KHI2, "fp64", (16, 128), (32, 256)),
    (FD2, "fp32", (16, 128), (32, 256))],
    ids=["mms3_p3_fp64", "mms3_p3_fp32", "fourier3_p4_fp64", "khi2_hg_fp64", "vortex2_fp32"])
def test_tiles_agree_bitwise_and_autotune_keeps_the_state(spec, precision, one, alt):
    """one: the one-wave-per-direction tile; alt: two waves per direction."""
    from trixi_hip.autotune import autotune
    base, _ = build(dict(spec, epb=one[0], nt=one[1]), precision)
    assert base.tile_used() == one and base.stage_tile_used() == one
    u0 = base.compute_coefficients(0.0)
    du_ref, u_ref = _rhs_and_steps(base, u0)
    base.close()
    other, _ = build(dict(spec, epb=alt[0], nt=alt[1]), precision)
    assert other.tile_used() == alt and other.stage_tile_used() == alt
    du_alt, u_alt = _rhs_and_steps(other, u0)
    np.testing.assert_array_equal(du_alt, du_ref)
    np.testing.assert_array_equal(u_alt, u_ref)
    other.close()
    # autotuning on the resident state: the state is unchanged, results equal
    tuned, _ = build(spec, precision)
    assert tuned._lib.trixi_hip_upload_u(tuned.ctx, u0.ctypes.data_as(ctypes.c_void_p)) == 0
    r = autotune(tuned, repetitions=4)
    assert {(t["stage"], t["elements_per_block"], t["threads_per_block"]) for t in r["timings"]} >= \
        {(0,) + alt, (1,) + alt}
    assert all(t["us"] > 0 for t in r["timings"])
    np.testing.assert_array_equal(_state(tuned), u0)
    assert tuned.tile_used() == r["rhs"] and tuned.stage_tile_used() == r["stage"]
    du_t, u_t = _rhs_and_steps(tuned, u0)
    np.testing.assert_array_equal(du_t, du_ref)
    np.testing.assert_array_equal(u_t, u_ref)
    tuned.close()
    # a context created later in the process takes the tuned tiles as its default
    later, _ = build(spec, precision)
    assert later.tile_used() == r["rhs"] and later.stage_tile_used() == r["stage"]
    later.close()
