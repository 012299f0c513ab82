"""FP32 mode: fused stages on meshes without mortars and slip walls read the FP64
register and leave the FP32 shadow of the state behind; readers of the shadow
(the indicator entry point) refresh it first. After graph-replayed and eager
steps, the indicator of the resident state must equal the indicator of the same
FP64 state uploaded afresh (whose shadow is converted on upload)."""
import ctypes

import numpy as np
import pytest

from gpu_helpers import build

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("periodic", [True, False])
def test_shadow_refreshed_after_fused_steps(periodic):
    import trixi_hip as th
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    spec = dict(ndims=2, polydeg=3, level=4, cmin=(-1, -1), cmax=(1, 1), ic="weak_blast_wave",
                volume_integral="hg", surface_flux="llf", fv_flux="llf", alpha_max=0.5, alpha_min=0.001)
    if not periodic:
        spec["periodicity"] = False
    semi, ic = build(spec, "fp32")
    lib, ctx = semi._lib, semi.ctx
    u0 = semi.compute_coefficients(0.0)
    vp = lambda x: x.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    pd = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    assert lib.trixi_hip_upload_u(ctx, vp(u0)) == 0
    rk = th.CarpenterKennedy2N54()
    a, b, c = (np.ascontiguousarray(x, dtype=np.float64) for x in (rk.a, rk.b, rk.c))
    E = u0.size // (4 * 16)  # 2D Euler (4 variables), polydeg 3 (16 nodes per element)
    for use_graph in (1, 0, 1):
        for k in range(3):
            assert lib.trixi_hip_rk_step(ctx, pd(a), pd(b), pd(c), len(a), 1e-3 * k, 1e-3, use_graph) == 0
        u = np.empty_like(u0)
        assert lib.trixi_hip_download_u(ctx, vp(u)) == 0
        assert np.all(np.isfinite(u))
        got = np.zeros(E)
        assert lib.trixi_hip_indicator(ctx, 0, vp(got)) == 0
        assert lib.trixi_hip_upload_u(ctx, vp(u)) == 0  # same FP64 state, shadow converted on upload
        ref = np.zeros_like(got)
        assert lib.trixi_hip_indicator(ctx, 0, vp(ref)) == 0
        assert np.array_equal(got, ref), (use_graph, float(np.max(np.abs(got - ref))))
        assert np.max(ref) > 0  # the blast wave is flagged
