"""HIP rhs_hyperbolic! against the FP64 oracle on the same inputs.
Bar: FP64 within 1e-12 (max-norm, relative to max|du|); FP32 bulk within the
per-case tolerance written below (FP32 fluxes; measured 1e-7..7e-5)."""
import ctypes

import numpy as np
import pytest

from gpu_helpers import build, fp32_abs_bound, fp64_abs_bound, oracle_for, relerr

torch = pytest.importorskip("torch")

MMS3 = dict(ndims=3, polydeg=3, level=2, cmin=(0, 0, 0), cmax=(2, 2, 2), ic="convergence_test", source="convergence_test")
BLAST = dict(ndims=3, polydeg=3, level=2, cmin=(-2, -2, -2), cmax=(2, 2, 2), ic="weak_blast_wave")
CASES = {
    "3d_p3_fd_ranocha_llf_mms": (dict(MMS3, volume_integral="fd", surface_flux="llf"), 1e-4),
    "3d_p3_fd_ranocha_llfnaive_mms": (dict(MMS3, volume_integral="fd", surface_flux="llf_naive"), 1e-4),
    "3d_p3_weak_mms": (dict(MMS3, volume_integral="weak", surface_flux="llf_naive"), 1e-4),
    "3d_p3_fd_central_mms": (dict(MMS3, volume_integral="fd", volume_flux="central", surface_flux="llf_naive"), 1e-4),
    "3d_p3_ec_blast": (dict(BLAST, volume_integral="fd", surface_flux="ranocha"), 1e-5),
    "3d_p3_hg_blast": (dict(BLAST, volume_integral="hg", surface_flux="ranocha", fv_flux="ranocha"), 1e-5),
    "3d_p4_fd_fourier": (dict(ndims=3, polydeg=4, level=2, cmin=(-1,) * 3, cmax=(1,) * 3, ic="fourier_modes",
                              ic_params=(0,)), 1e-5),
    "3d_p2_fd_fourier": (dict(ndims=3, polydeg=2, level=2, cmin=(-1,) * 3, cmax=(1,) * 3, ic="fourier_modes",
                              ic_params=(1,)), 1e-5),
    "2d_p3_hg_khi": (dict(ndims=2, polydeg=3, level=4, cmin=(-1, -1), cmax=(1, 1), ic="kelvin_helmholtz",
                          ic_params=(1e-3, 0), volume_integral="hg", surface_flux="llf_naive", fv_flux="llf_naive",
                          alpha_max=0.002, alpha_min=1e-4), 2e-4),
    "2d_p3_fd_vortex_p1": (dict(ndims=2, polydeg=3, level=4, cmin=(-10, -10), cmax=(10, 10), ic="isentropic_vortex",
                                ic_params=(1.0,)), 1e-4),
    "2d_p3_weak_vortex": (dict(ndims=2, polydeg=3, level=4, cmin=(-10, -10), cmax=(10, 10), ic="isentropic_vortex",
                               ic_params=(25.0,), volume_integral="weak", surface_flux="llf_naive"), 2e-4),
    "2d_p4_fd_mms": (dict(ndims=2, polydeg=4, level=3, cmin=(0, 0), cmax=(2, 2), ic="convergence_test",
                          source="convergence_test"), 2e-4),
    "2d_p2_fd_mms": (dict(ndims=2, polydeg=2, level=3, cmin=(0, 0), cmax=(2, 2), ic="convergence_test",
                          source="convergence_test"), 2e-4),
    "2d_p3_pure_fv": (dict(ndims=2, polydeg=3, level=3, cmin=(0, 0), cmax=(2, 2), ic="convergence_test",
                           source="convergence_test", volume_integral="fv"), 1e-4),
    # non-periodic meshes: BoundaryConditionDirichlet(initial_condition) at time t
    "2d_p3_weak_mms_bounded": (dict(ndims=2, polydeg=3, level=3, cmin=(0, 0), cmax=(2, 2), ic="convergence_test",
                                    source="convergence_test", volume_integral="weak", surface_flux="llf_naive",
                                    periodicity=False), 2e-4),
    "3d_p3_fd_mms_bounded": (dict(MMS3, volume_integral="fd", surface_flux="llf", periodicity=False), 1e-4),
    "3d_p3_hg_blast_bounded_y": (dict(BLAST, volume_integral="hg", surface_flux="ranocha", fv_flux="ranocha",
                                      periodicity=(True, False, True)), 1e-5),
    "3d_p4_fd_mms_bounded": (dict(ndims=3, polydeg=4, level=1, cmin=(0,) * 3, cmax=(2,) * 3, ic="convergence_test",
                                  source="convergence_test", periodicity=False), 1e-4),
    "3d_p2_weak_mms_bounded": (dict(ndims=3, polydeg=2, level=2, cmin=(0,) * 3, cmax=(2,) * 3, ic="convergence_test",
                                    source="convergence_test", volume_integral="weak", periodicity=False), 1e-4),
    "2d_p3_hg_khi_bounded": (dict(ndims=2, polydeg=3, level=4, cmin=(-1, -1), cmax=(1, 1), ic="kelvin_helmholtz",
                                  ic_params=(0.0, 0), volume_integral="hg", surface_flux="llf_naive",
                                  fv_flux="llf_naive", alpha_max=0.002, alpha_min=1e-4, periodicity=False), 2e-4),
    # boundary_condition_slip_wall: the wall flux from the inner state
    "2d_p3_ec_blast_slip": (dict(ndims=2, polydeg=3, level=3, cmin=(-1, -1), cmax=(1, 1), ic="weak_blast_wave",
                                 volume_integral="fd", surface_flux="ranocha", periodicity=False,
                                 boundary_condition="slip_wall"), 2e-5),
    "3d_p3_fd_blast_slip": (dict(BLAST, cmin=(-1,) * 3, cmax=(1,) * 3, volume_integral="fd", surface_flux="llf",
                                 periodicity=False, boundary_condition="slip_wall"), 1e-5),
    "3d_p3_hg_blast_slip_x": (dict(BLAST, cmin=(-1,) * 3, cmax=(1,) * 3, volume_integral="hg", surface_flux="ranocha",
                                   fv_flux="ranocha", periodicity=(False, True, True),
                                   boundary_condition="slip_wall"), 1e-5),
    "2d_p4_fd_khi_channel_slip": (dict(ndims=2, polydeg=4, level=3, cmin=(-1, -1), cmax=(1, 1), ic="kelvin_helmholtz",
                                       ic_params=(0.0, 0), surface_flux="llf_naive", periodicity=(True, False),
                                       boundary_condition="slip_wall"), 2e-4),
    "3d_p2_weak_mms_slip_mortar": (dict(MMS3, polydeg=2, volume_integral="weak", periodicity=False,
                                        boundary_condition="slip_wall",
                                        refinement_patches=[((0.5, 0.5, 0.5), (1.5, 1.5, 1.5))]), 1e-4),
    # refined meshes (box refinement patches): L2 mortars at the level changes
    "3d_p3_weak_mms_mortar": (dict(MMS3, volume_integral="weak", surface_flux="llf_naive",
                                   refinement_patches=[((0.5, 0.5, 0.5), (1.5, 1.5, 1.5))]), 1e-4),
    "3d_p3_fd_mms_mortar": (dict(MMS3, volume_integral="fd", surface_flux="llf",
                                 refinement_patches=[((0.5, 0.5, 0.5), (1.5, 1.5, 1.5))]), 1e-4),
    "3d_p4_fd_mms_mortar": (dict(ndims=3, polydeg=4, level=1, cmin=(0,) * 3, cmax=(2,) * 3, ic="convergence_test",
                                 source="convergence_test", refinement_patches=[((0, 0, 0), (1, 1, 1))]), 1e-4),
    "3d_p2_fd_ec_mortar_bounded": (dict(BLAST, polydeg=2, volume_integral="fd", surface_flux="ranocha",
                                        periodicity=False, refinement_patches=[((-1, -1, -1), (1, 1, 1))]), 1e-5),
    "2d_p3_fd_mms_mortar_nested": (dict(ndims=2, polydeg=3, level=3, cmin=(0, 0), cmax=(2, 2), ic="convergence_test",
                                        source="convergence_test",
                                        refinement_patches=[((0.6, 0.6), (1.4, 1.4)), ((0.8, 0.8), (1.2, 1.2))]), 2e-4),
    "2d_p3_hg_khi_mortar": (dict(ndims=2, polydeg=3, level=3, cmin=(-1, -1), cmax=(1, 1), ic="kelvin_helmholtz",
                                 ic_params=(1e-3, 0), volume_integral="hg", surface_flux="llf_naive",
                                 fv_flux="llf_naive", alpha_max=0.002, alpha_min=1e-4,
                                 refinement_patches=[((-1, -0.5), (1, 0.5))]), 2e-4),
    # flux_hllc (pure finite volume): subcell, interface, boundary and mortar fluxes
    "3d_p3_pure_fv_hllc_mms": (dict(MMS3, volume_integral="fv", surface_flux="hllc", fv_flux="hllc"), 1e-4),
    "3d_p3_pure_fv_hllc_blast_mortar": (dict(BLAST, volume_integral="fv", surface_flux="hllc", fv_flux="hllc",
                                             refinement_patches=[((-1, -1, -1), (1, 1, 1))]), 1e-5),
    "2d_p3_pure_fv_hllc_mms_bounded": (dict(ndims=2, polydeg=3, level=3, cmin=(0, 0), cmax=(2, 2), ic="convergence_test",
                                            source="convergence_test", volume_integral="fv", surface_flux="hllc",
                                            fv_flux="hllc", periodicity=False), 2e-4),
    "2d_p3_pure_fv_hllc_blast_slip": (dict(ndims=2, polydeg=3, level=3, cmin=(-1, -1), cmax=(1, 1),
                                           ic="weak_blast_wave", volume_integral="fv", surface_flux="hllc",
                                           fv_flux="hllc", periodicity=False, boundary_condition="slip_wall"), 2e-5),
}


def perturbed(u, seed, eps=0.02):
    rng = np.random.default_rng(seed)
    return u * (1 + eps * rng.standard_normal(u.shape))


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_rhs_parity_fp64_and_fp32(name):
    import trixi_hip as th
    spec, tol32 = CASES[name]
    o = oracle_for(spec)
    for variant in ("ic", "perturbed"):
        semi64, ic = build(spec, "fp64")
        u = semi64.compute_coefficients(0.0)
        if variant == "perturbed" and spec.get("ic") != "constant":
            u = perturbed(u, 11)
        t = 0.37
        du_ref = o.rhs(u, t)
        # host pointers
        du = np.empty_like(u)
        th.rhs_hyperbolic_(du, u, semi64, t)
        assert relerr(du, du_ref) < 1e-12, (name, variant, "fp64 host")
        # device pointers (torch tensors on torch's current stream)
        ud = torch.from_numpy(u).cuda()
        dud = torch.empty_like(ud)
        th.rhs_hyperbolic_(dud, ud, semi64, t)
        torch.cuda.synchronize()
        assert relerr(dud.cpu().numpy(), du_ref) < 1e-12, (name, variant, "fp64 device")
        semi64.close()
        semi32, _ = build(spec, "fp32")
        du32 = np.empty_like(u)
        th.rhs_hyperbolic_(du32, u, semi32, t)
        assert relerr(du32, du_ref) < tol32, (name, variant, "fp32", relerr(du32, du_ref))
        semi32.close()


@pytest.mark.gpu
def test_native_fp32_entry_point_and_determinism():
    """trixi_hip_rhs_native on float32 device arrays (the benchmarked path) equals
    trixi_hip_rhs in FP32 mode bitwise, and repeated calls are bitwise identical."""
    import trixi_hip as th
    spec = CASES["3d_p3_fd_ranocha_llf_mms"][0]
    semi, _ = build(spec, "fp32")
    u = perturbed(semi.compute_coefficients(0.0), 3)
    du_a = np.empty_like(u)
    th.rhs_hyperbolic_(du_a, u, semi, 0.2)
    u32 = torch.from_numpy(u).cuda().float()
    outs = []
    for _ in range(3):
        d32 = torch.empty_like(u32)
        rc = semi._lib.trixi_hip_rhs_native(semi.ctx, ctypes.c_void_p(d32.data_ptr()), ctypes.c_void_p(u32.data_ptr()),
                                            0.2)
        assert rc == 0
        semi._lib.trixi_hip_synchronize(semi.ctx)
        outs.append(d32.cpu().numpy())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[1], outs[2])
    assert np.array_equal(outs[0].astype(np.float64), du_a)
    semi.close()


@pytest.mark.gpu
@pytest.mark.parametrize("D,per", [(2, True), (3, True), (2, False), (3, False)])
def test_free_stream_preservation(D, per):
    """A constant state gives du = 0 to round-off, also with Dirichlet boundaries
    that impose the same constant state."""
    import trixi_hip as th
    spec = dict(ndims=D, polydeg=3, level=3, cmin=(-1,) * D, cmax=(1,) * D, ic="constant", surface_flux="llf",
                periodicity=per)
    for prec in ("fp64", "fp32"):
        semi, _ = build(spec, prec)
        u = semi.compute_coefficients(0.0)
        du = np.empty_like(u)
        th.rhs_hyperbolic_(du, u, semi, 0.0)
        # round-off level: K eps J^-1 max|f| with the working precision's eps
        tol = fp64_abs_bound(semi, u) if prec == "fp64" else fp32_abs_bound(semi, u)
        assert np.max(np.abs(du)) < tol, (prec, np.max(np.abs(du)), tol)
        semi.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["3d_p3_hg_blast", "2d_p3_hg_khi", "2d_p3_hg_khi_bounded", "2d_p3_hg_khi_mortar"])
def test_hg_indicator_matches_oracle(name):
    spec = CASES[name][0]
    o = oracle_for(spec)
    semi, _ = build(spec, "fp64")
    u = perturbed(semi.compute_coefficients(0.0), 5, 0.01)
    lib, ctx = semi._lib, semi.ctx
    lib.trixi_hip_upload_u(ctx, u.ctypes.data_as(ctypes.c_void_p))
    a = np.empty(semi.nelements())
    assert lib.trixi_hip_indicator(ctx, 0, a.ctypes.data_as(ctypes.c_void_p)) == 0
    a_ref = o.indicator_hg(u)
    np.testing.assert_allclose(a, a_ref, rtol=1e-12, atol=1e-15)
    semi.close()


@pytest.mark.gpu
@pytest.mark.parametrize("D,vi,per", [(2, "fd", True), (3, "fd", True), (2, "hg", False), (3, "weak", False)])
def test_rhs_parity_on_randomly_adapted_meshes(D, vi, per):
    """The RHS on meshes adapted three times with random controller values
    (levels 1-4: mortars in every orientation, on both sides, next to
    boundaries): FP64 to 1e-12 of the oracle, FP32 within the flux-difference
    round-off bound."""
    import trixi_hip as th
    spec = dict(ndims=D, polydeg=3, level=2, cmin=(0,) * D, cmax=(2,) * D, volume_integral=vi,
                volume_flux="ranocha", surface_flux="llf", ic="convergence_test", periodicity=per)
    o = oracle_for(spec)
    rng = np.random.default_rng(40 + D)
    lams = []
    u0 = np.zeros(o.nvalues)
    for _ in range(3):
        lam = rng.integers(-1, 2, o.nelements).astype(np.int32)
        lam[o.levels() >= 4] = np.minimum(lam[o.levels() >= 4], 0)
        lams.append(lam)
        _, u0, _, _ = o.amr_adapt(u0, lam)
    assert o.nmortars > 0
    o.set_initial_condition("convergence_test")
    u = perturbed(o.compute_coefficients(0.0), 5)
    t = 0.21
    du_ref = o.rhs(u, t)
    for prec in ("fp64", "fp32"):
        print(f"adapted-mesh RHS D={D} {vi} periodic={per} {prec}", flush=True)
        semi, _ = build(spec, prec)
        for lam in lams:
            changed, kind, src = semi.mesh.adapt(lam)
        semi.remesh(kind, src)
        assert semi.nelements() == o.nelements and semi.mesh.nmortars == o.nmortars
        du = np.empty_like(u)
        th.rhs_hyperbolic_(du, u, semi, t)
        if prec == "fp64":
            assert relerr(du, du_ref) < 1e-12, relerr(du, du_ref)
        else:
            assert np.max(np.abs(du - du_ref)) < fp32_abs_bound(semi, u)
        semi.close()

