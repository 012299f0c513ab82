"""Adaptive mesh refinement with the device doing the per-node work:
IndicatorMax and the L2 solution transfer against the oracle, and the
reference's AMR known-answer run (elixir_euler_amr.jl) end to end."""
import ctypes

import numpy as np
import pytest

from conftest import golden
from gpu_helpers import build, oracle_for, relerr

torch = pytest.importorskip("torch")


def upload(semi, u):
    assert semi._lib.trixi_hip_upload_u(semi.ctx, u.ctypes.data_as(ctypes.c_void_p)) == 0


def download(semi):
    u = np.empty(semi.nvalues())
    assert semi._lib.trixi_hip_download_u(semi.ctx, u.ctypes.data_as(ctypes.c_void_p)) == 0
    return u


def _spec(D, patches=True, periodic=True):
    cmin, cmax = (0.0,) * D, (2.0,) * D
    box = [((0.5,) * D, (1.5,) * D)] if patches else []
    return dict(ndims=D, polydeg=3, level=2, cmin=cmin, cmax=cmax, volume_integral="weak", surface_flux="llf",
                ic="fourier_modes" if D == 3 else "kelvin_helmholtz", refinement_patches=box, periodicity=periodic)


def _state(semi, seed):
    u = semi.compute_coefficients(0.0)
    return u * (1 + 0.05 * np.random.default_rng(seed).standard_normal(u.shape))


@pytest.mark.gpu
@pytest.mark.parametrize("D,precision", [(2, "fp64"), (3, "fp64"), (3, "fp32")])
def test_indicator_max_matches_oracle(D, precision):
    """IndicatorMax(variable = first) and of the energy: maxima are exact."""
    import trixi_hip as th
    spec = _spec(D)
    print(f"indicator_max D={D} {precision}", flush=True)
    semi, _ = build(spec, precision)
    o = oracle_for(spec)
    u = _state(semi, 3)
    upload(semi, u)
    for var in (th.first, D + 1):
        alpha = th.IndicatorMax(semi, variable=var)(semi)
        np.testing.assert_array_equal(alpha, o.indicator_max(u, 0 if var is th.first else var))
    semi.close()


@pytest.mark.gpu
@pytest.mark.parametrize("D,periodic,precision", [(2, True, "fp64"), (2, False, "fp64"), (3, True, "fp64"),
                                                  (3, False, "fp32")])
def test_transfer_matches_oracle(D, periodic, precision):
    """Random controller values, three adaptations in a row: the device's
    projected state equals the oracle's (refine: forward interpolation; coarsen:
    L2 projection of the 2^D children) to FP64 round-off, for FP64 and FP32
    contexts (the FP64 register is what is projected)."""
    spec = _spec(D, periodic=periodic)
    print(f"transfer D={D} periodic={periodic} {precision}", flush=True)
    semi, _ = build(spec, precision)
    o = oracle_for(spec)
    u = _state(semi, 5)
    upload(semi, u)
    rng = np.random.default_rng(17 + D)
    for it in range(3):
        lam = rng.integers(-1, 2, semi.nelements()).astype(np.int32)
        lam[semi.mesh.levels >= 4] = np.minimum(lam[semi.mesh.levels >= 4], 0)
        changed, kind, src = semi.mesh.adapt(lam)
        assert changed
        semi.remesh(kind, src)
        _, u_ref, okind, osrc = o.amr_adapt(u, lam)
        np.testing.assert_array_equal(kind, okind)
        np.testing.assert_array_equal(src, osrc)
        u = download(semi)
        assert u.size == u_ref.size
        assert relerr(u, u_ref) < 1e-14
        u = u_ref
    semi.close()


@pytest.mark.gpu
def test_transferred_fp32_state_steps_like_an_uploaded_one():
    """After a transfer into an FP32 context the FP32 shadow matches the FP64
    register: one RK step from the transferred state equals one RK step from
    the same state uploaded into a fresh context on the same mesh (bitwise)."""
    import trixi_hip as th
    spec = _spec(3)
    semi, _ = build(spec, "fp32")
    upload(semi, _state(semi, 9))
    lam = np.zeros(semi.nelements(), np.int32)
    lam[::7] = 1
    _, kind, src = semi.mesh.adapt(lam)
    semi.remesh(kind, src)
    u = download(semi)
    fresh = th.SemidiscretizationHyperbolic(semi.mesh, semi.equations, semi.initial_condition, semi.solver,
                                            precision="fp32")
    upload(fresh, u)
    alg = th.CarpenterKennedy2N54()
    a = (ctypes.c_double * 5)(*alg.a)
    b = (ctypes.c_double * 5)(*alg.b)
    c = (ctypes.c_double * 5)(*alg.c)
    for s in (semi, fresh):
        assert s._lib.trixi_hip_rk_step(s.ctx, a, b, c, 5, 0.0, 1e-3, 1) == 0
        assert s._lib.trixi_hip_synchronize(s.ctx) == 0
    assert np.array_equal(download(semi), download(fresh))
    semi.close()
    fresh.close()


@pytest.mark.gpu
def test_transfer_rejects_out_of_range_maps():
    """A map entry that would read outside the old context is refused on the
    host (no launch), as are contexts of another polydeg."""
    import trixi_hip as th
    spec = _spec(2, patches=False)
    semi, _ = build(spec, "fp64")
    other, _ = build(dict(spec, level=3), "fp64")
    lib = semi._lib
    E_new = other.nelements()
    for kind_val, src_val in ((0, semi.nelements()), (-1, semi.nelements() - 3), (5, 0), (0, -1)):
        kind = np.zeros(E_new, np.int32)
        src = np.zeros(E_new, np.int32)
        kind[-1], src[-1] = kind_val, src_val
        rc = lib.trixi_hip_transfer(other.ctx, semi.ctx, kind.ctypes.data_as(ctypes.c_void_p),
                                    src.ctypes.data_as(ctypes.c_void_p))
        assert rc == th._lib.EINVAL
    p4, _ = build(dict(spec, polydeg=4), "fp64")
    kind = np.zeros(p4.nelements(), np.int32)
    src = np.arange(p4.nelements(), dtype=np.int32)
    assert lib.trixi_hip_transfer(p4.ctx, semi.ctx, kind.ctypes.data_as(ctypes.c_void_p),
                                  src.ctypes.data_as(ctypes.c_void_p)) == th._lib.EINVAL
    for s in (semi, other, p4):
        s.close()


def run_amr_kta(precision, device_steps=True):
    import trixi_hip as th
    c = golden("kta.json")["amr_3d"]
    a = c["amr"]
    spec = dict(c["oracle"], ic=c["ic"], ic_params=tuple(c["ic_params"]))
    semi, _ = build(spec, precision)
    ode = th.semidiscretize(semi, tuple(c["tspan"]))
    ind = th.IndicatorMax(semi, variable=th.first)
    ctl = th.ControllerThreeLevel(semi, ind, base_level=a["base_level"], med_level=a["med_level"],
                                  med_threshold=a["med_threshold"], max_level=a["max_level"],
                                  max_threshold=a["max_threshold"])
    amr = th.AMRCallback(semi, ctl, interval=a["interval"], adapt_initial_condition=a["adapt_initial_condition"],
                         adapt_initial_condition_only_refine=a["adapt_initial_condition_only_refine"])
    ac = th.AnalysisCallback(semi, interval=100)
    sol = th.solve(ode, th.CarpenterKennedy2N54(),
                   callback=th.CallbackSet(ac, amr, th.StepsizeCallback(cfl=c["cfl"])), device_steps=device_steps)
    semi.close()
    return c, sol, amr


@pytest.mark.gpu
def test_amr_kta_on_gpu_fp64():
    """elixir_euler_amr.jl on the device: the same sequence of adapted meshes
    as the oracle (which reproduces the reference to 2e-15) and the
    reference's final errors. The errors are dominated by the displacement of
    the pulse, so the element-count history is what pins the mesh decisions."""
    from oracle import ControllerThreeLevel
    print("amr kta fp64", flush=True)
    c, sol, amr = run_amr_kta("fp64")
    a = c["amr"]
    o = oracle_for(dict(c["oracle"]))
    o.set_initial_condition(c["ic"])
    ctl = ControllerThreeLevel(base_level=a["base_level"], med_level=a["med_level"], med_threshold=a["med_threshold"],
                               max_level=a["max_level"], max_threshold=a["max_threshold"])
    u_ref, t_ref, n_ref, hist_ref = o.solve_amr(c["tspan"][0], c["tspan"][1], c["cfl"], ctl, interval=a["interval"])
    assert amr.nelements_history == hist_ref
    assert sol.nsteps == n_ref and sol.t == pytest.approx(t_ref, rel=1e-15, abs=0)
    assert relerr(sol.u, u_ref) < 1e-12
    l2, linf = np.array(sol.analysis[-1]["l2"]), np.array(sol.analysis[-1]["linf"])
    np.testing.assert_allclose(l2, c["l2"], rtol=1e-11, atol=0)
    np.testing.assert_allclose(linf, c["linf"], rtol=1e-11, atol=0)


@pytest.mark.gpu
def test_amr_kta_on_gpu_fp32():
    print("amr kta fp32", flush=True)
    c, sol, amr = run_amr_kta("fp32")
    l2, linf = np.array(sol.analysis[-1]["l2"]), np.array(sol.analysis[-1]["linf"])
    np.testing.assert_allclose(l2, c["l2"], rtol=2e-4)
    np.testing.assert_allclose(linf, c["linf"], rtol=2e-3)


@pytest.mark.gpu
def test_amr_device_chunks_equal_host_loop():
    """With an AMRCallback the steps between adaptations run as device-controlled
    solves; the run equals the host-driven loop (device_steps=False) bitwise:
    same adapted meshes, steps, final time and state."""
    print("amr chunks vs host loop", flush=True)
    _, sol_dev, amr_dev = run_amr_kta("fp64")
    _, sol_host, amr_host = run_amr_kta("fp64", device_steps=False)
    assert amr_dev.nelements_history == amr_host.nelements_history
    assert sol_dev.nsteps == sol_host.nsteps and sol_dev.t == sol_host.t
    assert np.array_equal(sol_dev.u, sol_host.u)


@pytest.mark.gpu
def test_blast_wave_amr_kta_on_gpu_fp64():
    """elixir_euler_blast_wave_amr.jl on the device: HG shock capturing on the
    adapted meshes and the device HG indicator driving ControllerThreeLevel
    (206 steps, 42 adaptations): the same element-count history as the oracle
    (which reproduces the reference to 7e-15) and the reference's errors."""
    import trixi_hip as th
    from oracle import ControllerThreeLevel
    print("blast wave amr kta fp64", flush=True)
    c = golden("kta.json")["blast_wave_amr_2d"]
    a = c["amr"]
    spec = dict(c["oracle"], ic=c["ic"], ic_params=())
    semi, _ = build(spec, "fp64")
    vi = semi.solver.volume_integral
    ind = th.IndicatorHennemannGassner(semi, alpha_max=vi.indicator.alpha_max, alpha_min=vi.indicator.alpha_min,
                                       alpha_smooth=True, variable=th.density_pressure)
    ctl = th.ControllerThreeLevel(semi, ind, base_level=a["base_level"], max_level=a["max_level"],
                                  max_threshold=a["max_threshold"])
    amr = th.AMRCallback(semi, ctl, interval=a["interval"], adapt_initial_condition=True,
                         adapt_initial_condition_only_refine=True)
    sol = th.solve(th.semidiscretize(semi, tuple(c["tspan"])), th.CarpenterKennedy2N54(),
                   callback=th.CallbackSet(th.AnalysisCallback(semi, interval=100), amr,
                                           th.StepsizeCallback(cfl=c["cfl"])))
    o = oracle_for(dict(c["oracle"]))
    o.set_initial_condition(c["ic"])
    octl = ControllerThreeLevel(base_level=a["base_level"], max_level=a["max_level"], max_threshold=a["max_threshold"])
    u_ref, t_ref, n_ref, hist_ref = o.solve_amr(c["tspan"][0], c["tspan"][1], c["cfl"], octl, interval=a["interval"],
                                                indicator="hg")
    assert amr.nelements_history == hist_ref
    assert sol.nsteps == n_ref
    l2, linf = np.array(sol.analysis[-1]["l2"]), np.array(sol.analysis[-1]["linf"])
    np.testing.assert_allclose(l2, c["l2"], rtol=1e-10, atol=0)
    np.testing.assert_allclose(linf, c["linf"], rtol=1e-10, atol=0)
    semi.close()


@pytest.mark.gpu
def test_khi_amr_kta_on_gpu_fp64():
    """elixir_euler_kelvin_helmholtz_instability_amr.jl on the device: an AMR-only
    HG indicator with its own parameters (variable density, alpha_max 1, no
    smoothing; trixi_hip_indicator_hg), med_level = current level, an
    adaptation after every step: the oracle's element-count history and the
    reference's errors."""
    import trixi_hip as th
    from oracle import ControllerThreeLevel
    print("khi amr kta fp64", flush=True)
    c = golden("kta.json")["khi_amr_2d"]
    a = c["amr"]
    spec = dict(c["oracle"], ic=c["ic"], ic_params=())
    semi, _ = build(spec, "fp64")
    am, amin, smooth, var = a["indicator"]
    ind = th.IndicatorHennemannGassner(semi, alpha_max=am, alpha_min=amin, alpha_smooth=smooth, variable=th.density)
    ctl = th.ControllerThreeLevel(semi, ind, base_level=a["base_level"], med_level=a["med_level"],
                                  med_threshold=a["med_threshold"], max_level=a["max_level"],
                                  max_threshold=a["max_threshold"])
    amr = th.AMRCallback(semi, ctl, interval=a["interval"], adapt_initial_condition=True,
                         adapt_initial_condition_only_refine=True)
    sol = th.solve(th.semidiscretize(semi, tuple(c["tspan"])), th.CarpenterKennedy2N54(),
                   callback=th.CallbackSet(th.AnalysisCallback(semi, interval=100), amr,
                                           th.StepsizeCallback(cfl=c["cfl"])))
    o = oracle_for(dict(c["oracle"]))
    octl = ControllerThreeLevel(base_level=a["base_level"], med_level=a["med_level"], med_threshold=a["med_threshold"],
                                max_level=a["max_level"], max_threshold=a["max_threshold"])
    o.set_initial_condition(c["ic"])
    _, _, n_ref, hist_ref = o.solve_amr(c["tspan"][0], c["tspan"][1], c["cfl"], octl, interval=a["interval"],
                                        indicator=tuple(a["indicator"]))
    assert amr.nelements_history == hist_ref and sol.nsteps == n_ref
    l2, linf = np.array(sol.analysis[-1]["l2"]), np.array(sol.analysis[-1]["linf"])
    np.testing.assert_allclose(l2, c["l2"], rtol=1e-10, atol=0)
    np.testing.assert_allclose(linf, c["linf"], rtol=1e-10, atol=0)
    semi.close()


@pytest.mark.gpu
@pytest.mark.parametrize("variable,smooth", [("density", False), ("pressure", True), ("density_pressure", True)])
def test_indicator_hg_with_own_parameters_matches_oracle(variable, smooth):
    """trixi_hip_indicator_hg on a refined mesh (smoothing across mortars)."""
    import trixi_hip as th
    spec = _spec(2)
    semi, _ = build(spec, "fp64")
    o = oracle_for(spec)
    u = _state(semi, 21)
    upload(semi, u)
    ind = th.IndicatorHennemannGassner(semi, alpha_max=0.7, alpha_min=0.002, alpha_smooth=smooth, variable=variable)
    np.testing.assert_allclose(ind(semi), o.indicator_hg_params(u, 0.7, 0.002, smooth, variable), rtol=1e-12,
                               atol=1e-14)
    semi.close()


@pytest.mark.gpu
def test_blob_amr_kta_on_gpu_fp64():
    """elixir_euler_blob_amr.jl on the device: gamma 5/3, polydeg 4, shock
    capturing on pressure, ControllerThreeLevelCombined with the device HG
    indicators (AMR-only density indicator + the shock-capturing one), an
    adaptation after every step: the oracle's element-count history and the
    reference's errors."""
    import trixi_hip as th
    from oracle import ControllerThreeLevelCombined as OracleCombined
    print("blob amr kta fp64", flush=True)
    c = golden("kta.json")["blob_amr_2d"]
    a = c["amr"]
    spec = dict(c["oracle"], ic=c["ic"], ic_params=())
    semi, _ = build(spec, "fp64")
    am, amin, smooth, var = a["indicator"]
    primary = th.IndicatorHennemannGassner(semi, alpha_max=am, alpha_min=amin, alpha_smooth=smooth, variable=var)
    secondary = semi.solver.volume_integral.indicator
    ctl = th.ControllerThreeLevelCombined(semi, primary, secondary, base_level=a["base_level"],
                                          med_level=a["med_level"], med_threshold=a["med_threshold"],
                                          max_level=a["max_level"], max_threshold=a["max_threshold"],
                                          max_threshold_secondary=a["max_threshold_secondary"])
    amr = th.AMRCallback(semi, ctl, interval=a["interval"], adapt_initial_condition=True,
                         adapt_initial_condition_only_refine=True)
    sol = th.solve(th.semidiscretize(semi, tuple(c["tspan"])), th.CarpenterKennedy2N54(),
                   callback=th.CallbackSet(th.AnalysisCallback(semi, interval=100), amr,
                                           th.StepsizeCallback(cfl=c["cfl"])))
    o = oracle_for(dict(c["oracle"]))
    o.set_initial_condition(c["ic"])
    octl = OracleCombined(base_level=a["base_level"], med_level=a["med_level"], med_threshold=a["med_threshold"],
                          max_level=a["max_level"], max_threshold=a["max_threshold"],
                          max_threshold_secondary=a["max_threshold_secondary"])
    _, _, n_ref, hist_ref = o.solve_amr(c["tspan"][0], c["tspan"][1], c["cfl"], octl, interval=a["interval"],
                                        indicator=tuple(a["indicator"]))
    assert amr.nelements_history == hist_ref and sol.nsteps == n_ref
    l2, linf = np.array(sol.analysis[-1]["l2"]), np.array(sol.analysis[-1]["linf"])
    np.testing.assert_allclose(l2, c["l2"], rtol=1e-10, atol=0)
    np.testing.assert_allclose(linf, c["linf"], rtol=1e-10, atol=0)
    semi.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11)])
def test_amr_with_dirichlet_boundaries_matches_oracle(precision, tol):
    """AMR on a non-periodic mesh (Dirichlet boundaries from the initial
    condition), HG shock capturing and the HG indicator driving the controller:
    30 steps with an adaptation every 2 steps give the oracle's element-count
    history and final state (no reference KTA covers this combination)."""
    import trixi_hip as th
    from oracle import ControllerThreeLevel
    print(f"amr bounded {precision}", flush=True)
    spec = dict(ndims=2, polydeg=3, level=5, cmin=(-2, -2), cmax=(2, 2), volume_integral="hg", volume_flux="ranocha",
                surface_flux="llf_naive", fv_flux="llf_naive", alpha_max=0.5, alpha_min=0.001, ic="blast_wave",
                periodicity=False)
    semi, _ = build(spec, precision)
    vi = semi.solver.volume_integral
    ctl = th.ControllerThreeLevel(semi, vi.indicator, base_level=3, max_level=6, max_threshold=0.01)
    amr = th.AMRCallback(semi, ctl, interval=2, adapt_initial_condition=True, adapt_initial_condition_only_refine=True)
    sol = th.solve(th.semidiscretize(semi, (0.0, 1.0)), th.CarpenterKennedy2N54(),
                   callback=th.CallbackSet(amr, th.StepsizeCallback(cfl=0.9)), maxiters=30)
    o = oracle_for(spec)
    u_ref, t_ref, n_ref, hist_ref = o.solve_amr(0.0, 1.0, 0.9, ControllerThreeLevel(base_level=3, max_level=6,
                                                                                     max_threshold=0.01),
                                                interval=2, indicator="hg", maxiters=30)
    assert semi.mesh.nboundaries > 0 and semi.mesh.nmortars > 0
    assert amr.nelements_history == hist_ref and sol.nsteps == n_ref == 30
    assert sol.t == pytest.approx(t_ref, rel=1e-10 if precision == "fp64" else 1e-4)
    assert relerr(sol.u, u_ref) < tol
    semi.close()
