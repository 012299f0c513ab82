"""Device-resident time integration and reductions: max_dt, error norms,
integrals, the 2N RK stage and the hipGraph-captured step, and Trixi.jl's
known-answer trajectories run entirely on the GPU."""
import ctypes

import numpy as np
import pytest

from conftest import golden
from gpu_helpers import build, oracle_for, relerr

torch = pytest.importorskip("torch")

MMS3 = dict(ndims=3, polydeg=3, level=2, cmin=(0, 0, 0), cmax=(2, 2, 2), ic="convergence_test",
            source="convergence_test", surface_flux="llf")


def upload(semi, u):
    assert semi._lib.trixi_hip_upload_u(semi.ctx, u.ctypes.data_as(ctypes.c_void_p)) == 0


def download(semi):
    u = np.empty(semi.nvalues())
    assert semi._lib.trixi_hip_download_u(semi.ctx, u.ctypes.data_as(ctypes.c_void_p)) == 0
    return u


@pytest.mark.gpu
def test_reductions_match_oracle():
    import trixi_hip as th
    for spec in (MMS3, dict(ndims=2, polydeg=3, level=4, cmin=(-10, -10), cmax=(10, 10), ic="isentropic_vortex",
                            ic_params=(25.0,))):
        o = oracle_for(spec)
        o.set_initial_condition(spec["ic"], spec.get("ic_params", ()))
        semi, ic = build(spec, "fp64")
        u = semi.compute_coefficients(0.0)
        u = u * (1 + 0.01 * np.random.default_rng(1).standard_normal(u.shape))
        upload(semi, u)
        dt = th.StepsizeCallback(cfl=0.7).calculate_dt(semi)
        assert dt == pytest.approx(o.max_dt(u, 0.7), rel=1e-14)
        ac = th.AnalysisCallback(semi)
        rec = ac.analyze(0.25, 0)
        l2, li = o.error_norms(u, 0.25)
        np.testing.assert_allclose(rec["l2"], l2, rtol=1e-11)
        np.testing.assert_allclose(rec["linf"], li, rtol=1e-12)
        cons, ent = ac.integrals()
        cons_ref, ent_ref = o.integrals(u)
        np.testing.assert_allclose(cons, cons_ref, rtol=1e-12)
        assert ent == pytest.approx(ent_ref, rel=1e-12)
        semi.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 5e-6)])
def test_rk_step_graph_eager_and_oracle(precision, tol):
    """One CK2N54 step: graph replay == eager launches (bitwise), and both equal
    the oracle's step (FP64 to 1e-12; FP32 bulk within tol of the state)."""
    import trixi_hip as th
    o = oracle_for(MMS3)
    semi, _ = build(MMS3, precision)
    lib, ctx = semi._lib, semi.ctx
    alg = th.CarpenterKennedy2N54()
    na = 5
    a = (ctypes.c_double * na)(*alg.a)
    b = (ctypes.c_double * na)(*alg.b)
    c = (ctypes.c_double * na)(*alg.c)
    u0 = semi.compute_coefficients(0.0)
    t, dt = 0.1, 2e-3
    results = {}
    for graph in (0, 1):
        upload(semi, u0)
        for rep in range(3):  # replay the graph several times with host work between
            assert lib.trixi_hip_rk_step(ctx, a, b, c, na, t + rep * dt, dt, graph) == 0
            assert lib.trixi_hip_synchronize(ctx) == 0
            results[(graph, rep)] = download(semi)
    for rep in range(3):
        assert np.array_equal(results[(0, rep)], results[(1, rep)])
    u_ref = o.rk_step(u0.copy(), t, dt)
    assert relerr(results[(1, 0)], u_ref) < tol
    semi.close()


def run_kta(case, precision):
    import trixi_hip as th
    c = golden("kta.json")[case]
    spec = dict(c["oracle"], ic=c["ic"], ic_params=tuple(c["ic_params"]))
    semi, _ = build(spec, precision)
    ode = th.semidiscretize(semi, tuple(c["tspan"]))
    ac = th.AnalysisCallback(semi, interval=0)
    sol = th.solve(ode, th.CarpenterKennedy2N54(), callback=th.CallbackSet(th.StepsizeCallback(cfl=c["cfl"]), ac),
                   maxiters=c.get("maxiters"))
    semi.close()
    return c, np.array(sol.analysis[-1]["l2"]), np.array(sol.analysis[-1]["linf"])


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["mms_3d_weak", "mms_3d_split_central", "ec_3d", "ec_3d_constant", "shockcapturing_3d",
                                  "khi_2d", "vortex_2d", "vortex_2d_constant_one_step", "mms_2d_nonperiodic",
                                  "mms_3d_mortar", "ec_2d_slip_wall", "pure_fv_hllc_3d", "pure_fv_hllc_2d",
                                  "blast_wave_pure_fv_2d", "density_pulse_3d"])
def test_kta_trajectory_on_gpu_fp64(case):
    """The whole run (RK stages, dt, analysis) on the device reproduces the
    reference's stored L2/Linf errors."""
    c, l2, linf = run_kta(case, "fp64")
    if c.get("roundoff"):
        assert np.all(l2 < 1e-13) and np.all(linf < 1e-12)
    else:
        np.testing.assert_allclose(l2, c["l2"], rtol=1e-9, atol=500 * np.finfo(float).eps)
        np.testing.assert_allclose(linf, c["linf"], rtol=1e-9, atol=500 * np.finfo(float).eps)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["mms_3d_weak", "ec_3d", "khi_2d", "density_pulse_3d"])
def test_kta_trajectory_on_gpu_fp32(case):
    c, l2, linf = run_kta(case, "fp32")
    np.testing.assert_allclose(l2, c["l2"], rtol=2e-4)
    np.testing.assert_allclose(linf, c["linf"], rtol=2e-3)


@pytest.mark.gpu
@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_slip_wall_reflection_conserves_and_matches_oracle(D, precision):
    """A weak blast wave reflected by boundary_condition_slip_wall on every side
    (to t = 0.8, past the first reflection): total mass and energy stay fixed
    (the wall flux has neither), and the final state equals the oracle's
    trajectory (FP64 to 1e-10 of max|u|; FP32 to 2e-5)."""
    import trixi_hip as th
    spec = dict(ndims=D, polydeg=3, level=3 if D == 2 else 2, cmin=(-1,) * D, cmax=(1,) * D, ic="weak_blast_wave",
                volume_integral="fd", surface_flux="ranocha", periodicity=False, boundary_condition="slip_wall")
    semi, _ = build(spec, precision)
    ac = th.AnalysisCallback(semi, interval=0)
    upload(semi, semi.compute_coefficients(0.0))
    first = ac.integrals()[0]
    sol = th.solve(th.semidiscretize(semi, (0.0, 0.8)), th.CarpenterKennedy2N54(),
                   callback=th.CallbackSet(th.StepsizeCallback(cfl=0.5), ac))
    last = sol.analysis[-1]["integrals"]
    drift = [abs(last[k] - first[k]) / abs(first[k]) for k in (0, D + 1)]
    print(f"slip wall {D}D {precision}: mass / energy drift {drift[0]:.2e} / {drift[1]:.2e}", flush=True)
    # FP64: round-off of the sums; FP32: the FP32 volume fluxes cancel to
    # FP32 round-off only (the face fluxes are the same bits on both sides).
    # Measured: FP64 <= 9e-16, FP32 <= 1.9e-10 (bounds about 3x that)
    bound = 3e-15 if precision == "fp64" else 6e-10
    assert max(drift) < bound, drift
    o = oracle_for(spec)
    u_ref, t_ref, n_ref = o.solve(o.compute_coefficients(0.0), 0.0, 0.8, 0.5)
    assert sol.nsteps == n_ref and sol.t == pytest.approx(t_ref, rel=1e-14)
    tol = 1e-10 if precision == "fp64" else 2e-5
    assert relerr(sol.u, u_ref) < tol, relerr(sol.u, u_ref)
    semi.close()


@pytest.mark.gpu
def test_nan_state_stops_the_run():
    """A non-physical state (negative pressure) gives NaN wave speeds -> NaN dt ->
    the integrator raises, as step! does on isnan(dt) (methods_2N.jl:138-140)."""
    import trixi_hip as th
    semi, _ = build(MMS3, "fp64")
    u = semi.compute_coefficients(0.0)
    u[4] = -10.0  # rho e of node 0 of element 0: pressure < 0
    ode = th.semidiscretize(semi, (0.0, 1.0))
    ode.u0 = u
    with pytest.raises(FloatingPointError):
        th.solve(ode, th.CarpenterKennedy2N54(), callback=th.CallbackSet(th.StepsizeCallback(cfl=0.5)))
    semi.close()


@pytest.mark.gpu
def test_torch_device_arrays_after_a_library_context():
    """A process that creates a context before it touches torch can still use
    torch device arrays with the library (both share one HIP runtime)."""
    import subprocess
    import sys
    from conftest import PKG
    code = (
        "import sys, numpy as np\n"
        f"sys.path.insert(0, {PKG!r})\n"
        "import trixi_hip as th\n"
        "m = th.TreeMesh((0, 0), (2, 2), initial_refinement_level=2, periodicity=True)\n"
        "s = th.SemidiscretizationHyperbolic(m, th.CompressibleEulerEquations2D(1.4),\n"
        "    th.initial_condition_convergence_test, th.DGSEM(polydeg=3, surface_flux=th.flux_lax_friedrichs))\n"
        "s.ctx\n"
        "import torch\n"
        "u = torch.from_numpy(s.compute_coefficients(0.0)).cuda()\n"
        "du = torch.empty_like(u)\n"
        "th.rhs_hyperbolic_(du, u, s, 0.0)\n"
        "torch.cuda.synchronize()\n"
        "assert bool(torch.isfinite(du).all())\n"
        "print('ok')\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stderr[-2000:]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_device_step_control_equals_host_loop(precision):
    """trixi_hip_solve (dt and limit_dt on the device, no host synchronisation
    per step) takes the same steps as the host loop that reads max_dt after
    every step (stepsize.jl:93-155, time_integration.jl:46-55): same number of
    steps, the same final time and a bitwise-equal state, for any poll interval;
    maxiters stops both after the same step. Analysis every 5 steps (run on the
    device during the device-controlled solve) gives the same records: the
    initial state, every 5th step and the final state (analysis.jl:121-122)."""
    import trixi_hip as th
    c = golden("kta.json")["khi_2d"]
    spec = dict(c["oracle"], ic=c["ic"], ic_params=tuple(c["ic_params"]))
    semi, _ = build(spec, precision)
    runs = {}
    # maxiters 7: a short solve, launched without graphs
    for tspan, maxiters in (((0.0, 0.37), None), ((0.0, 5.0), 23), ((0.0, 5.0), 7)):
        for dev, poll in ((False, 1), (True, 1), (True, 7), (True, 64)):
            ode = th.semidiscretize(semi, tspan)
            ac = th.AnalysisCallback(semi, interval=5)
            sol = th.solve(ode, th.CarpenterKennedy2N54(),
                           callback=th.CallbackSet(th.StepsizeCallback(cfl=0.9), ac),
                           maxiters=maxiters, device_steps=dev, poll_every=poll)
            runs[(tspan, dev, poll)] = sol
        host = runs[(tspan, False, 1)]
        assert host.nsteps == (maxiters if maxiters else host.nsteps) and host.nsteps > 5
        if maxiters == 7:
            assert all(runs[(tspan, True, p)].nsteps == 7 for p in (1, 7, 64))
        if maxiters is None:
            assert host.t == tspan[1]
        iters = [r["iter"] for r in host.analysis]
        assert iters == sorted(set(list(range(0, host.nsteps, 5)) + [host.nsteps])), iters
        assert host.analysis[-1]["t"] == host.t
        for poll in (1, 7, 64):
            d = runs[(tspan, True, poll)]
            assert (d.nsteps, d.t) == (host.nsteps, host.t), (poll, d.nsteps, host.nsteps, d.t, host.t)
            assert np.array_equal(d.u, host.u), poll
            assert len(d.analysis) == len(host.analysis), (poll, [r["iter"] for r in d.analysis], iters)
            for rd, rh in zip(d.analysis, host.analysis):
                assert (rd["t"], rd["iter"]) == (rh["t"], rh["iter"])
                for key in ("l2", "linf", "integrals"):
                    assert np.array_equal(rd[key], rh[key]), (poll, rd["iter"], key)
                assert rd["entropy"] == rh["entropy"]
    # an empty time span takes one step of size 0, as step! with limit_dt! does
    for dev in (False, True):
        ode = th.semidiscretize(semi, (0.2, 0.2))
        sol = th.solve(ode, th.CarpenterKennedy2N54(), callback=th.CallbackSet(th.StepsizeCallback(cfl=0.9)),
                       device_steps=dev)
        assert sol.nsteps == 1 and sol.t == 0.2 and np.array_equal(sol.u, ode.u0), dev
    semi.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("case", ["khi_2d", "mms_2d_nonperiodic", "shockcapturing_3d", "mms_3d_weak", "mms_3d_mortar"])
def test_fused_stage_equals_rhs_plus_update(case, precision, monkeypatch):
    """The element-kernel variant that applies the 2N update in its epilogue
    (the state alternates between two buffers) runs bitwise the same steps as
    the RHS kernel followed by rk2n_update_kernel: device-controlled solve with
    an odd step count, per-step graph replays and eager steps, and single stages."""
    import trixi_hip as th
    c = golden("kta.json")[case]
    spec = dict(c["oracle"], ic=c["ic"], ic_params=tuple(c["ic_params"]))
    alg = th.CarpenterKennedy2N54()
    na = 5
    a = (ctypes.c_double * na)(*alg.a)
    b = (ctypes.c_double * na)(*alg.b)
    cc = (ctypes.c_double * na)(*alg.c)
    out = {}
    for unfused in ("1", "0"):
        monkeypatch.setenv("TRIXI_HIP_UNFUSED_STAGE", unfused)
        semi, _ = build(spec, precision)
        lib, ctx = semi._lib, semi.ctx
        ode = th.semidiscretize(semi, (0.0, 5.0))
        sol = th.solve(ode, th.CarpenterKennedy2N54(), callback=th.CallbackSet(th.StepsizeCallback(cfl=0.9)),
                       maxiters=13)
        res = [sol.nsteps, sol.t, sol.u]
        u0 = semi.compute_coefficients(0.0)
        for graph in (1, 0):
            upload(semi, u0)
            for rep in range(3):
                assert lib.trixi_hip_rk_step(ctx, a, b, cc, na, 0.1 + rep * 1e-3, 1e-3, graph) == 0
            res.append(download(semi))
        upload(semi, u0)
        assert lib.trixi_hip_rk_stage(ctx, 0.0, 2e-3, 0.05, 1) == 0
        assert lib.trixi_hip_rk_stage(ctx, alg.a[1], alg.b[1] * 1e-3, 0.06, 0) == 0
        res.append(download(semi))
        out[unfused] = res
        semi.close()
    ref, fused = out["1"], out["0"]
    assert ref[:2] == fused[:2] and ref[0] == 13
    for k in range(2, len(ref)):
        assert np.array_equal(ref[k], fused[k]), k


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 5e-6)])
def test_rk_steps_with_time_independent_boundaries(precision, tol):
    """A time-independent Dirichlet boundary function (weak blast wave) is
    evaluated once per context: eager steps, graph replays and a
    device-controlled solve agree bitwise, and a step equals the oracle's."""
    import trixi_hip as th
    spec = dict(ndims=3, polydeg=3, level=2, cmin=(-2, -2, -2), cmax=(2, 2, 2), ic="weak_blast_wave",
                volume_integral="fd", surface_flux="llf", periodicity=(True, False, False))
    o = oracle_for(spec)
    alg = th.CarpenterKennedy2N54()
    na = 5
    a = (ctypes.c_double * na)(*alg.a)
    b = (ctypes.c_double * na)(*alg.b)
    c = (ctypes.c_double * na)(*alg.c)
    res = {}
    for graph in (0, 1):
        semi, _ = build(spec, precision)
        lib, ctx = semi._lib, semi.ctx
        u0 = semi.compute_coefficients(0.0)
        upload(semi, u0)
        outs = []
        for rep in range(3):
            assert lib.trixi_hip_rk_step(ctx, a, b, c, na, 0.1 + rep * 2e-3, 2e-3, graph) == 0
            outs.append(download(semi))
        res[graph] = outs
        if graph == 1:
            assert relerr(outs[0], o.rk_step(u0.copy(), 0.1, 2e-3)) < tol
            sols = []
            for dev in (False, True):
                ode = th.semidiscretize(semi, (0.0, 0.05))
                sols.append(th.solve(ode, th.CarpenterKennedy2N54(),
                                     callback=th.CallbackSet(th.StepsizeCallback(cfl=0.8)), device_steps=dev))
            assert sols[0].nsteps == sols[1].nsteps and np.array_equal(sols[0].u, sols[1].u)
        semi.close()
    for rep in range(3):
        assert np.array_equal(res[0][rep], res[1][rep]), rep


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("D,ic,params", [(3, "convergence_test", ()), (3, "density_pulse", ()),
                                         (3, "weak_blast_wave", ()), (3, "fourier_modes", (1,)),
                                         (2, "isentropic_vortex", (25.0,)), (2, "blast_wave", ()),
                                         (2, "blob", ()), (2, "weak_blast_wave", ()),
                                         (2, "kelvin_helmholtz", (0.0, 0))])
def test_device_compute_coefficients_matches_host(D, ic, params, precision):
    """trixi_hip_compute_coefficients (the reference's GPU compute_coefficients!)
    gives the host evaluation of the initial condition at every node up to the
    last bits of the transcendental functions (relative 1e-14 of the state's
    magnitude), on a refined mesh; the FP32 shadow is the rounded FP64 state."""
    cmin, cmax = (-1.0,) * D, (1.0,) * D
    patch = [((-0.5,) * D, (0.5,) * D)]
    spec = dict(ndims=D, polydeg=3, level=2 if D == 3 else 3, cmin=cmin, cmax=cmax, ic=ic, ic_params=params,
                volume_integral="fd", surface_flux="llf", refinement_patches=patch)
    if ic == "blob":
        spec["gamma"] = 5.0 / 3.0
    semi, _ = build(spec, precision)
    u_host = semi.compute_coefficients(0.3)
    semi.set_initial_state(0.3)
    u_dev = download(semi)
    scale = np.max(np.abs(u_host))
    assert np.max(np.abs(u_dev - u_host)) <= 1e-14 * scale, float(np.max(np.abs(u_dev - u_host)) / scale)
    # the native state (FP32 shadow in FP32 mode) is the one an upload of the
    # same FP64 state gives: a stage from either is bitwise the same
    lib, ctx = semi._lib, semi.ctx
    assert lib.trixi_hip_rk_stage(ctx, 0.0, 1e-3, 0.3, 1) == 0
    a = download(semi)
    upload(semi, u_dev)
    assert lib.trixi_hip_rk_stage(ctx, 0.0, 1e-3, 0.3, 1) == 0
    np.testing.assert_array_equal(a, download(semi))
    semi.close()
