"""The FP64 oracle reproduces Trixi.jl's own known-answer tests (full
trajectories: CarpenterKennedy2N54 + StepsizeCallback + AnalysisCallback
errors). Values and parameters: tests/golden/kta.json (cited per case)."""
import numpy as np
import pytest

from conftest import make_oracle

CASES = ["mms_3d_weak", "mms_3d_split_central", "ec_3d", "ec_3d_constant", "shockcapturing_3d", "khi_2d",
         "vortex_2d", "vortex_2d_constant_one_step", "mms_2d_nonperiodic", "mms_3d_mortar", "ec_2d_slip_wall",
         "pure_fv_hllc_3d", "pure_fv_hllc_2d", "blast_wave_pure_fv_2d", "density_pulse_3d"]


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_trixi_kta(kta, name):
    c = kta[name]
    o = make_oracle(c["oracle"])
    o.set_initial_condition(c["ic"], c["ic_params"])
    u = o.compute_coefficients(c["tspan"][0])
    u, t, nsteps = o.solve(u, c["tspan"][0], c["tspan"][1], c["cfl"], c.get("maxiters", 0))
    l2, linf = o.error_norms(u, t)
    if c.get("roundoff"):
        # free-stream preservation: the reference errors are round-off (~1e-16..1e-14)
        assert np.all(l2 < 1e-13) and np.all(linf < 1e-12)
    else:
        assert t == pytest.approx(c["tspan"][1], abs=0, rel=1e-15)
        # rtol 1e-9 (tighter than the reference tests' sqrt(eps)) plus their
        # absolute floor of 500 eps for errors near round-off
        np.testing.assert_allclose(l2, c["l2"], rtol=1e-9, atol=500 * np.finfo(float).eps)
        np.testing.assert_allclose(linf, c["linf"], rtol=1e-9, atol=500 * np.finfo(float).eps)


def test_planted_mistake_is_caught(kta):
    """A wrong LLF wave-speed variant (max_abs_speed instead of the elixir's
    max_abs_speed_naive) moves the MMS errors far outside the KTA tolerance."""
    c = kta["mms_3d_weak"]
    spec = dict(c["oracle"], surface_flux="llf")
    o = make_oracle(spec)
    o.set_initial_condition(c["ic"], c["ic_params"])
    u, t, _ = o.solve(o.compute_coefficients(0.0), 0.0, 5.0, c["cfl"])
    l2, _ = o.error_norms(u, t)
    assert np.max(np.abs(l2 / np.array(c["l2"]) - 1)) > 1e-6


def test_planted_boundary_mistake_is_caught(kta):
    """Periodic faces in place of the Dirichlet boundaries of the non-periodic
    MMS run move its errors far outside the KTA tolerance."""
    c = kta["mms_2d_nonperiodic"]
    spec = dict(c["oracle"], periodicity=True)
    o = make_oracle(spec)
    o.set_initial_condition(c["ic"], c["ic_params"])
    u, t, _ = o.solve(o.compute_coefficients(0.0), 0.0, 2.0, c["cfl"])
    l2, _ = o.error_norms(u, t)
    assert np.max(np.abs(l2 / np.array(c["l2"]) - 1)) > 1e-6


def test_planted_hllc_mistake_is_caught(kta):
    """Local Lax-Friedrichs subcell fluxes in place of flux_hllc move the pure
    finite-volume errors far outside the KTA tolerance."""
    c = kta["pure_fv_hllc_2d"]
    o = make_oracle(dict(c["oracle"], fv_flux="llf"))
    o.set_initial_condition(c["ic"], c["ic_params"])
    u, t, _ = o.solve(o.compute_coefficients(0.0), 0.0, c["tspan"][1], c["cfl"])
    l2, _ = o.error_norms(u, t)
    assert np.max(np.abs(l2 / np.array(c["l2"]) - 1)) > 1e-6


def _amr_run(c, controller_overrides=None):
    from oracle import ControllerThreeLevel, ControllerThreeLevelCombined
    a = dict(c["amr"], **(controller_overrides or {}))
    o = make_oracle(c["oracle"])
    o.set_initial_condition(c["ic"], c["ic_params"])
    kw = dict(base_level=a["base_level"], med_level=a["med_level"], med_threshold=a["med_threshold"],
              max_level=a["max_level"], max_threshold=a["max_threshold"])
    ctl = ControllerThreeLevelCombined(max_threshold_secondary=a["max_threshold_secondary"], **kw) \
        if a.get("combined") else ControllerThreeLevel(**kw)
    u, t, nsteps, history = o.solve_amr(c["tspan"][0], c["tspan"][1], c["cfl"], ctl, interval=a["interval"],
                                        adapt_initial_condition=a["adapt_initial_condition"],
                                        adapt_initial_condition_only_refine=a["adapt_initial_condition_only_refine"],
                                        variable=a["variable"],
                                        indicator=tuple(a["indicator"]) if isinstance(a.get("indicator"), list)
                                        else a.get("indicator", "max"))
    return o, u, t, history


def test_oracle_reproduces_trixi_amr_kta(kta):
    """elixir_euler_amr.jl: the initial condition is adapted until the mesh
    settles, then the mesh is refined and coarsened every 5 steps while the
    pulse moves. The errors are dominated by the displacement of the pulse
    (the exact solution is the unmoved initial condition), so mesh decisions
    move them by as little as 3e-10 relative: rtol 1e-12 here."""
    c = kta["amr_3d"]
    o, u, t, history = _amr_run(c)
    # initial adaptation (4096 -> ... settled), then 3 adaptations during the run
    assert len(history) >= 5 and history[-1] == o.nelements
    assert len(set(history[-3:])) > 1  # the run adapts the mesh (refinement and coarsening)
    assert t == pytest.approx(c["tspan"][1], abs=0, rel=1e-15)
    l2, linf = o.error_norms(u, t)
    np.testing.assert_allclose(l2, c["l2"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(linf, c["linf"], rtol=1e-12, atol=0)


@pytest.mark.parametrize("mistake", [{"interval": 0}, {"max_threshold": 2.0}, {"med_threshold": 1.2}])
def test_planted_amr_mistake_is_caught(kta, mistake):
    """No adaptation during the run, a controller that never reaches the
    maximum level, or a shifted medium threshold each move the AMR errors
    outside the rtol of the KTA test above."""
    c = kta["amr_3d"]
    o, u, t, _ = _amr_run(c, mistake)
    l2, linf = o.error_norms(u, t)
    worst = max(np.max(np.abs(l2 / np.array(c["l2"]) - 1)), np.max(np.abs(linf / np.array(c["linf"]) - 1)))
    assert worst > 1e-11


def test_oracle_reproduces_trixi_blast_wave_amr_kta(kta):
    """elixir_euler_blast_wave_amr.jl: HG-indicator-driven AMR with shock
    capturing on the adapted meshes (206 steps, 42 adaptations, mostly
    coarsening from level 6 towards 4 away from the blast)."""
    c = kta["blast_wave_amr_2d"]
    o, u, t, history = _amr_run(c)
    assert len(history) > 30 and min(history) < 1000 < max(history)
    assert t == pytest.approx(c["tspan"][1], abs=0, rel=1e-15)
    l2, linf = o.error_norms(u, t)
    np.testing.assert_allclose(l2, c["l2"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(linf, c["linf"], rtol=1e-12, atol=0)


def test_planted_blast_wave_amr_mistake_is_caught(kta):
    """AMR decisions from IndicatorMax(rho) at the same thresholds instead of the
    HG indicator move the errors far outside the tolerance."""
    c = dict(kta["blast_wave_amr_2d"])
    c["amr"] = dict(c["amr"], indicator="max", max_threshold=1.01)
    o, u, t, _ = _amr_run(c)
    l2, _ = o.error_norms(u, t)
    assert np.max(np.abs(l2 / np.array(kta["blast_wave_amr_2d"]["l2"]) - 1)) > 1e-8


def test_oracle_reproduces_trixi_khi_amr_kta(kta):
    """elixir_euler_kelvin_helmholtz_instability_amr.jl: an AMR-only HG indicator
    (own parameters, variable density, no smoothing), med_level = current level,
    an adaptation after every step."""
    c = kta["khi_amr_2d"]
    o, u, t, history = _amr_run(c)
    assert len(history) > 30 and len(set(history)) > 2
    l2, linf = o.error_norms(u, t)
    np.testing.assert_allclose(l2, c["l2"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(linf, c["linf"], rtol=1e-12, atol=0)


def test_planted_khi_amr_mistake_is_caught(kta):
    """The volume integral's HG indicator (density_pressure, alpha_max 0.002,
    smoothing) in place of the AMR indicator's own parameters moves the errors
    outside the tolerance."""
    c = dict(kta["khi_amr_2d"])
    c["amr"] = dict(c["amr"], indicator="hg")
    o, u, t, _ = _amr_run(c)
    l2, _ = o.error_norms(u, t)
    assert np.max(np.abs(l2 / np.array(kta["khi_amr_2d"]["l2"]) - 1)) > 1e-8


def test_oracle_reproduces_trixi_blob_amr_kta(kta):
    """elixir_euler_blob_amr.jl: gamma 5/3, polydeg 4, the shock-capturing
    indicator on pressure, ControllerThreeLevelCombined (an AMR-only density HG
    indicator plus the shock-capturing one), an adaptation after every step."""
    c = kta["blob_amr_2d"]
    o, u, t, history = _amr_run(c)
    assert len(history) > 100
    l2, linf = o.error_norms(u, t)
    np.testing.assert_allclose(l2, c["l2"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(linf, c["linf"], rtol=1e-12, atol=0)


@pytest.mark.parametrize("mistake", [{"combined": False}, {"hg_variable": "density_pressure"}])
def test_planted_blob_amr_mistake_is_caught(kta, mistake):
    """Dropping the secondary indicator (the run then blows up), or the
    shock-capturing indicator on density_pressure instead of pressure, moves
    the errors outside the tolerance."""
    c = dict(kta["blob_amr_2d"])
    if "hg_variable" in mistake:
        c["oracle"] = dict(c["oracle"], **mistake)
        o, u, t, _ = _amr_run(c)
    else:
        o, u, t, _ = _amr_run(c, mistake)
    l2, _ = o.error_norms(u, t)
    worst = np.max(np.abs(l2 / np.array(kta["blob_amr_2d"]["l2"]) - 1))
    assert not np.isfinite(worst) or worst > 1e-8
