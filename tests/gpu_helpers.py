"""Shared set-up for the GPU tests: one description of a case builds both the
HIP semidiscretisation and the matching FP64 oracle."""
import numpy as np

import trixi_hip as th
from oracle import Oracle

FLUX = {"llf": th.flux_lax_friedrichs, "llf_naive": th.FluxLaxFriedrichs(th.max_abs_speed_naive),
        "ranocha": th.flux_ranocha, "central": th.flux_central, "hllc": th.flux_hllc}
IC_ORACLE = {"constant": "constant", "convergence_test": "convergence_test", "weak_blast_wave": "weak_blast_wave",
             "isentropic_vortex": "isentropic_vortex", "kelvin_helmholtz": "kelvin_helmholtz",
             "fourier_modes": "fourier_modes", "density_pulse": "density_pulse", "blast_wave": "blast_wave",
             "blob": "blob"}


def make_ic(name, params=()):
    if name == "constant":
        return th.initial_condition_constant
    if name == "convergence_test":
        return th.initial_condition_convergence_test
    if name == "weak_blast_wave":
        return th.initial_condition_weak_blast_wave
    if name == "isentropic_vortex":
        return th.initial_condition_isentropic_vortex(*(params or (25.0,)))
    if name == "kelvin_helmholtz":
        return th.initial_condition_kelvin_helmholtz_instability(*params)
    if name == "fourier_modes":
        return th.initial_condition_fourier_modes(*(params or (0,)))
    if name == "density_pulse":
        return th.initial_condition_density_pulse
    if name == "blast_wave":
        return th.initial_condition_blast_wave
    if name == "blob":
        return th.initial_condition_blob
    raise KeyError(name)


def build(spec, precision="fp64"):
    """spec: ndims, polydeg, level, cmin, cmax, volume_integral, volume_flux, surface_flux,
    fv_flux, alpha_max, alpha_min, source, ic, ic_params."""
    D, N = spec["ndims"], spec["polydeg"]
    gamma = spec.get("gamma", 1.4)
    eq = th.CompressibleEulerEquations3D(gamma) if D == 3 else th.CompressibleEulerEquations2D(gamma)
    basis = th.LobattoLegendreBasis(N)
    vi = spec.get("volume_integral", "fd")
    vflux = spec.get("volume_flux", "ranocha")
    fvflux = spec.get("fv_flux", "llf")
    amax, amin = spec.get("alpha_max", 0.5), spec.get("alpha_min", 0.001)
    if vi == "weak":
        VI = th.VolumeIntegralWeakForm()
    elif vi == "fd":
        VI = th.VolumeIntegralFluxDifferencing(FLUX[vflux])
    elif vi == "hg":
        VI = th.VolumeIntegralShockCapturingHG(th.IndicatorHennemannGassner(eq, basis, alpha_max=amax, alpha_min=amin,
                                                                            variable=spec.get("hg_variable",
                                                                                              "density_pressure")),
                                               volume_flux_dg=FLUX[vflux], volume_flux_fv=FLUX[fvflux])
    else:
        VI = th.VolumeIntegralPureLGLFiniteVolume(FLUX[fvflux])
    solver = th.DGSEM(basis, FLUX[spec.get("surface_flux", "llf")], VI)
    mesh = th.TreeMesh(spec["cmin"], spec["cmax"], initial_refinement_level=spec["level"],
                       periodicity=spec.get("periodicity", True),
                       refinement_patches=spec.get("refinement_patches", ()))
    ic = make_ic(spec["ic"], spec.get("ic_params", ()))
    src = th.source_terms_convergence_test if spec.get("source") == "convergence_test" else None
    if mesh.nboundaries == 0:
        bc = th.boundary_condition_periodic
    elif spec.get("boundary_condition") == "slip_wall":
        bc = th.boundary_condition_slip_wall
    else:
        bc = th.BoundaryConditionDirichlet(ic)
    semi = th.SemidiscretizationHyperbolic(mesh, eq, ic, solver, source_terms=src, boundary_conditions=bc,
                                           precision=precision, elements_per_block=spec.get("epb", 0),
                                           threads_per_block=spec.get("nt", 0))
    return semi, ic


def oracle_for(spec):
    o = Oracle(spec["ndims"], spec["polydeg"], spec["level"], spec["cmin"], spec["cmax"],
               volume_integral=spec.get("volume_integral", "fd"), volume_flux=spec.get("volume_flux", "ranocha"),
               surface_flux=spec.get("surface_flux", "llf"), fv_flux=spec.get("fv_flux", "llf"),
               alpha_max=spec.get("alpha_max", 0.5), alpha_min=spec.get("alpha_min", 0.001),
               source=spec.get("source", "none"), periodicity=spec.get("periodicity", True),
               refinement_patches=spec.get("refinement_patches", ()), gamma=spec.get("gamma", 1.4),
               hg_variable=spec.get("hg_variable", "density_pressure"),
               boundary_condition=spec.get("boundary_condition", "dirichlet"))
    if spec.get("periodicity", True) is not True:  # the boundary values come from the initial condition
        ic = make_ic(spec["ic"], spec.get("ic_params", ()))
        o.set_initial_condition(spec["ic"], tuple(ic.params), ic.modes.reshape(-1) if ic.modes is not None else None)
    return o


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def flux_scale(semi, u):
    """inverse_jacobian_max * max |physical flux component| over nodes and
    directions: the magnitude of the terms whose O(h) difference du is."""
    D = semi.mesh.ndims
    U = u.reshape(-1, D + 2)
    w = semi.equations.cons2prim(U)
    p = w[:, -1]
    fmax = 0.0
    for o in range(D):
        vo = w[:, 1 + o]
        f = [U[:, 1 + o]] + [U[:, 1 + o] * w[:, 1 + d] + (p if d == o else 0) for d in range(D)] + [(U[:, -1] + p) * vo]
        fmax = max(fmax, max(float(np.max(np.abs(x))) for x in f))
    return float(np.max(semi.inverse_jacobian)) * fmax


def fp32_abs_bound(semi, u, K=200.0):
    """Round-off bound of the FP32 bulk RHS: K * eps32 * J^-1 * max|f| (absolute)."""
    return K * float(np.finfo(np.float32).eps) * flux_scale(semi, u)


def fp64_abs_bound(semi, u, K=200.0):
    """The same round-off bound with eps64 (free-stream residuals of the FP64 RHS)."""
    return K * float(np.finfo(np.float64).eps) * flux_scale(semi, u)
