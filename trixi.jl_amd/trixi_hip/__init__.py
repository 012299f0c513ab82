"""trixi_hip -- host-side mirror of Trixi.jl's interface for the MI355X DGSEM
right-hand side of the compressible Euler equations on a TreeMesh.

    mesh   = TreeMesh((0, 0, 0), (2, 2, 2), initial_refinement_level=5, periodicity=True)
    solver = DGSEM(polydeg=3, surface_flux=flux_lax_friedrichs,
                   volume_integral=VolumeIntegralFluxDifferencing(flux_ranocha))
    semi   = SemidiscretizationHyperbolic(mesh, CompressibleEulerEquations3D(1.4),
                                          initial_condition_convergence_test, solver,
                                          source_terms=source_terms_convergence_test,
                                          precision="fp32")
    ode    = semidiscretize(semi, (0.0, 1.0))
    sol    = solve(ode, CarpenterKennedy2N54(), callback=CallbackSet(StepsizeCallback(cfl=0.6)))

All numerical work runs in libtrixi_hip.so (HIP, gfx950) through the C ABI of
include/trixi_hip.h.
"""
from . import _lib
from .amr import AMRCallback, ControllerThreeLevel, ControllerThreeLevelCombined, IndicatorMax, first
from .basis import LobattoLegendreBasis
from .equations import (BoundaryConditionDirichlet, CompressibleEulerEquations2D, CompressibleEulerEquations3D,
                        cons2cons, cons2prim, varnames,
                        FluxLaxFriedrichs, boundary_condition_periodic, boundary_condition_slip_wall,
                        flux_central, flux_hllc, flux_lax_friedrichs, flux_ranocha, fourier_modes,
                        initial_condition_constant, initial_condition_convergence_test,
                        initial_condition_blast_wave, initial_condition_blob, initial_condition_density_pulse,
                        initial_condition_fourier_modes, initial_condition_isentropic_vortex,
                        initial_condition_kelvin_helmholtz_instability, initial_condition_weak_blast_wave,
                        max_abs_speed, max_abs_speed_naive, source_terms_convergence_test)
from .io import (SaveRestartCallback, SaveSolutionCallback, load_dt, load_mesh, load_mesh_file, load_restart_file,
                 load_time, load_timestep, read_solution_header, save_mesh_file, save_restart_file,
                 save_solution_file)
from .mesh import TreeMesh
from .semidiscretization import ODEProblem, SemidiscretizationHyperbolic, rhs_hyperbolic_, semidiscretize
from .solver import (DGSEM, IndicatorHennemannGassner, VolumeIntegralFluxDifferencing,
                     VolumeIntegralPureLGLFiniteVolume, VolumeIntegralShockCapturingHG, VolumeIntegralWeakForm,
                     density, density_pressure, pressure)
from .time_integration import AnalysisCallback, CallbackSet, CarpenterKennedy2N54, StepsizeCallback, solve

__all__ = [n for n in dir() if not n.startswith("_")]
