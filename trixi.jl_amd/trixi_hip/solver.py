"""DGSEM solver description: volume integrals, indicator, precision.

Mirrors src/solvers/dg.jl:135-221, 574-582, 913-918 and
src/solvers/dgsem/dgsem.jl:28-73, src/solvers/dgsem/indicators.jl:50-100.
The objects only describe the discretisation; the HIP library evaluates it.
"""
from . import _lib
from .basis import LobattoLegendreBasis
from .equations import _Flux, flux_central, flux_lax_friedrichs


class VolumeIntegralWeakForm:
    code = _lib.VI_WEAK

    def __repr__(self):
        return "VolumeIntegralWeakForm()"


class VolumeIntegralFluxDifferencing:
    code = _lib.VI_FLUX_DIFFERENCING

    def __init__(self, volume_flux):
        if not isinstance(volume_flux, _Flux) or volume_flux.code not in (_lib.FLUX_RANOCHA, _lib.FLUX_CENTRAL):
            raise ValueError("flux differencing supports flux_ranocha and flux_central")
        self.volume_flux = volume_flux

    def __repr__(self):
        return f"VolumeIntegralFluxDifferencing({self.volume_flux})"


class VolumeIntegralPureLGLFiniteVolume:
    code = _lib.VI_PURE_FV

    def __init__(self, volume_flux_fv):
        self.volume_flux_fv = volume_flux_fv

    def __repr__(self):
        return f"VolumeIntegralPureLGLFiniteVolume({self.volume_flux_fv})"


class IndicatorHennemannGassner:
    """`IndicatorHennemannGassner(equations, basis; alpha_max, alpha_min, alpha_smooth, variable)`
    (indicators.jl:50-117), or `IndicatorHennemannGassner(semi; ...)` as the
    indicator of an AMR controller. Called with a semidiscretization it returns
    the blending factors of the resident state with these parameters, computed
    on the device (trixi_hip_indicator_hg)."""

    def __init__(self, equations, basis=None, alpha_max=0.5, alpha_min=0.001, alpha_smooth=True,
                 variable="density_pressure"):
        if variable not in _HG_VARIABLES:
            raise ValueError("variable must be density_pressure, density or pressure")
        self.alpha_max = float(alpha_max)
        self.alpha_min = float(alpha_min)
        self.alpha_smooth = bool(alpha_smooth)
        self.variable = variable

    def __call__(self, semi):
        import ctypes
        import numpy as np
        alpha = np.empty(semi.nelements())
        _lib.check(semi._lib.trixi_hip_indicator_hg(semi.ctx, self.alpha_max, self.alpha_min, int(self.alpha_smooth),
                                                    _HG_VARIABLES[self.variable],
                                                    alpha.ctypes.data_as(ctypes.c_void_p)),
                   semi.ctx, "trixi_hip_indicator_hg")
        return alpha

    def __repr__(self):
        return (f"IndicatorHennemannGassner({self.variable}, alpha_max={self.alpha_max}, "
                f"alpha_min={self.alpha_min}, alpha_smooth={self.alpha_smooth})")


density_pressure = "density_pressure"
density = "density"
pressure = "pressure"
_HG_VARIABLES = {"density_pressure": _lib.VAR_DENSITY_PRESSURE, "density": _lib.VAR_DENSITY,
                 "pressure": _lib.VAR_PRESSURE}


class VolumeIntegralShockCapturingHG:
    code = _lib.VI_SHOCK_CAPTURING_HG

    def __init__(self, indicator, volume_flux_dg=flux_central, volume_flux_fv=flux_lax_friedrichs):
        self.indicator = indicator
        self.volume_flux_dg = volume_flux_dg
        self.volume_flux_fv = volume_flux_fv

    def __repr__(self):
        return (f"VolumeIntegralShockCapturingHG({self.indicator}; volume_flux_dg={self.volume_flux_dg}, "
                f"volume_flux_fv={self.volume_flux_fv})")


class DGSEM:
    """`DGSEM(basis, surface_flux, volume_integral)` / `DGSEM(; polydeg, surface_flux, volume_integral)`."""

    def __init__(self, basis=None, surface_flux=flux_central, volume_integral=None, polydeg=None):
        if basis is None:
            if polydeg is None:
                raise ValueError("give a basis or polydeg")
            basis = LobattoLegendreBasis(polydeg)
        elif isinstance(basis, int):
            basis = LobattoLegendreBasis(basis)
        self.basis = basis
        self.surface_flux = surface_flux
        self.volume_integral = volume_integral if volume_integral is not None else VolumeIntegralWeakForm()
        # the HIP element kernels compile flux_hllc into the pure finite-volume
        # variants only (as one more run-time branch it would cost every other
        # variant registers and occupancy)
        fluxes = (surface_flux, getattr(self.volume_integral, "volume_flux_fv", None))
        if any(getattr(f, "code", None) == _lib.FLUX_HLLC for f in fluxes) and \
                not isinstance(self.volume_integral, VolumeIntegralPureLGLFiniteVolume):
            raise ValueError("flux_hllc is supported with VolumeIntegralPureLGLFiniteVolume only")

    def polydeg(self):
        return self.basis.polydeg

    def nnodes(self):
        return self.basis.nnodes()

    def __repr__(self):
        return f"DGSEM(polydeg={self.polydeg()}, {self.surface_flux}, {self.volume_integral})"
