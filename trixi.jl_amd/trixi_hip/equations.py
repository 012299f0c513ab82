"""CompressibleEulerEquations{2,3}D, numerical-flux markers and the initial
conditions / source terms of the supported setups.

Numerical fluxes are markers: the HIP kernels implement them (they are not
Python callables here). Initial conditions are evaluated on the host during
setup (compute_coefficients!, src/solvers/dg.jl:1280-1325); each returns the
conservative state at the given node coordinates (arrays of shape [..., D]).
"""
import math

import numpy as np

from . import _lib


class CompressibleEulerEquations:
    """`CompressibleEulerEquations{2,3}D(gamma)` (compressible_euler_3d.jl:45-54)."""

    def __init__(self, ndims, gamma=1.4):
        self.ndims = ndims
        self.gamma = float(gamma)
        self.inv_gamma_minus_one = 1.0 / (self.gamma - 1.0)

    def nvariables(self):
        return self.ndims + 2

    def varnames(self):
        if self.ndims == 3:
            return ("rho", "rho_v1", "rho_v2", "rho_v3", "rho_e_total")
        return ("rho", "rho_v1", "rho_v2", "rho_e_total")

    def prim2cons(self, prim):
        prim = np.asarray(prim, dtype=np.float64)
        u = np.empty_like(prim)
        rho = prim[..., 0]
        u[..., 0] = rho
        ke2 = np.zeros_like(rho)
        for d in range(self.ndims):
            u[..., 1 + d] = rho * prim[..., 1 + d]
            ke2 += u[..., 1 + d] * prim[..., 1 + d]
        u[..., -1] = prim[..., -1] * self.inv_gamma_minus_one + 0.5 * ke2
        return u

    def cons2prim(self, u):
        u = np.asarray(u, dtype=np.float64)
        w = np.empty_like(u)
        rho = u[..., 0]
        w[..., 0] = rho
        ke2 = np.zeros_like(rho)
        for d in range(self.ndims):
            w[..., 1 + d] = u[..., 1 + d] / rho
            ke2 += u[..., 1 + d] * w[..., 1 + d]
        w[..., -1] = (self.gamma - 1) * (u[..., -1] - 0.5 * ke2)
        return w

    def __repr__(self):
        return f"CompressibleEulerEquations{self.ndims}D(gamma={self.gamma})"


def cons2cons(u, equations):
    """`cons2cons`: the conserved variables themselves."""
    return np.asarray(u, dtype=np.float64)


def cons2prim(u, equations):
    """`cons2prim` (compressible_euler_3d.jl:1783-1794): rho, v, p."""
    return equations.cons2prim(u)


def varnames(solution_variables, equations):
    """`varnames(cons2cons | cons2prim, equations)` (compressible_euler_3d.jl:65-70)."""
    if solution_variables is cons2cons:
        return equations.varnames()
    if solution_variables is cons2prim:
        return ("rho",) + tuple(f"v{d + 1}" for d in range(equations.ndims)) + ("p",)
    raise ValueError("solution_variables must be cons2cons or cons2prim")


def CompressibleEulerEquations2D(gamma=1.4):
    return CompressibleEulerEquations(2, gamma)


def CompressibleEulerEquations3D(gamma=1.4):
    return CompressibleEulerEquations(3, gamma)


# ---------------------------------------------------------------------------
# Numerical fluxes (markers) -- numerical_fluxes.jl, compressible_euler_*d.jl
# ---------------------------------------------------------------------------
class _Flux:
    def __init__(self, name, code):
        self.name = name
        self.code = code

    def __repr__(self):
        return self.name


def max_abs_speed(*_):
    """Marker for `max_abs_speed` (compressible_euler_3d.jl:1156-1177)."""
    raise TypeError("max_abs_speed is evaluated inside the HIP kernels")


def max_abs_speed_naive(*_):
    """Marker for `max_abs_speed_naive` (compressible_euler_3d.jl:1112-1133)."""
    raise TypeError("max_abs_speed_naive is evaluated inside the HIP kernels")


def FluxLaxFriedrichs(speed=max_abs_speed):
    """`FluxLaxFriedrichs(max_abs_speed = max_abs_speed)` (numerical_fluxes.jl:229-253)."""
    if speed is max_abs_speed:
        return _Flux("FluxLaxFriedrichs(max_abs_speed)", _lib.FLUX_LLF)
    if speed is max_abs_speed_naive:
        return _Flux("FluxLaxFriedrichs(max_abs_speed_naive)", _lib.FLUX_LLF_NAIVE)
    raise ValueError("unsupported wave-speed estimate")


flux_lax_friedrichs = FluxLaxFriedrichs()
flux_ranocha = _Flux("flux_ranocha", _lib.FLUX_RANOCHA)
flux_central = _Flux("flux_central", _lib.FLUX_CENTRAL)
flux_hllc = _Flux("flux_hllc", _lib.FLUX_HLLC)  # surface and FV flux (compressible_euler_3d.jl:1417-1541)


# ---------------------------------------------------------------------------
# Initial conditions (callables f(x, t, equations) -> u, vectorised over x[..., D])
# ---------------------------------------------------------------------------
class InitialCondition:
    """A pointwise initial condition. `code`/`params`/`modes` name the same
    function for the device (error norms, Dirichlet boundary values);
    device_exact is False when the device evaluates only part of it (the seeded
    noise of the KHI setup)."""

    def __init__(self, name, func, code, params=(), modes=None, device_exact=True):
        self.name = name
        self.func = func
        self.code = code
        self.params = tuple(params)
        self.modes = modes
        self.device_exact = device_exact

    def __call__(self, x, t, equations):
        return self.func(np.asarray(x, dtype=np.float64), float(t), equations)

    def __repr__(self):
        return self.name


def _ic_constant(x, t, eq):
    # compressible_euler_3d.jl:72-80, compressible_euler_2d.jl:71-78
    shp = x.shape[:-1]
    u = np.empty(shp + (eq.nvariables(),))
    u[..., 0] = 1.0
    u[..., 1] = 0.1
    u[..., 2] = -0.2
    if eq.ndims == 3:
        u[..., 3] = 0.7
    u[..., -1] = 10.0
    return u


def _ic_convergence(x, t, eq):
    # compressible_euler_3d.jl:94-111, compressible_euler_2d.jl:87-106
    c, A, L = 2.0, 0.1, 2.0
    f = 1.0 / L
    om = 2 * math.pi * f
    ini = c + A * np.sin(om * (np.sum(x, axis=-1) - t))
    u = np.empty(x.shape[:-1] + (eq.nvariables(),))
    for v in range(eq.nvariables() - 1):
        u[..., v] = ini
    u[..., -1] = ini ** 2
    return u


def _ic_weak_blast_wave(x, t, eq):
    if eq.ndims == 2:
        # compressible_euler_2d.jl:181-200 (polar coordinates)
        r = np.sqrt(x[..., 0] ** 2 + x[..., 1] ** 2)
        phi = np.arctan2(x[..., 1], x[..., 0])
        inside = r <= 0.5
        prim = np.empty(x.shape[:-1] + (4,))
        prim[..., 0] = np.where(inside, 1.1691, 1.0)
        prim[..., 1] = np.where(inside, 0.1882 * np.cos(phi), 0.0)
        prim[..., 2] = np.where(inside, 0.1882 * np.sin(phi), 0.0)
        prim[..., 3] = np.where(inside, 1.245, 1.0)
        return eq.prim2cons(prim)
    # compressible_euler_3d.jl:163-183
    xn, yn, zn = x[..., 0], x[..., 1], x[..., 2]
    r = np.sqrt(xn ** 2 + yn ** 2 + zn ** 2)
    phi = np.arctan2(yn, xn)
    with np.errstate(invalid="ignore", divide="ignore"):
        theta = np.where(r == 0, 0.0, np.arccos(np.where(r == 0, 1.0, zn / np.where(r == 0, 1.0, r))))
    inside = r <= 0.5
    prim = np.empty(x.shape[:-1] + (5,))
    prim[..., 0] = np.where(inside, 1.1691, 1.0)
    prim[..., 1] = np.where(inside, 0.1882 * np.cos(phi) * np.sin(theta), 0.0)
    prim[..., 2] = np.where(inside, 0.1882 * np.sin(phi) * np.sin(theta), 0.0)
    prim[..., 3] = np.where(inside, 0.1882 * np.cos(theta), 0.0)
    prim[..., 4] = np.where(inside, 1.245, 1.0)
    return eq.prim2cons(prim)


def _make_vortex(p_background):
    def f(x, t, eq):
        # examples/tree_2d_dgsem/elixir_euler_vortex.jl:18-50
        rho, v1, v2, p = 1.0, 1.0, 1.0, p_background
        rt = p / rho
        c0, c1 = -x[..., 1], x[..., 0]
        r2 = c0 ** 2 + c1 ** 2
        du = 5.0 / (2 * math.pi) * np.exp(0.5 * (1 - r2))
        dtemp = -(eq.gamma - 1) / (2 * eq.gamma * rt) * du ** 2
        prim = np.empty(x.shape[:-1] + (4,))
        prim[..., 0] = rho * (1 + dtemp) ** (1 / (eq.gamma - 1))
        prim[..., 1] = v1 + du * c0
        prim[..., 2] = v2 + du * c1
        prim[..., 3] = p * (1 + dtemp) ** (eq.gamma / (eq.gamma - 1))
        return eq.prim2cons(prim)
    return f


def _hash_noise(x, seed):
    """Deterministic noise in [-1, 1) keyed on coordinates rounded to 1e-12
    (shared face nodes of neighbouring elements get the same value)."""
    q = np.rint(x * 1e12).astype(np.int64).view(np.uint64)
    with np.errstate(over="ignore"):
        h = np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + np.uint64(0x632BE59BD9B4E019)
        for d in range(x.shape[-1]):
            h = h ^ (q[..., d] + np.uint64(0x9E3779B97F4A7C15) + (h << np.uint64(6)) + (h >> np.uint64(2)))
            z = h
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            h = z ^ (z >> np.uint64(31))
    return (h >> np.uint64(11)).astype(np.float64) * (2.0 ** -53) * 2.0 - 1.0


def _make_khi(noise_amplitude=0.0, seed=0):
    def f(x, t, eq):
        # examples/tree_2d_dgsem/elixir_euler_kelvin_helmholtz_instability.jl:18-31
        B = np.tanh(15 * x[..., 1] + 7.5) - np.tanh(15 * x[..., 1] - 7.5)
        prim = np.empty(x.shape[:-1] + (4,))
        prim[..., 0] = 0.5 + 0.75 * B
        prim[..., 1] = 0.5 * (B - 1)
        prim[..., 2] = 0.1 * np.sin(math.pi * 2 * x[..., 0])
        if noise_amplitude:
            prim[..., 2] += noise_amplitude * _hash_noise(x, seed)
        prim[..., 3] = 1.0
        return eq.prim2cons(prim)
    return f


def _make_fourier(modes, a_rho, a_v):
    modes = np.asarray(modes, dtype=np.float64).reshape(4, 8, 4)

    def f(x, t, eq):
        D = eq.ndims
        prim = np.empty(x.shape[:-1] + (D + 2,))
        for fld in range(1 + D):
            s = np.zeros(x.shape[:-1])
            for m in range(8):
                k = modes[fld, m]
                arg = np.zeros(x.shape[:-1])
                for d in range(D):
                    arg = arg + k[d] * (x[..., d] + 1.0)
                s = s + np.sin(math.pi * arg + k[3])
            prim[..., fld] = 1.0 + a_rho * s if fld == 0 else a_v * s
        prim[..., D + 1] = 1.0
        return eq.prim2cons(prim)
    return f


initial_condition_constant = InitialCondition("initial_condition_constant", _ic_constant, _lib.IC_CONSTANT)
initial_condition_convergence_test = InitialCondition(
    "initial_condition_convergence_test", _ic_convergence, _lib.IC_CONVERGENCE_TEST)
initial_condition_weak_blast_wave = InitialCondition(
    "initial_condition_weak_blast_wave", _ic_weak_blast_wave, _lib.IC_WEAK_BLAST_WAVE)


def _ic_density_pulse(x, t, eq):
    # examples/tree_3d_dgsem/elixir_euler_amr.jl:15-26: a Gaussian density pulse
    # advected with v = 1 in every direction at p = 1
    rho = 1 + np.exp(-np.sum(x ** 2, axis=-1)) / 2
    u = np.empty(x.shape[:-1] + (eq.nvariables(),))
    u[..., 0] = rho
    for d in range(eq.ndims):
        u[..., 1 + d] = rho
    u[..., -1] = 1 / (eq.gamma - 1) + 1 / 2 * rho * eq.ndims
    return u


def _ic_blast_wave(x, t, eq):
    # examples/tree_2d_dgsem/elixir_euler_blast_wave_amr.jl:14-33: the medium blast
    # wave of Hennemann & Gassner (2D; p = 1e-3 outside r = 0.5)
    r = np.sqrt(x[..., 0] ** 2 + x[..., 1] ** 2)
    phi = np.arctan2(x[..., 1], x[..., 0])
    out = r > 0.5
    prim = np.empty(x.shape[:-1] + (4,))
    prim[..., 0] = np.where(out, 1.0, 1.1691)
    prim[..., 1] = np.where(out, 0.0, 0.1882 * np.cos(phi))
    prim[..., 2] = np.where(out, 0.0, 0.1882 * np.sin(phi))
    prim[..., 3] = np.where(out, 1.0e-3, 1.245)
    return eq.prim2cons(prim)


initial_condition_blast_wave = InitialCondition(
    "initial_condition_blast_wave", _ic_blast_wave, _lib.IC_BLAST_WAVE)

def _ic_blob(x, t, eq):
    # examples/tree_2d_dgsem/elixir_euler_blob_amr.jl:13-58: a dense blob (density
    # contrast 10, radius 1, centre (-15, 0)) in a Mach 2.7 stream (Agertz et al. 2006)
    R, dens0, Chi = 1.0, 1.0, 10.0
    tau_cr = 1.0 / 1.6
    velx0 = 2 * R * math.sqrt(Chi) / tau_cr
    c = velx0 / 2.7
    p0 = c * c * dens0 / eq.gamma
    r = np.sqrt((x[..., 0] - (-15)) ** 2 + (x[..., 1] - 0) ** 2)
    shape = 0.5 * (1 + (np.tanh(2 * (r + R)) - (np.tanh(2 * (r - R)) + 1)))
    prim = np.empty(x.shape[:-1] + (4,))
    prim[..., 0] = dens0 + (Chi - 1) * shape
    prim[..., 1] = velx0 - velx0 * shape
    prim[..., 2] = 0.0
    prim[..., 3] = p0
    return eq.prim2cons(prim)


initial_condition_blob = InitialCondition("initial_condition_blob", _ic_blob, _lib.IC_BLOB)

initial_condition_density_pulse = InitialCondition(
    "initial_condition_density_pulse", _ic_density_pulse, _lib.IC_DENSITY_PULSE)


def initial_condition_isentropic_vortex(p_background=25.0):
    return InitialCondition("initial_condition_isentropic_vortex", _make_vortex(float(p_background)),
                            _lib.IC_ISENTROPIC_VORTEX, params=(float(p_background),))


def initial_condition_kelvin_helmholtz_instability(noise_amplitude=0.0, seed=0):
    return InitialCondition("initial_condition_kelvin_helmholtz_instability",
                            _make_khi(float(noise_amplitude), int(seed)), _lib.IC_KELVIN_HELMHOLTZ,
                            device_exact=not noise_amplitude)


def fourier_modes(seed, nmodes=8, kmax=4):
    """Seeded mode table [4 fields][8 modes][kx, ky, kz, phase]: integer
    wavenumbers in 1..kmax, uniform phases, numpy.random.default_rng(seed)."""
    rng = np.random.default_rng(seed)
    k = rng.integers(1, kmax + 1, size=(4, nmodes, 3)).astype(np.float64)
    ph = rng.uniform(0.0, 2 * math.pi, size=(4, nmodes, 1))
    return np.concatenate([k, ph], axis=-1)


def initial_condition_fourier_modes(seed=0, a_rho=0.1, a_v=0.2):
    modes = fourier_modes(seed)
    return InitialCondition("initial_condition_fourier_modes", _make_fourier(modes, a_rho, a_v),
                            _lib.IC_FOURIER_MODES, params=(a_rho, a_v), modes=modes)


class SourceTerms:
    def __init__(self, name, code):
        self.name = name
        self.code = code

    def __repr__(self):
        return self.name


source_terms_convergence_test = SourceTerms("source_terms_convergence_test", _lib.SRC_CONVERGENCE_TEST)


# ---------------------------------------------------------------------------
# Boundary conditions (src/equations/equations.jl:120-182)
# ---------------------------------------------------------------------------
class _PeriodicBC:
    def __repr__(self):
        return "boundary_condition_periodic"


boundary_condition_periodic = _PeriodicBC()


class BoundaryConditionDirichlet:
    """`BoundaryConditionDirichlet(boundary_value_function)` (equations.jl:159-182):
    the outer state at a boundary node x is boundary_value_function(x, t). The
    device evaluates it as the semidiscretization's initial condition, so the
    function must be that initial condition."""

    code = _lib.BC_DIRICHLET

    def __init__(self, boundary_value_function):
        if not isinstance(boundary_value_function, InitialCondition):
            raise TypeError("the boundary value function must be an initial condition known to the device")
        if not boundary_value_function.device_exact:
            raise ValueError(f"{boundary_value_function} is not evaluated exactly on the device (seeded noise)")
        self.boundary_value_function = boundary_value_function

    def __repr__(self):
        return f"BoundaryConditionDirichlet({self.boundary_value_function})"


class _SlipWallBC:
    """`boundary_condition_slip_wall` (compressible_euler_2d.jl:322-420,
    compressible_euler_3d.jl:319-415, TreeMesh methods): the wall flux from the
    inner state through the pressure Riemann problem (Toro 2009, sec. 6.3.3);
    the surface flux of the solver is not used at the wall."""

    code = _lib.BC_SLIP_WALL

    def __repr__(self):
        return "boundary_condition_slip_wall"


boundary_condition_slip_wall = _SlipWallBC()
