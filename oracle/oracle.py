"""ctypes wrapper of oracle/libtrixi_oracle.so (the C++ FP64 restatement of
Trixi.jl's TreeMesh DGSEM Euler right-hand side). Test infrastructure only."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtrixi_oracle.so")

VI = {"weak": 0, "fd": 1, "hg": 2, "fv": 3}
FLUX = {"llf": 0, "llf_naive": 1, "ranocha": 2, "central": 3, "hllc": 4}
IC = {"constant": 0, "convergence_test": 1, "weak_blast_wave": 2, "isentropic_vortex": 3,
      "kelvin_helmholtz": 4, "fourier_modes": 5, "density_pulse": 6, "blast_wave": 7, "blob": 8}
HG_VARIABLE = {"density_pressure": 0, "density": 1, "pressure": 2}


class _Config(ctypes.Structure):
    _fields_ = [("ndims", ctypes.c_int), ("polydeg", ctypes.c_int), ("initial_refinement_level", ctypes.c_int),
                ("coordinates_min", ctypes.c_double * 3), ("coordinates_max", ctypes.c_double * 3),
                ("gamma", ctypes.c_double), ("volume_integral", ctypes.c_int), ("volume_flux", ctypes.c_int),
                ("surface_flux", ctypes.c_int), ("fv_flux", ctypes.c_int), ("alpha_max", ctypes.c_double),
                ("alpha_min", ctypes.c_double), ("alpha_smooth", ctypes.c_int), ("source_terms", ctypes.c_int),
                ("periodicity", ctypes.c_int * 3), ("boundary_condition", ctypes.c_int),
                ("npatches", ctypes.c_int), ("patch_min", (ctypes.c_double * 3) * 8),
                ("patch_max", (ctypes.c_double * 3) * 8), ("hg_variable", ctypes.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        L.oracle_create.restype = P
        L.oracle_create.argtypes = [ctypes.POINTER(_Config)]
        L.oracle_destroy.argtypes = [P]
        for nm in ("oracle_nelements", "oracle_ninterfaces", "oracle_nboundaries", "oracle_nmortars", "oracle_nnodes"):
            getattr(L, nm).restype = I
            getattr(L, nm).argtypes = [P]
        L.oracle_ndofs.restype = ctypes.c_long
        L.oracle_ndofs.argtypes = [P]
        L.oracle_get_mesh.argtypes = [P] + [P] * 6
        L.oracle_get_boundaries.argtypes = [P] + [P] * 3
        L.oracle_get_mortars.argtypes = [P] + [P] * 3
        L.oracle_get_mortar_l2.argtypes = [P] + [P] * 4
        L.oracle_get_basis.argtypes = [P] + [P] * 6
        L.oracle_set_initial_condition.argtypes = [P, I, P, I, P]
        L.oracle_compute_coefficients.argtypes = [P, D, P]
        L.oracle_rhs.argtypes = [P, P, P, D]
        L.oracle_rhs_ld.argtypes = [P, P, P, D]
        L.oracle_indicator_hg.argtypes = [P, P, P]
        L.oracle_set_alpha.argtypes = [P, P]
        L.oracle_max_dt.restype = D
        L.oracle_max_dt.argtypes = [P, P]
        L.oracle_error_norms.argtypes = [P, P, D, P, P]
        L.oracle_integrals.argtypes = [P, P, P, P]
        L.oracle_rk_step.argtypes = [P, P, P, P, D, D]
        L.oracle_solve.restype = I
        L.oracle_solve.argtypes = [P, P, D, D, D, I, P]
        L.oracle_indicator_max.argtypes = [P, P, I, P]
        L.oracle_indicator_hg_params.argtypes = [P, P, D, D, I, I, P]
        L.oracle_amr_adapt.restype = I
        L.oracle_amr_adapt.argtypes = [P, P, P, I, I]
        L.oracle_get_adapted_u.argtypes = [P, P]
        L.oracle_get_transfer.argtypes = [P, P, P]
        L.oracle_get_levels.argtypes = [P, P]
        L.oracle_set_threads.argtypes = [I]
        L.oracle_get_threads.restype = I
        L.oracle_ck2n54_coefficients.argtypes = [P, P, P]
        L.oracle_ln_mean.restype = D
        L.oracle_ln_mean.argtypes = [D, D]
        L.oracle_inv_ln_mean.restype = D
        L.oracle_inv_ln_mean.argtypes = [D, D]
        L.oracle_flux.argtypes = [I, I, D, P, P, I, P]
        L.oracle_slip_wall_flux.argtypes = [I, D, P, I, I, P]
        L.oracle_cons2entropy.argtypes = [I, D, P, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """One oracle discretisation: Oracle(ndims=3, polydeg=3, level=2, cmin=..., cmax=...,
    volume_integral='fd', volume_flux='ranocha', surface_flux='llf', fv_flux='llf',
    alpha_max=0.5, alpha_min=0.001, alpha_smooth=True, source='none'|'convergence_test',
    gamma=1.4, periodicity=True, refinement_patches=()). A refinement patch is a
    box (cmin, cmax) whose leaf cells are refined once (2:1 balanced, L2
    mortars at the level changes). Non-periodic boundaries carry
    BoundaryConditionDirichlet(initial_condition) with the initial condition set
    by set_initial_condition (boundary_condition='dirichlet'), or
    boundary_condition_slip_wall (boundary_condition='slip_wall')."""

    def __init__(self, ndims, polydeg, level, cmin, cmax, volume_integral="fd", volume_flux="ranocha",
                 surface_flux="llf", fv_flux="llf", alpha_max=0.5, alpha_min=0.001, alpha_smooth=True,
                 source="none", gamma=1.4, periodicity=True, refinement_patches=(), hg_variable="density_pressure",
                 boundary_condition="dirichlet"):
        L = _load()
        c = _Config()
        c.ndims, c.polydeg, c.initial_refinement_level = ndims, polydeg, level
        for d in range(ndims):
            c.coordinates_min[d] = cmin[d]
            c.coordinates_max[d] = cmax[d]
        c.gamma = gamma
        c.volume_integral = VI[volume_integral]
        c.volume_flux = FLUX[volume_flux]
        c.surface_flux = FLUX[surface_flux]
        c.fv_flux = FLUX[fv_flux]
        c.alpha_max, c.alpha_min, c.alpha_smooth = alpha_max, alpha_min, int(alpha_smooth)
        c.source_terms = 1 if source == "convergence_test" else 0
        per = (periodicity,) * ndims if isinstance(periodicity, bool) else tuple(periodicity)
        for d in range(ndims):
            c.periodicity[d] = int(bool(per[d]))
        c.boundary_condition = 0 if all(per) else {"dirichlet": 1, "slip_wall": 2}[boundary_condition]
        if len(refinement_patches) > 8:
            raise ValueError("at most 8 refinement patches")
        c.npatches = len(refinement_patches)
        for k, (pmin, pmax) in enumerate(refinement_patches):
            for d in range(ndims):
                c.patch_min[k][d] = pmin[d]
                c.patch_max[k][d] = pmax[d]
        c.hg_variable = HG_VARIABLE[hg_variable]
        self._cfg = c
        self._c = L.oracle_create(ctypes.byref(c))
        if not self._c:
            raise ValueError("oracle_create rejected the configuration")
        self.L = L
        self.ndims, self.polydeg = ndims, polydeg
        self.n = polydeg + 1
        self.nvars = ndims + 2
        self._sizes()

    def _sizes(self):
        L = self.L
        self.nelements = L.oracle_nelements(self._c)
        self.ninterfaces = L.oracle_ninterfaces(self._c)
        self.nboundaries = L.oracle_nboundaries(self._c)
        self.nmortars = L.oracle_nmortars(self._c)
        self.ndofs = L.oracle_ndofs(self._c)
        self.nvalues = self.ndofs * self.nvars

    def __del__(self):
        try:
            if self._c:
                self.L.oracle_destroy(self._c)
                self._c = None
        except Exception:
            pass

    def mesh(self):
        E, I, D, NN = self.nelements, self.ninterfaces, self.ndims, self.n ** self.ndims
        nid = np.empty(2 * I, np.int32)
        ori = np.empty(I, np.int32)
        ij = np.empty(E)
        cen = np.empty(D * E)
        xc = np.empty(D * NN * E)
        en = np.empty(2 * D * E, np.int32)
        self.L.oracle_get_mesh(self._c, _p(nid), _p(ori), _p(ij), _p(cen), _p(xc), _p(en))
        return dict(neighbor_ids=nid.reshape(I, 2), orientations=ori, inverse_jacobian=ij,
                    centers=cen.reshape(E, D), node_coordinates=xc.reshape(E, NN, D),
                    element_neighbors=en.reshape(E, 2 * D))

    def mortars(self):
        """(neighbor_ids [M, 2^(D-1) + 1], large_sides [M], orientations [M]) of the
        mortars container, 0-based ids/orientations."""
        M, ns = self.nmortars, 1 << (self.ndims - 1)
        ids, side, ori = np.empty(M * (ns + 1), np.int32), np.empty(M, np.int32), np.empty(M, np.int32)
        self.L.oracle_get_mortars(self._c, _p(ids), _p(side), _p(ori))
        return ids.reshape(M, ns + 1), side, ori

    def mortar_l2(self):
        """forward_upper, forward_lower, reverse_upper, reverse_lower (n x n)."""
        n = self.n
        m = [np.empty((n, n)) for _ in range(4)]
        self.L.oracle_get_mortar_l2(self._c, *[_p(a) for a in m])
        return m

    def boundaries(self):
        """(neighbor_ids, orientations, neighbor_sides) of the boundaries container, 0-based."""
        B = self.nboundaries
        el, ori, side = np.empty(B, np.int32), np.empty(B, np.int32), np.empty(B, np.int32)
        self.L.oracle_get_boundaries(self._c, _p(el), _p(ori), _p(side))
        return el, ori, side

    def basis(self):
        n = self.n
        out = [np.empty(n), np.empty(n), np.empty(n * n), np.empty(n * n), np.empty(n * n), np.empty(n * n)]
        self.L.oracle_get_basis(self._c, *[_p(a) for a in out])
        return dict(nodes=out[0], weights=out[1], D=out[2].reshape(n, n), Dsplit=out[3].reshape(n, n),
                    Dhat=out[4].reshape(n, n), Vinv=out[5].reshape(n, n))

    def set_initial_condition(self, kind, params=(), modes=None):
        pa = np.array(list(params) + [0.0] * 8, dtype=np.float64)[:8]
        mo = np.ascontiguousarray(modes, dtype=np.float64) if modes is not None else np.zeros(128)
        self._ic_keep = (pa, mo)
        self.L.oracle_set_initial_condition(self._c, IC[kind], _p(pa), len(params), _p(mo))

    def compute_coefficients(self, t=0.0):
        u = np.empty(self.nvalues)
        self.L.oracle_compute_coefficients(self._c, float(t), _p(u))
        return u

    def rhs(self, u, t=0.0):
        u = np.ascontiguousarray(u, dtype=np.float64)
        du = np.empty_like(u)
        self.L.oracle_rhs(self._c, _p(du), _p(u), float(t))
        return du

    def rhs_extended(self, u, t=0.0):
        """The RHS evaluated in 80-bit long double on the same FP64 inputs."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        du = np.empty_like(u)
        self.L.oracle_rhs_ld(self._c, _p(du), _p(u), float(t))
        return du

    def indicator_hg(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        a = np.empty(self.nelements)
        self.L.oracle_indicator_hg(self._c, _p(u), _p(a))
        return a

    def set_alpha(self, alpha):
        """Blending factors [E] (after smoothing) for the shock-capturing RHS
        instead of its HG indicator; None returns to the HG indicator."""
        if alpha is None:
            self.L.oracle_set_alpha(self._c, None)
        else:
            self._alpha_given = np.ascontiguousarray(alpha, dtype=np.float64)
            assert self._alpha_given.shape == (self.nelements,)
            self.L.oracle_set_alpha(self._c, _p(self._alpha_given))

    def max_dt(self, u, cfl=1.0):
        u = np.ascontiguousarray(u, dtype=np.float64)
        return cfl * self.L.oracle_max_dt(self._c, _p(u))

    def error_norms(self, u, t):
        u = np.ascontiguousarray(u, dtype=np.float64)
        l2, li = np.empty(self.nvars), np.empty(self.nvars)
        self.L.oracle_error_norms(self._c, _p(u), float(t), _p(l2), _p(li))
        return l2, li

    def integrals(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        cons = np.empty(self.nvars)
        ent = np.empty(1)
        self.L.oracle_integrals(self._c, _p(u), _p(cons), _p(ent))
        return cons, float(ent[0])

    def rk_step(self, u, t, dt):
        du = np.empty_like(u)
        ut = np.empty_like(u)
        self.L.oracle_rk_step(self._c, _p(u), _p(du), _p(ut), float(t), float(dt))
        return u

    def solve(self, u, t0, t_end, cfl, maxiters=0):
        u = np.array(u, dtype=np.float64, copy=True)
        tout = np.empty(1)
        nsteps = self.L.oracle_solve(self._c, _p(u), float(t0), float(t_end), float(cfl), int(maxiters), _p(tout))
        if nsteps < 0:
            raise FloatingPointError("time step size `dt` is NaN")
        return u, float(tout[0]), nsteps

    def levels(self):
        lv = np.empty(self.nelements, np.int32)
        self.L.oracle_get_levels(self._c, _p(lv))
        return lv

    def indicator_hg_params(self, u, alpha_max, alpha_min, alpha_smooth, variable="density_pressure"):
        """IndicatorHennemannGassner with its own parameters (an AMR indicator)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        a = np.empty(self.nelements)
        var = HG_VARIABLE[variable]
        self.L.oracle_indicator_hg_params(self._c, _p(u), float(alpha_max), float(alpha_min), int(alpha_smooth), var,
                                          _p(a))
        return a

    def indicator_max(self, u, variable=0):
        """IndicatorMax(semi; variable): per-element maximum of a conserved variable."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        a = np.empty(self.nelements)
        self.L.oracle_indicator_max(self._c, _p(u), int(variable), _p(a))
        return a

    def amr_adapt(self, u, lam, only_refine=False, only_coarsen=False):
        """Adapt the tree with controller values lam (>0 refine, <0 coarsen);
        returns (changed, projected u, transfer kind, transfer src)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        lam = np.ascontiguousarray(lam, dtype=np.int32)
        rc = self.L.oracle_amr_adapt(self._c, _p(u), _p(lam), int(only_refine), int(only_coarsen))
        if rc < 0:
            raise ValueError("oracle_amr_adapt: the adapted tree is not 2:1 balanced")
        self._sizes()
        un = np.empty(self.nvalues)
        kind, src = np.empty(self.nelements, np.int32), np.empty(self.nelements, np.int32)
        self.L.oracle_get_adapted_u(self._c, _p(un))
        self.L.oracle_get_transfer(self._c, _p(kind), _p(src))
        return bool(rc), un, kind, src

    def solve_amr(self, t0, t_end, cfl, controller, interval=5, adapt_initial_condition=True,
                  adapt_initial_condition_only_refine=True, variable=0, maxiters=0, indicator="max"):
        """solve(ode, CarpenterKennedy2N54; callback = (AMRCallback(controller;
        interval, adapt_initial_condition), StepsizeCallback(cfl))) on the
        initial condition set by set_initial_condition. The AMR callback
        adapts the initial condition until the mesh settles (amr.jl:143-180),
        then after every `interval`-th step except the last one (t_end or
        maxiters reached, amr.jl:52-58);
        the step size is computed after it (callback order of the elixirs).
        indicator: "max" (IndicatorMax of conserved variable `variable`), "hg"
        (IndicatorHennemannGassner with the configured alpha_max/min/smooth), or
        a tuple (alpha_max, alpha_min, alpha_smooth, variable) for an HG
        indicator with its own parameters.
        Returns (u, t, nsteps, history of element counts)."""
        history = []

        def adapt(u, **kw):
            if indicator == "max":
                alpha = self.indicator_max(u, variable)
            elif indicator == "hg":
                alpha = self.indicator_hg(u)
            else:  # (alpha_max, alpha_min, alpha_smooth, variable) of an AMR-only HG indicator
                alpha = self.indicator_hg_params(u, *indicator)
            if getattr(controller, "combined", False):  # the shock-capturing indicator as the secondary one
                lam = controller(alpha, self.levels(), self.indicator_hg(u))
            else:
                lam = controller(alpha, self.levels())
            changed, un, _, _ = self.amr_adapt(u, lam, **kw)
            history.append(self.nelements)
            return changed, un

        u = self.compute_coefficients(t0)
        if adapt_initial_condition:
            changed, u = adapt(u, only_refine=adapt_initial_condition_only_refine)
            iterations = 1
            while changed:
                u = self.compute_coefficients(t0)
                changed, u = adapt(u, only_refine=adapt_initial_condition_only_refine)
                iterations += 1
                if iterations > max(10, controller.max_level):
                    break
        t = float(t0)
        dt = self.max_dt(u, cfl)
        it = 0
        eps = float(np.sqrt(np.finfo(np.float64).eps))
        while True:
            if np.isnan(dt):
                raise FloatingPointError("time step size `dt` is NaN")
            te = t + dt
            final = te > t_end or te == t_end or abs(te - t_end) <= eps * max(abs(te), abs(t_end))
            if final:
                dt = t_end - t
            self.rk_step(u, t, dt)
            it += 1
            t += dt
            # isfinished (callbacks_step.jl:25-31): t_end reached or maxiters steps taken
            last = final or bool(maxiters and it >= maxiters)
            if interval > 0 and it % interval == 0 and not last:
                _, u = adapt(u)
            dt = self.max_dt(u, cfl)
            if last:
                break
        return u, t, it, history

    @staticmethod
    def set_threads(n):
        _load().oracle_set_threads(int(n))

    @staticmethod
    def get_threads():
        return _load().oracle_get_threads()


def ln_mean(x, y):
    return _load().oracle_ln_mean(float(x), float(y))


def inv_ln_mean(x, y):
    return _load().oracle_inv_ln_mean(float(x), float(y))


def flux(ndims, kind, ul, ur, orientation, gamma=1.4):
    """kind: 'llf' | 'llf_naive' | 'ranocha' | 'central' | 'hllc' | 'physical' (orientation 0-based)."""
    ul = np.ascontiguousarray(ul, dtype=np.float64)
    ur = np.ascontiguousarray(ur if ur is not None else ul, dtype=np.float64)
    f = np.empty(ndims + 2)
    k = -1 if kind == "physical" else FLUX[kind]
    _load().oracle_flux(ndims, k, gamma, _p(ul), _p(ur), int(orientation), _p(f))
    return f


def slip_wall_flux(ndims, u, orientation, plus, gamma=1.4):
    """boundary_condition_slip_wall flux of the inner state u at a face of the
    given orientation (0-based); plus: the element is left of the wall."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    f = np.empty(ndims + 2)
    _load().oracle_slip_wall_flux(ndims, gamma, _p(u), int(orientation), int(bool(plus)), _p(f))
    return f


def cons2entropy(ndims, u, gamma=1.4):
    u = np.ascontiguousarray(u, dtype=np.float64)
    w = np.empty(ndims + 2)
    _load().oracle_cons2entropy(ndims, gamma, _p(u), _p(w))
    return w


def ck2n54():
    a, b, c = np.empty(5), np.empty(5), np.empty(5)
    _load().oracle_ck2n54_coefficients(_p(a), _p(b), _p(c))
    return a, b, c


class ControllerThreeLevel:
    """ControllerThreeLevel(semi, indicator; base_level, med_level, med_threshold,
    max_level, max_threshold), amr.jl:1116-1153: target level max_level above
    max_threshold, med_level above med_threshold (the current level if
    med_level <= 0), base_level otherwise; +1 / -1 / 0 as the element level is
    below / above / at the target."""

    def __init__(self, base_level=1, med_level=None, med_threshold=0.0, max_level=None, max_threshold=1.0):
        self.base_level = int(base_level)
        self.med_level = int(base_level if med_level is None else med_level)
        self.max_level = int(base_level if max_level is None else max_level)
        self.med_threshold = float(med_threshold)
        self.max_threshold = float(max_threshold)

    def __call__(self, alpha, levels):
        alpha = np.asarray(alpha, dtype=np.float64)
        levels = np.asarray(levels)
        target = np.where(alpha > self.max_threshold, self.max_level,
                          np.where(alpha > self.med_threshold,
                                   self.med_level if self.med_level > 0 else levels, self.base_level))
        return np.sign(target - levels).astype(np.int32)


class ControllerThreeLevelCombined(ControllerThreeLevel):
    """ControllerThreeLevelCombined(semi, indicator_primary, indicator_secondary;
    ..., max_threshold_secondary), amr.jl:1254-1297: the three-level target of
    the primary indicator, raised to max_level where the secondary indicator is
    at least max_threshold_secondary."""
    combined = True

    def __init__(self, base_level=1, med_level=None, med_threshold=0.0, max_level=None, max_threshold=1.0,
                 max_threshold_secondary=1.0):
        super().__init__(base_level, med_level, med_threshold, max_level, max_threshold)
        self.max_threshold_secondary = float(max_threshold_secondary)

    def __call__(self, alpha, levels, alpha_secondary):
        alpha = np.asarray(alpha, dtype=np.float64)
        levels = np.asarray(levels)
        target = np.where(alpha > self.max_threshold, self.max_level,
                          np.where(alpha > self.med_threshold,
                                   self.med_level if self.med_level > 0 else levels, self.base_level))
        target = np.where(np.asarray(alpha_secondary) >= self.max_threshold_secondary, self.max_level, target)
        return np.sign(target - levels).astype(np.int32)
