"""SemidiscretizationHyperbolic, semidiscretize and rhs_hyperbolic_ (the drop-in).

Mirrors src/semidiscretization/semidiscretization_hyperbolic.jl:14-76, 578-597
and src/semidiscretization/semidiscretization.jl:102-153, 224-242. The
right-hand side is evaluated by the HIP library through the C ABI; there is no
CPU path.
"""
import ctypes

import numpy as np

from . import _lib
from .equations import BoundaryConditionDirichlet, boundary_condition_periodic, boundary_condition_slip_wall
from .mesh import element_geometry
from .solver import (_HG_VARIABLES, VolumeIntegralFluxDifferencing, VolumeIntegralPureLGLFiniteVolume,
                     VolumeIntegralShockCapturingHG, VolumeIntegralWeakForm)


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class SemidiscretizationHyperbolic:
    """`SemidiscretizationHyperbolic(mesh, equations, initial_condition, solver;
    source_terms, boundary_conditions)` with a HIP backend.

    precision: "fp64" (plain FP64) or "fp32" (FP32 fluxes and 2N register
    u_tmp; the state u, the stage arithmetic and reductions in FP64). elements_per_block / threads_per_block override the
    element-kernel tile (0 = default). boundary_conditions: periodic (for a
    periodic mesh) or, for every boundary of a non-periodic mesh, one
    BoundaryConditionDirichlet(initial_condition) or boundary_condition_slip_wall."""

    def __init__(self, mesh, equations, initial_condition, solver, source_terms=None,
                 boundary_conditions=boundary_condition_periodic, precision="fp64", elements_per_block=0,
                 threads_per_block=0):
        periodic_bc = boundary_conditions in ("periodic", None) or boundary_conditions is boundary_condition_periodic
        if mesh.nboundaries == 0 and not periodic_bc:
            raise ValueError("a periodic mesh takes boundary_condition_periodic")
        if mesh.nboundaries > 0:
            if boundary_conditions is boundary_condition_slip_wall:
                pass
            elif not isinstance(boundary_conditions, BoundaryConditionDirichlet):
                raise ValueError("a non-periodic mesh needs BoundaryConditionDirichlet(initial_condition) or "
                                 "boundary_condition_slip_wall")
            else:
                f = boundary_conditions.boundary_value_function
                if (f.code, tuple(f.params)) != (initial_condition.code, tuple(initial_condition.params)):
                    raise ValueError("the Dirichlet boundary value function must be the initial condition "
                                     "(the device evaluates one function for both)")
        self.boundary_conditions = boundary_conditions
        if mesh.ndims != equations.ndims:
            raise ValueError("mesh and equations dimension differ")
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.mesh = mesh
        self.equations = equations
        self.initial_condition = initial_condition
        self.solver = solver
        self.source_terms = source_terms
        self.precision = precision
        self.tile = (int(elements_per_block), int(threads_per_block))
        self._set_geometry()
        self._ctx = None
        self._lib = _lib.load()
        self.performance_counter = {"ncalls": 0, "time": 0.0}

    def _set_geometry(self):
        # inverse_jacobian[E] (init_elements!); the node coordinates [E, NN, D]
        # only serve host-side evaluations of the initial condition and are
        # computed when first asked for
        b = self.solver.basis
        self.inverse_jacobian = 1.0 / (self.mesh.cell_lengths / float(np.sum(b.weights)))  # as element_geometry
        self._node_coordinates = None

    @property
    def node_coordinates(self):
        if self._node_coordinates is None:
            inv_j, self._node_coordinates = element_geometry(self.mesh, self.solver.basis)
            assert np.array_equal(inv_j, self.inverse_jacobian)
        return self._node_coordinates

    # -- sizes ---------------------------------------------------------------
    def nvariables(self):
        return self.equations.nvariables()

    def nnodes(self):
        return self.solver.nnodes()

    def nelements(self):
        return self.mesh.nelements

    def ndofs(self):
        """`ndofs(mesh, dg, cache)` = nelements * nnodes^ndims (dg.jl:969-971)."""
        return self.mesh.nelements * self.nnodes() ** self.mesh.ndims

    def nvalues(self):
        return self.ndofs() * self.nvariables()

    # -- C ABI configuration ---------------------------------------------------
    def config(self):
        vi = self.solver.volume_integral
        cfg = _lib.Config()
        cfg.ndims = self.mesh.ndims
        cfg.polydeg = self.solver.polydeg()
        cfg.nelements = self.mesh.nelements
        cfg.ninterfaces = self.mesh.ninterfaces
        cfg.gamma = self.equations.gamma
        cfg.precision = _lib.FP32 if self.precision == "fp32" else _lib.FP64
        cfg.volume_integral = vi.code
        cfg.volume_flux = _lib.FLUX_RANOCHA
        cfg.fv_flux = _lib.FLUX_LLF
        cfg.alpha_max, cfg.alpha_min, cfg.alpha_smooth = 0.5, 0.001, 1
        if isinstance(vi, VolumeIntegralFluxDifferencing):
            cfg.volume_flux = vi.volume_flux.code
        elif isinstance(vi, VolumeIntegralShockCapturingHG):
            cfg.volume_flux = vi.volume_flux_dg.code
            cfg.fv_flux = vi.volume_flux_fv.code
            ind = vi.indicator
            cfg.alpha_max, cfg.alpha_min, cfg.alpha_smooth = ind.alpha_max, ind.alpha_min, int(ind.alpha_smooth)
            cfg.hg_variable = _HG_VARIABLES[ind.variable]
        elif isinstance(vi, VolumeIntegralPureLGLFiniteVolume):
            cfg.fv_flux = vi.volume_flux_fv.code
        elif not isinstance(vi, VolumeIntegralWeakForm):
            raise ValueError(f"unsupported volume integral {vi}")
        cfg.surface_flux = self.solver.surface_flux.code
        cfg.source_terms = self.source_terms.code if self.source_terms is not None else _lib.SRC_NONE
        cfg.elements_per_block, cfg.threads_per_block = self.tile
        cfg.nboundaries = self.mesh.nboundaries
        cfg.boundary_condition = self.boundary_conditions.code if self.mesh.nboundaries else _lib.BC_NONE
        cfg.nmortars = self.mesh.nmortars
        return cfg

    @property
    def ctx(self):
        """The HIP context (created on first use; needs a GPU)."""
        if self._ctx is None:
            lib = _lib.load()
            b = self.solver.basis
            arrays = dict(
                nbr=np.ascontiguousarray(self.mesh.neighbor_ids, dtype=np.int32),
                ori=np.ascontiguousarray(self.mesh.orientations, dtype=np.int32),
                invj=np.ascontiguousarray(self.inverse_jacobian, dtype=np.float64),
                cen=np.ascontiguousarray(self.mesh.cell_centers, dtype=np.float64),
                bel=np.ascontiguousarray(self.mesh.boundary_neighbor_ids, dtype=np.int32),
                bor=np.ascontiguousarray(self.mesh.boundary_orientations, dtype=np.int32),
                bsd=np.ascontiguousarray(self.mesh.boundary_neighbor_sides, dtype=np.int32),
                mid=np.ascontiguousarray(self.mesh.mortar_neighbor_ids, dtype=np.int32),
                msd=np.ascontiguousarray(self.mesh.mortar_large_sides, dtype=np.int32),
                mor=np.ascontiguousarray(self.mesh.mortar_orientations, dtype=np.int32),
                mfu=np.ascontiguousarray(b.mortar_forward_upper), mfl=np.ascontiguousarray(b.mortar_forward_lower),
                mru=np.ascontiguousarray(b.mortar_reverse_upper), mrl=np.ascontiguousarray(b.mortar_reverse_lower),
                nodes=np.ascontiguousarray(b.nodes), weights=np.ascontiguousarray(b.weights),
                ds=np.ascontiguousarray(b.derivative_split), dh=np.ascontiguousarray(b.derivative_hat),
                vinv=np.ascontiguousarray(b.inverse_vandermonde_legendre),
                aw=np.ascontiguousarray(b.analysis_weights), av=np.ascontiguousarray(b.analysis_vandermonde))
            i32 = ctypes.POINTER(ctypes.c_int32)
            mesh = _lib.Mesh(arrays["nbr"].ctypes.data_as(i32), arrays["ori"].ctypes.data_as(i32),
                             _ptr(arrays["invj"]), _ptr(arrays["cen"]), arrays["bel"].ctypes.data_as(i32),
                             arrays["bor"].ctypes.data_as(i32), arrays["bsd"].ctypes.data_as(i32),
                             arrays["mid"].ctypes.data_as(i32), arrays["msd"].ctypes.data_as(i32),
                             arrays["mor"].ctypes.data_as(i32))
            basis = _lib.Basis(_ptr(arrays["nodes"]), _ptr(arrays["weights"]), _ptr(arrays["ds"]),
                               _ptr(arrays["dh"]), _ptr(arrays["vinv"]), len(b.analysis_nodes),
                               _ptr(arrays["aw"]), _ptr(arrays["av"]), _ptr(arrays["mfu"]), _ptr(arrays["mfl"]),
                               _ptr(arrays["mru"]), _ptr(arrays["mrl"]))
            cfg = self.config()
            out = ctypes.c_void_p()
            rc = lib.trixi_hip_create(ctypes.byref(cfg), ctypes.byref(mesh), ctypes.byref(basis),
                                      ctypes.byref(out))
            _lib.check(rc, None, "trixi_hip_create")
            self._ctx = out
            ic = self.initial_condition
            params = np.array(list(ic.params) + [0.0] * 4, dtype=np.float64)[:4]
            modes = np.ascontiguousarray(ic.modes if ic.modes is not None else np.zeros((4, 8, 4)),
                                         dtype=np.float64)
            _lib.check(lib.trixi_hip_set_exact_solution(out, ic.code, _ptr(params), len(ic.params),
                                                        _ptr(modes)), out, "set_exact_solution")
        return self._ctx

    def remesh(self, kind, src):
        """Carry the resident state over to the adapted mesh: a context is
        created on self.mesh (already adapted) and trixi_hip_transfer projects
        the old context's state into it (refine!/coarsen! of the DG solver,
        amr_dg3d.jl:138-430); the old context is destroyed."""
        old = self._ctx
        self._set_geometry()
        self._ctx = None
        new = self.ctx
        if old is None:
            return
        try:
            k = np.ascontiguousarray(kind, dtype=np.int32)
            s = np.ascontiguousarray(src, dtype=np.int32)
            if k.shape != (self.mesh.nelements,) or s.shape != k.shape:
                raise ValueError("one transfer entry per element of the adapted mesh")
            _lib.check(self._lib.trixi_hip_transfer(new, old, k.ctypes.data_as(ctypes.c_void_p),
                                                    s.ctypes.data_as(ctypes.c_void_p)), new, "trixi_hip_transfer")
        finally:
            self._lib.trixi_hip_destroy(old)

    def tile_used(self):
        """(elements per workgroup, threads) of the element kernel's RHS variant."""
        e, t = ctypes.c_int(), ctypes.c_int()
        _lib.check(self._lib.trixi_hip_get_tile(self.ctx, ctypes.byref(e), ctypes.byref(t)), self.ctx, "get_tile")
        return e.value, t.value

    def stage_tile_used(self):
        """The same for the fused 2N-stage variant (None when the stage runs unfused)."""
        e, t = ctypes.c_int(), ctypes.c_int()
        _lib.check(self._lib.trixi_hip_get_stage_tile(self.ctx, ctypes.byref(e), ctypes.byref(t)), self.ctx,
                   "get_stage_tile")
        return (e.value, t.value) if e.value > 0 else None

    def close(self):
        if self._ctx is not None:
            self._lib.trixi_hip_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- setup -----------------------------------------------------------------
    def compute_coefficients(self, t):
        """`compute_coefficients(t, semi)` (semidiscretization.jl:224-242): the
        initial condition at every node, as the flat u_ode vector [V, NN, E]."""
        u = self.initial_condition(self.node_coordinates, t, self.equations)  # [E, NN, V]
        return np.ascontiguousarray(u, dtype=np.float64).reshape(-1)

    def set_initial_state(self, t):
        """The resident state := compute_coefficients(t): evaluated on the device
        (trixi_hip_compute_coefficients, as the reference's GPU backend does,
        dg.jl:1293-1324) when the device knows the initial condition, else on
        the host and uploaded."""
        ctx = self.ctx
        if getattr(self.initial_condition, "device_exact", False):
            _lib.check(self._lib.trixi_hip_compute_coefficients(ctx, float(t)), ctx, "compute_coefficients")
        else:
            u0 = self.compute_coefficients(t)
            _lib.check(self._lib.trixi_hip_upload_u(ctx, u0.ctypes.data_as(ctypes.c_void_p)), ctx, "upload_u")

    def __repr__(self):
        return (f"SemidiscretizationHyperbolic({self.mesh}, {self.equations}, {self.initial_condition}, "
                f"{self.solver}, precision={self.precision})")


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def rhs_hyperbolic_(du_ode, u_ode, semi, t):
    """`rhs_hyperbolic!(du_ode, u_ode, semi, t)` (semidiscretization_hyperbolic.jl:578-597).

    u_ode / du_ode: flat float64 arrays of length nvariables*ndofs, either numpy
    (host memory) or torch CUDA tensors (device memory, torch's current stream)."""
    ctx = semi.ctx
    lib = semi._lib
    n = semi.nvalues()
    if _is_torch(u_ode):
        import torch
        if not (u_ode.is_cuda and du_ode.is_cuda and u_ode.dtype == torch.float64 and du_ode.dtype == torch.float64):
            raise TypeError("device arrays must be float64 CUDA tensors")
        if u_ode.numel() != n or du_ode.numel() != n or not (u_ode.is_contiguous() and du_ode.is_contiguous()):
            raise ValueError("u_ode/du_ode must be contiguous with nvariables*ndofs entries")
        _lib.check(lib.trixi_hip_set_stream(ctx, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
                   ctx, "set_stream")
        rc = lib.trixi_hip_rhs(ctx, ctypes.c_void_p(du_ode.data_ptr()), ctypes.c_void_p(u_ode.data_ptr()),
                               float(t), _lib.DEVICE)
        _lib.check(rc, ctx, "trixi_hip_rhs")
    else:
        if u_ode.dtype != np.float64 or du_ode.dtype != np.float64 or u_ode.size != n or du_ode.size != n:
            raise ValueError("u_ode/du_ode must be float64 with nvariables*ndofs entries")
        if not (u_ode.flags.c_contiguous and du_ode.flags.c_contiguous):
            raise ValueError("u_ode/du_ode must be contiguous")
        _lib.check(lib.trixi_hip_set_stream(ctx, None), ctx, "set_stream")
        rc = lib.trixi_hip_rhs(ctx, du_ode.ctypes.data_as(ctypes.c_void_p), u_ode.ctypes.data_as(ctypes.c_void_p),
                               float(t), _lib.HOST)
        _lib.check(rc, ctx, "trixi_hip_rhs")
    semi.performance_counter["ncalls"] += 1
    return None


class ODEProblem:
    """`ODEProblem(rhs_hyperbolic!, u0_ode, tspan, semi)` (semidiscretization.jl:151)."""

    def __init__(self, f, u0, tspan, p):
        self.f = f
        self.u0 = u0
        self.tspan = (float(tspan[0]), float(tspan[1]))
        self.p = p


def semidiscretize(semi, tspan, restart_file=None):
    """`semidiscretize(semi, tspan[, restart_file])` (semidiscretization.jl:102-153,
    156-184): u0 from the initial condition, or read from a restart file."""
    if restart_file is not None:
        from .io import load_restart_file
        return ODEProblem(rhs_hyperbolic_, load_restart_file(semi, restart_file), tspan, semi)
    u0_ode = semi.compute_coefficients(tspan[0])
    return ODEProblem(rhs_hyperbolic_, u0_ode, tspan, semi)
