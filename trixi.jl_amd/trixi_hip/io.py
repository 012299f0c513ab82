"""Solution, restart and mesh files (HDF5) in the reference's layout, and the
callbacks that write them.

  save_solution_file / SaveSolutionCallback  src/callbacks_step/save_solution_dg.jl:8-90,
                                             save_solution.jl:127-300
  save_restart_file / load_restart_file /    src/callbacks_step/save_restart_dg.jl:8-80,
  SaveRestartCallback, load_time/dt/timestep save_restart.jl:75-168
  save_mesh_file / load_mesh (TreeMesh)      src/meshes/mesh_io.jl:15-51, 400-420, 603-628

The files are written and read by libtrixi_hip_io.so (C ABI in
include/trixi_hip_io.h, linked against libhdf5); the state is read from the
device context. There is no Python HDF5 path: without the library every call
raises.
"""
import ctypes
import os

import numpy as np

from . import _lib
from .equations import cons2cons, cons2prim, varnames
from .mesh import TreeMesh

IO_LIB_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "libtrixi_hip_io.so")


class Header(ctypes.Structure):
    _fields_ = [("ndims", ctypes.c_int64), ("polydeg", ctypes.c_int64), ("n_vars", ctypes.c_int64),
                ("n_elements", ctypes.c_int64), ("timestep", ctypes.c_int64), ("time", ctypes.c_double),
                ("dt", ctypes.c_double), ("equations", ctypes.c_char * 64), ("mesh_type", ctypes.c_char * 32),
                ("mesh_file", ctypes.c_char * 256)]


class TreeHeader(ctypes.Structure):
    _fields_ = [("ndims", ctypes.c_int64), ("n_cells", ctypes.c_int64), ("n_leaf_cells", ctypes.c_int64),
                ("minimum_level", ctypes.c_int64), ("maximum_level", ctypes.c_int64),
                ("center_level_0", ctypes.c_double * 3), ("length_level_0", ctypes.c_double),
                ("periodicity", ctypes.c_uint8 * 3)]


_io = None
_P = ctypes.c_void_p


def _load():
    global _io
    if _io is None:
        if not os.path.exists(IO_LIB_PATH):
            raise RuntimeError(f"libtrixi_hip_io.so not found at {IO_LIB_PATH}: build it with "
                               "__graft_entry__.build() (needs the HDF5 headers and library)")
        L = ctypes.CDLL(IO_LIB_PATH)
        L.trixi_io_last_error.restype = ctypes.c_char_p
        L.trixi_io_write_solution.argtypes = [ctypes.c_char_p, ctypes.POINTER(Header), ctypes.c_int64, _P, _P,
                                              ctypes.c_int, _P, _P]
        L.trixi_io_read_header.argtypes = [ctypes.c_char_p, ctypes.POINTER(Header)]
        L.trixi_io_read_variables.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _P, _P]
        L.trixi_io_write_tree.argtypes = [ctypes.c_char_p, ctypes.POINTER(TreeHeader), _P, _P, _P, _P, _P]
        L.trixi_io_read_tree_header.argtypes = [ctypes.c_char_p, ctypes.POINTER(TreeHeader)]
        L.trixi_io_read_tree.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int64, _P, _P, _P, _P, _P]
        for f in ("write_solution", "read_header", "read_variables", "write_tree", "read_tree_header", "read_tree"):
            getattr(L, "trixi_io_" + f).restype = ctypes.c_int
        _io = L
    return _io


class TrixiIOError(RuntimeError):
    pass


def _check(rc, what):
    if rc != 0:
        raise TrixiIOError(f"{what}: {_load().trixi_io_last_error().decode()} (code {rc})")


def _strings(names):
    arr = (ctypes.c_char_p * len(names))(*[n.encode() for n in names])
    return arr


def _p(a):
    return a.ctypes.data_as(_P)


def download_state(semi):
    """The resident state of semi's context as the flat u_ode vector."""
    u = np.empty(semi.nvalues())
    _lib.check(semi._lib.trixi_hip_download_u(semi.ctx, u.ctypes.data_as(_P)), semi.ctx, "download_u")
    return u


# ---------------------------------------------------------------------------
# mesh files
# ---------------------------------------------------------------------------
def save_mesh_file(mesh, output_directory, timestep=0):
    """`save_mesh_file(mesh::TreeMesh, output_directory, timestep)`: mesh.h5, or
    mesh_<timestep>.h5 for timestep > 0. Returns the file name."""
    os.makedirs(output_directory, exist_ok=True)
    name = f"mesh_{timestep:09d}.h5" if timestep > 0 else "mesh.h5"
    filename = os.path.join(output_directory, name)
    parent, child, nbr, levels, coords = mesh.tree_arrays()
    leaf = np.all(child == 0, axis=1)
    h = TreeHeader()
    h.ndims = mesh.ndims
    h.n_cells = len(parent)
    h.n_leaf_cells = int(leaf.sum())
    h.minimum_level = int(levels[leaf].min())
    h.maximum_level = int(levels[leaf].max())
    for d in range(mesh.ndims):
        h.center_level_0[d] = mesh.center_level_0[d]
        h.periodicity[d] = int(mesh.periodicity[d])
    h.length_level_0 = mesh.length_level_0
    arrays = [np.ascontiguousarray(a) for a in (parent, child, nbr, levels, coords)]
    _check(_load().trixi_io_write_tree(filename.encode(), ctypes.byref(h), *[_p(a) for a in arrays]),
           "save_mesh_file")
    return filename


def load_mesh_file(mesh_file):
    """`load_mesh_serial` for a TreeMesh (mesh_io.jl:410-420, 603-628)."""
    L = _load()
    h = TreeHeader()
    _check(L.trixi_io_read_tree_header(mesh_file.encode(), ctypes.byref(h)), "load_mesh")
    D, n = int(h.ndims), int(h.n_cells)
    parent = np.empty(n, np.int64)
    child = np.empty((n, 1 << D), np.int64)
    nbr = np.empty((n, 2 * D), np.int64)
    levels = np.empty(n, np.int64)
    coords = np.empty((n, D))
    _check(L.trixi_io_read_tree(mesh_file.encode(), n, D, _p(parent), _p(child), _p(nbr), _p(levels), _p(coords)),
           "load_mesh")
    mesh = TreeMesh.from_tree(tuple(h.center_level_0[:D]), h.length_level_0, tuple(h.periodicity[:D]), levels,
                              coords, child)
    mesh.current_filename = mesh_file
    return mesh


def load_mesh(restart_file):
    """`load_mesh(restart_file)` (mesh_io.jl:400-408): the mesh file named by the
    restart file's `mesh_file` attribute, in the restart file's directory."""
    h = _read_header(restart_file)
    return load_mesh_file(os.path.join(os.path.dirname(restart_file), h.mesh_file.decode()))


def _save_mesh(semi, output_directory, timestep=0):
    """`save_mesh(semi, output_directory, timestep)` (save_solution.jl:188-203)."""
    mesh = semi.mesh
    if mesh.unsaved_changes:
        mesh.current_filename = save_mesh_file(mesh, output_directory, timestep)
        mesh.unsaved_changes = False
    return mesh.current_filename


# ---------------------------------------------------------------------------
# solution and restart files
# ---------------------------------------------------------------------------
def _header(semi, n_vars, t, dt, timestep):
    h = Header()
    h.ndims = semi.mesh.ndims
    h.polydeg = semi.solver.polydeg()
    h.n_vars = n_vars
    h.n_elements = semi.nelements()
    h.timestep = int(timestep)
    h.time = float(t)
    h.dt = float(dt)
    h.equations = f"CompressibleEulerEquations{semi.mesh.ndims}D".encode()
    h.mesh_type = b"TreeMesh"
    h.mesh_file = os.path.basename(semi.mesh.current_filename).encode()
    return h


def _write(filename, h, semi, data, names, element_variables):
    ev = list((element_variables or {}).items())
    vals = np.ascontiguousarray(np.array([np.asarray(v, np.float64) for _, v in ev]).reshape(len(ev), -1)) \
        if ev else np.zeros(1)
    if ev and vals.shape[1] != semi.nelements():
        raise ValueError("element variables need one value per element")
    data = np.ascontiguousarray(data, dtype=np.float64)
    _check(_load().trixi_io_write_solution(filename.encode(), ctypes.byref(h), semi.nnodes() ** semi.mesh.ndims,
                                           _p(data), _strings(names), len(ev), _p(vals),
                                           _strings([k for k, _ in ev]) if ev else None), "write")
    return filename


def save_solution_file(u, t, dt, timestep, semi, output_directory, solution_variables=cons2prim,
                       element_variables=None):
    """`save_solution_file(u, time, dt, timestep, mesh::TreeMesh, equations, dg,
    cache, solution_callback, element_variables)`: solution_<timestep>.h5 with
    the solution variables (cons2prim by default, as SaveSolutionCallback) and
    the given element variables (name -> [nelements] values)."""
    os.makedirs(output_directory, exist_ok=True)
    eq = semi.equations
    V = eq.nvariables()
    data = solution_variables(np.asarray(u, np.float64).reshape(-1, V), eq).reshape(-1)
    names = varnames(solution_variables, eq)
    filename = os.path.join(output_directory, f"solution_{int(timestep):09d}.h5")
    return _write(filename, _header(semi, len(names), t, dt, timestep), semi, data, names, element_variables)


def save_restart_file(u, t, dt, timestep, semi, output_directory):
    """`save_restart_file(u, time, dt, timestep, mesh, equations, dg, cache,
    restart_callback)`: restart_<timestep>.h5 with the conserved variables."""
    os.makedirs(output_directory, exist_ok=True)
    eq = semi.equations
    filename = os.path.join(output_directory, f"restart_{int(timestep):09d}.h5")
    return _write(filename, _header(semi, eq.nvariables(), t, dt, timestep), semi, u, eq.varnames(), None)


def _read_header(filename):
    h = Header()
    _check(_load().trixi_io_read_header(filename.encode(), ctypes.byref(h)), "read header")
    return h


def load_restart_file(semi, restart_file):
    """`load_restart_file(semi, restart_file)` (save_restart_dg.jl:42-80): u_ode of
    the file, after the reference's checks of ndims, equations, polydeg,
    n_elements and the variable names."""
    h = _read_header(restart_file)
    eq = semi.equations
    if h.ndims != semi.mesh.ndims:
        raise ValueError("restart mismatch: ndims differs from value in restart file")
    if h.equations.decode() != f"CompressibleEulerEquations{semi.mesh.ndims}D":
        raise ValueError("restart mismatch: equations differ from value in restart file")
    if h.polydeg != semi.solver.polydeg():
        raise ValueError("restart mismatch: polynomial degree in solver differs from value in restart file")
    if h.n_elements != semi.nelements():
        raise ValueError("restart mismatch: number of elements in solver differs from value in restart file")
    u = np.empty(semi.nvalues())
    names = eq.varnames()
    _check(_load().trixi_io_read_variables(restart_file.encode(), len(names), semi.nnodes() ** semi.mesh.ndims,
                                           semi.nelements(), _strings(names), _p(u)), "load_restart_file")
    return u


def load_time(restart_file):
    return _read_header(restart_file).time


def load_dt(restart_file):
    return _read_header(restart_file).dt


def load_timestep(restart_file):
    return _read_header(restart_file).timestep


def read_solution_header(filename):
    """The file-level attributes of a solution or restart file, as a dict."""
    h = _read_header(filename)
    return dict(ndims=h.ndims, equations=h.equations.decode(), polydeg=h.polydeg, n_vars=h.n_vars,
                n_elements=h.n_elements, mesh_type=h.mesh_type.decode(), mesh_file=h.mesh_file.decode(),
                time=h.time, dt=h.dt, timestep=h.timestep)


# ---------------------------------------------------------------------------
# callbacks
# ---------------------------------------------------------------------------
def _element_variables(semi, callbacks):
    """`get_element_variables!` of the solver and the callbacks: the HG blending
    factor (:indicator_shock_capturing, dg.jl:366-375 / indicators.jl:23) and the
    AMR indicator (:indicator_amr, amr.jl:1047-1059)."""
    from .amr import AMRCallback
    from .solver import VolumeIntegralShockCapturingHG
    ev = {}
    if isinstance(semi.solver.volume_integral, VolumeIntegralShockCapturingHG):
        alpha = np.empty(semi.nelements())
        _lib.check(semi._lib.trixi_hip_indicator(semi.ctx, _lib.IND_HG, alpha.ctypes.data_as(_P)), semi.ctx,
                   "indicator")
        ev["indicator_shock_capturing"] = alpha
    for cb in callbacks:
        if isinstance(cb, AMRCallback):
            ev["indicator_amr"] = cb.controller.indicator(semi)
    return ev


class SaveSolutionCallback:
    """`SaveSolutionCallback(; interval, save_initial_solution, save_final_solution,
    output_directory, solution_variables)`: every interval-th accepted step, at the
    start and at the end, the mesh file (when it changed) and a solution file."""

    def __init__(self, interval=0, save_initial_solution=True, save_final_solution=True, output_directory="out",
                 solution_variables=cons2prim):
        if solution_variables not in (cons2cons, cons2prim):
            raise ValueError("solution_variables must be cons2cons or cons2prim")
        self.interval = int(interval)
        self.save_initial_solution = bool(save_initial_solution)
        self.save_final_solution = bool(save_final_solution)
        self.output_directory = output_directory
        self.solution_variables = solution_variables
        self.files = []

    def condition(self, naccept, final):
        return self.interval > 0 and (naccept % self.interval == 0 or (self.save_final_solution and final))

    def initialize(self, semi, t, dt, callbacks):
        os.makedirs(self.output_directory, exist_ok=True)
        _save_mesh(semi, self.output_directory)
        if self.save_initial_solution:
            self(semi, t, dt, 0, callbacks)

    def __call__(self, semi, t, dt, naccept, callbacks):
        _save_mesh(semi, self.output_directory, naccept)
        ev = _element_variables(semi, callbacks)
        self.files.append(save_solution_file(download_state(semi), t, dt, naccept, semi, self.output_directory,
                                             self.solution_variables, ev))


class SaveRestartCallback:
    """`SaveRestartCallback(; interval, save_final_restart, output_directory)`:
    every interval-th accepted step and at the end, the mesh file (when it
    changed) and a restart file."""

    def __init__(self, interval=0, save_final_restart=True, output_directory="out"):
        self.interval = int(interval)
        self.save_final_restart = bool(save_final_restart)
        self.output_directory = output_directory
        self.files = []

    def condition(self, naccept, final):
        return self.interval > 0 and (naccept % self.interval == 0 or (self.save_final_restart and final))

    def initialize(self, semi, t, dt, callbacks):
        # initialize! (save_restart.jl:59-72): the directory and the mesh file
        os.makedirs(self.output_directory, exist_ok=True)
        _save_mesh(semi, self.output_directory)

    def __call__(self, semi, t, dt, naccept, callbacks):
        _save_mesh(semi, self.output_directory, naccept)
        self.files.append(save_restart_file(download_state(semi), t, dt, naccept, semi, self.output_directory))
