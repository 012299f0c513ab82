"""Solution, restart and mesh files (HDF5): the attributes, datatypes, shapes
and values the reference writes (save_solution_dg.jl, save_restart_dg.jl,
mesh_io.jl), read back with h5dump (an independent HDF5 reader) and with the
library's own readers. CPU only: the files are written from host arrays."""
import os
import re
import subprocess

import numpy as np
import pytest

import trixi_hip as th

H5DUMP = "/opt/conda/bin/h5dump"
pytestmark = pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not in this image")


def dump(path, obj, attr=False):
    """Values of a dataset or attribute as h5dump prints them (17 digits)."""
    out = subprocess.run([H5DUMP, "-m", "%.17g", "-a" if attr else "-d", obj, path], capture_output=True, text=True,
                         check=True).stdout
    body = out[out.index("DATA {") + 6:]
    body = body[:body.index("}")]
    body = re.sub(r"\([0-9,]+\):", " ", body)
    vals = [v.strip() for v in body.replace("\n", " ").split(",") if v.strip()]
    return [v.strip('"') if v.startswith('"') else int(v, 16) if v.startswith("0x") else float(v) for v in vals]


def header(path):
    """{name: (datatype summary, dataspace)} of the attributes and datasets (h5dump -H)."""
    out = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True).stdout
    items = {}
    for m in re.finditer(r'(ATTRIBUTE|DATASET) "([^"]+)" \{\s*DATATYPE\s+(H5T_STRING \{[^}]*\}|\S+)\s*DATASPACE\s+'
                         r'(SCALAR|SIMPLE \{[^}]*\})', out):
        items[m.group(2)] = (m.group(3), m.group(4))
    return items


def semi_for(mesh, N=3):
    eq = th.CompressibleEulerEquations3D(1.4) if mesh.ndims == 3 else th.CompressibleEulerEquations2D(1.4)
    ic = th.initial_condition_convergence_test
    bc = th.BoundaryConditionDirichlet(ic) if mesh.nboundaries else th.boundary_condition_periodic
    return th.SemidiscretizationHyperbolic(mesh, eq, ic, th.DGSEM(polydeg=N, surface_flux=th.flux_lax_friedrichs),
                                           boundary_conditions=bc)


@pytest.mark.parametrize("D", [2, 3])
def test_solution_file_layout_and_values(tmp_path, D):
    mesh = th.TreeMesh((0,) * D, (2,) * D, initial_refinement_level=1, periodicity=True,
                       refinement_patches=[((0.2,) * D, (0.8,) * D)])
    semi = semi_for(mesh)
    mesh.current_filename = th.save_mesh_file(mesh, str(tmp_path))
    u = semi.compute_coefficients(0.3) * (1 + 0.01 * np.random.default_rng(D).standard_normal(semi.nvalues()))
    alpha = np.linspace(0, 1, semi.nelements())
    f = th.save_solution_file(u, 0.3, 0.0125, 42, semi, str(tmp_path), element_variables={"indicator_amr": alpha})
    assert os.path.basename(f) == "solution_000000042.h5"
    h = header(f)
    V = D + 2
    for name in ("ndims", "polydeg", "n_vars", "n_elements", "timestep"):
        assert h[name] == ("H5T_STD_I64LE", "SCALAR")
    for name in ("time", "dt"):
        assert h[name] == ("H5T_IEEE_F64LE", "SCALAR")
    for name, size in (("equations", 28), ("mesh_type", 8), ("mesh_file", 7)):
        assert h[name][0].startswith("H5T_STRING") and f"STRSIZE {size};" in h[name][0]
        assert "H5T_CSET_UTF8" in h[name][0] and h[name][1] == "SCALAR"
    assert dump(f, "/equations", True) == [f"CompressibleEulerEquations{D}D"]
    assert dump(f, "/mesh_type", True) == ["TreeMesh"] and dump(f, "/mesh_file", True) == ["mesh.h5"]
    assert dump(f, "/n_vars", True) == [V] and dump(f, "/n_elements", True) == [semi.nelements()]
    assert dump(f, "/time", True) == [0.3] and dump(f, "/dt", True) == [0.0125] and dump(f, "/timestep", True) == [42]
    prim = semi.equations.cons2prim(u.reshape(-1, V))
    names = ["rho"] + [f"v{d + 1}" for d in range(D)] + ["p"]
    n = semi.nvalues() // V
    for v in range(V):
        assert h[f"variables_{v + 1}"] == ("H5T_IEEE_F64LE", f"SIMPLE {{ ( {n} ) / ( {n} ) }}")
        assert dump(f, f"/variables_{v + 1}/name", True) == [names[v]]
        assert np.array_equal(np.array(dump(f, f"/variables_{v + 1}")), prim[:, v])
    assert dump(f, "/element_variables_1/name", True) == ["indicator_amr"]
    assert np.array_equal(np.array(dump(f, "/element_variables_1")), alpha)


def test_restart_round_trip_and_mismatches(tmp_path):
    mesh = th.TreeMesh((0, 0, 0), (2, 2, 2), initial_refinement_level=1, periodicity=True)
    semi = semi_for(mesh)
    mesh.current_filename = th.save_mesh_file(mesh, str(tmp_path))
    u = np.random.default_rng(1).uniform(0.5, 2.0, semi.nvalues())
    f = th.save_restart_file(u, 0.75, 1e-3, 12, semi, str(tmp_path))
    assert os.path.basename(f) == "restart_000000012.h5"
    assert [dump(f, f"/variables_{v + 1}/name", True)[0] for v in range(5)] == \
        ["rho", "rho_v1", "rho_v2", "rho_v3", "rho_e_total"]
    assert np.array_equal(th.load_restart_file(semi, f), u)
    assert (th.load_time(f), th.load_dt(f), th.load_timestep(f)) == (0.75, 1e-3, 12)
    assert np.array_equal(th.semidiscretize(semi, (0.75, 1.0), restart_file=f).u0, u)
    with pytest.raises(ValueError, match="polynomial degree"):
        th.load_restart_file(semi_for(mesh, N=2), f)
    with pytest.raises(ValueError, match="number of elements"):
        th.load_restart_file(semi_for(th.TreeMesh((0, 0, 0), (2, 2, 2), initial_refinement_level=2)), f)
    with pytest.raises(ValueError, match="ndims"):
        th.load_restart_file(semi_for(th.TreeMesh((0, 0), (2, 2), initial_refinement_level=1)), f)


@pytest.mark.parametrize("periodic", [True, False])
def test_mesh_file_of_a_one_level_tree(tmp_path, periodic):
    """The 2D tree after one refinement: root 1 with children 2..5 in child
    order (x bit, then y bit); same-level face neighbours, wrapping through the
    periodic root (its own neighbour), none across a non-periodic boundary."""
    mesh = th.TreeMesh((0, 0), (2, 2), initial_refinement_level=1, periodicity=periodic)
    f = th.save_mesh_file(mesh, str(tmp_path))
    h = header(f)
    assert h["periodicity"] == ("H5T_STD_B8LE", "SIMPLE { ( 2 ) / ( 2 ) }")
    assert h["child_ids"] == ("H5T_STD_I64LE", "SIMPLE { ( 5, 4 ) / ( 5, 4 ) }")
    assert h["neighbor_ids"][1] == "SIMPLE { ( 5, 4 ) / ( 5, 4 ) }"
    assert h["coordinates"] == ("H5T_IEEE_F64LE", "SIMPLE { ( 5, 2 ) / ( 5, 2 ) }")
    assert dump(f, "/n_cells", True) == [5] and dump(f, "/n_leaf_cells", True) == [4]
    assert dump(f, "/parent_ids") == [0, 1, 1, 1, 1]
    assert dump(f, "/child_ids") == [2, 3, 4, 5] + [0] * 16
    assert dump(f, "/levels") == [0, 1, 1, 1, 1]
    assert dump(f, "/coordinates") == [1, 1, 0.5, 0.5, 1.5, 0.5, 0.5, 1.5, 1.5, 1.5]
    p = 1 if periodic else 0
    # directions -x, +x, -y, +y
    expect = [p, p, p, p,
              3 if periodic else 0, 3, 4 if periodic else 0, 4,
              2, 2 if periodic else 0, 5 if periodic else 0, 5,
              5 if periodic else 0, 5, 2, 2 if periodic else 0,
              4, 4 if periodic else 0, 3, 3 if periodic else 0]
    assert dump(f, "/neighbor_ids") == expect
    assert dump(f, "/periodicity", True) == [p, p]


@pytest.mark.parametrize("D,periodic", [(2, True), (2, False), (3, True), (3, (True, False, True))])
def test_mesh_file_round_trip_after_adaptation(tmp_path, D, periodic):
    """A patch-refined, then randomly adapted tree: the written tree is a valid
    pre-order tree (parent/child links, symmetric same-level neighbours, leaf
    centres = element centres), and loading it gives back the same elements,
    interfaces, mortars and boundaries."""
    mesh = th.TreeMesh((0,) * D, (2,) * D, initial_refinement_level=2, periodicity=periodic,
                       refinement_patches=[((0.5,) * D, (1.5,) * D)])
    rng = np.random.default_rng(3 + D)
    for _ in range(3):
        lam = rng.integers(-1, 2, mesh.nelements).astype(np.int32)
        lam[mesh.levels >= 4] = np.minimum(lam[mesh.levels >= 4], 0)
        mesh.adapt(lam)
    parent, child, nbr, levels, coords = mesh.tree_arrays()
    n = len(parent)
    ids = np.arange(1, n + 1)
    inner = child[:, 0] > 0
    assert np.all((child[inner] > ids[inner, None]))  # children after their parent (pre-order)
    for c in ids[inner]:
        assert np.all(parent[child[c - 1] - 1] == c) and np.all(levels[child[c - 1] - 1] == levels[c - 1] + 1)
    for dirn in range(2 * D):
        has = nbr[:, dirn] > 0
        back = nbr[nbr[has, dirn] - 1, dirn ^ 1]
        assert np.array_equal(back, ids[has]) and np.all(levels[nbr[has, dirn] - 1] == levels[has])
    leaf = ~inner
    assert np.array_equal(coords[leaf], mesh.cell_centers) and leaf.sum() == mesh.nelements
    f = th.save_mesh_file(mesh, str(tmp_path), timestep=7)
    assert os.path.basename(f) == "mesh_000000007.h5"
    m2 = th.load_mesh_file(f)
    for name in ("cell_centers", "levels", "neighbor_ids", "orientations", "mortar_neighbor_ids",
                 "mortar_large_sides", "mortar_orientations", "boundary_neighbor_ids", "boundary_orientations",
                 "boundary_neighbor_sides"):
        assert np.array_equal(getattr(m2, name), getattr(mesh, name)), name
    assert m2.periodicity == mesh.periodicity and not m2.unsaved_changes


def test_load_mesh_follows_the_restart_files_mesh_attribute(tmp_path):
    mesh = th.TreeMesh((-1, -1), (1, 1), initial_refinement_level=2, periodicity=False,
                       refinement_patches=[((-0.5, -0.5), (0.5, 0.5))])
    semi = semi_for(mesh)
    mesh.current_filename = th.save_mesh_file(mesh, str(tmp_path), timestep=3)
    f = th.save_restart_file(np.zeros(semi.nvalues()), 0.1, 0.01, 3, semi, str(tmp_path))
    assert dump(f, "/mesh_file", True) == ["mesh_000000003.h5"]
    m2 = th.load_mesh(f)
    assert np.array_equal(m2.cell_centers, mesh.cell_centers) and m2.nboundaries == mesh.nboundaries


def test_restart_callback_initialize_writes_the_mesh_file(tmp_path):
    """SaveRestartCallback's initialize! (save_restart.jl:59-72) writes mesh.h5, so
    the restart files of a run whose only writer is this callback name mesh.h5."""
    mesh = th.TreeMesh((-1, -1), (1, 1), initial_refinement_level=2, periodicity=True)
    semi = semi_for(mesh)
    cb = th.SaveRestartCallback(interval=5, output_directory=str(tmp_path / "out"))
    cb.initialize(semi, 0.0, 0.01, None)
    assert os.path.exists(tmp_path / "out" / "mesh.h5") and not mesh.unsaved_changes
    f = th.save_restart_file(np.zeros(semi.nvalues()), 0.1, 0.01, 5, semi, cb.output_directory)
    assert dump(f, "/mesh_file", True) == ["mesh.h5"]


def test_io_library_exports_every_declared_symbol():
    import ctypes
    from conftest import ROOT
    from trixi_hip import io
    src = open(os.path.join(ROOT, "include", "trixi_hip_io.h")).read()
    syms = sorted(set(re.findall(r"\b(trixi_io_[a-z_]+)\s*\(", src)))
    assert len(syms) == 7
    lib = ctypes.CDLL(io.IO_LIB_PATH)
    for s in syms:
        assert hasattr(lib, s), s


def test_io_errors_are_reported(tmp_path):
    from trixi_hip import io
    with pytest.raises(io.TrixiIOError, match="cannot open"):
        th.load_time(str(tmp_path / "missing.h5"))
    mesh = th.TreeMesh((0, 0), (2, 2), initial_refinement_level=1, periodicity=True)
    f = th.save_mesh_file(mesh, str(tmp_path))
    with pytest.raises(io.TrixiIOError, match="lacks a solution/restart attribute"):
        th.load_time(f)  # a mesh file is not a restart file
