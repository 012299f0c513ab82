"""Host-side mirror (numpy) against the oracle: TreeMesh element order and
geometry, interfaces, initial conditions; and the C ABI of the HIP library
(loads and exports every declared symbol; argument validation needs no GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT
import trixi_hip as th
from trixi_hip import _lib
from trixi_hip.mesh import element_geometry


@pytest.mark.parametrize("D,N,level,cmin,cmax,per", [
    (3, 3, 2, (0, 0, 0), (2, 2, 2), True), (2, 3, 4, (-10, -10), (10, 10), True),
    (3, 4, 1, (-1, -1, -1), (1, 1, 1), True), (2, 2, 3, (-1, -1), (1, 1), True),
    (2, 3, 3, (0, 0), (2, 2), False), (3, 3, 2, (0, 0, 0), (2, 2, 2), False),
    (3, 2, 2, (0, 0, 0), (2, 2, 2), (True, False, True)), (2, 3, 0, (0, 0), (1, 1), False)])
def test_mesh_and_geometry_match_oracle(D, N, level, cmin, cmax, per):
    from oracle import Oracle
    o = Oracle(D, N, level, cmin, cmax, periodicity=per)
    m = o.mesh()
    mesh = th.TreeMesh(cmin, cmax, initial_refinement_level=level, periodicity=per)
    basis = th.LobattoLegendreBasis(N)
    ij, xc = element_geometry(mesh, basis)
    assert mesh.nelements == o.nelements and mesh.ninterfaces == o.ninterfaces
    assert mesh.nboundaries == o.nboundaries
    assert 2 * mesh.ninterfaces + mesh.nboundaries == 2 * D * mesh.nelements
    b_el, b_ori, b_side = o.boundaries()
    np.testing.assert_array_equal(mesh.boundary_neighbor_ids, b_el)
    np.testing.assert_array_equal(mesh.boundary_orientations, b_ori)
    np.testing.assert_array_equal(mesh.boundary_neighbor_sides, b_side)
    np.testing.assert_array_equal(mesh.cell_centers, m["centers"])
    np.testing.assert_array_equal(mesh.neighbor_ids, m["neighbor_ids"])
    np.testing.assert_array_equal(mesh.orientations, m["orientations"])
    np.testing.assert_array_equal(mesh.element_neighbors, m["element_neighbors"])
    np.testing.assert_array_equal(ij, m["inverse_jacobian"])
    np.testing.assert_array_equal(xc, m["node_coordinates"])


def test_tree_order_follows_child_sign_table():
    """Element 1 (0-based) of a level-1 tree is child 2 of abstract_tree.jl:118-128: (+, -, -)."""
    mesh = th.TreeMesh((0, 0, 0), (2, 2, 2), initial_refinement_level=1)
    np.testing.assert_array_equal(mesh.cell_centers[:4], [[0.5, 0.5, 0.5], [1.5, 0.5, 0.5], [0.5, 1.5, 0.5],
                                                          [1.5, 1.5, 0.5]])
    # periodic neighbours: +x neighbour centre is +dx modulo the domain
    mesh = th.TreeMesh((0, 0, 0), (2, 2, 2), initial_refinement_level=3, periodicity=True)
    c = mesh.cell_centers
    for d in range(3):
        nb = mesh.element_neighbors[:, 2 * d + 1]
        expect = c.copy()
        expect[:, d] = (expect[:, d] + mesh.cell_length) % 2.0
        np.testing.assert_allclose(c[nb], expect)


@pytest.mark.parametrize("name,D,ic,params", [
    ("constant", 3, th.initial_condition_constant, ()),
    ("convergence_test", 3, th.initial_condition_convergence_test, ()),
    ("convergence_test", 2, th.initial_condition_convergence_test, ()),
    ("weak_blast_wave", 3, th.initial_condition_weak_blast_wave, ()),
    ("weak_blast_wave", 2, th.initial_condition_weak_blast_wave, ()),
    ("isentropic_vortex", 2, th.initial_condition_isentropic_vortex(25.0), (25.0,)),
    ("kelvin_helmholtz", 2, th.initial_condition_kelvin_helmholtz_instability(), ()),
    ("fourier_modes", 3, th.initial_condition_fourier_modes(2), (0.1, 0.2)),
    ("density_pulse", 3, th.initial_condition_density_pulse, ()),
    ("density_pulse", 2, th.initial_condition_density_pulse, ()),
    ("blast_wave", 2, th.initial_condition_blast_wave, ()),
    ("blob", 2, th.initial_condition_blob, ()),
])
def test_initial_conditions_match_oracle(name, D, ic, params):
    from oracle import Oracle
    cmin, cmax = ((-2,) * D, (2,) * D)
    o = Oracle(D, 3, 2, cmin, cmax)
    o.set_initial_condition(name, params, ic.modes.reshape(-1) if ic.modes is not None else None)
    eq = th.CompressibleEulerEquations3D() if D == 3 else th.CompressibleEulerEquations2D()
    mesh = th.TreeMesh(cmin, cmax, initial_refinement_level=2)
    _, xc = element_geometry(mesh, th.LobattoLegendreBasis(3))
    u = ic(xc, 0.3, eq).reshape(-1)
    np.testing.assert_allclose(u, o.compute_coefficients(0.3), rtol=1e-13, atol=1e-14)


def test_khi_noise_is_seeded_and_shared_on_faces():
    eq = th.CompressibleEulerEquations2D()
    mesh = th.TreeMesh((-1, -1), (1, 1), initial_refinement_level=3)
    _, xc = element_geometry(mesh, th.LobattoLegendreBasis(3))
    a = th.initial_condition_kelvin_helmholtz_instability(1e-3, 5)(xc, 0, eq)
    b = th.initial_condition_kelvin_helmholtz_instability(1e-3, 5)(xc, 0, eq)
    c = th.initial_condition_kelvin_helmholtz_instability(1e-3, 6)(xc, 0, eq)
    np.testing.assert_array_equal(a, b)
    assert np.max(np.abs(a - c)) > 1e-5
    # element 0's +x face equals element (+x neighbour)'s -x face
    nb = mesh.element_neighbors[0, 1]
    n = 4
    right = a[0].reshape(n, n, 4)[:, n - 1]
    left = a[nb].reshape(n, n, 4)[:, 0]
    np.testing.assert_array_equal(right, left)


def header_symbols():
    src = open(os.path.join(ROOT, "include", "trixi_hip.h")).read()
    return sorted(set(re.findall(r"\b(trixi_hip_[a-z_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s
    # the Python binding declares every entry point of the header
    assert sorted(name for name, _, _ in _lib.SIGNATURES) == syms


def test_create_rejects_bad_configuration_without_gpu():
    """Argument validation happens before any device call."""
    lib = _lib.load()
    cfg = _lib.Config()
    cfg.ndims, cfg.polydeg, cfg.nelements, cfg.ninterfaces = 4, 3, 8, 32
    out = ctypes.c_void_p()
    rc = lib.trixi_hip_create(ctypes.byref(cfg), ctypes.byref(_lib.Mesh()), ctypes.byref(_lib.Basis()),
                              ctypes.byref(out))
    assert rc == _lib.EINVAL and not out.value
    cfg.ndims, cfg.ninterfaces = 3, 5  # not a conforming periodic tree (I != D E)
    rc = lib.trixi_hip_create(ctypes.byref(cfg), ctypes.byref(_lib.Mesh()), ctypes.byref(_lib.Basis()),
                              ctypes.byref(out))
    assert rc == _lib.EINVAL
    assert lib.trixi_hip_create(None, None, None, None) == _lib.EINVAL


def test_hllc_only_with_pure_finite_volume(capfd):
    """flux_hllc is compiled into the pure finite-volume kernels only: the Python
    solver and trixi_hip_create both refuse it with any other volume integral
    (before any device call), and the pure-FV configuration carries its code."""
    eq = th.CompressibleEulerEquations3D(1.4)
    basis = th.LobattoLegendreBasis(3)
    with pytest.raises(ValueError, match="flux_hllc"):
        th.DGSEM(basis, th.flux_hllc, th.VolumeIntegralFluxDifferencing(th.flux_ranocha))
    ind = th.IndicatorHennemannGassner(eq, basis, alpha_max=0.5, alpha_min=0.001, alpha_smooth=True,
                                       variable=th.density_pressure)
    with pytest.raises(ValueError, match="flux_hllc"):
        th.DGSEM(basis, th.flux_lax_friedrichs,
                 th.VolumeIntegralShockCapturingHG(ind, volume_flux_dg=th.flux_ranocha, volume_flux_fv=th.flux_hllc))
    solver = th.DGSEM(basis, th.flux_hllc, th.VolumeIntegralPureLGLFiniteVolume(th.flux_hllc))
    mesh = th.TreeMesh((0, 0, 0), (2, 2, 2), initial_refinement_level=1, periodicity=True)
    semi = th.SemidiscretizationHyperbolic(mesh, eq, th.initial_condition_convergence_test, solver)
    c = semi.config()
    assert c.volume_integral == _lib.VI_PURE_FV and c.surface_flux == c.fv_flux == _lib.FLUX_HLLC
    lib = _lib.load()
    cfg = _lib.Config()
    cfg.ndims, cfg.polydeg, cfg.nelements, cfg.ninterfaces = 3, 3, 8, 24
    cfg.volume_integral, cfg.volume_flux, cfg.surface_flux, cfg.fv_flux = \
        _lib.VI_FLUX_DIFFERENCING, _lib.FLUX_RANOCHA, _lib.FLUX_HLLC, _lib.FLUX_LLF
    out = ctypes.c_void_p()
    rc = lib.trixi_hip_create(ctypes.byref(cfg), ctypes.byref(_lib.Mesh()), ctypes.byref(_lib.Basis()),
                              ctypes.byref(out))
    assert rc == _lib.EINVAL and not out.value
    assert "flux_hllc" in capfd.readouterr().err  # create reports its refusal on stderr


def test_host_api_config_mapping():
    eq = th.CompressibleEulerEquations2D(1.4)
    basis = th.LobattoLegendreBasis(3)
    ind = th.IndicatorHennemannGassner(eq, basis, alpha_max=0.002, alpha_min=1e-4, alpha_smooth=True,
                                       variable=th.density_pressure)
    vi = th.VolumeIntegralShockCapturingHG(ind, volume_flux_dg=th.flux_ranocha,
                                           volume_flux_fv=th.FluxLaxFriedrichs(th.max_abs_speed_naive))
    solver = th.DGSEM(basis, th.FluxLaxFriedrichs(th.max_abs_speed_naive), vi)
    mesh = th.TreeMesh((-1, -1), (1, 1), initial_refinement_level=3, periodicity=True)
    semi = th.SemidiscretizationHyperbolic(mesh, eq, th.initial_condition_kelvin_helmholtz_instability(), solver,
                                           precision="fp32")
    c = semi.config()
    assert (c.ndims, c.polydeg, c.nelements, c.ninterfaces) == (2, 3, 64, 128)
    assert c.volume_integral == _lib.VI_SHOCK_CAPTURING_HG and c.volume_flux == _lib.FLUX_RANOCHA
    assert c.fv_flux == _lib.FLUX_LLF_NAIVE and c.surface_flux == _lib.FLUX_LLF_NAIVE
    assert c.precision == _lib.FP32 and c.alpha_max == 0.002 and c.alpha_min == 1e-4 and c.alpha_smooth == 1
    assert semi.ndofs() == 64 * 16
    with pytest.raises(ValueError):
        th.TreeMesh((0, 0), (1, 2), initial_refinement_level=1)
    assert c.nboundaries == 0 and c.boundary_condition == _lib.BC_NONE


def test_boundaries_of_a_bounded_mesh():
    """TreeMesh defaults to periodicity=false (tree_mesh.jl:108-130): the domain
    faces become boundaries (init_boundaries!, containers_2d.jl:475-543), which
    need BoundaryConditionDirichlet(initial_condition)."""
    mesh = th.TreeMesh((0, 0), (2, 2), initial_refinement_level=2)
    assert mesh.periodicity == (False, False)
    assert (mesh.nelements, mesh.ninterfaces, mesh.nboundaries) == (16, 24, 16)
    # direction-major: 4 boundaries per direction; side 2 on -faces, 1 on +faces
    np.testing.assert_array_equal(mesh.boundary_orientations, [0] * 8 + [1] * 8)
    np.testing.assert_array_equal(mesh.boundary_neighbor_sides, [2] * 4 + [1] * 4 + [2] * 4 + [1] * 4)
    xc = mesh.cell_centers
    assert np.all(xc[mesh.boundary_neighbor_ids[:4], 0] == 0.25) and np.all(xc[mesh.boundary_neighbor_ids[4:8], 0] == 1.75)
    eq = th.CompressibleEulerEquations2D(1.4)
    solver = th.DGSEM(polydeg=3, surface_flux=th.flux_lax_friedrichs)
    ic = th.initial_condition_convergence_test
    with pytest.raises(ValueError):
        th.SemidiscretizationHyperbolic(mesh, eq, ic, solver)
    with pytest.raises(ValueError):  # the boundary function must be the initial condition
        th.SemidiscretizationHyperbolic(mesh, eq, ic, solver,
                                        boundary_conditions=th.BoundaryConditionDirichlet(th.initial_condition_constant))
    semi = th.SemidiscretizationHyperbolic(mesh, eq, ic, solver, boundary_conditions=th.BoundaryConditionDirichlet(ic))
    c = semi.config()
    assert (c.ninterfaces, c.nboundaries, c.boundary_condition) == (24, 16, _lib.BC_DIRICHLET)
    with pytest.raises(ValueError):
        th.SemidiscretizationHyperbolic(th.TreeMesh((0, 0), (2, 2), initial_refinement_level=2, periodicity=True),
                                        eq, ic, solver, boundary_conditions=th.BoundaryConditionDirichlet(ic))


@pytest.mark.parametrize("D,N,level,cmin,cmax,per,patches", [
    (3, 3, 2, (0, 0, 0), (2, 2, 2), True, [((0.5, 0.5, 0.5), (1.5, 1.5, 1.5))]),
    (2, 3, 3, (-1, -1), (1, 1), True, [((-0.3, -0.3), (0.3, 0.3)), ((-0.1, -0.1), (0.1, 0.1))]),
    (2, 2, 2, (0, 0), (2, 2), False, [((0.0, 0.0), (1.0, 2.0))]),
    (3, 2, 1, (-1, -1, -1), (1, 1, 1), (True, False, True), [((-1, -1, -1), (0, 0, 0))]),
])
def test_refined_mesh_matches_oracle(D, N, level, cmin, cmax, per, patches):
    """TreeMesh with box refinement patches (refine_box!, rebalance!): the same
    depth-first elements, geometry, interfaces, L2 mortars and boundaries as the
    oracle; every element face is covered exactly once."""
    from oracle import Oracle
    o = Oracle(D, N, level, cmin, cmax, periodicity=per, refinement_patches=patches)
    m = o.mesh()
    mesh = th.TreeMesh(cmin, cmax, initial_refinement_level=level, periodicity=per, refinement_patches=patches)
    assert (mesh.nelements, mesh.ninterfaces, mesh.nmortars, mesh.nboundaries) == \
        (o.nelements, o.ninterfaces, o.nmortars, o.nboundaries)
    assert mesh.nmortars > 0
    np.testing.assert_array_equal(mesh.cell_centers, m["centers"])
    np.testing.assert_array_equal(mesh.neighbor_ids, m["neighbor_ids"])
    np.testing.assert_array_equal(mesh.orientations, m["orientations"])
    ids, side, ori = o.mortars()
    np.testing.assert_array_equal(mesh.mortar_neighbor_ids, ids)
    np.testing.assert_array_equal(mesh.mortar_large_sides, side)
    np.testing.assert_array_equal(mesh.mortar_orientations, ori)
    b_el, b_ori, b_side = o.boundaries()
    np.testing.assert_array_equal(mesh.boundary_neighbor_ids, b_el)
    ij, xc = element_geometry(mesh, th.LobattoLegendreBasis(N))
    np.testing.assert_array_equal(ij, m["inverse_jacobian"])
    np.testing.assert_array_equal(xc, m["node_coordinates"])
    ns = 1 << (D - 1)
    faces = 2 * mesh.ninterfaces + mesh.nboundaries + mesh.nmortars * (ns + 1)
    assert faces == 2 * D * mesh.nelements
    assert set(np.diff(np.sort(mesh.levels))) <= {0, 1}


def _compare_mesh_with_oracle(mesh, o, N):
    m = o.mesh()
    assert (mesh.nelements, mesh.ninterfaces, mesh.nmortars, mesh.nboundaries) == \
        (o.nelements, o.ninterfaces, o.nmortars, o.nboundaries)
    np.testing.assert_array_equal(mesh.levels, o.levels())
    np.testing.assert_array_equal(mesh.cell_centers, m["centers"])
    np.testing.assert_array_equal(mesh.neighbor_ids, m["neighbor_ids"])
    np.testing.assert_array_equal(mesh.orientations, m["orientations"])
    ids, side, ori = o.mortars()
    np.testing.assert_array_equal(mesh.mortar_neighbor_ids, ids)
    np.testing.assert_array_equal(mesh.mortar_large_sides, side)
    np.testing.assert_array_equal(mesh.mortar_orientations, ori)
    b_el, b_ori, b_side = o.boundaries()
    np.testing.assert_array_equal(mesh.boundary_neighbor_ids, b_el)
    np.testing.assert_array_equal(mesh.boundary_orientations, b_ori)
    np.testing.assert_array_equal(mesh.boundary_neighbor_sides, b_side)
    ij, xc = element_geometry(mesh, th.LobattoLegendreBasis(N))
    np.testing.assert_array_equal(ij, m["inverse_jacobian"])
    np.testing.assert_array_equal(xc, m["node_coordinates"])


@pytest.mark.parametrize("D,per,patches", [(2, True, ()), (2, False, ()), (3, True, ()), (3, False, ()),
                                           (3, True, [((0.5, 0.5, 0.5), (1.5, 1.5, 1.5))])])
def test_amr_adapt_matches_oracle(D, per, patches):
    """TreeMesh.adapt (refine! with rebalancing, guarded coarsen!) on random
    controller values: after every adaptation the same elements, interfaces,
    mortars and boundaries as the oracle's tree, and the same transfer map."""
    from oracle import Oracle
    N, level = 3, 2
    cmin, cmax = (0.0,) * D, (2.0,) * D
    patches = [(p[0][:D], p[1][:D]) for p in patches]
    o = Oracle(D, N, level, cmin, cmax, periodicity=per, refinement_patches=patches)
    if not per:
        o.set_initial_condition("constant")
    mesh = th.TreeMesh(cmin, cmax, initial_refinement_level=level, periodicity=per, refinement_patches=patches)
    rng = np.random.default_rng(100 + D + 10 * per)
    u = np.zeros(o.nvalues)
    for it in range(10):
        lam = rng.integers(-1, 2, mesh.nelements).astype(np.int32)
        lam[mesh.levels >= 5] = np.minimum(lam[mesh.levels >= 5], 0)
        only_refine = it == 0
        changed, kind, src = mesh.adapt(lam, only_refine=only_refine)
        ochanged, u, okind, osrc = o.amr_adapt(u, lam, only_refine=only_refine)
        assert changed == ochanged
        np.testing.assert_array_equal(kind, okind)
        np.testing.assert_array_equal(src, osrc)
        _compare_mesh_with_oracle(mesh, o, N)


def test_native_tree_refuses_invalid_input():
    """trixi_hip_tree_*: leaves outside their level's range and refinement
    beyond level 19 are refused with an error instead of corrupting the keys."""
    from trixi_hip.mesh import _Leaves, _NativeTree
    one = lambda lv, ic: _Leaves(np.array([lv]), np.array([ic], dtype=np.int64), np.zeros((1, 2)),
                                 np.zeros(1, np.int32), np.zeros(1, np.int32))
    with pytest.raises(ValueError):
        _NativeTree(one(2, (4, 0)), 2, 1.0, (True, True))  # coordinate 4 at level 2
    with pytest.raises(ValueError):
        _NativeTree(one(20, (0, 0)), 2, 1.0, (True, True))
    nt = _NativeTree(one(19, (0, 0)), 2, 1.0, (True, True))
    with pytest.raises(ValueError, match="level 19"):
        nt.refine(np.array([True]))
    nt = _NativeTree(one(0, (0, 0)), 2, 1.0, (False, True))
    assert nt.refine(np.array([True])) == 1
    t = nt.leaves()
    np.testing.assert_array_equal(t.lv, [1, 1, 1, 1])
    np.testing.assert_array_equal(t.ic, [[0, 0], [1, 0], [0, 1], [1, 1]])
    np.testing.assert_array_equal(t.kind, [1, 2, 3, 4])
    np.testing.assert_array_equal(t.src, [0, 0, 0, 0])
    c = nt.connectivity()
    # x bounded, y periodic: two x interfaces, four y interfaces (wrapping), four x boundaries
    assert len(c["orientations"]) == 6 and len(c["boundary_neighbor_ids"]) == 4
    np.testing.assert_array_equal(c["boundary_orientations"], [0, 0, 0, 0])
    np.testing.assert_array_equal(c["boundary_neighbor_ids"], [0, 2, 1, 3])
    np.testing.assert_array_equal(c["boundary_neighbor_sides"], [2, 2, 1, 1])
    assert nt.coarsen(t.lv, t.ic) == 1
    t = nt.leaves()
    np.testing.assert_array_equal(t.lv, [0])
    assert (t.kind[0], t.src[0]) == (-1, 0)


def test_spec_features_closed_forms():
    """spec_features (the configuration's input set, used for the feature
    comparison): a constant state gives rho*p / mean = 1 and energy ratios
    (0, 0, 1); a state whose rho*p is one highest mode plus a mean gives
    (E - E1) / E equal to the mode's energy share, and a mode of degree N - 1
    gives (E1 - E2) / E1 the same share."""
    from trixi_hip.learned_indicator import khi_semidiscretization, spec_features
    semi = khi_semidiscretization(level=1)
    E, n = semi.nelements(), 4
    gamma = semi.equations.gamma
    basis = semi.solver.basis
    x = basis.nodes
    u = np.zeros((E, n * n, 4))
    u[..., 0] = 1.0
    u[..., 3] = 2.0 / (gamma - 1)  # p = 2, v = 0
    f = spec_features(semi, u.reshape(-1))
    np.testing.assert_allclose(f[:, :16], 1.0, rtol=1e-6)
    np.testing.assert_allclose(f[:, 16:19], np.tile([0.0, 0.0, 1.0], (E, 1)), atol=1e-6)
    from numpy.polynomial import legendre
    for deg, col in ((3, 16), (2, 17)):
        # rho = 1, p(x, y) = 2 + 0.5 P_deg(x): modal energies 2^2 (mean) and
        # (0.5 c)^2 with c the normalised Legendre coefficient of the nodal values
        px = 2.0 + 0.5 * legendre.legval(x, [0] * deg + [1])
        pnode = np.tile(px, n)  # node = i + n j, x index fastest
        w = u.copy()
        w[..., 3] = pnode[None, :] / (gamma - 1)
        f = spec_features(semi, w.reshape(-1))
        vinv = np.asarray(basis.inverse_vandermonde_legendre)
        m1 = vinv @ px  # modal coefficients along x
        # rows are constant in y, so only y-mode 0 carries energy: the shares
        # are those of the x transform
        e_mode, e_rest = m1[deg] ** 2, (m1[:deg] ** 2).sum()
        np.testing.assert_allclose(f[:, col], e_mode / (e_mode + e_rest), rtol=1e-5)
