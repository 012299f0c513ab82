"""Size-independent properties of the oracle's rhs_hyperbolic!: free-stream
preservation, discrete conservation, design order of the MMS residual, the
HG indicator range, and the order of the CarpenterKennedy2N54 recurrence."""
import numpy as np
import pytest

from conftest import make_oracle
import oracle.oracle as O


def spec(D, N, level, vi="fd", vflux="ranocha", sflux="llf", **kw):
    cmin, cmax = ((0,) * D, (2,) * D)
    return dict(ndims=D, polydeg=N, level=level, cmin=cmin, cmax=cmax, volume_integral=vi, volume_flux=vflux,
                surface_flux=sflux, **kw)


@pytest.mark.parametrize("D,vi,vflux,sflux", [(3, "fd", "ranocha", "llf"), (3, "weak", "ranocha", "llf_naive"),
                                               (2, "fd", "ranocha", "llf"), (2, "hg", "ranocha", "llf_naive")])
def test_free_stream(D, vi, vflux, sflux):
    o = make_oracle(spec(D, 3, 2, vi, vflux, sflux))
    o.set_initial_condition("constant")
    du = o.rhs(o.compute_coefficients(0.0))
    assert np.max(np.abs(du)) < 1e-12


@pytest.mark.parametrize("D", [2, 3])
def test_conservation(D):
    """sum_e sum_nodes w J du = 0 for every conserved variable (periodic, no sources)."""
    o = make_oracle(spec(D, 3, 2))
    o.set_initial_condition("fourier_modes" if D == 3 else "kelvin_helmholtz", (0.1, 0.2),
                            np.random.default_rng(3).uniform(0.5, 3, 128) if D == 3 else None)
    u = o.compute_coefficients(0.0)
    du = o.rhs(u)
    b = o.basis()
    n = o.n
    w = b["weights"]
    wn = np.einsum("i,j,k->kji", w, w, w).reshape(-1) if D == 3 else np.outer(w, w).reshape(-1)
    ij = o.mesh()["inverse_jacobian"]
    tot = np.einsum("env,n,e->v", du.reshape(o.nelements, n ** D, D + 2), wn, 1.0 / ij ** D)
    scale = np.max(np.abs(du)) * o.nelements
    assert np.max(np.abs(tot)) < 1e-12 * scale


def test_mms_residual_converges_at_design_order():
    """|rhs(u_exact) - du_exact/dt| decays like h^N (here N = 3) on the MMS."""
    errs = []
    for level in (1, 2, 3):
        o = make_oracle(spec(3, 3, level, "fd", "ranocha", "llf", source="convergence_test"))
        o.set_initial_condition("convergence_test")
        t = 0.3
        u = o.compute_coefficients(t)
        du = o.rhs(u, t)
        h = 1e-6
        dudt = (o.compute_coefficients(t + h) - o.compute_coefficients(t - h)) / (2 * h)
        errs.append(np.max(np.abs(du - dudt)))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > 2.5), rates


def test_hg_indicator_range():
    o = make_oracle(dict(ndims=2, polydeg=3, level=4, cmin=(-1, -1), cmax=(1, 1), volume_integral="hg",
                         volume_flux="ranocha", surface_flux="llf_naive", fv_flux="llf_naive", alpha_max=0.002,
                         alpha_min=1e-4))
    o.set_initial_condition("kelvin_helmholtz")
    a = o.indicator_hg(o.compute_coefficients(0.0))
    assert np.all(a >= 0) and np.all(a <= 0.002)
    # smooth elements sit at 1/(1 + exp(s)) = 1e-4 = alpha_min (not below it, so
    # they stay blended); the shear layers reach alpha_max
    assert np.min(a) == pytest.approx(1e-4, rel=1e-9) and np.max(a) == 0.002


def test_given_blending_factors():
    """With the blending factors given (oracle_set_alpha), the shock-capturing
    RHS is the one its own HG indicator gives for those factors, pure flux
    differencing at alpha = 0 and the pure LGL finite volume at alpha = 1."""
    kw = dict(ndims=2, polydeg=3, level=4, cmin=(-1, -1), cmax=(1, 1), volume_flux="ranocha",
              surface_flux="llf_naive", fv_flux="llf_naive")
    o = make_oracle(dict(kw, volume_integral="hg", alpha_max=0.002, alpha_min=1e-4))
    o.set_initial_condition("kelvin_helmholtz", (1e-3, 0))
    u = o.compute_coefficients(0.0)
    du_hg = o.rhs(u)
    o.set_alpha(o.indicator_hg(u))
    assert np.array_equal(o.rhs(u), du_hg)
    o.set_alpha(np.zeros(o.nelements))
    fd = make_oracle(dict(kw, volume_integral="fd"))
    np.testing.assert_allclose(o.rhs(u), fd.rhs(u), rtol=0, atol=1e-13 * np.max(np.abs(du_hg)))
    o.set_alpha(np.ones(o.nelements))
    fv = make_oracle(dict(kw, volume_integral="fv"))
    np.testing.assert_allclose(o.rhs(u), fv.rhs(u), rtol=0, atol=1e-13 * np.max(np.abs(du_hg)))
    o.set_alpha(None)
    assert np.array_equal(o.rhs(u), du_hg)


def test_ck2n54_fourth_order():
    """The 2N recurrence u_tmp = du - a u_tmp; u += b dt u_tmp (methods_2N.jl:153-157)
    with the oracle's coefficients is 4th order on y' = lambda y."""
    a, b, c = O.ck2n54()
    lam = -1.3 + 0.7j

    def integrate(nsteps, T=1.0):
        y, dt = 1.0 + 0j, T / nsteps
        for _ in range(nsteps):
            tmp = 0j
            for s in range(5):
                du = lam * y
                tmp = du - a[s] * tmp
                y = y + b[s] * dt * tmp
        return y

    errs = [abs(integrate(n) - np.exp(lam)) for n in (10, 20, 40)]
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > 3.8), rates
    # consistency: sum of the effective weights equals one (exact for y' = 1)
    assert abs(integrate(1, 0.5) - np.exp(lam * 0.5)) < 1e-2


@pytest.mark.parametrize("D", [2, 3])
def test_integrals_closed_forms(D):
    """integrate_via_indices (normalised): a constant state integrates to itself
    and its entropy_math S = -rho s/(gamma-1), s = ln p - gamma ln rho; the MMS
    density integrates to its mean c = 2 (LGL quadrature is exact for the
    periodic sine up to the quadrature error, < 1e-6 at level 3)."""
    import math
    o = make_oracle(spec(D, 3, 3))
    o.set_initial_condition("constant")
    u = o.compute_coefficients(0.0)
    cons, ent = o.integrals(u)
    u0 = u[:D + 2]
    np.testing.assert_allclose(cons, u0, rtol=1e-12)  # plain summation over all nodes
    p = 0.4 * (u0[-1] - 0.5 * np.sum(u0[1:-1] ** 2) / u0[0])
    s = math.log(p) - 1.4 * math.log(u0[0])
    assert ent == pytest.approx(-s * u0[0] / 0.4, rel=1e-12)
    o.set_initial_condition("convergence_test")
    cons, _ = o.integrals(o.compute_coefficients(0.3))
    assert cons[0] == pytest.approx(2.0, abs=1e-6)


@pytest.mark.parametrize("D", [2, 3])
def test_entropy_gradient_is_entropy_variables(D):
    """entropy_math S(u) and cons2entropy w(u) are tied by w = dS/du; w is pinned
    on its own by Tadmor's entropy-conservation identity of flux_ranocha, so a
    central difference of the integrated S of constant states pins S."""
    from oracle.oracle import cons2entropy
    o = make_oracle(spec(D, 2, 1))
    o.set_initial_condition("constant")
    nn = o.compute_coefficients(0.0).size // (D + 2)
    rng = np.random.default_rng(7)

    def S(state):
        return o.integrals(np.tile(state, nn))[1]

    for _ in range(4):
        rho = rng.uniform(0.5, 2.0)
        v = rng.uniform(-0.5, 0.5, D)
        p = rng.uniform(0.5, 2.0)
        u = np.concatenate([[rho], rho * v, [p / 0.4 + 0.5 * rho * v @ v]])
        grad = np.empty(D + 2)
        for k in range(D + 2):
            h = 1e-5 * max(1.0, abs(u[k]))
            up, um = u.copy(), u.copy()
            up[k] += h
            um[k] -= h
            grad[k] = (S(up) - S(um)) / (2 * h)
        np.testing.assert_allclose(grad, cons2entropy(D, u), rtol=1e-7, atol=1e-8)


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5])
def test_mortar_l2_matrices(N):
    """LobattoLegendreMortarL2 (l2projection.jl): the forward matrices
    interpolate polynomials of degree N exactly onto the two halves, and the
    reverse (L2) projections undo the split: R_u F_u + R_l F_l = I. A reverse
    matrix built from the wrong half or transposed fails the second check."""
    from oracle import Oracle
    o = Oracle(2, N, 1, (0, 0), (2, 2))
    fu, fl, ru, rl = o.mortar_l2()
    b = o.basis()
    x = b["nodes"]
    for k in range(N + 1):
        p = x ** k
        np.testing.assert_allclose(fu @ p, (0.5 * (x + 1)) ** k, atol=1e-13)
        np.testing.assert_allclose(fl @ p, (0.5 * (x - 1)) ** k, atol=1e-13)
    np.testing.assert_allclose(ru @ fu + rl @ fl, np.eye(N + 1), atol=1e-13)
    assert np.abs(ru.T @ fu + rl.T @ fl - np.eye(N + 1)).max() > 1e-3 or N == 1


# ---------------------------------------------------------------------------
# AMR: solution transfer (L2 adaptor) and tree adaptation
# ---------------------------------------------------------------------------
def _poly(x, N, seed):
    """A random polynomial of total degree <= N per variable at points x[..., D]."""
    rng = np.random.default_rng(seed)
    D = x.shape[-1]
    out = np.zeros(x.shape[:-1] + (D + 2,))
    for v in range(D + 2):
        for _ in range(6):
            p = rng.integers(0, N + 1, D)
            while p.sum() > N:
                p[np.argmax(p)] -= 1
            out[..., v] += rng.normal() * np.prod(x ** p, axis=-1)
    return out


def _adapt_oracle(D, level=2, N=3, periodicity=True):
    return make_oracle(spec(D, N, level, "weak", "ranocha", "llf", periodicity=periodicity))


@pytest.mark.parametrize("D", [2, 3])
def test_amr_refine_interpolates_polynomials_exactly(D):
    """refine_element!: the children hold the parent's polynomial, so a field of
    degree <= N stays equal to the polynomial at the children's nodes (pins the
    forward matrices and the child placement: bit d of the child index = upper half in d)."""
    o = _adapt_oracle(D)
    x = o.mesh()["node_coordinates"]
    u = _poly(x, o.polydeg, 1).reshape(-1)
    lam = np.zeros(o.nelements, np.int32)
    lam[np.random.default_rng(2).choice(o.nelements, o.nelements // 3, replace=False)] = 1
    changed, un, kind, src = o.amr_adapt(u, lam)
    assert changed and (kind > 0).any()
    xn = o.mesh()["node_coordinates"]
    np.testing.assert_allclose(un, _poly(xn, o.polydeg, 1).reshape(-1), rtol=0, atol=1e-12)


@pytest.mark.parametrize("D", [2, 3])
def test_amr_coarsen_projects_polynomials_exactly_and_conserves(D):
    """coarsen_elements!: the L2 projection of 2^D children holding one global
    polynomial of degree <= N is that polynomial; for arbitrary data the
    projection conserves the LGL-quadrature integral of every variable."""
    o = _adapt_oracle(D)
    u0 = np.zeros(o.nvalues)
    _, _, _, _ = o.amr_adapt(u0, np.ones(o.nelements, np.int32))  # refine everything once
    x = o.mesh()["node_coordinates"]
    u = _poly(x, o.polydeg, 4).reshape(-1)
    lam = -np.ones(o.nelements, np.int32)
    lam[: 2 << D] = 0  # keep the first two parents refined
    changed, un, kind, _ = o.amr_adapt(u, lam)
    assert changed and (kind < 0).sum() == o.nelements - 2 * (1 << D) and (kind == 0).sum() == 2 * (1 << D)
    np.testing.assert_allclose(un, _poly(o.mesh()["node_coordinates"], o.polydeg, 4).reshape(-1), rtol=0, atol=1e-12)
    # conservation for random data: refine all, coarsen all
    o = _adapt_oracle(D)
    rng = np.random.default_rng(5)
    _, ur, _, _ = o.amr_adapt(np.zeros(o.nvalues), np.ones(o.nelements, np.int32))
    ur = rng.uniform(0.5, 2.0, o.nvalues)
    before, _ = o.integrals(ur)
    changed, uc, _, _ = o.amr_adapt(ur, -np.ones(o.nelements, np.int32))
    after, _ = o.integrals(uc)
    assert changed and o.nelements == 1 << (2 * D)
    np.testing.assert_allclose(after, before, rtol=1e-13)


@pytest.mark.parametrize("D", [2, 3])
def test_amr_refine_then_coarsen_is_identity(D):
    """reverse_{lower,upper} after forward_{lower,upper}, summed over the
    children, is the identity (L2 projection of the parent's own polynomial)."""
    o = _adapt_oracle(D)
    u = np.random.default_rng(7).normal(size=o.nvalues)
    lv0 = o.levels().copy()
    _, ur, _, _ = o.amr_adapt(u, np.ones(o.nelements, np.int32))
    _, uc, kind, src = o.amr_adapt(ur, -np.ones(o.nelements, np.int32))
    np.testing.assert_array_equal(o.levels(), lv0)
    np.testing.assert_array_equal(src, np.arange(o.nelements) << D)
    np.testing.assert_allclose(uc, u, rtol=0, atol=1e-12)


@pytest.mark.parametrize("D,periodic", [(2, True), (2, False), (3, True), (3, False)])
def test_amr_random_adaptations_stay_balanced(D, periodic):
    """Random controller values over many adaptations: the rebalancing and the
    coarsening guard keep the tree 2:1 balanced (the oracle raises otherwise),
    the leaves tile the domain, and a coarsened parent is never next to a cell
    two levels finer."""
    o = _adapt_oracle(D, level=2, periodicity=periodic)
    if not periodic:
        o.set_initial_condition("constant")
    rng = np.random.default_rng(11 + D)
    u = np.zeros(o.nvalues)
    vol0 = 2.0 ** D
    for it in range(12):
        lv = o.levels()
        lam = rng.integers(-1, 2, o.nelements).astype(np.int32)
        lam[lv >= 5] = np.minimum(lam[lv >= 5], 0)
        _, u, kind, src = o.amr_adapt(u, lam)
        lv = o.levels()
        ij = o.mesh()["inverse_jacobian"]
        h = 2.0 / ij  # LGL reference length 2
        assert np.isclose(np.sum(h ** D), vol0)
        # transfer consistency: copies keep the level, children are one finer, parents one coarser
        nb = o.mesh()["element_neighbors"]
        ok = nb >= 0
        e_idx = np.repeat(np.arange(o.nelements)[:, None], 2 * D, axis=1)
        assert np.all(np.abs(lv[e_idx[ok]] - lv[nb[ok]]) == 0)  # same-level interfaces
        ids, _, _ = o.mortars()
        if len(ids):
            assert np.all(lv[ids[:, :-1]] == lv[ids[:, -1:]] + 1)


@pytest.mark.parametrize("D", [2, 3])
def test_slip_walls_conserve_mass_and_energy(D):
    """A weak blast wave reflected by boundary_condition_slip_wall on every side:
    the wall flux has no mass or energy component, so the totals stay fixed to
    round-off through the reflection. With Dirichlet(initial condition)
    boundaries the same run exchanges mass with the outside."""
    def run(bc):
        o = make_oracle(dict(ndims=D, polydeg=3, level=3 if D == 2 else 2, cmin=(-1,) * D, cmax=(1,) * D,
                             volume_integral="fd", volume_flux="ranocha", surface_flux="ranocha",
                             periodicity=False, boundary_condition=bc))
        o.set_initial_condition("weak_blast_wave")
        u0 = o.compute_coefficients(0.0)
        u, t, _ = o.solve(u0, 0.0, 0.8, 0.5)
        return o.integrals(u0)[0], o.integrals(u)[0], u
    m0, m1, u = run("slip_wall")
    assert np.isfinite(u).all()
    assert abs(m1[0] - m0[0]) < 1e-13 * abs(m0[0])
    assert abs(m1[D + 1] - m0[D + 1]) < 1e-13 * abs(m0[D + 1])
    d0, d1, _ = run("dirichlet")
    assert abs(d1[0] - d0[0]) > 1e-6 * abs(d0[0])
