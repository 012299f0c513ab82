"""The learned indicator's FP64 reference (oracle/mlp_ref.py) pinned on the CPU:
analytic backward against central differences, BF16 rounding against exact
cases, features against closed forms, layout constants against the C ABI."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT

from oracle import mlp_ref


def _params(seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, mlp_ref.NP)
    return p


def test_layout_constants_match_header():
    h = open(os.path.join(ROOT, "include", "trixi_hip.h")).read()
    assert int(re.search(r"TRIXI_HIP_MLP_NPARAMS (\d+)", h).group(1)) == mlp_ref.NP == 5505
    assert int(re.search(r"TRIXI_HIP_MLP_NFEATURES (\d+)", h).group(1)) == mlp_ref.NF
    assert int(re.search(r"TRIXI_HIP_MLP_FEATURE_STRIDE (\d+)", h).group(1)) == mlp_ref.KF


def test_backward_matches_central_differences():
    rng = np.random.default_rng(1)
    p = _params(2)
    x = rng.uniform(-1, 1, (37, mlp_ref.KF))
    t = rng.uniform(0, 1, 37)
    _, g = mlp_ref.loss_grad(p, x, t)
    idx = np.concatenate([rng.choice(mlp_ref.NP, 60, replace=False), [mlp_ref.NP - 1]])
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp_ref.loss_grad(pp, x, t)[0] - mlp_ref.loss_grad(pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_planted_sign_error_in_backward_is_caught():
    """A wrong sign in dz2 (a plausible slip) breaks the central-difference check."""
    rng = np.random.default_rng(3)
    p = _params(4)
    x = rng.uniform(-1, 1, (20, mlp_ref.KF))
    t = rng.uniform(0, 1, 20)
    _, g = mlp_ref.loss_grad(p, x, t)
    o = mlp_ref.OFF["W2"][0]
    h = 1e-6
    pp, pm = p.copy(), p.copy()
    pp[o] += h
    pm[o] -= h
    fd = (mlp_ref.loss_grad(pp, x, t)[0] - mlp_ref.loss_grad(pm, x, t)[0]) / (2 * h)
    assert abs(fd - g[o]) < 1e-6 * max(1.0, abs(fd))
    if abs(fd) > 1e-8:
        assert abs(fd - (-g[o])) > 1e-6 * abs(fd)


def test_bf16_rounding():
    # exactly representable values are unchanged; ties go to even
    assert mlp_ref.bf16(1.0) == 1.0
    assert mlp_ref.bf16(1.0 + 2.0 ** -8) == 1.0           # tie between 1 and 1 + 2^-7 -> even (1.0)
    assert mlp_ref.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6  # tie -> even mantissa
    assert mlp_ref.bf16(1.0 + 2.0 ** -7) == 1.0 + 2.0 ** -7
    x = np.random.default_rng(5).normal(size=1000)
    r = mlp_ref.bf16(x)
    assert np.all(np.abs(r - x) <= 2.0 ** -8 * np.abs(x) * (1 + 1e-7))


def test_forward_closed_form():
    """Zero weights: y = sigmoid(b3) for every input."""
    p = np.zeros(mlp_ref.NP)
    p[mlp_ref.OFF["b3"][0]] = 0.7
    y = mlp_ref.forward(p, np.ones((3, mlp_ref.KF)))["y"]
    np.testing.assert_allclose(y, 1 / (1 + np.exp(-0.7)), rtol=1e-15)


def test_features_closed_forms():
    """Constant state: centred values 0, no high modes (HG alpha 0). A state
    whose rho*p carries a large highest-degree mode in x: the first energy
    ratio is far above the threshold (feature 16 clipped at 2, alpha =
    alpha_max), and the centred values reach +-1."""
    from trixi_hip.basis import LobattoLegendreBasis
    b = LobattoLegendreBasis(3)
    E = 2
    u = np.zeros((E, 16, 4))
    u[..., 0] = 1.2
    u[..., 1] = 0.1
    u[..., 2] = -0.2
    u[..., 3] = 2.5
    f, a = mlp_ref.features(u.ravel(), 2, 3, b.inverse_vandermonde_legendre, 1.4, 0.5, 0.001)
    assert np.all(f[:, :16] == 0) and np.all(a == 0)
    # rho = 1, v = 0, p = 1 + eps * L3(x): rho*p has a highest-mode component in x
    from numpy.polynomial import legendre
    L3 = legendre.legval(b.nodes, [0, 0, 0, 1])
    p = 1.0 + 0.5 * np.tile(L3, 4)
    u2 = np.zeros((1, 16, 4))
    u2[..., 0] = 1.0
    u2[..., 3] = p / 0.4
    f2, a2 = mlp_ref.features(u2.ravel(), 2, 3, b.inverse_vandermonde_legendre, 1.4, 0.5, 0.001)
    assert f2[0, 16] == 2.0 and a2[0] == 0.5
    assert np.max(np.abs(f2[0, :16])) == pytest.approx(1.0, abs=1e-6)


def test_committed_network_tracks_hg_on_khi_states():
    """The committed weights, evaluated by the FP64 reference on Kelvin-Helmholtz
    states of the oracle (level 4, t = 0 and t = 0.4; no noise, unlike the
    training runs): the clipped alpha close to the training target made from
    the HG indicator's raw alpha."""
    from conftest import PKG
    from oracle import Oracle
    path = os.path.join(PKG, "trixi_hip", "data", "indicator_nn_khi_p3.npz")
    if not os.path.exists(path):
        pytest.skip("trained weights not present")
    from trixi_hip.learned_indicator import training_targets
    with np.load(path, allow_pickle=False) as z:
        params = z["params"].astype(np.float64)
        assert params.shape == (mlp_ref.NP,) and z["feature_lo"].shape == (mlp_ref.NF,)
        dead_zone = float(z["dead_zone"]) if "dead_zone" in z.files else 1.0
    o = Oracle(2, 3, 4, (-1, -1), (1, 1), volume_integral="hg", volume_flux="ranocha", surface_flux="llf_naive",
               fv_flux="llf_naive", alpha_max=0.002, alpha_min=0.0001, alpha_smooth=True)
    o.set_initial_condition("kelvin_helmholtz")
    u = o.compute_coefficients(0.0)
    vinv = o.basis()["Vinv"]
    for state in (u, o.solve(u, 0.0, 0.4, 1.3)[0]):
        f, a_hg = mlp_ref.features(state, 2, 3, vinv, 1.4, 0.002, 0.0001)
        a_nn = mlp_ref.alpha(params, f, 0.002, 0.0001, bf=False)
        err = np.abs(a_nn - 0.002 * training_targets(a_hg, 0.002, 0.0001, dead_zone)) / 0.002
        assert err.mean() < 0.05 and np.quantile(err, 0.95) < 0.2, (err.mean(), np.quantile(err, 0.95))


def test_training_targets():
    """HG factors divided by alpha_max, 0 inside the dead zone below
    dead_zone * alpha_min; dead_zone 1 keeps HG's clipped factors."""
    from trixi_hip.learned_indicator import training_targets
    a = np.array([0.0, 1e-4, 1.2e-4, 1.6e-4, 1e-3, 2e-3])
    np.testing.assert_array_equal(training_targets(a, 0.002, 1e-4, 1.0), (a / 0.002).astype(np.float32))
    np.testing.assert_array_equal(training_targets(a, 0.002, 1e-4, 1.5),
                                  np.array([0, 0, 0, 0.08, 0.5, 1.0], dtype=np.float32))
