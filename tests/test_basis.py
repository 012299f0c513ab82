"""LGL basis of the oracle and of the host package against mpmath fixtures,
closed forms and the SBP invariants the DGSEM operators rely on
(basis_lobatto_legendre.jl:392-727)."""
import math

import numpy as np
import pytest

from conftest import golden


def oracle_basis(N):
    from oracle import Oracle
    return Oracle(2, N, 0, (0, 0), (1, 1)).basis()


@pytest.mark.parametrize("N", range(1, 8))
def test_nodes_weights_vs_mpmath(N):
    ref = golden("basis_mpmath.json")[str(N)]
    b = oracle_basis(N)
    np.testing.assert_allclose(b["nodes"], [float(x) for x in ref["nodes"]], rtol=0, atol=4.5e-16)
    np.testing.assert_allclose(b["weights"], [float(x) for x in ref["weights"]], rtol=2e-15, atol=0)


def test_closed_forms():
    b3 = oracle_basis(3)
    np.testing.assert_allclose(b3["nodes"], [-1, -1 / math.sqrt(5), 1 / math.sqrt(5), 1], atol=1e-16)
    np.testing.assert_allclose(b3["weights"], [1 / 6, 5 / 6, 5 / 6, 1 / 6], rtol=1e-15)
    b4 = oracle_basis(4)
    np.testing.assert_allclose(b4["nodes"], [-1, -math.sqrt(3 / 7), 0, math.sqrt(3 / 7), 1], atol=1e-16)
    np.testing.assert_allclose(b4["weights"], [1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10], rtol=1e-15)


@pytest.mark.parametrize("N", range(2, 8))
def test_operator_invariants(N):
    b = oracle_basis(N)
    x, w, D, Ds, Dh, Vi = b["nodes"], b["weights"], b["D"], b["Dsplit"], b["Dhat"], b["Vinv"]
    n = N + 1
    # exact differentiation of polynomials up to degree N
    for k in range(N + 1):
        np.testing.assert_allclose(D @ x ** k, k * x ** max(k - 1, 0) if k else 0 * x, atol=1e-12)
    np.testing.assert_allclose(D.sum(axis=1), 0, atol=1e-12)
    # SBP: M D + D^T M = B = diag(-1, 0, ..., 0, 1)
    M = np.diag(w)
    B = np.zeros((n, n))
    B[0, 0], B[-1, -1] = -1, 1
    np.testing.assert_allclose(M @ D + D.T @ M, B, atol=1e-13)
    # Dsplit = 2D - M^-1 B has a zero diagonal; Dhat = -M^-1 D^T M
    np.testing.assert_allclose(np.diag(Ds), 0, atol=1e-13)
    np.testing.assert_allclose(Ds, 2 * D - np.diag(1 / w) @ B, atol=1e-13)
    np.testing.assert_allclose(Dh, -np.diag(1 / w) @ D.T @ M, atol=1e-13)
    # Vinv inverts the normalised Legendre Vandermonde matrix
    V = np.array([[np.polynomial.legendre.legval(xi, [0] * m + [1]) * math.sqrt(m + 0.5) for m in range(n)] for xi in x])
    np.testing.assert_allclose(Vi @ V, np.eye(n), atol=1e-12)


@pytest.mark.parametrize("N", range(1, 8))
def test_host_basis_matches_oracle(N):
    from trixi_hip.basis import LobattoLegendreBasis
    h = LobattoLegendreBasis(N)
    b = oracle_basis(N)
    np.testing.assert_array_equal(h.nodes, b["nodes"])
    np.testing.assert_array_equal(h.weights, b["weights"])
    np.testing.assert_array_equal(h.derivative_split, b["Dsplit"])
    np.testing.assert_array_equal(h.derivative_hat, b["Dhat"])
    np.testing.assert_allclose(h.inverse_vandermonde_legendre, b["Vinv"], rtol=1e-13, atol=1e-14)
    # analysis interpolation matrix reproduces polynomials of degree N exactly
    xa = h.analysis_nodes
    for k in range(N + 1):
        np.testing.assert_allclose(h.analysis_vandermonde @ h.nodes ** k, xa ** k, atol=1e-13)
