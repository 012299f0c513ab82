"""LobattoLegendreBasis and the SolutionAnalyzer basis (host-side setup).

Restates src/solvers/dgsem/basis_lobatto_legendre.jl of Trixi.jl in numpy:
  gauss_lobatto_nodes_weights   :570-640 (Kopriva alg. 25)
  calc_q_and_l                  :642-667 (Kopriva alg. 24)
  barycentric_weights           :500-513
  polynomial_derivative_matrix  :427-440 (Kopriva alg. 37)
  calc_Dsplit / calc_Dhat       :399-424
  legendre_polynomial_and_derivative, vandermonde_legendre :677-727
  polynomial_interpolation_matrix :456-484 (Kopriva alg. 32)
  SolutionAnalyzer(analysis_polydeg = 2 polydeg) :274-290
Matrices are numpy arrays indexed M[r, c] with 0-based r, c.
"""
import math

import numpy as np


def _calc_q_and_l(N, x):
    L_Nm2, L_Nm1 = 1.0, x
    Lder_Nm2, Lder_Nm1 = 0.0, 1.0
    L = x
    for i in range(2, N + 1):
        L = ((2 * i - 1) * x * L_Nm1 - (i - 1) * L_Nm2) / i
        Lder = Lder_Nm2 + (2 * i - 1) * L_Nm1
        L_Nm2, L_Nm1 = L_Nm1, L
        Lder_Nm2, Lder_Nm1 = Lder_Nm1, Lder
    q = (2.0 * N + 1) / (N + 1) * (x * L - L_Nm2)
    qder = (2.0 * N + 1) * L
    return q, qder, L


def gauss_lobatto_nodes_weights(n_nodes):
    nodes = np.zeros(n_nodes)
    weights = np.zeros(n_nodes)
    if n_nodes == 1:
        nodes[0], weights[0] = 0.0, 2.0
        return nodes, weights
    N = n_nodes - 1
    tolerance = 2 * np.finfo(np.float64).eps
    nodes[0], nodes[N] = -1.0, 1.0
    weights[0] = 2.0 / (N * (N + 1))
    weights[N] = weights[0]
    if N > 1:
        cont1 = math.pi / N
        cont2 = 3.0 / (8 * N * math.pi)
        for i in range(1, (N + 1) // 2):
            x = -math.cos(cont1 * (i + 0.25) - cont2 / (i + 0.25))
            for _ in range(21):
                q, qder, _L = _calc_q_and_l(N, x)
                dx = -q / qder
                x += dx
                if abs(dx) < tolerance * abs(x):
                    break
            nodes[i] = x
            _, _, L = _calc_q_and_l(N, x)
            weights[i] = weights[0] / L ** 2
            nodes[N - i] = -x
            weights[N - i] = weights[i]
    if n_nodes % 2 == 1:
        _, _, L = _calc_q_and_l(N, 0.0)
        nodes[N // 2] = 0.0
        weights[N // 2] = weights[0] / L ** 2
    return nodes, weights


def barycentric_weights(nodes):
    n = len(nodes)
    w = np.ones(n)
    for j in range(1, n):
        for k in range(j):
            w[k] *= nodes[k] - nodes[j]
            w[j] *= nodes[j] - nodes[k]
    return 1.0 / w


def polynomial_derivative_matrix(nodes):
    n = len(nodes)
    D = np.zeros((n, n))
    wb = barycentric_weights(nodes)
    for i in range(n):
        for j in range(n):
            if j != i:
                D[i, j] = (wb[j] / wb[i]) * 1.0 / (nodes[i] - nodes[j])
                D[i, i] -= D[i, j]
    return D


def calc_dsplit(D, weights):
    Ds = 2.0 * D
    Ds[0, 0] += 1.0 / weights[0]
    Ds[-1, -1] -= 1.0 / weights[-1]
    return Ds


def calc_dhat(D, weights):
    n = len(weights)
    Dh = D.T.copy()
    for c in range(n):
        for r in range(n):
            Dh[r, c] *= -weights[c] / weights[r]
    return Dh


def legendre_normalized(N, x):
    if N == 0:
        poly = 1.0
    elif N == 1:
        poly = x
    else:
        p2, p1 = 1.0, x
        poly = 0.0
        for i in range(2, N + 1):
            poly = ((2 * i - 1) * x * p1 - (i - 1) * p2) / i
            p2, p1 = p1, poly
    return poly * math.sqrt(N + 0.5)


def vandermonde_legendre(nodes):
    n = len(nodes)
    V = np.array([[legendre_normalized(m, nodes[i]) for m in range(n)] for i in range(n)])
    return V, np.linalg.inv(V)


def polynomial_interpolation_matrix(nodes_in, nodes_out):
    wb = barycentric_weights(nodes_in)
    nin, nout = len(nodes_in), len(nodes_out)
    Vm = np.zeros((nout, nin))
    for k in range(nout):
        match = False
        for j in range(nin):
            a, b = nodes_out[k], nodes_in[j]
            if a == b or abs(a - b) <= math.sqrt(np.finfo(np.float64).eps) * max(abs(a), abs(b)):
                match = True
                Vm[k, j] = 1.0
        if not match:
            t = wb / (nodes_out[k] - nodes_in)
            Vm[k, :] = t / t.sum()
    return Vm


def legendre_and_derivative(N, x):
    """Normalized Legendre polynomial and derivative (basis_lobatto_legendre.jl:677-708)."""
    if N == 0:
        poly, deriv = 1.0, 0.0
    elif N == 1:
        poly, deriv = float(x), 1.0
    else:
        p2, p1, d2, d1 = 1.0, float(x), 0.0, 1.0
        poly = deriv = 0.0
        for i in range(2, N + 1):
            poly = ((2 * i - 1) * x * p1 - (i - 1) * p2) / i
            deriv = d2 + (2 * i - 1) * p1
            p2, p1, d2, d1 = p1, poly, d1, deriv
    s = math.sqrt(N + 0.5)
    return poly * s, deriv * s


def gauss_nodes_weights(n_nodes):
    """Legendre-Gauss nodes and weights (basis_gauss_legendre.jl:149-205):
    Newton from cos starting guesses, symmetric roots."""
    N = n_nodes - 1
    nodes, weights = np.ones(n_nodes), np.zeros(n_nodes)
    if N == 0:
        return np.zeros(1), np.full(1, 2.0)
    if N == 1:
        x = math.sqrt(1.0 / 3)
        return np.array([-x, x]), np.ones(2)
    for i in range((N + 1) // 2):
        x = -math.cos(math.pi / (2 * N + 2) * (2 * i + 1))
        for _ in range(21):
            poly, deriv = legendre_and_derivative(N + 1, x)
            dx = -poly / deriv
            x += dx
            if abs(dx) < 2 * np.finfo(float).eps * abs(x):
                break
        poly, deriv = legendre_and_derivative(N + 1, x)
        nodes[i], nodes[N - i] = x, -x
        weights[i] = weights[N - i] = (2 * N + 3) / ((1 - x * x) * deriv * deriv)
    if n_nodes % 2 == 1:
        poly, deriv = legendre_and_derivative(N + 1, 0.0)
        nodes[N // 2] = 0.0
        weights[N // 2] = (2 * N + 3) / (deriv * deriv)
    return nodes, weights


def mortar_l2_matrices(nodes):
    """LobattoLegendreMortarL2 (basis_lobatto_legendre.jl:176-205, l2projection.jl:26-116):
    forward_{upper,lower} interpolate to the two halves; reverse_{upper,lower}
    are the L2 projections back, evaluated on the Gauss nodes."""
    n = len(nodes)
    fu = polynomial_interpolation_matrix(nodes, 0.5 * (nodes + 1))
    fl = polynomial_interpolation_matrix(nodes, 0.5 * (nodes - 1))
    g, gw = gauss_nodes_weights(n)
    g2l = polynomial_interpolation_matrix(g, nodes)
    l2g = polynomial_interpolation_matrix(nodes, g)
    rev = []
    for shift in (1.0, -1.0):
        P = polynomial_interpolation_matrix(g, 0.5 * (g + shift))  # P[j, i] = lg_i(pts_j)
        op = 0.5 * P.T * gw[None, :] / gw[:, None]
        rev.append(g2l @ (op @ l2g))
    return fu, fl, rev[0], rev[1]


class LobattoLegendreBasis:
    """`LobattoLegendreBasis(polydeg)` (basis_lobatto_legendre.jl:55-86)."""

    def __init__(self, polydeg):
        self.polydeg = int(polydeg)
        n = self.polydeg + 1
        self.nodes, self.weights = gauss_lobatto_nodes_weights(n)
        self.inverse_weights = 1.0 / self.weights
        _, self.inverse_vandermonde_legendre = vandermonde_legendre(self.nodes)
        self.derivative_matrix = polynomial_derivative_matrix(self.nodes)
        self.derivative_split = calc_dsplit(self.derivative_matrix, self.weights)
        self.derivative_hat = calc_dhat(self.derivative_matrix, self.weights)
        # SolutionAnalyzer(basis; analysis_polydeg = 2 * polydeg)
        self.analysis_nodes, self.analysis_weights = gauss_lobatto_nodes_weights(2 * self.polydeg + 1)
        self.analysis_vandermonde = polynomial_interpolation_matrix(self.nodes, self.analysis_nodes)
        (self.mortar_forward_upper, self.mortar_forward_lower,
         self.mortar_reverse_upper, self.mortar_reverse_lower) = mortar_l2_matrices(self.nodes)

    def nnodes(self):
        return self.polydeg + 1

    def __repr__(self):
        return f"LobattoLegendreBasis{{Float64}}(polydeg={self.polydeg})"
