// trixi_oracle.cpp -- FP64 CPU restatement of Trixi.jl's DGSEM right-hand side
// for the compressible Euler equations on a uniform periodic TreeMesh (2D, 3D).
//
// TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP library: the
// tests, the smoke check and the CPU-baseline leg of bench.py call it; the
// product path never links, loads or calls it.
//
// Each routine follows the Julia reference loop by loop (file:line cited at
// each function, paths relative to the Trixi.jl source tree). Arithmetic is
// IEEE double throughout; `@muladd` contraction in the reference means results
// agree with Trixi to round-off (~1e-15 relative), not bitwise.
//
// Layout (Julia column-major, first index fastest), 0-based here:
//   u[v + V*(node + NN*e)],  node = i + n*j (+ n*n*k)
//   surface_flux_values sfv[v + V*(face_node + nf*(dir + 2d*e))]
//   face_node = l + n*m with (l,m) the two in-face indices in reference order.

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <vector>
#include <algorithm>
#include <array>
#include <type_traits>
#include <omp.h>

namespace {

// ---------------------------------------------------------------------------
// LGL basis -- src/solvers/dgsem/basis_lobatto_legendre.jl
// ---------------------------------------------------------------------------

// calc_q_and_l, basis_lobatto_legendre.jl:642-667 (Kopriva alg. 24)
void calc_q_and_l(int N, double x, double &q, double &qder, double &L) {
    double L_Nm2 = 1.0, L_Nm1 = x, Lder_Nm2 = 0.0, Lder_Nm1 = 1.0;
    L = x;  // N == 1 never reaches here with an unset L in practice
    for (int i = 2; i <= N; ++i) {
        L = ((2 * i - 1) * x * L_Nm1 - (i - 1) * L_Nm2) / i;
        double Lder = Lder_Nm2 + (2 * i - 1) * L_Nm1;
        L_Nm2 = L_Nm1;
        L_Nm1 = L;
        Lder_Nm2 = Lder_Nm1;
        Lder_Nm1 = Lder;
    }
    q = (2.0 * N + 1) / (N + 1) * (x * L - L_Nm2);
    qder = (2.0 * N + 1) * L;
}

// gauss_lobatto_nodes_weights, basis_lobatto_legendre.jl:570-640 (Kopriva alg. 25)
void gauss_lobatto_nodes_weights(int n_nodes, std::vector<double> &nodes,
                                 std::vector<double> &weights) {
    const int n_iterations = 20;
    const double tolerance = 2 * 2.220446049250313e-16;
    nodes.assign(n_nodes, 0.0);
    weights.assign(n_nodes, 0.0);
    if (n_nodes == 1) {
        nodes[0] = 0;
        weights[0] = 2;
        return;
    }
    int N = n_nodes - 1;
    nodes[0] = -1;
    nodes[N] = 1;
    weights[0] = 2.0 / (N * (N + 1));
    weights[N] = weights[0];
    if (N > 1) {
        double cont1 = M_PI / N;
        double cont2 = 3.0 / (8 * N * M_PI);
        for (int i = 1; i <= (N + 1) / 2 - 1; ++i) {
            nodes[i] = -std::cos(cont1 * (i + 0.25) - cont2 / (i + 0.25));
            for (int k = 0; k <= n_iterations; ++k) {
                double q, qder, L;
                calc_q_and_l(N, nodes[i], q, qder, L);
                double dx = -q / qder;
                nodes[i] += dx;
                if (std::fabs(dx) < tolerance * std::fabs(nodes[i])) break;
            }
            double q, qder, L;
            calc_q_and_l(N, nodes[i], q, qder, L);
            weights[i] = weights[0] / (L * L);
            nodes[N - i] = -nodes[i];
            weights[N - i] = weights[i];
        }
    }
    if (n_nodes % 2 == 1) {
        double q, qder, L;
        calc_q_and_l(N, 0.0, q, qder, L);
        nodes[N / 2] = 0;
        weights[N / 2] = weights[0] / (L * L);
    }
}

// Legendre polynomial and derivative, normalized by sqrt(N + 1/2)
// (legendre_polynomial_and_derivative, basis_lobatto_legendre.jl:677-708)
void legendre_and_derivative(int N, double x, double &poly, double &deriv) {
    if (N == 0) {
        poly = 1.0;
        deriv = 0.0;
    } else if (N == 1) {
        poly = x;
        deriv = 1.0;
    } else {
        double p2 = 1.0, p1 = x, d2 = 0.0, d1 = 1.0;
        poly = deriv = 0.0;
        for (int i = 2; i <= N; ++i) {
            poly = ((2 * i - 1) * x * p1 - (i - 1) * p2) / i;
            deriv = d2 + (2 * i - 1) * p1;
            p2 = p1;
            p1 = poly;
            d2 = d1;
            d1 = deriv;
        }
    }
    poly *= std::sqrt(N + 0.5);
    deriv *= std::sqrt(N + 0.5);
}

// Legendre-Gauss nodes and weights by Newton's method from the Chebyshev-like
// starting guesses, using the symmetry of the roots (gauss_nodes_weights,
// basis_gauss_legendre.jl:149-205).
void gauss_nodes_weights(int n_nodes, std::vector<double> &nodes, std::vector<double> &weights) {
    const int n_iterations = 20;
    const double tolerance = 2 * 2.220446049250313e-16;
    nodes.assign(n_nodes, 1.0);
    weights.assign(n_nodes, 0.0);
    const int N = n_nodes - 1;
    if (N == 0) {
        nodes[0] = 0;
        weights[0] = 2;
        return;
    }
    if (N == 1) {
        nodes[0] = -std::sqrt(1.0 / 3);
        nodes[1] = -nodes[0];
        weights[0] = weights[1] = 1;
        return;
    }
    double poly, deriv;
    for (int i = 0; i <= (N + 1) / 2 - 1; ++i) {
        nodes[i] = -std::cos(M_PI / (2 * N + 2) * (2 * i + 1));
        for (int k = 0; k <= n_iterations; ++k) {
            legendre_and_derivative(N + 1, nodes[i], poly, deriv);
            const double dx = -poly / deriv;
            nodes[i] += dx;
            if (std::fabs(dx) < tolerance * std::fabs(nodes[i])) break;
        }
        legendre_and_derivative(N + 1, nodes[i], poly, deriv);
        weights[i] = (2 * N + 3) / ((1 - nodes[i] * nodes[i]) * deriv * deriv);
        nodes[N - i] = -nodes[i];
        weights[N - i] = weights[i];
    }
    if (n_nodes % 2 == 1) {
        legendre_and_derivative(N + 1, 0.0, poly, deriv);
        nodes[N / 2] = 0;
        weights[N / 2] = (2 * N + 3) / (deriv * deriv);
    }
}

// barycentric_weights, basis_lobatto_legendre.jl:500-513
std::vector<double> barycentric_weights(const std::vector<double> &nodes) {
    int n = (int)nodes.size();
    std::vector<double> w(n, 1.0);
    for (int j = 1; j < n; ++j)
        for (int k = 0; k < j; ++k) {
            w[k] *= nodes[k] - nodes[j];
            w[j] *= nodes[j] - nodes[k];
        }
    for (int j = 0; j < n; ++j) w[j] = 1.0 / w[j];
    return w;
}

// polynomial_derivative_matrix, basis_lobatto_legendre.jl:427-440 (Kopriva alg. 37)
// Row-major: D[i*n + j] = D[i, j].
std::vector<double> polynomial_derivative_matrix(const std::vector<double> &nodes) {
    int n = (int)nodes.size();
    std::vector<double> D(n * n, 0.0);
    auto wbary = barycentric_weights(nodes);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if (j != i) {
                D[i * n + j] = (wbary[j] / wbary[i]) * 1.0 / (nodes[i] - nodes[j]);
                D[i * n + i] -= D[i * n + j];
            }
    return D;
}

// legendre_polynomial_and_derivative (normalized), basis_lobatto_legendre.jl:677-708
double legendre_normalized(int N, double x) {
    double poly;
    if (N == 0) poly = 1.0;
    else if (N == 1) poly = x;
    else {
        double p2 = 1.0, p1 = x;
        poly = 0.0;
        for (int i = 2; i <= N; ++i) {
            poly = ((2 * i - 1) * x * p1 - (i - 1) * p2) / i;
            p2 = p1;
            p1 = poly;
        }
    }
    return poly * std::sqrt(N + 0.5);
}

// Dense inverse by Gauss-Jordan with partial pivoting (the reference calls
// Julia's `inv`, basis_lobatto_legendre.jl:711-727).
std::vector<double> invert(const std::vector<double> &A, int n) {
    std::vector<double> M(A), I(n * n, 0.0);
    for (int i = 0; i < n; ++i) I[i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(M[r * n + c]) > std::fabs(M[p * n + c])) p = r;
        if (p != c)
            for (int k = 0; k < n; ++k) {
                std::swap(M[c * n + k], M[p * n + k]);
                std::swap(I[c * n + k], I[p * n + k]);
            }
        double piv = M[c * n + c];
        for (int k = 0; k < n; ++k) {
            M[c * n + k] /= piv;
            I[c * n + k] /= piv;
        }
        for (int r = 0; r < n; ++r)
            if (r != c) {
                double f = M[r * n + c];
                if (f != 0.0)
                    for (int k = 0; k < n; ++k) {
                        M[r * n + k] -= f * M[c * n + k];
                        I[r * n + k] -= f * I[c * n + k];
                    }
            }
    }
    return I;
}

// polynomial_interpolation_matrix, basis_lobatto_legendre.jl:456-484 (Kopriva alg. 32)
std::vector<double> interpolation_matrix(const std::vector<double> &nodes_in,
                                         const std::vector<double> &nodes_out) {
    int nin = (int)nodes_in.size(), nout = (int)nodes_out.size();
    auto wb = barycentric_weights(nodes_in);
    std::vector<double> V(nout * nin, 0.0);
    const double rtol = 1.4901161193847656e-08;  // isapprox default sqrt(eps)
    for (int k = 0; k < nout; ++k) {
        bool match = false;
        for (int j = 0; j < nin; ++j) {
            double a = nodes_out[k], b = nodes_in[j];
            if (a == b || std::fabs(a - b) <= rtol * std::max(std::fabs(a), std::fabs(b))) {
                match = true;
                V[k * nin + j] = 1;
            }
        }
        if (!match) {
            double s = 0;
            for (int j = 0; j < nin; ++j) {
                double t = wb[j] / (nodes_out[k] - nodes_in[j]);
                V[k * nin + j] = t;
                s += t;
            }
            for (int j = 0; j < nin; ++j) V[k * nin + j] /= s;
        }
    }
    return V;
}

struct Basis {
    int n = 0;
    std::vector<double> nodes, weights, inv_weights, D, Dsplit, Dhat, Vinv;
    // analysis basis (analysis_polydeg = 2*polydeg, basis_lobatto_legendre.jl:274-290)
    int na = 0;
    std::vector<double> a_nodes, a_weights, a_vdm;  // a_vdm[k*n + j], na x n

    void init(int polydeg) {
        n = polydeg + 1;
        gauss_lobatto_nodes_weights(n, nodes, weights);
        inv_weights.resize(n);
        for (int i = 0; i < n; ++i) inv_weights[i] = 1.0 / weights[i];
        std::vector<double> V(n * n);
        for (int i = 0; i < n; ++i)
            for (int m = 0; m < n; ++m) V[i * n + m] = legendre_normalized(m, nodes[i]);
        Vinv = invert(V, n);
        D = polynomial_derivative_matrix(nodes);
        // calc_Dsplit, basis_lobatto_legendre.jl:414-424
        Dsplit.resize(n * n);
        for (int k = 0; k < n * n; ++k) Dsplit[k] = 2.0 * D[k];
        Dsplit[0] += 1.0 / weights[0];
        Dsplit[n * n - 1] -= 1.0 / weights[n - 1];
        // calc_Dhat, basis_lobatto_legendre.jl:399-411: Dhat = D^T, scaled by -w[c]/w[r]
        Dhat.resize(n * n);
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < n; ++c) Dhat[r * n + c] = D[c * n + r];
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r) Dhat[r * n + c] *= -weights[c] / weights[r];
        na = 2 * polydeg + 1;
        gauss_lobatto_nodes_weights(na, a_nodes, a_weights);
        a_vdm = interpolation_matrix(nodes, a_nodes);
        init_mortar_l2();
    }

    // LobattoLegendreMortarL2 (basis_lobatto_legendre.jl:176-205, l2projection.jl:26-116):
    // forward_{upper,lower}[j, i] = l_i(0.5 (xi_j +- 1)) (interpolation to the
    // halves); reverse_{upper,lower} = gauss2lobatto * op * lobatto2gauss with
    // op[i, j] = 0.5 lg_i(0.5 (g_j +- 1)) wg_j / wg_i on the Gauss nodes g.
    std::vector<double> fwd_upper, fwd_lower, rev_upper, rev_lower;  // row-major n x n
    void init_mortar_l2() {
        std::vector<double> up(n), lo(n);
        for (int j = 0; j < n; ++j) {
            up[j] = 0.5 * (nodes[j] + 1);
            lo[j] = 0.5 * (nodes[j] - 1);
        }
        fwd_upper = interpolation_matrix(nodes, up);
        fwd_lower = interpolation_matrix(nodes, lo);
        std::vector<double> g, gw;
        gauss_nodes_weights(n, g, gw);
        auto g2l = interpolation_matrix(g, nodes);  // [lobatto k][gauss j]
        auto l2g = interpolation_matrix(nodes, g);  // [gauss k][lobatto j]
        for (int side = 0; side < 2; ++side) {
            std::vector<double> pts(n), op(n * n);
            for (int j = 0; j < n; ++j) pts[j] = 0.5 * (g[j] + (side == 0 ? 1.0 : -1.0));
            auto P = interpolation_matrix(g, pts);  // P[j * n + i] = lg_i(pts_j)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) op[i * n + j] = 0.5 * P[j * n + i] * gw[j] / gw[i];
            std::vector<double> t(n * n, 0.0), r(n * n, 0.0);
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j)
                    for (int k = 0; k < n; ++k) t[i * n + j] += op[i * n + k] * l2g[k * n + j];
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j)
                    for (int k = 0; k < n; ++k) r[i * n + j] += g2l[i * n + k] * t[k * n + j];
            (side == 0 ? rev_upper : rev_lower) = r;
        }
    }
};

// ---------------------------------------------------------------------------
// Equations -- src/equations/compressible_euler_{2,3}d.jl, numerical_fluxes.jl,
// src/auxiliary/math.jl. Templated on the number of space dimensions D
// (V = D + 2 conserved variables: rho, rho v_1..rho v_D, rho e_total).
// ---------------------------------------------------------------------------

// ln_mean, math.jl:198-210
template <typename R> inline R ln_mean(R x, R y) {
    const R epsilon_f2 = R(1.0e-4);
    R f2 = (x * (x - 2 * y) + y * y) / (x * (x + 2 * y) + y * y);
    if (f2 < epsilon_f2)
        return (x + y) / (2 + f2 * (R(2) / 3 + f2 * (R(2) / 5 + f2 * (R(2) / 7))));
    return (y - x) / std::log(y / x);
}

// inv_ln_mean, math.jl:238-250
template <typename R> inline R inv_ln_mean(R x, R y) {
    const R epsilon_f2 = R(1.0e-4);
    R f2 = (x * (x - 2 * y) + y * y) / (x * (x + 2 * y) + y * y);
    if (f2 < epsilon_f2)
        return (2 + f2 * (R(2) / 3 + f2 * (R(2) / 5 + f2 * (R(2) / 7)))) / (x + y);
    return std::log(y / x) / (y - x);
}

template <int D, typename R = double> struct Euler {
    static constexpr int V = D + 2;
    R gamma, inv_gamma_minus_one;
    explicit Euler(double g = 1.4) : gamma(R(g)), inv_gamma_minus_one(R(1) / (R(g) - 1)) {}

    // cons2prim: euler_3d.jl:1783-1794, euler_2d.jl:2038-2046
    void cons2prim(const R *u, R *w) const {
        R rho = u[0];
        R ke2 = 0;
        w[0] = rho;
        for (int d = 0; d < D; ++d) {
            w[1 + d] = u[1 + d] / rho;
            ke2 += u[1 + d] * w[1 + d];
        }
        w[V - 1] = (gamma - 1) * (u[V - 1] - R(0.5) * ke2);
    }

    // flux: euler_3d.jl:420-447, euler_2d.jl:415-434
    void flux(const R *u, int o, R *f) const {
        R w[V];
        cons2prim(u, w);
        R p = w[V - 1];
        f[0] = u[1 + o];
        for (int d = 0; d < D; ++d) f[1 + d] = u[1 + o] * w[1 + d];
        f[1 + o] += p;
        f[V - 1] = (u[V - 1] + p) * w[1 + o];
    }

    // flux_central: numerical_fluxes.jl:17-25
    void flux_central(const R *ul, const R *ur, int o, R *f) const {
        R fl[V], fr[V];
        flux(ul, o, fl);
        flux(ur, o, fr);
        for (int v = 0; v < V; ++v) f[v] = R(0.5) * (fl[v] + fr[v]);
    }

    // flux_ranocha: euler_3d.jl:746-793, euler_2d.jl:686-722
    void flux_ranocha(const R *ul, const R *ur, int o, R *f) const {
        R wl[V], wr[V];
        cons2prim(ul, wl);
        cons2prim(ur, wr);
        R rho_ll = wl[0], rho_rr = wr[0], p_ll = wl[V - 1], p_rr = wr[V - 1];
        R rho_mean = ln_mean(rho_ll, rho_rr);
        R inv_rho_p_mean = p_ll * p_rr * inv_ln_mean(rho_ll * p_rr, rho_rr * p_ll);
        R vavg[D];
        R velocity_square_avg_sum = 0;
        for (int d = 0; d < D; ++d) {
            vavg[d] = R(0.5) * (wl[1 + d] + wr[1 + d]);
            velocity_square_avg_sum += wl[1 + d] * wr[1 + d];
        }
        R p_avg = R(0.5) * (p_ll + p_rr);
        R velocity_square_avg = R(0.5) * velocity_square_avg_sum;
        R f1 = rho_mean * vavg[o];
        f[0] = f1;
        for (int d = 0; d < D; ++d) f[1 + d] = f1 * vavg[d];
        f[1 + o] += p_avg;
        f[V - 1] = f1 * (velocity_square_avg + inv_rho_p_mean * inv_gamma_minus_one) +
                   R(0.5) * (p_ll * wr[1 + o] + p_rr * wl[1 + o]);
    }

    // max_abs_speed_naive: euler_3d.jl:1112-1133, euler_2d.jl:1423-1440
    R max_abs_speed_naive(const R *ul, const R *ur, int o) const {
        R wl[V], wr[V];
        cons2prim(ul, wl);
        cons2prim(ur, wr);
        R c_ll = std::sqrt(gamma * wl[V - 1] / wl[0]);
        R c_rr = std::sqrt(gamma * wr[V - 1] / wr[0]);
        return std::max(std::fabs(wl[1 + o]), std::fabs(wr[1 + o])) + std::max(c_ll, c_rr);
    }

    // max_abs_speed: euler_3d.jl:1156-1177, euler_2d.jl:1464-1482
    R max_abs_speed(const R *ul, const R *ur, int o) const {
        R wl[V], wr[V];
        cons2prim(ul, wl);
        cons2prim(ur, wr);
        R c_ll = std::sqrt(gamma * wl[V - 1] / wl[0]);
        R c_rr = std::sqrt(gamma * wr[V - 1] / wr[0]);
        return std::max(std::fabs(wl[1 + o]) + c_ll, std::fabs(wr[1 + o]) + c_rr);
    }

    // FluxLaxFriedrichs = FluxPlusDissipation(flux_central, DissipationLocalLaxFriedrichs)
    // numerical_fluxes.jl:37-45, 166-178, 229-253
    void flux_llf(const R *ul, const R *ur, int o, bool naive, R *f) const {
        R fc[V];
        flux_central(ul, ur, o, fc);
        R lambda = naive ? max_abs_speed_naive(ul, ur, o) : max_abs_speed(ul, ur, o);
        for (int v = 0; v < V; ++v) f[v] = fc[v] + (-R(0.5) * lambda * (ur[v] - ul[v]));
    }

    // flux_hllc: euler_3d.jl:1417-1541, euler_2d.jl:1714-1816 (orientation
    // form). Toro's HLL with contact: signal speeds from the Roe averages of
    // the velocity normal to the face and of the enthalpy (Einfeldt), the
    // star-state speed from the jump conditions; the flux of the upwind side,
    // or that side's flux plus its signal speed times the jump to its star
    // state. Operation order as in the reference.
    void flux_hllc(const R *ul, const R *ur, int o, R *f) const {
        const R rho_ll = ul[0], rho_rr = ur[0];
        R v_ll[D], v_rr[D], ke_ll = 0, ke_rr = 0;
        for (int d = 0; d < D; ++d) {
            v_ll[d] = ul[1 + d] / rho_ll;
            v_rr[d] = ur[1 + d] / rho_rr;
            ke_ll += v_ll[d] * v_ll[d];
            ke_rr += v_rr[d] * v_rr[d];
        }
        const R e_ll = ul[V - 1] / rho_ll, e_rr = ur[V - 1] / rho_rr;
        const R p_ll = (gamma - 1) * (ul[V - 1] - R(0.5) * rho_ll * ke_ll);
        const R p_rr = (gamma - 1) * (ur[V - 1] - R(0.5) * rho_rr * ke_rr);
        const R c_ll = std::sqrt(gamma * p_ll / rho_ll), c_rr = std::sqrt(gamma * p_rr / rho_rr);
        R f_ll[V], f_rr[V];
        flux(ul, o, f_ll);
        flux(ur, o, f_rr);
        const R sr_ll = std::sqrt(rho_ll), sr_rr = std::sqrt(rho_rr);
        const R sum_sr = sr_ll + sr_rr;
        const R vel_L = v_ll[o], vel_R = v_rr[o];
        const R vel_roe = (sr_ll * vel_L + sr_rr * vel_R) / sum_sr;
        R mag = 0;
        for (int d = 0; d < D; ++d) {
            const R vr = sr_ll * v_ll[d] + sr_rr * v_rr[d];
            mag += vr * vr;
        }
        mag = mag / (sum_sr * sum_sr);
        const R H_ll = (ul[V - 1] + p_ll) / rho_ll, H_rr = (ur[V - 1] + p_rr) / rho_rr;
        const R H_roe = (sr_ll * H_ll + sr_rr * H_rr) / sum_sr;
        const R c_roe = std::sqrt((gamma - 1) * (H_roe - R(0.5) * mag));
        const R Ssl = std::min(vel_L - c_ll, vel_roe - c_roe);
        const R Ssr = std::max(vel_R + c_rr, vel_roe + c_roe);
        const R sMu_L = Ssl - vel_L, sMu_R = Ssr - vel_R;
        if (Ssl >= 0) {
            for (int v = 0; v < V; ++v) f[v] = f_ll[v];
        } else if (Ssr <= 0) {
            for (int v = 0; v < V; ++v) f[v] = f_rr[v];
        } else {
            const R SStar = (p_rr - p_ll + rho_ll * vel_L * sMu_L - rho_rr * vel_R * sMu_R) /
                            (rho_ll * sMu_L - rho_rr * sMu_R);
            const bool left = Ssl <= 0 && 0 <= SStar;
            const R *us = left ? ul : ur;
            const R *fs = left ? f_ll : f_rr;
            const R *vs = left ? v_ll : v_rr;
            const R rho_s = left ? rho_ll : rho_rr, sMu = left ? sMu_L : sMu_R, S = left ? Ssl : Ssr;
            const R vel = left ? vel_L : vel_R, p_s = left ? p_ll : p_rr, e_s = left ? e_ll : e_rr;
            const R densStar = rho_s * sMu / (S - SStar);
            const R enerStar = e_s + (SStar - vel) * (SStar + p_s / (rho_s * sMu));
            R ustar[V];
            ustar[0] = densStar;
            for (int d = 0; d < D; ++d) ustar[1 + d] = densStar * (d == o ? SStar : vs[d]);
            ustar[V - 1] = densStar * enerStar;
            for (int v = 0; v < V; ++v) f[v] = fs[v] + S * (ustar[v] - us[v]);
        }
    }

    // max_abs_speeds: euler_3d.jl:1770-1775, euler_2d.jl:2025-2030
    void max_abs_speeds(const R *u, R *lam) const {
        R w[V];
        cons2prim(u, w);
        R c = std::sqrt(gamma * w[V - 1] / w[0]);
        for (int d = 0; d < D; ++d) lam[d] = std::fabs(w[1 + d]) + c;
    }

    // pressure: euler_2d.jl (pressure(u, equations)): (gamma - 1) (rho e - |rho v|^2 / (2 rho))
    R pressure(const R *u) const {
        R m2 = 0;
        for (int d = 0; d < D; ++d) m2 += u[1 + d] * u[1 + d];
        return (gamma - 1) * (u[V - 1] - R(0.5) * m2 / u[0]);
    }

    // density_pressure: euler_3d.jl:1951-1956, euler_2d.jl:2205-2210
    R density_pressure(const R *u) const {
        R m2 = 0;
        for (int d = 0; d < D; ++d) m2 += u[1 + d] * u[1 + d];
        return (gamma - 1) * (u[0] * u[V - 1] - R(0.5) * m2);
    }

    // entropy_math: euler_3d.jl:1970-1976 (S = -rho s / (gamma-1), s = log p - gamma log rho)
    R entropy_math(const R *u) const {
        R m2 = 0;
        for (int d = 0; d < D; ++d) m2 += u[1 + d] * u[1 + d];
        R p = (gamma - 1) * (u[V - 1] - R(0.5) * m2 / u[0]);
        R s = std::log(p) - gamma * std::log(u[0]);
        return -s * u[0] * inv_gamma_minus_one;
    }

    // cons2entropy: euler_3d.jl (w = ((gamma - s)/(gamma-1) - 0.5 rho/p |v|^2, rho/p v, -rho/p))
    void cons2entropy(const R *u, R *wout) const {
        R w[V];
        cons2prim(u, w);
        R v2 = 0;
        for (int d = 0; d < D; ++d) v2 += w[1 + d] * w[1 + d];
        R p = w[V - 1];
        R s = std::log(p) - gamma * std::log(w[0]);
        R rho_p = w[0] / p;
        wout[0] = (gamma - s) * inv_gamma_minus_one - R(0.5) * rho_p * v2;
        for (int d = 0; d < D; ++d) wout[1 + d] = rho_p * w[1 + d];
        wout[V - 1] = -rho_p;
    }

    // prim2cons: euler_2d.jl:2120-2127 (and the 3D twin)
    void prim2cons(const R *w, R *u) const {
        u[0] = w[0];
        R ke2 = 0;
        for (int d = 0; d < D; ++d) {
            u[1 + d] = w[0] * w[1 + d];
            ke2 += u[1 + d] * w[1 + d];
        }
        u[V - 1] = w[V - 1] * inv_gamma_minus_one + R(0.5) * ke2;
    }
};

// Numerical-flux selector shared by the volume (two-point), interface and FV paths.
enum FluxKind { FLUX_LLF = 0, FLUX_LLF_NAIVE = 1, FLUX_RANOCHA = 2, FLUX_CENTRAL = 3, FLUX_HLLC = 4 };

template <int D, typename R>
inline void num_flux(const Euler<D, R> &eq, int kind, const R *ul, const R *ur, int o, R *f) {
    switch (kind) {
        case FLUX_LLF: eq.flux_llf(ul, ur, o, false, f); break;
        case FLUX_LLF_NAIVE: eq.flux_llf(ul, ur, o, true, f); break;
        case FLUX_RANOCHA: eq.flux_ranocha(ul, ur, o, f); break;
        case FLUX_HLLC: eq.flux_hllc(ul, ur, o, f); break;
        default: eq.flux_central(ul, ur, o, f); break;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Public C API
// ---------------------------------------------------------------------------

extern "C" {

// Volume-integral selector (src/solvers/dg.jl:135-221, 574-582).
enum { VI_WEAK = 0, VI_FLUX_DIFFERENCING = 1, VI_SHOCK_CAPTURING_HG = 2, VI_PURE_FV = 3 };
// Source-term selector.
enum { SRC_NONE = 0, SRC_CONVERGENCE_TEST = 1 };
// BC_DIRICHLET_IC: BoundaryConditionDirichlet(initial_condition), equations.jl:159-182
// BC_SLIP_WALL: boundary_condition_slip_wall (compressible_euler_2d.jl:322-420,
// compressible_euler_3d.jl:319-415, the TreeMesh methods)
enum { BC_NONE = 0, BC_DIRICHLET_IC = 1, BC_SLIP_WALL = 2 };
// Initial-condition selector.
enum {
    IC_CONSTANT = 0,
    IC_CONVERGENCE_TEST = 1,
    IC_WEAK_BLAST_WAVE = 2,
    IC_ISENTROPIC_VORTEX = 3,
    IC_KELVIN_HELMHOLTZ = 4,
    IC_FOURIER_MODES = 5,
    IC_DENSITY_PULSE = 6,
    IC_BLAST_WAVE = 7,
    IC_BLOB = 8
};

struct oracle_config {
    int ndims;
    int polydeg;
    int initial_refinement_level;
    double coordinates_min[3];
    double coordinates_max[3];
    double gamma;
    int volume_integral;  // VI_*
    int volume_flux;      // FluxKind for flux differencing / HG high-order part
    int surface_flux;     // FluxKind for interfaces
    int fv_flux;          // FluxKind for the FV subcell fluxes (HG / pure FV)
    double alpha_max;
    double alpha_min;
    int alpha_smooth;
    int source_terms;     // SRC_*
    int periodicity[3];   // TreeMesh(...; periodicity) per dimension (nonzero = periodic)
    int boundary_condition;  // BC_* for every non-periodic boundary
    int npatches;         // refinement_patches of type "box", applied in order
    double patch_min[8][3], patch_max[8][3];
    int hg_variable;      // shock-capturing indicator variable: 0 density_pressure, 1 density, 2 pressure
};

struct oracle_ctx;

}  // extern "C"

// A leaf cell of the TreeMesh: level, integer coordinates at that level, centre.
struct TreeCell {
    int level, ic[3];
    double center[3];
};

// Where a leaf comes from in an adaptation: kind 0 = the same cell as leaf src
// before; kind 1 + q = child q of leaf src; kind -1 = the parent of the 2^D
// consecutive leaves src .. src + 2^D - 1.
struct Origin {
    int kind, src;
};

struct oracle_ctx {
    oracle_config cfg;
    int D, n, NN, nf, V, E, I;
    Basis basis;
    // mesh (src/meshes/abstract_tree.jl, containers_{2,3}d.jl)
    std::vector<double> centers;        // [D, E]
    std::vector<double> lengths;        // [E]
    std::vector<double> inverse_jacobian;  // [E]
    std::vector<double> node_coords;    // [D, NN, E]
    std::vector<int32_t> neighbor_ids;  // [2, I] 0-based element ids
    std::vector<int32_t> orientations;  // [I] 0-based orientation
    std::vector<int32_t> elem_nbr;      // [2D, E] neighbour element per direction (0-based dir), -1 at a boundary
    std::vector<int32_t> bnd_elem, bnd_dir;  // boundaries: element and its 0-based direction (init_boundaries!)
    // L2 mortars (init_mortars!, containers_3d.jl:652-788): small elements in
    // positions q = (first free dim +) + 2 (second free dim +), then the large one
    std::vector<int32_t> mortar_ids;      // [M][2^(D-1) + 1]
    std::vector<int32_t> mortar_orient;   // [M] 0-based orientation
    std::vector<int32_t> mortar_side;     // [M] large_sides: 1 = large element on the left
    int M = 0;
    std::vector<int32_t> levels;          // [E] tree level of each element
    std::vector<TreeCell> leaves;         // [E] leaf cells in element (depth-first) order
    std::vector<double> u_adapted;        // solution projected by the last adaptation
    std::vector<Origin> transfer;         // [E] origin of each element in the last adaptation
    std::vector<double> sfv;            // surface_flux_values
    std::vector<double> alpha, alpha_tmp;
    // blending factors given from outside (oracle_set_alpha): the shock-capturing
    // volume integral uses them instead of running the HG indicator, as
    // VolumeIntegralShockCapturingHG uses whatever its indicator returns
    // (calc_volume_integral.jl:231-272)
    bool alpha_given = false;
    std::vector<double> alpha_given_v;
    int hg_variable = 0;  // HG indicator variable: 0 density_pressure, 1 density, 2 pressure
    // analysis exact-solution parameters
    int ic_kind = IC_CONSTANT;
    double ic_params[8] = {0};
    std::vector<double> ic_modes;       // IC_FOURIER_MODES: [4 fields][8 modes][kx,ky,kz,phase]
    double domain_length = 0;
};

namespace {

// Uniform refinement in depth-first order: refine_uniformly!, abstract_tree.jl:206-250;
// child_coordinates :668-674; child_sign table :118-128; leaves in id order :154-166.
void build_leaves(int D, int level, int max_level, const double *center, double length,
                  const int *icoord, std::vector<double> &centers, std::vector<double> &lengths,
                  std::vector<int> &icoords) {
    if (level >= max_level) {
        for (int d = 0; d < D; ++d) centers.push_back(center[d]);
        lengths.push_back(length);
        for (int d = 0; d < D; ++d) icoords.push_back(icoord[d]);
        return;
    }
    int nchild = 1 << D;
    for (int child = 0; child < nchild; ++child) {
        double cc[3];
        int ic[3];
        double child_length = length / 2;
        for (int d = 0; d < D; ++d) {
            int sgn = ((child >> d) & 1) ? 1 : -1;  // child_sign(child, d)
            cc[d] = center[d] + sgn * child_length / 2;
            ic[d] = 2 * icoord[d] + (sgn > 0 ? 1 : 0);
        }
        build_leaves(D, level + 1, max_level, cc, child_length, ic, centers, lengths, icoords);
    }
}

// ---------------------------------------------------------------------------
// Initial conditions (host-side setup) -- cited per case.
// ---------------------------------------------------------------------------
template <int D>
void initial_condition(const oracle_ctx &c, const Euler<D> &eq, const double *x, double t,
                       double *u) {
    const int V = D + 2;
    switch (c.ic_kind) {
        case IC_CONSTANT: {
            // euler_3d.jl:72-80, euler_2d.jl:71-78
            u[0] = 1;
            u[1] = 0.1;
            u[2] = -0.2;
            if (D == 3) u[3] = 0.7;
            u[V - 1] = 10;
            break;
        }
        case IC_CONVERGENCE_TEST: {
            // euler_3d.jl:94-111, euler_2d.jl:87-106
            double cst = 2, A = 0.1, L = 2, f = 1.0 / L, om = 2 * M_PI * f;
            double s = 0;
            for (int d = 0; d < D; ++d) s += x[d];
            double ini = cst + A * std::sin(om * (s - t));
            for (int v = 0; v < V - 1; ++v) u[v] = ini;
            u[V - 1] = ini * ini;
            break;
        }
        case IC_WEAK_BLAST_WAVE: {
            if (D == 2) {
                // euler_2d.jl:181-200 (polar coordinates, sincos of phi)
                const double xn = x[0], yn = x[1];
                const double r = std::sqrt(xn * xn + yn * yn);
                const double phi = std::atan2(yn, xn);
                const double w[4] = {r > 0.5 ? 1.0 : 1.1691, r > 0.5 ? 0.0 : 0.1882 * std::cos(phi),
                                     r > 0.5 ? 0.0 : 0.1882 * std::sin(phi), r > 0.5 ? 1.0 : 1.245};
                eq.prim2cons(w, u);
                break;
            }
            // euler_3d.jl:163-183
            double xn = x[0], yn = x[1], zn = D == 3 ? x[2] : 0.0;
            double r = std::sqrt(xn * xn + yn * yn + zn * zn);
            double phi = std::atan2(yn, xn);
            double theta = r == 0 ? 0.0 : std::acos(zn / r);
            double w[5];
            w[0] = r > 0.5 ? 1.0 : 1.1691;
            w[1] = r > 0.5 ? 0.0 : 0.1882 * std::cos(phi) * std::sin(theta);
            w[2] = r > 0.5 ? 0.0 : 0.1882 * std::sin(phi) * std::sin(theta);
            w[3] = r > 0.5 ? 0.0 : 0.1882 * std::cos(theta);
            w[4] = r > 0.5 ? 1.0 : 1.245;
            eq.prim2cons(w, u);
            break;
        }
        case IC_ISENTROPIC_VORTEX: {
            // examples/tree_2d_dgsem/elixir_euler_vortex.jl:18-50; background
            // pressure taken from ic_params[0] (25 in the elixir).
            double rho = 1, v1 = 1, v2 = 1, p = c.ic_params[0];
            double iniamplitude = 5;
            double rt = p / rho;
            double cx = x[0] - 0.0, cy = x[1] - 0.0;
            double c0 = -cy, c1 = cx;
            double r2 = c0 * c0 + c1 * c1;
            double du = iniamplitude / (2 * M_PI) * std::exp(0.5 * (1 - r2));
            double dtemp = -(eq.gamma - 1) / (2 * eq.gamma * rt) * du * du;
            rho = rho * std::pow(1 + dtemp, 1 / (eq.gamma - 1));
            v1 = v1 + du * c0;
            v2 = v2 + du * c1;
            p = p * std::pow(1 + dtemp, eq.gamma / (eq.gamma - 1));
            double w[4] = {rho, v1, v2, p};
            eq.prim2cons(w, u);
            break;
        }
        case IC_KELVIN_HELMHOLTZ: {
            // examples/tree_2d_dgsem/elixir_euler_kelvin_helmholtz_instability.jl:18-31
            double slope = 15;
            double B = std::tanh(slope * x[1] + 7.5) - std::tanh(slope * x[1] - 7.5);
            double rho = 0.5 + 0.75 * B;
            double v1 = 0.5 * (B - 1);
            double v2 = 0.1 * std::sin(M_PI * 2 * x[0]);
            double w[4] = {rho, v1, v2, 1.0};
            eq.prim2cons(w, u);
            break;
        }
        case IC_DENSITY_PULSE: {
            // examples/tree_3d_dgsem/elixir_euler_amr.jl:15-26: a Gaussian density
            // pulse advected with v = 1 in every direction at p = 1 (3D)
            double rho = 1 + std::exp(-(x[0] * x[0] + x[1] * x[1] + (D == 3 ? x[2] * x[2] : 0.0))) / 2;
            u[0] = rho;
            for (int d = 0; d < D; ++d) u[1 + d] = rho;
            u[V - 1] = 1 / (eq.gamma - 1) + 1.0 / 2 * rho * D;
            break;
        }
        case IC_BLAST_WAVE: {
            // examples/tree_2d_dgsem/elixir_euler_blast_wave_amr.jl:14-33: the medium
            // blast wave (2D; p = 1e-3 outside r = 0.5)
            double r = std::sqrt(x[0] * x[0] + x[1] * x[1]);
            double phi = std::atan2(x[1], x[0]);
            double w[4];
            w[0] = r > 0.5 ? 1.0 : 1.1691;
            w[1] = r > 0.5 ? 0.0 : 0.1882 * std::cos(phi);
            w[2] = r > 0.5 ? 0.0 : 0.1882 * std::sin(phi);
            w[3] = r > 0.5 ? 1.0e-3 : 1.245;
            if (D == 2) {
                double wp[4] = {w[0], w[1], w[2], w[3]};
                eq.prim2cons(wp, u);
            }
            break;
        }
        case IC_BLOB: {
            // examples/tree_2d_dgsem/elixir_euler_blob_amr.jl:13-58: a dense blob in a
            // supersonic stream (Agertz et al. 2006), tanh transition of slope 2
            const double R = 1, dens0 = 1, Chi = 10, tau_kh = 1, tau_cr = tau_kh / 1.6;
            const double velx0 = 2 * R * std::sqrt(Chi) / tau_cr, vely0 = 0, Ma0 = 2.7;
            const double cs = velx0 / Ma0, p0 = cs * cs * dens0 / eq.gamma;
            const double xr = x[0] - (-15), yr = x[1] - 0;
            const double r = std::sqrt(xr * xr + yr * yr), slope = 2;
            const double shape = 0.5 * (1 + (std::tanh(slope * (r + R)) - (std::tanh(slope * (r - R)) + 1)));
            const double dens = dens0 + (Chi - 1) * shape;
            const double velx = velx0 - velx0 * shape;
            double w[4] = {dens, velx, vely0, p0};
            if (D == 2) eq.prim2cons(w, u);
            break;
        }
        case IC_FOURIER_MODES: {
            // Seeded smooth state: rho = 1 + A_rho sum_m sin(pi k_m.(x+1) + phi_m),
            // v_d = A_v sum_m sin(...), p = 1 (modes passed by the caller).
            double w[5];
            for (int fld = 0; fld < 1 + D; ++fld) {
                double s = 0;
                for (int m = 0; m < 8; ++m) {
                    const double *md = &c.ic_modes[(fld * 8 + m) * 4];
                    double arg = 0;
                    for (int d = 0; d < D; ++d) arg += md[d] * (x[d] + 1.0);
                    s += std::sin(M_PI * arg + md[3]);
                }
                w[fld] = (fld == 0 ? 1.0 + c.ic_params[0] * s : c.ic_params[1] * s);
            }
            w[D + 1] = 1.0;
            eq.prim2cons(w, u);
            break;
        }
    }
}

// source_terms_convergence_test: euler_3d.jl:127-153, euler_2d.jl:120-146
template <int D, typename R> void source_convergence(const Euler<D, R> &eq, const double *x, double t, R *s) {
    const R cst = 2, A = R(0.1), L = 2, f = R(1) / L, om = 2 * R(M_PI) * f, g = eq.gamma;
    R xs = 0;
    for (int d = 0; d < D; ++d) xs += R(x[d]);
    R si = std::sin(om * (xs - R(t))), co = std::cos(om * (xs - R(t)));
    R rho = cst + A * si;
    R rho_x = om * A * co;
    if (D == 3) {
        R tmp = (2 * rho - R(1.5)) * (g - 1);
        s[0] = 2 * rho_x;
        s[1] = rho_x * (2 + tmp);
        s[2] = s[1];
        s[3] = s[1];
        s[4] = rho_x * (4 * rho + 3 * tmp);
    } else {
        R tmp = (2 * rho - 1) * (g - 1);
        s[0] = rho_x;
        s[1] = rho_x * (1 + tmp);
        s[2] = s[1];
        s[3] = 2 * rho_x * (rho + tmp);
    }
}

inline int ipow(int b, int e) {
    int r = 1;
    for (int i = 0; i < e; ++i) r *= b;
    return r;
}

// Node multi-index helpers: node = i + n*j (+ n*n*k).
template <int D> inline int node_index(int n, const int *idx) {
    int r = 0, s = 1;
    for (int d = 0; d < D; ++d) {
        r += idx[d] * s;
        s *= n;
    }
    return r;
}

// Face-node mapping: for orientation o (0-based) the face of nodes with index_o
// fixed; in-face indices are the remaining indices in increasing dimension order
// (prolong2interfaces!, dg_3d.jl:530-567, dg_2d.jl:515-541).
template <int D> inline int face_to_node(int n, int o, int side_index, int fnode) {
    int idx[3];
    int rest[2] = {fnode % n, fnode / n};
    int r = 0;
    for (int d = 0; d < D; ++d) {
        if (d == o) idx[d] = side_index;
        else idx[d] = rest[r++];
    }
    return node_index<D>(n, idx);
}

// boundary_condition_slip_wall on a TreeMesh face of orientation o
// (compressible_euler_2d.jl:322-420, compressible_euler_3d.jl:319-415): the
// inner state is rotated to the outward normal n = +-e_o (rotate_to_x with the
// tangents of the reference: 3D o = x: (e_z, e_y), o = y: (e_x, e_z), o = z:
// (e_y, e_x) up to sign; 2D o = x: e_y, o = y: e_x), the wall pressure p* of
// the pressure Riemann problem (Toro 2009, sec. 6.3.3, clipped at 0) gives the
// flux; only the normal momentum component is nonzero, and for either side of
// the element it is p* in the direction of +e_o.
template <int D, typename R> void slip_wall_flux(const Euler<D, R> &eq, const R *u, int o, bool plus, R *f) {
    constexpr int V = D + 2;
    const R rho = u[0];
    const R mn = plus ? u[1 + o] : -u[1 + o];
    R mt[2] = {R(0), R(0)};
    if (D == 3) {
        const int t1 = o == 0 ? 2 : (o == 1 ? 0 : 1), t2 = o == 0 ? 1 : (o == 1 ? 2 : 0);
        mt[0] = u[1 + t1];
        mt[1] = u[1 + t2];
    } else {
        mt[0] = u[1 + (1 - o)];
    }
    // cons2prim of the rotated state (component order n, t1, t2); the signs of
    // the tangential momenta do not change their products m * (m / rho)
    const R vn = mn / rho;
    R ke2 = mn * vn;
    for (int d = 0; d < D - 1; ++d) ke2 += mt[d] * (mt[d] / rho);
    const R p = (eq.gamma - 1) * (u[V - 1] - R(0.5) * ke2);
    R p_star;
    if (vn <= 0) {
        const R c = std::sqrt(eq.gamma * p / rho);
        const R base = 1 + R(0.5) * (eq.gamma - 1) * vn / c;
        if (base >= 0) p_star = p * std::pow(base, 2 * eq.gamma * eq.inv_gamma_minus_one);
        else p_star = R(0);
    } else {
        const R A = 2 / ((eq.gamma + 1) * rho);
        const R B = p * (eq.gamma - 1) / (eq.gamma + 1);
        p_star = p + R(0.5) * vn / A * (vn + std::sqrt(vn * vn + 4 * A * (p + B)));
    }
    for (int v = 0; v < V; ++v) f[v] = R(0);
    f[1 + o] = p_star;
}

// ---------------------------------------------------------------------------
// Volume kernels
// ---------------------------------------------------------------------------

// weak_form_kernel!: dg_3d.jl:133-164, dg_2d.jl:195-221
template <int D, typename R>
void weak_form_kernel(const oracle_ctx &c, const Euler<D, R> &eq, R *du, const R *u,
                      int e, R alpha) {
    const int V = D + 2, n = c.n, NN = c.NN;
    const double *Dhat = c.basis.Dhat.data();
    for (int node = 0; node < NN; ++node) {
        int idx[3] = {node % n, (node / n) % n, node / (n * n)};
        const R *un = &u[V * (node + NN * e)];
        for (int o = 0; o < D; ++o) {
            R f[V];
            eq.flux(un, o, f);
            for (int ii = 0; ii < n; ++ii) {
                int jdx[3] = {idx[0], idx[1], idx[2]};
                jdx[o] = ii;
                int tnode = node_index<D>(n, jdx);
                R fac = alpha * R(Dhat[ii * n + idx[o]]);
                for (int v = 0; v < V; ++v) du[V * (tnode + NN * e) + v] += fac * f[v];
            }
        }
    }
}

// flux_differencing_kernel!: dg_3d.jl:166-214, dg_2d.jl:223-259
template <int D, typename R>
void flux_differencing_kernel(const oracle_ctx &c, const Euler<D, R> &eq, int fluxkind, R *du,
                              const R *u, int e, R alpha) {
    const int V = D + 2, n = c.n, NN = c.NN;
    const double *Ds = c.basis.Dsplit.data();
    for (int node = 0; node < NN; ++node) {
        int idx[3] = {node % n, (node / n) % n, node / (n * n)};
        const R *un = &u[V * (node + NN * e)];
        for (int o = 0; o < D; ++o) {
            int i = idx[o];
            for (int ii = i + 1; ii < n; ++ii) {
                int jdx[3] = {idx[0], idx[1], idx[2]};
                jdx[o] = ii;
                int onode = node_index<D>(n, jdx);
                const R *uo = &u[V * (onode + NN * e)];
                R f[V];
                num_flux<D>(eq, fluxkind, un, uo, o, f);
                R fi = alpha * R(Ds[i * n + ii]), fii = alpha * R(Ds[ii * n + i]);
                for (int v = 0; v < V; ++v) {
                    du[V * (node + NN * e) + v] += fi * f[v];
                    du[V * (onode + NN * e) + v] += fii * f[v];
                }
            }
        }
    }
}

// fv_kernel! + calcflux_fv!: dg_3d.jl:268-306, 359-389; dg_2d.jl:407-464.
// Subcell fluxes at element faces (index 0 and n) are zero (create_f_threaded,
// dg_3d.jl:11-54).
template <int D, typename R>
void fv_kernel(const oracle_ctx &c, const Euler<D, R> &eq, int fluxkind, R *du,
               const R *u, int e, R alpha) {
    const int V = D + 2, n = c.n, NN = c.NN;
    const double *iw = c.basis.inv_weights.data();
    // fstar[o][v + V*(node index with index_o in 0..n)] -- store per orientation
    // as [n+1 along o] x [n^(D-1)] x V
    std::vector<R> fstar(D * (n + 1) * ipow(n, D - 1) * V, 0.0);
    auto fidx = [&](int o, int io, int rest) { return ((o * (n + 1) + io) * ipow(n, D - 1) + rest) * V; };
    for (int o = 0; o < D; ++o) {
        for (int rest = 0; rest < ipow(n, D - 1); ++rest) {
            for (int io = 1; io < n; ++io) {
                int nl = face_to_node<D>(n, o, io - 1, rest);
                int nr = face_to_node<D>(n, o, io, rest);
                R f[V];
                num_flux<D>(eq, fluxkind, &u[V * (nl + NN * e)], &u[V * (nr + NN * e)], o, f);
                for (int v = 0; v < V; ++v) fstar[fidx(o, io, rest) + v] = f[v];
            }
        }
    }
    for (int node = 0; node < NN; ++node) {
        int idx[3] = {node % n, (node / n) % n, node / (n * n)};
        for (int v = 0; v < V; ++v) {
            R acc = 0;
            for (int o = 0; o < D; ++o) {
                int rest = 0, s = 1;
                for (int d = 0; d < D; ++d)
                    if (d != o) {
                        rest += idx[d] * s;
                        s *= n;
                    }
                int io = idx[o];
                acc += R(iw[io]) * (fstar[fidx(o, io + 1, rest) + v] - fstar[fidx(o, io, rest) + v]);
            }
            du[V * (node + NN * e) + v] += alpha * acc;
        }
    }
}

// ---------------------------------------------------------------------------
// Hennemann-Gassner indicator: indicators.jl:119-156, indicators_2d.jl:30-135,
// indicators_3d.jl:39-173, interpolation.jl:207-235, 348-386.
// ---------------------------------------------------------------------------
template <int D>
void hg_indicator(oracle_ctx &c, const Euler<D> &eq, const double *u) {
    const int V = D + 2, n = c.n, NN = c.NN, E = c.E;
    const double threshold = 0.5 * std::pow(10.0, -1.8 * std::pow((double)n, 0.25));
    const double o_0001 = 0.0001;
    const double parameter_s = std::log((1 - o_0001) / o_0001);
    const double *Vi = c.basis.Vinv.data();
    c.alpha.assign(E, 0.0);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < E; ++e) {
        std::vector<double> ind(NN), t1(NN), t2(NN), modal(NN);
        for (int node = 0; node < NN; ++node) {
            const double *un = &u[V * (node + NN * e)];
            ind[node] = c.hg_variable == 1 ? un[0] : c.hg_variable == 2 ? eq.pressure(un) : eq.density_pressure(un);
        }
        // dimensionwise modal transform
        std::vector<double> *src = &ind;
        std::vector<double> *bufs[3] = {&t1, &t2, &modal};
        if (D == 2) bufs[1] = &modal;
        for (int o = 0; o < D; ++o) {
            std::vector<double> &dst = *bufs[o];
            for (int node = 0; node < NN; ++node) {
                int idx[3] = {node % n, (node / n) % n, node / (n * n)};
                double res = 0;
                for (int ii = 0; ii < n; ++ii) {
                    int jdx[3] = {idx[0], idx[1], idx[2]};
                    jdx[o] = ii;
                    res += Vi[idx[o] * n + ii] * (*src)[node_index<D>(n, jdx)];
                }
                dst[node] = res;
            }
            src = &dst;
        }
        auto m2 = [&](int i, int j, int k) {
            int idx[3] = {i, j, k};
            double m = modal[node_index<D>(n, idx)];
            return m * m;
        };
        double te2 = 0, te1, te;
        if (D == 2) {
            for (int j = 0; j < n - 2; ++j)
                for (int i = 0; i < n - 2; ++i) te2 += m2(i, j, 0);
            te1 = te2;
            for (int i = 0; i < n - 1; ++i) te1 += m2(i, n - 2, 0);
            for (int j = 0; j < n - 2; ++j) te1 += m2(n - 2, j, 0);
            te = te1;
            for (int i = 0; i < n; ++i) te += m2(i, n - 1, 0);
            for (int j = 0; j < n - 1; ++j) te += m2(n - 1, j, 0);
        } else {
            for (int k = 0; k < n - 2; ++k)
                for (int j = 0; j < n - 2; ++j)
                    for (int i = 0; i < n - 2; ++i) te2 += m2(i, j, k);
            te1 = te2;
            for (int j = 0; j < n - 1; ++j)
                for (int i = 0; i < n - 1; ++i) te1 += m2(i, j, n - 2);
            for (int k = 0; k < n - 2; ++k)
                for (int i = 0; i < n - 1; ++i) te1 += m2(i, n - 2, k);
            for (int k = 0; k < n - 2; ++k)
                for (int j = 0; j < n - 2; ++j) te1 += m2(n - 2, j, k);
            te = te1;
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) te += m2(i, j, n - 1);
            for (int k = 0; k < n - 1; ++k)
                for (int i = 0; i < n; ++i) te += m2(i, n - 1, k);
            for (int k = 0; k < n - 1; ++k)
                for (int j = 0; j < n - 1; ++j) te += m2(n - 1, j, k);
        }
        double ef1 = te != 0 ? (te - te1) / te : 0.0;
        double ef2 = te1 != 0 ? (te1 - te2) / te1 : 0.0;
        double energy = std::max(ef1, ef2);
        double a = 1 / (1 + std::exp(-parameter_s / threshold * (energy - threshold)));
        if (a < c.cfg.alpha_min) a = 0;
        if (a > 1 - c.cfg.alpha_min) a = 1;
        c.alpha[e] = std::min(c.cfg.alpha_max, a);
    }
    if (c.cfg.alpha_smooth) {
        // apply_smoothing!, indicators_2d.jl:104-135 / indicators_3d.jl:125-173
        c.alpha_tmp = c.alpha;
        for (int i = 0; i < c.I; ++i) {
            int l = c.neighbor_ids[2 * i], r = c.neighbor_ids[2 * i + 1];
            c.alpha[l] = std::max(std::max(c.alpha_tmp[l], 0.5 * c.alpha_tmp[r]), c.alpha[l]);
            c.alpha[r] = std::max(std::max(c.alpha_tmp[r], 0.5 * c.alpha_tmp[l]), c.alpha[r]);
        }
        // L2 mortars: every small element and the large one (indicators_3d.jl:145-171)
        const int ns = 1 << (D - 1);
        for (int m = 0; m < c.M; ++m) {
            const int32_t *ids = &c.mortar_ids[(size_t)m * (ns + 1)];
            const int large = ids[ns];
            for (int q = 0; q < ns; ++q)
                c.alpha[ids[q]] = std::max(std::max(c.alpha_tmp[ids[q]], 0.5 * c.alpha_tmp[large]), c.alpha[ids[q]]);
            for (int q = 0; q < ns; ++q)
                c.alpha[large] = std::max(std::max(c.alpha_tmp[large], 0.5 * c.alpha_tmp[ids[q]]), c.alpha[large]);
        }
    }
}

// ---------------------------------------------------------------------------
// rhs_hyperbolic!: dg_2d.jl:113-186 (shared by 2D/3D TreeMesh)
// ---------------------------------------------------------------------------
template <int D, typename R> void rhs(oracle_ctx &c, R *du, const R *u, double t) {
    const int V = D + 2, n = c.n, NN = c.NN, nf = c.nf, E = c.E, I = c.I;
    Euler<D, R> eq(c.cfg.gamma);
    std::vector<R> sfv_local;
    R *sfv;
    if constexpr (std::is_same<R, double>::value) sfv = c.sfv.data();
    else {
        sfv_local.assign(c.sfv.size(), R(0));
        sfv = sfv_local.data();
    }
    const size_t total = (size_t)V * NN * E;
    // set_zero!, solvers.jl:8-29
    for (size_t k = 0; k < total; ++k) du[k] = R(0);

    // calc_volume_integral!, calc_volume_integral.jl:180-272
    const int vi = c.cfg.volume_integral;
    if (vi == VI_SHOCK_CAPTURING_HG && c.alpha_given) {
        c.alpha = c.alpha_given_v;
    } else if (vi == VI_SHOCK_CAPTURING_HG) {
        if constexpr (std::is_same<R, double>::value) hg_indicator<D>(c, Euler<D>(c.cfg.gamma), u);
        else {
            std::vector<double> ud(u, u + total);
            hg_indicator<D>(c, Euler<D>(c.cfg.gamma), ud.data());
        }
    }
    const double atol = 1.8189894035458565e-12;  // max(100 eps, eps^0.75) for Float64
#pragma omp parallel for schedule(static)
    for (int e = 0; e < E; ++e) {
        if (vi == VI_WEAK) weak_form_kernel<D, R>(c, eq, du, u, e, 1.0);
        else if (vi == VI_FLUX_DIFFERENCING) flux_differencing_kernel<D, R>(c, eq, c.cfg.volume_flux, du, u, e, 1.0);
        else if (vi == VI_PURE_FV) fv_kernel<D, R>(c, eq, c.cfg.fv_flux, du, u, e, 1.0);
        else {
            R a = R(c.alpha[e]);
            bool dg_only = std::fabs(c.alpha[e]) <= atol;
            if (dg_only) flux_differencing_kernel<D, R>(c, eq, c.cfg.volume_flux, du, u, e, 1.0);
            else {
                flux_differencing_kernel<D, R>(c, eq, c.cfg.volume_flux, du, u, e, 1 - a);
                fv_kernel<D, R>(c, eq, c.cfg.fv_flux, du, u, e, a);
            }
        }
    }

    // prolong2interfaces! + calc_interface_flux!: dg_3d.jl:530-602, dg_2d.jl:515-632
#pragma omp parallel for schedule(static)
    for (int i = 0; i < I; ++i) {
        int left = c.neighbor_ids[2 * i], right = c.neighbor_ids[2 * i + 1];
        int o = c.orientations[i];
        int left_dir = 2 * o + 1, right_dir = 2 * o;  // 0-based: 2o (+side) and 2o-1 (-side)
        for (int fn = 0; fn < nf; ++fn) {
            int nl = face_to_node<D>(n, o, n - 1, fn);
            int nr = face_to_node<D>(n, o, 0, fn);
            R f[V];
            num_flux<D>(eq, c.cfg.surface_flux, &u[V * (nl + NN * left)], &u[V * (nr + NN * right)], o, f);
            for (int v = 0; v < V; ++v) {
                sfv[V * (fn + nf * (left_dir + 2 * D * left)) + v] = f[v];
                sfv[V * (fn + nf * (right_dir + 2 * D * right)) + v] = f[v];
            }
        }
    }

    // prolong2boundaries! + calc_boundary_flux!: dg_2d.jl:680-860, dg_3d.jl:651-768.
    // BoundaryConditionDirichlet(initial_condition) (equations.jl:164-182): the
    // outer state is the initial condition at the boundary node and time t; the
    // element is the left state at its + faces (even 1-based direction).
    // boundary_condition_slip_wall: the flux of slip_wall_flux from the inner state.
    {
        const Euler<D> eqd(c.cfg.gamma);
        const int B = (int)c.bnd_elem.size();
#pragma omp parallel for schedule(static)
        for (int b = 0; b < B; ++b) {
            const int e = c.bnd_elem[b], dir = c.bnd_dir[b], o = dir / 2;
            const bool plus = dir % 2 == 1;
            for (int fn = 0; fn < nf; ++fn) {
                const int node = face_to_node<D>(n, o, plus ? n - 1 : 0, fn);
                if (c.cfg.boundary_condition == BC_SLIP_WALL) {
                    R f[V];
                    slip_wall_flux<D>(eq, &u[V * (node + NN * e)], o, plus, f);
                    for (int v = 0; v < V; ++v) sfv[V * (fn + nf * (dir + 2 * D * e)) + v] = f[v];
                    continue;
                }
                double ub[D + 2];
                initial_condition<D>(c, eqd, &c.node_coords[D * (node + NN * e)], t, ub);
                R ubr[V], f[V];
                for (int v = 0; v < V; ++v) ubr[v] = R(ub[v]);
                const R *ui = &u[V * (node + NN * e)];
                if (plus) num_flux<D>(eq, c.cfg.surface_flux, ui, ubr, o, f);
                else num_flux<D>(eq, c.cfg.surface_flux, ubr, ui, o, f);
                for (int v = 0; v < V; ++v) sfv[V * (fn + nf * (dir + 2 * D * e)) + v] = f[v];
            }
        }
    }

    // prolong2mortars! + calc_mortar_flux! (dg_3d.jl:770-1335, dg_2d.jl:897-1243):
    // the large element's face trace is interpolated to each small face
    // (forward_{lower,upper} per free dimension, first index first), the
    // numerical flux is evaluated there (the large element is the left state
    // when large_sides == 1) and given to the small element, and the large
    // element receives the reverse L2 projection summed over the small faces.
    {
        const int ns = 1 << (D - 1);
        const int n1 = D == 3 ? n : 1;  // extent of the second free index
        const Basis &b = c.basis;
#pragma omp parallel for schedule(static)
        for (int m = 0; m < c.M; ++m) {
            const int32_t *ids = &c.mortar_ids[(size_t)m * (ns + 1)];
            const int large = ids[ns], o = c.mortar_orient[m];
            const bool large_left = c.mortar_side[m] == 1;
            const int large_dir = large_left ? 2 * o + 1 : 2 * o, small_dir = large_left ? 2 * o : 2 * o + 1;
            std::vector<R> ul((size_t)nf * V), tmp((size_t)nf * V), up((size_t)nf * V), fs((size_t)ns * nf * V),
                fl((size_t)nf * V, R(0));
            for (int fn = 0; fn < nf; ++fn) {
                const int node = face_to_node<D>(n, o, large_left ? n - 1 : 0, fn);
                for (int v = 0; v < V; ++v) ul[V * fn + v] = u[V * (node + NN * large) + v];
            }
            for (int q = 0; q < ns; ++q) {
                const std::vector<double> &F0 = (q & 1) ? b.fwd_upper : b.fwd_lower;
                const std::vector<double> &F1 = (q >> 1) ? b.fwd_upper : b.fwd_lower;
                for (int j = 0; j < n1; ++j)  // first free index
                    for (int i = 0; i < n; ++i)
                        for (int v = 0; v < V; ++v) {
                            R acc = 0;
                            for (int a = 0; a < n; ++a) acc += R(F0[i * n + a]) * ul[V * (a + n * j) + v];
                            tmp[V * (i + n * j) + v] = acc;
                        }
                for (int j = 0; j < n1; ++j)  // second free index (3D)
                    for (int i = 0; i < n; ++i)
                        for (int v = 0; v < V; ++v) {
                            R acc = 0;
                            if (D == 3)
                                for (int a = 0; a < n; ++a) acc += R(F1[j * n + a]) * tmp[V * (i + n * a) + v];
                            else
                                acc = tmp[V * (i + n * j) + v];
                            up[V * (i + n * j) + v] = acc;
                        }
                const int small = ids[q];
                for (int fn = 0; fn < nf; ++fn) {
                    const int node = face_to_node<D>(n, o, large_left ? 0 : n - 1, fn);
                    const R *us = &u[V * (node + NN * small)];
                    R *f = &fs[(size_t)V * (fn + nf * q)];
                    if (large_left) num_flux<D>(eq, c.cfg.surface_flux, &up[V * fn], us, o, f);
                    else num_flux<D>(eq, c.cfg.surface_flux, us, &up[V * fn], o, f);
                    for (int v = 0; v < V; ++v) sfv[V * (fn + nf * (small_dir + 2 * D * small)) + v] = f[v];
                }
            }
            // mortar_fluxes_to_elements!: reverse projections in the order
            // upper_left, upper_right, lower_left, lower_right (3D), upper, lower (2D)
            const int order3[4] = {2, 3, 0, 1}, order2[2] = {1, 0};
            for (int k = 0; k < ns; ++k) {
                const int q = D == 3 ? order3[k] : order2[k];
                const std::vector<double> &R0 = (q & 1) ? b.rev_upper : b.rev_lower;
                const std::vector<double> &R1 = (q >> 1) ? b.rev_upper : b.rev_lower;
                const R *fq = &fs[(size_t)V * nf * q];
                for (int j = 0; j < n1; ++j)
                    for (int i = 0; i < n; ++i)
                        for (int v = 0; v < V; ++v) {
                            R acc = 0;
                            for (int a = 0; a < n; ++a) acc += R(R0[i * n + a]) * fq[V * (a + n * j) + v];
                            tmp[V * (i + n * j) + v] = acc;
                        }
                for (int j = 0; j < n1; ++j)
                    for (int i = 0; i < n; ++i)
                        for (int v = 0; v < V; ++v) {
                            R acc = 0;
                            if (D == 3)
                                for (int a = 0; a < n; ++a) acc += R(R1[j * n + a]) * tmp[V * (i + n * a) + v];
                            else
                                acc = tmp[V * (i + n * j) + v];
                            fl[V * (i + n * j) + v] += acc;
                        }
            }
            for (int fn = 0; fn < nf; ++fn)
                for (int v = 0; v < V; ++v) sfv[V * (fn + nf * (large_dir + 2 * D * large)) + v] = fl[V * fn + v];
        }
    }

    // calc_surface_integral!: dg_3d.jl:1337-1394, dg_2d.jl:1245-1287
    const R factor = R(c.basis.inv_weights[0]);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < E; ++e) {
        for (int o = 0; o < D; ++o)
            for (int fn = 0; fn < nf; ++fn) {
                int nm = face_to_node<D>(n, o, 0, fn), np = face_to_node<D>(n, o, n - 1, fn);
                for (int v = 0; v < V; ++v) {
                    du[V * (nm + NN * e) + v] = du[V * (nm + NN * e) + v] - sfv[V * (fn + nf * (2 * o + 2 * D * e)) + v] * factor;
                    du[V * (np + NN * e) + v] = du[V * (np + NN * e) + v] + sfv[V * (fn + nf * (2 * o + 1 + 2 * D * e)) + v] * factor;
                }
            }
        // apply_jacobian!: dg_3d.jl:1396-1415
        R fac = -R(c.inverse_jacobian[e]);
        for (int k = 0; k < V * NN; ++k) du[(size_t)V * NN * e + k] *= fac;
        // calc_sources!: dg_3d.jl:1417-1438
        if (c.cfg.source_terms == SRC_CONVERGENCE_TEST) {
            for (int node = 0; node < NN; ++node) {
                double x[3];
                for (int d = 0; d < D; ++d) x[d] = c.node_coords[d + D * (node + NN * e)];
                R s[V];
                source_convergence<D, R>(eq, x, t, s);
                for (int v = 0; v < V; ++v) du[V * (node + NN * e) + v] += s[v];
            }
        }
    }
}

// max_dt: stepsize_dg3d.jl:8-32, stepsize_dg2d.jl:8-30
template <int D> double max_dt(const oracle_ctx &c, const double *u) {
    const int V = D + 2, NN = c.NN;
    Euler<D> eq(c.cfg.gamma);
    double max_scaled_speed = 4.9406564584124654e-324;  // nextfloat(0.0)
    for (int e = 0; e < c.E; ++e) {
        double ml[3] = {0, 0, 0};
        for (int node = 0; node < NN; ++node) {
            double lam[3];
            eq.max_abs_speeds(&u[V * (node + NN * e)], lam);
            for (int d = 0; d < D; ++d) ml[d] = std::max(ml[d], lam[d]);
        }
        double s = 0;
        for (int d = 0; d < D; ++d) s += ml[d];
        max_scaled_speed = std::max(max_scaled_speed, c.inverse_jacobian[e] * s);
    }
    return 2 / (c.n * max_scaled_speed);
}

// calc_error_norms: analysis_dg3d.jl:123-161 (and analysis_dg2d.jl twin), func = cons2cons
template <int D> void error_norms(const oracle_ctx &c, const double *u, double t, double *l2, double *linf) {
    const int V = D + 2, n = c.n, NN = c.NN, na = c.basis.na;
    Euler<D> eq(c.cfg.gamma);
    const double *vdm = c.basis.a_vdm.data();
    const int NA = ipow(na, D);
    for (int v = 0; v < V; ++v) l2[v] = linf[v] = 0;
    std::vector<double> ul(NA * V), xl(NA * D), t1(NA * (V > D ? V : D)), t2(NA * (V > D ? V : D));
    for (int e = 0; e < c.E; ++e) {
        // multiply_dimensionwise! for u and x: interpolate along dims one at a time
        auto interp = [&](const double *in, int nv, double *out) {
            // in: [nv, n^D]; out: [nv, na^D]; stage through t1/t2
            std::vector<int> cur = {n, n, n};
            std::vector<double> a(in, in + (size_t)nv * NN), b;
            for (int o = 0; o < D; ++o) {
                std::vector<int> nxt = cur;
                nxt[o] = na;
                int cnt = 1;
                for (int d = 0; d < D; ++d) cnt *= nxt[d];
                b.assign((size_t)cnt * nv, 0.0);
                for (int idx = 0; idx < cnt; ++idx) {
                    int id[3] = {0, 0, 0};
                    int rem = idx;
                    for (int d = 0; d < D; ++d) {
                        id[d] = rem % nxt[d];
                        rem /= nxt[d];
                    }
                    for (int vv = 0; vv < nv; ++vv) {
                        double res = 0;
                        for (int ii = 0; ii < n; ++ii) {
                            int jd[3] = {id[0], id[1], id[2]};
                            jd[o] = ii;
                            int src = 0, s = 1;
                            for (int d = 0; d < D; ++d) {
                                src += jd[d] * s;
                                s *= cur[d];
                            }
                            res += vdm[id[o] * n + ii] * a[(size_t)src * nv + vv];
                        }
                        b[(size_t)idx * nv + vv] = res;
                    }
                }
                a.swap(b);
                cur = nxt;
            }
            std::copy(a.begin(), a.end(), out);
        };
        interp(&u[(size_t)V * NN * e], V, ul.data());
        interp(&c.node_coords[(size_t)D * NN * e], D, xl.data());
        double vol = std::pow(1.0 / c.inverse_jacobian[e], D);
        for (int q = 0; q < NA; ++q) {
            int id[3] = {q % na, (q / na) % na, q / (na * na)};
            double w = 1;
            for (int d = 0; d < D; ++d) w *= c.basis.a_weights[id[d]];
            double ue[V];
            initial_condition<D>(c, eq, &xl[D * q], t, ue);
            for (int v = 0; v < V; ++v) {
                double diff = ue[v] - ul[V * q + v];
                l2[v] += diff * diff * (w * vol);
                linf[v] = std::max(linf[v], std::fabs(diff));
            }
        }
    }
    double tv = std::pow(c.domain_length, D);
    for (int v = 0; v < V; ++v) l2[v] = std::sqrt(l2[v] / tv);
}

}  // namespace

namespace {

// Cells of the tree addressed by (level, integer coordinates at that level).
// The map holds every cell of the tree: a leaf maps to its element index, an
// inner cell (one with children) to -1 (abstract_tree.jl: has_children,
// has_neighbor, neighbor_ids at the same level).
struct TreeIndex {
    int D;
    const int *periodic;
    std::unordered_map<long long, int> cells;
    TreeIndex(int D_, const int *periodic_) : D(D_), periodic(periodic_) {}
    long long key(int level, const int *ic) const {
        long long k = level;
        for (int d = 0; d < D; ++d) k |= (long long)ic[d] << (6 + 19 * d);
        return k;
    }
    void rebuild(const std::vector<TreeCell> &leaves) {
        cells.clear();
        for (size_t i = 0; i < leaves.size(); ++i) {
            const TreeCell &cl = leaves[i];
            cells[key(cl.level, cl.ic)] = (int)i;
            int ic[3] = {cl.ic[0], cl.ic[1], cl.ic[2]};
            for (int l = cl.level - 1; l >= 0; --l) {
                for (int d = 0; d < D; ++d) ic[d] >>= 1;
                if (!cells.emplace(key(l, ic), -1).second) break;
            }
        }
    }
    const int *find(int level, const int *ic) const {
        auto it = cells.find(key(level, ic));
        return it == cells.end() ? nullptr : &it->second;
    }
    // same-level cell coordinates across face dir (0-based: 2o = -o, 2o+1 = +o),
    // wrapping periodic dimensions; false outside a non-periodic domain
    bool neighbour(int level, const int *ic, int dir, int *out) const {
        const int o = dir / 2, nside = 1 << level;
        for (int d = 0; d < D; ++d) out[d] = ic[d];
        int k = ic[o] + (dir % 2 ? 1 : -1);
        if ((k < 0 || k >= nside) && !periodic[o]) return false;
        out[o] = (k + nside) % nside;
        return true;
    }
};

// refine_unbalanced!, abstract_tree.jl:317-353: every marked leaf is replaced
// by its 2^D children (child index bits: x = 1, y = 2, z = 4; child_sign
// :118-128, child coordinates :668-674), inserted in place (depth-first).
void refine_marked(int D, double root_length, std::vector<TreeCell> &leaves, std::vector<Origin> &orig,
                   const std::vector<char> &mark) {
    std::vector<TreeCell> out;
    std::vector<Origin> oout;
    for (size_t i = 0; i < leaves.size(); ++i) {
        const TreeCell &p = leaves[i];
        if (!mark[i]) {
            out.push_back(p);
            oout.push_back(orig[i]);
            continue;
        }
        const double child_length = root_length / (1 << p.level) / 2;
        for (int child = 0; child < (1 << D); ++child) {
            TreeCell ch{p.level + 1, {0, 0, 0}, {0, 0, 0}};
            for (int d = 0; d < D; ++d) {
                const int bit = (child >> d) & 1;
                ch.center[d] = p.center[d] + (bit ? 1 : -1) * child_length / 2;
                ch.ic[d] = 2 * p.ic[d] + bit;
            }
            out.push_back(ch);
            oout.push_back(Origin{1 + child, orig[i].src});
        }
    }
    leaves.swap(out);
    orig.swap(oout);
}

// refine! + rebalance!, abstract_tree.jl:367-400, 453-490: refine the marked
// leaves, then, as long as cells were refined, refine every coarse neighbour
// (the parent's face neighbour) of a refined cell that has no same-level face
// neighbour (Cartesian neighbours only). Returns the number of refined cells.
int refine_balanced(int D, double root_length, TreeIndex &idx, std::vector<TreeCell> &leaves,
                    std::vector<Origin> &orig, std::vector<char> mark) {
    int count = 0;
    while (true) {
        std::vector<TreeCell> refined;
        for (size_t i = 0; i < leaves.size(); ++i)
            if (mark[i]) refined.push_back(leaves[i]);
        if (refined.empty()) break;
        count += (int)refined.size();
        refine_marked(D, root_length, leaves, orig, mark);
        idx.rebuild(leaves);
        mark.assign(leaves.size(), 0);
        for (const TreeCell &p : refined)
            for (int dir = 0; dir < 2 * D; ++dir) {
                int nb[3];
                if (idx.neighbour(p.level, p.ic, dir, nb) && idx.find(p.level, nb)) continue;  // has_neighbor
                if (p.level == 0) continue;
                int par[3] = {0, 0, 0}, pnb[3];
                for (int d = 0; d < D; ++d) par[d] = p.ic[d] >> 1;
                if (!idx.neighbour(p.level - 1, par, dir, pnb)) continue;
                const int *f = idx.find(p.level - 1, pnb);  // has_coarse_neighbor
                if (f && *f >= 0) mark[*f] = 1;
            }
    }
    return count;
}

// AMRCallback coarsening (amr.jl:304-370) and coarsen!(tree, cell_ids)
// (abstract_tree.jl:516-600): a parent is coarsened when all 2^D children are
// leaves marked for coarsening, unless a child has a non-sibling same-level
// face neighbour with children (that would break the 2:1 balance). Parents
// are taken from the highest level down, each check on the tree as already
// coarsened. Returns the number of coarsened parents.
int coarsen_marked(int D, double root_length, TreeIndex &idx, std::vector<TreeCell> &leaves,
                   std::vector<Origin> &orig, const std::vector<long long> &leaf_keys) {
    const int nch = 1 << D;
    std::unordered_map<long long, int> count;
    std::vector<std::pair<int, std::array<int, 3>>> parents;
    for (long long k : leaf_keys) {
        const int level = (int)(k & 63);
        if (level == 0) continue;  // has_parent
        int ic[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) ic[d] = (int)((k >> (6 + 19 * d)) & ((1 << 19) - 1));
        const int *f = idx.find(level, ic);
        if (!f || *f < 0) continue;  // is_leaf (a cell refined by rebalancing is skipped)
        std::array<int, 3> pic{{0, 0, 0}};
        for (int d = 0; d < D; ++d) pic[d] = ic[d] >> 1;
        if (++count[idx.key(level - 1, pic.data())] == nch) parents.push_back({level - 1, pic});
    }
    std::stable_sort(parents.begin(), parents.end(),
                     [](const std::pair<int, std::array<int, 3>> &a, const std::pair<int, std::array<int, 3>> &b) {
                         return a.first > b.first;
                     });
    std::unordered_map<long long, char> coarsened;
    for (const auto &pc : parents) {
        const int level = pc.first;
        const int *pic = pc.second.data();
        bool skip = false;
        for (int q = 0; q < nch && !skip; ++q) {
            int ch[3] = {0, 0, 0};
            for (int d = 0; d < D; ++d) ch[d] = 2 * pic[d] + ((q >> d) & 1);
            for (int dir = 0; dir < 2 * D; ++dir) {
                const int bit = (q >> (dir / 2)) & 1;
                if ((dir % 2 == 0 && bit == 1) || (dir % 2 == 1 && bit == 0)) continue;  // has_sibling
                int nb[3];
                if (!idx.neighbour(level + 1, ch, dir, nb)) continue;
                const int *f = idx.find(level + 1, nb);
                if (f && *f < 0) {  // neighbour with children
                    skip = true;
                    break;
                }
            }
        }
        if (skip) continue;
        for (int q = 0; q < nch; ++q) {
            int ch[3] = {0, 0, 0};
            for (int d = 0; d < D; ++d) ch[d] = 2 * pic[d] + ((q >> d) & 1);
            idx.cells.erase(idx.key(level + 1, ch));
        }
        const long long pk = idx.key(level, pic);
        idx.cells[pk] = 0;  // now a leaf
        coarsened[pk] = 1;
    }
    if (coarsened.empty()) return 0;
    std::vector<TreeCell> out;
    std::vector<Origin> oout;
    for (size_t i = 0; i < leaves.size(); ++i) {
        const TreeCell &cl = leaves[i];
        if (cl.level > 0) {
            int pic[3] = {0, 0, 0};
            for (int d = 0; d < D; ++d) pic[d] = cl.ic[d] >> 1;
            if (coarsened.count(idx.key(cl.level - 1, pic))) {
                // the children follow their parent in depth-first order, child 0 first
                const double child_length = root_length / (1 << cl.level);
                TreeCell p{cl.level - 1, {pic[0], pic[1], pic[2]}, {0, 0, 0}};
                for (int d = 0; d < D; ++d) p.center[d] = cl.center[d] + child_length / 2;
                out.push_back(p);
                oout.push_back(Origin{-1, orig[i].src});
                i += nch - 1;
                continue;
            }
        }
        out.push_back(cl);
        oout.push_back(orig[i]);
    }
    leaves.swap(out);
    orig.swap(oout);
    idx.rebuild(leaves);
    return (int)coarsened.size();
}

// Everything the discretisation needs from the leaves: element geometry,
// interfaces, L2 mortars and boundaries. False if the tree is not 2:1
// balanced or has boundaries without a boundary condition.
bool finalize_mesh(oracle_ctx *c) {
    const int D = c->D;
    TreeIndex idx(D, c->cfg.periodicity);
    idx.rebuild(c->leaves);
    c->mortar_ids.clear();
    c->mortar_orient.clear();
    c->mortar_side.clear();
    c->bnd_elem.clear();
    c->bnd_dir.clear();
    c->E = (int)c->leaves.size();
    c->centers.clear();
    c->lengths.clear();
    c->levels.clear();
    for (const TreeCell &cl : c->leaves) {
        for (int d = 0; d < D; ++d) c->centers.push_back(cl.center[d]);
        c->lengths.push_back(c->domain_length / (1 << cl.level));
        c->levels.push_back(cl.level);
    }
    // neighbours, mortars, boundaries (init_neighbors!, init_interfaces!,
    // init_mortars!, init_boundaries!): a same-level leaf across a face is an
    // interface neighbour; a refined same-level cell makes a mortar with this
    // element as the large side; no same-level cell means this element is a
    // small element of the coarser neighbour's mortar; outside a non-periodic
    // domain the face is a boundary.
    c->elem_nbr.assign(2 * D * c->E, -1);
    std::vector<char> is_boundary(2 * D * c->E, 0);
    const int ns = 1 << (D - 1);
    for (int e = 0; e < c->E; ++e) {
        const TreeCell &cl = c->leaves[e];
        for (int dir = 0; dir < 2 * D; ++dir) {
            int nb[3];
            if (!idx.neighbour(cl.level, cl.ic, dir, nb)) {
                is_boundary[dir + 2 * D * e] = 1;
                continue;
            }
            const int *f = idx.find(cl.level, nb);
            if (!f) continue;  // small side of a mortar
            if (*f >= 0) {
                c->elem_nbr[dir + 2 * D * e] = *f;
                continue;
            }
            const int o = dir / 2, plus = dir % 2;
            int fd[2] = {0, 0}, r = 0;
            for (int d = 0; d < D; ++d)
                if (d != o) fd[r++] = d;
            for (int q = 0; q < ns; ++q) {
                int ch[3];
                for (int d = 0; d < D; ++d) ch[d] = 2 * nb[d];
                ch[o] += plus ? 0 : 1;
                ch[fd[0]] += q & 1;
                if (D == 3) ch[fd[1]] += q >> 1;
                const int *s = idx.find(cl.level + 1, ch);
                c->mortar_ids.push_back(s ? *s : -1);
            }
            c->mortar_ids.push_back(e);
            c->mortar_orient.push_back(o);
            c->mortar_side.push_back(plus ? 1 : 2);
        }
    }
    c->M = (int)c->mortar_orient.size();
    for (int32_t id : c->mortar_ids)
        if (id < 0) return false;  // a mortar partner that is not a leaf: the tree is not 2:1 balanced
    // init_boundaries!: containers_2d.jl:475-543 (direction-major, element order)
    for (int dir = 0; dir < 2 * D; ++dir)
        for (int e = 0; e < c->E; ++e)
            if (is_boundary[dir + 2 * D * e]) {
                c->bnd_elem.push_back(e);
                c->bnd_dir.push_back(dir);
            }
    if (!c->bnd_elem.empty() && c->cfg.boundary_condition != BC_DIRICHLET_IC &&
        c->cfg.boundary_condition != BC_SLIP_WALL)
        return false;
    // init_elements!: containers_3d.jl:87-128
    c->inverse_jacobian.resize(c->E);
    c->node_coords.resize((size_t)c->D * c->NN * c->E);
    double reference_length = 0;
    for (double w : c->basis.weights) reference_length += w;
    double reference_offset = (c->basis.nodes.front() + c->basis.nodes.back()) / 2;
    for (int e = 0; e < c->E; ++e) {
        double dx = c->lengths[e];
        double jacobian = dx / reference_length;
        c->inverse_jacobian[e] = 1.0 / jacobian;
        for (int node = 0; node < c->NN; ++node) {
            int idx[3] = {node % c->n, (node / c->n) % c->n, node / (c->n * c->n)};
            for (int d = 0; d < c->D; ++d)
                c->node_coords[d + c->D * (node + c->NN * e)] =
                    c->centers[c->D * e + d] + jacobian * (c->basis.nodes[idx[d]] - reference_offset);
        }
    }
    // init_interfaces!: containers_3d.jl:229-279 (positive directions, element order)
    c->neighbor_ids.clear();
    c->orientations.clear();
    for (int e = 0; e < c->E; ++e)
        for (int o = 0; o < c->D; ++o) {
            if (c->elem_nbr[2 * o + 1 + 2 * c->D * e] < 0) continue;
            c->neighbor_ids.push_back(e);
            c->neighbor_ids.push_back(c->elem_nbr[2 * o + 1 + 2 * c->D * e]);
            c->orientations.push_back(o);
        }
    c->I = (int)c->orientations.size();
    c->sfv.assign((size_t)c->V * c->nf * 2 * c->D * c->E, 0.0);
    c->alpha.assign(c->E, 0.0);
    return true;
}

// Solution transfer of the L2 adaptor (LobattoLegendreAdaptorL2,
// basis_lobatto_legendre.jl:334-360, the mortar matrices): a child q of a
// refined element is the tensor interpolation forward_{lower,upper} per
// dimension (bit 0 / 1 of q), refine_element!, amr_dg3d.jl:138-230 (2D:
// amr_dg2d.jl); a coarsened element is the sum over its 2^D children of the
// tensor reverse_{lower,upper} projections, coarsen_elements!,
// amr_dg3d.jl:337-430. Contractions in x, then y, then z
// (multiply_dimensionwise!, interpolation.jl:297-420).
void project_solution(const oracle_ctx &c, const double *u_old, const std::vector<Origin> &orig,
                      std::vector<double> &u_new) {
    const int D = c.D, n = c.n, NN = c.NN, V = c.V, nch = 1 << D;
    const size_t per = (size_t)V * NN;
    u_new.assign(per * orig.size(), 0.0);
    auto apply = [&](const double *in, double *out, const int *which, bool forward, bool add) {
        // which[d] = 0 lower, 1 upper
        std::vector<double> a(in, in + per), b(per);
        for (int d = 0; d < D; ++d) {
            const std::vector<double> &M = forward ? (which[d] ? c.basis.fwd_upper : c.basis.fwd_lower)
                                                   : (which[d] ? c.basis.rev_upper : c.basis.rev_lower);
            const int stride = ipow(n, d);
            for (int node = 0; node < NN; ++node) {
                const int i = (node / stride) % n, base = node - i * stride;
                for (int v = 0; v < V; ++v) {
                    double r = 0;
                    for (int ii = 0; ii < n; ++ii) r += M[i * n + ii] * a[v + V * (base + ii * stride)];
                    b[v + V * node] = r;
                }
            }
            a.swap(b);
        }
        for (size_t k = 0; k < per; ++k) out[k] = add ? out[k] + a[k] : a[k];
    };
#pragma omp parallel for schedule(static)
    for (long e = 0; e < (long)orig.size(); ++e) {
        const Origin o = orig[e];
        double *out = &u_new[per * e];
        if (o.kind == 0) {
            std::copy(u_old + per * o.src, u_old + per * (o.src + 1), out);
        } else if (o.kind > 0) {
            const int q = o.kind - 1;
            int which[3] = {q & 1, (q >> 1) & 1, (q >> 2) & 1};
            apply(u_old + per * o.src, out, which, true, false);
        } else {
            for (int q = 0; q < nch; ++q) {
                int which[3] = {q & 1, (q >> 1) & 1, (q >> 2) & 1};
                apply(u_old + per * (o.src + q), out, which, false, q > 0);
            }
        }
    }
}

}  // namespace

extern "C" {

void oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int oracle_get_threads(void) { return omp_get_max_threads(); }

oracle_ctx *oracle_create(const oracle_config *cfg) {
    if (!cfg || (cfg->ndims != 2 && cfg->ndims != 3) || cfg->polydeg < 1 || cfg->polydeg > 15)
        return nullptr;
    oracle_ctx *c = new oracle_ctx();
    c->cfg = *cfg;
    c->D = cfg->ndims;
    c->n = cfg->polydeg + 1;
    c->NN = ipow(c->n, c->D);
    c->nf = ipow(c->n, c->D - 1);
    c->V = c->D + 2;
    c->basis.init(cfg->polydeg);
    c->hg_variable = cfg->hg_variable;
    // TreeMesh(coordinates_min, coordinates_max; initial_refinement_level,
    // refinement_patches, periodicity), tree_mesh.jl:128-215: uniform refinement,
    // then every box patch refines the leaf cells whose centre lies strictly
    // inside the box (refine_box!, abstract_tree.jl:405-418), rebalancing to
    // 2:1 (refine!, rebalance!, :356-490). Children are inserted right behind
    // their parent, so the leaves stay in depth-first order (:317-353).
    const int D = c->D;
    double center[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d) center[d] = (cfg->coordinates_min[d] + cfg->coordinates_max[d]) / 2;
    const double length = cfg->coordinates_max[0] - cfg->coordinates_min[0];
    c->domain_length = length;
    {
        std::vector<double> cen, len;
        std::vector<int> icoords;
        int ic0[3] = {0, 0, 0};
        build_leaves(D, 0, cfg->initial_refinement_level, center, length, ic0, cen, len, icoords);
        for (size_t e = 0; e < len.size(); ++e) {
            TreeCell cl{cfg->initial_refinement_level, {0, 0, 0}, {0, 0, 0}};
            for (int d = 0; d < D; ++d) {
                cl.ic[d] = icoords[D * e + d];
                cl.center[d] = cen[D * e + d];
            }
            c->leaves.push_back(cl);
        }
    }
    TreeIndex idx(D, c->cfg.periodicity);
    idx.rebuild(c->leaves);
    std::vector<Origin> orig;
    for (int pt = 0; pt < cfg->npatches; ++pt) {
        std::vector<char> mark(c->leaves.size(), 0);
        for (size_t i = 0; i < c->leaves.size(); ++i) {
            bool inside = true;
            for (int d = 0; d < D; ++d)
                inside = inside && cfg->patch_min[pt][d] < c->leaves[i].center[d] &&
                         cfg->patch_max[pt][d] > c->leaves[i].center[d];
            mark[i] = inside;
        }
        orig.assign(c->leaves.size(), Origin{0, 0});
        refine_balanced(D, length, idx, c->leaves, orig, mark);
    }
    if (!finalize_mesh(c)) {
        delete c;
        return nullptr;
    }
    return c;
}

void oracle_destroy(oracle_ctx *c) { delete c; }

// IndicatorMax(semi; variable), indicators_3d.jl:239-259 (2D: indicators_2d.jl):
// alpha[e] = maximum of conserved variable `variable` over the nodes of element e.
void oracle_indicator_max(const oracle_ctx *c, const double *u, int variable, double *alpha) {
    for (int e = 0; e < c->E; ++e) {
        double m = u[variable + c->V * (size_t)c->NN * e];
        for (int node = 1; node < c->NN; ++node) m = std::max(m, u[variable + c->V * (node + (size_t)c->NN * e)]);
        alpha[e] = m;
    }
}

// One AMRCallback adaptation of a TreeMesh (amr.jl:244-392) with the controller
// values `lambda` (> 0 refine, < 0 coarsen, per element): refinement with 2:1
// rebalancing, then coarsening of the parents whose children are all still
// marked leaves. Returns 1 if the mesh changed, 0 if not, -1 if the adapted
// tree could not be discretised. The projected solution is read back with
// oracle_get_adapted_u and the element origins with oracle_get_transfer.
int oracle_amr_adapt(oracle_ctx *c, const double *u, const int32_t *lambda, int only_refine, int only_coarsen) {
    const int D = c->D;
    TreeIndex idx(D, c->cfg.periodicity);
    idx.rebuild(c->leaves);
    std::vector<Origin> orig(c->E);
    for (int e = 0; e < c->E; ++e) orig[e] = Origin{0, e};
    std::vector<long long> to_coarsen;
    std::vector<char> to_refine(c->E, 0);
    bool any_refine = false;
    for (int e = 0; e < c->E; ++e) {
        if (lambda[e] > 0) to_refine[e] = 1, any_refine = true;
        else if (lambda[e] < 0) to_coarsen.push_back(idx.key(c->leaves[e].level, c->leaves[e].ic));
    }
    int changed = 0;
    if (!only_coarsen && any_refine)
        changed += refine_balanced(D, c->domain_length, idx, c->leaves, orig, to_refine);
    if (!only_refine && !to_coarsen.empty())
        changed += coarsen_marked(D, c->domain_length, idx, c->leaves, orig, to_coarsen);
    project_solution(*c, u, orig, c->u_adapted);
    c->transfer = orig;
    if (!changed) return 0;
    return finalize_mesh(c) ? 1 : -1;
}

void oracle_get_adapted_u(const oracle_ctx *c, double *u) { std::copy(c->u_adapted.begin(), c->u_adapted.end(), u); }

void oracle_get_transfer(const oracle_ctx *c, int32_t *kind, int32_t *src) {
    for (size_t e = 0; e < c->transfer.size(); ++e) {
        kind[e] = c->transfer[e].kind;
        src[e] = c->transfer[e].src;
    }
}

void oracle_get_levels(const oracle_ctx *c, int32_t *levels) { std::copy(c->levels.begin(), c->levels.end(), levels); }

int oracle_nelements(const oracle_ctx *c) { return c->E; }
int oracle_ninterfaces(const oracle_ctx *c) { return c->I; }
int oracle_nboundaries(const oracle_ctx *c) { return (int)c->bnd_elem.size(); }
int oracle_nmortars(const oracle_ctx *c) { return c->M; }
// mortars container (0-based): neighbor_ids [M][2^(D-1) + 1], large_sides, orientations
void oracle_get_mortars(const oracle_ctx *c, int32_t *ids, int32_t *large_sides, int32_t *orientations) {
    if (ids) std::copy(c->mortar_ids.begin(), c->mortar_ids.end(), ids);
    if (large_sides) std::copy(c->mortar_side.begin(), c->mortar_side.end(), large_sides);
    if (orientations) std::copy(c->mortar_orient.begin(), c->mortar_orient.end(), orientations);
}
// L2 mortar matrices, row-major n x n each: forward_upper, forward_lower, reverse_upper, reverse_lower
void oracle_get_mortar_l2(const oracle_ctx *c, double *fu, double *fl, double *ru, double *rl) {
    const Basis &b = c->basis;
    std::copy(b.fwd_upper.begin(), b.fwd_upper.end(), fu);
    std::copy(b.fwd_lower.begin(), b.fwd_lower.end(), fl);
    std::copy(b.rev_upper.begin(), b.rev_upper.end(), ru);
    std::copy(b.rev_lower.begin(), b.rev_lower.end(), rl);
}
// boundaries container (0-based): element, orientation, neighbor_side (1: element
// left of the boundary, 2: right), as init_boundaries! fills them
void oracle_get_boundaries(const oracle_ctx *c, int32_t *elem, int32_t *orientation, int32_t *side) {
    for (size_t b = 0; b < c->bnd_elem.size(); ++b) {
        if (elem) elem[b] = c->bnd_elem[b];
        if (orientation) orientation[b] = c->bnd_dir[b] / 2;
        if (side) side[b] = c->bnd_dir[b] % 2 == 1 ? 1 : 2;
    }
}
int oracle_nnodes(const oracle_ctx *c) { return c->n; }
long oracle_ndofs(const oracle_ctx *c) { return (long)c->NN * c->E; }

// Mesh / geometry export (0-based ids/orientations).
void oracle_get_mesh(const oracle_ctx *c, int32_t *neighbor_ids, int32_t *orientations,
                     double *inverse_jacobian, double *centers, double *node_coords,
                     int32_t *elem_nbr) {
    if (neighbor_ids) std::copy(c->neighbor_ids.begin(), c->neighbor_ids.end(), neighbor_ids);
    if (orientations) std::copy(c->orientations.begin(), c->orientations.end(), orientations);
    if (inverse_jacobian) std::copy(c->inverse_jacobian.begin(), c->inverse_jacobian.end(), inverse_jacobian);
    if (centers) std::copy(c->centers.begin(), c->centers.end(), centers);
    if (node_coords) std::copy(c->node_coords.begin(), c->node_coords.end(), node_coords);
    if (elem_nbr) std::copy(c->elem_nbr.begin(), c->elem_nbr.end(), elem_nbr);
}

// Basis export: matrices row-major [r*n + c] = M[r, c].
void oracle_get_basis(const oracle_ctx *c, double *nodes, double *weights, double *D,
                      double *Dsplit, double *Dhat, double *Vinv) {
    const Basis &b = c->basis;
    if (nodes) std::copy(b.nodes.begin(), b.nodes.end(), nodes);
    if (weights) std::copy(b.weights.begin(), b.weights.end(), weights);
    if (D) std::copy(b.D.begin(), b.D.end(), D);
    if (Dsplit) std::copy(b.Dsplit.begin(), b.Dsplit.end(), Dsplit);
    if (Dhat) std::copy(b.Dhat.begin(), b.Dhat.end(), Dhat);
    if (Vinv) std::copy(b.Vinv.begin(), b.Vinv.end(), Vinv);
}

void oracle_set_initial_condition(oracle_ctx *c, int kind, const double *params, int nparams,
                                  const double *modes) {
    c->ic_kind = kind;
    for (int i = 0; i < 8; ++i) c->ic_params[i] = i < nparams ? params[i] : 0.0;
    if (kind == IC_FOURIER_MODES && modes) c->ic_modes.assign(modes, modes + 4 * 8 * 4);
}

// compute_coefficients!: dg.jl:1280-1325 (pointwise initial condition at nodes)
void oracle_compute_coefficients(const oracle_ctx *c, double t, double *u) {
    for (int e = 0; e < c->E; ++e)
        for (int node = 0; node < c->NN; ++node) {
            const double *x = &c->node_coords[c->D * (node + c->NN * e)];
            double *un = &u[c->V * (node + c->NN * e)];
            if (c->D == 2) initial_condition<2>(*c, Euler<2>(c->cfg.gamma), x, t, un);
            else initial_condition<3>(*c, Euler<3>(c->cfg.gamma), x, t, un);
        }
}

void oracle_rhs(oracle_ctx *c, double *du, const double *u, double t) {
    if (c->D == 2) rhs<2, double>(*c, du, u, t);
    else rhs<3, double>(*c, du, u, t);
}

// The same operator evaluated in 80-bit extended precision (long double) on
// the given FP64 inputs; the result is rounded to FP64. Used to measure the
// round-off of FP64 evaluations (cancellation in du grows as the mesh is refined).
void oracle_rhs_ld(oracle_ctx *c, double *du, const double *u, double t) {
    const size_t total = (size_t)c->V * c->NN * c->E;
    std::vector<long double> ul(u, u + total), dul(total);
    if (c->D == 2) rhs<2, long double>(*c, dul.data(), ul.data(), t);
    else rhs<3, long double>(*c, dul.data(), ul.data(), t);
    for (size_t k = 0; k < total; ++k) du[k] = (double)dul[k];
}

// Blending factors (after smoothing) that the shock-capturing RHS uses instead
// of its HG indicator; alpha == NULL returns to the HG indicator.
void oracle_set_alpha(oracle_ctx *c, const double *alpha) {
    c->alpha_given = alpha != nullptr;
    if (alpha) c->alpha_given_v.assign(alpha, alpha + c->E);
    else c->alpha_given_v.clear();
}

// Indicator only (alpha[E]); runs the HG indicator regardless of volume integral.
void oracle_indicator_hg(oracle_ctx *c, const double *u, double *alpha) {
    if (c->D == 2) hg_indicator<2>(*c, Euler<2>(c->cfg.gamma), u);
    else hg_indicator<3>(*c, Euler<3>(c->cfg.gamma), u);
    std::copy(c->alpha.begin(), c->alpha.end(), alpha);
}

// IndicatorHennemannGassner with its own parameters (an AMR indicator, e.g.
// elixir_euler_kelvin_helmholtz_instability_amr.jl): variable 0 density_pressure,
// 1 density, 2 pressure.
void oracle_indicator_hg_params(oracle_ctx *c, const double *u, double alpha_max, double alpha_min, int alpha_smooth,
                                int variable, double *alpha) {
    const oracle_config saved = c->cfg;
    const int saved_var = c->hg_variable;
    c->cfg.alpha_max = alpha_max;
    c->cfg.alpha_min = alpha_min;
    c->cfg.alpha_smooth = alpha_smooth;
    c->hg_variable = variable;
    oracle_indicator_hg(c, u, alpha);
    c->cfg = saved;
    c->hg_variable = saved_var;
}

double oracle_max_dt(const oracle_ctx *c, const double *u) {
    return c->D == 2 ? max_dt<2>(*c, u) : max_dt<3>(*c, u);
}

void oracle_error_norms(const oracle_ctx *c, const double *u, double t, double *l2, double *linf) {
    if (c->D == 2) error_norms<2>(*c, u, t, l2, linf);
    else error_norms<3>(*c, u, t, l2, linf);
}

// integrate_via_indices with normalize (analysis_dg3d.jl:364-386): returns the
// volume-normalised integral of each conserved variable and of entropy_math.
void oracle_integrals(const oracle_ctx *c, const double *u, double *cons_int, double *entropy_int) {
    const int V = c->V, n = c->n, NN = c->NN, D = c->D;
    std::vector<double> acc(V, 0.0);
    double S = 0;
    Euler<2> e2(c->cfg.gamma);
    Euler<3> e3(c->cfg.gamma);
    for (int e = 0; e < c->E; ++e) {
        double vol = std::pow(1.0 / c->inverse_jacobian[e], D);
        for (int node = 0; node < NN; ++node) {
            int idx[3] = {node % n, (node / n) % n, node / (n * n)};
            double w = vol;
            for (int d = 0; d < D; ++d) w *= c->basis.weights[idx[d]];
            const double *un = &u[V * (node + NN * e)];
            for (int v = 0; v < V; ++v) acc[v] += w * un[v];
            S += w * (D == 2 ? e2.entropy_math(un) : e3.entropy_math(un));
        }
    }
    double tv = std::pow(c->domain_length, D);
    for (int v = 0; v < V; ++v) cons_int[v] = acc[v] / tv;
    *entropy_int = S / tv;
}

// Pointwise functions (for unit tests of the restatement itself).
double oracle_ln_mean(double x, double y) { return ln_mean(x, y); }
double oracle_inv_ln_mean(double x, double y) { return inv_ln_mean(x, y); }

// Two-point numerical flux `kind` (0 LLF, 1 LLF naive, 2 ranocha, 3 central) or,
// for kind < 0, the physical flux of ul, in orientation o (0-based).
void oracle_flux(int ndims, int kind, double gamma, const double *ul, const double *ur, int o, double *f) {
    if (ndims == 2) {
        Euler<2> eq(gamma);
        if (kind < 0) eq.flux(ul, o, f);
        else num_flux<2>(eq, kind, ul, ur, o, f);
    } else {
        Euler<3> eq(gamma);
        if (kind < 0) eq.flux(ul, o, f);
        else num_flux<3>(eq, kind, ul, ur, o, f);
    }
}

// boundary_condition_slip_wall flux of the inner state u at a face of
// orientation o (0-based); plus != 0: the element is left of the wall (+ face).
void oracle_slip_wall_flux(int ndims, double gamma, const double *u, int o, int plus, double *f) {
    if (ndims == 2) slip_wall_flux<2>(Euler<2>(gamma), u, o, plus != 0, f);
    else slip_wall_flux<3>(Euler<3>(gamma), u, o, plus != 0, f);
}

void oracle_cons2entropy(int ndims, double gamma, const double *u, double *w) {
    if (ndims == 2) Euler<2>(gamma).cons2entropy(u, w);
    else Euler<3>(gamma).cons2entropy(u, w);
}

// CarpenterKennedy2N54 coefficients: methods_2N.jl:24-48
static const double CK_A[5] = {0.0, 567301805773.0 / 1357537059087.0,
                               2404267990393.0 / 2016746695238.0,
                               3550918686646.0 / 2091501179385.0,
                               1275806237668.0 / 842570457699.0};
static const double CK_B[5] = {1432997174477.0 / 9575080441755.0,
                               5161836677717.0 / 13612068292357.0,
                               1720146321549.0 / 2090206949498.0,
                               3134564353537.0 / 4481467310338.0,
                               2277821191437.0 / 14882151754819.0};
static const double CK_C[5] = {0.0, 1432997174477.0 / 9575080441755.0,
                               2526269341429.0 / 6820363962896.0,
                               2006345519317.0 / 3224310063776.0,
                               2802321613138.0 / 2924317926251.0};

void oracle_ck2n54_coefficients(double *a, double *b, double *cc) {
    for (int i = 0; i < 5; ++i) {
        a[i] = CK_A[i];
        b[i] = CK_B[i];
        cc[i] = CK_C[i];
    }
}

// One 2N low-storage step (step!, methods_2N.jl:131-168).
void oracle_rk_step(oracle_ctx *c, double *u, double *du, double *u_tmp, double t, double dt) {
    const size_t total = (size_t)c->V * c->NN * c->E;
    std::memset(u_tmp, 0, total * sizeof(double));
    for (int s = 0; s < 5; ++s) {
        double t_stage = t + dt * CK_C[s];
        oracle_rhs(c, du, u, t_stage);
        double a_stage = CK_A[s], b_stage_dt = CK_B[s] * dt;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)total; ++i) {
            u_tmp[i] = du[i] - u_tmp[i] * a_stage;
            u[i] += u_tmp[i] * b_stage_dt;
        }
    }
}

// solve(ode, CarpenterKennedy2N54; callback = StepsizeCallback(cfl)) from t0 to t_end:
// stepsize callback at initialisation and after each step (stepsize.jl:71-124),
// limit_dt! at each step (time_integration.jl:46-55). Returns the number of steps,
// or -1 if dt became NaN. maxiters <= 0 means unlimited.
int oracle_solve(oracle_ctx *c, double *u, double t0, double t_end, double cfl, int maxiters,
                 double *t_out) {
    const size_t total = (size_t)c->V * c->NN * c->E;
    std::vector<double> du(total), u_tmp(total);
    double t = t0;
    double dt = cfl * oracle_max_dt(c, u);
    int iter = 0;
    while (true) {
        if (std::isnan(dt)) return -1;
        bool final_step = false;
        double te = t + dt;
        if (te > t_end || te == t_end || std::fabs(te - t_end) <= 1.4901161193847656e-08 * std::max(std::fabs(te), std::fabs(t_end))) {
            dt = t_end - t;
            final_step = true;
        }
        oracle_rk_step(c, u, du.data(), u_tmp.data(), t, dt);
        ++iter;
        t += dt;
        dt = cfl * oracle_max_dt(c, u);
        if (final_step || (maxiters > 0 && iter >= maxiters)) break;
    }
    if (t_out) *t_out = t;
    return iter;
}

}  // extern "C"
