"""Generate the golden fixtures of tests/golden with mpmath (50 digits).

These vectors are computed from the mathematics, not from Trixi.jl (which is
Julia and not runnable here):
  basis_mpmath.json   LGL nodes (+-1 and the roots of P_N') and weights
                      2 / (N (N+1) P_N(x)^2) for N = 1..7
  ln_mean_mpmath.json (y - x) / (ln y - ln x) and its inverse on a grid that
                      straddles the series threshold f^2 = ((y-x)/(y+x))^2 = 1e-4
  flux_ranocha_mpmath.json  Ranocha's entropy-conserving flux (Ranocha 2018,
                      flux_ranocha in src/equations/compressible_euler_{2,3}d.jl)
                      evaluated with exact logarithmic means on seeded random states
Run: python tests/golden/make_fixtures.py
"""
import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50
HERE = os.path.dirname(os.path.abspath(__file__))


def legendre_coeffs(N):
    # exact coefficients (highest degree first) from (n+1) P_{n+1} = (2n+1) x P_n - n P_{n-1}
    p0, p1 = [mp.mpf(1)], [mp.mpf(1), mp.mpf(0)]
    if N == 0:
        return p0
    for n in range(1, N):
        a = [(2 * n + 1) * c for c in p1] + [mp.mpf(0)]
        b = [mp.mpf(0), mp.mpf(0)] + [n * c for c in p0]
        p0, p1 = p1, [(x - y) / (n + 1) for x, y in zip(a, b)]
    return p1


def lgl(N):
    c = legendre_coeffs(N)
    deg = len(c) - 1
    dc = [c[i] * (deg - i) for i in range(deg)]
    interior = [] if N < 2 else [mp.re(r) for r in mp.polyroots(dc, maxsteps=200, extraprec=200)]
    nodes = sorted([mp.mpf(-1)] + interior + [mp.mpf(1)])
    P = lambda x: mp.legendre(N, x)
    weights = [mp.mpf(2) / (N * (N + 1) * P(x) ** 2) for x in nodes]
    return nodes, weights


def lnmean(x, y):
    x, y = mp.mpf(x), mp.mpf(y)
    if x == y:
        return x
    return (y - x) / (mp.log(y) - mp.log(x))


def ranocha(u_ll, u_rr, o, D, gamma=mp.mpf("1.4")):
    def prim(u):
        rho = u[0]
        v = [u[1 + d] / rho for d in range(D)]
        p = (gamma - 1) * (u[D + 1] - mp.mpf(1) / 2 * sum(u[1 + d] * v[d] for d in range(D)))
        return rho, v, p
    rl, vl, pl = prim(u_ll)
    rr, vr, pr = prim(u_rr)
    rho_mean = lnmean(rl, rr)
    inv_rho_p_mean = pl * pr / lnmean(rl * pr, rr * pl)
    vavg = [(vl[d] + vr[d]) / 2 for d in range(D)]
    p_avg = (pl + pr) / 2
    vsq = sum(vl[d] * vr[d] for d in range(D)) / 2
    f1 = rho_mean * vavg[o]
    f = [f1] + [f1 * vavg[d] + (p_avg if d == o else 0) for d in range(D)]
    f.append(f1 * (vsq + inv_rho_p_mean / (gamma - 1)) + (pl * vr[o] + pr * vl[o]) / 2)
    return f


def main():
    basis = {}
    for N in range(1, 8):
        x, w = lgl(N)
        basis[str(N)] = {"nodes": [mp.nstr(v, 30) for v in x], "weights": [mp.nstr(v, 30) for v in w]}
    json.dump(basis, open(os.path.join(HERE, "basis_mpmath.json"), "w"), indent=1)

    grid = []
    for x in ("0.1", "1", "3.7"):
        for r in ("1", "1.000001", "1.001", "1.01", "1.0199", "1.0201", "1.02", "1.05", "1.5", "2", "10", "1000",
                  "0.999", "0.98", "0.5"):
            xx = mp.mpf(x)
            yy = xx * mp.mpf(r)
            L = lnmean(xx, yy)
            grid.append({"x": mp.nstr(xx, 30), "y": mp.nstr(yy, 30), "ln_mean": mp.nstr(L, 30),
                         "inv_ln_mean": mp.nstr(1 / L, 30)})
    json.dump(grid, open(os.path.join(HERE, "ln_mean_mpmath.json"), "w"), indent=1)

    rng = np.random.default_rng(20261202)
    cases = []
    for D in (2, 3):
        for c in range(32):
            def state(scale):
                rho = 0.5 + rng.uniform(0, 1.5)
                v = rng.uniform(-1, 1, D) * scale
                p = 0.3 + rng.uniform(0, 2)
                return [float(rho)] + [float(rho * vi) for vi in v] + [float(p / 0.4 + 0.5 * rho * float(np.dot(v, v)))]
            ul = state(1.0)
            # half of the pairs are close (series branch), half far apart
            ur = [float(a * (1 + 1e-3 * rng.standard_normal())) for a in ul] if c % 2 else state(1.0)
            o = int(rng.integers(0, D))
            f = ranocha([mp.mpf(repr(a)) for a in ul], [mp.mpf(repr(a)) for a in ur], o, D)
            cases.append({"ndims": D, "orientation": o, "u_ll": ul, "u_rr": ur, "flux": [mp.nstr(v, 30) for v in f]})
    json.dump(cases, open(os.path.join(HERE, "flux_ranocha_mpmath.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
