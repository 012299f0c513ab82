"""Reference for the learned troubled-cell indicator (test infrastructure only).

Only tests/ use this module, as the checker of the HIP perceptron kernels and
of the feature extraction; the product path never imports it.

* features(u, ...)   the 19 inputs per 2D p3 element, recomputed in FP64 from
                     the state with the HG indicator's definitions
                     (indicators.jl:119-156, indicators_2d.jl:30-102,
                     density_pressure compressible_euler_2d.jl).
* hg_clip, alpha     HG's clipping of a blending factor (indicators_2d.jl:86-99)
                     and the network's clipped blending factor.
* forward/backward   the 19 -> 64 -> 64 -> 1 perceptron (ReLU, ReLU, sigmoid)
                     and the gradient of L = sum (y - target)^2 / n, in FP64.
                     With bf16=True the operands that the MFMA kernels round to
                     BF16 (inputs, W1, W2, first hidden layer, the two back-
                     propagated pre-activation gradients) are rounded the same
                     way (round to nearest even), everything else stays FP64.
The analytic backward is pinned by central differences (bf16=False).
"""
import numpy as np

NF, KF, NH = 19, 32, 64
OFF = {}
_o = 0
for _name, _shape in (("W1", (NH, NF)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("w3", (NH,)), ("b3", (1,))):
    OFF[_name] = (_o, _shape)
    _o += int(np.prod(_shape))
NP = _o  # 5505


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return {k: p[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in OFF.items()}


def pack(d):
    p = np.zeros(NP)
    for k, (o, s) in OFF.items():
        p[o:o + int(np.prod(s))] = np.asarray(d[k], dtype=np.float64).reshape(-1)
    return p


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    b = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) >> 16 << 16
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def _r(x, on):
    return bf16(x) if on else np.asarray(x, dtype=np.float64)


def forward(p, x, bf=False):
    """x [n, >=19] -> dict of intermediates and y [n]."""
    w = unpack(np.asarray(p, dtype=np.float32).astype(np.float64) if bf else p)
    X = _r(np.asarray(x, dtype=np.float64)[:, :NF], bf)
    W1, W2 = _r(w["W1"], bf), _r(w["W2"], bf)
    z1 = X @ W1.T + w["b1"]
    h1 = _r(np.maximum(z1, 0), bf)
    z2 = h1 @ W2.T + w["b2"]
    h2 = np.maximum(z2, 0)
    o = h2 @ w["w3"] + w["b3"][0]
    y = 1 / (1 + np.exp(-o))
    return dict(X=X, W1=W1, W2=W2, w=w, z1=z1, h1=h1, z2=z2, h2=h2, y=y)


def hg_clip(alpha, alpha_min, alpha_max):
    """The clipping of calc_indicator_hennemann_gassner! (indicators_2d.jl:86-99):
    0 below alpha_min, 1 above 1 - alpha_min, then min(alpha_max, .)."""
    a = np.asarray(alpha, dtype=np.float64)
    a = np.where(a < alpha_min, 0.0, a)
    a = np.where(a > 1 - alpha_min, 1.0, a)
    return np.minimum(a, alpha_max)


def alpha(p, x, alpha_max, alpha_min, bf=True):
    """Blending factor of the learned indicator without fallback:
    hg_clip(alpha_max * y)."""
    return hg_clip(alpha_max * forward(p, x, bf)["y"], alpha_min, alpha_max)


def loss_grad(p, x, target, bf=False):
    f = forward(p, x, bf)
    n = len(f["y"])
    y = f["y"]
    res = y - np.asarray(target, dtype=np.float64)
    loss = np.sum(res * res) / n
    dout = 2 * res * y * (1 - y) / n
    w = f["w"]
    g = {"b3": np.array([dout.sum()]), "w3": f["h2"].T @ dout}
    dz2 = np.outer(dout, w["w3"]) * (f["z2"] > 0)
    g["b2"] = dz2.sum(0)
    dz2r = _r(dz2, bf)
    g["W2"] = dz2r.T @ f["h1"]
    dh1 = dz2r @ f["W2"]
    dz1 = dh1 * (f["z1"] > 0)
    g["b1"] = dz1.sum(0)
    g["W1"] = _r(dz1, bf).T @ f["X"]
    return loss, pack(g)


def features(u, ndims, polydeg, vinv, gamma, alpha_max, alpha_min):
    """Features [E, 32] and raw HG alpha [E] of a 2D p3 state u (flat, reference layout)."""
    assert ndims == 2 and polydeg == 3
    n = polydeg + 1
    NN = n * n
    V = 4
    u = np.asarray(u, dtype=np.float64).reshape(-1, NN, V)
    E = u.shape[0]
    rho = u[:, :, 0]
    q = (gamma - 1) * (rho * u[:, :, 3] - 0.5 * (u[:, :, 1] ** 2 + u[:, :, 2] ** 2))  # rho * p
    # modal coefficients: Vinv along x then y (node = i + n j)
    Q = q.reshape(E, n, n)  # [e, j, i]
    M = np.einsum("ab,ejb->eja", vinv, Q)
    M = np.einsum("ab,ebi->eai", vinv, M)  # [e, j, i] modal
    m2 = M ** 2
    te = m2.sum(axis=(1, 2))
    te1 = m2[:, :n - 1, :n - 1].sum(axis=(1, 2))
    te2 = m2[:, :n - 2, :n - 2].sum(axis=(1, 2))
    with np.errstate(divide="ignore", invalid="ignore"):
        ef1 = np.where(te != 0, (te - te1) / te, 0.0)
        ef2 = np.where(te1 != 0, (te1 - te2) / te1, 0.0)
    T = 0.5 * 10 ** (-1.8 * n ** 0.25)
    s = np.log((1 - 0.0001) / 0.0001)
    energy = np.maximum(ef1, ef2)
    alpha = hg_clip(1 / (1 + np.exp(-s / T * (energy - T))), alpha_min, alpha_max)
    feat = np.zeros((E, KF))
    mean = q.mean(axis=1, keepdims=True)
    dev = np.abs(q - mean).max(axis=1, keepdims=True)
    feat[:, :NN] = (q - mean) / (dev + 1e-7 * np.abs(mean) + 1e-300)
    k = s / T
    feat[:, NN] = np.clip(k * (ef1 - T) / 8, -2, 2)
    feat[:, NN + 1] = np.clip(k * (ef2 - T) / 8, -2, 2)
    smooth = np.where(te > 0, te2 / np.where(te > 0, te, 1), 1.0)
    feat[:, NN + 2] = np.clip(np.log10(smooth + 1e-16) / 8, -2, 0)
    return feat, alpha
