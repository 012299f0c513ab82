"""Learned troubled-cell indicator for 2D p3 shock capturing.

A 19 -> 64 -> 64 -> 1 perceptron (ReLU, ReLU, sigmoid) whose output times
alpha_max replaces the Hennemann-Gassner blending factor of
VolumeIntegralShockCapturingHG (src/solvers/dgsem/indicators.jl:119-156).
Trixi.jl's own neural-network indicators were moved out of its source tree
(NEWS.md:495); this one is trained on the HG indicator itself, from seeded
Kelvin-Helmholtz runs (examples/tree_2d_dgsem/elixir_euler_kelvin_helmholtz_instability.jl),
so that it is a drop-in for the HG factor.

Inputs (computed on the device from the state, see the HIP kernel):
  0..15  nodal density*pressure, centred per element and scaled by the largest
         deviation; 16, 17 the HG logits of the two modal energy ratios / 8,
         clipped to [-2, 2]; 18 log10 of the smoothest-modes energy share / 8.
Output: alpha = alpha_max * sigmoid(net), clipped as HG clips its factor
(indicators_2d.jl:86-99): 0 below alpha_min (the element kernel then skips the
FV subcells of that element), at most alpha_max. With fallback (the default)
an element whose inputs leave the range seen in training uses the HG value.

Everything numerical runs in libtrixi_hip.so (BF16 MFMA forward and backward,
Adam on FP32 master weights). Device arrays are torch tensors (memory only).
"""
import ctypes
import os

import numpy as np

from . import _lib

NPARAMS = _lib.MLP_NPARAMS
NFEATURES = _lib.MLP_NFEATURES
FEATURE_STRIDE = _lib.MLP_FEATURE_STRIDE
NHIDDEN = 64

DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "indicator_nn_khi_p3.npz")


def init_params(seed=0):
    """He-normal weights for the ReLU layers, zero biases (FP32 vector in the
    C ABI order W1, b1, W2, b2, w3, b3)."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0.0, np.sqrt(2.0 / NFEATURES), (NHIDDEN, NFEATURES))
    W2 = rng.normal(0.0, np.sqrt(2.0 / NHIDDEN), (NHIDDEN, NHIDDEN))
    w3 = rng.normal(0.0, np.sqrt(1.0 / NHIDDEN), NHIDDEN)
    p = np.concatenate([W1.ravel(), np.zeros(NHIDDEN), W2.ravel(), np.zeros(NHIDDEN), w3, [-2.0]])
    assert p.size == NPARAMS
    return p.astype(np.float32)


def _fptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class IndicatorNeuralNetwork:
    """Learned replacement of IndicatorHennemannGassner's raw blending factor.

    params: FP32 [NPARAMS]; feature_lo/hi: FP32 [NFEATURES] training range;
    fallback: use the HG factor where an input leaves that range."""

    def __init__(self, params, feature_lo=None, feature_hi=None, fallback=True):
        self.params = np.ascontiguousarray(params, dtype=np.float32)
        if self.params.size != NPARAMS:
            raise ValueError(f"expected {NPARAMS} parameters")
        big = np.float32(3.0e38)
        self.feature_lo = np.ascontiguousarray(feature_lo if feature_lo is not None else np.full(NFEATURES, -big),
                                               dtype=np.float32)
        self.feature_hi = np.ascontiguousarray(feature_hi if feature_hi is not None else np.full(NFEATURES, big),
                                               dtype=np.float32)
        self.fallback = bool(fallback)

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS, fallback=True):
        with np.load(path, allow_pickle=False) as z:
            return cls(z["params"], z["feature_lo"], z["feature_hi"], fallback)

    def save(self, path, **meta):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        extra = {k: np.asarray(v) for k, v in meta.items()}
        np.savez(path, params=self.params, feature_lo=self.feature_lo, feature_hi=self.feature_hi, **extra)

    @property
    def mode(self):
        return _lib.IND_MLP_WITH_FALLBACK if self.fallback else _lib.IND_MLP

    def upload(self, semi):
        """Copy the parameters into the context of `semi` (resets the Adam state)."""
        lib, ctx = semi._lib, semi.ctx
        _lib.check(lib.trixi_hip_mlp_set_params(ctx, _fptr(self.params), _fptr(self.feature_lo),
                                                _fptr(self.feature_hi)), ctx, "mlp_set_params")

    def attach(self, semi):
        """Make the shock-capturing RHS of `semi` blend with this indicator."""
        self.upload(semi)
        _lib.check(semi._lib.trixi_hip_set_indicator(semi.ctx, self.mode), semi.ctx, "set_indicator")
        return semi

    def download(self, semi):
        _lib.check(semi._lib.trixi_hip_mlp_get_params(semi.ctx, _fptr(self.params)), semi.ctx, "mlp_get_params")
        return self


def use_hg(semi):
    _lib.check(semi._lib.trixi_hip_set_indicator(semi.ctx, _lib.IND_HG), semi.ctx, "set_indicator")


def features(semi):
    """Features [E, 32] (FP32) and raw HG alpha [E] of the resident state."""
    E = semi.nelements()
    f = np.empty((E, FEATURE_STRIDE), dtype=np.float32)
    a = np.empty(E, dtype=np.float64)
    _lib.check(semi._lib.trixi_hip_mlp_features(semi.ctx, _fptr(f), _fptr(a)), semi.ctx, "mlp_features")
    return f, a


def indicator_alpha(semi, mode):
    """Smoothed blending factors [E] of the resident state for indicator `mode`."""
    a = np.empty(semi.nelements(), dtype=np.float64)
    _lib.check(semi._lib.trixi_hip_indicator(semi.ctx, int(mode), _fptr(a)), semi.ctx, "indicator")
    return a


def forward(semi, x_dev, y_dev):
    """y = sigmoid(net(x)) on device tensors (x: float32 [n, 32], y: float32 [n])."""
    n = x_dev.shape[0]
    _lib.check(semi._lib.trixi_hip_mlp_forward(semi.ctx, ctypes.c_void_p(x_dev.data_ptr()), int(n),
                                               ctypes.c_void_p(y_dev.data_ptr())), semi.ctx, "mlp_forward")


def loss_grad(semi, x_dev, t_dev):
    """(loss, gradient [NPARAMS]) of sum (y - t)^2 / n."""
    g = np.empty(NPARAMS, dtype=np.float32)
    loss = ctypes.c_float()
    _lib.check(semi._lib.trixi_hip_mlp_loss_grad(semi.ctx, ctypes.c_void_p(x_dev.data_ptr()),
                                                 ctypes.c_void_p(t_dev.data_ptr()), int(x_dev.shape[0]),
                                                 ctypes.byref(loss), _fptr(g)), semi.ctx, "mlp_loss_grad")
    return float(loss.value), g


def train_step(semi, x_dev, t_dev, learning_rate, want_loss=False):
    loss = ctypes.c_float()
    _lib.check(semi._lib.trixi_hip_mlp_train_step(semi.ctx, ctypes.c_void_p(x_dev.data_ptr()),
                                                  ctypes.c_void_p(t_dev.data_ptr()), int(x_dev.shape[0]),
                                                  float(learning_rate), ctypes.byref(loss) if want_loss else None),
               semi.ctx, "mlp_train_step")
    return float(loss.value) if want_loss else None


# ---------------------------------------------------------------------------
# Kelvin-Helmholtz training data and training
# ---------------------------------------------------------------------------
def khi_semidiscretization(level=5, noise_amplitude=1e-3, seed=0, precision="fp64"):
    """The setup of elixir_euler_kelvin_helmholtz_instability.jl (HG alpha_max
    0.002, alpha_min 1e-4, flux_ranocha volume flux, LLF with the naive wave
    speed), with a seeded perturbation of v2 of the given amplitude."""
    from .equations import (CompressibleEulerEquations2D, FluxLaxFriedrichs, flux_ranocha,
                            initial_condition_kelvin_helmholtz_instability, max_abs_speed_naive)
    from .basis import LobattoLegendreBasis
    from .mesh import TreeMesh
    from .semidiscretization import SemidiscretizationHyperbolic
    from .solver import DGSEM, IndicatorHennemannGassner, VolumeIntegralShockCapturingHG
    eq = CompressibleEulerEquations2D(1.4)
    mesh = TreeMesh((-1.0, -1.0), (1.0, 1.0), initial_refinement_level=level, periodicity=True)
    surface_flux = FluxLaxFriedrichs(max_abs_speed_naive)
    basis = LobattoLegendreBasis(3)
    ind = IndicatorHennemannGassner(eq, basis, alpha_max=0.002, alpha_min=0.0001, alpha_smooth=True)
    solver = DGSEM(basis, surface_flux,
                   VolumeIntegralShockCapturingHG(ind, volume_flux_dg=flux_ranocha, volume_flux_fv=surface_flux))
    ic = initial_condition_kelvin_helmholtz_instability(noise_amplitude, seed)
    return SemidiscretizationHyperbolic(mesh, eq, ic, solver, precision=precision)


def run_segments(semi, t_end, nsnap, cfl=1.3, on_snapshot=None):
    """Integrate from t = 0 to t_end with CarpenterKennedy2N54 and the CFL
    StepsizeCallback, stopping at nsnap equally spaced times (including 0) to
    call on_snapshot(semi, t) on the resident state. Returns the final u."""
    from .semidiscretization import ODEProblem, rhs_hyperbolic_, semidiscretize
    from .time_integration import CallbackSet, CarpenterKennedy2N54, StepsizeCallback, solve
    u = semidiscretize(semi, (0.0, t_end)).u0
    times = np.linspace(0.0, t_end, nsnap)
    lib, ctx = semi._lib, semi.ctx
    _lib.check(lib.trixi_hip_upload_u(ctx, u.ctypes.data_as(ctypes.c_void_p)), ctx, "upload_u")
    if on_snapshot:
        on_snapshot(semi, 0.0)
    for t0, t1 in zip(times[:-1], times[1:]):
        sol = solve(ODEProblem(rhs_hyperbolic_, u, (t0, t1), semi), CarpenterKennedy2N54(),
                    callback=CallbackSet(StepsizeCallback(cfl=cfl)))
        u = sol.u
        if on_snapshot:
            on_snapshot(semi, t1)
    return u


def collect_khi_data(seeds, level=5, t_end=3.0, nsnap=21, precision="fp64", log=None):
    """Features and HG targets (raw alpha / alpha_max) of HG-driven KHI runs."""
    X, Y = [], []
    for s in seeds:
        semi = khi_semidiscretization(level, 1e-3, s, precision)

        def grab(sm, t):
            f, a = features(sm)
            X.append(f)
            Y.append((a / 0.002).astype(np.float32))

        run_segments(semi, t_end, nsnap, on_snapshot=grab)
        semi.close()
        if log:
            log(f"seed {s}: {len(X) * X[-1].shape[0]} samples")
    return np.concatenate(X), np.concatenate(Y)


def dataset_from_files(semi, files, alpha_max=0.002):
    """Training samples from restart files (the reference's (state, alpha) logging
    path, save_restart_dg.jl, save_solution_dg.jl): each file's conserved state
    is loaded into semi's context, and the features and HG targets (raw alpha /
    alpha_max) are computed on the device as during a run."""
    from .io import load_restart_file
    X, Y = [], []
    lib, ctx = semi._lib, semi.ctx
    for f in files:
        u = load_restart_file(semi, f)
        _lib.check(lib.trixi_hip_upload_u(ctx, u.ctypes.data_as(ctypes.c_void_p)), ctx, "upload_u")
        x, a = features(semi)
        X.append(x)
        Y.append((a / alpha_max).astype(np.float32))
    return np.concatenate(X), np.concatenate(Y)


def training_targets(alpha_hg, alpha_max=0.002, alpha_min=0.0001, dead_zone=1.5):
    """Network targets from HG blending factors: alpha / alpha_max, and 0 where
    alpha < dead_zone * alpha_min. With the KHI parameters HG never reaches its
    own clipping (its sigmoid is alpha_min at zero modal energy,
    indicators.jl:130-135 and indicators_2d.jl:80-99), so every element is
    blended, most of them at alpha ~ alpha_min where the FV part weighs 1e-4.
    The dead zone maps that floor to 0: the network learns pure DG there, and
    its clipped output (0 below alpha_min) lets the element kernel skip the FV
    subcells of those elements. dead_zone = 1 keeps HG's targets unchanged."""
    a = np.asarray(alpha_hg, dtype=np.float64)
    return np.where(a < dead_zone * alpha_min, 0.0, a / alpha_max).astype(np.float32)


def train(semi, X, Y, steps=4000, batch=8192, lr=2e-3, seed=0, log=None, log_every=500):
    """Adam on the GPU from init_params(seed); minibatches drawn with
    numpy.random.default_rng(seed). Returns the trained IndicatorNeuralNetwork
    and the loss history [(step, loss)]."""
    import torch
    rng = np.random.default_rng(seed)
    lo = X[:, :NFEATURES].min(axis=0)
    hi = X[:, :NFEATURES].max(axis=0)
    net = IndicatorNeuralNetwork(init_params(seed), lo, hi, fallback=True)
    net.upload(semi)
    Xd = torch.from_numpy(np.ascontiguousarray(X, dtype=np.float32)).cuda()
    Yd = torch.from_numpy(np.ascontiguousarray(Y, dtype=np.float32)).cuda()
    n = Xd.shape[0]
    hist = []
    for it in range(1, steps + 1):
        idx = torch.from_numpy(rng.integers(0, n, batch)).cuda()
        xb = Xd.index_select(0, idx).contiguous()
        yb = Yd.index_select(0, idx).contiguous()
        rate = lr * (0.5 * (1 + np.cos(np.pi * (it - 1) / steps)))  # cosine decay
        want = it % log_every == 0 or it == steps
        loss = train_step(semi, xb, yb, rate, want_loss=want)
        if want:
            hist.append((it, loss))
            if log:
                log(f"step {it}: loss {loss:.3e}")
    net.download(semi)
    return net, hist


def evaluate(net, seed=1000, level=5, t_end=1.0, nsnap=5):
    """Run the KHI setup twice from the same seeded state, blending with HG and
    with `net`, and compare along the run: mean / max |alpha_net - alpha_HG|
    (in units of alpha_max) per snapshot, the share of elements that fell back
    to HG, and the final density difference (RMS, max)."""
    alphas, fellback, finals = {}, [], {}
    for name in ("hg", "nn"):
        semi = khi_semidiscretization(level, 1e-3, seed)
        if name == "nn":
            net.attach(semi)
        rec = []

        def snap(sm, t):
            rec.append(indicator_alpha(sm, _lib.IND_HG if name == "hg" else net.mode))
            if name == "nn" and net.fallback:
                f, _ = features(sm)
                out = np.any((f[:, :NFEATURES] < net.feature_lo) | (f[:, :NFEATURES] > net.feature_hi), axis=1)
                fellback.append(float(np.mean(out)))

        finals[name] = run_segments(semi, t_end, nsnap, on_snapshot=snap)
        alphas[name] = rec
        semi.close()
    d = [np.abs(a - b) / 0.002 for a, b in zip(alphas["nn"], alphas["hg"])]
    drho = finals["nn"].reshape(-1, 4)[:, 0] - finals["hg"].reshape(-1, 4)[:, 0]
    return dict(seed=seed, level=level, t_end=t_end,
                alpha_mean_abs_diff=[float(x.mean()) for x in d], alpha_max_abs_diff=[float(x.max()) for x in d],
                fallback_share=fellback, rho_rms_diff=float(np.sqrt(np.mean(drho ** 2))),
                rho_max_diff=float(np.abs(drho).max()))


# ---------------------------------------------------------------------------
# Judging the learned indicator by the HG run's solution error (KHI, level 7)
# ---------------------------------------------------------------------------
def transfer_uniform(to, frm):
    """Carry the resident state of `frm` into `to` on the device, for two
    uniform meshes of the same domain one level apart: refinement (each element
    of frm interpolated to its 2^D children, which follow each other in child
    order in the finer mesh's element order) or the L2 projection of the 2^D
    children onto their parent (amr_dg2d.jl; trixi_hip_transfer)."""
    D = to.mesh.ndims
    nch = 1 << D
    if to.nelements() == nch * frm.nelements():
        E = frm.nelements()
        kind = np.tile(np.arange(1, nch + 1, dtype=np.int32), E)
        src = np.repeat(np.arange(E, dtype=np.int32), nch)
    elif frm.nelements() == nch * to.nelements():
        E = to.nelements()
        kind = np.full(E, -1, dtype=np.int32)
        src = (nch * np.arange(E)).astype(np.int32)
    else:
        raise ValueError("meshes are not one uniform refinement level apart")
    _lib.check(to._lib.trixi_hip_transfer(to.ctx, frm.ctx, kind.ctypes.data_as(ctypes.c_void_p),
                                          src.ctypes.data_as(ctypes.c_void_p)), to.ctx, "transfer")


def _download(semi):
    u = np.empty(semi.nvalues())
    _lib.check(semi._lib.trixi_hip_download_u(semi.ctx, u.ctypes.data_as(ctypes.c_void_p)), semi.ctx, "download_u")
    return u


def _upload(semi, u):
    _lib.check(semi._lib.trixi_hip_upload_u(semi.ctx, np.ascontiguousarray(u).ctypes.data_as(ctypes.c_void_p)),
               semi.ctx, "upload_u")


def l2_density(semi, u_a, u_b):
    """L2 norm of the density difference over the domain, by the LGL quadrature
    of the DG solution (the integral calc_error_norms evaluates, here on the
    solution nodes), normalised by the domain volume."""
    b = semi.solver.basis
    D = semi.mesh.ndims
    w = np.asarray(b.weights)
    W = w if D == 1 else (np.outer(w, w).reshape(-1) if D == 2 else np.einsum("i,j,k->kji", w, w, w).reshape(-1))
    V = semi.nvariables()
    d = (np.asarray(u_a) - np.asarray(u_b)).reshape(semi.nelements(), -1, V)[:, :, 0]
    jac = (1.0 / semi.inverse_jacobian) ** D  # volume of the reference element's image / 2^D
    return float(np.sqrt(np.sum(jac[:, None] * W[None, :] * d * d) / np.sum(jac[:, None] * W[None, :])))


def rhs_time(semi, u, mode, launches=200):
    """Time of one RHS (indicator kernels and element kernel) of the shock-
    capturing operator with blending factors from indicator `mode`, replayed
    from a HIP graph on device copies of u in the context's precision."""
    import torch
    lib, ctx = semi._lib, semi.ctx
    _lib.check(lib.trixi_hip_set_indicator(ctx, int(mode)), ctx, "set_indicator")
    dtype = torch.float32 if semi.precision == "fp32" else torch.float64
    ud = torch.from_numpy(u).to(device="cuda", dtype=dtype)
    dud = torch.empty_like(ud)
    stream = torch.cuda.Stream()
    _lib.check(lib.trixi_hip_set_stream(ctx, ctypes.c_void_p(stream.cuda_stream)), ctx, "set_stream")

    def launch():
        _lib.check(lib.trixi_hip_rhs_native(ctx, ctypes.c_void_p(dud.data_ptr()), ctypes.c_void_p(ud.data_ptr()),
                                            0.0), ctx, "rhs_native")

    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        for _ in range(20):
            launch()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        for _ in range(10):
            launch()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    with torch.cuda.stream(stream):
        g.replay()
        ev[0].record(stream)
        for _ in range(launches // 10):
            g.replay()
        ev[1].record(stream)
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / (10 * (launches // 10))
    _lib.check(lib.trixi_hip_set_stream(ctx, None), ctx, "set_stream")
    _lib.check(lib.trixi_hip_set_indicator(ctx, _lib.IND_HG), ctx, "set_indicator")
    del g
    return ms


def blended_share(semi, alpha, elements_per_tile=16):
    """Share of elements whose blending factor is not dg_only
    (calc_volume_integral.jl:240-255: |alpha| > atol of the RHS precision) and
    share of element-kernel workgroup tiles with at least one such element."""
    atol = 1.1920928955078125e-5 if semi.precision == "fp32" else 1.8189894035458565e-12
    b = np.abs(alpha) > atol
    E = b.size
    pad = (-E) % elements_per_tile
    tiles = np.concatenate([b, np.zeros(pad, bool)]).reshape(-1, elements_per_tile).any(axis=1)
    return float(b.mean()), float(tiles.mean())


def evaluate_config5(net, seed=1000, level=7, t_end=3.0, ref_level=8, cfl=1.3, log=None):
    """The learned indicator judged by the HG run's solution error on a KHI
    seed not used in training: three FP64 runs from the same seeded initial
    state to t_end, (1) HG at `level`, (2) the network (with HG fallback) at
    `level`, (3) the reference, HG at `ref_level` started from the level-`level`
    initial state interpolated to the finer mesh. The errors are the L2 norms of
    the density differences to the reference projected (L2) onto the coarser
    mesh. Also: the direct density difference of runs (1) and (2), the RHS time
    of both operators (FP64 and FP32) on the final HG state, the blended shares
    and the share of elements that fell back to HG."""
    from .semidiscretization import ODEProblem, rhs_hyperbolic_
    from .time_integration import CallbackSet, CarpenterKennedy2N54, StepsizeCallback, solve

    def say(msg):
        if log:
            log(msg)

    def run(semi, u0):
        sol = solve(ODEProblem(rhs_hyperbolic_, u0, (0.0, t_end), semi), CarpenterKennedy2N54(),
                    callback=CallbackSet(StepsizeCallback(cfl=cfl)))
        return sol

    out = dict(seed=seed, level=level, ref_level=ref_level, t_end=t_end, cfl=cfl)
    hg = khi_semidiscretization(level, 1e-3, seed, "fp64")
    u0 = hg.compute_coefficients(0.0)
    s_hg = run(hg, u0)
    out["steps_hg"] = s_hg.nsteps
    out["wall_s_hg"] = s_hg.wall_time
    say(f"HG level {level}: {s_hg.nsteps} steps, {s_hg.wall_time:.2f} s")
    nn = khi_semidiscretization(level, 1e-3, seed, "fp64")
    net.attach(nn)
    s_nn = run(nn, u0)
    out["steps_nn"] = s_nn.nsteps
    out["wall_s_nn"] = s_nn.wall_time
    say(f"NN level {level}: {s_nn.nsteps} steps, {s_nn.wall_time:.2f} s")
    # reference: the same initial polynomial state on the finer mesh
    ref = khi_semidiscretization(ref_level, 1e-3, seed, "fp64")
    _upload(hg, u0)
    transfer_uniform(ref, hg)
    s_ref = run(ref, _download(ref))
    out["steps_ref"] = s_ref.nsteps
    say(f"HG level {ref_level}: {s_ref.nsteps} steps, {s_ref.wall_time:.2f} s")
    transfer_uniform(hg, ref)  # the reference projected onto the coarser mesh
    u_ref = _download(hg)
    out["l2_rho_error_hg"] = l2_density(hg, s_hg.u, u_ref)
    out["l2_rho_error_nn"] = l2_density(hg, s_nn.u, u_ref)
    out["error_ratio_nn_over_hg"] = out["l2_rho_error_nn"] / out["l2_rho_error_hg"]
    out["l2_rho_nn_minus_hg"] = l2_density(hg, s_nn.u, s_hg.u)
    out["l2_rho_hg_norm"] = l2_density(hg, s_hg.u, 0 * s_hg.u)
    # fallback share and blended shares on the final states
    _upload(nn, s_nn.u)
    f, _ = features(nn)
    outside = np.any((f[:, :NFEATURES] < net.feature_lo) | (f[:, :NFEATURES] > net.feature_hi), axis=1)
    out["fallback_share_final"] = float(np.mean(outside))
    ref.close()
    nn.close()
    # RHS time of both operators on the final HG state, FP64 and FP32
    for prec in ("fp64", "fp32"):
        sm = khi_semidiscretization(level, 1e-3, seed, prec)
        _upload(sm, s_hg.u)
        net.upload(sm)
        for name, mode in (("hg", _lib.IND_HG), ("nn", net.mode)):
            out[f"rhs_ms_{name}_{prec}"] = rhs_time(sm, s_hg.u, mode)
            el, tl = blended_share(sm, indicator_alpha(sm, mode))
            out[f"blended_elements_{name}_{prec}"] = el
            out[f"blended_tiles_{name}_{prec}"] = tl
        sm.close()
    hg.close()
    say(str(out))
    return out


def spec_features(semi, u):
    """The input set the learned-indicator configuration describes, computed on
    the host for the feature comparison (compare_feature_sets): the 16 nodal
    values of rho*p divided by their element mean, and three ratios of the
    modal energies of rho*p (HG's dimensionwise modal transform,
    indicators_2d.jl:59-79): (E - E1) / E, (E1 - E2) / E1 and E2 / E, with E, E1,
    E2 the energies of all modes, of the modes below N and of those below N - 1.
    u: the FP64 state [E * 16 * 4]. Returns FP32 [E, 32] (zero padded)."""
    gamma = semi.equations.gamma
    n = semi.solver.basis.nnodes()
    w = np.asarray(u, dtype=np.float64).reshape(-1, n * n, 4)
    rho = w[..., 0]
    p = (gamma - 1) * (w[..., 3] - 0.5 * (w[..., 1] ** 2 + w[..., 2] ** 2) / rho)
    q = rho * p  # [E, n*n], node = i + n j
    vinv = np.asarray(semi.solver.basis.inverse_vandermonde_legendre, dtype=np.float64)
    m = np.einsum("ai,bj,eji->eba", vinv, vinv, q.reshape(-1, n, n))  # modal [E, j-mode b, i-mode a]
    m2 = m ** 2
    e_all = m2.sum(axis=(1, 2))
    e1 = m2[:, : n - 1, : n - 1].sum(axis=(1, 2))
    e2 = m2[:, : n - 2, : n - 2].sum(axis=(1, 2))
    tiny = 1e-300
    x = np.zeros((q.shape[0], FEATURE_STRIDE), dtype=np.float32)
    x[:, :16] = q / q.mean(axis=1, keepdims=True)
    x[:, 16] = (e_all - e1) / np.maximum(e_all, tiny)
    x[:, 17] = (e1 - e2) / np.maximum(e1, tiny)
    x[:, 18] = e2 / np.maximum(e_all, tiny)
    return x


def compare_feature_sets(train_seeds=(0, 1, 2, 3), val_seed=4, level=5, t_end=3.0, nsnap=11, steps=3000,
                         batch=8192, lr=2e-3, dead_zone=1.5, log=None):
    """The same network trained on the two input sets, same targets, seeds and
    schedule: the features of indicator_2d_p3_kernel (rho*p centred by the mean
    and scaled by the largest deviation, HG logits of the energy ratios, log of
    the low-mode share) and spec_features (rho*p / mean, three energy ratios).
    Snapshots of HG-driven KHI runs; validation on a seed not trained on: mean
    squared error against the targets, mean |alpha - alpha_target| / alpha_max
    of the clipped output, and the share of elements whose blend decision
    (clipped alpha > 0) differs from the target's."""
    import torch
    data = {}
    for s in list(train_seeds) + [val_seed]:
        semi = khi_semidiscretization(level, 1e-3, s)
        rec = {"ours": [], "spec": [], "y": []}

        def grab(sm, t):
            f, a = features(sm)
            rec["ours"].append(f)
            rec["spec"].append(spec_features(sm, _download(sm)))
            rec["y"].append(training_targets(a, dead_zone=dead_zone))

        run_segments(semi, t_end, nsnap, on_snapshot=grab)
        semi.close()
        data[s] = {k: np.concatenate(v) for k, v in rec.items()}
    semi = khi_semidiscretization(level, 1e-3, 0)
    out = dict(level=level, t_end=t_end, snapshots=nsnap, train_seeds=list(train_seeds), val_seed=val_seed,
               steps=steps, batch=batch)
    yv = data[val_seed]["y"]
    for name in ("ours", "spec"):
        X = np.concatenate([data[s][name] for s in train_seeds])
        Y = np.concatenate([data[s]["y"] for s in train_seeds])
        net, hist = train(semi, X, Y, steps, batch, lr, log=log)
        net.upload(semi)
        xv = torch.from_numpy(np.ascontiguousarray(data[val_seed][name])).cuda()
        y = torch.empty(xv.shape[0], dtype=torch.float32, device="cuda")
        forward(semi, xv, y)
        yn = y.cpu().numpy().astype(np.float64)
        alpha = np.where(yn * 0.002 < 0.0001, 0.0, np.minimum(yn * 0.002, 0.002))  # hg_clip
        out[name] = dict(train_loss=float(hist[-1][1]), val_mse=float(np.mean((yn - yv) ** 2)),
                         val_alpha_mean_abs_diff=float(np.mean(np.abs(alpha / 0.002 - yv))),
                         val_blend_decision_mismatch=float(np.mean((alpha > 0) != (yv > 0))))
        if log:
            log(f"{name}: {out[name]}")
    semi.close()
    return out


def main(argv=None):
    import argparse
    import json
    import sys
    p = argparse.ArgumentParser(description="train the learned indicator on seeded KHI runs (needs a GPU)")
    p.add_argument("--out", default=DEFAULT_WEIGHTS)
    p.add_argument("--seeds", type=int, default=20)
    # the defaults reproduce the committed weights (their metadata records them)
    p.add_argument("--level", type=int, default=7)
    p.add_argument("--t-end", type=float, default=3.0)
    p.add_argument("--snapshots", type=int, default=21)
    p.add_argument("--steps", type=int, default=6000)
    p.add_argument("--batch", type=int, default=8192)
    p.add_argument("--lr", type=float, default=2e-3)
    p.add_argument("--dead-zone", type=float, default=1.5,
                   help="targets are 0 where the HG factor is below this multiple of alpha_min")
    p.add_argument("--evaluate", action="store_true", help="evaluate the weights at --out instead of training")
    p.add_argument("--config5", action="store_true",
                   help="judge the weights at --out by the HG run's density error (evaluate_config5)")
    p.add_argument("--seed", type=int, default=1000, help="KHI seed of --evaluate / --config5")
    p.add_argument("--compare-features", action="store_true",
                   help="train on this kernel's features and on spec_features and compare (compare_feature_sets)")
    a = p.parse_args(argv)

    def log(msg):
        print(f"[learned_indicator] {msg}", file=sys.stderr, flush=True)

    if a.evaluate:
        print(json.dumps(evaluate(IndicatorNeuralNetwork.load(a.out), seed=a.seed, level=a.level)))
        return
    if a.compare_features:
        print(json.dumps(compare_feature_sets(level=a.level, t_end=a.t_end, nsnap=a.snapshots, steps=a.steps,
                                              batch=a.batch, lr=a.lr, dead_zone=a.dead_zone, log=log)))
        return
    if a.config5:
        print(json.dumps(evaluate_config5(IndicatorNeuralNetwork.load(a.out), seed=a.seed, level=a.level,
                                          t_end=a.t_end, ref_level=a.level + 1, log=log)))
        return

    X, Y = collect_khi_data(range(a.seeds), a.level, a.t_end, a.snapshots, log=log)
    Y = training_targets(Y * 0.002, dead_zone=a.dead_zone)
    log(f"{X.shape[0]} samples, {np.mean(Y > 0):.3f} with target > 0 (dead zone {a.dead_zone} alpha_min)")
    semi = khi_semidiscretization(a.level, 1e-3, 0)
    net, hist = train(semi, X, Y, a.steps, a.batch, a.lr, log=log)
    net.save(a.out, loss_history=np.array(hist), seeds=np.arange(a.seeds), level=a.level, t_end=a.t_end,
             snapshots=a.snapshots, dead_zone=a.dead_zone, steps=a.steps, batch=a.batch, lr=a.lr)
    print(json.dumps({"out": a.out, "samples": int(X.shape[0]), "final_loss": hist[-1][1]}))


if __name__ == "__main__":
    main()
