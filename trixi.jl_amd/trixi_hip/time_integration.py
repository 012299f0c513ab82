"""2N low-storage Runge-Kutta integration on the device-resident state, with
the StepsizeCallback and AnalysisCallback semantics Trixi.jl uses.

  CarpenterKennedy2N54        src/time_integration/methods_2N.jl:24-48
  SimpleIntegrator2N / step!  methods_2N.jl:95-168 (one graph replay per step)
  limit_dt!                   src/time_integration/time_integration.jl:46-55
  StepsizeCallback            src/callbacks_step/stepsize.jl:31-155
  AnalysisCallback (errors)   src/callbacks_step/analysis.jl, analysis_dg3d.jl:123-161
  AMRCallback                 src/callbacks_step/amr.jl (trixi_hip.amr)
"""
import ctypes
import math
import time

import numpy as np

from . import _lib
from .amr import AMRCallback


class CarpenterKennedy2N54:
    def __init__(self, williamson_condition=False):
        self.a = np.array([0.0,
                           567301805773.0 / 1357537059087.0,
                           2404267990393.0 / 2016746695238.0,
                           3550918686646.0 / 2091501179385.0,
                           1275806237668.0 / 842570457699.0])
        self.b = np.array([1432997174477.0 / 9575080441755.0,
                           5161836677717.0 / 13612068292357.0,
                           1720146321549.0 / 2090206949498.0,
                           3134564353537.0 / 4481467310338.0,
                           2277821191437.0 / 14882151754819.0])
        self.c = np.array([0.0,
                           1432997174477.0 / 9575080441755.0,
                           2526269341429.0 / 6820363962896.0,
                           2006345519317.0 / 3224310063776.0,
                           2802321613138.0 / 2924317926251.0])


class StepsizeCallback:
    """`StepsizeCallback(; cfl)`: dt = cfl * max_dt(u) at initialisation and after every step."""

    def __init__(self, cfl=1.0, interval=1):
        self.cfl = float(cfl)
        self.interval = int(interval)

    def calculate_dt(self, semi):
        dt = ctypes.c_double()
        _lib.check(semi._lib.trixi_hip_max_dt(semi.ctx, self.cfl, ctypes.byref(dt)), semi.ctx, "max_dt")
        return dt.value


class AnalysisCallback:
    """Error norms (l2, linf against the initial condition at time t) every
    `interval` steps and at the end (analysis.jl:46-236, 380-470)."""

    def __init__(self, semi, interval=100, capacity=1 << 16):
        self.semi = semi
        self.interval = int(interval)
        self.capacity = int(capacity)  # records kept on the device during a device-controlled solve
        self.history = []

    def analyze(self, t, it):
        semi = self.semi
        V = semi.nvariables()
        l2 = (ctypes.c_double * V)()
        linf = (ctypes.c_double * V)()
        _lib.check(semi._lib.trixi_hip_error_norms(semi.ctx, float(t), l2, linf), semi.ctx, "error_norms")
        cons, ent = self.integrals()
        rec = {"t": t, "iter": it, "l2": list(l2), "linf": list(linf), "integrals": cons, "entropy": ent}
        self.history.append(rec)
        return rec

    def device_records(self):
        """Records the device made during the last trixi_hip_solve."""
        semi = self.semi
        V = semi.nvariables()
        W = 3 + 3 * V
        cap = self.capacity
        buf = (ctypes.c_double * (cap * W))()
        n = ctypes.c_int()
        _lib.check(semi._lib.trixi_hip_get_analysis(semi.ctx, buf, cap, ctypes.byref(n)), semi.ctx, "get_analysis")
        a = np.frombuffer(buf, dtype=np.float64).reshape(cap, W)[:n.value]
        for r in a:
            self.history.append({"t": float(r[0]), "iter": int(r[1]), "l2": list(r[2:2 + V]),
                                 "linf": list(r[2 + V:2 + 2 * V]), "integrals": list(r[2 + 2 * V:2 + 3 * V]),
                                 "entropy": float(r[2 + 3 * V])})

    def integrals(self):
        semi = self.semi
        V = semi.nvariables()
        cons = (ctypes.c_double * V)()
        ent = ctypes.c_double()
        _lib.check(semi._lib.trixi_hip_integrals(semi.ctx, cons, ctypes.byref(ent)), semi.ctx, "integrals")
        return list(cons), ent.value


class CallbackSet:
    def __init__(self, *callbacks):
        self.callbacks = [c for c in callbacks if c is not None]

    def of(self, kind):
        return [c for c in self.callbacks if isinstance(c, kind)]


class Solution:
    def __init__(self, u, t, nsteps, analysis, wall, semi):
        self.u = u
        self.t = t
        self.nsteps = nsteps
        self.analysis = analysis
        self.wall_time = wall
        self.semi = semi


def _isapprox(a, b):
    return a == b or abs(a - b) <= math.sqrt(np.finfo(np.float64).eps) * max(abs(a), abs(b))


def solve(ode, alg, dt=None, callback=None, maxiters=None, use_graph=True, save_final=True, device_steps=True,
          poll_every=16, initial_step=0):
    """`solve(ode, CarpenterKennedy2N54(); dt, callback)` on the device.

    The state lives in the HIP context for the whole run. Three loops:
    * a StepsizeCallback (interval 1) and no AMR or file-writing callback: the
      whole run is one trixi_hip_solve (step graph with device-side dt and
      limit_dt, periodic analysis enqueued on the device, the host reads the
      device every `poll_every` steps);
    * the same with an AMRCallback: device-controlled solves up to the next
      adaptation or analysis step, which the host then runs;
    * otherwise (file writers, a fixed dt, `device_steps=False`): one
      trixi_hip_rk_step (a hipGraph replay of all stages) per step, the host
      computes dt and runs the callbacks after each step in the order of the
      CallbackSet (as OrdinaryDiffEq applies them; the elixirs order analysis,
      restart, solution, AMR, step size).
    Callbacks are initialised in CallbackSet order before the first step. All
    loops take the same steps. An AnalysisCallback records the initial state,
    every `interval`-th step and the final state (analysis.jl:121-122,
    143-230). `initial_step` continues the accepted-step count of a restarted
    run (load_timestep!, save_restart.jl:150-153)."""
    from .io import SaveRestartCallback, SaveSolutionCallback
    semi = ode.p
    lib = _lib.load()
    ctx = semi.ctx
    _lib.check(lib.trixi_hip_set_stream(ctx, None), ctx, "set_stream")
    _lib.check(lib.trixi_hip_upload_u(ctx, ode.u0.ctypes.data_as(ctypes.c_void_p)), ctx, "upload_u")
    cbs = callback if isinstance(callback, CallbackSet) else CallbackSet(*(callback or []))
    stepsize = cbs.of(StepsizeCallback)
    analysis = cbs.of(AnalysisCallback)
    amr = cbs.of(AMRCallback)
    writers = cbs.of(SaveSolutionCallback) + cbs.of(SaveRestartCallback)
    t0, t_end = ode.tspan
    t = t0
    na = len(alg.a)
    a = (ctypes.c_double * na)(*alg.a)
    b = (ctypes.c_double * na)(*alg.b)
    c = (ctypes.c_double * na)(*alg.c)
    it = int(initial_step)
    wall0 = time.perf_counter()
    on_device = bool(device_steps and use_graph and stepsize and stepsize[0].interval == 1 and not amr
                     and not writers and it == 0)
    records_on_device = bool(on_device and analysis and analysis[0].interval > 0)
    # with AMR (and no file writers) the steps between adaptations and analyses
    # run as device-controlled solves; the host adapts between them
    chunked = bool(device_steps and use_graph and stepsize and stepsize[0].interval == 1 and amr and not writers)
    for cb in cbs.callbacks:  # initialisation in callback order
        if isinstance(cb, AnalysisCallback) and not records_on_device:
            cb.analyze(t0, it)
        elif isinstance(cb, (SaveSolutionCallback, SaveRestartCallback)):
            cb.initialize(semi, t0, dt if dt is not None else 0.0, cbs.callbacks)
        elif isinstance(cb, AMRCallback):
            cb.initialize(semi, t0)
            ctx = semi.ctx
    if on_device:
        _lib.check(lib.trixi_hip_set_analysis(ctx, analysis[0].interval if records_on_device else 0,
                                              analysis[0].capacity if records_on_device else 0), ctx, "set_analysis")
        tout, nsteps = ctypes.c_double(), ctypes.c_int()
        rc = lib.trixi_hip_solve(ctx, a, b, c, na, float(t0), float(t_end), stepsize[0].cfl,
                                 int(maxiters) if maxiters is not None else 2 ** 31 - 1, int(poll_every),
                                 ctypes.byref(tout), ctypes.byref(nsteps))
        if rc == _lib.ESTATE and b"NaN" in lib.trixi_hip_last_error(ctx):
            raise FloatingPointError("time step size `dt` is NaN")
        _lib.check(rc, ctx, "solve")
        t, it = tout.value, nsteps.value
        if records_on_device:
            analysis[0].device_records()
    elif not chunked:
        if stepsize:
            dt = stepsize[0].calculate_dt(semi)
        if dt is None:
            raise ValueError("give dt or a StepsizeCallback")
    steps = 0
    while chunked:
        left = [2 ** 31 - 1]
        if amr[0].interval > 0:
            left.append(amr[0].interval - it % amr[0].interval)
        if analysis and analysis[0].interval > 0:
            left.append(analysis[0].interval - it % analysis[0].interval)
        if maxiters is not None:
            left.append(maxiters - steps)
        chunk = min(left)
        _lib.check(lib.trixi_hip_set_analysis(ctx, 0, 0), ctx, "set_analysis")
        tout, nsteps = ctypes.c_double(), ctypes.c_int()
        rc = lib.trixi_hip_solve(ctx, a, b, c, na, float(t), float(t_end), stepsize[0].cfl, int(chunk),
                                 int(poll_every), ctypes.byref(tout), ctypes.byref(nsteps))
        if rc == _lib.ESTATE and b"NaN" in lib.trixi_hip_last_error(ctx):
            raise FloatingPointError("time step size `dt` is NaN")
        _lib.check(rc, ctx, "solve")
        t, it, steps = tout.value, it + nsteps.value, steps + nsteps.value
        # a solve stops early only at t_end; a step that is not the last one ends
        # farther than sqrt(eps) from t_end (limit_dt!)
        last = nsteps.value < chunk or t >= t_end or _isapprox(t, t_end) or \
            (maxiters is not None and steps >= maxiters)
        if analysis and analysis[0].interval > 0 and it % analysis[0].interval == 0 and not last:
            analysis[0].analyze(t, it)
        if amr[0].condition(it, last):
            amr[0](semi)
            ctx = semi.ctx
        if last:
            break
    while not on_device and not chunked:
        if math.isnan(dt):
            raise FloatingPointError("time step size `dt` is NaN")
        final = False
        if t + dt > t_end or _isapprox(t + dt, t_end):
            dt = t_end - t
            final = True
        _lib.check(lib.trixi_hip_rk_step(ctx, a, b, c, na, float(t), float(dt), int(use_graph)), ctx, "rk_step")
        it += 1
        steps += 1
        t += dt
        # isfinished (callbacks_step.jl:25-31): t_end reached or maxiters steps taken
        last = final or (maxiters is not None and steps >= maxiters)
        for cb in cbs.callbacks:
            if isinstance(cb, AnalysisCallback):
                if cb.interval > 0 and it % cb.interval == 0 and not last:
                    cb.analyze(t, it)
            elif isinstance(cb, (SaveSolutionCallback, SaveRestartCallback)):
                if cb.condition(it, last):
                    cb(semi, t, dt, it, cbs.callbacks)
            elif isinstance(cb, AMRCallback):
                if cb.condition(it, last):
                    cb(semi)
                    ctx = semi.ctx
            elif isinstance(cb, StepsizeCallback):
                if it % cb.interval == 0:
                    dt = cb.calculate_dt(semi)
        if last:
            break
    _lib.check(lib.trixi_hip_synchronize(ctx), ctx, "synchronize")
    wall = time.perf_counter() - wall0
    if analysis and not records_on_device:
        analysis[0].analyze(t, it)
    u = None
    if save_final:
        u = np.empty(semi.nvalues())
        _lib.check(lib.trixi_hip_download_u(ctx, u.ctypes.data_as(ctypes.c_void_p)), ctx, "download_u")
    return Solution(u, t, it, analysis[0].history if analysis else [], wall, semi)
