"""Adaptive mesh refinement on a TreeMesh: IndicatorMax, ControllerThreeLevel
and AMRCallback, with the per-node work on the device.

  IndicatorMax(semi; variable)      src/solvers/dgsem/indicators.jl (IndicatorMax),
                                    indicators_3d.jl:239-259 -> trixi_hip_indicator_max
  ControllerThreeLevel(semi, ind; base_level, med_level, med_threshold,
                       max_level, max_threshold)
                                    src/callbacks_step/amr.jl:1116-1153 (host, per element)
  ControllerThreeLevelCombined      amr.jl:1254-1297
  AMRCallback(semi, controller; interval, adapt_initial_condition,
              adapt_initial_condition_only_refine)
                                    amr.jl:24-76 (condition), :143-198 (initialisation),
                                    :244-392 (one adaptation on a TreeMesh)
  tree adaptation                   TreeMesh.adapt (abstract_tree.jl:367-600)
  solution transfer                 trixi_hip_transfer (amr_dg3d.jl:138-430)

The tree is host data, as in the reference; the controller values are computed
from the device indicator, the adapted mesh gets a new device context, and the
state is projected from the old context to the new one on the device.
"""
import ctypes
import time

import numpy as np

from . import _lib


def first(u, equations=None):
    """`first`: the first conserved variable (the density)."""
    return u[..., 0]


class IndicatorMax:
    """`IndicatorMax(semi; variable)`: alpha = maximum of `variable` over the
    nodes of each element. `variable` is `first` or the 0-based index of a
    conserved variable (the device evaluates conserved variables only)."""

    def __init__(self, semi, variable=first):
        if variable is first:
            variable = 0
        if not isinstance(variable, (int, np.integer)) or not 0 <= variable < semi.nvariables():
            raise ValueError("IndicatorMax takes `first` or a conserved-variable index")
        self.variable = int(variable)

    def __call__(self, semi):
        alpha = np.empty(semi.nelements())
        _lib.check(semi._lib.trixi_hip_indicator_max(semi.ctx, self.variable, alpha.ctypes.data_as(ctypes.c_void_p)),
                   semi.ctx, "trixi_hip_indicator_max")
        return alpha


class ControllerThreeLevel:
    """`ControllerThreeLevel(semi, indicator; base_level=1, med_level=base_level,
    med_threshold=0.0, max_level=base_level, max_threshold=1.0)`: the target
    level is max_level above max_threshold, med_level above med_threshold (the
    current level if med_level <= 0), base_level otherwise; the controller
    value is +1 / -1 / 0 as the element is coarser / finer / at the target."""

    def __init__(self, semi, indicator, base_level=1, med_level=None, med_threshold=0.0, max_level=None,
                 max_threshold=1.0):
        self.indicator = indicator
        self.base_level = int(base_level)
        self.med_level = int(base_level if med_level is None else med_level)
        self.max_level = int(base_level if max_level is None else max_level)
        self.med_threshold = float(med_threshold)
        self.max_threshold = float(max_threshold)

    def __call__(self, semi):
        alpha = self.indicator(semi)
        levels = semi.mesh.levels
        med = self.med_level if self.med_level > 0 else levels
        target = np.where(alpha > self.max_threshold, self.max_level,
                          np.where(alpha > self.med_threshold, med, self.base_level))
        return np.sign(target - levels).astype(np.int32)


class ControllerThreeLevelCombined(ControllerThreeLevel):
    """`ControllerThreeLevelCombined(semi, indicator_primary, indicator_secondary;
    base_level, med_level, med_threshold, max_level, max_threshold,
    max_threshold_secondary)` (amr.jl:1254-1297): the three-level target of the
    primary indicator, raised to max_level where the secondary indicator is at
    least max_threshold_secondary."""

    def __init__(self, semi, indicator_primary, indicator_secondary, base_level=1, med_level=None,
                 med_threshold=0.0, max_level=None, max_threshold=1.0, max_threshold_secondary=1.0):
        super().__init__(semi, indicator_primary, base_level, med_level, med_threshold, max_level, max_threshold)
        self.indicator_secondary = indicator_secondary
        self.max_threshold_secondary = float(max_threshold_secondary)

    def __call__(self, semi):
        alpha = self.indicator(semi)
        alpha_secondary = self.indicator_secondary(semi)
        levels = semi.mesh.levels
        med = self.med_level if self.med_level > 0 else levels
        target = np.where(alpha > self.max_threshold, self.max_level,
                          np.where(alpha > self.med_threshold, med, self.base_level))
        target = np.where(alpha_secondary >= self.max_threshold_secondary, self.max_level, target)
        return np.sign(target - levels).astype(np.int32)


class AMRCallback:
    """`AMRCallback(semi, controller; interval, adapt_initial_condition=true,
    adapt_initial_condition_only_refine=true)`. `solve` calls it after every
    `interval`-th accepted step except the final one, before the step size is
    computed, and adapts the initial condition before the first step."""

    def __init__(self, semi, controller, interval=5, adapt_initial_condition=True,
                 adapt_initial_condition_only_refine=True):
        if not (isinstance(interval, int) and interval >= 0):
            raise ValueError("`interval` must be a non-negative integer")
        self.controller = controller
        self.interval = interval
        self.adapt_initial_condition = bool(adapt_initial_condition)
        self.adapt_initial_condition_only_refine = bool(adapt_initial_condition_only_refine)
        self.nelements_history = []
        # seconds spent per phase over all adaptations (the reference's timer
        # sections "indicator", "mesh" and "solver" of an AMR step)
        self.timer = {"indicator": 0.0, "mesh": 0.0, "solver": 0.0, "calls": 0}

    def __call__(self, semi, only_refine=False, only_coarsen=False):
        """One adaptation of semi's mesh and resident state; True if the mesh changed."""
        t0 = time.perf_counter()
        lam = self.controller(semi)
        t1 = time.perf_counter()
        changed, kind, src = semi.mesh.adapt(lam, only_refine=only_refine, only_coarsen=only_coarsen)
        t2 = time.perf_counter()
        if changed:
            semi.remesh(kind, src)
        t3 = time.perf_counter()
        self.timer["indicator"] += t1 - t0
        self.timer["mesh"] += t2 - t1
        self.timer["solver"] += t3 - t2
        self.timer["calls"] += 1
        self.nelements_history.append(semi.nelements())
        return changed

    def condition(self, naccept, final):
        return self.interval > 0 and naccept % self.interval == 0 and not final

    def initialize(self, semi, t0):
        """Adapt the initial condition until the mesh no longer changes,
        re-evaluating the initial condition on every new mesh (amr.jl:143-180)."""
        if not self.adapt_initial_condition:
            return
        only_refine = self.adapt_initial_condition_only_refine
        changed = self(semi, only_refine=only_refine)
        iterations = 1
        while changed:
            semi.set_initial_state(t0)  # the initial condition on the new mesh
            changed = self(semi, only_refine=only_refine)
            iterations += 1
            if iterations > max(10, self.controller.max_level):
                break
