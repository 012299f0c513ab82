"""Benchmark: DOF-updates/s of the DGSEM compressible-Euler RHS on one MI355X.

Workload (the metric's configuration): 3D Euler, periodic TreeMesh on [0,2]^3
at initial_refinement_level 5 (32^3 = 32768 elements), polydeg 3 (64 LGL
nodes per element, 2,097,152 DOFs), VolumeIntegralFluxDifferencing(flux_ranocha),
surface flux_lax_friedrichs, source_terms_convergence_test, state
initial_condition_convergence_test at t = 0.

One "step" = one right-hand-side evaluation over the whole mesh with inputs
resident in HBM. The FP32 RHS (FP32 fluxes, graph-replayed) is the headline;
the plain FP64 RHS is measured beside it, as is the equal-L2-error protocol
(both precisions integrated to the same final time, density L2 error against
the exact manufactured solution), and the FP64 CPU oracle on the host cores
(a port of Trixi's rhs_hyperbolic!, not Trixi itself).

To keep the 256 MB Infinity Cache from serving the state between calls, the
timed loop rotates over R input/output buffer pairs totalling > 512 MB.

Prints ONE JSON line.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "trixi.jl_amd"))

METRIC = "DOF-updates/s per RHS (3D Euler p3, 32^3): FP32 vs FP64 plain at equal L2 err"
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    # ~40 ms (FP32) / ~110 ms (FP64) of warm-up: enough for the clocks of a
    # freshly started GPU to settle before the timed region
    p.add_argument("--steps", type=int, default=1000)
    p.add_argument("--warmup", type=int, default=1000)
    p.add_argument("--level", type=int, default=5)
    p.add_argument("--polydeg", type=int, default=3)
    p.add_argument("--cpu-seconds", type=float, default=10.0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-l2", action="store_true")
    p.add_argument("--only", choices=["fp32", "fp64"], default=None,
                   help="time one precision only (profiling runs)")
    p.add_argument("--copy-calibration", action="store_true",
                   help="also run one plain 512 MiB device copy (calibrates FETCH_SIZE/WRITE_SIZE)")
    p.add_argument("--no-bounded", action="store_true",
                   help="skip the same RHS on the non-periodic mesh with Dirichlet boundaries")
    p.add_argument("--no-refined", action="store_true",
                   help="skip the RHS on the mesh with the central box refined once (L2 mortars)")
    p.add_argument("--no-amr", action="store_true",
                   help="skip the AMR run (elixir_euler_amr.jl to t = 0.1)")
    p.add_argument("--no-io", action="store_true",
                   help="skip the restart-file write/read timing")
    p.add_argument("--profile-amr", action="store_true",
                   help="only run the FP64 AMR workload once (for rocprofv3)")
    p.add_argument("--profile-step", choices=["fp32", "fp64"], default=None,
                   help="profiling mode: only run K device-controlled CK2N54 steps (fused stages) of the "
                        "workload in one precision and print the time per step (no JSON line)")
    p.add_argument("--no-configs", action="store_true",
                   help="skip the other benchmark configurations (3D p4 Fourier modes, 2D KHI with HG and with "
                        "the learned indicator)")
    p.add_argument("--profile-config", choices=["c3", "c4", "c5"], default=None,
                   help="profiling mode: only time that configuration's RHS (both precisions; no JSON line)")
    return p.parse_args()


def make_semi(th, level, polydeg, precision, tile=(0, 0), periodic=True, patches=()):
    mesh = th.TreeMesh((0.0, 0.0, 0.0), (2.0, 2.0, 2.0), initial_refinement_level=level, periodicity=periodic,
                       refinement_patches=patches)
    solver = th.DGSEM(polydeg=polydeg, surface_flux=th.flux_lax_friedrichs,
                      volume_integral=th.VolumeIntegralFluxDifferencing(th.flux_ranocha))
    ic = th.initial_condition_convergence_test
    bc = th.boundary_condition_periodic if periodic else th.BoundaryConditionDirichlet(ic)
    return th.SemidiscretizationHyperbolic(mesh, th.CompressibleEulerEquations3D(1.4), ic, solver,
                                           source_terms=th.source_terms_convergence_test, boundary_conditions=bc,
                                           precision=precision, elements_per_block=tile[0],
                                           threads_per_block=tile[1])


def make_semi_c3(th, precision, seed=0):
    """3D Euler p4 on [-1,1]^3 at level 5, seeded Fourier-mode state, flux_ranocha
    FD volume integral, flux_lax_friedrichs surface flux, no sources."""
    mesh = th.TreeMesh((-1.0,) * 3, (1.0,) * 3, initial_refinement_level=5, periodicity=True)
    solver = th.DGSEM(polydeg=4, surface_flux=th.flux_lax_friedrichs,
                      volume_integral=th.VolumeIntegralFluxDifferencing(th.flux_ranocha))
    return th.SemidiscretizationHyperbolic(mesh, th.CompressibleEulerEquations3D(1.4),
                                           th.initial_condition_fourier_modes(seed), solver, precision=precision)


def step_time(th, semi, steps):
    """Device-controlled CK2N54 steps (fused stages, step graph, device dt) from
    the initial state: microseconds per step, and the tiles used."""
    ode = th.semidiscretize(semi, (0.0, 1e9))
    th.solve(ode, th.CarpenterKennedy2N54(), callback=th.CallbackSet(th.StepsizeCallback(cfl=1.0)),
             maxiters=min(20, steps), save_final=False)  # graph capture
    sol = th.solve(ode, th.CarpenterKennedy2N54(), callback=th.CallbackSet(th.StepsizeCallback(cfl=1.0)),
                   maxiters=steps, save_final=False)
    return dict(us_per_step=1e6 * sol.wall_time / sol.nsteps, steps=sol.nsteps, stage_tile=semi.stage_tile_used())


def roofline_line(ndofs, nvars, bytes_per_value, ms):
    alg = 2 * nvars * ndofs * bytes_per_value  # read u once, write du once
    achieved = alg / (ms * 1e-3) / 1e9
    return dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                algorithmic_bytes=alg)


def config_c3(th, torch, steps, warmup, with_steps=True):
    """3D p4 32^3: RHS throughput (FP32, FP64) and the fused-stage time."""
    out = dict(workload="3D Euler, polydeg 4 (125 nodes/element), TreeMesh level 5 (32^3 elements) on [-1,1]^3, "
                        "seeded Fourier-mode state (seed 0), flux_ranocha FD volume, flux_lax_friedrichs surface")
    for prec in ("fp32", "fp64"):
        semi = make_semi_c3(th, prec)
        u0 = semi.compute_coefficients(0.0)
        r = time_rhs(th, torch, semi, steps, warmup, u0)
        nd = semi.ndofs()
        out[prec] = dict(value=nd / (r["ms_per_step"] * 1e-3), ms_per_step=r["ms_per_step"], tile=semi.tile_used(),
                         roofline=roofline_line(nd, 5, 4 if prec == "fp32" else 8, r["kernel_avg_ms"]),
                         finite=r["finite"])
        if with_steps:
            # the seeded velocity field steepens into shocks near t = 0.1, where
            # the plain flux-differencing solver fails (the oracle too: NaN at
            # step 19 on the level-3 mesh); 30 steps reach t = 0.05
            out[prec]["step"] = step_time(th, semi, 30)
        semi.close()
    out["fp32_over_fp64"] = out["fp32"]["value"] / out["fp64"]["value"]
    out["ndofs"] = nd
    return out


def config_c4_c5(th, torch, steps, t_state=1.0):
    """2D KHI p3 128^2 with HG shock capturing (C4) and with the learned
    indicator in its place (C5): RHS time (indicator kernels + element kernel,
    graph-replayed) on the FP64 HG run's state at t_state, the blended shares
    and, for C5, the share of elements that fell back to HG."""
    import numpy as np
    from trixi_hip import _lib
    from trixi_hip.learned_indicator import (IndicatorNeuralNetwork, blended_share, features, indicator_alpha,
                                             khi_semidiscretization, rhs_time, NFEATURES)
    semi = khi_semidiscretization(7, 1e-3, 0, "fp64")
    sol = th.solve(th.semidiscretize(semi, (0.0, t_state)), th.CarpenterKennedy2N54(),
                   callback=th.CallbackSet(th.StepsizeCallback(cfl=1.3)))
    u = sol.u
    semi.close()
    net = IndicatorNeuralNetwork.load()
    out = dict(workload=f"2D Euler Kelvin-Helmholtz, polydeg 3, TreeMesh level 7 (128^2 elements), HG shock "
                        f"capturing (alpha_max 0.002, alpha_min 1e-4, density_pressure), flux_ranocha volume, "
                        f"LLF (naive wave speed) surface and FV fluxes; state of the FP64 HG run at t = "
                        f"{t_state} ({sol.nsteps} steps, seed 0)", ndofs=None)
    for prec in ("fp32", "fp64"):
        sm = khi_semidiscretization(7, 1e-3, 0, prec)
        sm._lib.trixi_hip_upload_u(sm.ctx, u.ctypes.data_as(ctypes.c_void_p))
        net.upload(sm)
        nd = sm.ndofs()
        out["ndofs"] = nd
        for name, mode in (("c4_hg", _lib.IND_HG), ("c5_learned", net.mode)):
            ms = rhs_time(sm, u, mode, launches=max(steps, 100))
            el, tl = blended_share(sm, indicator_alpha(sm, mode))
            out.setdefault(name, {})[prec] = dict(value=nd / (ms * 1e-3), ms_per_rhs=ms,
                                                  ms_per_step=solve_step_ms(th, sm, u, max(steps, 100)),
                                                  blended_elements=el, blended_tiles=tl,
                                                  roofline=roofline_line(nd, 4, 4 if prec == "fp32" else 8, ms))
        f, _ = features(sm)
        out["c5_learned"]["fallback_share"] = float(np.mean(np.any(
            (f[:, :NFEATURES] < net.feature_lo) | (f[:, :NFEATURES] > net.feature_hi), axis=1)))
        sm.close()
    for prec in ("fp32", "fp64"):
        for what in ("rhs", "step"):
            out[f"learned_over_hg_{what}_time_{prec}"] = (out["c5_learned"][prec][f"ms_per_{what}"] /
                                                          out["c4_hg"][prec][f"ms_per_{what}"])
    return out


def solve_step_ms(th, semi, u, steps, reps=3):
    """Time per CarpenterKennedy2N54 step (fused stages, wave-speed reduction
    and device step control, replayed step graphs) from state u: the best of
    reps device-controlled solves of `steps` steps at CFL 1.3."""
    ode = th.semidiscretize(semi, (0.0, 1.0e3))
    ode.u0 = u
    best = math.inf
    for _ in range(reps):
        sol = th.solve(ode, th.CarpenterKennedy2N54(), callback=th.CallbackSet(th.StepsizeCallback(cfl=1.3)),
                       maxiters=steps, save_final=False)
        best = min(best, 1e3 * sol.wall_time / sol.nsteps)
    return best


def time_rhs(th, torch, semi, steps, warmup, u0):
    """Graph-replayed RHS throughput and per-launch kernel time (HIP events on
    the launch stream)."""
    ctx = semi.ctx
    lib = semi._lib
    dtype = torch.float32 if semi.precision == "fp32" else torch.float64
    nbytes_pair = 2 * semi.nvalues() * (4 if dtype == torch.float32 else 8)
    R = max(2, min(16, math.ceil(600e6 / nbytes_pair)))
    us = [torch.from_numpy(u0).to(device="cuda", dtype=dtype) for _ in range(R)]
    dus = [torch.empty_like(us[0]) for _ in range(R)]
    stream = torch.cuda.Stream()
    lib.trixi_hip_set_stream(ctx, ctypes.c_void_p(stream.cuda_stream))

    def launch(k):
        rc = lib.trixi_hip_rhs_native(ctx, ctypes.c_void_p(dus[k].data_ptr()), ctypes.c_void_p(us[k].data_ptr()), 0.0)
        if rc != 0:
            raise RuntimeError(lib.trixi_hip_last_error(ctx))

    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        for k in range(warmup):
            launch(k % R)
    torch.cuda.synchronize()
    # eager launches: HIP events on the launch stream bracket batches of LB
    # back-to-back launches (includes the host launch gaps)
    LB = 8
    nev = max(1, min(steps, 256) // LB)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(nev)]
    with torch.cuda.stream(stream):
        for k in range(nev):
            ev[k][0].record(stream)
            for j in range(LB):
                launch((k * LB + j) % R)
            ev[k][1].record(stream)
    torch.cuda.synchronize()
    kern_ms = sorted(a.elapsed_time(b) / LB for a, b in ev)
    kern_avg_ms = sum(kern_ms) / len(kern_ms)
    # graph of one pass over the R buffer pairs (plus a remainder graph)
    g_full = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g_full, stream=stream):
        for k in range(R):
            launch(k)
    rem = steps % R
    g_rem = None
    if rem:
        g_rem = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g_rem, stream=stream):
            for k in range(rem):
                launch(k)
    lib.trixi_hip_set_stream(ctx, ctypes.c_void_p(stream.cuda_stream))
    with torch.cuda.stream(stream):
        g_full.replay()
    torch.cuda.synchronize()
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    with torch.cuda.stream(stream):
        t_start.record(stream)
        for _ in range(steps // R):
            g_full.replay()
        if g_rem is not None:
            g_rem.replay()
        t_end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    total_ms = t_start.elapsed_time(t_end)
    # correctness guard on the timed buffers: finite output
    finite = bool(torch.isfinite(dus[0]).all().item())
    lib.trixi_hip_set_stream(ctx, None)
    del g_full, g_rem
    # average launch duration over the timed region: HIP events on the launch
    # stream around the graph replays (one element-kernel launch per step)
    return dict(ms_per_step=total_ms / steps, kernel_avg_ms=total_ms / steps, eager_launch_ms=kern_avg_ms,
                eager_launch_min_ms=kern_ms[0], wall_s=wall, rotate=R, finite=finite)


def equal_l2(th, semi64, semi32, nsteps=100, cfl=0.6):
    """Integrate both precisions to the same t (about nsteps CK2N54 steps) and
    compare the density L2 errors against the exact manufactured solution."""
    out = {}
    dt0 = th.StepsizeCallback(cfl=cfl)
    u0 = semi64.compute_coefficients(0.0)
    lib = semi64._lib
    lib.trixi_hip_upload_u(semi64.ctx, u0.ctypes.data_as(ctypes.c_void_p))
    t_end = nsteps * dt0.calculate_dt(semi64)
    for name, semi in (("fp64", semi64), ("fp32", semi32)):
        ode = th.semidiscretize(semi, (0.0, t_end))
        ac = th.AnalysisCallback(semi, interval=0)
        sol = th.solve(ode, th.CarpenterKennedy2N54(), callback=th.CallbackSet(th.StepsizeCallback(cfl=cfl), ac))
        out[name] = dict(l2_rho=sol.analysis[-1]["l2"][0], linf_rho=sol.analysis[-1]["linf"][0],
                         nsteps=sol.nsteps, t=sol.t, u=sol.u, wall_s=sol.wall_time)
    import numpy as np
    d = out["fp32"]["u"] - out["fp64"]["u"]
    nv = semi64.nvariables()
    diff_l2 = float(np.sqrt(np.mean(d.reshape(-1, nv)[:, 0] ** 2)))
    diff_linf = float(np.max(np.abs(d.reshape(-1, nv)[:, 0])))
    return dict(t_end=t_end, steps_fp64=out["fp64"]["nsteps"], steps_fp32=out["fp32"]["nsteps"],
                l2_rho_fp64=out["fp64"]["l2_rho"], l2_rho_fp32=out["fp32"]["l2_rho"],
                l2_rho_abs_diff=abs(out["fp32"]["l2_rho"] - out["fp64"]["l2_rho"]),
                rho_fp32_vs_fp64_rms=diff_l2, rho_fp32_vs_fp64_max=diff_linf,
                step_wall_ms_fp64=1e3 * out["fp64"]["wall_s"] / out["fp64"]["nsteps"],
                step_wall_ms_fp32=1e3 * out["fp32"]["wall_s"] / out["fp32"]["nsteps"])


# The process's affinity mask, read before any OpenMP runtime starts: with
# OMP_PROC_BIND set, the runtime binds the calling thread to one place, and the
# mask read afterwards is that single core.
START_AFFINITY = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)


def host_cores():
    """Cores this process may run on (its affinity mask at start), the
    machine's count, and the OpenMP setting the job was started with."""
    return dict(available=START_AFFINITY, cpu_count=os.cpu_count(), omp_num_threads=os.environ.get("OMP_NUM_THREADS"),
                omp_proc_bind=os.environ.get("OMP_PROC_BIND"), omp_places=os.environ.get("OMP_PLACES"))


def cpu_baseline(level, polydeg, seconds, u0):
    """The oracle's FP64 RHS on every core this process may use, one OpenMP
    thread per core, threads pinned (OMP_PROC_BIND/OMP_PLACES are set before
    the OpenMP runtime starts, see main)."""
    from oracle import Oracle
    hc = host_cores()
    threads = hc["available"]
    Oracle.set_threads(threads)
    o = Oracle(3, polydeg, level, (0, 0, 0), (2, 2, 2), volume_integral="fd", volume_flux="ranocha",
               surface_flux="llf", source="convergence_test")
    o.rhs(u0, 0.0)
    n = 0
    t0 = time.perf_counter()
    while True:
        o.rhs(u0, 0.0)
        n += 1
        el = time.perf_counter() - t0
        if el > seconds:
            break
    t_rhs = el / n
    return dict(value=o.ndofs / t_rhs, unit="DOF-updates/s", cores=threads, kind="port",
                sample=f"{n} full RHS calls of the same 3D p3 32^3 workload (FP64 C++ OpenMP restatement "
                       f"of Trixi rhs_hyperbolic!, not Trixi/Julia), {el:.1f} s, {threads} threads on the "
                       f"{hc['available']} cores of the process affinity mask at start (os.cpu_count() {hc['cpu_count']}, "
                       f"OMP_PROC_BIND={hc['omp_proc_bind']}, OMP_PLACES={hc['omp_places']})",
                time_per_dof_per_rhs_s=t_rhs * threads / o.ndofs, host=hc)


def measured_traffic(level, polydeg, precision):
    """HBM bytes per launch of the element kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (KiB per dispatch), each scaled by what the
    same counter reads for a plain 512 MiB copy. None for workloads that were
    not profiled."""
    path = os.path.join(ROOT, "profiles", "dg_element_kernel_traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        t = json.load(f)
    w = t["workloads"].get(f"euler3d_p{polydeg}_level{level}_{precision}")
    if w is None:
        return None, None
    c = t["calibration"]
    kib = w["fetch_kib"] * c["fetch_factor"] + w["write_kib"] * c["write_factor"]
    return kib * 1024.0, t["summary"]


def copy_bandwidth(torch, nbytes=1 << 30, reps=20):
    """On-box HBM bandwidth of a plain device-to-device copy of nbytes (read +
    write counted), best of reps launches timed with events; far above the
    256 MB Infinity Cache, so it measures HBM."""
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for a, b in ev:
        a.record()
        dst.copy_(src)
        b.record()
    torch.cuda.synchronize()
    ms = min(a.elapsed_time(b) for a, b in ev)
    del src, dst
    return 2 * nbytes / (ms * 1e-3) / 1e9


def amr_run(th, precision):
    """elixir_euler_amr.jl (test_tree_3d_euler.jl:145-165): 3D Euler p3 density
    pulse on [-5, 5]^3, TreeMesh level 4 adapted by ControllerThreeLevel
    (IndicatorMax(rho); levels 4/5/6 at 1.05/1.3) with the initial condition
    adapted first and an AMRCallback every 5 steps, CFL 0.9, t = 0 .. 0.1."""
    mesh = th.TreeMesh((-5.0, -5.0, -5.0), (5.0, 5.0, 5.0), initial_refinement_level=4, periodicity=True)
    semi = th.SemidiscretizationHyperbolic(mesh, th.CompressibleEulerEquations3D(1.4),
                                           th.initial_condition_density_pulse,
                                           th.DGSEM(polydeg=3, surface_flux=th.FluxLaxFriedrichs(th.max_abs_speed_naive)),
                                           precision=precision)
    ctl = th.ControllerThreeLevel(semi, th.IndicatorMax(semi, variable=th.first), base_level=4, med_level=5,
                                  med_threshold=1.05, max_level=6, max_threshold=1.3)
    amr = th.AMRCallback(semi, ctl, interval=5, adapt_initial_condition=True, adapt_initial_condition_only_refine=True)
    ac = th.AnalysisCallback(semi, interval=100)
    t0 = time.perf_counter()
    sol = th.solve(th.semidiscretize(semi, (0.0, 0.1)), th.CarpenterKennedy2N54(),
                   callback=th.CallbackSet(ac, amr, th.StepsizeCallback(cfl=0.9)))
    wall = time.perf_counter() - t0
    tm = amr.timer
    out = dict(wall_s=wall, steps=sol.nsteps, adaptations=tm["calls"], nelements=amr.nelements_history,
               ms_per_adaptation={k: 1e3 * tm[k] / max(tm["calls"], 1) for k in ("indicator", "mesh", "solver")},
               ms_per_step_excluding_adaptation=1e3 * (wall - tm["indicator"] - tm["mesh"] - tm["solver"]) / sol.nsteps,
               l2_rho=sol.analysis[-1]["l2"][0], linf_rho=sol.analysis[-1]["linf"][0])
    semi.close()
    return out


def amr_cpu_baseline():
    """The oracle's run of the same AMR workload on the host cores."""
    from oracle import ControllerThreeLevel, Oracle
    o = Oracle(3, 3, 4, (-5, -5, -5), (5, 5, 5), volume_integral="weak", surface_flux="llf_naive", periodicity=True)
    o.set_initial_condition("density_pulse")
    ctl = ControllerThreeLevel(base_level=4, med_level=5, med_threshold=1.05, max_level=6, max_threshold=1.3)
    t0 = time.perf_counter()
    _, _, nsteps, _ = o.solve_amr(0.0, 0.1, 0.9, ctl, interval=5)
    return dict(wall_s=time.perf_counter() - t0, steps=nsteps, threads=Oracle.get_threads(), kind="port")


def io_timing(th, semi, u0, directory):
    """Write and read one restart file of the state (device download included)."""
    import numpy as np
    semi.mesh.current_filename = th.save_mesh_file(semi.mesh, directory)
    assert semi._lib.trixi_hip_upload_u(semi.ctx, u0.ctypes.data_as(ctypes.c_void_p)) == 0
    from trixi_hip.io import download_state
    t0 = time.perf_counter()
    f = th.save_restart_file(download_state(semi), 0.0, 0.0, 1, semi, directory)
    t1 = time.perf_counter()
    u = th.load_restart_file(semi, f)
    t2 = time.perf_counter()
    nbytes = os.path.getsize(f)
    os.remove(f)
    return dict(file_bytes=nbytes, write_s=t1 - t0, read_s=t2 - t1, write_GBps=nbytes / (t1 - t0) / 1e9,
                read_GBps=nbytes / (t2 - t1) / 1e9, round_trip_exact=bool(np.array_equal(u, u0)))


def main():
    args = parse()
    # pin the CPU baseline's OpenMP threads, one per core; this must happen
    # before the first OpenMP runtime is loaded (torch loads one)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU path is not built "
              "in this round.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import trixi_hip as th
    torch.cuda.set_device(0)
    if args.profile_step:
        semi = make_semi(th, args.level, args.polydeg, args.profile_step)
        ode = th.semidiscretize(semi, (0.0, 10.0))
        sol = th.solve(ode, th.CarpenterKennedy2N54(), callback=th.CallbackSet(th.StepsizeCallback(cfl=1.0)),
                       maxiters=args.steps, save_final=False)
        print(f"[bench] {args.profile_step}: {sol.nsteps} CK2N54 steps, {1e6 * sol.wall_time / sol.nsteps:.1f} us "
              "per step (device step control, fused stages)", flush=True)
        semi.close()
        return

    if args.profile_config:
        if args.profile_config == "c3":
            r = config_c3(th, torch, args.steps, args.warmup, with_steps=False)
        else:
            r = config_c4_c5(th, torch, args.steps)
        print(f"[bench] {args.profile_config}: {r}", flush=True)
        return

    if args.profile_amr:
        for rep in ("cold", "warm"):  # the second run finds the kernels loaded
            r = amr_run(th, "fp64")
            print(f"[bench] amr fp64 ({rep}): {r}", flush=True)
        return

    res = {}
    semis = {}
    u0 = None
    for prec in ("fp32", "fp64"):
        semi = make_semi(th, args.level, args.polydeg, prec)
        semis[prec] = semi
        if u0 is None:
            u0 = semi.compute_coefficients(0.0)
        if args.only and prec != args.only:
            continue
        res[prec] = time_rhs(th, torch, semi, args.steps, args.warmup, u0)
        res[prec]["tile"] = semi.tile_used()
        res[prec]["stage_tile"] = semi.stage_tile_used()
        print(f"[bench] {prec}: {res[prec]}", file=sys.stderr, flush=True)

    if args.copy_calibration:
        src = torch.empty(128 * 1024 * 1024, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] copy calibration: {src.numel() * 4} bytes read, {dst.numel() * 4} bytes written",
              file=sys.stderr, flush=True)
        del src, dst

    bounded = {}
    if not args.no_bounded and not args.only:
        # the same RHS on [0,2]^3 with periodicity=false and
        # BoundaryConditionDirichlet(initial_condition) on all six sides
        for prec in ("fp32", "fp64"):
            sb = make_semi(th, args.level, args.polydeg, prec, periodic=False)
            rb = time_rhs(th, torch, sb, args.steps, args.warmup, u0)
            bounded[prec] = dict(value=sb.ndofs() / (rb["ms_per_step"] * 1e-3), ms_per_step=rb["ms_per_step"],
                                 nboundaries=sb.mesh.nboundaries, finite=rb["finite"])
            sb.close()
        print(f"[bench] bounded mesh: {bounded}", file=sys.stderr, flush=True)
    refined = {}
    if not args.no_refined and not args.only:
        # the same RHS with the leaf cells inside (0.5, 1.5)^3 refined once:
        # twice the resolution in the centre, L2 mortars around it
        for prec in ("fp32", "fp64"):
            sr = make_semi(th, args.level, args.polydeg, prec, patches=[((0.5, 0.5, 0.5), (1.5, 1.5, 1.5))])
            ur = sr.compute_coefficients(0.0)
            rr = time_rhs(th, torch, sr, args.steps, args.warmup, ur)
            refined[prec] = dict(value=sr.ndofs() / (rr["ms_per_step"] * 1e-3), ms_per_step=rr["ms_per_step"],
                                 nelements=sr.nelements(), nmortars=sr.mesh.nmortars, finite=rr["finite"])
            sr.close()
        print(f"[bench] refined mesh: {refined}", file=sys.stderr, flush=True)

    amr = {}
    if not args.no_amr and not args.only:
        for prec in ("fp64", "fp32"):
            amr[prec] = amr_run(th, prec)
            print(f"[bench] amr {prec}: {amr[prec]}", file=sys.stderr, flush=True)
        if not args.no_cpu:
            amr["cpu_oracle"] = amr_cpu_baseline()
            print(f"[bench] amr cpu: {amr['cpu_oracle']}", file=sys.stderr, flush=True)
    io_res = None
    if not args.no_io and not args.only:
        import tempfile
        with tempfile.TemporaryDirectory() as d:
            io_res = io_timing(th, semis["fp64"], u0, d)
        print(f"[bench] io: {io_res}", file=sys.stderr, flush=True)

    configs = {}
    if not args.no_configs and not args.only:
        configs["c3_3d_p4_32cubed"] = config_c3(th, torch, args.steps, args.warmup)
        print(f"[bench] c3: {configs['c3_3d_p4_32cubed']}", file=sys.stderr, flush=True)
        configs["c4_c5_khi_128sq"] = config_c4_c5(th, torch, args.steps)
        print(f"[bench] c4/c5: {configs['c4_c5_khi_128sq']}", file=sys.stderr, flush=True)
        from trixi_hip.learned_indicator import IndicatorNeuralNetwork, evaluate_config5
        # the learned indicator judged by the HG run's density error at t = 3
        # (seed 1000, not a training seed; reference: HG on the level-8 mesh)
        configs["c5_vs_hg_level7_t3"] = evaluate_config5(IndicatorNeuralNetwork.load(), seed=1000, level=7,
                                                         t_end=3.0, ref_level=8)
        print(f"[bench] c5 vs hg at t = 3: {configs['c5_vs_hg_level7_t3']}", file=sys.stderr, flush=True)

    ndofs = semis["fp32"].ndofs()
    nv = semis["fp32"].nvariables()
    copy_gbs = copy_bandwidth(torch)
    print(f"[bench] HBM copy: {copy_gbs:.0f} GB/s", file=sys.stderr, flush=True)
    l2 = None
    if not args.no_l2 and not args.only:
        l2 = equal_l2(th, semis["fp64"], semis["fp32"])
        print(f"[bench] equal-L2: {l2}", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu and not args.only:
        cpu = cpu_baseline(args.level, args.polydeg, args.cpu_seconds, u0)
        print(f"[bench] cpu: {cpu}", file=sys.stderr, flush=True)

    head = "fp32" if "fp32" in res else "fp64"
    r = res[head]
    bpv = 4 if head == "fp32" else 8
    alg_bytes = 2 * nv * ndofs * bpv  # read u once + write du once
    achieved = alg_bytes / (r["kernel_avg_ms"] * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic(args.level, args.polydeg, head)
    value = ndofs / (r["ms_per_step"] * 1e-3)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head == "fp32" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"3D compressible Euler MMS, TreeMesh level {args.level} "
                        f"({(2 ** args.level) ** 3} elements), polydeg {args.polydeg}, flux_ranocha FD volume, "
                        "flux_lax_friedrichs surface, source terms, FP32 fluxes, graph-replayed RHS",
            "ndofs": ndofs,
            "nelements": (2 ** args.level) ** 3,
            "element_tile": list(r["tile"]),
            "stage_tile": list(r["stage_tile"]) if r["stage_tile"] else None,
            "rotating_buffers": r["rotate"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": "dg_element_kernel (fused volume + interface + surface + Jacobian + source)",
            "kernel_avg_ms": r["kernel_avg_ms"],
            "algorithmic_bytes": alg_bytes,
            # the same achieved bandwidth against a plain 1 GiB copy measured on this box
            "measured_copy_gbs": copy_gbs,
            "frac_of_measured_copy": achieved / copy_gbs,
        },
        "cpu_baseline": None if cpu is None else {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")},
    }
    if "fp64" in res and "fp32" in res:
        v64 = ndofs / (res["fp64"]["ms_per_step"] * 1e-3)
        out["fp64_plain"] = {"value": v64, "ms_per_step": res["fp64"]["ms_per_step"],
                             "kernel_avg_ms": res["fp64"]["kernel_avg_ms"],
                             "tile": list(res["fp64"]["tile"]),
                             "roofline": roofline_line(ndofs, nv, 8, res["fp64"]["kernel_avg_ms"])}
        out["fp32_over_fp64"] = value / v64
    if bounded:
        out["bounded_mesh"] = dict(workload="same RHS, periodicity=false, BoundaryConditionDirichlet("
                                            "initial_condition) on all sides (boundary-state kernel included)",
                                   **bounded)
    if refined:
        out["refined_mesh"] = dict(workload="same RHS with refinement_patches = box (0.5,0.5,0.5)-(1.5,1.5,1.5): "
                                            "the central cells refined once, L2 mortars (mortar-flux kernel included)",
                                   **refined)
    if amr:
        out["amr"] = dict(workload="elixir_euler_amr.jl: 3D Euler p3 density pulse, TreeMesh levels 4-6 by "
                                   "ControllerThreeLevel(IndicatorMax(rho)), AMR every 5 steps, CFL 0.9, t = 0..0.1 "
                                   "(host-driven steps; adaptation = device indicator, host tree, new context + "
                                   "device transfer)", **amr)
    if io_res is not None:
        out["restart_io"] = dict(workload="restart file (HDF5, FP64 conserved variables) of the state above: "
                                          "device download + write, read + checks", **io_res)
    if configs:
        out["configs"] = configs
    if l2 is not None:
        out["equal_l2"] = l2
    if cpu is not None:
        out["gpu_fp32_over_cpu"] = value / cpu["value"]
        out["cpu_time_per_dof_per_rhs_s"] = cpu["time_per_dof_per_rhs_s"]
    ok = all(v["finite"] for v in res.values()) and all(v["finite"] for v in bounded.values()) and \
        all(v["finite"] for v in refined.values())
    out["outputs_finite"] = ok
    for s in semis.values():
        s.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
