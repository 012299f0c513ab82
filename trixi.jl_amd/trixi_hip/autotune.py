"""Workgroup-tile autotuner of the element kernel (needs a GPU).

`autotune(semi)` times every instantiated tile of the context's element kernel
(the RHS variant and the fused 2N-stage variant separately) on its resident
state and makes the fastest the context's tile and this process's default for
later contexts of the same kernel variant (trixi_hip_autotune in the library).

`python -m trixi_hip.autotune` sweeps the benchmark workloads and writes the
measured table the library compiles in (csrc/tile_table.inc: the default tile
of each variant when a configuration asks for elements_per_block = 0) and a
summary of every candidate's time. Rebuild the library afterwards.
"""
import argparse
import ctypes
import datetime
import json
import os
import sys

from . import _lib

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE = os.path.join(PKG_ROOT, "csrc", "tile_table.inc")

def autotune(semi, repetitions=20):
    """Tune the element-kernel tiles of `semi` on its resident state. Returns
    {"rhs": (epb, nt), "stage": (epb, nt) or None, "timings": [dict]}."""
    lib, ctx = semi._lib, semi.ctx
    buf = (_lib.TileTiming * 64)()
    n = ctypes.c_int()
    _lib.check(lib.trixi_hip_autotune(ctx, int(repetitions), buf, 64, ctypes.byref(n)), ctx, "autotune")
    timings = [dict(stage=int(t.stage), elements_per_block=int(t.elements_per_block),
                    threads_per_block=int(t.threads_per_block), us=float(t.microseconds))
               for t in buf[:min(n.value, 64)]]
    return dict(rhs=semi.tile_used(), stage=semi.stage_tile_used(), timings=timings)


def _workloads(th):
    """(name, description, builder(precision) -> semi) of the swept workloads."""
    from .learned_indicator import khi_semidiscretization

    def euler3d(polydeg, fourier):
        def make(prec):
            mesh = th.TreeMesh((-1.0,) * 3 if fourier else (0.0,) * 3, (1.0,) * 3 if fourier else (2.0,) * 3,
                               initial_refinement_level=5, periodicity=True)
            solver = th.DGSEM(polydeg=polydeg, surface_flux=th.flux_lax_friedrichs,
                              volume_integral=th.VolumeIntegralFluxDifferencing(th.flux_ranocha))
            if fourier:
                ic, src = th.initial_condition_fourier_modes(0), None
            else:
                ic, src = th.initial_condition_convergence_test, th.source_terms_convergence_test
            return th.SemidiscretizationHyperbolic(mesh, th.CompressibleEulerEquations3D(1.4), ic, solver,
                                                   source_terms=src, precision=prec)
        return make

    def khi(prec):
        return khi_semidiscretization(level=7, noise_amplitude=1e-3, seed=0, precision=prec)

    return [
        ("euler3d_p3_32cubed_mms", "3D Euler MMS, p3, 32^3 elements, flux_ranocha FD (the headline RHS)",
         euler3d(3, False)),
        ("euler3d_p4_32cubed_fourier", "3D Euler seeded Fourier modes, p4, 32^3 elements, flux_ranocha FD",
         euler3d(4, True)),
        ("khi2d_p3_128sq_hg", "2D Euler Kelvin-Helmholtz, p3, 128^2 elements, HG shock capturing (alpha_max "
                              "0.002)", khi),
    ]


def vol_kind(cfg):
    """The element kernel's volume-kernel code (VolumeKind, csrc/dg_common.hpp)
    of a configuration, as the library derives it."""
    ran = cfg.volume_flux == _lib.FLUX_RANOCHA
    return {_lib.VI_WEAK: 0, _lib.VI_FLUX_DIFFERENCING: 1 if ran else 2,
            _lib.VI_SHOCK_CAPTURING_HG: 3 if ran else 4, _lib.VI_PURE_FV: 5}[cfg.volume_integral]


def _upload(semi):
    u = semi.compute_coefficients(0.0)
    _lib.check(semi._lib.trixi_hip_upload_u(semi.ctx, u.ctypes.data_as(ctypes.c_void_p)), semi.ctx, "upload_u")


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("--repetitions", type=int, default=40)
    p.add_argument("--table", default=TABLE, help="where to write the compiled-in tile table")
    p.add_argument("--summary", default=None, help="markdown summary of every candidate's time")
    p.add_argument("--json", default=None, help="raw results as JSON")
    a = p.parse_args(argv)
    import torch
    import trixi_hip as th
    torch.cuda.set_device(0)
    rows, entries = [], {}
    for name, desc, make in _workloads(th):
        for prec in ("fp32", "fp64"):
            semi = make(prec)
            _upload(semi)
            default = (semi.tile_used(), semi.stage_tile_used())
            r = autotune(semi, a.repetitions)
            cfg = semi.config()
            key = (1 if prec == "fp32" else 0, cfg.ndims, cfg.polydeg, vol_kind(cfg),
                   1 if cfg.nboundaries + cfg.nmortars > 0 else 0)
            for st, tile in ((0, r["rhs"]), (1, r["stage"])):
                if tile is not None:
                    entries[key + (st,)] = tile
            rows.append(dict(workload=name, description=desc, precision=prec, default_rhs=default[0],
                             default_stage=default[1], chosen_rhs=r["rhs"], chosen_stage=r["stage"],
                             timings=r["timings"]))
            print(f"[autotune] {name} {prec}: rhs {r['rhs']} stage {r['stage']} "
                  + ", ".join(f"{'S' if x['stage'] else 'R'}{x['elements_per_block']}x{x['threads_per_block']}="
                              f"{x['us']:.2f}us" for x in r["timings"]), flush=True)
            semi.close()
    stamp = datetime.datetime.now(datetime.timezone.utc).strftime("%Y-%m-%d")
    lines = [f"// Measured by python -m trixi_hip.autotune --repetitions {a.repetitions} on MI355X (gfx950), "
             f"{stamp}.",
             "// {fp32, ndims, polydeg, volume kernel, boundaries, stage, elements per block, threads}"]
    for k in sorted(entries):
        lines.append("{%d, %d, %d, %d, %d, %d, %d, %d}," % (k + tuple(entries[k])))
    with open(a.table, "w") as f:
        f.write("\n".join(lines) + "\n")
    if a.json:
        with open(a.json, "w") as f:
            json.dump(rows, f, indent=1)
    if a.summary:
        with open(a.summary, "w") as f:
            f.write(summary_markdown(rows, a.repetitions))
    print(json.dumps({"table": a.table, "entries": len(entries)}))


def summary_markdown(rows, repetitions):
    cmd = "python3 -m trixi_hip.autotune " + " ".join(sys.argv[1:])
    out = ["# Element-kernel tile sweep (`python -m trixi_hip.autotune`)", "", "Command (repository root, "
           "trixi.jl_amd on PYTHONPATH, one MI355X):", "```", cmd, "```", "",
           f"Each candidate: 3 rounds of {repetitions} launches on the workload's resident state, best round, "
           "HIP events on the context's stream. R = RHS kernel (`trixi_hip_rhs_native`), S = fused 2N-stage "
           "kernel (first stage with a = b = 0, which leaves the state unchanged). Tile = elements per "
           "workgroup x threads.", "",
           "| workload | precision | candidate | µs per launch | chosen |", "|---|---|---|---|---|"]
    for r in rows:
        for t in sorted(r["timings"], key=lambda x: (x["stage"], x["elements_per_block"])):
            tile = (t["elements_per_block"], t["threads_per_block"])
            chosen = tile == tuple(r["chosen_rhs"] if t["stage"] == 0 else (r["chosen_stage"] or (0, 0)))
            out.append(f"| {r['workload']} | {r['precision']} | {'S' if t['stage'] else 'R'} {tile[0]}x{tile[1]} | "
                       f"{t['us']:.2f} | {'yes' if chosen else ''} |")
    return "\n".join(out) + "\n"


if __name__ == "__main__":
    sys.exit(main())
