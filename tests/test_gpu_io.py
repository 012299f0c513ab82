"""Restart and solution files written from device runs: a run resumed from a
restart file (and its mesh file, after AMR) continues bitwise like the run
that wrote it; a solution file holds the state and the HG indicator of the
step it was written at."""
import ctypes
import os

import numpy as np
import pytest

from conftest import golden
from gpu_helpers import build

torch = pytest.importorskip("torch")


def _run(semi, tspan, callbacks, u0=None, initial_step=0):
    import trixi_hip as th
    ode = th.semidiscretize(semi, tspan)
    if u0 is not None:
        ode.u0 = u0
    return th.solve(ode, th.CarpenterKennedy2N54(), callback=th.CallbackSet(*callbacks), initial_step=initial_step)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_restart_continues_bitwise(tmp_path, precision):
    """elixir_euler_source_terms.jl setup: 20 steps with a restart file every 10
    steps; the run restarted from step 10 (u, t, step count from the file) ends
    with the same state, time and step count, bit for bit."""
    import trixi_hip as th
    print(f"restart {precision}", flush=True)
    spec = dict(ndims=3, polydeg=3, level=2, cmin=(0, 0, 0), cmax=(2, 2, 2), ic="convergence_test",
                source="convergence_test", surface_flux="llf", volume_integral="weak")
    out = str(tmp_path / "out")
    semi, _ = build(spec, precision)
    rs = th.SaveRestartCallback(interval=10, save_final_restart=False, output_directory=out)
    sol = th.solve(th.semidiscretize(semi, (0.0, 1.0)), th.CarpenterKennedy2N54(),
                   callback=th.CallbackSet(rs, th.StepsizeCallback(cfl=0.6)), maxiters=20)
    assert [os.path.basename(f) for f in rs.files] == ["restart_000000010.h5", "restart_000000020.h5"]
    restart = rs.files[0]
    semi2, _ = build(spec, precision)
    t10 = th.load_time(restart)
    ode2 = th.semidiscretize(semi2, (t10, 1.0), restart_file=restart)
    sol2 = th.solve(ode2, th.CarpenterKennedy2N54(), callback=th.CallbackSet(th.StepsizeCallback(cfl=0.6)),
                    maxiters=10, initial_step=th.load_timestep(restart))
    assert sol2.nsteps == 20 and sol2.t == sol.t
    assert np.array_equal(sol2.u, sol.u)
    assert np.array_equal(th.load_restart_file(semi, rs.files[1]), sol.u)
    semi.close()
    semi2.close()


@pytest.mark.gpu
def test_amr_restart_continues_bitwise(tmp_path):
    """elixir_euler_amr.jl with a restart file every 5 steps written after the
    adaptation of that step: the run restarted from step 10 reads the adapted
    mesh (mesh_000000010.h5) and the state, continues with the AMR callback
    (no initial adaptation) and ends bitwise where the uninterrupted run ends."""
    import trixi_hip as th
    print("amr restart", flush=True)
    c = golden("kta.json")["amr_3d"]
    a = c["amr"]
    spec = dict(c["oracle"], ic=c["ic"], ic_params=())
    out = str(tmp_path / "out")

    def amr_for(semi, initial):
        ctl = th.ControllerThreeLevel(semi, th.IndicatorMax(semi, variable=th.first), base_level=a["base_level"],
                                      med_level=a["med_level"], med_threshold=a["med_threshold"],
                                      max_level=a["max_level"], max_threshold=a["max_threshold"])
        return th.AMRCallback(semi, ctl, interval=a["interval"], adapt_initial_condition=initial)

    semi, _ = build(spec, "fp64")
    amr = amr_for(semi, True)
    rs = th.SaveRestartCallback(interval=5, save_final_restart=False, output_directory=out)
    sol = _run(semi, tuple(c["tspan"]), [amr, rs, th.StepsizeCallback(cfl=c["cfl"])])
    restart = os.path.join(out, "restart_000000010.h5")
    assert th.read_solution_header(restart)["mesh_file"] == "mesh_000000010.h5"
    mesh2 = th.load_mesh(restart)
    semi2 = th.SemidiscretizationHyperbolic(mesh2, semi.equations, semi.initial_condition, semi.solver)
    assert semi2.nelements() == th.read_solution_header(restart)["n_elements"]
    ode2 = th.semidiscretize(semi2, (th.load_time(restart), c["tspan"][1]), restart_file=restart)
    sol2 = th.solve(ode2, th.CarpenterKennedy2N54(),
                    callback=th.CallbackSet(amr_for(semi2, False), th.StepsizeCallback(cfl=c["cfl"])),
                    initial_step=th.load_timestep(restart))
    assert sol2.nsteps == sol.nsteps and sol2.t == sol.t
    assert semi2.nelements() == semi.nelements()
    assert np.array_equal(sol2.u, sol.u)
    semi.close()
    semi2.close()


@pytest.mark.gpu
def test_solution_files_hold_state_and_hg_indicator(tmp_path):
    """SaveSolutionCallback on the KHI setup with HG shock capturing: files at
    steps 0, 2, 4; the last holds cons2prim of the final state and the blending
    factor of that state as element variable indicator_shock_capturing."""
    import trixi_hip as th
    from test_io import H5DUMP, dump
    if not os.path.exists(H5DUMP):
        pytest.skip("h5dump not in this image")
    print("solution files", flush=True)
    spec = dict(ndims=2, polydeg=3, level=4, cmin=(-1, -1), cmax=(1, 1), ic="kelvin_helmholtz",
                volume_integral="hg", volume_flux="ranocha", surface_flux="llf_naive", alpha_max=0.002,
                alpha_min=0.0001)
    out = str(tmp_path / "out")
    semi, _ = build(spec, "fp64")
    ss = th.SaveSolutionCallback(interval=2, output_directory=out)
    sol = th.solve(th.semidiscretize(semi, (0.0, 1.0)), th.CarpenterKennedy2N54(),
                   callback=th.CallbackSet(ss, th.StepsizeCallback(cfl=1.0)), maxiters=4)
    assert [os.path.basename(f) for f in ss.files] == ["solution_000000000.h5", "solution_000000002.h5",
                                                       "solution_000000004.h5"]
    assert os.path.exists(os.path.join(out, "mesh.h5"))
    f = ss.files[-1]
    prim = semi.equations.cons2prim(sol.u.reshape(-1, 4))
    for v in range(4):
        assert np.array_equal(np.array(dump(f, f"/variables_{v + 1}")), prim[:, v])
    alpha = np.empty(semi.nelements())
    assert semi._lib.trixi_hip_indicator(semi.ctx, 0, alpha.ctypes.data_as(ctypes.c_void_p)) == 0
    assert dump(f, "/element_variables_1/name", True) == ["indicator_shock_capturing"]
    assert np.array_equal(np.array(dump(f, "/element_variables_1")), alpha)
    assert dump(f, "/timestep", True) == [4] and dump(f, "/time", True) == [sol.t]
    semi.close()
