"""Sim directories on the GPU: the reference's own input.yml files
(tests/golden/sim_inputs) over a generated SU2 mesh whose markers carry
their patch names, run by the front end's Driver loop and checked against
the oracle stepping the same mesh; VTK files per write_stride."""
import os
import shutil

import numpy as np
import pytest

import netflux_amd as nf
from netflux_amd import frontend as fe
from tests.oracle_bind import Oracle
from tests.test_frontend import three_patch_box

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "sim_inputs")


def _sim_dir(tmp_path, name, n_timesteps, names):
    d = tmp_path / name
    d.mkdir()
    text = open(os.path.join(GOLD, name + ".yml")).read()
    lines = [f"n_timesteps: {n_timesteps}" if l.startswith("n_timesteps:") else l for l in text.split("\n")]
    (d / "input.yml").write_text("\n".join(lines))
    return d


@pytest.mark.parametrize("name,n_steps", [("small_brick_domain", 20), ("very_fine_brick_domain_first", 40),
                                          ("fine_box_domain", 12)])
def test_sim_dir_matches_oracle(tmp_path, name, n_steps):
    print(f"[case] sim dir {name}, {n_steps} steps", flush=True)
    cfg, opts = fe.parse_input(os.path.join(GOLD, name + ".yml"))
    names = list(opts["patches"])
    mesh = three_patch_box(6, jitter=0.12, seed=17)
    if len(names) == 1:
        nodes, cells, pn, pf, _ = mesh.arrays()
        mesh = nf.Mesh.from_arrays(nodes, cells, [pf], ["SlipWall"])
    d = _sim_dir(tmp_path, name, n_steps, names)
    mesh.write_su2(d / "mesh.su2", names)
    logs = []
    solver, run_cfg = fe.run_sim_dir(str(d), log=logs.append)
    assert run_cfg.timestep == n_steps
    # the oracle on the same mesh and configuration, Driver loop order
    cfg.n_timesteps = n_steps
    ora = Oracle.from_mesh(mesh, cfg)
    ora.set_freestream()
    t = 0.0
    for _ in range(n_steps):
        dt = ora.calc_timestep()
        ora.step(dt)
        t += dt
    U = solver.get_solver_data().get_solution()
    np.testing.assert_array_equal(U, ora.get("U"))
    assert run_cfg.time == t
    stride = opts["write_stride"]
    files = sorted(os.listdir(d / "output"))
    assert files == sorted(f"n={k}.vtk" for k in range(0, n_steps + 1, stride))
    last = fe.read_vtk_cell_data(d / "output" / f"n={n_steps - n_steps % stride}.vtk")
    if n_steps % stride == 0:
        np.testing.assert_allclose(last["density"], U[:, 0], rtol=5e-6)
    if opts["write_vtk_debug"]:
        assert "limiter_EQ4" in last and "gradient_EQ0" in last
    shutil.rmtree(d / "output")


def test_sim_dir_cli(tmp_path):
    d = _sim_dir(tmp_path, "cyprien_hex_tutorial", 3, ["walls"])
    nodes, cells, pn, pf, _ = nf.Mesh.box(4, jitter=0.1, seed=5).arrays()
    nf.Mesh.from_arrays(nodes, cells, [pf], ["SlipWall"]).write_su2(d / "mesh.su2", ["walls"])
    print("[case] frontend main cyprien_hex_tutorial 3 steps", flush=True)
    assert fe.main([str(d), "--full-precision"]) == 0
    assert sorted(os.listdir(d / "output")) == ["n=0.vtk"]  # write_stride 10
    # a marker missing from input.yml is an error
    nf.Mesh.from_arrays(nodes, cells, [pf], ["SlipWall"]).write_su2(d / "mesh.su2", ["other"])
    assert fe.main([str(d), "--no-output"]) == 1
