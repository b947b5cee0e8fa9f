"""The benchmark workload at its full size: 128^3 hexahedra split into 6
tetrahedra (12,582,912 cells, 25,264,128 faces), interior nodes jittered by
+-15 % of the spacing, Mach 0.5 freestream plus a Gaussian density pulse,
farfield on all six sides.  At this size the packed neighbour words (29-bit
cell index + slot), the 64-cell tiles, the per-wave node tables and the
block / reduction tables are exercised at the shapes the benchmark times.

Checks: two TVD-RK3 steps against the oracle (the reference's first-step
zero-U_old quirk included), held to 1e-12 relative L2 (farfield faces use
the device pow, 1e-13 level); a uniform freestream stays uniform to
round-off; with slip walls the total mass and energy stay constant to
round-off (each interior face flux enters its two cells exactly negated).
The mixed-element path is checked against the oracle at the size its own
benchmark times (96^3 hexahedra)."""
import math

import numpy as np
import pytest

import netflux_amd as nf
from tests.cases import prim_to_cons, pulse_config, pulse_state, rel_l2
from tests.oracle_bind import Oracle

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

SEED = 20260101


@pytest.fixture(scope="module")
def farfield_mesh():
    return nf.Mesh.box(128, jitter=0.15, seed=SEED, patches=("FarField",) * 6)


def test_benchmark_mesh_two_steps_vs_oracle(farfield_mesh):
    mesh = farfield_mesh
    print(f"[case] 128^3x6 = {mesh.n_cells} cells, pulse, 2 RK3 steps vs oracle", flush=True)
    assert mesh.n_cells == 12_582_912
    U0 = pulse_state(mesh)
    cfg = pulse_config()
    g = nf.EulerSolver(cfg, mesh)
    assert g.sizes.n_total_faces == 25_264_128
    g.set_state(U0)
    dts = []
    for _ in range(2):
        g.calc_timestep(cfg)
        dts.append(cfg.delta_time)
        g.step(cfg)
    Ug = g.get_solver_data().get_solution()
    del g
    print("[case] GPU steps done; oracle", flush=True)
    o = Oracle.from_mesh(mesh, pulse_config())
    o.set_state(U0)
    dto = []
    for _ in range(2):
        dto.append(o.calc_timestep())
        o.step(dto[-1])
    Uo = o.get("U")
    err = rel_l2(Ug, Uo)
    print(f"dt gpu {dts} oracle {dto}; rel L2 {err:.3e}", flush=True)
    assert dts == dto
    assert err < 1e-12


def test_benchmark_mesh_freestream_preserved(farfield_mesh):
    """rho = 1, u = (0.5, 0, 0), p = 1/gamma everywhere with farfield at the
    same state: every face flux balances, 10 graph-replayed steps leave the
    state at round-off distance."""
    mesh = farfield_mesh
    print("[case] 128^3x6 freestream, 10 steps", flush=True)
    V = np.tile([1.0, 0.5, 0.0, 0.0, 1 / 1.4], (mesh.n_cells, 1))
    U0 = prim_to_cons(V)
    cfg = pulse_config(rk3_zero_old_first_step=False)
    g = nf.EulerSolver(cfg, mesh)
    g.set_state(U0, U0)
    st = g.run(10, use_graph=True, check_every=5)
    assert st["steps_done"] == 10
    U = g.get_solver_data().get_solution()
    d = np.abs(U - U0).max()
    print(f"max |U - U0| after 10 steps: {d:.2e}", flush=True)
    assert d < 1e-12


def test_benchmark_mesh_slip_walls_conserve():
    mesh = nf.Mesh.box(128, jitter=0.15, seed=SEED)  # slip walls on all six sides
    print("[case] 128^3x6 slip walls, pulse, 10 steps: conservation", flush=True)
    U0 = pulse_state(mesh)
    vol = mesh.fv_grid()["volume"][: mesh.n_cells]
    g = nf.EulerSolver(pulse_config(rk3_zero_old_first_step=False), mesh)
    g.set_state(U0, U0)
    st = g.run(10, use_graph=True, check_every=5)
    assert st["steps_done"] == 10
    U = g.get_solver_data().get_solution()
    assert np.abs(U - U0).max() > 1e-6  # the pulse has moved
    for k in (0, 4):
        m0, m1 = math.fsum(vol * U0[:, k]), math.fsum(vol * U[:, k])  # exact sums of the products
        print(f"variable {k}: total {m0:.17g} -> {m1:.17g}, relative drift {abs(m1 - m0) / m0:.2e}", flush=True)
        assert abs(m1 - m0) / m0 < 1e-12


def test_hex_benchmark_mesh_two_steps_vs_oracle():
    """The mixed-element path at the size `bench.py --mesh hex` times: 96^3
    jittered hexahedra (884,736 cells, six face slots each), Mach 0.5 +
    pulse, farfield; two TVD-RK3 + Barth steps against the oracle's
    mixed-element grid, held to 1e-12 relative L2 (farfield faces use the
    device pow), with the same time steps."""
    mesh = nf.Mesh.hex_box(96, jitter=0.15, seed=SEED, patches=("FarField",) * 6)
    print(f"[case] hex 96^3 = {mesh.n_cells} cells, pulse, 2 RK3 steps vs oracle", flush=True)
    assert mesh.n_cells == 884_736
    U0 = pulse_state(mesh)
    cfg = pulse_config(mach=0.5)
    g = nf.EulerSolver(cfg, mesh)
    g.set_state(U0)
    dts = []
    for _ in range(2):
        g.calc_timestep(cfg)
        dts.append(cfg.delta_time)
        g.step(cfg)
    Ug = g.get_solver_data().get_solution()
    del g
    print("[case] GPU steps done; oracle", flush=True)
    o = Oracle.from_mesh(mesh, pulse_config(mach=0.5))
    o.set_state(U0)
    dto = []
    for _ in range(2):
        dto.append(o.calc_timestep())
        o.step(dto[-1])
    err = rel_l2(Ug, o.get("U"))
    print(f"dt gpu {dts} oracle {dto}; rel L2 {err:.3e}", flush=True)
    assert dts == dto
    assert err < 1e-12
