"""GPU (HIP, gfx950) against the CPU oracle on identical meshes and states.

The FP64 path evaluates every formula in the oracle's operation order and
gathers face contributions in the reference's face order, so the expected
result is bit-identical; the parity bar is 1e-12 relative L2 of U after
100 TVD-RK3 steps, asserted on top of the exact comparison."""
import numpy as np
import pytest

import netflux_amd as nf
from tests.cases import pulse_config, pulse_state, rel_l2, sod_config, sod_state
from tests.oracle_bind import Oracle

pytestmark = pytest.mark.gpu

TOL_REL_L2 = 1e-12


def _pair(mesh, cfg, U0, uold_equals_u=False):
    print(f"[case] cells={mesh.n_cells} scheme={cfg.time_scheme} order={cfg.spatial_order} "
          f"limiter={cfg.limiter}", flush=True)
    gpu = nf.EulerSolver(cfg, mesh)
    gpu.set_state(U0, U0 if uold_equals_u else None)
    ora = Oracle.from_mesh(mesh, cfg)
    ora.set_state(U0, uold_equals_u)
    return gpu, ora


@pytest.mark.parametrize("jitter", [0.0, 0.15])
def test_flux_balance_matches_oracle(jitter):
    mesh = nf.Mesh.box(8, jitter=jitter, seed=7)
    cfg = sod_config()
    gpu, ora = _pair(mesh, cfg, sod_state(mesh))
    gpu.eval_flux_balance()
    ora.eval_flux_balance()
    d = gpu.get_solver_data()
    np.testing.assert_array_equal(d.get_primvars_gradient(), ora.get("grad"))
    np.testing.assert_array_equal(d.get_primvars_limiter(), ora.get("psi"))
    np.testing.assert_array_equal(d.get_flux_balance(), ora.get("R"))


def test_farfield_flux_balance_matches_oracle():
    mesh = nf.Mesh.box(8, jitter=0.15, seed=3, patches=("FarField",) * 6)
    cfg = pulse_config()
    gpu, ora = _pair(mesh, cfg, pulse_state(mesh))
    gpu.eval_flux_balance()
    ora.eval_flux_balance()
    # pow() of the farfield ghost differs between the device and glibc in the
    # last bit at most; everything else is exact
    R_g, R_o = gpu.get_solver_data().get_flux_balance(), ora.get("R")
    assert np.max(np.abs(R_g - R_o)) <= 1e-13 * np.max(np.abs(R_o))


def test_sod_16cubed_100_rk3_steps():
    mesh = nf.Mesh.box(16)
    cfg = sod_config()
    U0 = sod_state(mesh)
    gpu, ora = _pair(mesh, cfg, U0)
    for n in range(100):
        gpu.calc_timestep(cfg)
        dt_o = ora.calc_timestep()
        assert cfg.delta_time == dt_o, f"dt differs at step {n}"
        gpu.step(cfg)
        ora.step(dt_o)
    Ug, Uo = gpu.get_solver_data().get_solution(), ora.get("U")
    err = rel_l2(Ug, Uo)
    print(f"rel L2 after 100 steps: {err:.3e}, identical: {np.array_equal(Ug, Uo)}")
    assert err <= TOL_REL_L2
    np.testing.assert_array_equal(Ug, Uo)


@pytest.mark.parametrize("scheme,order,limiter", [
    ("ExplicitEuler", "Second", "Barth"),
    ("TVD_RK3", "First", "NONE"),
    ("TVD_RK3", "Second", "NONE"),
])
def test_scheme_variants(scheme, order, limiter):
    mesh = nf.Mesh.box(6, jitter=0.1, seed=11)
    cfg = sod_config(time_scheme=scheme, spatial_order=order, limiter=limiter)
    gpu, ora = _pair(mesh, cfg, sod_state(mesh))
    for n in range(20):
        gpu.calc_timestep(cfg)
        gpu.step(cfg)
    ora.run(20)
    np.testing.assert_array_equal(gpu.get_solver_data().get_solution(), ora.get("U"))


def test_farfield_pulse_run_graph():
    mesh = nf.Mesh.box(8, jitter=0.15, seed=5, patches=("FarField",) * 6)
    cfg = pulse_config()
    gpu, ora = _pair(mesh, cfg, pulse_state(mesh))
    st = gpu.run(40, check_every=8, use_graph=True)
    assert st["steps_done"] == 40
    t_o = ora.run(40)
    assert abs(st["time"] - t_o) <= 1e-14 * t_o
    assert rel_l2(gpu.get_solver_data().get_solution(), ora.get("U")) <= TOL_REL_L2


def test_graph_equals_eager():
    mesh = nf.Mesh.box(6, jitter=0.1, seed=2)
    cfg = sod_config()
    U0 = sod_state(mesh)
    a = nf.EulerSolver(cfg, mesh)
    a.set_state(U0)
    b = nf.EulerSolver(cfg, mesh)
    b.set_state(U0)
    sa = a.run(30, check_every=7, use_graph=True)
    sb = b.run(30, check_every=7, use_graph=False)
    assert sa["steps_done"] == sb["steps_done"] == 30
    np.testing.assert_array_equal(a.get_solver_data().get_solution(), b.get_solver_data().get_solution())
    np.testing.assert_array_equal(a.residuals(), b.residuals())


def test_cpp_driver_matches_oracle(tmp_path):
    """The C++ host classes (netflux_solver.hpp) in the Driver loop order."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(nf.LIB_PATH), "netflux_driver")
    out = tmp_path / "U.bin"
    print("[case] netflux_driver 16 100 sod", flush=True)
    r = subprocess.run([exe, "16", "100", "sod", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    mesh = nf.Mesh.box(16)
    Ug = np.fromfile(out, dtype=np.float64).reshape(-1, 5)
    ora = Oracle.from_mesh(mesh, sod_config())
    ora.set_state(sod_state(mesh))
    ora.run(100)
    np.testing.assert_array_equal(Ug, ora.get("U"))


def test_mixed_patches_noslip_farfield():
    """All three boundary types in one mesh, explicit and RK3 steps."""
    mesh = nf.Mesh.box((6, 5, 4), jitter=0.12, seed=31,
                       patches=("FarField", "NoSlipWall", "SlipWall", "NoSlipWall", "FarField", "SlipWall"))
    cfg = pulse_config(mach=0.3)
    gpu, ora = _pair(mesh, cfg, pulse_state(mesh, mach=0.3), uold_equals_u=True)
    cfg.rk3_zero_old_first_step = False
    for _ in range(15):
        gpu.calc_timestep(cfg)
        gpu.step(cfg)
    ora.run(15)
    assert rel_l2(gpu.get_solver_data().get_solution(), ora.get("U")) <= TOL_REL_L2


def test_residual_history_matches_oracle():
    """r_k = ||R_rho||_2 at the first stage of step k."""
    mesh = nf.Mesh.box(6, jitter=0.1, seed=4)
    cfg = sod_config()
    U0 = sod_state(mesh)
    gpu, ora = _pair(mesh, cfg, U0)
    gpu.run(5, use_graph=False)
    res = gpu.residuals()
    assert len(res) == 5
    expect = []
    for _ in range(5):
        dt = ora.calc_timestep()  # from the primvars as last set (stale), before R(U^n)
        ora.eval_flux_balance()
        expect.append(np.linalg.norm(ora.get("R")[:, 0]))
        ora.step(dt)
    np.testing.assert_allclose(res, expect, rtol=1e-13)


@pytest.mark.parametrize("scheme", ["TVD_RK3", "ExplicitEuler"])
def test_flux_balance_between_steps_leaves_run_unchanged(scheme):
    """Evaluating the flux balance between steps (debug output) keeps the
    primvars and the device dt the last step left, so the trajectory is
    the oracle's."""
    mesh = nf.Mesh.box(6, jitter=0.1, seed=9)
    cfg = sod_config(time_scheme=scheme)
    gpu, ora = _pair(mesh, cfg, sod_state(mesh))
    for n in range(6):
        gpu.eval_flux_balance()
        gpu.calc_timestep(cfg)
        dt = ora.calc_timestep()
        assert cfg.delta_time == dt, f"dt differs at step {n}"
        gpu.step(cfg)
        ora.step(dt)
    np.testing.assert_array_equal(gpu.get_solver_data().get_solution(), ora.get("U"))


def _quantised_state(mesh, seed):
    """Primitive values drawn from two or three levels per variable, so that
    many neighbour differences, limiter numerators and gradients are exact
    zeros or exact ties."""
    from tests.cases import prim_to_cons
    rng = np.random.default_rng(seed)
    N = mesh.n_cells
    V = np.zeros((N, 5))
    V[:, 0] = rng.choice([1.0, 1.5], N)
    V[:, 1] = rng.choice([0.0, 0.25, -0.25], N)
    V[:, 2] = rng.choice([0.0, 0.5], N)
    V[:, 3] = 0.0
    V[:, 4] = rng.choice([1.0, 2.0], N)
    return prim_to_cons(V)


@pytest.mark.parametrize("seed", [1, 2])
def test_limiter_ties_and_zeros_match_oracle(seed):
    """Barth with exact zero numerators, ties between faces and zero
    gradient rows (the grouped limiter's edge cases) against the oracle's
    face-by-face limiter, and 10 RK3 steps from there."""
    mesh = nf.Mesh.box(8, jitter=0.1, seed=seed)
    cfg = sod_config()
    gpu, ora = _pair(mesh, cfg, _quantised_state(mesh, seed))
    gpu.eval_flux_balance()
    ora.eval_flux_balance()
    d = gpu.get_solver_data()
    np.testing.assert_array_equal(d.get_primvars_limiter(), ora.get("psi"))
    np.testing.assert_array_equal(d.get_primvars_gradient(), ora.get("grad"))
    np.testing.assert_array_equal(d.get_flux_balance(), ora.get("R"))
    for _ in range(10):
        gpu.calc_timestep(cfg)
        gpu.step(cfg)
    ora.run(10)
    np.testing.assert_array_equal(d.get_solution(), ora.get("U"))


def test_limiter_tiny_gradients_match_oracle():
    """Velocity components of order 1e-250: limiter denominators Delta_2 =
    d + eps round to eps, which the grouped limiter must see as positive."""
    from tests.cases import prim_to_cons
    mesh = nf.Mesh.box(6, jitter=0.1, seed=3)
    N = mesh.n_cells
    rng = np.random.default_rng(5)
    V = np.zeros((N, 5))
    V[:, 0] = 1.0
    V[:, 1] = 1e-250 * rng.random(N)
    V[:, 2] = -1e-250 * rng.random(N)
    V[:, 4] = 1.0
    cfg = sod_config()
    gpu, ora = _pair(mesh, cfg, prim_to_cons(V))
    gpu.eval_flux_balance()
    ora.eval_flux_balance()
    d = gpu.get_solver_data()
    psi = ora.get("psi")
    assert np.any((psi[:, 1] > 0) & (psi[:, 1] < 1))  # the tiny rows are limited
    np.testing.assert_array_equal(d.get_primvars_limiter(), psi)
    np.testing.assert_array_equal(d.get_primvars_gradient(), ora.get("grad"))


def test_primvars_export_after_graph_replays():
    """Steps do not store the primvars; the export evaluates prim(U) of the
    current state, also after graph replays that follow an earlier export."""
    mesh = nf.Mesh.box(6, jitter=0.1, seed=8)
    cfg = sod_config()
    a = nf.EulerSolver(cfg, mesh)
    a.set_state(sod_state(mesh))
    a.run(8, check_every=4, use_graph=True)  # capture + replay
    a.get_solver_data().get_primvars()
    a.run(8, check_every=4, use_graph=True)  # replays only
    Va = a.get_solver_data().get_primvars()
    b = nf.EulerSolver(cfg, mesh)
    b.set_state(a.get_solver_data().get_solution())
    np.testing.assert_array_equal(Va, b.get_solver_data().get_primvars())


# ---- HLLC (Toro's scheme; the reference's is a stub) ----

@pytest.mark.parametrize("scheme,order,limiter", [
    ("TVD_RK3", "Second", "Barth"),
    ("ExplicitEuler", "First", "NONE"),
])
def test_hllc_sod_matches_oracle(scheme, order, limiter):
    mesh = nf.Mesh.box(10, jitter=0.1, seed=21)
    cfg = sod_config(time_scheme=scheme, spatial_order=order, limiter=limiter, inviscid_flux_scheme="HLLC")
    gpu, ora = _pair(mesh, cfg, sod_state(mesh))
    gpu.eval_flux_balance()
    ora.eval_flux_balance()
    np.testing.assert_array_equal(gpu.get_solver_data().get_flux_balance(), ora.get("R"))
    for _ in range(30):
        gpu.calc_timestep(cfg)
        gpu.step(cfg)
    ora.run(30)
    np.testing.assert_array_equal(gpu.get_solver_data().get_solution(), ora.get("U"))


def test_hllc_farfield_pulse_run_graph():
    mesh = nf.Mesh.box(8, jitter=0.15, seed=6, patches=("FarField",) * 6)
    cfg = pulse_config(mach=0.7, inviscid_flux_scheme="HLLC")
    gpu, ora = _pair(mesh, cfg, pulse_state(mesh, mach=0.7))
    st = gpu.run(30, check_every=10, use_graph=True)
    assert st["steps_done"] == 30
    ora.run(30)
    assert rel_l2(gpu.get_solver_data().get_solution(), ora.get("U")) <= TOL_REL_L2


def test_hllc_needs_fp64():
    mesh = nf.Mesh.box(4)
    with pytest.raises(RuntimeError):
        nf.EulerSolver(sod_config(inviscid_flux_scheme="HLLC", precision="mixed"), mesh)
    s = nf.EulerSolver(sod_config(inviscid_flux_scheme="HLLC"), mesh)
    with pytest.raises(RuntimeError, match="FP64 only"):
        s.set_precision("mixed")


def test_oblique_supersonic_farfield_matches_oracle():
    """Every characteristic regime of the farfield boundary on the device:
    freestream velocity (1.3, 0.75, 0) with c = 1, so the x-faces see
    supersonic inflow and outflow (|v_n| = 1.3), the y-faces subsonic inflow
    and outflow (|v_n| = 0.75, tangential velocity nonzero) and the z-faces
    v_n = 0; a density pulse carried through the boundaries makes the domain
    state differ from the freestream there.  80 TVD-RK3 steps, 1e-12 relative
    L2 against the oracle (src/Numerics.cpp:83-171)."""
    mesh = nf.Mesh.box(8, jitter=0.12, seed=41, patches=("FarField",) * 6)
    cfg = pulse_config(velocity_fs=(1.3, 0.75, 0.0))
    U0 = pulse_state(mesh, centre=(0.7, 0.6, 0.5))
    V = np.zeros((mesh.n_cells, 5))
    V[:, 0] = U0[:, 0]
    V[:, 1], V[:, 2] = 1.3, 0.75
    V[:, 4] = 1 / 1.4
    from tests.cases import prim_to_cons
    U0 = prim_to_cons(V)
    gpu, ora = _pair(mesh, cfg, U0)
    errs = []
    for n in range(80):
        gpu.calc_timestep(cfg)
        dt_o = ora.calc_timestep()
        assert abs(cfg.delta_time - dt_o) <= 1e-13 * dt_o
        gpu.step(cfg)
        ora.step(cfg.delta_time)
        if n in (9, 39, 79):
            errs.append(rel_l2(gpu.get_solver_data().get_solution(), ora.get("U")))
    print("rel L2 after 10, 40, 80 steps:", errs)
    assert max(errs) <= TOL_REL_L2
    assert np.abs(ora.get("U") - U0).max() > 1e-3  # the flow has moved


def test_single_tetrahedron_and_ragged_box():
    """Ragged extremes of the tetrahedral sweeps: one cell in a 256-lane block
    (255 padding lanes, every face a boundary face, no interior face so the
    reference's time step is DBL_MAX), and a 3 x 2 x 5 box (180 cells: the
    block's third wave is partly padding) stepped 20 times against the
    oracle, bit for bit."""
    nodes = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.1, 0.2, 0.9]], float)
    cells = np.array([[0, 1, 2, 3]], np.uint32)
    faces = [[[0, 2, 1]], [[0, 1, 3]], [[0, 3, 2]], [[1, 2, 3]]]
    one = nf.Mesh.from_arrays(nodes, cells, faces, ["SlipWall", "FarField", "SlipWall", "FarField"])
    cfg = pulse_config(mach=0.4)
    gpu, ora = _pair(one, cfg, pulse_state(one))
    gpu.calc_timestep(cfg)
    assert cfg.delta_time == ora.calc_timestep() == np.finfo(float).max
    gpu.eval_flux_balance()
    ora.eval_flux_balance()
    R_g, R_o = gpu.get_solver_data().get_flux_balance(), ora.get("R")
    assert np.all(np.isfinite(R_g)) and np.max(np.abs(R_g - R_o)) <= 1e-13 * np.max(np.abs(R_o))
    box = nf.Mesh.box((3, 2, 5), lengths=(0.6, 0.4, 1.0), jitter=0.15, seed=5)
    assert box.n_cells == 180
    cfg = sod_config()
    gpu, ora = _pair(box, cfg, sod_state(box))
    for _ in range(20):
        gpu.calc_timestep(cfg)
        dt = ora.calc_timestep()
        assert cfg.delta_time == dt
        gpu.step(cfg)
        ora.step(dt)
    np.testing.assert_array_equal(gpu.get_solver_data().get_solution(), ora.get("U"))
