"""Error paths of the tetrahedral sweeps against the oracle.

ValidityChecker::check_flux_balance_validity (src/SolverData.cpp:66-79)
throws at the end of evaluate_flux_balance, before that stage's update
(src/Solver.cpp:80, 115-132), with "<n> cells with unphysical values detected
in the flux_balance field of the Euler solver"; calc_timestep throws
"Invalid dt calculated (dt = ...)" (src/Solver.cpp:267-268).  The GPU path
must raise the same message, with the count of that one evaluation, leave
the state, time and residual history where the failing step found them, and
do so from a single step, from a graph-replayed run and in mixed precision."""
import numpy as np
import pytest

import netflux_amd as nf
from tests.cases import prim_to_cons, sod_config, sod_state
from tests.oracle_bind import Oracle

pytestmark = pytest.mark.gpu


def _nan_case():
    mesh = nf.Mesh.box(8, jitter=0.1, seed=3)
    U0 = sod_state(mesh)
    U0[137, 0] = np.nan  # one cell; its neighbours' gradients, limiters and face values follow
    return mesh, U0


def _oracle_error(mesh, cfg, U0):
    o = Oracle.from_mesh(mesh, cfg)
    o.set_state(U0)
    with pytest.raises(RuntimeError) as e:
        o.run(1)
    return str(e.value)


def _state_time(g):
    st = g.run(0)  # no steps: reports the run counters
    return g.get_solver_data().get_solution(), st["time"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nan_flux_balance_single_step(precision):
    print(f"[case] NaN cell, nf_step, {precision}", flush=True)
    mesh, U0 = _nan_case()
    cfg = sod_config(check_physical_validity=True, precision=precision)
    msg_o = _oracle_error(mesh, sod_config(check_physical_validity=True), U0)
    assert msg_o.endswith("cells with unphysical values detected in the flux_balance field of the Euler solver")
    assert int(msg_o.split()[0]) > 1
    g = nf.EulerSolver(cfg, mesh)
    g.set_state(U0)
    g.calc_timestep(cfg)
    with pytest.raises(RuntimeError) as e:
        g.step(cfg)
    print("oracle:", msg_o, "| gpu:", e.value)
    assert str(e.value) == msg_o
    U, t = _state_time(g)
    np.testing.assert_array_equal(U, U0)  # NaN compares equal to NaN here
    assert t == 0.0
    assert g.residuals().size == 0


def test_nan_flux_balance_graph_run():
    """The error surfaces after the graph replay; the count is the first
    failing evaluation's, not a sum over the stages and steps replayed."""
    print("[case] NaN cell, graph-replayed nf_run", flush=True)
    mesh, U0 = _nan_case()
    cfg = sod_config(check_physical_validity=True)
    msg_o = _oracle_error(mesh, cfg, U0)
    g = nf.EulerSolver(cfg, mesh)
    g.set_state(U0)
    with pytest.raises(RuntimeError) as e:
        g.run(10, use_graph=True, check_every=5)
    assert str(e.value) == msg_o
    U, t = _state_time(g)
    np.testing.assert_array_equal(U, U0)
    assert t == 0.0


def test_nan_after_good_steps():
    """Three valid steps, then a NaN injected into the state: the fourth step
    fails with the oracle's count and leaves the injected state."""
    print("[case] NaN after 3 good steps", flush=True)
    mesh, U0 = _nan_case()
    good = sod_state(mesh)
    cfg = sod_config(check_physical_validity=True)
    g = nf.EulerSolver(cfg, mesh)
    g.set_state(good)
    st = g.run(3, use_graph=False)
    assert st["steps_done"] == 3
    o = Oracle.from_mesh(mesh, cfg)
    o.set_state(good)
    o.run(3)
    np.testing.assert_array_equal(g.get_solver_data().get_solution(), o.get("U"))
    U3 = o.get("U")
    U3[137, 0] = np.nan
    msg_o = _oracle_error(mesh, cfg, U3)
    g.set_state(U3)
    with pytest.raises(RuntimeError) as e:
        g.run(5, use_graph=True, check_every=5)
    assert str(e.value) == msg_o
    np.testing.assert_array_equal(g.get_solver_data().get_solution(), U3)


def test_checked_run_equals_unchecked():
    """With valid states the check changes nothing: the checked stages (R
    stored, then the update pass) give the unchecked run's bits."""
    print("[case] checked vs unchecked, 20 steps", flush=True)
    mesh = nf.Mesh.box(8, jitter=0.12, seed=4)
    U0 = sod_state(mesh)
    out = []
    for chk in (False, True):
        cfg = sod_config(check_physical_validity=chk)
        g = nf.EulerSolver(cfg, mesh)
        g.set_state(U0)
        st = g.run(20, use_graph=True, check_every=10)
        out.append((g.get_solver_data().get_solution(), g.residuals(), st["time"]))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


def test_invalid_dt_message():
    """The running minimum starts at DBL_MAX and skips NaN and inf, so an
    invalid dt needs a negative or zero CFL step: a negative CFL number makes
    every cell's step negative, and calc_timestep / nf_run raise the
    reference's message with the value (std::to_string, i.e. %f)."""
    print("[case] invalid dt", flush=True)
    mesh = nf.Mesh.box(6, jitter=0.1, seed=8)
    U0 = sod_state(mesh)
    cfg = sod_config(CFL=-0.5)
    o = Oracle.from_mesh(mesh, cfg)
    o.set_state(U0)
    with pytest.raises(RuntimeError) as eo:
        o.calc_timestep()
    assert str(eo.value).startswith("Invalid dt calculated (dt = -0.0")
    g = nf.EulerSolver(cfg, mesh)
    g.set_state(U0)
    with pytest.raises(RuntimeError) as e:
        g.calc_timestep(cfg)
    assert str(e.value) == str(eo.value)
    h = nf.EulerSolver(cfg, mesh)
    h.set_state(U0)
    with pytest.raises(RuntimeError) as e2:
        h.run(10, use_graph=True, check_every=5)
    assert str(e2.value) == str(eo.value)
    np.testing.assert_array_equal(h.get_solver_data().get_solution(), U0)
    # p = 0 at rest is not an invalid dt: every cell's step is inf and the
    # minimum stays DBL_MAX (the reference only warns that dt is high)
    V = np.zeros((mesh.n_cells, 5))
    V[:, 0] = 1.0
    o2 = Oracle.from_mesh(mesh, sod_config())
    o2.set_state(prim_to_cons(V))
    g2 = nf.EulerSolver(sod_config(), mesh)
    g2.set_state(prim_to_cons(V))
    c2 = sod_config()
    g2.calc_timestep(c2)
    assert c2.delta_time == o2.calc_timestep() == np.finfo(np.float64).max


def test_nan_cell_ignored_by_dt_min():
    """A NaN cell drops out of the dt minimum, as in the reference's running
    std::min, wherever the cell lands in a block (each position tried)."""
    print("[case] NaN cell in the dt minimum", flush=True)
    mesh = nf.Mesh.box(6, jitter=0.1, seed=2)
    base = sod_state(mesh)
    cfg = sod_config()
    o = Oracle.from_mesh(mesh, cfg)
    g = nf.EulerSolver(cfg, mesh)
    for c in (0, 1, 63, 64, 255, 700, mesh.n_cells - 1):
        U = base.copy()
        U[c, 0] = np.nan
        o.set_state(U)
        g.set_state(U)
        g.calc_timestep(cfg)
        assert cfg.delta_time == o.calc_timestep(), c
