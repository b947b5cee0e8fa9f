"""Warm start plumbing on the GPU: network features, prediction -> state,
the freestream fallback for a non-physical prediction and for a stalled
residual, the residual reference of the drop, and residual evaluation."""
import numpy as np
import pytest

import netflux_amd as nf
from netflux_amd.mlp import Mlp
from tests.cases import GAMMA, prim_to_cons

pytestmark = pytest.mark.gpu


def _cfg(mach=0.5):
    return nf.Config(density_fs=1.0, velocity_fs=(mach, 0.0, 0.0), pressure_fs=1 / GAMMA, CFL=0.5,
                     rk3_zero_old_first_step=False)


def _const_net(V):
    """A network whose output is V for every input (zero weights)."""
    net = Mlp(0)
    net.params = np.zeros(net.n_params, np.float32)
    net.set_norm(np.zeros(6), np.ones(6), np.asarray(V, np.float32), np.ones(5))
    return net


def test_features_match_geometry():
    print("[case] features of a 10x5x4 bump channel", flush=True)
    mesh = nf.Mesh.bump_channel(10, 5, 4)
    b = nf.BatchSolver(_cfg(), [mesh])
    x = b.features(0, (0.55, 0.07))
    g = mesh.fv_grid()
    N = mesh.n_cells
    assert np.all(x[:, 0] == np.float32(0.55)) and np.all(x[:, 1] == np.float32(0.07))
    np.testing.assert_allclose(x[:, 2:5], g["centroid"][:N], rtol=0, atol=1e-6)
    # walls: every patch but the two x-planes; distance to the nearest wall-face centroid
    nodes, _, pn, pf, pbc = mesh.arrays()
    first = np.r_[0, np.cumsum(pn)[:-1]]
    wall = np.concatenate([nodes[pf[f:f + n].astype(np.int64)].mean(1)
                           for f, n, bc in zip(first, pn, pbc) if bc != nf.solver.BC_NAMES["FarField"]])
    d = np.sqrt(((g["centroid"][:N, None, :] - wall[None]) ** 2).sum(-1).min(1))
    np.testing.assert_allclose(x[:, 5], d, rtol=1e-5, atol=1e-6)


def test_warm_start_sets_predicted_state():
    print("[case] warm start from a constant network", flush=True)
    mesh = nf.Mesh.bump_channel(10, 5, 4)
    V = np.array([1.1, 0.45, 0.01, 0.0, 0.8], np.float32)
    b = nf.BatchSolver(_cfg(), [mesh])
    assert b.warm_start(0, _const_net(V), (0.5, 0.05)) is False
    U = b.get_solution(0)
    expect = prim_to_cons(np.tile(V.astype(np.float64), (mesh.n_cells, 1)))
    np.testing.assert_array_equal(U, expect)


def test_nonphysical_prediction_falls_back_to_freestream():
    print("[case] warm start fallback (rho < 0)", flush=True)
    mesh = nf.Mesh.bump_channel(10, 5, 4)
    cfg = _cfg()
    b = nf.BatchSolver(cfg, [mesh])
    assert b.warm_start(0, _const_net([-1.0, 0.5, 0, 0, 0.7]), (0.5, 0.05)) is True
    np.testing.assert_array_equal(b.get_solution(0), prim_to_cons(np.tile(cfg.primvars_inf(), (mesh.n_cells, 1))))


def test_stalled_warm_start_restarts_from_freestream():
    print("[case] stall fallback", flush=True)
    mesh = nf.Mesh.bump_channel(10, 5, 4)
    cfg = _cfg()
    b = nf.BatchSolver(cfg, [mesh])
    b.warm_start(0, _const_net([1.0, 0.5, 0.0, 0.0, 1 / GAMMA + 0.01]), (0.5, 0.05))
    # a ratio nothing meets: the first check (after 5 steps) restarts it
    st = b.run(20, check_every=5, stall_window=3, stall_ratio=1e-30)[0]
    assert st["fell_back"] == 1 and st["steps_done"] == 20
    plain = nf.BatchSolver(cfg, [mesh])
    U0 = prim_to_cons(np.tile(cfg.primvars_inf(), (mesh.n_cells, 1)))
    plain.set_state(0, U0, U0)
    plain.run(15, check_every=5)
    np.testing.assert_array_equal(b.get_solution(0), plain.get_solution(0))


def test_residual_reference_and_evaluation():
    print("[case] residual reference / evaluation", flush=True)
    mesh = nf.Mesh.bump_channel(10, 5, 4)
    cfg = _cfg()
    b = nf.BatchSolver([cfg, cfg], [mesh, mesh])
    U0 = prim_to_cons(np.tile(cfg.primvars_inf(), (mesh.n_cells, 1)))
    for i in range(2):
        b.set_state(i, U0, U0)
    r = b.eval_residuals()
    assert r[0] > 0 and r[0] == r[1]
    b.set_residual_ref(1, 1e6 * r[1])
    st = b.run(30, residual_drop=1e-3, check_every=1)
    assert st[1]["converged"] == 1 and st[1]["steps_done"] == 1
    assert st[0]["res_first"] == r[0] and st[0]["steps_done"] == 30
