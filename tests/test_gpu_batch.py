"""Batched contexts: several uneven meshes advanced by one launch per sweep.

Each instance of a batch must evolve exactly as the same mesh in a context of
its own (same per-instance dt, same fixed-order reductions), and therefore as
the oracle; the launch configuration chosen by the autotuner must not change
the state; an instance that reaches its residual drop stops while the others
continue."""
import numpy as np
import pytest

import netflux_amd as nf
from tests.cases import pulse_config, pulse_state, sod_config, sod_state
from tests.oracle_bind import Oracle

pytestmark = pytest.mark.gpu


def _cases():
    """Three meshes of different sizes, boundary types and freestreams."""
    m0 = nf.Mesh.box(6, jitter=0.1, seed=5)
    m1 = nf.Mesh.box((10, 7, 5), lengths=(1.4, 1.0, 0.7), jitter=0.12, seed=9,
                     patches=("FarField", "FarField", "SlipWall", "SlipWall", "NoSlipWall", "SlipWall"))
    m2 = nf.Mesh.box(9, jitter=0.15, seed=3, patches=("FarField",) * 6)
    c0, c1, c2 = sod_config(), pulse_config(mach=0.3), pulse_config(mach=0.5)
    U = [sod_state(m0), pulse_state(m1, mach=0.3, centre=(0.7, 0.5, 0.35)), pulse_state(m2)]
    return [m0, m1, m2], [c0, c1, c2], U


def _singles(meshes, cfgs, U, n_steps, uold_equals_u=False, **run):
    out = []
    for m, c, u in zip(meshes, cfgs, U):
        s = nf.EulerSolver(c, m)
        s.set_state(u, u if uold_equals_u else None)
        st = s.run(n_steps, **run)
        out.append((s.get_solver_data().get_solution(), s.residuals(), st))
    return out


def test_batch_equals_single_contexts():
    meshes, cfgs, U = _cases()
    print("[case] batch of 3, 20 steps", flush=True)
    b = nf.BatchSolver(cfgs, meshes)
    assert b.N == [m.n_cells for m in meshes]
    for i, u in enumerate(U):
        b.set_state(i, u)
    stats = b.run(20, check_every=5)
    ref = _singles(meshes, cfgs, U, 20, check_every=5)
    for i, (Us, rs, ss) in enumerate(ref):
        np.testing.assert_array_equal(b.get_solution(i), Us)
        np.testing.assert_array_equal(b.residuals(i), rs)
        assert stats[i]["steps_done"] == 20
        assert stats[i]["time"] == ss["time"] and stats[i]["last_dt"] == ss["last_dt"]


def test_batch_instance_matches_oracle():
    meshes, cfgs, U = _cases()
    print("[case] batch of 3 vs oracle, 10 steps", flush=True)
    b = nf.BatchSolver(cfgs, meshes)
    for i, u in enumerate(U):
        b.set_state(i, u)
    b.run(10, check_every=5)
    for i in (0, 2):
        o = Oracle.from_mesh(meshes[i], cfgs[i])
        o.set_state(U[i])
        o.run(10)
        Ug, Uo = b.get_solution(i), o.get("U")
        if i == 0:
            np.testing.assert_array_equal(Ug, Uo)
        else:  # farfield pow() may differ from glibc in the last bit
            assert np.linalg.norm(Ug - Uo) <= 1e-12 * np.linalg.norm(Uo)


def test_autotune_keeps_state_and_results():
    meshes, cfgs, U = _cases()
    print("[case] autotune over 4 launch configurations", flush=True)
    b = nf.BatchSolver(cfgs, meshes)
    for i, u in enumerate(U):
        b.set_state(i, u)
    before = [b.get_solution(i) for i in range(3)]
    r = b.autotune(steps=4)
    assert r["block"] in (128, 256) and r["xcd_remap"] in (0, 1)
    assert len(r["tried_ms"]) == 4 and min(r["tried_ms"]) == r["ms_per_step"] > 0
    for i in range(3):
        np.testing.assert_array_equal(b.get_solution(i), before[i])
        assert len(b.residuals(i)) == 0
    results = []
    for block, remap in ((256, 1), (256, 0), (128, 1), (128, 0)):
        for i, u in enumerate(U):
            b.set_state(i, u)
        b.set_launch(block, remap)
        b.run(8, check_every=4)
        results.append(([b.get_solution(i) for i in range(3)], [b.residuals(i) for i in range(3)]))
    for sol, res in results[1:]:
        for i in range(3):
            np.testing.assert_array_equal(sol[i], results[0][0][i])
            # the residual sums group cells by block: equal to rounding
            np.testing.assert_allclose(res[i], results[0][1][i], rtol=1e-13, atol=0)


def test_instances_stop_independently_at_residual_drop():
    meshes, _, U = _cases()
    meshes, U = meshes[1:], U[1:]
    cfgs = [pulse_config(mach=0.3), pulse_config(mach=0.5)]
    n = 60
    print(f"[case] batch of 2 farfield pulses, {n} steps with a residual drop", flush=True)
    hist = []
    for i in range(2):
        s = nf.EulerSolver(cfgs[i], meshes[i])
        s.set_state(U[i], U[i])
        s.run(n, check_every=1)
        hist.append(s.residuals() / s.residuals()[0])
    # the instance whose residual falls furthest by step 40 sets the drop at
    # its minimum there; the batch must stop it at its first hit
    a = int(np.argmin([h[5:41].min() for h in hist]))
    k = 5 + int(np.argmin(hist[a][5:41]))
    drop = hist[a][k]
    assert drop < 1.0
    b = nf.BatchSolver(cfgs, meshes)
    for i, u in enumerate(U):
        b.set_state(i, u, u)
    stats = b.run(n, residual_drop=drop, check_every=4)
    for i in range(2):
        r = b.residuals(i)
        hit = np.nonzero(r / r[0] <= drop)[0]
        if len(hit):
            assert stats[i]["converged"] == 1
            assert stats[i]["steps_done"] == hit[0] + 1 == len(r)
        else:
            assert stats[i]["converged"] == 0 and stats[i]["steps_done"] == n
        s = nf.EulerSolver(cfgs[i], meshes[i])
        s.set_state(U[i], U[i])
        s.run(stats[i]["steps_done"], check_every=1)
        np.testing.assert_array_equal(b.get_solution(i), s.get_solver_data().get_solution())
    assert stats[a]["converged"] == 1 and stats[a]["steps_done"] <= k + 1
    # a converged instance stays stopped when the run continues with a drop
    again = b.run(8, residual_drop=drop, check_every=4)
    assert again[a]["steps_done"] == 0 and again[a]["converged"] == 1


def test_bump_channel_batch_mixed_then_fp64():
    """The batched workload's shape, scaled down: 16 bump channels of uneven
    size (log-uniform 3k-40k cells), Mach U[0.3, 0.8] and bump height
    U[0.02, 0.1] from a seed, local time steps, mixed precision then FP64 in
    the same context, launch picked by the autotuner.  Every instance ends
    bit-identical to its mesh advanced alone through the same phases."""
    from netflux_amd import steady as SD
    cases = SD.sample_cases(16, seed=11, cells_lo=3e3, cells_hi=4e4)
    meshes = [SD.mesh_of(c) for c in cases]
    cfgs = [SD.config_of(c, "mixed") for c in cases]
    U0 = [SD.freestream_state(cf, m.n_cells) for cf, m in zip(cfgs, meshes)]
    print(f"[case] 16 bump channels, {sum(m.n_cells for m in meshes)} cells, mixed 30 + fp64 20 steps", flush=True)
    assert len({m.n_cells for m in meshes}) > 4
    b = nf.BatchSolver(cfgs, meshes)
    for i, u in enumerate(U0):
        b.set_state(i, u, u)
    tune = b.autotune(steps=4)
    print("launch", tune, flush=True)
    st1 = b.run(30, check_every=10)
    b.set_precision("fp64")
    st2 = b.run(20, check_every=10)
    assert all(s["steps_done"] == 30 for s in st1) and all(s["steps_done"] == 20 for s in st2)
    for i, (m, cf, u) in enumerate(zip(meshes, cfgs, U0)):
        s = nf.EulerSolver(cf, m)
        s.set_state(u, u)
        s.run(30, check_every=10)
        s.set_precision("fp64")
        s.run(20, check_every=10)
        Us = s.get_solver_data().get_solution()
        Ub = b.get_solution(i)
        assert np.abs(Us - u).max() > 1e-6  # the flow has developed
        np.testing.assert_array_equal(Ub, Us)
        np.testing.assert_allclose(b.residuals(i), s.residuals(), rtol=1e-13, atol=0)
