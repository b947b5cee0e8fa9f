"""Mixed precision (FP32 gradient/limiter/flux, FP64 state/RK/norms) against
the FP64 oracle.  The error bound is stated in each test: FP32 face
arithmetic (unit round-off 6e-8) accumulated over the steps run."""
import numpy as np
import pytest

import netflux_amd as nf
from tests.cases import pulse_config, pulse_state, rel_l2, sod_config, sod_state
from tests.oracle_bind import Oracle

pytestmark = pytest.mark.gpu


def _run(mesh, cfg, U0, steps, uold_equals_u=True):
    print(f"[case] mixed cells={mesh.n_cells} steps={steps}", flush=True)
    cfg.rk3_zero_old_first_step = not uold_equals_u
    g = nf.EulerSolver(cfg, mesh)
    g.set_state(U0, U0 if uold_equals_u else None)
    st = g.run(steps, use_graph=True, check_every=10)
    return g, st


def test_mixed_sod_error_vs_fp64_oracle():
    mesh = nf.Mesh.box(16)
    U0 = sod_state(mesh)
    g, _ = _run(mesh, sod_config(precision="mixed"), U0, 50)
    ora = Oracle.from_mesh(mesh, sod_config())
    ora.set_state(U0, uold_equals_u=True)
    ora.run(50)
    Ug, Uo = g.get_solver_data().get_solution(), ora.get("U")
    errs = [rel_l2(Ug[:, k], Uo[:, k]) if np.linalg.norm(Uo[:, k]) > 0 else float(np.abs(Ug[:, k]).max())
            for k in range(5)]
    print("per-variable relative L2 vs FP64 oracle:", errs)
    assert rel_l2(Ug, Uo) < 2e-5  # FP32 flux/gradient/limiter, 150 stages
    assert errs[0] < 2e-5 and errs[4] < 2e-5


def test_mixed_conserves_mass_and_energy():
    mesh = nf.Mesh.box(10, jitter=0.15, seed=12)
    U0 = sod_state(mesh)
    vol = mesh.fv_grid()["volume"][: mesh.n_cells]
    drift = {}
    for prec in ("fp64", "mixed"):
        g, _ = _run(mesh, sod_config(precision=prec), U0, 40)
        U = g.get_solver_data().get_solution()
        drift[prec] = [abs(vol @ U[:, k] - vol @ U0[:, k]) / (vol @ U0[:, k]) for k in (0, 4)]
    print("relative drift of mass and energy after 40 steps:", drift)
    # Interior faces cancel exactly in both modes (each face flux is evaluated
    # bit-identically by its two cells and summed in FP64).  The slip walls
    # leak at the rounding level of the reflected ghost state: FP64 there,
    # FP32 in mixed mode (the unit normal S/|S| is an FP32 quantity).
    assert max(drift["fp64"]) < 1e-13
    assert max(drift["mixed"]) < 1e-9


def test_mixed_interior_fluxes_cancel_exactly():
    """At rest (u = 0) next to slip walls the wall mass and energy fluxes are
    exactly zero, so the cell flux balances sum to exactly zero iff each
    interior face flux enters its two cells with opposite signs bit for bit.
    The per-cell balances are FP64 sums of FP32 fluxes of similar magnitude,
    so math.fsum of them is exact."""
    import math
    mesh = nf.Mesh.box(10, jitter=0.15, seed=12)
    g = nf.EulerSolver(sod_config(precision="mixed"), mesh)
    g.set_state(sod_state(mesh))
    g.eval_flux_balance()
    R = g.get_solver_data().get_flux_balance()
    assert np.abs(R[:, 0]).max() > 0
    assert math.fsum(R[:, 0]) == 0.0 and math.fsum(R[:, 4]) == 0.0


def test_mixed_farfield_pulse():
    mesh = nf.Mesh.box(12, jitter=0.15, seed=5, patches=("FarField",) * 6)
    U0 = pulse_state(mesh)
    g, _ = _run(mesh, pulse_config(precision="mixed"), U0, 30)
    ora = Oracle.from_mesh(mesh, pulse_config())
    ora.set_state(U0, uold_equals_u=True)
    ora.run(30)
    assert rel_l2(g.get_solver_data().get_solution(), ora.get("U")) < 1e-5


def test_precision_switch_continues_in_fp64():
    """Mixed for 10 steps, then FP64 for 10 in the same context, equals a FP64
    context started from the mixed state (primvars and dt recomputed from U)."""
    mesh = nf.Mesh.box(8, jitter=0.1, seed=21, patches=("FarField",) * 6)
    U0 = pulse_state(mesh)
    a, _ = _run(mesh, pulse_config(precision="mixed"), U0, 10)
    U10 = a.get_solver_data().get_solution()
    a.set_precision("fp64")
    a.run(10, use_graph=True, check_every=5)
    b, _ = _run(mesh, pulse_config(precision="fp64"), U10, 10)
    np.testing.assert_array_equal(a.get_solver_data().get_solution(), b.get_solver_data().get_solution())
