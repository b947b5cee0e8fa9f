"""Mixed-element meshes (hexahedra, wedges, pyramids) on the GPU against the
CPU oracle: the polyhedral FP64 path gathers each cell's faces in the
reference's face order and evaluates the oracle's operations in the same
order, so walls give bit-identical states; farfield ghosts call pow(), whose
last bit can differ between the device and glibc, so those cases are held to
the 1e-12 relative L2 bar of the tetrahedral path."""
import os

import numpy as np
import pytest

import netflux_amd as nf
from netflux_amd import frontend as fe
from tests.cases import pulse_config, pulse_state, rel_l2, sod_config, sod_state
from tests.mixed_meshes import mixed_box
from tests.oracle_bind import Oracle

pytestmark = pytest.mark.gpu

TOL_REL_L2 = 1e-12
GOLD = os.path.join(os.path.dirname(__file__), "golden", "sim_inputs")


def _pair(mesh, cfg, U0):
    print(f"[case] mixed cells={mesh.n_cells} scheme={cfg.time_scheme} order={cfg.spatial_order} "
          f"limiter={cfg.limiter} flux={cfg.inviscid_flux_scheme}", flush=True)
    gpu = nf.EulerSolver(cfg, mesh)
    gpu.set_state(U0)
    ora = Oracle.from_mesh(mesh, cfg)
    ora.set_state(U0, False)
    return gpu, ora


def _steps(gpu, ora, cfg, n, exact=True):
    for _ in range(n):
        gpu.calc_timestep(cfg)
        dt_o = ora.calc_timestep()
        if exact:
            assert cfg.delta_time == dt_o
        gpu.step(cfg)
        ora.step(dt_o)


@pytest.mark.parametrize("case,scheme,order,limiter,flux", [
    ("sod", "TVD_RK3", "Second", "Barth", "Rusanov"),
    ("sod", "ExplicitEuler", "First", "NONE", "Rusanov"),
    ("pulse", "TVD_RK3", "Second", "NONE", "Rusanov"),   # unlimited: a smooth field (Sod goes nonphysical)
    ("sod", "TVD_RK3", "Second", "Barth", "HLLC"),
])
def test_mixed_walls_match_oracle(case, scheme, order, limiter, flux):
    """Slip walls on every side: bit-identical flux balance, gradient,
    limiter, primvars and state after 25 steps."""
    mesh = mixed_box((6, 4, 4), kinds=("hex", "wedge", "pyr"), jitter=0.12, seed=5)
    make = sod_config if case == "sod" else pulse_config
    cfg = make(time_scheme=scheme, spatial_order=order, limiter=limiter, inviscid_flux_scheme=flux,
               check_physical_validity=True)
    gpu, ora = _pair(mesh, cfg, sod_state(mesh) if case == "sod" else pulse_state(mesh))
    gpu.eval_flux_balance()
    ora.eval_flux_balance()
    d = gpu.get_solver_data()
    np.testing.assert_array_equal(d.get_flux_balance(), ora.get("R"))
    if order == "Second":
        np.testing.assert_array_equal(d.get_primvars_gradient(), ora.get("grad"))
        np.testing.assert_array_equal(d.get_primvars_limiter(), ora.get("psi"))
    _steps(gpu, ora, cfg, 25)
    np.testing.assert_array_equal(d.get_solution(), ora.get("U"))
    ora.eval_flux_balance()   # primvars = prim(U) in the oracle
    np.testing.assert_array_equal(d.get_primvars(), ora.get("V")[:ora.N])


def test_mixed_farfield_pulse_100_rk3_steps():
    """The headline parity bar on a mixed mesh: 100 TVD-RK3 steps from
    freestream plus a pulse, farfield on all sides, first step with the
    reference's zero U_old."""
    mesh = mixed_box((6, 6, 5), kinds=("pyr", "hex", "wedge"), jitter=0.15, seed=12, patches=("FarField",) * 6,
                     lengths=(1.0, 1.0, 1.0))
    cfg = pulse_config()
    gpu, ora = _pair(mesh, cfg, pulse_state(mesh))
    _steps(gpu, ora, cfg, 100, exact=False)
    U_g, U_o = gpu.get_solver_data().get_solution(), ora.get("U")
    assert rel_l2(U_g, U_o) <= TOL_REL_L2


def test_mixed_walls_and_freestream_start():
    """Slip and no-slip walls plus farfield on one mesh, started from
    set_freestream as the Driver does."""
    mesh = mixed_box((5, 4, 3), kinds=("wedge", "pyr", "hex"), jitter=0.1, seed=19,
                     patches=("FarField", "FarField", "SlipWall", "NoSlipWall", "SlipWall", "SlipWall"))
    cfg = pulse_config(mach=0.6)
    print(f"[case] mixed walls cells={mesh.n_cells}", flush=True)
    gpu = nf.EulerSolver(cfg, mesh)
    ora = Oracle.from_mesh(mesh, cfg)
    ora.set_freestream()
    _steps(gpu, ora, cfg, 20, exact=False)
    assert rel_l2(gpu.get_solver_data().get_solution(), ora.get("U")) <= TOL_REL_L2


def test_mixed_unsupported_calls_and_errors():
    mesh = mixed_box((3, 3, 2), kinds=("hex", "pyr"), jitter=0.05, seed=1)
    print("[case] mixed unsupported calls", flush=True)
    with pytest.raises(RuntimeError):
        nf.EulerSolver(sod_config(precision="mixed"), mesh)
    s = nf.EulerSolver(sod_config(), mesh)
    with pytest.raises(RuntimeError, match="mixed-element"):
        s.autotune(2)
    kt = s.time_kernels(2)   # HIP-event durations of the polyhedral kernels
    assert set(kt) == {"poly_prim", "poly_ghost", "poly_sweepA", "poly_sweepB"}
    assert all(calls == 6 and avg > 0 for avg, calls in kt.values())
    with pytest.raises(RuntimeError, match="mixed-element"):
        s.set_precision("mixed")
    with pytest.raises(RuntimeError):
        nf.BatchSolver([sod_config()] * 2, [mesh, mesh])
    # a non-physical state is reported with the reference's message
    U = sod_state(mesh)
    U[3, 0] = np.nan
    cfg = sod_config(check_physical_validity=True)
    s2 = nf.EulerSolver(cfg, mesh)
    s2.set_state(U)
    with pytest.raises(RuntimeError, match="cells with unphysical values"):
        s2.eval_flux_balance()


def test_mixed_sim_dir_matches_oracle(tmp_path):
    """A reference sim directory (input.yml from the reference's own runs)
    over a mixed SU2 mesh, through the front end's Driver loop."""
    name = "small_brick_domain"
    cfg, opts = fe.parse_input(os.path.join(GOLD, name + ".yml"))
    names = list(opts["patches"])    # inlet, outlet, sides
    mesh = mixed_box((5, 3, 3), kinds=("hex", "wedge", "pyr"), jitter=0.1, seed=23)
    ct, co, cn, pn, ft, fo, fn, pbc = mesh.elements()
    faces, e = [], 0
    for k in pn:
        faces.append([fn[fo[q]:fo[q + 1]] for q in range(e, e + k)])
        e += k
    X = np.ctypeslib.as_array(mesh.desc.nodes, shape=(mesh.desc.n_nodes, 3)).copy()
    three = nf.Mesh.from_elements(X, [cn[co[q]:co[q + 1]] for q in range(len(ct))],
                                  [faces[0], faces[1], faces[2] + faces[3] + faces[4] + faces[5]], ["SlipWall"] * 3)
    d = tmp_path / name
    d.mkdir()
    text = open(os.path.join(GOLD, name + ".yml")).read()
    n_steps = 10
    d.joinpath("input.yml").write_text("\n".join(f"n_timesteps: {n_steps}" if l.startswith("n_timesteps:") else l
                                                 for l in text.split("\n")))
    three.write_su2(d / "mesh.su2", names)
    print(f"[case] mixed sim dir {name}, {n_steps} steps", flush=True)
    solver, run_cfg = fe.run_sim_dir(str(d), log=lambda *_: None)
    cfg.n_timesteps = n_steps
    ora = Oracle.from_mesh(nf.Mesh.read_su2(d / "mesh.su2", opts["patches"]), cfg)
    ora.set_freestream()
    t = 0.0
    for _ in range(n_steps):
        dt = ora.calc_timestep()
        ora.step(dt)
        t += dt
    np.testing.assert_array_equal(solver.get_solver_data().get_solution(), ora.get("U"))
    assert run_cfg.time == t
    assert len(os.listdir(d / "output")) == n_steps + 1


def test_mixed_run_loop_and_residuals():
    """nf_run on a mixed mesh is the Driver loop: the same state as stepping
    by hand, the residual norm ||R_rho||_2 of the flux balance of U^n at each
    step (the oracle's R of that state), and the residual-drop stop."""
    mesh = mixed_box((5, 4, 3), kinds=("hex", "wedge", "pyr"), jitter=0.1, seed=31)
    cfg = sod_config()
    U0 = sod_state(mesh)
    a, ora = _pair(mesh, cfg, U0)
    st = a.run(30)
    assert st["steps_done"] == 30 and not st["converged"]
    b = nf.EulerSolver(cfg, mesh)
    b.set_state(U0)
    t = 0.0
    for _ in range(30):
        b.calc_timestep(cfg)
        b.step(cfg)
        t += cfg.delta_time
    np.testing.assert_array_equal(a.get_solver_data().get_solution(), b.get_solver_data().get_solution())
    assert st["time"] == t
    # device-resident run and Driver loop: the same residual bits
    np.testing.assert_array_equal(a.residuals(), b.residuals())
    # with the flux-balance check the run loop synchronises per stage: same state
    c = nf.EulerSolver(sod_config(check_physical_validity=True), mesh)
    c.set_state(U0)
    assert c.run(30, check_every=7)["steps_done"] == 30
    np.testing.assert_array_equal(c.get_solver_data().get_solution(), b.get_solver_data().get_solution())
    ora.run(29)
    ora.eval_flux_balance()   # R(U^29), the 30th step's residual
    r = a.residuals()
    assert len(r) == 30 and st["res_last"] == r[-1] and st["res_first"] == r[0]
    assert r[-1] == pytest.approx(np.linalg.norm(ora.get("R")[:, 0]), rel=1e-13)
    # stop at a drop of 0.5 of the first step's residual: a pulse leaving
    # through farfield boundaries
    ff = mixed_box((5, 4, 3), kinds=("hex", "wedge", "pyr"), jitter=0.1, seed=31, patches=("FarField",) * 6)
    a = nf.EulerSolver(pulse_config(), ff)
    a.set_state(pulse_state(ff))
    st = a.run(5000, residual_drop=0.5)
    r = a.residuals()
    assert st["converged"] and st["steps_done"] == len(r) < 5000
    assert r[-1] / r[0] <= 0.5 < r[-2] / r[0]
    assert a.run(10, residual_drop=0.5)["steps_done"] == 0   # stays stopped


@pytest.mark.parametrize("flux,order", [("Rusanov", "morton"), ("HLLC", "rcm"), ("Rusanov", "input")])
def test_hex_box_24cubed_matches_oracle(flux, order):
    """54 blocks of 256 cells: a jittered 24^3 hexahedral box with slip
    walls, 20 TVD-RK3 steps of Sod, bit-identical to the oracle in every
    device cell numbering (each cell keeps the reference's face order)."""
    mesh = nf.Mesh.hex_box(24, jitter=0.2, seed=41)
    cfg = sod_config(inviscid_flux_scheme=flux, cell_order=order)
    gpu, ora = _pair(mesh, cfg, sod_state(mesh))
    _steps(gpu, ora, cfg, 20)
    np.testing.assert_array_equal(gpu.get_solver_data().get_solution(), ora.get("U"))


def test_single_pyramid_and_thin_mesh():
    """Ragged extremes: one cell in a 256-lane block (a pyramid with a
    farfield base and slip sides), and a one-block-wide column of wedges.
    The single cell has no interior face, so its Delta S is zero and the
    reference's time step is DBL_MAX (only interior faces count); its flux
    balance is still well defined."""
    nodes = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.4, 0.6, 0.9]], float)
    one = nf.Mesh.from_elements(nodes, [[0, 1, 2, 3, 4]], [[[0, 3, 2, 1]], [[0, 1, 4], [1, 2, 4], [2, 3, 4], [3, 0, 4]]],
                                ["FarField", "SlipWall"])
    cfg = pulse_config(mach=0.4)
    print("[case] single pyramid", flush=True)
    gpu, ora = _pair(one, cfg, pulse_state(one, centre=(0.5, 0.5, 0.3)))
    gpu.calc_timestep(cfg)
    assert cfg.delta_time == ora.calc_timestep() == np.finfo(float).max
    gpu.eval_flux_balance()
    ora.eval_flux_balance()
    R_g, R_o = gpu.get_solver_data().get_flux_balance(), ora.get("R")
    assert np.all(np.isfinite(R_g)) and np.max(np.abs(R_g - R_o)) <= 1e-13 * np.max(np.abs(R_o))
    thin = mixed_box((1, 1, 7), kinds=("wedge",), jitter=0.0, seed=2, patches=("FarField",) * 2 + ("SlipWall",) * 4)
    print(f"[case] thin wedge column cells={thin.n_cells}", flush=True)
    gpu, ora = _pair(thin, cfg, pulse_state(thin))
    _steps(gpu, ora, cfg, 5, exact=False)
    assert rel_l2(gpu.get_solver_data().get_solution(), ora.get("U")) <= TOL_REL_L2


@pytest.mark.parametrize("name,n_steps", [("cyprien_hex_tutorial", 20), ("mixed_tiny", 12)])
def test_reference_hex_and_mixed_sim_dirs(tmp_path, name, n_steps):
    """The reference's hexahedral tutorial and mixed-element input.yml files
    (their meshes are not shipped) over generated SU2 meshes of that kind,
    run unchanged through the front end and checked against the oracle."""
    cfg, opts = fe.parse_input(os.path.join(GOLD, name + ".yml"))
    names = list(opts["patches"])
    if name == "cyprien_hex_tutorial":
        m = nf.Mesh.hex_box((8, 4, 4), lengths=(2.0, 1.0, 1.0), jitter=0.15, seed=7)
        groups = [[0, 1, 2, 3, 4, 5]]
    else:
        m = mixed_box((6, 3, 3), kinds=("hex", "wedge", "pyr"), jitter=0.1, seed=9, lengths=(2.0, 1.0, 1.0))
        groups = [[0], [1], [2, 3, 4, 5]]
    ct, co, cn, pn, ft, fo, fn, pbc = m.elements()
    faces, e = [], 0
    for k in pn:
        faces.append([fn[fo[q]:fo[q + 1]] for q in range(e, e + k)])
        e += k
    X = np.ctypeslib.as_array(m.desc.nodes, shape=(m.desc.n_nodes, 3)).copy()
    mesh = nf.Mesh.from_elements(X, [cn[co[q]:co[q + 1]] for q in range(len(ct))],
                                 [sum((faces[g] for g in grp), []) for grp in groups], ["SlipWall"] * len(groups))
    d = tmp_path / name
    d.mkdir()
    text = open(os.path.join(GOLD, name + ".yml")).read()
    d.joinpath("input.yml").write_text("\n".join(f"n_timesteps: {n_steps}" if l.startswith("n_timesteps:") else l
                                                 for l in text.split("\n")))
    mesh.write_su2(d / "mesh.su2", names)
    print(f"[case] reference sim dir {name} on a {'hexahedral' if len(groups) == 1 else 'mixed'} mesh, "
          f"{n_steps} steps", flush=True)
    solver, run_cfg = fe.run_sim_dir(str(d), log=lambda *_: None)
    cfg.n_timesteps = n_steps
    ora = Oracle.from_mesh(nf.Mesh.read_su2(d / "mesh.su2", opts["patches"]), cfg)
    ora.set_freestream()
    for _ in range(n_steps):
        ora.step(ora.calc_timestep())
    np.testing.assert_array_equal(solver.get_solver_data().get_solution(), ora.get("U"))
    stride = opts["write_stride"]
    assert sorted(os.listdir(d / "output")) == sorted(f"n={k}.vtk" for k in range(0, n_steps + 1, stride))


def test_cpp_driver_on_mixed_su2(tmp_path):
    """The C++ host classes (netflux_solver.hpp) on a mixed-element SU2 file:
    netflux::Driver with the Sod state, against the oracle."""
    import subprocess
    mesh = mixed_box((5, 3, 3), kinds=("hex", "wedge", "pyr"), jitter=0.1, seed=13)
    path = tmp_path / "mesh.su2"
    mesh.write_su2(path, ["xm", "xp", "ym", "yp", "zm", "zp"])
    exe = os.path.join(os.path.dirname(nf.LIB_PATH), "netflux_driver")
    out = tmp_path / "U.bin"
    print("[case] netflux_driver su2 mixed 30 sod", flush=True)
    r = subprocess.run([exe, f"su2:{path}", "30", "sod", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    m = nf.Mesh.read_su2(path, {n: "SlipWall" for n in ["xm", "xp", "ym", "yp", "zm", "zp"]})
    ora = Oracle.from_mesh(m, sod_config())
    ora.set_state(sod_state(m), False)
    ora.run(30)
    np.testing.assert_array_equal(np.fromfile(out, dtype=np.float64).reshape(-1, 5), ora.get("U"))


def _hex_bump_channel(nx=24, ny=8, nz=4, h=0.06):
    """Hexahedral channel [0,3]x[0,1]x[0,0.5] with a Gaussian bump on the
    floor (nodes lifted, fading to the ceiling); farfield x-planes."""
    m = nf.Mesh.hex_box((nx, ny, nz), lengths=(3.0, 1.0, 0.5))
    ct, co, cn, pn, ft, fo, fn, pbc = m.elements()
    X = np.ctypeslib.as_array(m.desc.nodes, shape=(m.desc.n_nodes, 3)).copy()
    X[:, 2] += h * np.exp(-((X[:, 0] - 1.5) / 0.3) ** 2) * (1.0 - X[:, 2] / 0.5)
    faces, e = [], 0
    for k in pn:
        faces.append(fn[fo[e]:fo[e + k]].reshape(k, 4))
        e += k
    return nf.Mesh.from_elements(X, cn.reshape(-1, 8), faces, ["FarField", "FarField"] + ["SlipWall"] * 4)


def test_mixed_local_time_steps_reach_the_same_steady_state():
    """Local time stepping (each cell its own CFL step, from prim(U^n)) on a
    hexahedral bump channel reaches the steady state of global steps in
    fewer steps."""
    mesh = _hex_bump_channel()
    print(f"[case] hexahedral bump channel cells={mesh.n_cells}, global vs local steps", flush=True)
    out = {}
    for lts in (False, True):
        cfg = pulse_config(mach=0.5, CFL=0.8, limiter="NONE", local_time_step=lts, rk3_zero_old_first_step=False)
        s = nf.EulerSolver(cfg, mesh)
        s.set_freestream()
        st = s.run(40000, residual_drop=1e-7)
        assert st["converged"], st
        out[lts] = (st["steps_done"], s.get_solver_data().get_solution())
    assert out[True][0] < out[False][0]
    U_g, U_l = out[False][1], out[True][1]
    assert rel_l2(U_l, U_g) < 1e-5
    assert rel_l2(U_g, np.tile(U_g[0], (len(U_g), 1))) > 1e-3   # a non-trivial flow


def test_mixed_continued_run_with_zero_first_residual():
    """At rest with uniform density and slip walls the first step's density
    residual is exactly 0.  A continued run with a drop measures against that
    first residual, as the Driver loop on the host does, and so never stops
    early; the device-resident loop and the host loop (the flux-balance check
    on) agree step for step."""
    mesh = mixed_box((4, 3, 3), kinds=("hex", "wedge", "pyr"), jitter=0.1, seed=33)
    V = np.zeros((mesh.n_cells, 5))
    V[:, 0] = 1.0
    V[:, 4] = 1.0
    from tests.cases import prim_to_cons
    U0 = prim_to_cons(V)
    runs = []
    for chk in (False, True):
        s = nf.EulerSolver(sod_config(check_physical_validity=chk), mesh)
        s.set_state(U0)
        first = s.run(1)
        assert first["res_first"] == 0.0
        st = s.run(6, residual_drop=0.5, check_every=3)
        runs.append((st, s.residuals(), s.get_solver_data().get_solution()))
    for st, r, _ in runs:
        assert st["steps_done"] == 6 and not st["converged"] and r[0] == 0.0 and len(r) == 7
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
    np.testing.assert_array_equal(runs[0][2], runs[1][2])


def test_mixed_aux_fields_follow_the_state():
    """Gradient, limiter and flux balance are exported only for the state
    nf_eval_flux_balance saw: setting a state, stepping or running makes them
    unavailable until the next evaluation."""
    mesh = mixed_box((4, 3, 3), kinds=("hex", "wedge", "pyr"), jitter=0.1, seed=34)
    s = nf.EulerSolver(sod_config(), mesh)
    U0 = sod_state(mesh)
    for change in ("set_state", "step", "run"):
        s.set_state(U0)
        s.eval_flux_balance()
        s.get_solver_data().get_primvars_limiter()
        if change == "set_state":
            s.set_state(U0)
        elif change == "step":
            cfg = sod_config()
            s.calc_timestep(cfg)
            s.step(cfg)
        else:
            s.run(2)
        for get in ("get_primvars_limiter", "get_primvars_gradient", "get_flux_balance"):
            with pytest.raises(RuntimeError, match="nf_eval_flux_balance first"):
                getattr(s.get_solver_data(), get)()
