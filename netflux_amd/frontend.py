"""Simulation-directory front end: a reference sim directory (input.yml +
mesh.su2) runs through the GPU solver with the reference's loop order and
VTK output.

- input.yml: the key set of ConfigParser::parse_yaml_file_options
  (src/ConfigParser.cpp:38-94) -- solver, time_scheme, inviscid_flux_scheme,
  spatial_order, limiter and grad_scheme (second order), initial_cond,
  density_fs / velocity_fs / pressure_fs (standard-air defaults),
  n_timesteps, CFL, write_stride, check_physical_validity,
  check_grid_validity, write_vtk_debug, patches (src/ConfigParser.cpp:
  110-137); enum strings as include/Utilities.hpp:15-98.
- mesh: the one file named mesh.su2 in the directory (set_mesh_name,
  src/ConfigParser.cpp:138-183; the native .nf reader of the reference is a
  stub, src/geometry/PrimalGrid.cpp:54-56).
- loop: Driver::solve (src/Driver.cpp:24-61): VTK at step 0, then
  calc_timestep, step, t += dt, VTK every write_stride steps, until
  n_timesteps.
- output: <sim_dir>/output/n=<step>.vtk (Config::get_unsteady_vtk_filename,
  include/Config.hpp:106-108), legacy ASCII as Output::write_vtk_ascii
  (src/Output.cpp:17-126): grid, then density, velocity, pressure per cell;
  6 significant digits like the reference's default stream precision, or 17
  with full_precision=True.  With write_vtk_debug the limiter and gradient
  are appended (Output::write_vtk_ascii_debug, src/Output.cpp:127-153);
  here they are evaluated at the written state, whereas the reference
  writes those left over from the last stage of the step.

Usage: python -m netflux_amd.frontend <sim_dir> [--full-precision] [--no-output]
"""
import argparse
import os
import shutil
import sys
import time

import numpy as np
import yaml

from . import solver as S

# standard air defaults of density_fs / pressure_fs (include/Utilities.hpp:100-107)
STANDARD_AIR_PRESSURE = 101325.0
STANDARD_AIR_DENSITY = 1.225

_ENUMS = {
    "solver": {"Euler"},
    "time_scheme": {"ExplicitEuler", "TVD_RK3"},
    "inviscid_flux_scheme": {"Rusanov", "HLLC"},
    "spatial_order": {"First", "Second"},
    "limiter": {"NONE", "Barth"},
    "grad_scheme": {"GreenGauss"},
    "initial_cond": {"Freestream"},
}
_BCS = {"NoSlipWall", "SlipWall", "FarField"}


class SimDirError(RuntimeError):
    pass


def _required(d, key):
    if key not in d:
        raise SimDirError(f"required option \"{key}\" missing from the configuration file")
    return d[key]


def _enum(d, key):
    v = str(_required(d, key))
    if v not in _ENUMS[key]:
        raise SimDirError(f"illegal value \"{v}\" of option \"{key}\"; legal values: {sorted(_ENUMS[key])}")
    return v


def parse_input(path):
    """input.yml -> (Config, options dict with n_timesteps, write_stride, patches...)."""
    with open(path) as f:
        d = yaml.safe_load(f) or {}
    if not isinstance(d, dict):
        raise SimDirError("configuration file is not a key-value map")
    _enum(d, "solver")
    opts = {"time_scheme": _enum(d, "time_scheme"), "inviscid_flux_scheme": _enum(d, "inviscid_flux_scheme"),
            "spatial_order": _enum(d, "spatial_order"), "initial_cond": _enum(d, "initial_cond")}
    # HLLC is an assert(false) stub in the reference (src/Numerics.cpp:20-23);
    # here it is Toro's published scheme (FP64)
    opts["limiter"] = _enum(d, "limiter") if opts["spatial_order"] == "Second" else "NONE"
    if opts["spatial_order"] == "Second":
        _enum(d, "grad_scheme")
    vel = d.get("velocity_fs", [0.0, 0.0, 0.0])
    if not (isinstance(vel, (list, tuple)) and len(vel) == 3):
        raise SimDirError("velocity_fs must be a list of three numbers")
    opts["n_timesteps"] = int(_required(d, "n_timesteps"))
    opts["write_stride"] = int(d.get("write_stride", 1))
    if opts["write_stride"] < 1:
        raise SimDirError("write_stride must be positive")
    opts["write_vtk_debug"] = bool(d.get("write_vtk_debug", False))
    opts["check_grid_validity"] = bool(d.get("check_grid_validity", False))
    patches = _required(d, "patches")
    if not isinstance(patches, list):
        raise SimDirError("\"patches\" setting in configuration file is not a valid sequence")
    bc = {}
    for p in patches:
        if not isinstance(p, dict) or len(p) != 1:
            raise SimDirError("Entry in \"patches\" setting in configuration file is not a key-value pair")
        (name, t), = p.items()
        name, t = str(name), str(t)
        if name in bc:
            raise SimDirError(f"Duplicate patch name \"{name}\" specified in config file")
        if t not in _BCS:
            raise SimDirError(f"illegal boundary type \"{t}\" for patch \"{name}\"")
        bc[name] = t
    opts["patches"] = bc
    cfg = S.Config(time_scheme=opts["time_scheme"], spatial_order=opts["spatial_order"], limiter=opts["limiter"],
                   inviscid_flux_scheme=opts["inviscid_flux_scheme"], CFL=float(_required(d, "CFL")),
                   density_fs=float(d.get("density_fs", STANDARD_AIR_DENSITY)),
                   velocity_fs=tuple(float(v) for v in vel),
                   pressure_fs=float(d.get("pressure_fs", STANDARD_AIR_PRESSURE)),
                   check_physical_validity=bool(d.get("check_physical_validity", False)),
                   n_timesteps=opts["n_timesteps"])
    return cfg, opts


def find_mesh(sim_dir):
    found = [f for f in sorted(os.listdir(sim_dir)) if f in ("mesh.su2", "mesh.nf") and
             os.path.isfile(os.path.join(sim_dir, f))]
    if not found:
        raise SimDirError(f"No files in the sim-dir '{sim_dir}' has a valid mesh name. Valid mesh names are: "
                          "'mesh.su2', 'mesh.nf'")
    if len(found) > 1:
        raise SimDirError("More than one file with valid mesh name found in sim-directory: " + ", ".join(found))
    if found[0].endswith(".nf"):
        raise SimDirError("the native .nf mesh format is a stub in the reference; use mesh.su2")
    return os.path.join(sim_dir, found[0])


def grid_validity_report(mesh):
    """Elements with negative volume or a centroid on its bounding box
    (PrimalGrid::partial_validity_check, src/geometry/PrimalGrid.cpp:269-301,
    volume = (n1-n0) x (n2-n0) . (n3-n0) / 6 in VTK order).  The reference
    constructs these errors without throwing them, so a run continues; here
    they are returned and printed as warnings."""
    if mesh.is_mixed:
        # element volumes and centroids by the reference's formulas
        # (volume_element_calc_geometry_properties), from the face graph
        ct, co, cn, *_ = mesh.elements()
        d = mesh.desc
        nodes = np.ctypeslib.as_array(d.nodes, shape=(d.n_nodes, 3))
        g = mesh.fv_grid()
        N = len(ct)
        vol, cen = g["volume"][:N], g["centroid"][:N]
        hi, lo = cen.copy(), cen.copy()
        for e in range(N):
            P = nodes[cn[co[e]:co[e + 1]].astype(np.int64)]
            hi[e] = np.maximum(hi[e], P.max(axis=0))
            lo[e] = np.minimum(lo[e], P.min(axis=0))
    else:
        nodes, cells, *_ = mesh.arrays()
        P = nodes[cells.astype(np.int64)]                       # [N][4][3]
        a, b, c = P[:, 1] - P[:, 0], P[:, 2] - P[:, 0], P[:, 3] - P[:, 0]
        vol = np.einsum("ij,ij->i", np.cross(a, b), c) / 6.0
        cen = P.mean(axis=1)
        hi = np.maximum(P.max(axis=1), cen)
        lo = np.minimum(P.min(axis=1), cen)
    flat = np.any(np.isclose(hi, cen, rtol=1e-12, atol=0.0) | np.isclose(lo, cen, rtol=1e-12, atol=0.0), axis=1)
    return {"negative_volume": np.nonzero(vol < 0.0)[0], "centroid_on_bbox": np.nonzero(flat)[0]}


def _fmt(full):
    return "%.17g" if full else "%.6g"


def write_vtk(path, mesh, U, full_precision=False, limiter=None, gradient=None):
    """Legacy ASCII VTK of the mesh (tetrahedra or mixed elements, VTK cell
    types) with cell density, velocity and pressure (src/Output.cpp:48-126)."""
    ct, co, cn, *_ = mesh.elements()
    d = mesh.desc
    nodes = np.ctypeslib.as_array(d.nodes, shape=(d.n_nodes, 3))
    f = _fmt(full_precision)
    N = len(ct)
    rho = U[:, 0]
    vel = U[:, 1:4] / U[:, :1]
    p = (1.4 - 1.0) * (U[:, 4] - 0.5 * ((U[:, 1] * U[:, 1] + U[:, 2] * U[:, 2]) + U[:, 3] * U[:, 3]) / U[:, 0])
    with open(path, "w") as o:
        o.write("# vtk DataFile Version 3.0\nNetFlux\nASCII\n\nDATASET UNSTRUCTURED_GRID\n\n")
        o.write(f"POINTS {len(nodes)} double\n")
        np.savetxt(o, nodes, fmt=f)
        o.write(f"\nCELLS {N} {N + len(cn)}\n")
        if mesh.is_mixed:
            for e in range(N):
                row = cn[co[e]:co[e + 1]]
                o.write(f"{len(row)} " + " ".join(str(int(x)) for x in row) + "\n")
        else:
            np.savetxt(o, np.column_stack([np.full(N, 4), cn.reshape(N, 4)]), fmt="%d")
        o.write(f"\nCELL_TYPES {N}\n")
        np.savetxt(o, ct.astype(np.int64), fmt="%d")
        o.write(f"\nCELL_DATA {N}\n\nSCALARS density double 1\nLOOKUP_TABLE default\n")
        np.savetxt(o, rho, fmt=f)
        o.write("\nVECTORS velocity double\n")
        np.savetxt(o, vel, fmt=f)
        o.write("\nSCALARS pressure double 1\nLOOKUP_TABLE default\n")
        np.savetxt(o, p, fmt=f)
        if limiter is not None and gradient is not None:
            for k in range(5):
                o.write(f"\nSCALARS limiter_EQ{k} double 1\nLOOKUP_TABLE default\n")
                np.savetxt(o, limiter[:, k], fmt=f)
            for k in range(5):
                o.write(f"\nVECTORS gradient_EQ{k} double\n")
                np.savetxt(o, gradient[:, [k, k + 5, k + 10]], fmt=f)


def read_vtk_cell_data(path):
    """Cell arrays of a file written by write_vtk (for checks)."""
    out = {}
    with open(path) as f:
        lines = f.read().split("\n")
    i = lines.index(next(l for l in lines if l.startswith("CELL_DATA")))
    n = int(lines[i].split()[1])
    i += 1
    while i < len(lines):
        l = lines[i]
        if l.startswith("SCALARS") or l.startswith("VECTORS"):
            name, vec = l.split()[1], l.startswith("VECTORS")
            i += 1 if vec else 2
            out[name] = np.loadtxt(lines[i:i + n])
            i += n
        else:
            i += 1
    return out


def run_sim_dir(sim_dir, full_precision=False, output=True, log=print):
    """Runs a sim directory; returns (EulerSolver, Config)."""
    cfg, opts = parse_input(os.path.join(sim_dir, "input.yml"))
    mesh = S.Mesh.read_su2(find_mesh(sim_dir), opts["patches"])
    if opts["check_grid_validity"]:
        rep = grid_validity_report(mesh)
        for k, idx in rep.items():
            if len(idx):
                log(f"warning: {len(idx)} element(s) with {k.replace('_', ' ')} (first {int(idx[0])})")
    solver = S.EulerSolver(cfg, mesh)  # freestream initial condition
    out_dir = os.path.join(sim_dir, "output")
    if output:
        if os.path.exists(out_dir):
            shutil.rmtree(out_dir)
        os.makedirs(out_dir)

    def write(step):
        if not output or step % opts["write_stride"]:
            return
        lim = grad = None
        if opts["write_vtk_debug"]:
            solver.eval_flux_balance()
            lim = solver.get_solver_data().get_primvars_limiter()
            grad = solver.get_solver_data().get_primvars_gradient()
        write_vtk(os.path.join(out_dir, f"n={step}.vtk"), mesh, solver.get_solver_data().get_solution(),
                  full_precision, lim, grad)

    t0 = time.perf_counter()
    cfg.timestep, cfg.time = 0, 0.0
    write(0)
    while True:
        solver.calc_timestep(cfg)
        solver.step(cfg)
        cfg.timestep += 1
        cfg.time += cfg.delta_time
        write(cfg.timestep)
        if cfg.timestep >= opts["n_timesteps"]:
            break
    log(f"Solver finished\nElapsed time: {time.perf_counter() - t0:.3f} s ({cfg.timestep} steps, t = {cfg.time:g})")
    return solver, cfg


def main(argv=None):
    ap = argparse.ArgumentParser(description="run a NetFlux sim directory on the GPU")
    ap.add_argument("sim_dir")
    ap.add_argument("--full-precision", action="store_true")
    ap.add_argument("--no-output", action="store_true")
    a = ap.parse_args(argv)
    try:
        run_sim_dir(a.sim_dir, a.full_precision, not a.no_output)
    except (SimDirError, RuntimeError) as e:
        print(f"Error: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
