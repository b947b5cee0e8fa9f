"""Test cases shared by the CPU and GPU tests (sizes, initial states)."""
import numpy as np

import netflux_amd as nf

GAMMA = 1.4


def prim_to_cons(V):
    """EulerEqs::prim_to_cons (include/SolverData.hpp:166-177), vectorised,
    same association order."""
    U = np.empty_like(V)
    U[:, 0] = V[:, 0]
    U[:, 1] = V[:, 0] * V[:, 1]
    U[:, 2] = V[:, 0] * V[:, 2]
    U[:, 3] = V[:, 0] * V[:, 3]
    gm1_inv = 1 / (GAMMA - 1)
    U[:, 4] = gm1_inv * V[:, 4] + 0.5 * V[:, 0] * ((V[:, 1] * V[:, 1] + V[:, 2] * V[:, 2]) + V[:, 3] * V[:, 3])
    return U


def sod_state(mesh):
    """Sod tube along x by cell centroid: x<0.5 (1,0,1), else (0.125,0,0.1)."""
    g = mesh.fv_grid()
    N = mesh.n_cells
    x = g["centroid"][:N, 0]
    V = np.zeros((N, 5))
    left = x < 0.5
    V[:, 0] = np.where(left, 1.0, 0.125)
    V[:, 4] = np.where(left, 1.0, 0.1)
    return prim_to_cons(V)


def pulse_state(mesh, mach=0.5, amp=0.1, sigma=0.1, centre=(0.5, 0.5, 0.5)):
    """Freestream rho=1, p=1/gamma (c=1), u=(mach,0,0) plus a Gaussian density pulse."""
    g = mesh.fv_grid()
    N = mesh.n_cells
    d = g["centroid"][:N] - np.asarray(centre)
    r2 = (d * d).sum(1)
    V = np.zeros((N, 5))
    V[:, 0] = 1.0 + amp * np.exp(-r2 / (2 * sigma * sigma))
    V[:, 1] = mach
    V[:, 4] = 1.0 / GAMMA
    return prim_to_cons(V)


def sod_config(**kw):
    c = nf.Config(time_scheme="TVD_RK3", spatial_order="Second", limiter="Barth", CFL=0.5,
                  density_fs=1.0, velocity_fs=(0.0, 0.0, 0.0), pressure_fs=1.0)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def pulse_config(mach=0.5, **kw):
    c = nf.Config(time_scheme="TVD_RK3", spatial_order="Second", limiter="Barth", CFL=0.5,
                  density_fs=1.0, velocity_fs=(mach, 0.0, 0.0), pressure_fs=1.0 / GAMMA)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
