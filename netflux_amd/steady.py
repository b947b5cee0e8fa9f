"""Steady bump-channel solves: plain FP64 from freestream, batched mixed
precision, and learned warm start.

A case is a channel [0,3]x[0,1]x[0,0.5] with a Gaussian floor bump of height
h (sigma 0.2), farfield x-planes at Mach M and slip walls elsewhere, meshed
with (nx, ny, nz) proportional to (5, 2, 1) hexahedra split into 6 tets.
Steady runs start with U_old = U (the first-step zero-U_old quirk of the
reference would scale any start by about 1/6) and stop at a density-residual
drop of `drop` measured against the residual of the freestream state of the
same case, so plain and warm-started solves aim at the same absolute
residual.

Time to solution = wall time of the solve from its start (context creation
and mesh upload excluded), including feature evaluation and network
inference for warm starts.
"""
import math
import time
from dataclasses import dataclass

import numpy as np

from . import solver as S
from .mlp import Mlp

GAMMA = 1.4


@dataclass
class Case:
    nx: int
    ny: int
    nz: int
    mach: float
    h: float

    @property
    def n_cells(self):
        return 6 * self.nx * self.ny * self.nz


def dims_for(cells):
    """(5k, 2k, k) hexahedra with 6*10*k^3 tets nearest `cells`."""
    k = max(1, int(round((cells / 60.0) ** (1.0 / 3.0))))
    return 5 * k, 2 * k, k


def sample_cases(n, seed, cells_lo, cells_hi, mach=(0.3, 0.8), h=(0.02, 0.1)):
    """n cases with log-uniform cell counts and uniform Mach and bump height."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        c = math.exp(rng.uniform(math.log(cells_lo), math.log(cells_hi)))
        nx, ny, nz = dims_for(c)
        out.append(Case(nx, ny, nz, float(rng.uniform(*mach)), float(rng.uniform(*h))))
    return out


def mesh_of(case):
    return S.Mesh.bump_channel(case.nx, case.ny, case.nz, h=case.h)


def config_of(case, precision="fp64", local_time_step=True, cfl=0.5):
    return S.Config(density_fs=1.0, velocity_fs=(case.mach, 0.0, 0.0), pressure_fs=1.0 / GAMMA, CFL=cfl,
                    rk3_zero_old_first_step=False, precision=precision, local_time_step=local_time_step)


def freestream_state(cfg, n):
    V = np.asarray(cfg.primvars_inf(), np.float64)
    U = np.empty(5)
    U[0] = V[0]
    U[1:4] = V[0] * V[1:4]
    U[4] = V[4] / (GAMMA - 1) + 0.5 * V[0] * ((V[1] * V[1] + V[2] * V[2]) + V[3] * V[3])
    return np.tile(U, (n, 1))


def cons_to_prim(U):
    V = np.empty_like(U)
    V[:, 0] = U[:, 0]
    V[:, 1:4] = U[:, 1:4] / U[:, :1]
    V[:, 4] = (GAMMA - 1) * (U[:, 4] - 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2 + U[:, 3] ** 2) / U[:, 0])
    return V


def wall_pressure(mesh, U):
    """Pressure of the cells that own a floor-patch (y-) face, in reference order."""
    V = cons_to_prim(U)
    g = mesh.fv_grid()
    nodes, cells, pn, pf, pbc = mesh.arrays()
    # cells touching the floor: all three face nodes at the lowest logical layer
    first = np.r_[0, np.cumsum(pn)[:-1]]
    floor = pf[first[2]:first[2] + pn[2]]
    floor_nodes = np.zeros(len(nodes), bool)
    floor_nodes[floor.ravel()] = True
    owner = np.nonzero(floor_nodes[cells].sum(1) >= 3)[0]
    return owner, V[owner, 4]


class BatchRun:
    """One batched context over `cases`; solve() runs them all to the drop."""

    def __init__(self, cases, precision="mixed", local_time_step=True, autotune_steps=0):
        self.cases = cases
        self.meshes = [mesh_of(c) for c in cases]
        self.cfgs = [config_of(c, precision, local_time_step) for c in cases]
        self.batch = S.BatchSolver(self.cfgs, self.meshes)
        self.tune = None
        self.autotune_steps = autotune_steps

    def reset_freestream(self):
        for i, (cfg, N) in enumerate(zip(self.cfgs, self.batch.N)):
            U = freestream_state(cfg, N)
            self.batch.set_state(i, U, U)

    def solve(self, drop=1e-8, max_steps=200000, check_every=200, net=None, stall_window=0, stall_ratio=0.0):
        """Returns per-instance stats and the wall time (s) of the solve."""
        b = self.batch
        self.reset_freestream()
        r_fs = b.eval_residuals()
        for i, r in enumerate(r_fs):
            b.set_residual_ref(i, r)
        if self.autotune_steps and self.tune is None:
            self.tune = b.autotune(self.autotune_steps)
        t0 = time.perf_counter()
        fell = [False] * b.n
        if net is not None:
            for i, c in enumerate(self.cases):
                fell[i] = b.warm_start(i, net, (c.mach, c.h))
        st = b.run(max_steps, residual_drop=drop, check_every=check_every, stall_window=stall_window,
                   stall_ratio=stall_ratio)
        wall = time.perf_counter() - t0
        for i in range(b.n):
            st[i]["fell_back"] = int(st[i]["fell_back"] or fell[i])
            st[i]["r_ref"] = float(r_fs[i])
        return st, wall

    def solutions(self):
        return [self.batch.get_solution(i) for i in range(self.batch.n)]

    def features(self, i):
        c = self.cases[i]
        return self.batch.features(i, (c.mach, c.h))


def training_set(run, per_case, seed=0):
    """(features, primitive variables) of `per_case` random cells of each
    converged solution of a BatchRun."""
    rng = np.random.default_rng(seed)
    xs, ys = [], []
    for i, U in enumerate(run.solutions()):
        x = run.features(i)
        V = cons_to_prim(U)
        pick = rng.choice(len(U), size=min(per_case, len(U)), replace=False)
        xs.append(x[pick])
        ys.append(V[pick].astype(np.float32))
    return np.concatenate(xs), np.concatenate(ys)


def train_network(x, y, epochs=30, batch=4096, lr=2e-3, lr_decay=0.92, seed=0):
    net = Mlp(seed)
    stats = net.train(x, y, epochs=epochs, batch=batch, lr=lr, lr_decay=lr_decay, seed=seed)
    return net, stats
