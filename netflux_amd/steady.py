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


# Steady defaults: CFL 1.0 is stable for the unlimited scheme (1.5 is not);
# the Barth limiter stalls these flows near a 5e-4 drop, so 1e-8 targets use
# the unlimited second-order reconstruction (limiter NONE of the reference).
STEADY_CFL = 1.0
# a residual plateau counts as the FP32 noise floor only below this drop
STALL_GUARD = 1e-3
STEADY_LIMITER = "NONE"


def config_of(case, precision="fp64", local_time_step=True, cfl=STEADY_CFL, limiter=STEADY_LIMITER):
    return S.Config(density_fs=1.0, velocity_fs=(case.mach, 0.0, 0.0), pressure_fs=1.0 / GAMMA, CFL=cfl,
                    limiter=limiter, rk3_zero_old_first_step=False, precision=precision,
                    local_time_step=local_time_step)


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

    def __init__(self, cases, precision="mixed", local_time_step=True, autotune_steps=0, cfl=STEADY_CFL,
                 limiter=STEADY_LIMITER):
        self.cases = cases
        self.meshes = [mesh_of(c) for c in cases]
        self.cfgs = [config_of(c, precision, local_time_step, cfl, limiter) for c in cases]
        self.batch = S.BatchSolver(self.cfgs, self.meshes)
        self.precision = precision
        self.keep_switch = False  # keep the states reached when mixed precision hands over to FP64
        self.switch_solutions = None
        self.switch_steps = None
        self.switch_drops = None
        self.tune = None
        self.autotune_steps = autotune_steps

    def reset_freestream(self):
        for i, (cfg, N) in enumerate(zip(self.cfgs, self.batch.N)):
            U = freestream_state(cfg, N)
            self.batch.set_state(i, U, U)

    def solve(self, drop=1e-8, max_steps=200000, check_every=200, net=None, stall_window=0, stall_ratio=0.0,
              chunk_steps=10000, switch_drop=1e-6):
        """Returns per-instance stats and the wall time (s) of the solve.
        Mixed precision runs to `switch_drop` (above its FP32 noise floor),
        then the same batch continues in FP64 to `drop`."""
        b = self.batch
        b.set_precision(self.precision)
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
        st = None
        done = 0
        phases = [("mixed", switch_drop), ("fp64", drop)] if self.precision == "mixed" and switch_drop > drop \
            else [(self.precision, drop)]
        phase = 0
        b.set_precision(phases[0][0])
        while done < max_steps:
            k = min(chunk_steps, max_steps - done)
            part = b.run(k, residual_drop=phases[phase][1], check_every=check_every, stall_window=stall_window,
                         stall_ratio=stall_ratio)
            if st is None:
                st = part
            else:
                for s0, s1 in zip(st, part):
                    s0["steps_done"] += s1["steps_done"]
                    s0["converged"] = s1["converged"]
                    s0["fell_back"] = int(s0["fell_back"] or s1["fell_back"])
                    s0["res_last"] = s1["res_last"] if s1["steps_done"] else s0["res_last"]
            done += k
            n_conv = sum(s["converged"] for s in st)
            # an unconverged instance whose residual fell by less than half over
            # the chunk, once below a 1e-3 drop (earlier plateaus are transients),
            # has reached the noise floor of the current precision
            stalled = sum(1 for s, p, r in zip(st, part, r_fs) if not s["converged"] and p["steps_done"] and
                          p["res_last"] > 0.5 * s.get("res_prev", np.inf) and p["res_last"] < STALL_GUARD * r)
            for s, p in zip(st, part):
                if p["steps_done"]:
                    s["res_prev"] = p["res_last"]
            _log(f"[solve] {done} steps, {n_conv}/{b.n} converged ({phases[phase][0]}, drop "
                 f"{phases[phase][1]:.0e}), {time.perf_counter() - t0:.1f} s")
            if n_conv == b.n or (phase + 1 < len(phases) and n_conv + stalled == b.n):
                if phase + 1 == len(phases):
                    break
                self.switch_steps = done
                self.switch_drops = [s["res_last"] / r for s, r in zip(st, r_fs)]
                self.switch_solutions = self.solutions() if self.keep_switch else None
                phase += 1
                b.set_precision(phases[phase][0])
                _log(f"[solve] switching to {phases[phase][0]} after {done} steps")
                for s0 in st:
                    s0["converged"] = 0
                    s0.pop("res_prev", None)
        wall = time.perf_counter() - t0
        for i in range(b.n):
            st[i]["fell_back"] = int(st[i]["fell_back"] or fell[i])
            st[i]["r_ref"] = float(r_fs[i])
            st[i].pop("res_prev", None)
        return st, wall

    def solutions(self):
        return [self.batch.get_solution(i) for i in range(self.batch.n)]

    def features(self, i):
        c = self.cases[i]
        return self.batch.features(i, (c.mach, c.h))


def training_set(run, per_case, stats, seed=0):
    """(features, primitive variables) of `per_case` random cells of each
    converged, finite solution of a BatchRun."""
    rng = np.random.default_rng(seed)
    xs, ys = [], []
    for i, U in enumerate(run.solutions()):
        if not stats[i]["converged"] or not np.all(np.isfinite(U)):
            continue
        x = run.features(i)
        V = cons_to_prim(U)
        pick = rng.choice(len(U), size=min(per_case, len(U)), replace=False)
        xs.append(x[pick])
        ys.append(V[pick].astype(np.float32))
    return np.concatenate(xs), np.concatenate(ys)


def train_network(x, y, epochs=200, batch=2048, lr=2e-3, lr_decay=0.98, seed=0):
    """Inputs normalised by their mean / standard deviation; outputs by
    their mean and max(std, 0.01): the spanwise velocity is zero up to
    discretisation noise, which a unit-variance target would amplify."""
    net = Mlp(seed)
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    sx = x.std(0)
    net.set_norm(x.mean(0), np.where(sx > 0, 1.0 / np.where(sx > 0, sx, 1.0), 1.0), y.mean(0),
                 np.maximum(y.std(0), 0.01))
    stats = net.train(x, y, epochs=epochs, batch=batch, lr=lr, lr_decay=lr_decay, seed=seed)
    return net, stats


def save_network(net, path, meta=None):
    """Parameters, normalisation and a JSON description in one .npz file."""
    import json
    sh, sc, osh, osc = net.get_norm()
    np.savez(path, params=net.params, in_shift=sh, in_scale=sc, out_shift=osh, out_scale=osc,
             meta=np.array(json.dumps(meta or {})))


def load_network(path):
    """A network saved by save_network (plain arrays; nothing is unpickled)."""
    import json
    with np.load(path, allow_pickle=False) as z:
        net = Mlp(0)
        net.params = z["params"]
        net.set_norm(z["in_shift"], z["in_scale"], z["out_shift"], z["out_scale"])
        meta = json.loads(str(z["meta"]))
    return net, meta


def solve_training_chunk(train_cases, train_cells, seed, per_case=4000, train_drop=1e-6, max_steps=60000,
                         check_every=200):
    """One chunk of training data: steady FP64 solutions of `train_cases`
    cases sampled with `seed` (batched, to `train_drop`) and `per_case`
    samples of each converged one; returns (x, y, description)."""
    tcases = sample_cases(train_cases, seed, train_cells, train_cells)
    trun = BatchRun(tcases, precision="fp64")
    tst, twall = trun.solve(train_drop, max_steps, check_every)
    x, y = training_set(trun, per_case, tst, seed)
    del trun
    info = {"cases": train_cases, "cells": int(tcases[0].n_cells), "case_seed": seed, "drop": train_drop,
            "solves": _summary(tst), "solve_wall_s": twall, "samples": int(len(x))}
    _log(f"[train data] {info}")
    return x.astype(np.float32), y.astype(np.float32), info


def save_training_chunk(path, x, y, info):
    import json
    np.savez(path, x=x, y=y, meta=np.array(json.dumps(info)))


def train_from_chunks(paths, epochs=200):
    """A network trained by the library's kernels on the samples of saved
    training chunks (plain arrays, nothing unpickled)."""
    import json
    xs, ys, infos = [], [], []
    for p in paths:
        with np.load(p, allow_pickle=False) as z:
            xs.append(z["x"])
            ys.append(z["y"])
            infos.append(json.loads(str(z["meta"])))
    x, y = np.concatenate(xs), np.concatenate(ys)
    t0 = time.perf_counter()
    net, tstats = train_network(x, y, epochs=epochs)
    _log(f"[train] {tstats}")
    info = {"cases": int(sum(i["cases"] for i in infos)),
            "converged": int(sum(i["solves"]["converged"] for i in infos)),
            "cells": infos[0]["cells"], "case_seeds": [i["case_seed"] for i in infos], "drop": infos[0]["drop"],
            "solve_wall_s": float(sum(i["solve_wall_s"] for i in infos)), "samples": int(len(x)),
            "epochs": epochs, "train": tstats, "train_s": time.perf_counter() - t0, "chunks": infos}
    return net, info


def train_from_solves(train_cases, train_cells, seed, per_case=4000, train_drop=1e-6, max_steps=60000,
                      check_every=200, epochs=200):
    """Steady FP64 solutions of `train_cases` sampled cases of about
    `train_cells` cells (batched, to `train_drop`), `per_case` cells of each as
    samples, and a network trained on them by the library's kernels."""
    tcases = sample_cases(train_cases, seed, train_cells, train_cells)
    t0 = time.perf_counter()
    trun = BatchRun(tcases, precision="fp64")
    tst, twall = trun.solve(train_drop, max_steps, check_every)
    x, y = training_set(trun, per_case, tst, seed)
    del trun
    _log(f"[train data] {_summary(tst)} {twall:.2f} s, {len(x)} samples")
    net, tstats = train_network(x, y, epochs=epochs)
    _log(f"[train] {tstats}")
    info = {"cases": train_cases, "cells": int(tcases[0].n_cells), "case_seed": seed, "drop": train_drop,
            "solves": _summary(tst), "solve_wall_s": twall, "samples": int(len(x)), "epochs": epochs,
            "train": tstats, "total_s": time.perf_counter() - t0}
    return net, info


def _log(msg):
    import sys
    print(msg, file=sys.stderr, flush=True)


def _summary(stats):
    steps = [s["steps_done"] for s in stats]
    drops = [s["res_last"] / s["r_ref"] if s.get("r_ref") else float("nan") for s in stats]
    return {"converged": int(sum(s["converged"] for s in stats)), "n": len(stats),
            "final_drop_max": float(np.max(drops)), "final_drop_median": float(np.median(drops)),
            "steps_min": int(min(steps)), "steps_median": float(np.median(steps)), "steps_max": int(max(steps)),
            "fell_back": int(sum(s.get("fell_back", 0) for s in stats))}


def study(n_cases=8, seed=1, cells_lo=2e4, cells_hi=3e5, drop=1e-8, max_steps=60000, check_every=200,
          train_cases=0, train_cells=2e5, per_case=4000, epochs=200, train_drop=1e-6, plain=True,
          plain_lts=False, partial=None, train_max_steps=60000, net_path=None, mach=(0.3, 0.8),
          exclude_unconverged=True):
    """Plain FP64 solves one by one (the reference's global time step; with
    plain_lts also, or only, with local time steps, which separates the gain
    of local time stepping from that of batching), the batched FP64 and
    mixed-precision solves with local time steps, and (train_cases > 0, or a
    saved network at net_path) the same batches from a learned warm start."""
    cases = sample_cases(n_cases, seed, cells_lo, cells_hi, mach=mach)
    rep = {"cases": [c.__dict__ | {"cells": c.n_cells} for c in cases], "drop": drop, "plain_lts": plain_lts}

    def save():
        if partial:
            import json
            with open(partial, "w") as f:
                json.dump(rep, f)

    save()
    plain_U = []
    modes = {"both": [False, True], True: [True], False: [False]}[plain_lts] if plain else []
    for lts in modes:
        plain_wall, plain_stats, sols = [], [], []
        for i, c in enumerate(cases):
            run = BatchRun([c], precision="fp64", local_time_step=lts)
            st, wall = run.solve(drop, max_steps, check_every)
            plain_stats.append(st[0])
            plain_wall.append(wall)
            sols.append(run.solutions()[0])
            _log(f"[plain{' lts' if lts else ''}] case {i} cells {c.n_cells} steps {st[0]['steps_done']} "
                 f"conv {st[0]['converged']} {wall:.2f} s")
            del run
        if not plain_U:
            plain_U = sols  # the first plain mode is the reference for the comparisons
        rep["plain_fp64_lts" if lts else "plain_fp64"] = _summary(plain_stats) | {
            "wall_total_s": sum(plain_wall), "per_instance_s": sum(plain_wall) / n_cases,
            "steps": [s["steps_done"] for s in plain_stats], "wall_s": plain_wall}
        save()
        if exclude_unconverged and not lts:
            # a case the reference's method does not converge (a transonic bump)
            # is left out of every mode, and listed
            keep = [i for i, st in enumerate(plain_stats) if st["converged"]]
            if len(keep) < len(cases):
                rep["excluded"] = [cases[i].__dict__ | {"cells": cases[i].n_cells}
                                   for i in range(len(cases)) if i not in keep]
                cases = [cases[i] for i in keep]
                plain_U = [plain_U[i] for i in keep]
                n_cases = len(cases)
                rep["plain_fp64"] = _summary([plain_stats[i] for i in keep]) | {
                    "wall_total_s": sum(plain_wall[i] for i in keep),
                    "per_instance_s": sum(plain_wall[i] for i in keep) / n_cases,
                    "steps": [plain_stats[i]["steps_done"] for i in keep], "wall_s": [plain_wall[i] for i in keep]}
                save()

    def compare(sols):
        """Wall pressure and state against the plain FP64 converged solutions."""
        dp, du, dl = [], [], []
        for i, (c, U) in enumerate(zip(cases, sols)):
            if not plain_U:
                break
            m = mesh_of(c)
            _, p = wall_pressure(m, U)
            _, p0 = wall_pressure(m, plain_U[i])
            dp.append(float(np.max(np.abs(p - p0))))
            du.append(float(np.max(np.abs(U - plain_U[i]))))
            dl.append(float(np.linalg.norm(U - plain_U[i]) / np.linalg.norm(plain_U[i])))
        if not dp:
            return {}
        return {"wall_p_maxabs_max": max(dp), "wall_p_maxabs_median": float(np.median(dp)),
                "U_maxabs_max": max(du), "U_relL2_max": max(dl)}

    run64 = BatchRun(cases, precision="fp64", autotune_steps=20)
    st, wall = run64.solve(drop, max_steps, check_every)
    _log(f"[batch fp64] {_summary(st)} {wall:.2f} s, tune {run64.tune}")
    rep["batch_fp64"] = _summary(st) | {"wall_s": wall, "per_instance_s": wall / n_cases, "tune": run64.tune,
                                        "steps": [s["steps_done"] for s in st]} | compare(run64.solutions())
    save()
    run = BatchRun(cases, precision="mixed", autotune_steps=20)
    st, wall = run.solve(drop, max_steps, check_every)
    _log(f"[batch mixed] {_summary(st)} {wall:.2f} s, tune {run.tune}")
    rep["batch_mixed"] = _summary(st) | {"wall_s": wall, "per_instance_s": wall / n_cases, "tune": run.tune,
                                         "steps": [s["steps_done"] for s in st]} | compare(run.solutions())
    save()
    net = None
    if net_path:
        net, rep["training"] = load_network(net_path)
        rep["training"]["file"] = net_path
    elif train_cases:
        net, rep["training"] = train_from_solves(train_cases, train_cells, seed + 1000, per_case, train_drop,
                                                 train_max_steps, check_every, epochs)
        save()
    if net is not None:
        st, wall = run.solve(drop, max_steps, check_every, net=net, stall_window=4000, stall_ratio=0.5)
        _log(f"[batch mixed warm] {_summary(st)} {wall:.2f} s")
        rep["batch_mixed_warm"] = _summary(st) | {"wall_s": wall, "per_instance_s": wall / n_cases,
                                                   "steps": [s["steps_done"] for s in st]} | compare(run.solutions())
        save()
        st, wall = run64.solve(drop, max_steps, check_every, net=net, stall_window=4000, stall_ratio=0.5)
        _log(f"[batch fp64 warm] {_summary(st)} {wall:.2f} s")
        rep["batch_fp64_warm"] = _summary(st) | {"wall_s": wall, "per_instance_s": wall / n_cases,
                                                  "steps": [s["steps_done"] for s in st]} | compare(run64.solutions())
        save()
    return rep


def time_to_drop(net, n_cases=8, seed=7, cells_lo=2e4, cells_hi=1e5, drop=1e-8, max_steps=120000,
                 check_every=200, mach=(0.3, 0.7)):
    """Time to a `drop` density-residual drop per instance, three ways, on
    `n_cases` sampled bump channels:
    plain    -- FP64, one context per case, the reference's global time step,
                from the freestream (the reference's method on the GPU, with
                the launch configuration the autotuner picks for that mesh);
    batched  -- all cases in one context, mixed precision to its FP32 floor
                then FP64, local time steps, from the freestream;
    learned  -- the same batch from the network's warm start.
    Per-instance time is the batch wall time / the number of cases for the
    batched modes.  Wall pressure of the learned solves is compared with the
    plain ones.  A case whose plain solve does not reach the drop in
    max_steps (a transonic bump: the unlimited scheme oscillates at the
    shock, Barth stalls) is left out of the comparison and listed."""
    sampled = sample_cases(n_cases, seed, cells_lo, cells_hi, mach=mach)
    cases, excluded = [], []
    plain_wall, plain_steps, plain_U = [], [], []
    for c in sampled:
        # the plain solve gets the autotuned launch too (picked before the timed solve)
        run = BatchRun([c], precision="fp64", local_time_step=False, autotune_steps=20)
        st, wall = run.solve(drop, max_steps, check_every)
        if not st[0]["converged"]:
            fd = st[0]["res_last"] / st[0]["r_ref"]
            excluded.append(c.__dict__ | {"cells": c.n_cells, "final_drop": fd if math.isfinite(fd) else None})
            continue
        cases.append(c)
        plain_wall.append(wall)
        plain_steps.append(int(st[0]["steps_done"]))
        plain_U.append(run.solutions()[0])
        del run
    n_cases = len(cases)
    if n_cases == 0:
        raise RuntimeError("no case reached the drop in the plain solve")
    run = BatchRun(cases, precision="mixed", autotune_steps=20)
    st_b, wall_b = run.solve(drop, max_steps, check_every)
    st_l, wall_l = run.solve(drop, max_steps, check_every, net=net, stall_window=4000, stall_ratio=0.5)
    for st in (st_b, st_l):
        if not all(s["converged"] for s in st):
            raise RuntimeError("a batched solve did not reach the drop")
    dp = []
    for i, c in enumerate(cases):
        m = run.meshes[i]
        _, p = wall_pressure(m, run.batch.get_solution(i))
        _, p0 = wall_pressure(m, plain_U[i])
        dp.append(float(np.max(np.abs(p - p0))))
    plain_per = sum(plain_wall) / n_cases
    return {
        "cases": n_cases, "case_seed": seed, "mach_range": list(mach), "excluded": excluded,
        "cells": [c.n_cells for c in cases],
        "mach": [round(c.mach, 4) for c in cases], "bump_h": [round(c.h, 4) for c in cases], "drop": drop,
        "plain_fp64": {"per_instance_s": plain_per, "wall_s": plain_wall, "steps": plain_steps},
        "batched_freestream": {"per_instance_s": wall_b / n_cases, "wall_s": wall_b,
                               "steps": [int(s["steps_done"]) for s in st_b], "launch": run.tune},
        "learned_batched": {"per_instance_s": wall_l / n_cases, "wall_s": wall_l,
                            "steps": [int(s["steps_done"]) for s in st_l],
                            "fell_back": int(sum(s["fell_back"] for s in st_l))},
        "speedup_per_instance": plain_per / (wall_l / n_cases),
        "speedup_batched_freestream": plain_per / (wall_b / n_cases),
        "warm_start_gain": wall_b / wall_l,
        "wall_p_maxabs_diff_vs_plain": max(dp),
    }


def c5_scaled(net_path, n_cases=2, cells=1.5e6, seed=11, drop=1e-6, max_steps=400000, check_every=500,
              precision="fp64", freestream=True):
    """The learned warm start at the inference resolution of the learned
    workload: `n_cases` unseen bump channels of about `cells` cells (Mach
    U[0.3, 0.7], h U[0.02, 0.1]), batched, local time steps, solved to `drop`
    from the freestream and from the network's prediction in the same mode;
    the warm start's own cost (features, inference, state upload) is timed
    separately as well as inside the solve.  freestream=False skips the
    freestream solve (it does not depend on the network)."""
    net, meta = load_network(net_path)
    cases = sample_cases(n_cases, seed, cells, cells, mach=(0.3, 0.7))
    run = BatchRun(cases, precision=precision)
    t0 = time.perf_counter()
    for i, c in enumerate(cases):
        run.batch.warm_start(i, net, (c.mach, c.h))
    run.batch.get_solution(0)  # synchronises
    t_warm = time.perf_counter() - t0
    rep = {"cases": [c.__dict__ | {"cells": c.n_cells} for c in cases], "drop": drop, "precision": precision,
           "network": {"file": net_path, "trained_on_cells": meta.get("cells"), "samples": meta.get("samples")},
           "warm_start_s": t_warm}
    if freestream:
        st_fs, wall_fs = run.solve(drop, max_steps, check_every)
        rep["freestream"] = {"wall_s": wall_fs, "steps": [int(s["steps_done"]) for s in st_fs],
                             "converged": [int(s["converged"]) for s in st_fs]}
    st_w, wall_w = run.solve(drop, max_steps, check_every, net=net, stall_window=20000, stall_ratio=0.5)
    rep["learned"] = {"wall_s": wall_w, "steps": [int(s["steps_done"]) for s in st_w],
                      "converged": [int(s["converged"]) for s in st_w],
                      "fell_back": [int(s["fell_back"]) for s in st_w]}
    if freestream:
        rep["gain"] = wall_fs / wall_w
    return rep


def c3(drop=1e-8, max_steps=500000, check_every=200):
    """The 2M-tet channel (160 x 64 x 32, Mach 0.5, h = 0.05): FP64 with the
    reference's global time step against mixed precision with local steps --
    at its FP32 noise floor and after the FP64 finish -- at steady state."""
    case = Case(160, 64, 32, 0.5, 0.05)
    rep = {"case": case.__dict__ | {"cells": case.n_cells}, "drop": drop}
    ref = BatchRun([case], precision="fp64", local_time_step=False)
    st, wall = ref.solve(drop, max_steps, check_every)
    U64 = ref.solutions()[0]
    rep["fp64_global_dt"] = _summary(st) | {"wall_s": wall}
    _log(f"[c3 fp64] {rep['fp64_global_dt']}")
    del ref
    mx = BatchRun([case], precision="mixed")
    mx.keep_switch = True
    st, wall = mx.solve(drop, max_steps, check_every)
    rep["mixed_then_fp64"] = _summary(st) | {"wall_s": wall, "switch_steps": mx.switch_steps,
                                             "drop_at_switch": mx.switch_drops[0]}
    _log(f"[c3 mixed] {rep['mixed_then_fp64']}")
    mesh = mx.meshes[0]
    _, p64 = wall_pressure(mesh, U64)

    def err(U):
        _, p = wall_pressure(mesh, U)
        return {"U_maxabs": float(np.max(np.abs(U - U64))),
                "U_relL2": float(np.linalg.norm(U - U64) / np.linalg.norm(U64)),
                "wall_p_maxabs": float(np.max(np.abs(p - p64))),
                "wall_p_relL2": float(np.linalg.norm(p - p64) / np.linalg.norm(p64))}

    rep["error_mixed_at_floor"] = err(mx.switch_solutions[0])
    rep["error_mixed_then_fp64"] = err(mx.solutions()[0])
    return rep


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="steady bump-channel study")
    ap.add_argument("--cases", type=int, default=8)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cells-lo", type=float, default=2e4)
    ap.add_argument("--cells-hi", type=float, default=3e5)
    ap.add_argument("--drop", type=float, default=1e-8)
    ap.add_argument("--max-steps", type=int, default=60000)
    ap.add_argument("--check-every", type=int, default=200)
    ap.add_argument("--train-cases", type=int, default=0)
    ap.add_argument("--train-cells", type=float, default=2e5)
    ap.add_argument("--per-case", type=int, default=4000)
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--plain-lts", choices=["no", "yes", "both"], default="no",
                    help="plain solves with local time steps too (both) or only (yes)")
    ap.add_argument("--net", default=None, help="a saved network instead of training one")
    ap.add_argument("--mach", type=float, nargs=2, default=(0.3, 0.8))
    ap.add_argument("--partial", default=None, help="write the report after each phase to this file")
    ap.add_argument("--c3", action="store_true", help="the 2M-tet mixed-precision channel case only")
    ap.add_argument("--c5-scaled", action="store_true",
                    help="unseen cases at the learned workload's inference size, freestream vs learned start (--net)")
    ap.add_argument("--c5-seed", type=int, default=11, help="case seed of --c5-scaled")
    ap.add_argument("--c5-no-freestream", action="store_true", help="--c5-scaled: the warm-started solve only")
    ap.add_argument("--train-chunk", default=None, metavar="PATH",
                    help="only solve one chunk of training cases (train-cases at train-cells, case seed "
                         "--seed) and save its samples")
    ap.add_argument("--train-from", nargs="+", default=None, metavar="PATH",
                    help="train on saved chunks (with --train-save PATH for the network)")
    ap.add_argument("--train-save", default=None, metavar="PATH",
                    help="only train the warm-start network (train-cases at train-cells) and save it")
    a = ap.parse_args(argv)
    if a.c5_scaled:
        print(json.dumps(c5_scaled(a.net, n_cases=a.cases, cells=a.cells_hi, seed=a.c5_seed, drop=a.drop,
                                   max_steps=a.max_steps, check_every=a.check_every,
                                   freestream=not a.c5_no_freestream)))
        return
    if a.train_chunk:
        x, y, info = solve_training_chunk(a.train_cases, a.train_cells, a.seed, a.per_case,
                                          max_steps=a.max_steps, check_every=a.check_every)
        save_training_chunk(a.train_chunk, x, y, info)
        print(json.dumps(info))
        return
    if a.train_from:
        if not a.train_save:
            ap.error("--train-from needs --train-save PATH for the network")
        net, info = train_from_chunks(a.train_from, epochs=a.epochs)
        save_network(net, a.train_save, info)
        print(json.dumps(info))
        return
    if a.train_save:
        net, info = train_from_solves(a.train_cases, a.train_cells, a.seed + 1000, a.per_case,
                                      max_steps=a.max_steps, check_every=a.check_every, epochs=a.epochs)
        save_network(net, a.train_save, info)
        print(json.dumps(info))
        return
    if a.c3:
        print(json.dumps(c3(a.drop, a.max_steps, a.check_every)))
        return
    rep = study(a.cases, a.seed, a.cells_lo, a.cells_hi, a.drop, a.max_steps, a.check_every, a.train_cases,
                a.train_cells, a.per_case, a.epochs, plain=not a.no_plain,
                plain_lts={"no": False, "yes": True, "both": "both"}[a.plain_lts], partial=a.partial,
                train_max_steps=a.max_steps, net_path=a.net, mach=tuple(a.mach))
    print(json.dumps(rep))


if __name__ == "__main__":
    main()
