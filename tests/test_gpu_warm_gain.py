"""What the trained network gains: the warm-start network is trained by the
library's BF16-MFMA kernels on steady solutions the plain solver produced for
other cases, then unseen cases are solved to a fixed residual drop from the
freestream and from the network's prediction in the same mode (batched FP64,
the same local time steps, the same stop rule, the drop measured against the
freestream residual of each case).  The hook is the reference's only initial
condition, the freestream (src/SolverData.cpp:27-34)."""
import numpy as np
import pytest

from netflux_amd import steady as SD

pytestmark = pytest.mark.gpu


def test_trained_network_saves_steps():
    print("[case] train on 10 small channels, solve 3 unseen ones to a 1e-6 drop", flush=True)
    net, info = SD.train_from_solves(train_cases=10, train_cells=6000, seed=101, per_case=3000,
                                     train_drop=1e-6, max_steps=40000, check_every=100, epochs=60)
    print("training:", info["solves"], info["train"], flush=True)
    assert info["solves"]["converged"] >= 8
    assert info["train"]["loss_last"] < 0.05 * info["train"]["loss_first"]
    cases = SD.sample_cases(3, seed=202, cells_lo=6000, cells_hi=6000)
    run = SD.BatchRun(cases, precision="fp64")
    st_fs, _ = run.solve(1e-6, 40000, 100)
    U_fs = run.solutions()
    st_w, _ = run.solve(1e-6, 40000, 100, net=net)
    U_w = run.solutions()
    n_fs = [s["steps_done"] for s in st_fs]
    n_w = [s["steps_done"] for s in st_w]
    print("steps from the freestream:", n_fs, "from the warm start:", n_w, flush=True)
    assert all(s["converged"] for s in st_fs) and all(s["converged"] for s in st_w)
    assert not any(s["fell_back"] for s in st_w)
    assert sum(n_w) < 0.85 * sum(n_fs)
    assert all(w <= 1.02 * f for w, f in zip(n_w, n_fs))
    # both reach the same steady state (to the accuracy a 1e-6 drop leaves)
    for a, b in zip(U_fs, U_w):
        assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-4


def test_committed_network_loads_and_predicts_physical_states():
    """The network bench.py uses for its time-to-1e-8 comparison: it loads
    without unpickling, and its prediction on a case of the bench's range is
    a physical state (no freestream fallback)."""
    import os
    path = os.path.join(os.path.dirname(SD.__file__), "data", "warm_start_bump.npz")
    net, meta = SD.load_network(path)
    assert meta["samples"] > 0
    case = SD.sample_cases(1, seed=7, cells_lo=2e4, cells_hi=2e4)[0]
    run = SD.BatchRun([case], precision="fp64")
    assert run.batch.warm_start(0, net, (case.mach, case.h)) is False
    V = SD.cons_to_prim(run.batch.get_solution(0))
    assert (V[:, 0] > 0).all() and (V[:, 4] > 0).all()


def test_training_in_chunks(tmp_path):
    """Training data solved in chunks (one GPU job each for large sets) and
    a network trained on the saved chunks: every converged case contributes
    its samples, the loss falls, and predictions are physical."""
    print("[case] two chunks of 4 channels at 6000 cells, trained from the saved chunks", flush=True)
    paths = []
    for seed in (301, 302):
        x, y, info = SD.solve_training_chunk(4, 6000, seed, per_case=2000, max_steps=40000, check_every=200)
        assert x.shape == (2000 * info["solves"]["converged"], 6) and y.shape == (len(x), 5)
        assert info["solves"]["converged"] >= 3
        p = tmp_path / f"chunk_{seed}.npz"
        SD.save_training_chunk(p, x, y, info)
        paths.append(p)
    net, info = SD.train_from_chunks(paths, epochs=40)
    print("training:", info["samples"], info["train"], flush=True)
    assert info["cases"] == 8 and info["samples"] == 2000 * info["converged"]
    assert info["train"]["loss_last"] < 0.1 * info["train"]["loss_first"]
    case = SD.sample_cases(1, seed=303, cells_lo=6000, cells_hi=6000)[0]
    run = SD.BatchRun([case], precision="fp64")
    assert not run.batch.warm_start(0, net, (case.mach, case.h))  # no fallback: a physical prediction
    V = SD.cons_to_prim(run.batch.get_solution(0))
    assert np.all(V[:, 0] > 0) and np.all(V[:, 4] > 0)
