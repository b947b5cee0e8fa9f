"""Host logic of the steady bump-channel study (no GPU): case sampling,
mesh dimensions, freestream state and floor-cell selection."""
import numpy as np

import netflux_amd as nf
from netflux_amd import steady as S
from tests.cases import prim_to_cons


def test_dims_follow_5_2_1_and_target():
    assert S.dims_for(1966080) == (160, 64, 32)  # the 2M-tet channel: 6*160*64*32
    for target in (2e4, 2e5, 1.5e6):
        nx, ny, nz = S.dims_for(target)
        assert nx == 5 * nz and ny == 2 * nz
        k = nz
        # nearest of the neighbouring k
        errs = [abs(60 * q ** 3 - target) for q in (k - 1, k, k + 1)]
        assert errs[1] == min(errs)


def test_sampled_cases_are_reproducible_and_in_range():
    a = S.sample_cases(64, 7, 2e4, 1.5e6)
    b = S.sample_cases(64, 7, 2e4, 1.5e6)
    assert a == b
    assert all(0.3 <= c.mach <= 0.8 and 0.02 <= c.h <= 0.1 for c in a)
    cells = np.array([c.n_cells for c in a])
    assert cells.min() >= 6 * 10 * 2 ** 3 and cells.max() <= 2.2e6
    # log-uniform: the median is near the geometric mean of the range
    assert 0.5 * np.sqrt(2e4 * 1.5e6) < np.median(cells) < 2 * np.sqrt(2e4 * 1.5e6)


def test_freestream_state_and_prims():
    cfg = S.config_of(S.Case(10, 4, 2, 0.6, 0.05))
    U = S.freestream_state(cfg, 3)
    np.testing.assert_allclose(U, prim_to_cons(np.tile(cfg.primvars_inf(), (3, 1))), rtol=1e-15)
    np.testing.assert_allclose(S.cons_to_prim(U), np.tile(cfg.primvars_inf(), (3, 1)), rtol=1e-14)
    assert cfg.rk3_zero_old_first_step is False and cfg.local_time_step is True


def test_floor_cells_touch_the_bump():
    case = S.Case(10, 4, 2, 0.5, 0.08)
    mesh = S.mesh_of(case)
    U = S.freestream_state(S.config_of(case), mesh.n_cells)
    owner, p = S.wall_pressure(mesh, U)
    assert len(owner) == 2 * case.nx * case.nz  # two floor triangles per floor quad
    g = mesh.fv_grid()
    yb = 0.08 * np.exp(-((g["centroid"][owner, 0] - 1.5) ** 2) / (2 * 0.2 ** 2))
    assert np.all(g["centroid"][owner, 1] - yb < 1.0 / case.ny)
    np.testing.assert_allclose(p, 1 / 1.4, rtol=1e-14)


def test_training_chunk_round_trip(tmp_path):
    """A saved chunk of training samples loads back as plain arrays (no
    pickles) with its description."""
    import json
    rng = np.random.default_rng(0)
    x = rng.random((10, 6)).astype(np.float32)
    y = rng.random((10, 5)).astype(np.float32)
    info = {"cases": 2, "case_seed": 5, "samples": 10, "solves": {"converged": 2}}
    p = tmp_path / "chunk.npz"
    S.save_training_chunk(p, x, y, info)
    with np.load(p, allow_pickle=False) as z:
        np.testing.assert_array_equal(z["x"], x)
        np.testing.assert_array_equal(z["y"], y)
        assert json.loads(str(z["meta"])) == info
