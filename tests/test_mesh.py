"""Host mesh pipeline (product) against the oracle's FV_Grid restatement and
against closed forms."""
import numpy as np
import pytest

import netflux_amd as nf
from tests.cases import sod_config
from tests.oracle_bind import Oracle


def _compare(mesh):
    g = mesh.fv_grid()
    o = Oracle.from_mesh(mesh, sod_config()).grid()
    for k in ("ij", "S", "r_i", "r_j", "volume", "centroid", "patch_first", "patch_nfaces", "delta_S"):
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    return g


@pytest.mark.parametrize("n,jitter,bump", [(4, 0.0, 0.0), (5, 0.15, 0.0), (6, 0.1, 0.05), ((7, 3, 4), 0.12, 0.0)])
def test_fv_grid_bit_identical_to_oracle(n, jitter, bump):
    patches = ("FarField", "FarField") + ("SlipWall",) * 4 if bump else ("SlipWall",) * 6
    mesh = nf.Mesh.box(n, jitter=jitter, seed=99, bump_h=bump, patches=patches)
    _compare(mesh)


def test_sizes_match_kuhn_counts():
    # 16^3 hexes x 6 tets: N=24,576, F=50,688, G=3,072
    g = nf.Mesh.box(16).fv_grid()
    assert g["sizes"] == (24576, 27648, 47616, 50688, 6)


def test_closed_forms_unit_cube():
    mesh = nf.Mesh.box(5, jitter=0.15, seed=4)
    g = mesh.fv_grid()
    N = mesh.n_cells
    F = g["sizes"][3]
    assert abs(g["volume"][:N].sum() - 1.0) < 1e-14
    assert (g["volume"][:N] > 0).all()
    assert (g["volume"][N:] == 0).all()
    # sum of outward normals of every closed cell vanishes
    acc = np.zeros((N, 3))
    ij = g["ij"].astype(np.int64)
    np.add.at(acc, ij[:, 0], g["S"])
    inner = ij[:, 1] < N
    np.add.at(acc, ij[inner, 1], -g["S"][inner])
    assert np.abs(acc).max() < 1e-15
    # S oriented from i towards j
    c = g["centroid"]
    assert (np.einsum("fd,fd->f", g["S"], c[ij[:, 1]] - c[ij[:, 0]]) > 0).all()
    # interior faces sorted by (i, j) with i < j
    Fi = g["sizes"][2]
    assert (ij[:Fi, 0] < ij[:Fi, 1]).all()
    key = ij[:Fi, 0] * (1 << 32) + ij[:Fi, 1]
    assert (np.diff(key) > 0).all()
    # total boundary area = 6, ghost centroid = reflection of i through the face
    assert abs(np.linalg.norm(g["S"][Fi:], axis=1).sum() - 6.0) < 1e-12
    fc = c[ij[Fi:, 0]] + g["r_i"][Fi:]
    np.testing.assert_allclose(c[ij[Fi:, 1]], 2 * fc - c[ij[Fi:, 0]], atol=1e-15)


def test_rejects_uncovered_boundary():
    mesh = nf.Mesh.box(2)
    nodes, cells, pn, pf, pbc = mesh.arrays()
    with pytest.raises(RuntimeError):
        nf.Mesh.from_arrays(nodes, cells, [pf[: pn[0]]], ["SlipWall"]).fv_grid()


def test_from_arrays_roundtrip_single_tet():
    nodes = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    cells = np.array([[0, 1, 2, 3]], np.uint32)
    faces = [[[0, 1, 2]], [[0, 1, 3]], [[0, 2, 3]], [[1, 2, 3]]]
    mesh = nf.Mesh.from_arrays(nodes, cells, faces, ["SlipWall"] * 4)
    g = _compare(mesh)
    assert abs(g["volume"][0] - 1 / 6) < 1e-16
    np.testing.assert_allclose(g["centroid"][0], [0.25, 0.25, 0.25])
