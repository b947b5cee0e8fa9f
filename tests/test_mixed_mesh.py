"""Mixed-element meshes on the CPU: the library's face graph against the
oracle's (two independent restatements of FV_Grid and the element geometry
of src/geometry/GeometryUtilities.cpp:242-441, compared bit for bit), pinned
by closed forms (volumes and centroids of undistorted and affinely sheared
hexahedra, wedges and pyramids), closed-cell normal sums, freestream
preservation of the oracle's solver, SU2 round trips and the VTK writer."""
import numpy as np
import pytest

import netflux_amd as nf
from netflux_amd import frontend as fe
from tests.mixed_meshes import mixed_box
from tests.oracle_bind import Oracle, prim_to_cons

GRID_KEYS = ["ij", "S", "r_i", "r_j", "volume", "centroid", "patch_first", "patch_nfaces", "delta_S"]


def _cfg(**kw):
    base = dict(density_fs=1.0, velocity_fs=(0.5, 0.1, -0.05), pressure_fs=1 / 1.4)
    base.update(kw)
    return nf.Config(**base)


@pytest.mark.parametrize("kinds,jitter,seed", [(("hex", "wedge", "pyr"), 0.12, 3), (("pyr", "hex"), 0.2, 11),
                                               (("wedge",), 0.15, 5)])
def test_mixed_grid_matches_oracle(kinds, jitter, seed):
    m = mixed_box((4, 3, 3), kinds=kinds, jitter=jitter, seed=seed)
    assert m.is_mixed
    g = m.fv_grid()
    go = Oracle.from_mesh(m, _cfg()).grid()
    for k in GRID_KEYS:
        np.testing.assert_array_equal(g[k], go[k], err_msg=k)
    N, NT, Fint, F, P = g["sizes"]
    # every face is oriented from i to j, ghosts follow the cells
    assert np.all(g["ij"][:, 0] < g["ij"][:, 1]) and np.all(g["ij"][Fint:, 1] >= N)
    assert np.all(g["volume"][:N] > 0) and np.all(g["volume"][N:] == 0)


def test_element_closed_forms():
    """Undistorted blocks: hexahedron h^3, wedge h^3/2, pyramid h^3/6 with
    the textbook centroids; the same after an affine shear (volume times
    det A, centroid mapped), which keeps every face planar."""
    n = (3, 3, 2)
    L = (1.5, 1.2, 0.8)
    h = np.prod(np.array(L) / np.array(n))
    m = mixed_box(n, kinds=("hex", "wedge", "pyr"), jitter=0.0, lengths=L)
    ct, co, cn, *_ = m.elements()
    g = m.fv_grid()
    N = len(ct)
    vol, cen = g["volume"][:N], g["centroid"][:N]
    X = np.ctypeslib.as_array(m.desc.nodes, shape=(m.desc.n_nodes, 3))
    expect = {12: h, 13: h / 2, 14: h / 6}
    for e in range(N):
        P = X[cn[co[e]:co[e + 1]].astype(np.int64)]
        assert vol[e] == pytest.approx(expect[int(ct[e])], rel=1e-13)
        if ct[e] == 14:   # a quarter of the way from the base centroid to the apex
            ref = 0.75 * P[:4].mean(axis=0) + 0.25 * P[4]
        else:             # hexahedron and right prism: the vertex mean
            ref = P.mean(axis=0)
        np.testing.assert_allclose(cen[e], ref, rtol=0, atol=1e-14)
    assert vol.sum() == pytest.approx(np.prod(L), rel=1e-13)
    # the same mesh sheared by A: volumes scale by det A, centroids map by A
    A = np.array([[1.0, 0.3, 0.1], [0.0, 0.9, -0.2], [0.05, 0.0, 1.1]])
    Xs = X @ A.T
    ms = nf.Mesh.from_elements(Xs, [cn[co[e]:co[e + 1]] for e in range(N)],
                               *_patches(m))
    gs = ms.fv_grid()
    np.testing.assert_allclose(gs["volume"][:N], vol * np.linalg.det(A), rtol=1e-12)
    np.testing.assert_allclose(gs["centroid"][:N], cen @ A.T, rtol=0, atol=1e-13)


def _patches(m):
    ct, co, cn, pn, ft, fo, fn, pbc = m.elements()
    faces, e = [], 0
    for k in pn:
        faces.append([fn[fo[q]:fo[q + 1]] for q in range(e, e + k)])
        e += k
    return faces, list(pbc)


def test_closed_cell_normal_sums():
    """The outward face vectors of every cell (S on its i side, -S on its j
    side) sum to zero: each quadrilateral is two triangles on both sides."""
    m = mixed_box((4, 4, 3), kinds=("hex", "wedge", "pyr"), jitter=0.2, seed=9)
    g = m.fv_grid()
    N = g["sizes"][0]
    acc = np.zeros((g["sizes"][1], 3))
    np.add.at(acc, g["ij"][:, 0].astype(np.int64), g["S"])
    np.add.at(acc, g["ij"][:, 1].astype(np.int64), -g["S"])
    scale = np.abs(g["S"]).max()
    assert np.abs(acc[:N]).max() < 1e-14 * scale


def test_all_tetrahedra_stored_in_tet_layout():
    m = mixed_box((3, 2, 2), kinds=("tet",), jitter=0.1, seed=4)
    assert not m.is_mixed
    nodes, cells, pn, pf, pbc = m.arrays()
    assert cells.shape[1] == 4 and pf.shape[1] == 3
    g = m.fv_grid()
    go = Oracle.from_mesh(m, _cfg()).grid()
    for k in GRID_KEYS:
        np.testing.assert_array_equal(g[k], go[k], err_msg=k)


@pytest.mark.parametrize("order", ["First", "Second"])
def test_oracle_freestream_preserved_on_mixed_mesh(order):
    """A uniform state with farfield boundaries has a zero flux balance (up
    to rounding) on any closed cell: the oracle's face loops see every face."""
    m = mixed_box((4, 3, 3), kinds=("hex", "wedge", "pyr"), jitter=0.15, seed=2, patches=("FarField",) * 6)
    cfg = _cfg(spatial_order=order, limiter="Barth" if order == "Second" else "NONE", time_scheme="ExplicitEuler")
    o = Oracle.from_mesh(m, cfg)
    o.set_freestream()
    dt = o.calc_timestep()
    o.step(dt)
    U = o.get("U")
    U0 = np.tile(prim_to_cons(np.array(cfg.primvars_inf(), float)), (o.N, 1))
    np.testing.assert_allclose(U, U0, rtol=0, atol=1e-13)


def test_su2_mixed_round_trip(tmp_path):
    m = mixed_box((3, 3, 2), kinds=("hex", "wedge", "pyr"), jitter=0.1, seed=6)
    path = tmp_path / "mesh.su2"
    names = ["xm", "xp", "ym", "yp", "zm", "zp"]
    m.write_su2(path, names)
    r = nf.Mesh.read_su2(path, {n: "SlipWall" for n in names})
    assert r.is_mixed
    for a, b in zip(m.elements(), r.elements()):
        np.testing.assert_array_equal(a, b)
    a, b = m.fv_grid(), r.fv_grid()
    for k in GRID_KEYS:
        np.testing.assert_array_equal(a[k], b[k])
    # hexahedra are written in Salome order: VTK nodes 1<->3 and 5<->7
    ct, co, cn, *_ = m.elements()
    e = int(np.nonzero(ct == 12)[0][0])
    v = cn[co[e]:co[e + 1]]
    line = next(l for l in path.read_text().split("\n")[2:] if l.split()[-1] == str(e) and l.startswith("12 "))
    assert [int(x) for x in line.split()[1:9]] == [v[0], v[3], v[2], v[1], v[4], v[7], v[6], v[5]]


def test_mixed_mesh_errors():
    nodes = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 0.5, 1.0]], float)
    pyr = [0, 1, 2, 3, 4]
    sides = [[0, 1, 4], [1, 2, 4], [2, 3, 4], [3, 0, 4]]
    m = nf.Mesh.from_elements(nodes, [pyr], [[[0, 3, 2, 1]], sides], ["SlipWall", "FarField"])
    g = m.fv_grid()
    assert g["sizes"] == (1, 6, 0, 5, 2)
    assert g["volume"][0] == pytest.approx(1 / 3, rel=1e-15)
    bad = nf.Mesh.from_elements(nodes, [pyr], [sides], ["SlipWall"])   # base not covered
    with pytest.raises(RuntimeError, match="not covered by any patch"):
        bad.fv_grid()


def test_vtk_and_validity_mixed(tmp_path):
    m = mixed_box((3, 2, 2), kinds=("hex", "wedge", "pyr"), jitter=0.1, seed=8)
    rep = fe.grid_validity_report(m)
    assert len(rep["negative_volume"]) == 0
    ct, co, cn, *_ = m.elements()
    N = len(ct)
    U = np.tile([1.0, 0.2, 0.0, 0.1, 2.5], (N, 1))
    path = tmp_path / "n=0.vtk"
    fe.write_vtk(path, m, U)
    lines = path.read_text().split("\n")
    assert f"CELLS {N} {N + len(cn)}" in lines
    i = lines.index(f"CELL_TYPES {N}")
    np.testing.assert_array_equal([int(x) for x in lines[i + 1:i + 1 + N]], ct)
    np.testing.assert_allclose(fe.read_vtk_cell_data(path)["density"], 1.0)


def test_hex_box():
    m = nf.Mesh.hex_box((4, 3, 2), lengths=(2.0, 1.5, 1.0))
    g = m.fv_grid()
    N, NT, Fint, F, P = g["sizes"]
    assert (N, P) == (24, 6) and NT - N == F - Fint == 2 * (4 * 3 + 4 * 2 + 3 * 2)
    np.testing.assert_allclose(g["volume"][:N], 0.125, rtol=1e-14)
    np.testing.assert_array_equal(g["patch_nfaces"], [6, 6, 8, 8, 12, 12])
    j = nf.Mesh.hex_box(5, jitter=0.2, seed=3, patches=("FarField",) * 6)
    gj, go = j.fv_grid(), Oracle.from_mesh(j, _cfg()).grid()
    for k in GRID_KEYS:
        np.testing.assert_array_equal(gj[k], go[k], err_msg=k)
    # non-planar faces: the six-pyramid volumes of neighbours do not tile the
    # box exactly (each hexahedron splits its faces about its own centre)
    assert np.all(gj["volume"][:125] > 0) and gj["volume"][:125].sum() == pytest.approx(1.0, rel=1e-3)


def test_mixed_descriptor_validation():
    """nf_mesh_from_desc rejects inconsistent element arrays with a message."""
    import ctypes as C
    from netflux_amd import _abi as A
    nodes = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 0.5, 1.0]], float)
    conn = np.array([0, 1, 2, 3, 4], np.uint32)
    fconn = np.array([0, 3, 2, 1, 0, 1, 4, 1, 2, 4, 2, 3, 4, 3, 0, 4], np.uint32)
    pn = np.array([5], np.uint32)
    pbc = np.array([1], np.int32)

    def make(ctype, coff, ftype, foff):
        keep = [np.ascontiguousarray(a) for a in (ctype, coff, ftype, foff)]
        d = A.MeshDesc(len(nodes), nodes.ctypes.data_as(A.f64p), len(keep[0]), conn.ctypes.data_as(A.u32p), 1,
                       pn.ctypes.data_as(A.u32p), fconn.ctypes.data_as(A.u32p), pbc.ctypes.data_as(A.i32p),
                       keep[0].ctypes.data_as(A.u8p), keep[1].ctypes.data_as(A.u32p),
                       keep[2].ctypes.data_as(A.u8p), keep[3].ctypes.data_as(A.u32p))
        h = C.c_void_p()
        rc = A.lib().nf_mesh_from_desc(C.byref(d), C.byref(h))
        if rc == A.NF_OK:
            A.lib().nf_mesh_destroy(h)
        return rc, A.lib().nf_mesh_last_error().decode()

    ok = (np.array([14], np.uint8), np.array([0, 5], np.uint32), np.array([9, 5, 5, 5, 5], np.uint8),
          np.array([0, 4, 7, 10, 13, 16], np.uint32))
    assert make(*ok)[0] == A.NF_OK
    bad = [
        ((np.array([12], np.uint8),) + ok[1:], "node count does not match"),           # hexahedron with 5 nodes
        ((np.array([5], np.uint8),) + ok[1:], "not a volume element"),                 # a triangle as a cell
        (ok[:2] + (np.array([9, 5, 5, 5, 10], np.uint8), ok[3]), "not a face type"),   # a tetrahedron as a face
        (ok[:3] + (np.array([0, 4, 7, 10, 13, 15], np.uint32),), "node count does not match"),
    ]
    for args, msg in bad:
        rc, err = make(*args)
        assert rc == A.NF_E_MESH and msg in err, (rc, err)
