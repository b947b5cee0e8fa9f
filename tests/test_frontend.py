"""Sim-directory front end on the CPU: input.yml parsing (against the
reference's own sim-directory inputs, tests/golden/sim_inputs), SU2 mesh
reading and writing, the grid validity report and the VTK writer."""
import glob
import os

import numpy as np
import pytest

import netflux_amd as nf
from netflux_amd import _abi as A
from netflux_amd import frontend as fe

GOLD = os.path.join(os.path.dirname(__file__), "golden", "sim_inputs")


def test_reference_inputs_parse():
    files = sorted(glob.glob(os.path.join(GOLD, "*.yml")))
    assert len(files) == 9
    for f in files:
        cfg, opts = fe.parse_input(f)
        assert cfg.density_fs == 1.2 and cfg.velocity_fs == (600.0, 0.0, 0.0) and cfg.pressure_fs == 100000.0
        assert opts["patches"] and all(v == "SlipWall" for v in opts["patches"].values())
    cfg, opts = fe.parse_input(os.path.join(GOLD, "very_fine_brick_domain_first.yml"))
    assert (cfg.time_scheme, cfg.spatial_order, cfg.limiter, cfg.CFL) == ("TVD_RK3", "First", "NONE", 0.9)
    assert (opts["n_timesteps"], opts["write_stride"], opts["patches"]) == (500, 5, {"all_walls": "SlipWall"})
    cfg, opts = fe.parse_input(os.path.join(GOLD, "small_brick_domain.yml"))
    assert (cfg.time_scheme, cfg.spatial_order, cfg.limiter) == ("ExplicitEuler", "Second", "Barth")
    assert cfg.check_physical_validity and opts["check_grid_validity"] and opts["write_stride"] == 1
    assert list(opts["patches"]) == ["inlet", "outlet", "sides"]
    _, opts = fe.parse_input(os.path.join(GOLD, "fine_box_domain.yml"))
    assert opts["write_vtk_debug"]


def _yml(tmp_path, text):
    p = tmp_path / "input.yml"
    p.write_text(text)
    return str(p)


BASE = """solver: Euler
n_timesteps: 3
CFL: 0.5
time_scheme: TVD_RK3
spatial_order: Second
grad_scheme: GreenGauss
inviscid_flux_scheme: Rusanov
limiter: Barth
initial_cond: Freestream
patches:
  - a: SlipWall
"""


def test_input_defaults_and_errors(tmp_path):
    cfg, opts = fe.parse_input(_yml(tmp_path, BASE))
    # standard-air freestream defaults (include/Utilities.hpp:100-107)
    assert (cfg.density_fs, cfg.velocity_fs, cfg.pressure_fs) == (1.225, (0.0, 0.0, 0.0), 101325.0)
    assert opts["write_stride"] == 1 and not cfg.check_physical_validity
    cfg, _ = fe.parse_input(_yml(tmp_path, BASE.replace("Rusanov", "HLLC")))
    assert cfg.inviscid_flux_scheme == "HLLC"
    bad = [
        (BASE.replace("Rusanov", "Roe"), "Roe"),
        (BASE.replace("limiter: Barth", "limiter: Venkat"), "Venkat"),
        (BASE.replace("CFL: 0.5\n", ""), "CFL"),
        (BASE + "  - a: FarField\n", "Duplicate patch"),
        (BASE.replace("a: SlipWall", "a: Inflow"), "Inflow"),
        (BASE.replace("patches:\n  - a: SlipWall\n", "patches: x\n"), "sequence"),
    ]
    for text, msg in bad:
        with pytest.raises(fe.SimDirError, match=msg):
            fe.parse_input(_yml(tmp_path, text))
    # first order: limiter and gradient scheme are not read
    cfg, _ = fe.parse_input(_yml(tmp_path, BASE.replace("Second", "First").replace("limiter: Barth\n", "")
                                 .replace("grad_scheme: GreenGauss\n", "")))
    assert cfg.limiter == "NONE"


def test_find_mesh(tmp_path):
    with pytest.raises(fe.SimDirError, match="valid mesh name"):
        fe.find_mesh(str(tmp_path))
    (tmp_path / "mesh.nf").write_text("")
    with pytest.raises(fe.SimDirError, match="stub"):
        fe.find_mesh(str(tmp_path))
    (tmp_path / "mesh.su2").write_text("")
    with pytest.raises(fe.SimDirError, match="More than one"):
        fe.find_mesh(str(tmp_path))


def three_patch_box(n=4, jitter=0.1, seed=3):
    """Box whose six sides form the patches inlet (x-), outlet (x+), sides."""
    m = nf.Mesh.box(n, jitter=jitter, seed=seed)
    nodes, cells, pn, pf, _ = m.arrays()
    first = np.concatenate([[0], np.cumsum(pn.astype(np.int64))])
    faces = [pf[first[0]:first[1]], pf[first[1]:first[2]], pf[first[2]:first[6]]]
    return nf.Mesh.from_arrays(nodes, cells, faces, ["SlipWall"] * 3)


def test_su2_round_trip(tmp_path):
    m = three_patch_box()
    path = tmp_path / "mesh.su2"
    m.write_su2(path, ["inlet", "outlet", "sides"])
    r = nf.Mesh.read_su2(path, {"inlet": "SlipWall", "outlet": "FarField", "sides": "NoSlipWall"})
    a, b = m.arrays(), r.arrays()
    for x, y in zip(a[:4], b[:4]):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(b[4], [A.NF_BC_SLIP_WALL, A.NF_BC_FARFIELD, A.NF_BC_NOSLIP_WALL])
    assert r.patch_names == ["inlet", "outlet", "sides"]
    # the file holds tetrahedra in Salome order: nodes 1 and 2 swapped
    lines = path.read_text().split("\n")
    first = [int(v) for v in lines[2].split()]
    assert first[0] == 10 and first[1:5] == [a[1][0][0], a[1][0][2], a[1][0][1], a[1][0][3]]


SU2_ONE_TET = """% one tetrahedron
NDIME=3
NELEM= 1
10 0 2 1 3 0
NPOIN= 4
0 0 0 0
1 0 0 1
0 1 0 2
0 0 1 3
NMARK= 2
MARKER_TAG= wall
MARKER_ELEMS= 3
5 0 1 3
5 0 2 3
5 1 2 3
MARKER_TAG=bottom
MARKER_ELEMS= 1
5 0 1 2
"""


def test_su2_read_conventions_and_errors(tmp_path):
    p = tmp_path / "m.su2"
    p.write_text(SU2_ONE_TET)
    m = nf.Mesh.read_su2(p, {"wall": "SlipWall", "bottom": "FarField"})
    nodes, cells, pn, pf, pbc = m.arrays()
    np.testing.assert_array_equal(cells, [[0, 1, 2, 3]])      # Salome -> VTK
    np.testing.assert_array_equal(pn, [3, 1])
    np.testing.assert_array_equal(pf[-1], [0, 1, 2])           # triangles as read
    assert m.patch_names == ["wall", "bottom"]
    # in VTK order this tetrahedron has positive volume in the reference's formula
    rep = fe.grid_validity_report(m)
    assert len(rep["negative_volume"]) == 0
    errors = [
        ({"wall": "SlipWall"}, SU2_ONE_TET, "bottom' is not named"),
        ({"wall": "SlipWall", "bottom": "FarField"}, SU2_ONE_TET.replace("10 0 2 1 3 0", "3 0 2 1 3 0"),
         r"Illegal su2 volume element identifier \(3\)"),
        ({"wall": "SlipWall", "bottom": "FarField"}, SU2_ONE_TET.replace("10 0 2 1 3 0", "5 0 2 1 0"),
         "Face element of type 5"),
        ({"wall": "SlipWall", "bottom": "FarField"}, SU2_ONE_TET.replace("5 0 1 2\n", "10 0 1 2 3\n"),
         "Volume element of type 10"),
        ({"wall": "SlipWall", "bottom": "FarField"}, SU2_ONE_TET.replace("NDIME=3", "NDIME= 2"), "NDIME=3"),
        ({"wall": "SlipWall", "bottom": "FarField"}, SU2_ONE_TET.replace("5 0 1 2\n", "5 0 1 9\n"), "out of range"),
        ({"wall": "SlipWall", "bottom": "FarField"}, SU2_ONE_TET.replace("NPOIN", "NPOINTS"), "NPOIN="),
    ]
    for bc, text, msg in errors:
        p.write_text(text)
        with pytest.raises(RuntimeError, match=msg):
            nf.Mesh.read_su2(p, bc)
    with pytest.raises(RuntimeError, match="open"):
        nf.Mesh.read_su2(tmp_path / "absent.su2", {})


def test_grid_validity_report_flags_inverted_element():
    m = nf.Mesh.box(2)
    nodes, cells, pn, pf, pbc = m.arrays()
    assert len(fe.grid_validity_report(m)["negative_volume"]) == 0
    cells = cells.copy()
    cells[5, [1, 2]] = cells[5, [2, 1]]
    first = np.concatenate([[0], np.cumsum(pn.astype(np.int64))])
    bad = nf.Mesh.from_arrays(nodes, cells, [pf[first[i]:first[i + 1]] for i in range(len(pn))], list(pbc))
    np.testing.assert_array_equal(fe.grid_validity_report(bad)["negative_volume"], [5])


def test_vtk_writer(tmp_path):
    m = nf.Mesh.box(2, jitter=0.1)
    N = m.n_cells
    rng = np.random.default_rng(0)
    U = np.column_stack([1 + rng.random(N), rng.standard_normal((N, 3)), 5 + rng.random(N)])
    path = tmp_path / "n=0.vtk"
    fe.write_vtk(path, m, U)
    text = path.read_text()
    nodes, cells, *_ = m.arrays()
    head = text.split("\n")
    assert head[:5] == ["# vtk DataFile Version 3.0", "NetFlux", "ASCII", "", "DATASET UNSTRUCTURED_GRID"]
    assert f"POINTS {len(nodes)} double" in head and f"CELLS {N} {5 * N}" in head and f"CELL_TYPES {N}" in head
    i = head.index(f"CELLS {N} {5 * N}")
    np.testing.assert_array_equal(np.loadtxt(head[i + 1:i + 1 + N], dtype=np.int64),
                                  np.column_stack([np.full(N, 4), cells]))
    d = fe.read_vtk_cell_data(path)
    p = 0.4 * (U[:, 4] - 0.5 * np.sum(U[:, 1:4] ** 2, axis=1) / U[:, 0])
    # 6 significant digits, the default precision of a C++ stream
    np.testing.assert_allclose(d["density"], U[:, 0], rtol=5e-6)
    np.testing.assert_allclose(d["velocity"], U[:, 1:4] / U[:, :1], rtol=5e-6, atol=1e-12)
    np.testing.assert_allclose(d["pressure"], p, rtol=5e-6)
    fe.write_vtk(path, m, U, full_precision=True, limiter=np.ones((N, 5)), gradient=np.arange(N * 15.0).reshape(N, 15))
    d = fe.read_vtk_cell_data(path)
    np.testing.assert_array_equal(d["density"], U[:, 0])
    assert sorted(k for k in d if k.startswith("limiter")) == [f"limiter_EQ{k}" for k in range(5)]
    np.testing.assert_array_equal(d["gradient_EQ1"][0], [1.0, 6.0, 11.0])
