"""Python mirror of NetFlux's Config / EulerSolver / SolverData / Driver
surface over the C ABI of libnetflux.so.

Same names, argument meaning and error behaviour as the reference:
  Config               include/Config.hpp:8-175 (keys of src/ConfigParser.cpp:38-83)
  EulerSolver          include/Solver.hpp:9-79  (step, calc_timestep, get_solver_data)
  SolverData           include/SolverData.hpp:25-73 (get_solution, get_primvars, ...)
  Driver.solve         src/Driver.cpp:31-57
Errors raise RuntimeError with the reference's message text
(src/Solver.cpp:267-268, src/SolverData.cpp:73-78).
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _abi as A


def _check(rc, ctx=None, what=""):
    if rc == A.NF_OK:
        return
    msg = ""
    if ctx is not None:
        msg = A.lib().nf_last_error(ctx).decode()
    if rc in (A.NF_E_INVALID_DT, A.NF_E_NONPHYSICAL):
        raise RuntimeError(msg)
    raise RuntimeError(f"{what} failed with status {rc}: {msg}")


def _f64(a):
    return a.ctypes.data_as(A.f64p)


def _u32(a):
    return a.ctypes.data_as(A.u32p)


def _i32(a):
    return a.ctypes.data_as(A.i32p)


BC_NAMES = {"NoSlipWall": A.NF_BC_NOSLIP_WALL, "SlipWall": A.NF_BC_SLIP_WALL,
            "FarField": A.NF_BC_FARFIELD}


@dataclass
class Config:
    """The option set of NetFlux's Config that the explicit step depends on."""
    time_scheme: str = "TVD_RK3"          # ExplicitEuler | TVD_RK3
    spatial_order: str = "Second"         # First | Second
    limiter: str = "Barth"                # NONE | Barth
    grad_scheme: str = "GreenGauss"
    inviscid_flux_scheme: str = "Rusanov"
    CFL: float = 0.5
    density_fs: float = 1.225             # standard_air defaults (include/Utilities.hpp:101-107)
    velocity_fs: tuple = (0.0, 0.0, 0.0)
    pressure_fs: float = 101325.0
    n_timesteps: int = 1
    check_physical_validity: bool = False
    # reference quirk: U_old is zero on the first TVD-RK3 step
    rk3_zero_old_first_step: bool = True
    precision: str = "fp64"               # fp64 | mixed
    cell_order: str = "morton"            # input | rcm | morton (device cell numbering)
    local_time_step: bool = False         # per-cell CFL step (steady runs; not the reference)
    # run-time state held by Config in the reference
    delta_time: float = 0.0
    time: float = 0.0
    timestep: int = 0

    def primvars_inf(self):
        return [self.density_fs, *self.velocity_fs, self.pressure_fs]

    def to_params(self):
        p = A.Params()
        p.time_scheme = {"ExplicitEuler": A.NF_EXPLICIT_EULER, "TVD_RK3": A.NF_TVD_RK3}[self.time_scheme]
        p.spatial_order = {"First": A.NF_FIRST_ORDER, "Second": A.NF_SECOND_ORDER}[self.spatial_order]
        p.limiter = {"NONE": A.NF_LIMITER_NONE, "Barth": A.NF_LIMITER_BARTH}[self.limiter]
        p.inv_flux_scheme = {"Rusanov": A.NF_FLUX_RUSANOV, "HLLC": A.NF_FLUX_HLLC}[self.inviscid_flux_scheme]
        p.CFL = self.CFL
        for k, v in enumerate(self.primvars_inf()):
            p.primvars_inf[k] = float(v)
        p.check_physical_validity = int(self.check_physical_validity)
        p.rk3_zero_old_first_step = int(self.rk3_zero_old_first_step)
        p.precision = {"fp64": A.NF_PREC_FP64, "mixed": A.NF_PREC_MIXED}[self.precision]
        p.cell_order = {"input": A.NF_ORDER_INPUT, "rcm": A.NF_ORDER_RCM,
                        "morton": A.NF_ORDER_MORTON}[self.cell_order]
        p.local_time_step = int(self.local_time_step)
        return p


class Mesh:
    """Primal mesh (what PrimalGrid holds), owned by the library: tetrahedra,
    or mixed elements (hexahedra, wedges, pyramids, tetrahedra)."""

    def __init__(self, handle):
        self._h = handle
        self._desc = A.MeshDesc()
        _check(A.lib().nf_mesh_get_desc(self._h, C.byref(self._desc)), what="nf_mesh_get_desc")

    @classmethod
    def box(cls, n, lengths=(1.0, 1.0, 1.0), jitter=0.0, seed=20260101, bump_h=0.0,
            bump_sigma=0.2, patches=("SlipWall",) * 6):
        nx, ny, nz = (n, n, n) if isinstance(n, int) else n
        bc = np.array([BC_NAMES[p] if isinstance(p, str) else int(p) for p in patches], dtype=np.int32)
        h = C.c_void_p()
        _check(A.lib().nf_mesh_box(nx, ny, nz, *map(float, lengths), float(jitter), int(seed),
                                   float(bump_h), float(bump_sigma), _i32(bc), C.byref(h)),
               what="nf_mesh_box")
        return cls(h)

    @classmethod
    def bump_channel(cls, nx, ny, nz, h=0.05, sigma=0.2, lengths=(3.0, 1.0, 0.5)):
        """Channel with a Gaussian floor bump; x-planes FarField, others SlipWall."""
        return cls.box((nx, ny, nz), lengths=lengths, bump_h=h, bump_sigma=sigma,
                       patches=("FarField", "FarField") + ("SlipWall",) * 4)

    @classmethod
    def read_su2(cls, path, patch_bc):
        """SU2 file with tetrahedra and triangle markers; patch_bc maps each
        marker tag to a boundary type (name or NF_BC_* value)."""
        names = list(patch_bc.keys())
        arr = (C.c_char_p * max(1, len(names)))(*[n.encode() for n in names])
        bcs = np.array([BC_NAMES[b] if isinstance(b, str) else int(b) for b in patch_bc.values()] or [0],
                       dtype=np.int32)
        h = C.c_void_p()
        rc = A.lib().nf_mesh_read_su2(str(path).encode(), len(names), arr, _i32(bcs), C.byref(h))
        if rc != A.NF_OK:
            raise RuntimeError(f"Error reading su2 mesh '{path}': {A.lib().nf_mesh_last_error().decode()}")
        return cls(h)

    def write_su2(self, path, names=None):
        names = list(names or [])
        arr = (C.c_char_p * max(1, len(names)))(*[n.encode() for n in names])
        rc = A.lib().nf_mesh_write_su2(self._h, str(path).encode(), len(names), arr)
        if rc != A.NF_OK:
            raise RuntimeError(f"nf_mesh_write_su2 failed ({rc}): {A.lib().nf_mesh_last_error().decode()}")

    @property
    def patch_names(self):
        out = []
        buf = C.create_string_buffer(256)
        for i in range(self._desc.n_patches):
            _check(A.lib().nf_mesh_patch_name(self._h, i, buf, 256), what="nf_mesh_patch_name")
            out.append(buf.value.decode())
        return out

    @classmethod
    def from_arrays(cls, nodes, cells, patch_faces, patch_bc):
        """nodes [n][3] f64, cells [N][4] u32 (VTK order), patch_faces list of [m][3]."""
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        cells = np.ascontiguousarray(cells, dtype=np.uint32)
        pn = np.array([len(p) for p in patch_faces], dtype=np.uint32)
        pf = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.uint32).reshape(-1, 3)
                                                  for p in patch_faces]), dtype=np.uint32)
        pbc = np.array([BC_NAMES[b] if isinstance(b, str) else int(b) for b in patch_bc], dtype=np.int32)
        d = A.MeshDesc(len(nodes), _f64(nodes), len(cells), _u32(cells), len(pn), _u32(pn), _u32(pf), _i32(pbc))
        h = C.c_void_p()
        _check(A.lib().nf_mesh_from_desc(C.byref(d), C.byref(h)), what="nf_mesh_from_desc")
        return cls(h)

    @classmethod
    def from_elements(cls, nodes, cells, patch_faces, patch_bc):
        """Mixed-element mesh (the reference's Elements): cells is a list of
        node lists in VTK order whose length gives the type (4 tetrahedron,
        5 pyramid, 6 wedge, 8 hexahedron); patch_faces holds per patch a list
        of node lists (3 triangle, 4 quadrilateral).  An all-tetrahedron input
        gives the tetrahedral layout."""
        ctype_of = {4: A.NF_ELEM_TETRA, 5: A.NF_ELEM_PYRAMID, 6: A.NF_ELEM_WEDGE, 8: A.NF_ELEM_HEXA}
        ftype_of = {3: A.NF_ELEM_TRIANGLE, 4: A.NF_ELEM_QUAD}

        def flat(rows, type_of):
            """(types, offsets, nodes) of a list of node lists or a 2-D array"""
            if isinstance(rows, np.ndarray) and rows.ndim == 2:
                n, k = rows.shape
                return (np.full(n, type_of[k], np.uint8), (np.arange(n + 1, dtype=np.int64) * k).astype(np.uint32),
                        np.ascontiguousarray(rows, dtype=np.uint32).ravel())
            if len(rows) == 0:
                return np.zeros(0, np.uint8), np.zeros(1, np.uint32), np.zeros(0, np.uint32)
            return (np.array([type_of[len(r)] for r in rows], dtype=np.uint8),
                    np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.uint32),
                    np.concatenate([np.asarray(r, dtype=np.uint32) for r in rows]))

        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        ct, co, cn = flat(cells, ctype_of)
        parts = [flat(p, ftype_of) for p in patch_faces]
        pn = np.array([len(t) for t, _, _ in parts], dtype=np.uint32)
        ft = np.concatenate([t for t, _, _ in parts]).astype(np.uint8)
        fn = np.ascontiguousarray(np.concatenate([v for _, _, v in parts]), dtype=np.uint32)
        fo = np.concatenate([[0], np.cumsum(np.concatenate([np.diff(o.astype(np.int64)) for _, o, _ in parts]))])
        fo = fo.astype(np.uint32)
        pbc = np.array([BC_NAMES[b] if isinstance(b, str) else int(b) for b in patch_bc], dtype=np.int32)
        d = A.MeshDesc(len(nodes), _f64(nodes), len(cells), _u32(cn), len(pn), _u32(pn), _u32(fn), _i32(pbc),
                       ct.ctypes.data_as(A.u8p), _u32(co), ft.ctypes.data_as(A.u8p), _u32(fo))
        h = C.c_void_p()
        rc = A.lib().nf_mesh_from_desc(C.byref(d), C.byref(h))
        if rc != A.NF_OK:
            raise RuntimeError(f"nf_mesh_from_desc failed ({rc}): {A.lib().nf_mesh_last_error().decode()}")
        return cls(h)

    @classmethod
    def hex_box(cls, n, lengths=(1.0, 1.0, 1.0), jitter=0.0, seed=20260101, patches=("SlipWall",) * 6):
        """nx x ny x nz hexahedra (VTK order) on a box, interior nodes moved
        uniformly by up to jitter x spacing; quadrilateral patches x-, x+,
        y-, y+, z-, z+."""
        nx, ny, nz = (n, n, n) if isinstance(n, int) else n
        h = np.array(lengths, float) / np.array([nx, ny, nz], float)
        I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
        X = np.stack([I * h[0], J * h[1], K * h[2]], -1).reshape(-1, 3)
        if jitter > 0:
            inner = ((I > 0) & (I < nx) & (J > 0) & (J < ny) & (K > 0) & (K < nz)).ravel()
            rng = np.random.default_rng(seed)
            X[inner] += (rng.random((int(inner.sum()), 3)) - 0.5) * 2 * jitter * h

        def idx(i, j, k):
            return ((i * (ny + 1) + j) * (nz + 1) + k).astype(np.int64)

        i, j, k = [a.ravel() for a in np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")]
        cells = np.stack([idx(i, j, k), idx(i + 1, j, k), idx(i + 1, j + 1, k), idx(i, j + 1, k),
                          idx(i, j, k + 1), idx(i + 1, j, k + 1), idx(i + 1, j + 1, k + 1), idx(i, j + 1, k + 1)], 1)

        def side(axis, end):
            dims = [nx, ny, nz]
            a, b = [d for d in range(3) if d != axis]
            u, v = [w.ravel() for w in np.meshgrid(np.arange(dims[a]), np.arange(dims[b]), indexing="ij")]
            c = np.full_like(u, 0 if end == 0 else dims[axis])

            def at(du, dv):
                q = [None] * 3
                q[axis], q[a], q[b] = c, u + du, v + dv
                return idx(*q)
            return np.stack([at(0, 0), at(1, 0), at(1, 1), at(0, 1)], 1)

        faces = [side(ax, e) for ax in range(3) for e in (0, 1)]
        return cls.from_elements(X, cells, faces, list(patches))

    @property
    def is_mixed(self):
        """True for a mesh with hexahedra, wedges or pyramids."""
        return bool(self._desc.cell_types)

    def elements(self):
        """(cell_types, cell_offsets, cell_nodes, patch_n, face_types,
        face_offsets, face_nodes, patch_bc) in either layout."""
        d = self._desc
        N, P = d.n_cells, d.n_patches
        pn = np.ctypeslib.as_array(d.patch_n_faces, shape=(P,)).copy() if P else np.zeros(0, np.uint32)
        pbc = np.ctypeslib.as_array(d.patch_bc, shape=(P,)).copy() if P else np.zeros(0, np.int32)
        nf_ = int(pn.sum())
        if not self.is_mixed:
            _, cells, _, pf, _ = self.arrays()
            return (np.full(N, A.NF_ELEM_TETRA, np.uint8), np.arange(N + 1, dtype=np.uint32) * 4, cells.ravel(),
                    pn, np.full(nf_, A.NF_ELEM_TRIANGLE, np.uint8), np.arange(nf_ + 1, dtype=np.uint32) * 3,
                    pf.ravel(), pbc)
        ct = np.ctypeslib.as_array(d.cell_types, shape=(N,)).copy()
        co = np.ctypeslib.as_array(d.cell_offsets, shape=(N + 1,)).copy()
        cn = np.ctypeslib.as_array(d.cells, shape=(int(co[-1]),)).copy()
        ft = np.ctypeslib.as_array(d.patch_face_types, shape=(nf_,)).copy() if nf_ else np.zeros(0, np.uint8)
        fo = np.ctypeslib.as_array(d.patch_face_offsets, shape=(nf_ + 1,)).copy()
        fn = np.ctypeslib.as_array(d.patch_faces, shape=(int(fo[-1]),)).copy() if nf_ else np.zeros(0, np.uint32)
        return ct, co, cn, pn, ft, fo, fn, pbc

    def __del__(self):
        if getattr(self, "_h", None):
            A.lib().nf_mesh_destroy(self._h)
            self._h = None

    @property
    def desc(self):
        return self._desc

    def arrays(self):
        if self.is_mixed:
            raise ValueError("mixed-element mesh: use elements()")
        d = self._desc
        nodes = np.ctypeslib.as_array(d.nodes, shape=(d.n_nodes, 3)).copy()
        cells = np.ctypeslib.as_array(d.cells, shape=(d.n_cells, 4)).copy()
        pn = np.ctypeslib.as_array(d.patch_n_faces, shape=(d.n_patches,)).copy()
        pf = np.ctypeslib.as_array(d.patch_faces, shape=(int(pn.sum()), 3)).copy()
        pbc = np.ctypeslib.as_array(d.patch_bc, shape=(d.n_patches,)).copy()
        return nodes, cells, pn, pf, pbc

    @property
    def n_cells(self):
        return self._desc.n_cells

    def fv_grid(self):
        """FV_Grid of this mesh: dict of numpy arrays in the reference face order."""
        s = A.FVSizes()
        rc = A.lib().nf_fv_sizes_of(self._h, C.byref(s))
        if rc != A.NF_OK:
            raise RuntimeError(f"Error creating grid: {A.lib().nf_mesh_last_error().decode()}")
        F, NT, P = s.n_total_faces, s.n_total_cells, s.n_patches
        out = dict(ij=np.zeros((F, 2), np.uint32), S=np.zeros((F, 3)), r_i=np.zeros((F, 3)),
                   r_j=np.zeros((F, 3)), volume=np.zeros(NT), centroid=np.zeros((NT, 3)),
                   patch_first=np.zeros(P, np.uint32), patch_nfaces=np.zeros(P, np.uint32),
                   delta_S=np.zeros((NT, 3)))
        _check(A.lib().nf_fv_export(self._h, _u32(out["ij"]), _f64(out["S"]), _f64(out["r_i"]),
                                    _f64(out["r_j"]), _f64(out["volume"]), _f64(out["centroid"]),
                                    _u32(out["patch_first"]), _u32(out["patch_nfaces"]),
                                    _f64(out["delta_S"])), what="nf_fv_export")
        out["sizes"] = (s.n_interior_cells, s.n_total_cells, s.n_interior_faces, s.n_total_faces, P)
        return out


class SolverData:
    """Read-only view of the solver fields (SolverData accessors); every
    accessor copies from the device in the reference layout."""

    def __init__(self, solver):
        self._s = solver

    def _field(self, which, cols):
        out = np.zeros((self._s.N, cols))
        _check(A.lib().nf_get_field(self._s._ctx, which, _f64(out)), self._s._ctx, "nf_get_field")
        return out

    def get_solution(self):
        out = np.zeros((self._s.N, 5))
        _check(A.lib().nf_get_state(self._s._ctx, _f64(out)), self._s._ctx, "nf_get_state")
        return out

    def get_solution_old(self):
        return self._field(A.NF_FIELD_STATE_OLD, 5)

    def get_primvars(self):
        return self._field(A.NF_FIELD_PRIM, 5)

    def get_primvars_gradient(self):
        """[N][15], per cell the 5x3 gradient column-major (offset k + 5 d)."""
        return self._field(A.NF_FIELD_GRADIENT, 15)

    def get_primvars_limiter(self):
        return self._field(A.NF_FIELD_LIMITER, 5)

    def get_flux_balance(self):
        return self._field(A.NF_FIELD_RESIDUAL, 5)


class EulerSolver:
    """Drop-in for EulerSolver(config, primal_grid, FV_grid) on one GPU."""

    def __init__(self, config: Config, mesh: Mesh):
        self.config = config
        self.mesh = mesh
        self._params = config.to_params()
        ctx = C.c_void_p()
        rc = A.lib().nf_create(C.byref(mesh.desc), C.byref(self._params), C.byref(ctx))
        if rc != A.NF_OK:
            raise RuntimeError(f"nf_create failed with status {rc}")
        self._ctx = ctx
        s = A.FVSizes()
        A.lib().nf_ctx_sizes(self._ctx, C.byref(s))
        self.sizes = s
        self.N = s.n_interior_cells
        self._data = SolverData(self)

    def __del__(self):
        if getattr(self, "_ctx", None):
            A.lib().nf_destroy(self._ctx)
            self._ctx = None

    # --- the reference's public surface ---
    def calc_timestep(self, config: Config):
        dt = C.c_double()
        _check(A.lib().nf_calc_timestep(self._ctx, C.byref(dt)), self._ctx, "nf_calc_timestep")
        config.delta_time = dt.value

    def step(self, config: Config):
        _check(A.lib().nf_step(self._ctx, config.delta_time), self._ctx, "nf_step")

    def get_solver_data(self):
        return self._data

    def get_solver_type(self):
        return "Euler"

    # --- additions ---
    def set_state(self, U, U_old=None):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == (self.N, 5)
        uo = None if U_old is None else _f64(np.ascontiguousarray(U_old, dtype=np.float64))
        _check(A.lib().nf_set_state(self._ctx, _f64(U), uo), self._ctx, "nf_set_state")

    def set_freestream(self):
        _check(A.lib().nf_set_freestream(self._ctx), self._ctx, "nf_set_freestream")

    def eval_flux_balance(self):
        _check(A.lib().nf_eval_flux_balance(self._ctx), self._ctx, "nf_eval_flux_balance")

    def run(self, n_steps, residual_drop=0.0, check_every=10, use_graph=True):
        o = A.RunOpts(n_steps, residual_drop, check_every, int(use_graph))
        st = A.RunStats()
        _check(A.lib().nf_run(self._ctx, C.byref(o), C.byref(st)), self._ctx, "nf_run")
        return {f: getattr(st, f) for f, _ in A.RunStats._fields_}

    def residuals(self, max_n=1 << 20):
        out = np.zeros(max_n)
        n = C.c_int32()
        _check(A.lib().nf_get_residuals(self._ctx, _f64(out), max_n, C.byref(n)), self._ctx, "nf_get_residuals")
        return out[: n.value].copy()

    def time_kernels(self, n_steps, max_k=32):
        arr = (A.KernelTime * max_k)()
        n = C.c_int32()
        _check(A.lib().nf_time_kernels(self._ctx, n_steps, arr, max_k, C.byref(n)), self._ctx, "nf_time_kernels")
        return {arr[i].name.decode(): (arr[i].avg_ms, arr[i].calls) for i in range(n.value)}

    def autotune(self, steps=10):
        return _autotune(self._ctx, steps)

    def set_launch(self, block, xcd_remap=True):
        _check(A.lib().nf_set_launch(self._ctx, block, int(xcd_remap)), self._ctx, "nf_set_launch")

    def set_precision(self, precision):
        p = {"fp64": A.NF_PREC_FP64, "mixed": A.NF_PREC_MIXED}.get(precision, precision)
        _check(A.lib().nf_set_precision(self._ctx, p), self._ctx, "nf_set_precision")


def _autotune(ctx, steps):
    r = A.TuneResult()
    _check(A.lib().nf_autotune(ctx, steps, C.byref(r)), ctx, "nf_autotune")
    return {"block": r.block, "xcd_remap": r.xcd_remap, "ms_per_step": r.ms_per_step,
            "tried_ms": list(r.tried_ms[: r.n_tried])}


class BatchSolver:
    """Several meshes advanced together: every sweep is one launch over the
    concatenated cells of all instances, each instance keeping its own time
    step, step counter, residual history and stop flag.

    configs may be one Config for all meshes or one per mesh; only the
    freestream state may differ between them.
    """

    def __init__(self, configs, meshes):
        meshes = list(meshes)
        if isinstance(configs, Config):
            configs = [configs] * len(meshes)
        configs = list(configs)
        if len(configs) != len(meshes) or not meshes:
            raise ValueError("one config per mesh")
        self.meshes = meshes
        n = len(meshes)
        descs = (A.MeshDesc * n)(*[m.desc for m in meshes])
        params = (A.Params * n)(*[c.to_params() for c in configs])
        ctx = C.c_void_p()
        rc = A.lib().nf_create_batch(n, descs, params, C.byref(ctx))
        if rc != A.NF_OK:
            raise RuntimeError(f"nf_create_batch failed with status {rc}")
        self._ctx = ctx
        self.n = n
        self.N = []
        for i in range(n):
            s = A.FVSizes()
            _check(A.lib().nf_inst_sizes(ctx, i, C.byref(s)), ctx, "nf_inst_sizes")
            self.N.append(s.n_interior_cells)

    def __del__(self):
        if getattr(self, "_ctx", None):
            A.lib().nf_destroy(self._ctx)
            self._ctx = None

    def set_state(self, i, U, U_old=None):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == (self.N[i], 5)
        uo = None if U_old is None else _f64(np.ascontiguousarray(U_old, dtype=np.float64))
        _check(A.lib().nf_inst_set_state(self._ctx, i, _f64(U), uo), self._ctx, "nf_inst_set_state")

    def set_freestream(self):
        _check(A.lib().nf_set_freestream(self._ctx), self._ctx, "nf_set_freestream")

    def get_solution(self, i):
        out = np.zeros((self.N[i], 5))
        _check(A.lib().nf_inst_get_state(self._ctx, i, _f64(out)), self._ctx, "nf_inst_get_state")
        return out

    def residuals(self, i, max_n=1 << 20):
        out = np.zeros(max_n)
        n = C.c_int32()
        _check(A.lib().nf_inst_residuals(self._ctx, i, _f64(out), max_n, C.byref(n)), self._ctx,
               "nf_inst_residuals")
        return out[: n.value].copy()

    def features(self, i, case_params):
        """Network features [N][6] of instance i (reference cell order)."""
        cp = np.ascontiguousarray(case_params, np.float64)
        x = np.zeros((self.N[i], 6), np.float32)
        _check(A.lib().nf_inst_features(self._ctx, i, _f64(cp), x.ctypes.data_as(A.f32p)), self._ctx,
               "nf_inst_features")
        return x

    def warm_start(self, i, net, case_params):
        """Network prediction as the state of instance i; True if it fell back to the freestream."""
        cp = np.ascontiguousarray(case_params, np.float64)
        fb = C.c_int32()
        _check(A.lib().nf_inst_warm_start(self._ctx, i, net._h, _f64(cp), C.byref(fb)), self._ctx,
               "nf_inst_warm_start")
        return bool(fb.value)

    def set_residual_ref(self, i, ref):
        _check(A.lib().nf_inst_set_residual_ref(self._ctx, i, float(ref)), self._ctx, "nf_inst_set_residual_ref")

    def eval_residuals(self):
        out = np.zeros(self.n)
        _check(A.lib().nf_eval_residuals(self._ctx, _f64(out)), self._ctx, "nf_eval_residuals")
        return out

    def run(self, n_steps, residual_drop=0.0, check_every=10, use_graph=True, stall_window=0, stall_ratio=0.0):
        o = A.RunOpts(n_steps, residual_drop, check_every, int(use_graph), stall_window, stall_ratio)
        st = (A.RunStats * self.n)()
        _check(A.lib().nf_run_batch(self._ctx, C.byref(o), st), self._ctx, "nf_run_batch")
        return [{f: getattr(st[i], f) for f, _ in A.RunStats._fields_} for i in range(self.n)]

    def autotune(self, steps=10):
        return _autotune(self._ctx, steps)

    def set_launch(self, block, xcd_remap=True):
        _check(A.lib().nf_set_launch(self._ctx, block, int(xcd_remap)), self._ctx, "nf_set_launch")

    def set_precision(self, precision):
        p = {"fp64": A.NF_PREC_FP64, "mixed": A.NF_PREC_MIXED}.get(precision, precision)
        _check(A.lib().nf_set_precision(self._ctx, p), self._ctx, "nf_set_precision")


class Driver:
    """The reference's time loop (src/Driver.cpp:31-57) without VTK output."""

    def __init__(self, config: Config, mesh: Mesh):
        self.config = config
        self.solver = EulerSolver(config, mesh)

    def solve(self):
        cfg = self.config
        cfg.timestep = 0
        cfg.time = 0.0
        while True:
            self.solver.calc_timestep(cfg)
            self.solver.step(cfg)
            cfg.timestep += 1
            cfg.time += cfg.delta_time
            if cfg.timestep >= cfg.n_timesteps:
                break
