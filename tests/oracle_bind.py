"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- tests only."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")


class OrParams(C.Structure):
    _fields_ = [("time_scheme", C.c_int32), ("spatial_order", C.c_int32), ("limiter", C.c_int32),
                ("CFL", C.c_double), ("V_inf", C.c_double * 5),
                ("check_physical_validity", C.c_int32), ("rk3_zero_old_first_step", C.c_int32),
                ("inv_flux_scheme", C.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        L = C.CDLL(LIB)
        f64p, u32p, i32p = C.POINTER(C.c_double), C.POINTER(C.c_uint32), C.POINTER(C.c_int32)
        L.or_create.restype = C.c_void_p
        L.or_create.argtypes = [C.c_uint32, f64p, C.c_uint32, u32p, C.c_uint32, u32p, u32p, i32p,
                                C.POINTER(OrParams), C.c_char_p, C.c_int]
        u8p = C.POINTER(C.c_uint8)
        L.or_create_mixed.restype = C.c_void_p
        L.or_create_mixed.argtypes = [C.c_uint32, f64p, C.c_uint32, u8p, u32p, u32p, C.c_uint32, u32p, u8p, u32p,
                                      u32p, i32p, C.POINTER(OrParams), C.c_char_p, C.c_int]
        L.or_destroy.argtypes = [C.c_void_p]
        L.or_sizes.argtypes = [C.c_void_p, u32p]
        L.or_export_grid.argtypes = [C.c_void_p, u32p, f64p, f64p, f64p, f64p, f64p, u32p, u32p, f64p]
        L.or_set_state.argtypes = [C.c_void_p, f64p, C.c_int]
        L.or_set_freestream.argtypes = [C.c_void_p]
        L.or_calc_timestep.argtypes = [C.c_void_p, f64p]
        L.or_step.argtypes = [C.c_void_p, C.c_double]
        L.or_run.argtypes = [C.c_void_p, C.c_int, f64p]
        L.or_eval_flux_balance.argtypes = [C.c_void_p]
        L.or_get.argtypes = [C.c_void_p, C.c_int, f64p]
        L.or_time.restype = C.c_double
        L.or_time.argtypes = [C.c_void_p]
        L.or_error.restype = C.c_char_p
        L.or_error.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Oracle:
    """Reference-restating CPU solver over the same primal mesh arrays."""

    def __init__(self, nodes, cells, patch_n, patch_faces, patch_bc, config, mixed=None):
        """cells [N][4] and patch_faces [m][3]; or, with mixed = (cell_types,
        cell_offsets, face_types, face_offsets), the flat node lists of
        mixed elements."""
        self._keep = [np.ascontiguousarray(nodes, np.float64), np.ascontiguousarray(cells, np.uint32),
                      np.ascontiguousarray(patch_n, np.uint32), np.ascontiguousarray(patch_faces, np.uint32),
                      np.ascontiguousarray(patch_bc, np.int32)]
        n, c, pn, pf, pb = self._keep
        p = OrParams()
        p.time_scheme = {"ExplicitEuler": 0, "TVD_RK3": 1}[config.time_scheme]
        p.spatial_order = {"First": 0, "Second": 1}[config.spatial_order]
        p.limiter = {"NONE": 0, "Barth": 1}[config.limiter]
        p.CFL = config.CFL
        for k, v in enumerate(config.primvars_inf()):
            p.V_inf[k] = v
        p.check_physical_validity = int(config.check_physical_validity)
        p.rk3_zero_old_first_step = int(config.rk3_zero_old_first_step)
        p.inv_flux_scheme = {"Rusanov": 0, "HLLC": 1}[config.inviscid_flux_scheme]
        err = C.create_string_buffer(256)
        if mixed is None:
            self._h = lib().or_create(len(n), _p(n, C.c_double), len(c), _p(c, C.c_uint32), len(pn),
                                      _p(pn, C.c_uint32), _p(pf, C.c_uint32), _p(pb, C.c_int32), C.byref(p), err, 256)
        else:
            ct, co, ft, fo = [np.ascontiguousarray(a, t) for a, t in
                              zip(mixed, (np.uint8, np.uint32, np.uint8, np.uint32))]
            self._keep += [ct, co, ft, fo]
            self._h = lib().or_create_mixed(len(n), _p(n, C.c_double), len(ct), _p(ct, C.c_uint8), _p(co, C.c_uint32),
                                            _p(c, C.c_uint32), len(pn), _p(pn, C.c_uint32), _p(ft, C.c_uint8),
                                            _p(fo, C.c_uint32), _p(pf, C.c_uint32), _p(pb, C.c_int32), C.byref(p),
                                            err, 256)
        if not self._h:
            raise RuntimeError(err.value.decode())
        s = np.zeros(5, np.uint32)
        lib().or_sizes(self._h, _p(s, C.c_uint32))
        self.N, self.NT, self.Fint, self.F, self.P = map(int, s)

    @classmethod
    def from_mesh(cls, mesh, config):
        if getattr(mesh, "is_mixed", False):
            ct, co, cn, pn, ft, fo, fn, pbc = mesh.elements()
            d = mesh.desc
            nodes = np.ctypeslib.as_array(d.nodes, shape=(d.n_nodes, 3)).copy()
            return cls(nodes, cn, pn, fn, pbc, config, mixed=(ct, co, ft, fo))
        nodes, cells, pn, pf, pbc = mesh.arrays()
        return cls(nodes, cells, pn, pf, pbc, config)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_destroy(self._h)
            self._h = None

    def grid(self):
        F, NT, P = self.F, self.NT, self.P
        g = dict(ij=np.zeros((F, 2), np.uint32), S=np.zeros((F, 3)), r_i=np.zeros((F, 3)),
                 r_j=np.zeros((F, 3)), volume=np.zeros(NT), centroid=np.zeros((NT, 3)),
                 patch_first=np.zeros(P, np.uint32), patch_nfaces=np.zeros(P, np.uint32),
                 delta_S=np.zeros((NT, 3)))
        lib().or_export_grid(self._h, _p(g["ij"], C.c_uint32), _p(g["S"], C.c_double), _p(g["r_i"], C.c_double),
                             _p(g["r_j"], C.c_double), _p(g["volume"], C.c_double), _p(g["centroid"], C.c_double),
                             _p(g["patch_first"], C.c_uint32), _p(g["patch_nfaces"], C.c_uint32),
                             _p(g["delta_S"], C.c_double))
        return g

    def set_state(self, U, uold_equals_u=False):
        U = np.ascontiguousarray(U, np.float64)
        lib().or_set_state(self._h, _p(U, C.c_double), int(uold_equals_u))

    def set_freestream(self):
        lib().or_set_freestream(self._h)

    def calc_timestep(self):
        dt = C.c_double()
        rc = lib().or_calc_timestep(self._h, C.byref(dt))
        if rc:
            raise RuntimeError(lib().or_error(self._h).decode())
        return dt.value

    def step(self, dt):
        if lib().or_step(self._h, dt):
            raise RuntimeError(lib().or_error(self._h).decode())

    def run(self, n):
        t = C.c_double()
        rc = lib().or_run(self._h, n, C.byref(t))
        if rc:
            raise RuntimeError(lib().or_error(self._h).decode())
        return t.value

    def eval_flux_balance(self):
        if lib().or_eval_flux_balance(self._h):
            raise RuntimeError(lib().or_error(self._h).decode())

    def get(self, which):
        rows = {"U": self.N, "V": self.NT, "psi": self.N, "grad": self.N, "R": self.N,
                "vmax": self.N, "vmin": self.N, "U_old": self.N}[which]
        cols = 15 if which == "grad" else 5
        out = np.zeros((rows, cols))
        idx = ["U", "V", "psi", "grad", "R", "vmax", "vmin", "U_old"].index(which)
        lib().or_get(self._h, idx, _p(out, C.c_double))
        return out

    def time(self):
        return lib().or_time(self._h)


def _vec(a):
    return np.ascontiguousarray(a, np.float64)


def prim_to_cons(V):
    V, U = _vec(V), np.zeros(5)
    lib().or_prim_to_cons(_p(V, C.c_double), _p(U, C.c_double))
    return U


def cons_to_prim(U):
    U, V = _vec(U), np.zeros(5)
    lib().or_cons_to_prim(_p(U, C.c_double), _p(V, C.c_double))
    return V


def rusanov(UL, UR, S):
    UL, UR, S, F = _vec(UL), _vec(UR), _vec(S), np.zeros(5)
    lib().or_rusanov(_p(UL, C.c_double), _p(UR, C.c_double), _p(S, C.c_double), _p(F, C.c_double))
    return F


def hllc(UL, UR, S):
    UL, UR, S, F = _vec(UL), _vec(UR), _vec(S), np.zeros(5)
    lib().or_hllc(_p(UL, C.c_double), _p(UR, C.c_double), _p(S, C.c_double), _p(F, C.c_double))
    return F


def ghost(bc_type, V_fs, V_domain, S):
    V_fs, Vd, S, Vg = _vec(V_fs), _vec(V_domain), _vec(S), np.zeros(5)
    lib().or_ghost(int(bc_type), _p(V_fs, C.c_double), _p(Vd, C.c_double), _p(S, C.c_double),
                   _p(Vg, C.c_double))
    return Vg
