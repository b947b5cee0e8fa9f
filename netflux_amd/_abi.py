"""ctypes binding of libnetflux.so (declarations in include/netflux.h).

The library is built in-tree by ``netflux_amd/Makefile``.  Loading fails
loudly when it is missing: there is no CPU fallback for the solver.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# NETFLUX_LIB overrides the library path (e.g. to compare builds)
LIB_PATH = os.environ.get("NETFLUX_LIB") or os.path.join(_HERE, "libnetflux.so")

NF_OK, NF_E_ARG, NF_E_MESH, NF_E_HIP, NF_E_INVALID_DT, NF_E_NONPHYSICAL, NF_E_NOGPU = range(7)
NF_EXPLICIT_EULER, NF_TVD_RK3 = 0, 1
NF_FIRST_ORDER, NF_SECOND_ORDER = 0, 1
NF_LIMITER_NONE, NF_LIMITER_BARTH = 0, 1
NF_FLUX_RUSANOV, NF_FLUX_HLLC = 0, 1
NF_BC_NOSLIP_WALL, NF_BC_SLIP_WALL, NF_BC_FARFIELD = 0, 1, 2
NF_PREC_FP64, NF_PREC_MIXED = 0, 1
NF_ORDER_INPUT, NF_ORDER_RCM, NF_ORDER_MORTON = 0, 1, 2
NF_FIELD_PRIM, NF_FIELD_LIMITER, NF_FIELD_GRADIENT, NF_FIELD_RESIDUAL, NF_FIELD_STATE_OLD = range(5)
NF_ELEM_TRIANGLE, NF_ELEM_QUAD, NF_ELEM_TETRA, NF_ELEM_HEXA, NF_ELEM_WEDGE, NF_ELEM_PYRAMID = 5, 9, 10, 12, 13, 14

u32p = C.POINTER(C.c_uint32)
u8p = C.POINTER(C.c_uint8)
i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)
f32p = C.POINTER(C.c_float)


class MeshDesc(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_uint32), ("nodes", f64p),
        ("n_cells", C.c_uint32), ("cells", u32p),
        ("n_patches", C.c_uint32), ("patch_n_faces", u32p),
        ("patch_faces", u32p), ("patch_bc", i32p),
        ("cell_types", u8p), ("cell_offsets", u32p),
        ("patch_face_types", u8p), ("patch_face_offsets", u32p),
    ]


class Params(C.Structure):
    _fields_ = [
        ("time_scheme", C.c_int32), ("spatial_order", C.c_int32),
        ("limiter", C.c_int32), ("inv_flux_scheme", C.c_int32),
        ("CFL", C.c_double), ("primvars_inf", C.c_double * 5),
        ("check_physical_validity", C.c_int32),
        ("rk3_zero_old_first_step", C.c_int32),
        ("precision", C.c_int32), ("cell_order", C.c_int32),
        ("local_time_step", C.c_int32),
    ]


class FVSizes(C.Structure):
    _fields_ = [("n_interior_cells", C.c_uint32), ("n_total_cells", C.c_uint32),
                ("n_interior_faces", C.c_uint32), ("n_total_faces", C.c_uint32),
                ("n_patches", C.c_uint32)]


class RunOpts(C.Structure):
    _fields_ = [("n_steps", C.c_int32), ("residual_drop", C.c_double),
                ("check_every", C.c_int32), ("use_graph", C.c_int32),
                ("stall_window", C.c_int32), ("stall_ratio", C.c_double)]


class RunStats(C.Structure):
    _fields_ = [("steps_done", C.c_int32), ("converged", C.c_int32),
                ("time", C.c_double), ("last_dt", C.c_double),
                ("res_first", C.c_double), ("res_last", C.c_double),
                ("elapsed_ms", C.c_double), ("fell_back", C.c_int32), ("pad", C.c_int32)]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("avg_ms", C.c_double), ("calls", C.c_int32)]


class TuneResult(C.Structure):
    _fields_ = [("block", C.c_int32), ("xcd_remap", C.c_int32), ("ms_per_step", C.c_double),
                ("n_tried", C.c_int32), ("tried_ms", C.c_double * 8)]


class TrainOpts(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("lr_decay", C.c_float), ("seed", C.c_uint64)]


class TrainStats(C.Structure):
    _fields_ = [("loss_first", C.c_float), ("loss_last", C.c_float), ("steps", C.c_int64),
                ("elapsed_ms", C.c_double)]


# (name, restype, argtypes) of every symbol declared in include/netflux.h
SIGNATURES = [
    ("nf_mesh_box", C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                              C.c_double, C.c_double, C.c_uint64, C.c_double, C.c_double,
                              i32p, C.POINTER(C.c_void_p)]),
    ("nf_mesh_from_desc", C.c_int, [C.POINTER(MeshDesc), C.POINTER(C.c_void_p)]),
    ("nf_mesh_get_desc", C.c_int, [C.c_void_p, C.POINTER(MeshDesc)]),
    ("nf_mesh_destroy", None, [C.c_void_p]),
    ("nf_mesh_read_su2", C.c_int, [C.c_char_p, C.c_int32, C.POINTER(C.c_char_p), i32p, C.POINTER(C.c_void_p)]),
    ("nf_mesh_write_su2", C.c_int, [C.c_void_p, C.c_char_p, C.c_int32, C.POINTER(C.c_char_p)]),
    ("nf_mesh_patch_name", C.c_int, [C.c_void_p, C.c_int32, C.c_char_p, C.c_int32]),
    ("nf_mesh_last_error", C.c_char_p, []),
    ("nf_fv_sizes_of", C.c_int, [C.c_void_p, C.POINTER(FVSizes)]),
    ("nf_fv_export", C.c_int, [C.c_void_p, u32p, f64p, f64p, f64p, f64p, f64p, u32p, u32p, f64p]),
    ("nf_create", C.c_int, [C.POINTER(MeshDesc), C.POINTER(Params), C.POINTER(C.c_void_p)]),
    ("nf_destroy", C.c_int, [C.c_void_p]),
    ("nf_last_error", C.c_char_p, [C.c_void_p]),
    ("nf_set_freestream", C.c_int, [C.c_void_p]),
    ("nf_set_state", C.c_int, [C.c_void_p, f64p, f64p]),
    ("nf_calc_timestep", C.c_int, [C.c_void_p, f64p]),
    ("nf_step", C.c_int, [C.c_void_p, C.c_double]),
    ("nf_run", C.c_int, [C.c_void_p, C.POINTER(RunOpts), C.POINTER(RunStats)]),
    ("nf_get_residuals", C.c_int, [C.c_void_p, f64p, C.c_int32, i32p]),
    ("nf_eval_flux_balance", C.c_int, [C.c_void_p]),
    ("nf_get_state", C.c_int, [C.c_void_p, f64p]),
    ("nf_get_field", C.c_int, [C.c_void_p, C.c_int32, f64p]),
    ("nf_time_kernels", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(KernelTime), C.c_int32, i32p]),
    ("nf_ctx_sizes", C.c_int, [C.c_void_p, C.POINTER(FVSizes)]),
    ("nf_create_batch", C.c_int, [C.c_int32, C.POINTER(MeshDesc), C.POINTER(Params),
                                  C.POINTER(C.c_void_p)]),
    ("nf_ctx_instances", C.c_int, [C.c_void_p, i32p]),
    ("nf_inst_sizes", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(FVSizes)]),
    ("nf_inst_set_state", C.c_int, [C.c_void_p, C.c_int32, f64p, f64p]),
    ("nf_inst_get_state", C.c_int, [C.c_void_p, C.c_int32, f64p]),
    ("nf_inst_residuals", C.c_int, [C.c_void_p, C.c_int32, f64p, C.c_int32, i32p]),
    ("nf_run_batch", C.c_int, [C.c_void_p, C.POINTER(RunOpts), C.POINTER(RunStats)]),
    ("nf_autotune", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(TuneResult)]),
    ("nf_set_launch", C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    ("nf_set_precision", C.c_int, [C.c_void_p, C.c_int32]),
    ("nf_inst_set_residual_ref", C.c_int, [C.c_void_p, C.c_int32, C.c_double]),
    ("nf_eval_residuals", C.c_int, [C.c_void_p, f64p]),
    ("nf_inst_features", C.c_int, [C.c_void_p, C.c_int32, f64p, f32p]),
    ("nf_inst_warm_start", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, f64p, i32p]),
    ("nf_mlp_n_params", C.c_int64, []),
    ("nf_mlp_create", C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    ("nf_mlp_destroy", C.c_int, [C.c_void_p]),
    ("nf_mlp_last_error", C.c_char_p, [C.c_void_p]),
    ("nf_mlp_get_params", C.c_int, [C.c_void_p, f32p]),
    ("nf_mlp_set_params", C.c_int, [C.c_void_p, f32p]),
    ("nf_mlp_set_norm", C.c_int, [C.c_void_p, f32p, f32p, f32p, f32p]),
    ("nf_mlp_get_norm", C.c_int, [C.c_void_p, f32p, f32p, f32p, f32p]),
    ("nf_mlp_predict", C.c_int, [C.c_void_p, f32p, C.c_int64, f32p]),
    ("nf_mlp_predict_noalloc", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    ("nf_mlp_loss_grad", C.c_int, [C.c_void_p, f32p, f32p, C.c_int64, f32p, f32p]),
    ("nf_mlp_train", C.c_int, [C.c_void_p, f32p, f32p, C.c_int64, C.POINTER(TrainOpts),
                               C.POINTER(TrainStats)]),
]

_lib = None


def lib():
    """Load libnetflux.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C netflux_amd` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
