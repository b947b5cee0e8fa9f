/*
 * netflux.h - C ABI of the MI355X explicit Euler solver (libnetflux.so).
 *
 * Every entry point takes plain pointers and sizes and returns an int status
 * (NF_OK = 0).  A solver context owns its device buffers and one HIP stream;
 * it is not thread-safe, like the reference solver object it replaces.
 * Arrays cross the boundary in the reference's own layouts:
 *   - per-cell vectors are AoS [cell][5] doubles (include/containers/
 *     DynamicContainer.hpp:147-157), in the reference's cell numbering
 *     (element order of the input mesh);
 *   - the gradient is [cell][5x3] column-major, offset k + 5*d
 *     (include/SolverData.hpp:83-86, Eigen default storage).
 */
#ifndef NETFLUX_H
#define NETFLUX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    NF_OK = 0,
    NF_E_ARG = 1,         /* bad argument / unsupported option              */
    NF_E_MESH = 2,        /* inconsistent mesh (unmatched face, bad patch)  */
    NF_E_HIP = 3,         /* HIP runtime error                              */
    NF_E_INVALID_DT = 4,  /* "Invalid dt calculated" (src/Solver.cpp:267)   */
    NF_E_NONPHYSICAL = 5, /* "cells with unphysical values detected in the
                              flux_balance field" (src/SolverData.cpp:73-78) */
    NF_E_NOGPU = 6        /* no HIP device available                        */
};

/* ---- option enums: same meaning as include/Utilities.hpp:26-98 ---------- */
enum { NF_EXPLICIT_EULER = 0, NF_TVD_RK3 = 1 };           /* TimeScheme      */
enum { NF_FIRST_ORDER = 0, NF_SECOND_ORDER = 1 };         /* SpatialOrder    */
enum { NF_LIMITER_NONE = 0, NF_LIMITER_BARTH = 1 };       /* Limiter         */
enum { NF_FLUX_RUSANOV = 0, NF_FLUX_HLLC = 1 };           /* InviscidFluxScheme */
enum { NF_BC_NOSLIP_WALL = 0, NF_BC_SLIP_WALL = 1, NF_BC_FARFIELD = 2 }; /* BoundaryType */
/* arithmetic of the face sweeps */
enum { NF_PREC_FP64 = 0,   /* everything FP64, bit-identical op order       */
       NF_PREC_MIXED = 1   /* FP32 flux/gradient/limiter, FP64 state/RK/norms */ };
/* cell numbering used on the device */
enum { NF_ORDER_INPUT = 0, NF_ORDER_RCM = 1, NF_ORDER_MORTON = 2 };
/* element types: VTK numbering, as ElementType (include/geometry/
 * GeometryUtilities.hpp:8-16) */
enum { NF_ELEM_TRIANGLE = 5, NF_ELEM_QUAD = 9, NF_ELEM_TETRA = 10, NF_ELEM_HEXA = 12,
       NF_ELEM_WEDGE = 13, NF_ELEM_PYRAMID = 14 };

/* ---- primal mesh: what PrimalGrid holds after read_su2_mesh ------------
 * (src/geometry/PrimalGrid.cpp:140-244, after salome_to_vtk_connectivity).
 * Tetrahedra in VTK node order; boundary triangles grouped by patch, patch
 * order = SU2 marker order, which fixes face and ghost numbering
 * (src/geometry/FV_Grid.cpp:93-115). */
typedef struct nf_mesh_desc {
    uint32_t n_nodes;
    const double *nodes;          /* [n_nodes][3]                          */
    uint32_t n_cells;
    const uint32_t *cells;        /* [n_cells][4] tetrahedra, VTK order    */
    uint32_t n_patches;
    const uint32_t *patch_n_faces;/* [n_patches] triangles per patch       */
    const uint32_t *patch_faces;  /* [sum patch_n_faces][3]                */
    const int32_t *patch_bc;      /* [n_patches] NF_BC_*                   */
    /* Mixed-element meshes (Elements, include/geometry/GeometryUtilities.hpp:
     * 123-160): all four NULL for the all-tetrahedron layout above.  Else
     * cell i has type cell_types[i] (NF_ELEM_TETRA/HEXA/WEDGE/PYRAMID) and
     * nodes cells[cell_offsets[i] .. cell_offsets[i+1]) in VTK order, and
     * patch face e has type patch_face_types[e] (NF_ELEM_TRIANGLE/QUAD) and
     * nodes patch_faces[patch_face_offsets[e] .. [e+1]).  A mesh of only
     * tetrahedra and triangles given this way is stored in the layout above. */
    const uint8_t *cell_types;        /* [n_cells]                         */
    const uint32_t *cell_offsets;     /* [n_cells + 1]                     */
    const uint8_t *patch_face_types;  /* [sum patch_n_faces]               */
    const uint32_t *patch_face_offsets; /* [sum patch_n_faces + 1]         */
} nf_mesh_desc;

/* Mirrors the Config keys read by src/ConfigParser.cpp:38-83 that the step
 * depends on, plus the reference-compatibility switch for the zero U_old of
 * the first TVD-RK3 step (src/SolverData.cpp:9, src/Solver.cpp:51). */
typedef struct nf_params {
    int32_t time_scheme;          /* NF_EXPLICIT_EULER | NF_TVD_RK3        */
    int32_t spatial_order;        /* NF_FIRST_ORDER | NF_SECOND_ORDER      */
    int32_t limiter;              /* NF_LIMITER_*                          */
    int32_t inv_flux_scheme;      /* NF_FLUX_RUSANOV, or NF_FLUX_HLLC with */
                                  /* NF_PREC_FP64 (Toro's scheme: the      */
                                  /* reference's is a stub)                */
    double CFL;
    double primvars_inf[5];       /* rho, u, v, w, p of the freestream     */
    int32_t check_physical_validity;
    int32_t rk3_zero_old_first_step; /* 1 = reference behaviour            */
    int32_t precision;            /* NF_PREC_*                             */
    int32_t cell_order;           /* NF_ORDER_*                            */
    /* 1: each cell advances with its own CFL step (steady-state runs).  The
     * reference always takes the global minimum (src/Solver.cpp:252-265). */
    int32_t local_time_step;
} nf_params;

/* ---- host mesh utilities (CPU only) ------------------------------------ */
typedef struct nf_mesh nf_mesh;   /* owned primal mesh                     */

/* Unit-box hexahedral grid [0,lx]x[0,ly]x[0,lz], nx*ny*nz hexes, each split
 * into 6 tetrahedra around the (0,0,0)-(1,1,1) diagonal.  Interior nodes are
 * moved by a uniform random offset in +-jitter*h per coordinate (splitmix64
 * seeded by `seed`).  bump_h > 0 maps y <- yb(x) + y (1 - yb(x)/ly) with
 * yb(x) = bump_h exp(-(x - lx/2)^2 / (2 bump_sigma^2)).
 * Six patches in the order x-, x+, y-, y+, z-, z+ with types patch_bc[6]. */
int nf_mesh_box(uint32_t nx, uint32_t ny, uint32_t nz,
                double lx, double ly, double lz,
                double jitter, uint64_t seed,
                double bump_h, double bump_sigma,
                const int32_t patch_bc[6], nf_mesh **out);
int nf_mesh_from_desc(const nf_mesh_desc *desc, nf_mesh **out);
int nf_mesh_get_desc(const nf_mesh *mesh, nf_mesh_desc *desc);
void nf_mesh_destroy(nf_mesh *mesh);
/* SU2 mesh file (PrimalGrid::read_su2_mesh, src/geometry/PrimalGrid.cpp:
 * 140-244): tetrahedra (10), hexahedra (12), wedges (13) and pyramids (14)
 * in Salome node order, stored in VTK order as src/geometry/
 * GeometryUtilities.cpp:443-466 does (tetrahedron nodes 1<->2, hexahedron
 * 1<->3 and 5<->7, pyramid 1<->3); triangle (5) and quadrilateral (9)
 * markers as read.  Only tetrahedra and triangles give the tetrahedral
 * layout of nf_mesh_desc, anything else the mixed-element fields.  Marker i
 * takes the boundary type bcs[j] of names[j] == its MARKER_TAG; a marker the
 * list does not name is an error, as in the reference.  Patch order = marker
 * order. */
int nf_mesh_read_su2(const char *path, int32_t n_names, const char *const *names, const int32_t *bcs,
                     nf_mesh **out);
/* Writes the elements in Salome order and one marker per patch, coordinates
 * with 17 significant digits; marker tags from names (or as read, or
 * "patch_<i>"). */
int nf_mesh_write_su2(const nf_mesh *mesh, const char *path, int32_t n_names, const char *const *names);
int nf_mesh_patch_name(const nf_mesh *mesh, int32_t patch, char *buf, int32_t len);
const char *nf_mesh_last_error(void);   /* text of the last NF_E_MESH of these calls */

/* The finite-volume face graph built exactly as FV_Grid does
 * (src/geometry/FV_Grid.cpp:21-267): interior faces sorted by (i,j), i<j,
 * then patch blocks; S oriented i->j; ghost volume 0, ghost centroid
 * reflected.  Sizes first, then caller-allocated copies. */
typedef struct nf_fv_sizes {
    uint32_t n_interior_cells, n_total_cells;
    uint32_t n_interior_faces, n_total_faces;
    uint32_t n_patches;
} nf_fv_sizes;
int nf_fv_sizes_of(const nf_mesh *mesh, nf_fv_sizes *sizes);
int nf_fv_export(const nf_mesh *mesh,
                 uint32_t *cell_ij,      /* [F][2]          */
                 double *S,              /* [F][3]          */
                 double *r_i,            /* [F][3]          */
                 double *r_j,            /* [F][3]          */
                 double *volume,         /* [N+G]           */
                 double *centroid,       /* [N+G][3]        */
                 uint32_t *patch_first,  /* [P]             */
                 uint32_t *patch_nfaces, /* [P]             */
                 double *delta_S);       /* [N+G][3]        */

/* ---- solver context ------------------------------------------------------
 * Replaces EulerSolver (include/Solver.hpp:52-79): constructed from the mesh
 * and config, stepped with calc_timestep + step in the Driver loop order
 * (src/Driver.cpp:31-57).  A mixed-element mesh (hexahedra, wedges,
 * pyramids) runs on the polyhedral FP64 path: one mesh, global or local time
 * steps, and the calls set_freestream, set_state, calc_timestep, step, run
 * (calc_timestep + step kept on the device, check_every steps per HIP
 * graph replay, the host synchronising between replays), get_residuals, time_kernels, eval_flux_balance, get_state,
 * get_field and ctx_sizes; the others return NF_E_ARG with a message, and
 * nf_create rejects NF_PREC_MIXED for such a mesh. */
typedef struct nf_ctx nf_ctx;

int nf_create(const nf_mesh_desc *mesh, const nf_params *params, nf_ctx **out);
int nf_destroy(nf_ctx *ctx);
const char *nf_last_error(const nf_ctx *ctx);

/* EulerSolverData::set_freestream_values (src/SolverData.cpp:50-58):
 * U = U_inf, primvars = V_inf, U_old = 0. */
int nf_set_freestream(nf_ctx *ctx);
/* Inject a state (AoS [N][5] conservative, reference cell order).  primvars
 * become prim(U).  U_old = U_old_or_null, or zero when NULL. */
int nf_set_state(nf_ctx *ctx, const double *U, const double *U_old_or_null);

/* EulerSolver::calc_timestep (src/Solver.cpp:232-274): dt from the primvars
 * as last set by the solver; NF_E_INVALID_DT if not finite and positive. */
int nf_calc_timestep(nf_ctx *ctx, double *dt_out);
/* Solver::step (src/Solver.cpp:34-52) with the given dt. */
int nf_step(nf_ctx *ctx, double dt);

/* Driver loop on the device: n_steps of (calc_timestep, step), dt kept on
 * the device, launched as a captured HIP graph.  If residual_drop > 0 the
 * loop stops after the first step k with r_k / r_1 <= residual_drop, where
 * r_k = ||R_rho||_2 over interior cells at the first stage of step k;
 * convergence is checked every check_every steps. */
typedef struct nf_run_opts {
    int32_t n_steps;
    double residual_drop;         /* <= 0: run all steps                   */
    int32_t check_every;          /* steps per graph replay (>=1)          */
    int32_t use_graph;            /* 1: hipGraph replay, 0: eager launches */
    /* stall fallback for warm-started instances (0: off): once per warm
     * start, at the first check between replays that is at least
     * stall_window steps after it, an instance whose residual did not fall
     * below stall_ratio times its value stall_window steps earlier restarts
     * from freestream; otherwise the warm start is accepted */
    int32_t stall_window;
    double stall_ratio;
} nf_run_opts;
typedef struct nf_run_stats {
    int32_t steps_done;
    int32_t converged;
    double time;                  /* sum of dt                             */
    double last_dt;
    double res_first, res_last;   /* r_1 and r_k                           */
    double elapsed_ms;            /* hipEvent time of the loop             */
    int32_t fell_back;            /* warm start replaced by the freestream */
    int32_t pad;
} nf_run_stats;
int nf_run(nf_ctx *ctx, const nf_run_opts *opts, nf_run_stats *stats);
/* Residual history r_k of the last nf_run / nf_step calls (up to max). */
int nf_get_residuals(nf_ctx *ctx, double *out, int32_t max, int32_t *n_out);

/* Solver::evaluate_flux_balance (src/Solver.cpp:54-81) of the current state,
 * once, without changing U: keeps R, the gradient and the limiter for
 * nf_get_field (the reference computes them inside every stage). */
int nf_eval_flux_balance(nf_ctx *ctx);

/* Read back (synchronising). */
int nf_get_state(nf_ctx *ctx, double *U);            /* [N][5] conservative */
enum { NF_FIELD_PRIM = 0,      /* [N][5] prim(U) of the current state      */
       NF_FIELD_LIMITER = 1,   /* [N][5] psi of the last stage             */
       NF_FIELD_GRADIENT = 2,  /* [N][15] column-major 5x3 of the last stage */
       NF_FIELD_RESIDUAL = 3,  /* [N][5] flux balance R of the last stage  */
       NF_FIELD_STATE_OLD = 4  /* [N][5] U_old                             */ };
int nf_get_field(nf_ctx *ctx, int32_t which, double *out);

/* Per-kernel timing of n_steps eager steps with HIP events recorded on the
 * context stream.  names/avg_ms/calls are filled for up to max kernels. */
typedef struct nf_kernel_time {
    char name[48];
    double avg_ms;
    int32_t calls;
} nf_kernel_time;
int nf_time_kernels(nf_ctx *ctx, int32_t n_steps, nf_kernel_time *out,
                    int32_t max, int32_t *n_out);

/* Sizes of a context (cells, faces) for callers sizing their buffers. */
int nf_ctx_sizes(const nf_ctx *ctx, nf_fv_sizes *sizes);

/* ---- batches: several meshes in one context -------------------------------
 * No counterpart in the reference, whose Driver holds one solver
 * (src/Driver.cpp:36 asserts solvers.size() == 1).  All instances share the
 * options of params[0] except the freestream (primvars_inf), which is per
 * instance; each instance has its own time step (global per instance, or
 * local), residual history, first-step flag and convergence state, and every
 * sweep is one launch over all instances.  The single-mesh functions above
 * act on a context of one instance. */
int nf_create_batch(int32_t n, const nf_mesh_desc *meshes, const nf_params *params, nf_ctx **out);
int nf_ctx_instances(const nf_ctx *ctx, int32_t *n);
int nf_inst_sizes(const nf_ctx *ctx, int32_t inst, nf_fv_sizes *sizes);
int nf_inst_set_state(nf_ctx *ctx, int32_t inst, const double *U, const double *U_old_or_null);
int nf_inst_get_state(nf_ctx *ctx, int32_t inst, double *U);
int nf_inst_residuals(nf_ctx *ctx, int32_t inst, double *out, int32_t max, int32_t *n_out);
/* Runs every instance until it converges (residual_drop) or n_steps steps
 * are done; stats[n] per instance, elapsed_ms is the batch wall time.  With
 * a residual drop, an instance that converged in an earlier call stays
 * stopped (runs can continue in chunks) until its state is set again. */
int nf_run_batch(nf_ctx *ctx, const nf_run_opts *opts, nf_run_stats *stats);

/* The drop of instance `inst` is measured against `ref` instead of its first
 * step's residual (ref <= 0 restores that); kept across state changes. */
int nf_inst_set_residual_ref(nf_ctx *ctx, int32_t inst, double ref);
/* ||R_rho||_2 of every instance's current state (no update); out[n] */
int nf_eval_residuals(nf_ctx *ctx, double *out);

/* Autotuner: times `steps` steps for each launch configuration (threads per
 * block 128 or 256, XCD-contiguous block order on or off) from a copy of the
 * current state, restores the state and keeps the fastest configuration. */
typedef struct nf_tune_result {
    int32_t block;
    int32_t xcd_remap;
    double ms_per_step;
    int32_t n_tried;
    double tried_ms[8];           /* per candidate, in the order tried      */
} nf_tune_result;
int nf_autotune(nf_ctx *ctx, int32_t steps, nf_tune_result *out);
int nf_set_launch(nf_ctx *ctx, int32_t block, int32_t xcd_remap);
/* Switches the face-sweep arithmetic of a context (NF_PREC_*), keeping the
 * FP64 states, run counters and residual histories: e.g. mixed precision
 * down to its FP32 noise floor, then FP64 to a tighter drop.  Primvars and
 * dt are recomputed from the states; converged flags are cleared. */
int nf_set_precision(nf_ctx *ctx, int32_t precision);

/* ---- learned warm start ---------------------------------------------------
 * MLP 6 -> 128 -> 128 -> 128 -> 5 with GELU: per-cell features (Mach, bump
 * height, centroid x, y, z, wall distance) -> primitive variables
 * (rho, u, v, w, p).  BF16 weights and activations as MFMA operands, FP32
 * accumulation, FP32 master weights and Adam state.  No counterpart in the
 * reference, whose only initial condition is the freestream
 * (src/SolverData.cpp:27-34). */
typedef struct nf_mlp nf_mlp;
/* FP32 parameters, flattened W1 [128][6], b1, W2 [128][128], b2, W3, b3,
 * W4 [5][128], b4 (each W row-major [out][in]) */
int64_t nf_mlp_n_params(void);
int nf_mlp_create(uint64_t seed, nf_mlp **out);
int nf_mlp_destroy(nf_mlp *m);
const char *nf_mlp_last_error(const nf_mlp *m);
int nf_mlp_get_params(nf_mlp *m, float *params);
int nf_mlp_set_params(nf_mlp *m, const float *params);   /* resets Adam state */
/* normalisation: x_n = (x - in_shift) * in_scale; y = y_n * out_scale + out_shift */
int nf_mlp_set_norm(nf_mlp *m, const float *in_shift, const float *in_scale, const float *out_shift,
                    const float *out_scale);
int nf_mlp_get_norm(nf_mlp *m, float *in_shift, float *in_scale, float *out_shift, float *out_scale);
/* raw features x [n][6] -> raw outputs y [n][5] (host arrays) */
int nf_mlp_predict(nf_mlp *m, const float *x, int64_t n, float *y);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops): nf_mlp_predict on
 * DEVICE arrays x [n][6] -> y [n][5], at most 16384 samples per call (more:
 * NF_E_ARG), enqueued on `stream` (a hipStream_t; null: the network's own
 * stream) without synchronising and without allocating.  The activations
 * live in a fixed device array of about 17 MB per GPU, shared by every call
 * on that GPU: calls that overlap on one GPU race.  The device that was
 * current at nf_mlp_create must be current.  Same results as nf_mlp_predict,
 * bit for bit. */
int nf_mlp_predict_noalloc(nf_mlp *m, const float *x, int64_t n, float *y, void *stream);
/* mean squared error of the normalised outputs over one batch of n <= 2^20
 * samples and its gradient with respect to the FP32 parameters */
int nf_mlp_loss_grad(nf_mlp *m, const float *x, const float *y, int64_t n, float *loss, float *grad);
typedef struct nf_train_opts {
    int32_t epochs;
    int32_t batch;                /* multiple of 256                        */
    float lr, beta1, beta2, eps;  /* Adam                                   */
    float lr_decay;               /* per-epoch factor (0 or 1: constant)    */
    uint64_t seed;                /* minibatch shuffling                    */
} nf_train_opts;
typedef struct nf_train_stats {
    float loss_first, loss_last;  /* first minibatch, last epoch's last one */
    int64_t steps;
    double elapsed_ms;
} nf_train_stats;
/* Adam over shuffled minibatches of the n samples x [n][6], y [n][5]; the
 * normalisation, unless set, becomes the data's mean and standard deviation */
int nf_mlp_train(nf_mlp *m, const float *x, const float *y, int64_t n, const nf_train_opts *opts,
                 nf_train_stats *stats);
/* Network features of instance `inst`: (case_params[0], case_params[1],
 * centroid x, y, z, distance to the nearest wall-face centroid), x [N][6] in
 * the instance's reference cell order (wall = slip or no-slip patch). */
int nf_inst_features(nf_ctx *ctx, int32_t inst, const double *case_params, float *x);
/* Warm start of instance `inst`: the network's primitive-variable prediction
 * from the features above becomes U (U_old = U).  If any cell has rho <= 0,
 * p <= 0 or a non-finite value, the instance starts from the freestream
 * instead and *fell_back = 1. */
int nf_inst_warm_start(nf_ctx *ctx, int32_t inst, nf_mlp *net, const double *case_params, int32_t *fell_back);

#ifdef __cplusplus
}
#endif
#endif
