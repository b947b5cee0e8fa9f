// Solver context and C ABI (include/netflux.h).
//
// The context replaces EulerSolver + EulerSolverData (include/Solver.hpp:52-79,
// include/SolverData.hpp:88-117): it owns the device fields and runs the
// Driver loop order (calc_timestep, then step; src/Driver.cpp:38-47) with dt
// kept on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/netflux.h"
#include "kernels_fp64.hpp"
#include "kernels_mixed.hpp"
#include "mesh.hpp"

using namespace nf;

namespace {

// First level of the per-block partial reductions: block b reduces the
// contiguous chunk [b*RCHUNK, (b+1)*RCHUNK) in a fixed order.
constexpr int RCHUNK = 2048;
template <bool MIN>
__global__ __launch_bounds__(BLOCK) void k_reduce(const double *in, int n, double *out, const int *stop) {
    __shared__ double sh[BLOCK / 64];
    if (*stop) return;
    const int base = blockIdx.x * RCHUNK + threadIdx.x * (RCHUNK / BLOCK);
    double v = MIN ? DBL_MAX : 0.0;
    for (int q = 0; q < RCHUNK / BLOCK; ++q) {
        const int i = base + q;
        if (i < n) v = MIN ? ((in[i] < v) ? in[i] : v) : v + in[i];
    }
    v = block_reduce<MIN>(v, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = v;
}

__global__ void k_finalize_dt(const double *part, int nb, double *dt, double *time_acc,
                              int *nstep, int *err, int *stop, int advance) {
    __shared__ double sh[BLOCK / 64];
    if (advance && *stop) return;
    double v = DBL_MAX;
    for (int b = threadIdx.x; b < nb; b += BLOCK) v = (part[b] < v) ? part[b] : v;
    v = block_reduce<true>(v, sh);
    if (threadIdx.x == 0) {
        if (advance) {
            *time_acc += *dt;
            *nstep += 1;
        }
        // num_is_valid_and_pos (include/Includes.hpp:131-137)
        const bool ok = (v - v == 0.0) && v > 0.0;
        *dt = v;
        if (!ok) {
            *err = NF_E_INVALID_DT;
            *stop = 1;
        }
    }
}

// r_k = sqrt(sum of per-block partials), summed in a fixed order.
__global__ void k_finalize_res(const double *part, int nb, double *hist, int hist_cap,
                               const int *nstep, int *stop, double drop, int *converged,
                               int check_now) {
    __shared__ double sh[BLOCK / 64];
    if (*stop) return;
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += BLOCK) v += part[b];
    v = block_reduce<false>(v, sh);
    if (threadIdx.x == 0) {
        const int k = *nstep;
        const double r = sqrt(v);
        if (k < hist_cap) hist[k] = r;
        const double r1 = hist[0];
        if (check_now && drop > 0.0 && k < hist_cap && r1 > 0.0 && r / r1 <= drop) *converged = 1;
    }
}

__global__ void k_stop_if_converged(const int *converged, int *stop) {
    if (*converged) *stop = 1;
}

// flags [first, first+n) = 0 and, if set_dt, scal[0] = dt
__global__ void k_set_flags(int *iscal, int first, int n, double *scal, int set_dt, double dt,
                            int clear_scal) {
    const int t = threadIdx.x;
    if (t < n) iscal[first + t] = 0;
    if (t == 0 && set_dt) scal[0] = dt;
    if (t < 2 && clear_scal) scal[t] = 0.0;
}


struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct nf_ctx {
    nf_params prm{};
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t N = 0, n_nodes = 0;
    nf_fv_sizes sizes{};
    std::vector<uint32_t> perm, iperm;  // device <-> reference cell
    FarField<double> fs{};
    int nb = 0;          // blocks of a sweep
    uint32_t npad = 0;   // N rounded up to whole 64-cell tiles
    uint32_t nbf = 0;    // boundary faces
    uint32_t *bface = nullptr;
    uint8_t *bbc = nullptr;
    double *Gc = nullptr, *Gf = nullptr;  // ghost states [nbf][5]
    bool mixed = false;                   // NF_PREC_MIXED: FP32 copies below
    float *V32 = nullptr, *FV32 = nullptr, *Gc32 = nullptr, *Gf32 = nullptr;
    double *dt_cell = nullptr;            // local time stepping: per-cell step
    FarField<float> fs32{};
    int xcd_remap = 1;
    // device arrays
    double *nodes = nullptr;
    uint32_t *cnodes = nullptr, *nbr = nullptr, *fcode = nullptr;
    double *Ua = nullptr, *Ub = nullptr, *V = nullptr, *FV = nullptr;
    double *grad_aux = nullptr, *psi_aux = nullptr, *R_aux = nullptr;
    double *dt_part = nullptr, *res_part = nullptr, *red2 = nullptr;
    int nb2 = 0;  // second-level partials
    double *scal = nullptr;  // [0] dt, [1] time
    int *iscal = nullptr;    // [0] nstep [1] err [2] stop [3] converged [4] bad_flux [5] bad_state
    double *hist = nullptr;
    int hist_cap = 0;
    bool uold_zero = true;   // U_old is all zeros (before the first step)
    bool dt_valid = false;   // scal[0] holds the dt of the primvars as last set
    bool aux_valid = false;
    std::vector<void *> owned;
    // graph cache
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    double graph_drop = -1.0;
    // timing
    bool timing = false;
    std::map<std::string, std::pair<double, int>> ktime;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
};

namespace {

int fail(nf_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPC(expr)                                                                     \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(ctx, NF_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int dalloc(nf_ctx *ctx, T **p, size_t n) {
    void *q = nullptr;
    if (n == 0) n = 1;
    HIPC(hipMalloc(&q, n * sizeof(T)));
    ctx->owned.push_back(q);
    *p = static_cast<T *>(q);
    return NF_OK;
}

DevMesh devmesh(const nf_ctx *c) { return DevMesh{c->N, 3.0, 6.0, c->nodes, c->cnodes, c->nbr, c->fcode}; }

// kernel launch with optional per-kernel event timing
template <typename K, typename... Args>
int launch(nf_ctx *ctx, const char *name, K kernel, dim3 grid, dim3 block, Args... args) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->timing) {
        HIPC(hipEventCreate(&e0));
        HIPC(hipEventCreate(&e1));
        HIPC(hipEventRecord(e0, ctx->stream));
    }
    hipLaunchKernelGGL(kernel, grid, block, 0, ctx->stream, args...);
    HIPC(hipGetLastError());
    if (ctx->timing) {
        HIPC(hipEventRecord(e1, ctx->stream));
        ctx->pending.push_back({name, {e0, e1}});
    }
    return NF_OK;
}

// same as launch() for a launcher callable returning hipError_t
template <typename F>
int launch_fn(nf_ctx *ctx, const char *name, F fn) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->timing) {
        HIPC(hipEventCreate(&e0));
        HIPC(hipEventCreate(&e1));
        HIPC(hipEventRecord(e0, ctx->stream));
    }
    HIPC(fn());
    if (ctx->timing) {
        HIPC(hipEventRecord(e1, ctx->stream));
        ctx->pending.push_back({name, {e0, e1}});
    }
    return NF_OK;
}

int flush_timing(nf_ctx *ctx) {
    HIPC(hipStreamSynchronize(ctx->stream));
    for (auto &p : ctx->pending) {
        float ms = 0;
        HIPC(hipEventElapsedTime(&ms, p.second.first, p.second.second));
        auto &slot = ctx->ktime[p.first];
        slot.first += ms;
        slot.second += 1;
        (void)hipEventDestroy(p.second.first);
        (void)hipEventDestroy(p.second.second);
    }
    ctx->pending.clear();
    return NF_OK;
}

SweepAArgs argsA(nf_ctx *ctx, bool dt, bool aux, bool dt_cell) {
    SweepAArgs a{};
    a.V = ctx->V;
    a.Gc = ctx->Gc;
    a.FV = ctx->FV;
    a.dt_part = dt ? ctx->dt_part : nullptr;
    a.CFL = ctx->prm.CFL;
    a.stop = ctx->iscal + 2;
    a.xcd_remap = ctx->xcd_remap;
    a.grad_aux = aux ? ctx->grad_aux : nullptr;
    a.psi_aux = aux ? ctx->psi_aux : nullptr;
    a.dt_cell = dt_cell ? ctx->dt_cell : nullptr;
    return a;
}

int run_ghost(nf_ctx *ctx, bool face) {
    if (ctx->nbf == 0) return NF_OK;
    dim3 g((ctx->nbf + BLOCK - 1) / BLOCK), b(BLOCK);
    if (ctx->mixed) {
        mx::GhostArgsMx a{ctx->bface, ctx->bbc, ctx->nbf, face ? ctx->FV32 : ctx->V32, face ? ctx->Gf32 : ctx->Gc32,
                          ctx->fs32, ctx->iscal + 2};
        return launch_fn(ctx, face ? "ghost_face_mx" : "ghost_cell_mx",
                         [&] { return mx::launch_ghost_mx(face, g, ctx->stream, devmesh(ctx), a); });
    }
    GhostArgs a{ctx->bface, ctx->bbc, ctx->nbf, face ? ctx->FV : ctx->V, face ? ctx->Gf : ctx->Gc, ctx->fs,
                ctx->iscal + 2};
    if (face) return launch(ctx, "ghost_face", k_ghost<true>, g, b, devmesh(ctx), a);
    return launch(ctx, "ghost_cell", k_ghost<false>, g, b, devmesh(ctx), a);
}

int run_sweepA(nf_ctx *ctx, bool dt, bool aux, bool dt_cell = false) {
    int rc;
    if ((rc = run_ghost(ctx, false))) return rc;
    const DevMesh m = devmesh(ctx);
    dim3 g(ctx->nb), b(BLOCK);
    if (ctx->mixed) {
        mx::SweepAArgsMx x{ctx->V32, ctx->Gc32, ctx->FV32, dt ? ctx->dt_part : nullptr, ctx->prm.CFL,
                           ctx->iscal + 2, ctx->xcd_remap, aux ? ctx->grad_aux : nullptr,
                           aux ? ctx->psi_aux : nullptr, dt_cell ? ctx->dt_cell : nullptr};
        const int order = ctx->prm.spatial_order == NF_FIRST_ORDER ? 0 : (ctx->prm.limiter == NF_LIMITER_NONE ? 1 : 2);
        return launch_fn(ctx, order == 2 ? "sweepA_barth_mx" : "sweepA_mx",
                         [&] { return mx::launch_sweepA_mx(order, g, ctx->stream, m, x); });
    }
    SweepAArgs a = argsA(ctx, dt, aux, dt_cell);
    if (ctx->prm.spatial_order == NF_FIRST_ORDER)
        return launch(ctx, "sweepA_first", k_sweepA<0>, g, b, m, a);
    if (ctx->prm.limiter == NF_LIMITER_NONE)
        return launch(ctx, "sweepA_nolim", k_sweepA<1>, g, b, m, a);
    return launch(ctx, "sweepA_barth", k_sweepA<2>, g, b, m, a);
}

int run_sweepB(nf_ctx *ctx, int kind, const double *U, const double *Uo, double *Uout, bool res,
               bool aux) {
    int rc;
    if ((rc = run_ghost(ctx, true))) return rc;
    if (ctx->mixed) {
        mx::SweepBArgsMx x{ctx->FV32, ctx->Gf32, U, Uo, Uout, ctx->V32, ctx->scal, kind,
                           res ? ctx->res_part : nullptr, ctx->iscal + 4, ctx->prm.check_physical_validity,
                           ctx->iscal + 5, ctx->iscal + 2, ctx->xcd_remap, aux ? ctx->R_aux : nullptr,
                           ctx->prm.local_time_step ? ctx->dt_cell : nullptr};
        return launch_fn(ctx, "sweepB_mx",
                         [&] { return mx::launch_sweepB_mx(dim3(ctx->nb), ctx->stream, devmesh(ctx), x); });
    }
    SweepBArgs a{};
    a.FV = ctx->FV;
    a.Gf = ctx->Gf;
    a.U = U;
    a.Uo = Uo;
    a.Uout = Uout;
    a.Vout = ctx->V;
    a.dt = ctx->scal;
    a.kind = kind;
    a.res_part = res ? ctx->res_part : nullptr;
    a.bad_flux = ctx->iscal + 4;
    a.check = ctx->prm.check_physical_validity;
    a.bad_state = ctx->iscal + 5;
    a.stop = ctx->iscal + 2;
    a.xcd_remap = ctx->xcd_remap;
    a.R_aux = aux ? ctx->R_aux : nullptr;
    a.dt_cell = ctx->prm.local_time_step ? ctx->dt_cell : nullptr;
    return launch(ctx, "sweepB", k_sweepB, dim3(ctx->nb), dim3(BLOCK), devmesh(ctx), a);
}

// One solver step on the device (Solver::step, src/Solver.cpp:34-52).  The
// last sweep A of the step evaluates the CFL dt of the primvars it reads --
// prim(U2) for TVD-RK3, prim(U^n) for explicit Euler -- which is exactly the
// state calc_timestep sees next (src/Solver.cpp:248).  k_finalize_dt then
// advances time and installs that dt for the following step.
int enqueue_step(nf_ctx *ctx, bool zero_old, double drop, bool check_now) {
    int rc;
    const bool rk3 = ctx->prm.time_scheme == NF_TVD_RK3;
    const bool lts = ctx->prm.local_time_step;
    int *is = ctx->iscal;
    // local time stepping: per-cell steps from the state entering the step
    if ((rc = run_sweepA(ctx, !rk3 || lts, false, lts))) return rc;
    if ((rc = run_sweepB(ctx, 1, ctx->Ua, nullptr, rk3 ? ctx->Ub : ctx->Ua, true, false))) return rc;
    if ((rc = launch(ctx, "reduce_res", k_reduce<false>, dim3(ctx->nb2), dim3(BLOCK), (const double *)ctx->res_part,
                     ctx->nb, ctx->red2 + ctx->nb2, (const int *)(is + 2))))
        return rc;
    if ((rc = launch(ctx, "finalize_res", k_finalize_res, dim3(1), dim3(BLOCK), (const double *)(ctx->red2 + ctx->nb2),
                     ctx->nb2, ctx->hist, ctx->hist_cap, (const int *)is, is + 2, drop, is + 3,
                     (int)check_now)))
        return rc;
    if (rk3) {
        const double *Uo = zero_old ? nullptr : ctx->Ua;
        if ((rc = run_sweepA(ctx, false, false))) return rc;
        if ((rc = run_sweepB(ctx, 2, ctx->Ub, Uo, ctx->Ub, false, false))) return rc;
        if ((rc = run_sweepA(ctx, !lts, false))) return rc;
        if ((rc = run_sweepB(ctx, 3, ctx->Ub, Uo, ctx->Ua, false, false))) return rc;
    }
    if ((rc = launch(ctx, "reduce_dt", k_reduce<true>, dim3(ctx->nb2), dim3(BLOCK), (const double *)ctx->dt_part,
                     ctx->nb, ctx->red2, (const int *)(is + 2))))
        return rc;
    if ((rc = launch(ctx, "finalize_dt", k_finalize_dt, dim3(1), dim3(BLOCK), (const double *)ctx->red2,
                     ctx->nb2, ctx->scal, ctx->scal + 1, is, is + 1, is + 2, 1)))
        return rc;
    if (drop > 0.0)
        if ((rc = launch(ctx, "stop_if_converged", k_stop_if_converged, dim3(1), dim3(1),
                         (const int *)(is + 3), is + 2)))
            return rc;
    return NF_OK;
}

// dt of the current primvars with a standalone sweep A (no face values used)
int enqueue_timestep(nf_ctx *ctx) {
    int rc;
    if ((rc = run_sweepA(ctx, true, false))) return rc;
    if ((rc = launch(ctx, "reduce_dt", k_reduce<true>, dim3(ctx->nb2), dim3(BLOCK), (const double *)ctx->dt_part,
                     ctx->nb, ctx->red2, (const int *)(ctx->iscal + 2))))
        return rc;
    return launch(ctx, "finalize_dt", k_finalize_dt, dim3(1), dim3(BLOCK), (const double *)ctx->red2,
                  ctx->nb2, ctx->scal, ctx->scal + 1, ctx->iscal, ctx->iscal + 1, ctx->iscal + 2, 0);
}

int check_errors(nf_ctx *ctx) {
    int is[6];
    HIPC(hipMemcpyAsync(is, ctx->iscal, sizeof is, hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    if (is[1] == NF_E_INVALID_DT) {
        double dt;
        HIPC(hipMemcpy(&dt, ctx->scal, sizeof dt, hipMemcpyDeviceToHost));
        char buf[96];
        snprintf(buf, sizeof buf, "Invalid dt calculated (dt = %f)", dt);
        return fail(ctx, NF_E_INVALID_DT, buf);
    }
    if (is[4] > 0)
        return fail(ctx, NF_E_NONPHYSICAL,
                    std::to_string(is[4]) +
                        " cells with unphysical values detected in the flux_balance field of the Euler solver");
    return NF_OK;
}

int clear_flags(nf_ctx *ctx) {
    hipLaunchKernelGGL(k_set_flags, dim3(1), dim3(64), 0, ctx->stream, ctx->iscal, 1, 5, ctx->scal, 0, 0.0, 0);
    HIPC(hipGetLastError());
    return NF_OK;
}

int ensure_hist(nf_ctx *ctx, int need) {
    if (need <= ctx->hist_cap) return NF_OK;
    int cap = std::max(need, 2 * ctx->hist_cap + 1024);
    double *h = nullptr;
    HIPC(hipMalloc(&h, (size_t)cap * sizeof(double)));
    HIPC(hipMemsetAsync(h, 0, (size_t)cap * sizeof(double), ctx->stream));
    if (ctx->hist) {
        HIPC(hipMemcpyAsync(h, ctx->hist, (size_t)ctx->hist_cap * sizeof(double), hipMemcpyDeviceToDevice,
                            ctx->stream));
        HIPC(hipStreamSynchronize(ctx->stream));
        auto it = std::find(ctx->owned.begin(), ctx->owned.end(), (void *)ctx->hist);
        if (it != ctx->owned.end()) ctx->owned.erase(it);
        HIPC(hipFree(ctx->hist));
    }
    ctx->owned.push_back(h);
    ctx->hist = h;
    ctx->hist_cap = cap;
    if (ctx->graph) {  // graph baked the old pointer
        (void)hipGraphExecDestroy(ctx->graph);
        ctx->graph = nullptr;
    }
    return NF_OK;
}

// reference-order AoS [N][5] <-> device-order tiles [tile][5][64]
int upload_aos(nf_ctx *ctx, const double *aos, double *dev) {
    const uint32_t N = ctx->N;
    std::vector<double> d((size_t)5 * ctx->npad, 0.0);
    for (uint32_t c = 0; c < N; ++c)
        for (int k = 0; k < 5; ++k)
            d[((size_t)(c >> 6) * 5 + k) * 64 + (c & 63u)] = aos[5 * (size_t)ctx->perm[c] + k];
    HIPC(hipMemcpy(dev, d.data(), d.size() * sizeof(double), hipMemcpyHostToDevice));
    return NF_OK;
}

int download_tiled(nf_ctx *ctx, const double *dev, double *aos) {
    const uint32_t N = ctx->N;
    std::vector<double> d((size_t)5 * ctx->npad);
    HIPC(hipStreamSynchronize(ctx->stream));
    HIPC(hipMemcpy(d.data(), dev, d.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (uint32_t c = 0; c < N; ++c)
        for (int k = 0; k < 5; ++k)
            aos[5 * (size_t)ctx->perm[c] + k] = d[((size_t)(c >> 6) * 5 + k) * 64 + (c & 63u)];
    return NF_OK;
}

int download_tiled_f32(nf_ctx *ctx, const float *dev, double *aos) {
    const uint32_t N = ctx->N;
    std::vector<float> d((size_t)5 * ctx->npad);
    HIPC(hipStreamSynchronize(ctx->stream));
    HIPC(hipMemcpy(d.data(), dev, d.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (uint32_t c = 0; c < N; ++c)
        for (int k = 0; k < 5; ++k)
            aos[5 * (size_t)ctx->perm[c] + k] = d[((size_t)(c >> 6) * 5 + k) * 64 + (c & 63u)];
    return NF_OK;
}

// device-order AoS [N][rows] -> reference-order AoS
int download_aos(nf_ctx *ctx, const double *dev, int rows, double *aos) {
    const uint32_t N = ctx->N;
    std::vector<double> d((size_t)rows * N);
    HIPC(hipStreamSynchronize(ctx->stream));
    HIPC(hipMemcpy(d.data(), dev, d.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (uint32_t c = 0; c < N; ++c)
        std::memcpy(aos + (size_t)rows * ctx->perm[c], &d[(size_t)rows * c], rows * sizeof(double));
    return NF_OK;
}

int after_state_set(nf_ctx *ctx) {
    if (ctx->mixed) {
        HIPC(mx::launch_prim_mx(dim3((ctx->N + BLOCK - 1) / BLOCK), ctx->stream, ctx->Ua, ctx->V32, ctx->N));
    } else {
        hipLaunchKernelGGL(k_prim, dim3((ctx->N + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, ctx->stream, ctx->Ua,
                           ctx->V, ctx->N);
        HIPC(hipGetLastError());
    }
    hipLaunchKernelGGL(k_set_flags, dim3(1), dim3(64), 0, ctx->stream, ctx->iscal, 0, 8, ctx->scal, 0, 0.0, 1);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(ctx->stream));
    ctx->dt_valid = false;
    ctx->aux_valid = false;
    return NF_OK;
}

}  // namespace

// ============================================================================
extern "C" {

int nf_create(const nf_mesh_desc *desc, const nf_params *prm, nf_ctx **out) {
    if (!desc || !prm || !out) return NF_E_ARG;
    *out = nullptr;
    if (prm->inv_flux_scheme != NF_FLUX_RUSANOV) return NF_E_ARG;  // HLLC: assert(false) in the reference
    if (prm->time_scheme != NF_TVD_RK3 && prm->time_scheme != NF_EXPLICIT_EULER) return NF_E_ARG;
    if (prm->precision != NF_PREC_FP64 && prm->precision != NF_PREC_MIXED) return NF_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return NF_E_NOGPU;
    std::unique_ptr<nf_ctx> hold(new nf_ctx);
    nf_ctx *ctx = hold.get();
    ctx->prm = *prm;
    HIPC(hipGetDevice(&ctx->device));
    nf::PrimalMesh pm;
    nf::FVGrid g;
    nf::CellLayout L;
    try {
        pm.nodes.assign(desc->nodes, desc->nodes + 3 * (size_t)desc->n_nodes);
        pm.cells.assign(desc->cells, desc->cells + 4 * (size_t)desc->n_cells);
        pm.patch_n.assign(desc->patch_n_faces, desc->patch_n_faces + desc->n_patches);
        size_t nt = 0;
        for (uint32_t p = 0; p < desc->n_patches; ++p) nt += desc->patch_n_faces[p];
        pm.patch_tri.assign(desc->patch_faces, desc->patch_faces + 3 * nt);
        pm.patch_bc.assign(desc->patch_bc, desc->patch_bc + desc->n_patches);
        for (int32_t b : pm.patch_bc)
            if (b < 0 || b > 2) return NF_E_ARG;
        nf::build_fv_grid(pm, g);
        nf::build_cell_layout(pm, g, prm->cell_order, L);
    } catch (const nf::MeshError &e) {
        return NF_E_MESH;
    }
    ctx->N = g.N;
    ctx->n_nodes = desc->n_nodes;
    ctx->sizes = {g.N, g.NT, g.Fint, g.F, (uint32_t)g.pfirst.size()};
    ctx->perm = L.perm;
    ctx->iperm = L.iperm;
    ctx->fs = make_farfield<double>(prm->primvars_inf);
    ctx->fs32 = make_farfield<float>(prm->primvars_inf);
    ctx->mixed = prm->precision == NF_PREC_MIXED;
    ctx->nb = (int)((g.N + BLOCK - 1) / BLOCK);
    ctx->npad = (g.N + TILE - 1) / TILE * TILE;
    HIPC(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    const uint32_t N = g.N;
    int rc;
    if ((rc = dalloc(ctx, &ctx->nodes, 4 * (size_t)ctx->n_nodes))) return rc;
    if ((rc = dalloc(ctx, &ctx->cnodes, 4 * (size_t)N))) return rc;
    if ((rc = dalloc(ctx, &ctx->nbr, 4 * (size_t)N))) return rc;
    if ((rc = dalloc(ctx, &ctx->fcode, (size_t)N))) return rc;
    const size_t NP = ctx->npad;
    if ((rc = dalloc(ctx, &ctx->Ua, 5 * NP))) return rc;
    if ((rc = dalloc(ctx, &ctx->Ub, 5 * NP))) return rc;
    if (ctx->mixed) {
        if ((rc = dalloc(ctx, &ctx->V32, 5 * NP))) return rc;
        if ((rc = dalloc(ctx, &ctx->FV32, 20 * NP))) return rc;
    } else {
        if ((rc = dalloc(ctx, &ctx->V, 5 * NP))) return rc;
        if ((rc = dalloc(ctx, &ctx->FV, 20 * NP))) return rc;
    }
    if (prm->local_time_step && (rc = dalloc(ctx, &ctx->dt_cell, NP))) return rc;
    ctx->nbf = (uint32_t)L.bface.size();
    if ((rc = dalloc(ctx, &ctx->bface, ctx->nbf))) return rc;
    if ((rc = dalloc(ctx, &ctx->bbc, ctx->nbf))) return rc;
    if (ctx->mixed) {
        if ((rc = dalloc(ctx, &ctx->Gc32, 5 * (size_t)ctx->nbf))) return rc;
        if ((rc = dalloc(ctx, &ctx->Gf32, 5 * (size_t)ctx->nbf))) return rc;
    } else {
        if ((rc = dalloc(ctx, &ctx->Gc, 5 * (size_t)ctx->nbf))) return rc;
        if ((rc = dalloc(ctx, &ctx->Gf, 5 * (size_t)ctx->nbf))) return rc;
    }
    if (ctx->nbf) {
        HIPC(hipMemcpy(ctx->bface, L.bface.data(), L.bface.size() * 4, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(ctx->bbc, L.bbc.data(), L.bbc.size(), hipMemcpyHostToDevice));
    }
    if ((rc = dalloc(ctx, &ctx->dt_part, (size_t)ctx->nb))) return rc;
    if ((rc = dalloc(ctx, &ctx->res_part, (size_t)ctx->nb))) return rc;
    ctx->nb2 = (ctx->nb + RCHUNK - 1) / RCHUNK;
    if ((rc = dalloc(ctx, &ctx->red2, 2 * (size_t)ctx->nb2))) return rc;
    if ((rc = dalloc(ctx, &ctx->scal, 4))) return rc;
    if ((rc = dalloc(ctx, &ctx->iscal, 8))) return rc;
    {
        std::vector<double> n4((size_t)4 * ctx->n_nodes, 0.0);
        for (uint32_t n = 0; n < ctx->n_nodes; ++n)
            for (int d = 0; d < 3; ++d) n4[4 * (size_t)n + d] = desc->nodes[3 * (size_t)n + d];
        HIPC(hipMemcpy(ctx->nodes, n4.data(), n4.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    HIPC(hipMemcpy(ctx->cnodes, L.cnodes.data(), L.cnodes.size() * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(ctx->nbr, L.nbr.data(), L.nbr.size() * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(ctx->fcode, L.fcode.data(), L.fcode.size() * 4, hipMemcpyHostToDevice));
    if ((rc = ensure_hist(ctx, 1024))) return rc;
    if ((rc = nf_set_freestream(ctx))) return rc;
    *out = hold.release();
    return NF_OK;
}

int nf_destroy(nf_ctx *ctx) {
    if (!ctx) return NF_OK;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->graph) (void)hipGraphExecDestroy(ctx->graph);
    for (void *p : ctx->owned) (void)hipFree(p);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return NF_OK;
}

const char *nf_last_error(const nf_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int nf_ctx_sizes(const nf_ctx *ctx, nf_fv_sizes *s) {
    if (!ctx || !s) return NF_E_ARG;
    *s = ctx->sizes;
    return NF_OK;
}

int nf_set_freestream(nf_ctx *ctx) {
    if (!ctx) return NF_E_ARG;
    double V[5], U[5];
    for (int k = 0; k < 5; ++k) V[k] = ctx->prm.primvars_inf[k];
    prim_to_cons(V, U);
    std::vector<double> aos((size_t)5 * ctx->N);
    for (uint32_t c = 0; c < ctx->N; ++c)
        for (int k = 0; k < 5; ++k) aos[5 * (size_t)c + k] = U[k];
    int rc = upload_aos(ctx, aos.data(), ctx->Ua);
    if (rc) return rc;
    if ((rc = after_state_set(ctx))) return rc;
    // primvars = V_inf exactly (set_freestream_values, src/SolverData.cpp:50-58)
    if (!ctx->mixed) {
        for (uint32_t c = 0; c < ctx->N; ++c)
            for (int k = 0; k < 5; ++k) aos[5 * (size_t)c + k] = V[k];
        if ((rc = upload_aos(ctx, aos.data(), ctx->V))) return rc;
    }
    ctx->uold_zero = true;
    return NF_OK;
}

int nf_set_state(nf_ctx *ctx, const double *U, const double *U_old) {
    if (!ctx || !U) return NF_E_ARG;
    int rc = upload_aos(ctx, U, ctx->Ua);
    if (rc) return rc;
    if (U_old) {
        // U_old is held in Ua between steps: a non-zero U_old different from U
        // can only be represented for the next step's stages 2/3 via Ub staging
        std::vector<double> d((size_t)5 * ctx->N);
        bool same = std::memcmp(U, U_old, d.size() * sizeof(double)) == 0;
        if (!same) return fail(ctx, NF_E_ARG, "U_old must be NULL (zero) or equal to U");
        ctx->uold_zero = false;
    } else {
        ctx->uold_zero = true;
    }
    return after_state_set(ctx);
}

int nf_calc_timestep(nf_ctx *ctx, double *dt_out) {
    if (!ctx) return NF_E_ARG;
    int rc;
    if (!ctx->dt_valid) {
        if ((rc = clear_flags(ctx))) return rc;
        if ((rc = enqueue_timestep(ctx))) return rc;
        ctx->dt_valid = true;
    }
    double dt = 0;
    HIPC(hipMemcpyAsync(&dt, ctx->scal, sizeof dt, hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    if (dt_out) *dt_out = dt;
    if (!((dt - dt == 0.0) && dt > 0.0)) {
        char buf[96];
        snprintf(buf, sizeof buf, "Invalid dt calculated (dt = %f)", dt);
        return fail(ctx, NF_E_INVALID_DT, buf);
    }
    return NF_OK;
}

}  // extern "C"

// One step with the dt currently held on the device.
static int step_device_dt(nf_ctx *ctx) {
    int rc;
    if ((rc = clear_flags(ctx))) return rc;
    int nstep = 0;
    HIPC(hipMemcpyAsync(&nstep, ctx->iscal, sizeof nstep, hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    if ((rc = ensure_hist(ctx, nstep + 2))) return rc;
    const bool zero_old = ctx->uold_zero && ctx->prm.rk3_zero_old_first_step;
    if ((rc = enqueue_step(ctx, zero_old, -1.0, false))) return rc;
    if (ctx->timing && (rc = flush_timing(ctx))) return rc;
    ctx->uold_zero = false;
    ctx->dt_valid = true;
    ctx->aux_valid = false;
    return check_errors(ctx);
}

extern "C" {

int nf_step(nf_ctx *ctx, double dt) {
    if (!ctx) return NF_E_ARG;
    hipLaunchKernelGGL(k_set_flags, dim3(1), dim3(64), 0, ctx->stream, ctx->iscal, 1, 0, ctx->scal, 1, dt, 0);
    HIPC(hipGetLastError());
    return step_device_dt(ctx);
}

int nf_run(nf_ctx *ctx, const nf_run_opts *o, nf_run_stats *st) {
    if (!ctx || !o || o->n_steps < 0) return NF_E_ARG;
    int rc;
    const int K = std::max(1, o->check_every);
    if ((rc = clear_flags(ctx))) return rc;
    if (!ctx->dt_valid) {
        if ((rc = enqueue_timestep(ctx))) return rc;
        ctx->dt_valid = true;
    }
    int n0 = 0;
    HIPC(hipMemcpyAsync(&n0, ctx->iscal, sizeof n0, hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    if ((rc = ensure_hist(ctx, n0 + o->n_steps + 2))) return rc;
    const double drop = o->residual_drop;
    hipEvent_t e0, e1;
    HIPC(hipEventCreate(&e0));
    HIPC(hipEventCreate(&e1));
    HIPC(hipEventRecord(e0, ctx->stream));
    int done = 0;
    bool stopped = false;
    if (done < o->n_steps && ctx->uold_zero && ctx->prm.rk3_zero_old_first_step) {
        if ((rc = enqueue_step(ctx, true, drop, true))) return rc;
        ctx->uold_zero = false;
        ++done;
    }
    ctx->uold_zero = false;
    while (done < o->n_steps && !stopped) {
        const int chunk = std::min(K, o->n_steps - done);
        if (o->use_graph && chunk == K) {
            if (!ctx->graph || ctx->graph_steps != K || ctx->graph_drop != drop) {
                if (ctx->graph) (void)hipGraphExecDestroy(ctx->graph);
                ctx->graph = nullptr;
                hipGraph_t gr;
                HIPC(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                for (int s = 0; s < K; ++s)
                    if ((rc = enqueue_step(ctx, false, drop, true))) {
                        (void)hipStreamEndCapture(ctx->stream, &gr);
                        return rc;
                    }
                HIPC(hipStreamEndCapture(ctx->stream, &gr));
                HIPC(hipGraphInstantiate(&ctx->graph, gr, nullptr, nullptr, 0));
                (void)hipGraphDestroy(gr);
                ctx->graph_steps = K;
                ctx->graph_drop = drop;
            }
            HIPC(hipGraphLaunch(ctx->graph, ctx->stream));
        } else {
            for (int s = 0; s < chunk; ++s)
                if ((rc = enqueue_step(ctx, false, drop, true))) return rc;
        }
        done += chunk;
        if (drop > 0.0 || done >= o->n_steps) {
            int is[4];
            HIPC(hipMemcpyAsync(is, ctx->iscal, sizeof is, hipMemcpyDeviceToHost, ctx->stream));
            HIPC(hipStreamSynchronize(ctx->stream));
            if (is[2]) stopped = true;
        }
    }
    HIPC(hipEventRecord(e1, ctx->stream));
    HIPC(hipEventSynchronize(e1));
    float ms = 0;
    HIPC(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->aux_valid = false;
    int is[6];
    double sc[2];
    HIPC(hipMemcpy(is, ctx->iscal, sizeof is, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(sc, ctx->scal, sizeof sc, hipMemcpyDeviceToHost));
    if (st) {
        st->steps_done = is[0] - n0;
        st->converged = is[3];
        st->time = sc[1];
        st->last_dt = sc[0];
        st->elapsed_ms = ms;
        double r[2] = {0, 0};
        HIPC(hipMemcpy(&r[0], ctx->hist, sizeof(double), hipMemcpyDeviceToHost));
        if (is[0] > 0)
            HIPC(hipMemcpy(&r[1], ctx->hist + std::min(is[0] - 1, ctx->hist_cap - 1), sizeof(double),
                           hipMemcpyDeviceToHost));
        st->res_first = r[0];
        st->res_last = r[1];
    }
    return check_errors(ctx);
}

int nf_get_residuals(nf_ctx *ctx, double *out, int32_t max, int32_t *n_out) {
    if (!ctx || !out) return NF_E_ARG;
    int n = 0;
    HIPC(hipStreamSynchronize(ctx->stream));
    HIPC(hipMemcpy(&n, ctx->iscal, sizeof n, hipMemcpyDeviceToHost));
    n = std::min(std::min(n, max), ctx->hist_cap);
    if (n > 0) HIPC(hipMemcpy(out, ctx->hist, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    if (n_out) *n_out = n;
    return NF_OK;
}

int nf_get_state(nf_ctx *ctx, double *U) {
    if (!ctx || !U) return NF_E_ARG;
    return download_tiled(ctx, ctx->Ua, U);
}

// Evaluates the flux balance of the current state once (Solver::
// evaluate_flux_balance, src/Solver.cpp:54-81) without updating U, keeping
// gradient, limiter and R for nf_get_field.
int nf_eval_flux_balance(nf_ctx *ctx) {
    if (!ctx) return NF_E_ARG;
    int rc;
    const uint32_t N = ctx->N;
    if (!ctx->R_aux) {
        if ((rc = dalloc(ctx, &ctx->grad_aux, 15 * (size_t)N))) return rc;
        if ((rc = dalloc(ctx, &ctx->psi_aux, 5 * (size_t)N))) return rc;
        if ((rc = dalloc(ctx, &ctx->R_aux, 5 * (size_t)N))) return rc;
    }
    if ((rc = clear_flags(ctx))) return rc;
    // primvars of the current state
    if (ctx->mixed) {
        HIPC(mx::launch_prim_mx(dim3((N + BLOCK - 1) / BLOCK), ctx->stream, ctx->Ua, ctx->V32, N));
    } else {
        hipLaunchKernelGGL(k_prim, dim3((N + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, ctx->stream, ctx->Ua, ctx->V, N);
        HIPC(hipGetLastError());
    }
    if ((rc = run_sweepA(ctx, false, true))) return rc;
    if ((rc = run_sweepB(ctx, 1, ctx->Ua, nullptr, ctx->Ub, false, true))) return rc;
    ctx->aux_valid = true;
    return check_errors(ctx);
}

int nf_get_field(nf_ctx *ctx, int32_t which, double *out) {
    if (!ctx || !out) return NF_E_ARG;
    switch (which) {
    case NF_FIELD_PRIM:
        if (ctx->mixed) return download_tiled_f32(ctx, ctx->V32, out);
        return download_tiled(ctx, ctx->V, out);
    case NF_FIELD_STATE_OLD: {
        if (ctx->uold_zero) {
            std::memset(out, 0, sizeof(double) * 5 * ctx->N);
            return NF_OK;
        }
        return download_tiled(ctx, ctx->Ua, out);
    }
    case NF_FIELD_LIMITER:
    case NF_FIELD_GRADIENT:
    case NF_FIELD_RESIDUAL:
        if (!ctx->aux_valid) return fail(ctx, NF_E_ARG, "call nf_eval_flux_balance first");
        if (which == NF_FIELD_LIMITER) return download_aos(ctx, ctx->psi_aux, 5, out);
        if (which == NF_FIELD_GRADIENT) return download_aos(ctx, ctx->grad_aux, 15, out);
        return download_aos(ctx, ctx->R_aux, 5, out);
    default:
        return fail(ctx, NF_E_ARG, "unknown field");
    }
}

int nf_time_kernels(nf_ctx *ctx, int32_t n_steps, nf_kernel_time *out, int32_t max, int32_t *n_out) {
    if (!ctx || n_steps < 1) return NF_E_ARG;
    int rc;
    if (!ctx->dt_valid) {
        double dt;
        if ((rc = nf_calc_timestep(ctx, &dt))) return rc;
    }
    ctx->ktime.clear();
    ctx->timing = true;
    for (int s = 0; s < n_steps; ++s) {
        if ((rc = step_device_dt(ctx))) {
            ctx->timing = false;
            return rc;
        }
    }
    ctx->timing = false;
    int n = 0;
    for (auto &kv : ctx->ktime) {
        if (n >= max) break;
        std::snprintf(out[n].name, sizeof out[n].name, "%s", kv.first.c_str());
        out[n].avg_ms = kv.second.first / kv.second.second;
        out[n].calls = kv.second.second;
        ++n;
    }
    if (n_out) *n_out = n;
    return NF_OK;
}

}  // extern "C"
