// Solver context and C ABI (include/netflux.h).
//
// The context replaces EulerSolver + EulerSolverData (include/Solver.hpp:52-79,
// include/SolverData.hpp:88-117): it owns the device fields of one or more
// meshes (instances) and runs the Driver loop order (calc_timestep, then
// step; src/Driver.cpp:38-47) with the time steps kept on the device.  All
// instances of a context advance together: one launch per sweep covers every
// instance, each with its own dt, run counters and stop flag.
#include <hip/hip_runtime.h>

#include <algorithm>
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
#include "mlp.hpp"
#include "poly.hpp"

using namespace nf;

namespace {

constexpr uint32_t RCHUNK = 2048;  // blocks per first-level reduction chunk

// First level of the per-block partial reductions: chunk q reduces blocks
// [first[q], first[q] + n[q]) of one instance in a fixed order.
template <bool MIN>
__global__ __launch_bounds__(BLOCK) void k_reduce(const double *in, const uint32_t *first, const uint32_t *n,
                                                  double *out) {
    __shared__ double sh[BLOCK / 64];
    const uint32_t q = blockIdx.x, b0 = first[q], nb = n[q];
    double v = MIN ? DBL_MAX : 0.0;
    for (uint32_t i = threadIdx.x; i < nb; i += BLOCK) {
        const double x = in[b0 + i];
        v = MIN ? ((x < v) ? x : v) : v + x;
    }
    v = block_reduce<MIN, BLOCK>(v, sh);
    if (threadIdx.x == 0) out[q] = v;
}

// Second level, one block per instance: dt = min over the instance's chunks.
// advance = 1 at the end of a step: time += dt, nstep += 1, the first-step
// flag clears, a converged instance stops, and the new dt (of the primvars
// the step's last sweep A read) is installed for the next step.
__global__ __launch_bounds__(BLOCK) void k_finalize_dt(const double *red, const uint32_t *c0, const uint32_t *nc,
                                                       InstState *st, int advance) {
    __shared__ double sh[BLOCK / 64];
    const uint32_t i = blockIdx.x;
    if (st[i].stop) return;
    double v = DBL_MAX;
    for (uint32_t q = threadIdx.x; q < nc[i]; q += BLOCK) {
        const double x = red[c0[i] + q];
        v = (x < v) ? x : v;
    }
    v = block_reduce<true, BLOCK>(v, sh);
    if (threadIdx.x == 0) {
        if (advance) {
            st[i].time += st[i].dt;
            st[i].nstep += 1;
            st[i].uold_zero = 0;
            if (st[i].converged) st[i].stop = 1;
        }
        // num_is_valid_and_pos (include/Includes.hpp:131-137)
        const bool ok = (v - v == 0.0) && v > 0.0;
        st[i].dt = v;
        if (!ok) {
            st[i].err = NF_E_INVALID_DT;
            st[i].stop = 1;
        }
    }
}

// r_k = sqrt(sum of the instance's partials), summed in a fixed order, into
// the instance's residual history; sets `converged` once r_k / ref <= drop,
// ref = rref if set, else r_0.  With `out`, r goes to out[i] instead.
__global__ __launch_bounds__(BLOCK) void k_finalize_res(const double *red, const uint32_t *c0, const uint32_t *nc,
                                                        InstState *st, double *hist, int cap, double drop,
                                                        double *out) {
    __shared__ double sh[BLOCK / 64];
    const uint32_t i = blockIdx.x;
    if (!out && st[i].stop) return;
    double v = 0.0;
    for (uint32_t q = threadIdx.x; q < nc[i]; q += BLOCK) v += red[c0[i] + q];
    v = block_reduce<false, BLOCK>(v, sh);
    if (threadIdx.x == 0) {
        const double r = sqrt(v);
        if (out) {
            out[i] = r;
            return;
        }
        const int k = st[i].nstep;
        double *h = hist + (size_t)i * cap;
        if (k < cap) h[k] = r;
        const double ref = st[i].rref > 0.0 ? st[i].rref : h[0];
        if (drop > 0.0 && k < cap && ref > 0.0 && r / ref <= drop) st[i].converged = 1;
    }
}

// op 0: clear the run flags of every instance (err, stop, converged, error
// counters); op 2: the same, but a converged instance stays converged and
// stopped; op 1: set dt of instance `which`
__global__ void k_flags(InstState *st, int n, int op, int which, double dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (op == 0 || op == 2) {
        const int keep = op == 2 && st[i].converged;
        st[i].err = 0;
        st[i].stop = keep;
        st[i].converged = keep;
        st[i].bad_flux = 0;
        st[i].bad_state = 0;
    } else if (i == which) {
        st[i].dt = dt;
    }
}

// fresh run state of instance i (nstep = 0, time = 0)
__global__ void k_reset_inst(InstState *st, int i, int uold_zero, double rref) {
    if (threadIdx.x == 0) st[i] = InstState{0.0, 0.0, rref, 0, 0, 0, 0, 0, 0, uold_zero, 0};
}

// Network features of the cells [0, n) of one instance, feature-major
// x[6][n]: (p0, p1, centroid, distance to the nearest wall point).  The wall
// points are streamed through LDS; every thread takes part in every pass.
__global__ __launch_bounds__(BLOCK) void k_features(DevMesh m, uint32_t cell0, uint32_t n, const float4 *wall,
                                                    uint32_t nw, float p0, float p1, float *x) {
    __shared__ float4 sh[BLOCK];
    const uint32_t c = blockIdx.x * BLOCK + threadIdx.x;
    const bool live = c < n;
    float cx = 0.f, cy = 0.f, cz = 0.f;
    if (live) {
        const uint4 q = *reinterpret_cast<const uint4 *>(m.cnodes + 4 * (size_t)(cell0 + c));
        const uint32_t id[4] = {q.x, q.y, q.z, q.w};
        double s[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < 4; ++k)
            for (int d = 0; d < 3; ++d) s[d] += m.nodes[4 * (size_t)id[k] + d];
        cx = (float)(0.25 * s[0]);
        cy = (float)(0.25 * s[1]);
        cz = (float)(0.25 * s[2]);
    }
    float d2 = INFINITY;
    for (uint32_t w0 = 0; w0 < nw; w0 += BLOCK) {
        if (w0 + threadIdx.x < nw) sh[threadIdx.x] = wall[w0 + threadIdx.x];
        __syncthreads();
        const uint32_t cnt = min((uint32_t)BLOCK, nw - w0);
        for (uint32_t k = 0; k < cnt; ++k) {
            const float dx = sh[k].x - cx, dy = sh[k].y - cy, dz = sh[k].z - cz;
            d2 = fminf(d2, dx * dx + dy * dy + dz * dz);
        }
        __syncthreads();
    }
    if (live) {
        x[c] = p0;
        x[n + c] = p1;
        x[2 * (size_t)n + c] = cx;
        x[3 * (size_t)n + c] = cy;
        x[4 * (size_t)n + c] = cz;
        x[5 * (size_t)n + c] = nw ? sqrtf(d2) : 0.f;
    }
}

// predicted primitives V[5][n] -> conservative state of the cells [0, n) of
// one instance (tiles starting at its first cell); cells with rho <= 0,
// p <= 0 or a non-finite value are counted in bad_state
__global__ __launch_bounds__(BLOCK) void k_warm_state(const float *Vp, uint32_t n, double *U, InstState *st, int inst) {
    const uint32_t c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= n) return;
    double V[5], Uc[5];
    for (int k = 0; k < 5; ++k) V[k] = (double)Vp[(size_t)k * n + c];
    prim_to_cons(V, Uc);
    bool ok = V[0] > 0.0 && V[4] > 0.0;
    for (int k = 0; k < 5; ++k) {
        ok = ok && (Uc[k] - Uc[k] == 0.0);
        U[tix(c, k)] = Uc[k];
    }
    if (!ok) atomicAdd(&st[inst].bad_state, 1);
}

struct Inst {
    uint32_t cell0 = 0, N = 0, Npad = 0;  // device cells [cell0, cell0 + N), block [cell0, cell0 + Npad)
    nf_fv_sizes sizes{};
    std::vector<uint32_t> perm;  // local device cell -> reference cell
    double Vinf[5] = {0, 0, 0, 0, 0};
    bool dt_valid = false;
    uint32_t wall0 = 0, nwall = 0;  // wall-face centroids of the instance
    double rref = 0.0;              // residual reference of the drop (0: r_0)
    bool warm = false;              // state from the network, not yet fallen back
    // FP64 primvars V of the instance differ from prim(U): steps do not store
    // them (each sweep A evaluates prim(U) itself); refreshed where V is read
    bool v_stale = false;
};

// reduction chunks of whole instances for one sweep block size
struct RedTable {
    int nchunk = 0;
    uint32_t *first = nullptr, *n = nullptr;          // [nchunk]
    uint32_t *inst_c0 = nullptr, *inst_nc = nullptr;  // [instances] chunk range
};

}  // namespace

struct nf_ctx {
    nf_params prm{};
    mutable std::string err;  // last error message (also set by const queries)
    int device = 0;
    hipStream_t stream = nullptr;
    int n_inst = 0;
    std::vector<Inst> inst;
    uint32_t Ncell = 0;  // device cells including padding (multiple of 256)
    int block = 256, xcd_remap = 1;  // sweep launch configuration
    bool mixed = false;
    double *nodes = nullptr;
    uint32_t *cnodes = nullptr, *nbr = nullptr, *fcode = nullptr;
    uint32_t *grp_off = nullptr;          // 64-cell group node tables
    double *gx = nullptr, *gy = nullptr, *gz = nullptr;
    uint2 *cell_ln = nullptr;
    uint32_t *blk_inst = nullptr;
    uint32_t *inst_end = nullptr;
    InstState *st = nullptr;
    double *Ua = nullptr, *Ub = nullptr, *V = nullptr, *FV = nullptr;
    float *V32 = nullptr, *FV32 = nullptr;
    uint32_t nbf = 0;
    uint32_t *bface = nullptr;
    uint8_t *bbc = nullptr;
    double *bS = nullptr;  // [3][nbf] S_ij of the boundary faces
    uint16_t *binst = nullptr;
    double *Gc = nullptr, *Gf = nullptr;
    float *Gc32 = nullptr, *Gf32 = nullptr;
    FarField<double> *fs = nullptr;
    FarField<float> *fs32 = nullptr;
    double *dt_cell = nullptr;
    double *grad_aux = nullptr, *psi_aux = nullptr, *R_aux = nullptr;
    double *Rst = nullptr, *fst = nullptr;           // checked stages: R and f for k_apply
    double *dt_part = nullptr, *res_part = nullptr;  // per sweep block
    RedTable red256, red128;
    double *red_dt = nullptr, *red_res = nullptr;  // per chunk
    double *hist = nullptr;                         // [instances][hist_cap]
    int hist_cap = 0;
    float4 *wall = nullptr;                         // wall-face centroids of all instances
    float *feat = nullptr, *vpred = nullptr;        // network features / prediction scratch
    uint32_t feat_cap = 0;
    double *r_eval = nullptr;                       // [instances]
    double *stage = nullptr;                        // [5 x largest instance] AoS state staging
    uint32_t *perm_dev = nullptr;                   // [Ncell] device cell -> reference cell of its instance
    bool aux_valid = false;
    std::vector<void *> owned;
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0, graph_block = 0, graph_xcd = -1;
    double graph_drop = -1.0;
    bool timing = false;
    std::map<std::string, std::pair<double, int>> ktime;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    nf::PolyPath *poly = nullptr;  // mixed-element mesh: the polyhedral path (n_inst = 0)
};

namespace {

int fail(nf_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPC(expr)                                                                         \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
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

DevMesh devmesh(const nf_ctx *c) {
    return DevMesh{3.0,         6.0,       c->nodes, c->cnodes, c->grp_off,  c->gx,       c->gy,
                   c->gz,       c->cell_ln, c->nbr,   c->fcode,  c->blk_inst, c->inst_end, c->st};
}

const RedTable &red(const nf_ctx *c) { return c->block == 128 ? c->red128 : c->red256; }

// kernel launch with optional per-kernel event timing
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
#define LAUNCH(name, kernel, grid, blk, ...)                                          \
    launch_fn(ctx, name, [&] {                                                        \
        hipLaunchKernelGGL(kernel, grid, blk, 0, ctx->stream, __VA_ARGS__);           \
        return hipGetLastError();                                                     \
    })

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

// V = prim(U) for the instances whose FP64 primvars are stale (what a step
// that stored them would have left)
int refresh_v(nf_ctx *ctx) {
    for (Inst &I : ctx->inst) {
        if (!I.v_stale) continue;
        hipLaunchKernelGGL(k_prim, dim3(I.Npad / BLOCK), dim3(BLOCK), 0, ctx->stream,
                           (const double *)(ctx->Ua + (size_t)5 * I.cell0), ctx->V + (size_t)5 * I.cell0, I.Npad);
        HIPC(hipGetLastError());
        I.v_stale = false;
    }
    return NF_OK;
}

// U: cell ghosts from prim(U) of this state instead of the stored primvars
int run_ghost(nf_ctx *ctx, bool face, const double *U = nullptr) {
    if (ctx->nbf == 0) return NF_OK;
    const dim3 g((ctx->nbf + BLOCK - 1) / BLOCK), b(BLOCK);
    const DevMesh m = devmesh(ctx);
    if (ctx->mixed) {
        const mx::GhostArgsMx a{ctx->bface, ctx->bbc, ctx->binst, ctx->nbf, face ? ctx->FV32 : ctx->V32,
                                face ? ctx->Gf32 : ctx->Gc32, ctx->fs32};
        return launch_fn(ctx, face ? "ghost_face_mx" : "ghost_cell_mx",
                         [&] { return mx::launch_ghost_mx(face, g, ctx->stream, m, a); });
    }
    const GhostArgs a{ctx->bface, ctx->bbc, ctx->binst, ctx->nbf, face ? ctx->FV : (U ? U : ctx->V),
                      face ? ctx->Gf : ctx->Gc, ctx->fs, (!face && U) ? 1 : 0, ctx->bS};
    if (face) return LAUNCH("ghost_face", k_ghost<true>, g, b, m, a);
    return LAUNCH("ghost_cell", k_ghost<false>, g, b, m, a);
}

int order_of(const nf_ctx *ctx) {
    if (ctx->prm.spatial_order == NF_FIRST_ORDER) return 0;
    return ctx->prm.limiter == NF_LIMITER_NONE ? 1 : 2;
}

template <int B, bool FROM_U>
hipError_t launch_sweepA(hipStream_t s, uint32_t ncell, int order, const DevMesh &m, const SweepAArgs &a) {
    const dim3 g(ncell / B), b(B);
    if (order == 0) hipLaunchKernelGGL((k_sweepA<0, B, FROM_U>), g, b, 0, s, m, a);
    else if (order == 1) hipLaunchKernelGGL((k_sweepA<1, B, FROM_U>), g, b, 0, s, m, a);
    else hipLaunchKernelGGL((k_sweepA<2, B, FROM_U>), g, b, 0, s, m, a);
    return hipGetLastError();
}

// sweep A of the current primvars, or (U set, FP64) of prim(U) of that state.
// dt: per-block CFL minima into dt_part; dtc: per-cell CFL steps into dt_cell
// (local time stepping); aux: keep the gradient and limiter fields.
int run_sweepA(nf_ctx *ctx, bool dt, bool aux, bool dtc = false, const double *U = nullptr) {
    int rc;
    if (ctx->mixed) U = nullptr;
    if (!U && !ctx->mixed && (rc = refresh_v(ctx))) return rc;
    if ((rc = run_ghost(ctx, false, U))) return rc;
    const DevMesh m = devmesh(ctx);
    const int order = order_of(ctx);
    if (ctx->mixed) {
        const mx::SweepAArgsMx x{ctx->V32, ctx->Gc32, ctx->FV32, dt ? ctx->dt_part : nullptr,
                                 dtc ? ctx->dt_cell : nullptr, ctx->prm.CFL, ctx->xcd_remap,
                                 aux ? ctx->grad_aux : nullptr, aux ? ctx->psi_aux : nullptr};
        return launch_fn(ctx, order == 2 ? "sweepA_barth_mx" : "sweepA_mx", [&] {
            return mx::launch_sweepA_mx(order, ctx->block, dim3(ctx->Ncell / ctx->block), ctx->stream, m, x);
        });
    }
    const SweepAArgs a{U ? U : ctx->V, ctx->Gc, ctx->FV, dt ? ctx->dt_part : nullptr, dtc ? ctx->dt_cell : nullptr,
                       ctx->prm.CFL, ctx->xcd_remap, aux ? ctx->grad_aux : nullptr, aux ? ctx->psi_aux : nullptr};
    const char *name = order == 2 ? "sweepA_barth" : (order == 1 ? "sweepA_nolim" : "sweepA_first");
    return launch_fn(ctx, name, [&] {
        if (U)
            return ctx->block == 128 ? launch_sweepA<128, true>(ctx->stream, ctx->Ncell, order, m, a)
                                     : launch_sweepA<256, true>(ctx->stream, ctx->Ncell, order, m, a);
        return ctx->block == 128 ? launch_sweepA<128, false>(ctx->stream, ctx->Ncell, order, m, a)
                                 : launch_sweepA<256, false>(ctx->stream, ctx->Ncell, order, m, a);
    });
}

// sweep B: stage `kind` of the update from U into Uout, prim(Uout) into V.
// res: per-block sums of R_rho^2; aux: keep R.  With check_physical_validity
// the sweep stores R and the stage factor, and k_apply updates only the
// instances whose flux balance is finite everywhere (the others stop with U
// untouched, as the reference throws before the update).
int run_sweepB(nf_ctx *ctx, int kind, const double *U, double *Uout, bool res, bool aux) {
    int rc;
    if ((rc = run_ghost(ctx, true))) return rc;
    const DevMesh m = devmesh(ctx);
    const double *dtc = ctx->prm.local_time_step ? ctx->dt_cell : nullptr;
    const bool checked = ctx->prm.check_physical_validity && !aux;
    double *Rst = checked ? ctx->Rst : nullptr, *fst = checked ? ctx->fst : nullptr;
    const ApplyArgs ap{U, ctx->Ua, Uout, ctx->mixed ? (void *)ctx->V32 : nullptr, Rst, fst, kind};
    if (ctx->mixed) {
        const mx::SweepBArgsMx x{ctx->FV32, ctx->Gf32, U, ctx->Ua, Uout, ctx->V32, dtc, kind,
                                 res ? ctx->res_part : nullptr, ctx->prm.check_physical_validity,
                                 ctx->xcd_remap, aux ? ctx->R_aux : nullptr, Rst, fst};
        if ((rc = launch_fn(ctx, "sweepB_mx", [&] {
                 return mx::launch_sweepB_mx(ctx->block, dim3(ctx->Ncell / ctx->block), ctx->stream, m, x);
             })))
            return rc;
        if (!checked) return NF_OK;
        return launch_fn(ctx, "apply_mx",
                         [&] { return mx::launch_apply_mx(dim3(ctx->Ncell / BLOCK), ctx->stream, m, ap); });
    }
    // (FP64: prim(Uout) is not stored; the next sweep A evaluates it from U)
    const SweepBArgs a{ctx->FV, ctx->Gf, U, ctx->Ua, Uout, nullptr, dtc, kind, res ? ctx->res_part : nullptr,
                       ctx->prm.check_physical_validity, ctx->xcd_remap, aux ? ctx->R_aux : nullptr, Rst, fst};
    const bool hllc = ctx->prm.inv_flux_scheme == NF_FLUX_HLLC;
    if ((rc = launch_fn(ctx, "sweepB", [&] {
             const dim3 g(ctx->Ncell / ctx->block), b(ctx->block);
             if (ctx->block == 128 && hllc) hipLaunchKernelGGL((k_sweepB<128, 1>), g, b, 0, ctx->stream, m, a);
             else if (ctx->block == 128) hipLaunchKernelGGL((k_sweepB<128, 0>), g, b, 0, ctx->stream, m, a);
             else if (hllc) hipLaunchKernelGGL((k_sweepB<256, 1>), g, b, 0, ctx->stream, m, a);
             else hipLaunchKernelGGL((k_sweepB<256, 0>), g, b, 0, ctx->stream, m, a);
             return hipGetLastError();
         })))
        return rc;
    if (!checked) return NF_OK;
    return LAUNCH("apply", k_apply, dim3(ctx->Ncell / BLOCK), dim3(BLOCK), m, ap);
}

int reduce_dt(nf_ctx *ctx, int advance) {
    int rc;
    const RedTable &r = red(ctx);
    if ((rc = LAUNCH("reduce_dt", k_reduce<true>, dim3(r.nchunk), dim3(BLOCK), (const double *)ctx->dt_part,
                     (const uint32_t *)r.first, (const uint32_t *)r.n, ctx->red_dt)))
        return rc;
    return LAUNCH("finalize_dt", k_finalize_dt, dim3(ctx->n_inst), dim3(BLOCK), (const double *)ctx->red_dt,
                  (const uint32_t *)r.inst_c0, (const uint32_t *)r.inst_nc, ctx->st, advance);
}

int reduce_res(nf_ctx *ctx, double drop, double *out = nullptr) {
    int rc;
    const RedTable &r = red(ctx);
    if ((rc = LAUNCH("reduce_res", k_reduce<false>, dim3(r.nchunk), dim3(BLOCK), (const double *)ctx->res_part,
                     (const uint32_t *)r.first, (const uint32_t *)r.n, ctx->red_res)))
        return rc;
    return LAUNCH("finalize_res", k_finalize_res, dim3(ctx->n_inst), dim3(BLOCK), (const double *)ctx->red_res,
                  (const uint32_t *)r.inst_c0, (const uint32_t *)r.inst_nc, ctx->st, ctx->hist, ctx->hist_cap,
                  drop, out);
}

// One step of every live instance (Solver::step, src/Solver.cpp:34-52).  The
// last sweep A of the step evaluates the CFL dt of the primvars it reads --
// prim(U2) for TVD-RK3, prim(U^n) for explicit Euler -- which is the state
// the reference's next calc_timestep sees (src/Solver.cpp:248);
// k_finalize_dt installs that dt for the following step.  U^n (= U_old after
// the first step) stays in Ua; stages write Ub, the last stage writes Ua.
int enqueue_step(nf_ctx *ctx, double drop) {
    int rc;
    const bool rk3 = ctx->prm.time_scheme == NF_TVD_RK3;
    const bool lts = ctx->prm.local_time_step;
    // each stage's sweep A evaluates the primvars of its input state
    if ((rc = run_sweepA(ctx, !rk3 || lts, false, lts, ctx->Ua))) return rc;
    if ((rc = run_sweepB(ctx, 1, ctx->Ua, rk3 ? ctx->Ub : ctx->Ua, true, false))) return rc;
    if ((rc = reduce_res(ctx, drop))) return rc;
    if (rk3) {
        if ((rc = run_sweepA(ctx, false, false, false, ctx->Ub))) return rc;
        if ((rc = run_sweepB(ctx, 2, ctx->Ub, ctx->Ub, false, false))) return rc;
        if ((rc = run_sweepA(ctx, !lts, false, false, ctx->Ub))) return rc;
        if ((rc = run_sweepB(ctx, 3, ctx->Ub, ctx->Ua, false, false))) return rc;
    }
    if (!ctx->mixed)
        for (Inst &I : ctx->inst) I.v_stale = true;
    return reduce_dt(ctx, 1);
}

// dt of the current primvars with a standalone sweep A
int enqueue_timestep(nf_ctx *ctx) {
    int rc;
    if ((rc = run_sweepA(ctx, true, false))) return rc;
    return reduce_dt(ctx, 0);
}

int clear_flags(nf_ctx *ctx, bool keep_converged = false) {
    hipLaunchKernelGGL(k_flags, dim3((ctx->n_inst + 63) / 64), dim3(64), 0, ctx->stream, ctx->st, ctx->n_inst,
                       keep_converged ? 2 : 0, 0, 0.0);
    HIPC(hipGetLastError());
    return NF_OK;
}

int read_states(nf_ctx *ctx, std::vector<InstState> &h) {
    h.resize(ctx->n_inst);
    HIPC(hipStreamSynchronize(ctx->stream));
    HIPC(hipMemcpy(h.data(), ctx->st, h.size() * sizeof(InstState), hipMemcpyDeviceToHost));
    return NF_OK;
}

// the reference's messages (src/Solver.cpp:267-268, src/SolverData.cpp:73-78);
// for a batch, the first failing instance is named
int check_errors(nf_ctx *ctx) {
    std::vector<InstState> h;
    int rc;
    if ((rc = read_states(ctx, h))) return rc;
    for (int i = 0; i < ctx->n_inst; ++i) {
        const std::string who = ctx->n_inst > 1 ? "instance " + std::to_string(i) + ": " : "";
        if (h[i].err == NF_E_INVALID_DT) {
            char buf[96];
            snprintf(buf, sizeof buf, "Invalid dt calculated (dt = %f)", h[i].dt);
            return fail(ctx, NF_E_INVALID_DT, who + buf);
        }
        if (h[i].bad_flux > 0)
            return fail(ctx, NF_E_NONPHYSICAL,
                        who + std::to_string(h[i].bad_flux) +
                            " cells with unphysical values detected in the flux_balance field of the Euler solver");
    }
    return NF_OK;
}

int ensure_hist(nf_ctx *ctx, int need) {
    if (need <= ctx->hist_cap) return NF_OK;
    const int cap = std::max(need, 2 * ctx->hist_cap + 1024);
    const size_t bytes = (size_t)cap * ctx->n_inst * sizeof(double);
    double *h = nullptr;
    HIPC(hipMalloc(&h, bytes));
    HIPC(hipMemsetAsync(h, 0, bytes, ctx->stream));
    if (ctx->hist) {
        HIPC(hipMemcpy2DAsync(h, (size_t)cap * sizeof(double), ctx->hist, (size_t)ctx->hist_cap * sizeof(double),
                              (size_t)ctx->hist_cap * sizeof(double), ctx->n_inst, hipMemcpyDeviceToDevice,
                              ctx->stream));
        HIPC(hipStreamSynchronize(ctx->stream));
        auto it = std::find(ctx->owned.begin(), ctx->owned.end(), (void *)ctx->hist);
        if (it != ctx->owned.end()) ctx->owned.erase(it);
        HIPC(hipFree(ctx->hist));
    }
    HIPC(hipStreamSynchronize(ctx->stream));
    ctx->owned.push_back(h);
    ctx->hist = h;
    ctx->hist_cap = cap;
    if (ctx->graph) {  // the graph holds the old pointer
        (void)hipGraphExecDestroy(ctx->graph);
        ctx->graph = nullptr;
    }
    return NF_OK;
}

size_t tile_index(uint32_t c, int k) { return ((size_t)(c >> 6) * 5 + k) * TILE + (c & 63u); }

// reference-order AoS [N][5] of instance i -> device tiles
// the staging buffer and the device copy of the cell maps, made on first use
int ensure_stage(nf_ctx *ctx) {
    if (ctx->stage) return NF_OK;
    uint32_t nmax = 1;
    std::vector<uint32_t> perm(ctx->Ncell, 0);
    for (const Inst &I : ctx->inst) {
        nmax = std::max(nmax, I.N);
        std::copy(I.perm.begin(), I.perm.end(), perm.begin() + I.cell0);
    }
    int rc;
    if ((rc = dalloc(ctx, &ctx->perm_dev, ctx->Ncell))) return rc;
    HIPC(hipMemcpy(ctx->perm_dev, perm.data(), perm.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    return dalloc(ctx, &ctx->stage, (size_t)5 * nmax);
}

// reference-order AoS [N][5] of instance i -> its device tiles: the array
// crosses as is and the device reorders it (the host never touches it)
int upload_state(nf_ctx *ctx, int i, const double *aos, double *dev) {
    const Inst &I = ctx->inst[i];
    int rc;
    if ((rc = ensure_stage(ctx))) return rc;
    HIPC(hipMemcpyAsync(ctx->stage, aos, (size_t)5 * I.N * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_state_in, dim3(I.Npad / BLOCK), dim3(BLOCK), 0, ctx->stream, (const double *)ctx->stage,
                       (const uint32_t *)(ctx->perm_dev + I.cell0), dev + (size_t)5 * I.cell0, I.N, I.Npad);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(ctx->stream));
    return NF_OK;
}

// device tiles of instance i -> reference-order AoS [N][5]
int download_state(nf_ctx *ctx, int i, const double *dev, double *aos) {
    const Inst &I = ctx->inst[i];
    int rc;
    if ((rc = ensure_stage(ctx))) return rc;
    hipLaunchKernelGGL(k_state_out, dim3(I.Npad / BLOCK), dim3(BLOCK), 0, ctx->stream,
                       dev + (size_t)5 * I.cell0, (const uint32_t *)(ctx->perm_dev + I.cell0), ctx->stage, I.N);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(aos, ctx->stage, (size_t)5 * I.N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    return NF_OK;
}
template <typename T>
int download_state(nf_ctx *ctx, int i, const T *dev, double *aos) {
    const Inst &I = ctx->inst[i];
    std::vector<T> d((size_t)5 * I.Npad);
    HIPC(hipStreamSynchronize(ctx->stream));
    HIPC(hipMemcpy(d.data(), dev + (size_t)5 * I.cell0, d.size() * sizeof(T), hipMemcpyDeviceToHost));
    for (uint32_t c = 0; c < I.N; ++c)
        for (int k = 0; k < 5; ++k) aos[5 * (size_t)I.perm[c] + k] = d[tile_index(c, k)];
    return NF_OK;
}

// device-cell AoS [cells][rows] of instance i -> reference-order AoS
int download_aos(nf_ctx *ctx, int i, const double *dev, int rows, double *aos) {
    const Inst &I = ctx->inst[i];
    std::vector<double> d((size_t)rows * I.N);
    HIPC(hipStreamSynchronize(ctx->stream));
    HIPC(hipMemcpy(d.data(), dev + (size_t)rows * I.cell0, d.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (uint32_t c = 0; c < I.N; ++c)
        std::memcpy(aos + (size_t)rows * I.perm[c], &d[(size_t)rows * c], rows * sizeof(double));
    return NF_OK;
}

// primvars of instance i from its state, fresh run state
int after_state_set(nf_ctx *ctx, int i, bool uold_zero) {
    Inst &I = ctx->inst[i];
    const dim3 g(I.Npad / BLOCK), b(BLOCK);
    if (ctx->mixed) {
        HIPC(mx::launch_prim_mx(g, ctx->stream, ctx->Ua + (size_t)5 * I.cell0, ctx->V32 + (size_t)5 * I.cell0,
                                I.Npad));
    } else {
        hipLaunchKernelGGL(k_prim, g, b, 0, ctx->stream, (const double *)(ctx->Ua + (size_t)5 * I.cell0),
                           ctx->V + (size_t)5 * I.cell0, I.Npad);
        HIPC(hipGetLastError());
    }
    hipLaunchKernelGGL(k_reset_inst, dim3(1), dim3(64), 0, ctx->stream, ctx->st, i, (int)uold_zero, I.rref);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(ctx->stream));
    I.dt_valid = false;
    I.warm = false;
    I.v_stale = false;
    ctx->aux_valid = false;
    return NF_OK;
}

// V = prim(U) over every device cell
int prim_all(nf_ctx *ctx) {
    if (ctx->mixed) {
        HIPC(mx::launch_prim_mx(dim3(ctx->Ncell / BLOCK), ctx->stream, ctx->Ua, ctx->V32, ctx->Ncell));
    } else {
        hipLaunchKernelGGL(k_prim, dim3(ctx->Ncell / BLOCK), dim3(BLOCK), 0, ctx->stream, (const double *)ctx->Ua,
                           ctx->V, ctx->Ncell);
        HIPC(hipGetLastError());
    }
    for (Inst &I : ctx->inst) I.v_stale = false;
    return NF_OK;
}

// ||R_rho||_2 of every instance's current state into out[n] (host); the
// update of the stage-1 sweep goes to Ub and is discarded, V = prim(U) after
int eval_residuals(nf_ctx *ctx, double *out) {
    int rc;
    if ((rc = clear_flags(ctx))) return rc;
    if ((rc = prim_all(ctx))) return rc;
    if ((rc = run_sweepA(ctx, false, false))) return rc;
    if ((rc = run_sweepB(ctx, 1, ctx->Ua, ctx->Ub, true, false))) return rc;
    if ((rc = reduce_res(ctx, 0.0, ctx->r_eval))) return rc;
    if ((rc = prim_all(ctx))) return rc;
    HIPC(hipStreamSynchronize(ctx->stream));
    HIPC(hipMemcpy(out, ctx->r_eval, ctx->n_inst * sizeof(double), hipMemcpyDeviceToHost));
    for (Inst &I : ctx->inst) I.dt_valid = false;
    ctx->aux_valid = false;
    return NF_OK;
}

// dt from the current primvars for the instances with need[i]; the others
// keep their device dt (the one their last step computed) untouched
int timestep_subset(nf_ctx *ctx, const std::vector<bool> &need) {
    int rc;
    if (std::all_of(need.begin(), need.end(), [](bool b) { return b; })) return enqueue_timestep(ctx);
    std::vector<InstState> h;
    if ((rc = read_states(ctx, h))) return rc;
    std::vector<InstState> masked = h;
    for (int i = 0; i < ctx->n_inst; ++i)
        if (!need[i]) masked[i].stop = 1;
    HIPC(hipMemcpy(ctx->st, masked.data(), h.size() * sizeof(InstState), hipMemcpyHostToDevice));
    if ((rc = enqueue_timestep(ctx))) return rc;
    std::vector<InstState> after;
    if ((rc = read_states(ctx, after))) return rc;
    for (int i = 0; i < ctx->n_inst; ++i)
        if (!need[i]) after[i] = h[i];
    HIPC(hipMemcpy(ctx->st, after.data(), h.size() * sizeof(InstState), hipMemcpyHostToDevice));
    return NF_OK;
}

// feature-major features of instance i into ctx->feat
int features(nf_ctx *ctx, int i, const double *cp) {
    const Inst &I = ctx->inst[i];
    if (I.N > ctx->feat_cap) {
        uint32_t cap = 0;
        for (const Inst &J : ctx->inst) cap = std::max(cap, J.N);
        int rc;
        if ((rc = dalloc(ctx, &ctx->feat, 6 * (size_t)cap))) return rc;
        if ((rc = dalloc(ctx, &ctx->vpred, 5 * (size_t)cap))) return rc;
        ctx->feat_cap = cap;
    }
    hipLaunchKernelGGL(k_features, dim3((I.N + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, ctx->stream, devmesh(ctx),
                       I.cell0, I.N, (const float4 *)(ctx->wall + I.wall0), I.nwall, (float)cp[0], (float)cp[1],
                       ctx->feat);
    HIPC(hipGetLastError());
    return NF_OK;
}

// U_old of the reference's first RK3 step is all zeros unless the quirk is
// switched off (Q1); a given U_old must equal U, as it does after any step
int set_state(nf_ctx *ctx, int i, const double *U, const double *U_old) {
    bool zero = ctx->prm.rk3_zero_old_first_step != 0;
    if (U_old) {
        if (std::memcmp(U, U_old, (size_t)5 * ctx->inst[i].N * sizeof(double)) != 0)
            return fail(ctx, NF_E_ARG, "U_old must be NULL (zero) or equal to U");
        zero = false;
    }
    int rc = upload_state(ctx, i, U, ctx->Ua);
    if (rc) return rc;
    return after_state_set(ctx, i, zero);
}

// set_freestream_values (src/SolverData.cpp:50-58): U = prim_to_cons(V_inf),
// primvars = V_inf exactly
int set_freestream(nf_ctx *ctx, int i) {
    const Inst &I = ctx->inst[i];
    double V[5], U[5];
    std::memcpy(V, I.Vinf, sizeof V);
    prim_to_cons(V, U);
    std::vector<double> aos((size_t)5 * I.N);
    for (uint32_t c = 0; c < I.N; ++c) std::memcpy(&aos[5 * (size_t)c], U, sizeof U);
    int rc = upload_state(ctx, i, aos.data(), ctx->Ua);
    if (rc) return rc;
    if ((rc = after_state_set(ctx, i, ctx->prm.rk3_zero_old_first_step != 0))) return rc;
    if (!ctx->mixed) {
        for (uint32_t c = 0; c < I.N; ++c) std::memcpy(&aos[5 * (size_t)c], V, sizeof V);
        if ((rc = upload_state(ctx, i, aos.data(), ctx->V))) return rc;
    }
    return NF_OK;
}

double d_node(const double *nodes, uint32_t v, int d) { return nodes[3 * (size_t)v + d]; }

int build_red_table(nf_ctx *ctx, uint32_t B, RedTable &r) {
    std::vector<uint32_t> first, n, c0, nc;
    for (const Inst &I : ctx->inst) {
        const uint32_t b0 = I.cell0 / B, nb = I.Npad / B;
        c0.push_back((uint32_t)first.size());
        for (uint32_t q = 0; q < nb; q += RCHUNK) {
            first.push_back(b0 + q);
            n.push_back(std::min(RCHUNK, nb - q));
        }
        nc.push_back((uint32_t)first.size() - c0.back());
    }
    r.nchunk = (int)first.size();
    int rc;
    if ((rc = dalloc(ctx, &r.first, first.size()))) return rc;
    if ((rc = dalloc(ctx, &r.n, n.size()))) return rc;
    if ((rc = dalloc(ctx, &r.inst_c0, c0.size()))) return rc;
    if ((rc = dalloc(ctx, &r.inst_nc, nc.size()))) return rc;
    HIPC(hipMemcpy(r.first, first.data(), first.size() * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(r.n, n.data(), n.size() * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(r.inst_c0, c0.data(), c0.size() * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(r.inst_nc, nc.data(), nc.size() * 4, hipMemcpyHostToDevice));
    return NF_OK;
}

int create(int32_t n, const nf_mesh_desc *descs, const nf_params *prms, nf_ctx **out) {
    if (!descs || !prms || !out || n < 1 || n > 65535) return NF_E_ARG;
    *out = nullptr;
    const nf_params &p0 = prms[0];
    for (int i = 0; i < n; ++i) {
        const nf_params &p = prms[i];
        // HLLC (an assert(false) stub in the reference, src/Numerics.cpp:20-23)
        // is the published scheme here, in the FP64 kernels only
        if (p.inv_flux_scheme != NF_FLUX_RUSANOV && p.inv_flux_scheme != NF_FLUX_HLLC) return NF_E_ARG;
        if (p.inv_flux_scheme == NF_FLUX_HLLC && p.precision != NF_PREC_FP64) return NF_E_ARG;
        if (p.time_scheme != NF_TVD_RK3 && p.time_scheme != NF_EXPLICIT_EULER) return NF_E_ARG;
        if (p.precision != NF_PREC_FP64 && p.precision != NF_PREC_MIXED) return NF_E_ARG;
        // one launch serves all instances: everything but the freestream is shared
        if (p.time_scheme != p0.time_scheme || p.spatial_order != p0.spatial_order || p.limiter != p0.limiter ||
            p.CFL != p0.CFL || p.check_physical_validity != p0.check_physical_validity ||
            p.rk3_zero_old_first_step != p0.rk3_zero_old_first_step || p.precision != p0.precision ||
            p.inv_flux_scheme != p0.inv_flux_scheme ||
            p.cell_order != p0.cell_order || p.local_time_step != p0.local_time_step)
            return NF_E_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return NF_E_NOGPU;
    // a mixed-element mesh (hexahedra, wedges, pyramids) takes the polyhedral path
    for (int i = 0; i < n; ++i) {
        if (!descs[i].cell_types) continue;
        nf::PrimalMesh pm;
        try {
            nf::primal_from_desc(descs[i], pm);
        } catch (const nf::MeshError &) {
            return NF_E_MESH;
        }
        if (!pm.mixed()) continue;
        if (n != 1) return NF_E_ARG;
        for (int32_t b : pm.patch_bc)
            if (b < 0 || b > 2) return NF_E_ARG;
        nf::FVGrid g;
        try {
            nf::build_fv_grid(pm, g);
        } catch (const nf::MeshError &) {
            return NF_E_MESH;
        }
        std::unique_ptr<nf_ctx> hold(new nf_ctx);
        nf_ctx *ctx = hold.get();
        ctx->prm = p0;
        ctx->n_inst = 0;
        HIPC(hipGetDevice(&ctx->device));
        HIPC(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        const int rc = nf::poly_create(pm, g, p0, ctx->stream, &ctx->poly, ctx->err);
        if (rc) {
            (void)hipStreamDestroy(ctx->stream);
            return rc;
        }
        *out = hold.release();
        return NF_OK;
    }
    std::unique_ptr<nf_ctx> hold(new nf_ctx);
    nf_ctx *ctx = hold.get();
    ctx->prm = p0;
    ctx->n_inst = n;
    ctx->mixed = p0.precision == NF_PREC_MIXED;
    HIPC(hipGetDevice(&ctx->device));

    // host pipeline per instance; the device arrays are the concatenation
    std::vector<uint32_t> cnodes, nbr, fcode, bface, inst_end;
    std::vector<float4> wall;
    std::vector<uint8_t> bbc;
    std::vector<double> bS3;  // [nb][3]
    std::vector<uint16_t> binst;
    std::vector<uint32_t> blk_inst;
    std::vector<double> nodes4;
    ctx->inst.resize(n);
    uint64_t cell0 = 0, node0 = 0;
    for (int i = 0; i < n; ++i) {
        const nf_mesh_desc *d = &descs[i];
        if (!d->nodes || !d->cells || d->n_cells == 0) return NF_E_MESH;
        nf::PrimalMesh pm;
        nf::FVGrid g;
        nf::CellLayout L;
        try {
            pm.nodes.assign(d->nodes, d->nodes + 3 * (size_t)d->n_nodes);
            pm.cells.assign(d->cells, d->cells + 4 * (size_t)d->n_cells);
            pm.patch_n.assign(d->patch_n_faces, d->patch_n_faces + d->n_patches);
            size_t nt = 0;
            for (uint32_t q = 0; q < d->n_patches; ++q) nt += d->patch_n_faces[q];
            pm.patch_tri.assign(d->patch_faces, d->patch_faces + 3 * nt);
            pm.patch_bc.assign(d->patch_bc, d->patch_bc + d->n_patches);
            for (int32_t b : pm.patch_bc)
                if (b < 0 || b > 2) return NF_E_ARG;
            nf::build_fv_grid(pm, g);
            nf::build_cell_layout(pm, g, p0.cell_order, L);
        } catch (const nf::MeshError &) {
            return NF_E_MESH;
        }
        Inst &I = ctx->inst[i];
        I.cell0 = (uint32_t)cell0;
        I.N = g.N;
        I.Npad = (g.N + BLOCK - 1) / BLOCK * BLOCK;
        if (cell0 + I.Npad > NBR_INDEX_MASK || node0 + d->n_nodes > 0xFFFFFFFFull) return NF_E_ARG;
        I.sizes = {g.N, g.NT, g.Fint, g.F, (uint32_t)g.pfirst.size()};
        I.perm = L.perm;
        std::memcpy(I.Vinf, prms[i].primvars_inf, sizeof I.Vinf);
        const uint32_t b0 = (uint32_t)bface.size();
        // padding cells repeat the last cell's topology; the sweeps never
        // store to them and clamp their lanes to the instance's last cell
        for (uint32_t c = 0; c < I.Npad; ++c) {
            const uint32_t src = std::min(c, g.N - 1);
            for (int q = 0; q < 4; ++q) {
                cnodes.push_back(L.cnodes[4 * (size_t)src + q] + (uint32_t)node0);
                const uint32_t w = L.nbr[4 * (size_t)src + q];
                const uint32_t off = (w & NBR_BOUNDARY) ? b0 : (uint32_t)cell0;
                nbr.push_back((w & ~NBR_INDEX_MASK) | ((w & NBR_INDEX_MASK) + off));
            }
            fcode.push_back(L.fcode[src]);
        }
        for (size_t q = 0; q < L.bface.size(); ++q) {
            const uint32_t w = L.bface[q];
            bface.push_back((w & ~NBR_INDEX_MASK) | ((w & NBR_INDEX_MASK) + (uint32_t)cell0));
            bbc.push_back(L.bbc[q]);
            for (int d = 0; d < 3; ++d) bS3.push_back(g.S[3 * (size_t)L.bfidx[q] + d]);
            binst.push_back((uint16_t)i);
        }
        // wall-face centroids (slip and no-slip patches) for the network features
        I.wall0 = (uint32_t)wall.size();
        for (size_t q = 0; q < g.pfirst.size(); ++q) {
            if (g.pbc[q] == NF_BC_FARFIELD) continue;
            for (uint32_t f = g.pfirst[q]; f < g.pfirst[q] + g.pn[q]; ++f) {
                double cf[3] = {0, 0, 0};
                for (int k = 0; k < 3; ++k)
                    for (int e = 0; e < 3; ++e) cf[e] += d_node(d->nodes, g.fnodes[3 * (size_t)f + k], e);
                wall.push_back(make_float4((float)(cf[0] / 3), (float)(cf[1] / 3), (float)(cf[2] / 3), 0.f));
            }
        }
        I.nwall = (uint32_t)wall.size() - I.wall0;
        for (uint32_t b = 0; b < I.Npad / BLOCK; ++b) blk_inst.push_back((uint32_t)i);
        inst_end.push_back((uint32_t)cell0 + g.N);
        for (uint32_t v = 0; v < d->n_nodes; ++v)
            nodes4.insert(nodes4.end(),
                          {d->nodes[3 * (size_t)v], d->nodes[3 * (size_t)v + 1], d->nodes[3 * (size_t)v + 2], 0.0});
        cell0 += I.Npad;
        node0 += d->n_nodes;
    }
    if (bface.size() > NBR_INDEX_MASK) return NF_E_ARG;
    // node tables of the 64-cell groups (one per wave; instances are padded
    // to 256 cells): distinct nodes in first-use order
    std::vector<uint32_t> grp_off(1, 0);
    std::vector<double> gx, gy, gz;
    std::vector<uint2> cell_ln(cell0);
    {
        std::vector<uint32_t> stamp(node0, UINT32_MAX), loc(node0);
        for (uint32_t gi = 0; gi < cell0 / 64; ++gi) {
            uint32_t cnt = 0;
            for (uint32_t p = gi * 64; p < gi * 64 + 64; ++p) {
                uint32_t l[4];
                for (int q = 0; q < 4; ++q) {
                    const uint32_t v = cnodes[4 * (size_t)p + q];
                    if (stamp[v] != gi) {
                        stamp[v] = gi;
                        loc[v] = cnt++;
                        gx.push_back(nodes4[4 * (size_t)v]);
                        gy.push_back(nodes4[4 * (size_t)v + 1]);
                        gz.push_back(nodes4[4 * (size_t)v + 2]);
                    }
                    l[q] = loc[v];
                }
                cell_ln[p] = make_uint2(l[0] | l[1] << 16, l[2] | l[3] << 16);
            }
            grp_off.push_back(grp_off.back() + cnt);
        }
    }
    ctx->Ncell = (uint32_t)cell0;
    ctx->nbf = (uint32_t)bface.size();

    HIPC(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    int rc;
    const size_t NC = ctx->Ncell;
    if ((rc = dalloc(ctx, &ctx->nodes, nodes4.size()))) return rc;
    if ((rc = dalloc(ctx, &ctx->cnodes, cnodes.size()))) return rc;
    if ((rc = dalloc(ctx, &ctx->nbr, nbr.size()))) return rc;
    if ((rc = dalloc(ctx, &ctx->fcode, fcode.size()))) return rc;
    if ((rc = dalloc(ctx, &ctx->grp_off, grp_off.size()))) return rc;
    if ((rc = dalloc(ctx, &ctx->gx, gx.size()))) return rc;
    if ((rc = dalloc(ctx, &ctx->gy, gy.size()))) return rc;
    if ((rc = dalloc(ctx, &ctx->gz, gz.size()))) return rc;
    if ((rc = dalloc(ctx, &ctx->cell_ln, cell_ln.size()))) return rc;
    if ((rc = dalloc(ctx, &ctx->blk_inst, blk_inst.size()))) return rc;
    if ((rc = dalloc(ctx, &ctx->inst_end, inst_end.size()))) return rc;
    if ((rc = dalloc(ctx, &ctx->st, (size_t)n))) return rc;
    if ((rc = dalloc(ctx, &ctx->Ua, 5 * NC))) return rc;
    if ((rc = dalloc(ctx, &ctx->Ub, 5 * NC))) return rc;
    if (ctx->mixed) {
        if ((rc = dalloc(ctx, &ctx->V32, 5 * NC))) return rc;
        if ((rc = dalloc(ctx, &ctx->FV32, 20 * NC))) return rc;
        if ((rc = dalloc(ctx, &ctx->Gc32, 5 * ((size_t)ctx->nbf + TILE - 1) / TILE * TILE))) return rc;
        if ((rc = dalloc(ctx, &ctx->Gf32, 5 * ((size_t)ctx->nbf + TILE - 1) / TILE * TILE))) return rc;
    } else {
        if ((rc = dalloc(ctx, &ctx->V, 5 * NC))) return rc;
        if ((rc = dalloc(ctx, &ctx->FV, 20 * NC))) return rc;
        if ((rc = dalloc(ctx, &ctx->Gc, 5 * ((size_t)ctx->nbf + TILE - 1) / TILE * TILE))) return rc;
        if ((rc = dalloc(ctx, &ctx->Gf, 5 * ((size_t)ctx->nbf + TILE - 1) / TILE * TILE))) return rc;
    }
    if (p0.local_time_step && (rc = dalloc(ctx, &ctx->dt_cell, NC))) return rc;
    if (p0.check_physical_validity) {
        if ((rc = dalloc(ctx, &ctx->Rst, 5 * NC))) return rc;
        if ((rc = dalloc(ctx, &ctx->fst, NC))) return rc;
    }
    if ((rc = dalloc(ctx, &ctx->bface, ctx->nbf))) return rc;
    if ((rc = dalloc(ctx, &ctx->bbc, ctx->nbf))) return rc;
    if ((rc = dalloc(ctx, &ctx->bS, 3 * (size_t)ctx->nbf))) return rc;
    if ((rc = dalloc(ctx, &ctx->binst, ctx->nbf))) return rc;
    if ((rc = dalloc(ctx, &ctx->fs, (size_t)n))) return rc;
    if ((rc = dalloc(ctx, &ctx->fs32, (size_t)n))) return rc;
    if ((rc = dalloc(ctx, &ctx->dt_part, NC / 128))) return rc;
    if ((rc = dalloc(ctx, &ctx->res_part, NC / 128))) return rc;
    HIPC(hipMemset(ctx->Ua, 0, 5 * NC * sizeof(double)));
    HIPC(hipMemset(ctx->Ub, 0, 5 * NC * sizeof(double)));
    HIPC(hipMemcpy(ctx->nodes, nodes4.data(), nodes4.size() * 8, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(ctx->cnodes, cnodes.data(), cnodes.size() * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(ctx->nbr, nbr.data(), nbr.size() * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(ctx->fcode, fcode.data(), fcode.size() * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(ctx->grp_off, grp_off.data(), grp_off.size() * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(ctx->gx, gx.data(), gx.size() * 8, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(ctx->gy, gy.data(), gy.size() * 8, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(ctx->gz, gz.data(), gz.size() * 8, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(ctx->cell_ln, cell_ln.data(), cell_ln.size() * sizeof(uint2), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(ctx->blk_inst, blk_inst.data(), blk_inst.size() * 4, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(ctx->inst_end, inst_end.data(), inst_end.size() * 4, hipMemcpyHostToDevice));
    if (ctx->nbf) {
        HIPC(hipMemcpy(ctx->bface, bface.data(), bface.size() * 4, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(ctx->bbc, bbc.data(), bbc.size(), hipMemcpyHostToDevice));
        std::vector<double> bS((size_t)3 * ctx->nbf);
        for (size_t q = 0; q < ctx->nbf; ++q)
            for (int d = 0; d < 3; ++d) bS[d * (size_t)ctx->nbf + q] = bS3[3 * q + d];
        HIPC(hipMemcpy(ctx->bS, bS.data(), bS.size() * 8, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(ctx->binst, binst.data(), binst.size() * 2, hipMemcpyHostToDevice));
    }
    {
        std::vector<FarField<double>> f64(n);
        std::vector<FarField<float>> f32(n);
        for (int i = 0; i < n; ++i) {
            f64[i] = make_farfield<double>(prms[i].primvars_inf);
            f32[i] = make_farfield<float>(prms[i].primvars_inf);
        }
        HIPC(hipMemcpy(ctx->fs, f64.data(), n * sizeof(FarField<double>), hipMemcpyHostToDevice));
        HIPC(hipMemcpy(ctx->fs32, f32.data(), n * sizeof(FarField<float>), hipMemcpyHostToDevice));
    }
    if ((rc = build_red_table(ctx, 256, ctx->red256))) return rc;
    if ((rc = build_red_table(ctx, 128, ctx->red128))) return rc;
    if ((rc = dalloc(ctx, &ctx->red_dt, (size_t)ctx->red128.nchunk))) return rc;
    if ((rc = dalloc(ctx, &ctx->red_res, (size_t)ctx->red128.nchunk))) return rc;
    if ((rc = dalloc(ctx, &ctx->r_eval, (size_t)n))) return rc;
    if ((rc = dalloc(ctx, &ctx->wall, wall.size()))) return rc;
    if (!wall.empty()) HIPC(hipMemcpy(ctx->wall, wall.data(), wall.size() * sizeof(float4), hipMemcpyHostToDevice));
    if ((rc = ensure_hist(ctx, 1024))) return rc;
    for (int i = 0; i < n; ++i)
        if ((rc = set_freestream(ctx, i))) return rc;
    *out = hold.release();
    return NF_OK;
}

const char *POLY_UNSUPPORTED =
    "not available for a mixed-element mesh (supported: set_freestream, set_state, calc_timestep, step, run, "
    "get_residuals, time_kernels, eval_flux_balance, get_state, get_field, ctx_sizes)";
bool single(const nf_ctx *ctx) {
    if (ctx && ctx->poly) ctx->err = POLY_UNSUPPORTED;
    return ctx && ctx->n_inst == 1;
}
bool valid_inst(const nf_ctx *ctx, int32_t i) {
    if (ctx && ctx->poly) ctx->err = POLY_UNSUPPORTED;
    return ctx && i >= 0 && i < ctx->n_inst;
}

// one step of a single-instance context with the dt held on the device
int step_device_dt(nf_ctx *ctx) {
    int rc;
    if ((rc = clear_flags(ctx))) return rc;
    std::vector<InstState> h;
    if ((rc = read_states(ctx, h))) return rc;
    if ((rc = ensure_hist(ctx, h[0].nstep + 2))) return rc;
    if ((rc = enqueue_step(ctx, -1.0))) return rc;
    if (ctx->timing && (rc = flush_timing(ctx))) return rc;
    ctx->inst[0].dt_valid = true;
    ctx->aux_valid = false;
    return check_errors(ctx);
}

// device-resident loop over all instances, K steps per graph launch; the host
// reads the run flags between launches when a residual drop or the stall
// fallback is set
int run_loop(nf_ctx *ctx, const nf_run_opts *o, nf_run_stats *stats) {
    if (!o || o->n_steps < 0) return NF_E_ARG;
    int rc;
    const int K = std::max(1, o->check_every);
    const int n = ctx->n_inst;
    // with a residual drop, an instance that already reached it stays stopped
    // (a run can be continued in chunks); its state must be set to run again
    if ((rc = clear_flags(ctx, o->residual_drop > 0.0))) return rc;
    std::vector<bool> need(n);
    for (int i = 0; i < n; ++i) need[i] = !ctx->inst[i].dt_valid;
    if (std::any_of(need.begin(), need.end(), [](bool b) { return b; }) && (rc = timestep_subset(ctx, need)))
        return rc;
    for (Inst &I : ctx->inst) I.dt_valid = true;
    std::vector<InstState> h0;
    if ((rc = read_states(ctx, h0))) return rc;
    int maxstep = 0;
    for (auto &s : h0) maxstep = std::max(maxstep, s.nstep);
    if ((rc = ensure_hist(ctx, maxstep + o->n_steps + 2))) return rc;
    const double drop = o->residual_drop;
    const bool stall = o->stall_window > 0;
    std::vector<int> base(n), acc(n, 0), fell(n, 0);
    for (int i = 0; i < n; ++i) base[i] = h0[i].nstep;
    hipEvent_t e0, e1;
    HIPC(hipEventCreate(&e0));
    HIPC(hipEventCreate(&e1));
    HIPC(hipEventRecord(e0, ctx->stream));
    int done = 0;
    bool stopped = false;
    while (done < o->n_steps && !stopped) {
        const int chunk = std::min(K, o->n_steps - done);
        if (o->use_graph && chunk == K) {
            if (!ctx->graph || ctx->graph_steps != K || ctx->graph_drop != drop || ctx->graph_block != ctx->block ||
                ctx->graph_xcd != ctx->xcd_remap) {
                if (ctx->graph) (void)hipGraphExecDestroy(ctx->graph);
                ctx->graph = nullptr;
                hipGraph_t gr;
                HIPC(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                for (int s = 0; s < K; ++s)
                    if ((rc = enqueue_step(ctx, drop))) {
                        (void)hipStreamEndCapture(ctx->stream, &gr);
                        return rc;
                    }
                HIPC(hipStreamEndCapture(ctx->stream, &gr));
                HIPC(hipGraphInstantiate(&ctx->graph, gr, nullptr, nullptr, 0));
                (void)hipGraphDestroy(gr);
                ctx->graph_steps = K;
                ctx->graph_drop = drop;
                ctx->graph_block = ctx->block;
                ctx->graph_xcd = ctx->xcd_remap;
            }
            HIPC(hipGraphLaunch(ctx->graph, ctx->stream));
            if (!ctx->mixed)  // (the capture marked them once; a replay steps again)
                for (Inst &I : ctx->inst) I.v_stale = true;
        } else {
            for (int s = 0; s < chunk; ++s)
                if ((rc = enqueue_step(ctx, drop))) return rc;
        }
        done += chunk;
        if ((drop > 0.0 || stall) && done < o->n_steps) {
            std::vector<InstState> h;
            if ((rc = read_states(ctx, h))) return rc;
            if (stall) {
                // one check per warm start, at the first sync point at least
                // stall_window steps after it: accepted or replaced
                std::vector<bool> reset(n, false);
                for (int i = 0; i < n; ++i) {
                    const int k = h[i].nstep;
                    if (!ctx->inst[i].warm || k <= o->stall_window || k > ctx->hist_cap) continue;
                    double r_now = 0, r_then = 0;
                    const double *hp = ctx->hist + (size_t)i * ctx->hist_cap;
                    HIPC(hipMemcpy(&r_now, hp + k - 1, sizeof(double), hipMemcpyDeviceToHost));
                    HIPC(hipMemcpy(&r_then, hp + k - 1 - o->stall_window, sizeof(double), hipMemcpyDeviceToHost));
                    if (r_now <= o->stall_ratio * r_then || h[i].converged) ctx->inst[i].warm = false;
                    else reset[i] = true;
                }
                if (std::any_of(reset.begin(), reset.end(), [](bool b) { return b; })) {
                    for (int i = 0; i < n; ++i) {
                        if (!reset[i]) continue;
                        acc[i] += h[i].nstep - base[i];
                        base[i] = 0;
                        fell[i] = 1;
                        if ((rc = set_freestream(ctx, i))) return rc;
                    }
                    if ((rc = timestep_subset(ctx, reset))) return rc;
                    for (Inst &I : ctx->inst) I.dt_valid = true;
                    if ((rc = read_states(ctx, h))) return rc;
                }
            }
            stopped = true;
            for (auto &s : h) stopped &= s.stop != 0;
        }
    }
    HIPC(hipEventRecord(e1, ctx->stream));
    HIPC(hipEventSynchronize(e1));
    float ms = 0;
    HIPC(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->aux_valid = false;
    std::vector<InstState> h;
    if ((rc = read_states(ctx, h))) return rc;
    if (stats) {
        for (int i = 0; i < n; ++i) {
            nf_run_stats &s = stats[i];
            s.steps_done = acc[i] + h[i].nstep - base[i];
            s.converged = h[i].converged;
            s.time = h[i].time;
            s.last_dt = h[i].dt;
            s.elapsed_ms = ms;
            s.fell_back = fell[i];
            s.pad = 0;
            double r[2] = {0, 0};
            const double *hp = ctx->hist + (size_t)i * ctx->hist_cap;
            HIPC(hipMemcpy(&r[0], hp, sizeof(double), hipMemcpyDeviceToHost));
            if (h[i].nstep > 0)
                HIPC(hipMemcpy(&r[1], hp + std::min(h[i].nstep - 1, ctx->hist_cap - 1), sizeof(double),
                               hipMemcpyDeviceToHost));
            s.res_first = r[0];
            s.res_last = r[1];
        }
    }
    return check_errors(ctx);
}

// saved copy of the state and run counters (autotuning runs on the copy)
struct Snapshot {
    std::vector<double> U, V;
    std::vector<float> V32;
    std::vector<InstState> st;
    std::vector<bool> dt_valid, v_stale;
};

int save(nf_ctx *ctx, Snapshot &s) {
    const size_t n5 = 5 * (size_t)ctx->Ncell;
    int rc;
    if ((rc = read_states(ctx, s.st))) return rc;
    s.U.resize(n5);
    HIPC(hipMemcpy(s.U.data(), ctx->Ua, n5 * sizeof(double), hipMemcpyDeviceToHost));
    if (ctx->mixed) {
        s.V32.resize(n5);
        HIPC(hipMemcpy(s.V32.data(), ctx->V32, n5 * sizeof(float), hipMemcpyDeviceToHost));
    } else {
        s.V.resize(n5);
        HIPC(hipMemcpy(s.V.data(), ctx->V, n5 * sizeof(double), hipMemcpyDeviceToHost));
    }
    s.dt_valid.clear();
    s.v_stale.clear();
    for (auto &I : ctx->inst) {
        s.dt_valid.push_back(I.dt_valid);
        s.v_stale.push_back(I.v_stale);
    }
    return NF_OK;
}

int restore(nf_ctx *ctx, const Snapshot &s) {
    const size_t n5 = 5 * (size_t)ctx->Ncell;
    HIPC(hipStreamSynchronize(ctx->stream));
    HIPC(hipMemcpy(ctx->Ua, s.U.data(), n5 * sizeof(double), hipMemcpyHostToDevice));
    if (ctx->mixed) HIPC(hipMemcpy(ctx->V32, s.V32.data(), n5 * sizeof(float), hipMemcpyHostToDevice));
    else HIPC(hipMemcpy(ctx->V, s.V.data(), n5 * sizeof(double), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(ctx->st, s.st.data(), s.st.size() * sizeof(InstState), hipMemcpyHostToDevice));
    for (size_t i = 0; i < ctx->inst.size(); ++i) {
        ctx->inst[i].dt_valid = s.dt_valid[i];
        ctx->inst[i].v_stale = s.v_stale[i];
    }
    ctx->aux_valid = false;
    return NF_OK;
}

}  // namespace

// ============================================================================
extern "C" {

int nf_create(const nf_mesh_desc *desc, const nf_params *prm, nf_ctx **out) { return create(1, desc, prm, out); }

int nf_create_batch(int32_t n, const nf_mesh_desc *meshes, const nf_params *params, nf_ctx **out) {
    return create(n, meshes, params, out);
}

int nf_destroy(nf_ctx *ctx) {
    if (!ctx) return NF_OK;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    nf::poly_destroy(ctx->poly);
    if (ctx->graph) (void)hipGraphExecDestroy(ctx->graph);
    for (void *p : ctx->owned) (void)hipFree(p);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return NF_OK;
}

const char *nf_last_error(const nf_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int nf_ctx_sizes(const nf_ctx *ctx, nf_fv_sizes *s) {
    if (!ctx || !s) return NF_E_ARG;
    if (ctx->poly) {
        *s = nf::poly_sizes(ctx->poly);
        return NF_OK;
    }
    *s = ctx->inst[0].sizes;
    return NF_OK;
}

int nf_ctx_instances(const nf_ctx *ctx, int32_t *n) {
    if (!ctx || !n) return NF_E_ARG;
    *n = ctx->n_inst;
    return NF_OK;
}

int nf_inst_sizes(const nf_ctx *ctx, int32_t i, nf_fv_sizes *s) {
    if (!valid_inst(ctx, i) || !s) return NF_E_ARG;
    *s = ctx->inst[i].sizes;
    return NF_OK;
}

int nf_set_freestream(nf_ctx *ctx) {
    if (!ctx) return NF_E_ARG;
    if (ctx->poly) return nf::poly_set_freestream(ctx->poly, ctx->err);
    for (int i = 0; i < ctx->n_inst; ++i) {
        int rc = set_freestream(ctx, i);
        if (rc) return rc;
    }
    return NF_OK;
}

int nf_set_state(nf_ctx *ctx, const double *U, const double *U_old) {
    if (ctx && ctx->poly) return U ? nf::poly_set_state(ctx->poly, U, U_old, ctx->err) : NF_E_ARG;
    if (!single(ctx) || !U) return NF_E_ARG;
    return set_state(ctx, 0, U, U_old);
}

int nf_inst_set_state(nf_ctx *ctx, int32_t i, const double *U, const double *U_old) {
    if (!valid_inst(ctx, i) || !U) return NF_E_ARG;
    return set_state(ctx, i, U, U_old);
}

int nf_calc_timestep(nf_ctx *ctx, double *dt_out) {
    if (ctx && ctx->poly) return nf::poly_calc_timestep(ctx->poly, dt_out, ctx->err);
    if (!single(ctx)) return NF_E_ARG;
    int rc;
    if (!ctx->inst[0].dt_valid) {
        if ((rc = clear_flags(ctx))) return rc;
        if ((rc = enqueue_timestep(ctx))) return rc;
        ctx->inst[0].dt_valid = true;
    }
    std::vector<InstState> h;
    if ((rc = read_states(ctx, h))) return rc;
    const double dt = h[0].dt;
    if (dt_out) *dt_out = dt;
    if (!((dt - dt == 0.0) && dt > 0.0)) {
        char buf[96];
        snprintf(buf, sizeof buf, "Invalid dt calculated (dt = %f)", dt);
        return fail(ctx, NF_E_INVALID_DT, buf);
    }
    return NF_OK;
}

int nf_step(nf_ctx *ctx, double dt) {
    if (ctx && ctx->poly) return nf::poly_step(ctx->poly, dt, ctx->err);
    if (!single(ctx)) return NF_E_ARG;
    hipLaunchKernelGGL(k_flags, dim3(1), dim3(64), 0, ctx->stream, ctx->st, 1, 1, 0, dt);
    HIPC(hipGetLastError());
    return step_device_dt(ctx);
}

int nf_run(nf_ctx *ctx, const nf_run_opts *o, nf_run_stats *st) {
    if (ctx && ctx->poly) return nf::poly_run(ctx->poly, o, st, ctx->err);
    if (!single(ctx)) return NF_E_ARG;
    return run_loop(ctx, o, st);
}

int nf_run_batch(nf_ctx *ctx, const nf_run_opts *o, nf_run_stats *stats) {
    if (ctx && ctx->poly) return fail(ctx, NF_E_ARG, POLY_UNSUPPORTED);
    if (!ctx) return NF_E_ARG;
    return run_loop(ctx, o, stats);
}

int nf_inst_residuals(nf_ctx *ctx, int32_t i, double *out, int32_t max, int32_t *n_out) {
    if (!valid_inst(ctx, i) || !out) return NF_E_ARG;
    std::vector<InstState> h;
    int rc;
    if ((rc = read_states(ctx, h))) return rc;
    const int n = std::min(std::min(h[i].nstep, max), ctx->hist_cap);
    if (n > 0)
        HIPC(hipMemcpy(out, ctx->hist + (size_t)i * ctx->hist_cap, (size_t)n * sizeof(double),
                       hipMemcpyDeviceToHost));
    if (n_out) *n_out = std::max(n, 0);
    return NF_OK;
}

int nf_get_residuals(nf_ctx *ctx, double *out, int32_t max, int32_t *n_out) {
    if (ctx && ctx->poly) return out ? nf::poly_get_residuals(ctx->poly, out, max, n_out) : NF_E_ARG;
    if (!single(ctx)) return NF_E_ARG;
    return nf_inst_residuals(ctx, 0, out, max, n_out);
}

int nf_get_state(nf_ctx *ctx, double *U) {
    if (ctx && ctx->poly) return U ? nf::poly_get_state(ctx->poly, U, ctx->err) : NF_E_ARG;
    if (!single(ctx) || !U) return NF_E_ARG;
    return download_state(ctx, 0, (const double *)ctx->Ua, U);
}

int nf_inst_get_state(nf_ctx *ctx, int32_t i, double *U) {
    if (!valid_inst(ctx, i) || !U) return NF_E_ARG;
    return download_state(ctx, i, (const double *)ctx->Ua, U);
}

// Evaluates the flux balance of the current state once (Solver::
// evaluate_flux_balance, src/Solver.cpp:54-81) without updating U, keeping
// gradient, limiter and R for nf_get_field.
int nf_eval_flux_balance(nf_ctx *ctx) {
    if (ctx && ctx->poly) return nf::poly_eval_flux_balance(ctx->poly, ctx->err);
    if (!single(ctx)) return NF_E_ARG;
    int rc;
    const size_t NC = ctx->Ncell;
    if (!ctx->R_aux) {
        if ((rc = dalloc(ctx, &ctx->grad_aux, 15 * NC))) return rc;
        if ((rc = dalloc(ctx, &ctx->psi_aux, 5 * NC))) return rc;
        if ((rc = dalloc(ctx, &ctx->R_aux, 5 * NC))) return rc;
    }
    if ((rc = clear_flags(ctx))) return rc;
    if ((rc = prim_all(ctx))) return rc;
    if ((rc = run_sweepA(ctx, false, true))) return rc;
    if ((rc = run_sweepB(ctx, 1, ctx->Ua, ctx->Ub, false, true))) return rc;
    // the sweep wrote prim of the discarded update: primvars = prim(U) again,
    // which is what they were before the call.  The dt the last step left on
    // the device (from prim(U2) for TVD-RK3) stays valid, so an evaluation
    // between steps does not change the run.
    if ((rc = prim_all(ctx))) return rc;
    ctx->aux_valid = true;
    return check_errors(ctx);
}

int nf_get_field(nf_ctx *ctx, int32_t which, double *out) {
    if (ctx && ctx->poly) return out ? nf::poly_get_field(ctx->poly, which, out, ctx->err) : NF_E_ARG;
    if (!single(ctx) || !out) return NF_E_ARG;
    switch (which) {
    case NF_FIELD_PRIM: {
        if (ctx->mixed) return download_state(ctx, 0, (const float *)ctx->V32, out);
        int rc;
        if ((rc = refresh_v(ctx))) return rc;
        return download_state(ctx, 0, (const double *)ctx->V, out);
    }
    case NF_FIELD_STATE_OLD: {
        std::vector<InstState> h;
        int rc;
        if ((rc = read_states(ctx, h))) return rc;
        if (h[0].uold_zero) {
            std::memset(out, 0, sizeof(double) * 5 * ctx->inst[0].N);
            return NF_OK;
        }
        return download_state(ctx, 0, (const double *)ctx->Ua, out);
    }
    case NF_FIELD_LIMITER:
    case NF_FIELD_GRADIENT:
    case NF_FIELD_RESIDUAL:
        if (!ctx->aux_valid) return fail(ctx, NF_E_ARG, "call nf_eval_flux_balance first");
        if (which == NF_FIELD_LIMITER) return download_aos(ctx, 0, ctx->psi_aux, 5, out);
        if (which == NF_FIELD_GRADIENT) return download_aos(ctx, 0, ctx->grad_aux, 15, out);
        return download_aos(ctx, 0, ctx->R_aux, 5, out);
    default:
        return fail(ctx, NF_E_ARG, "unknown field");
    }
}

int nf_time_kernels(nf_ctx *ctx, int32_t n_steps, nf_kernel_time *out, int32_t max, int32_t *n_out) {
    if (ctx && ctx->poly) return nf::poly_time_kernels(ctx->poly, n_steps, out, max, n_out, ctx->err);
    if (!single(ctx) || n_steps < 1 || !out) return NF_E_ARG;
    int rc;
    if (!ctx->inst[0].dt_valid) {
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

int nf_inst_set_residual_ref(nf_ctx *ctx, int32_t i, double ref) {
    if (!valid_inst(ctx, i)) return NF_E_ARG;
    ctx->inst[i].rref = ref > 0.0 ? ref : 0.0;
    std::vector<InstState> h;
    int rc;
    if ((rc = read_states(ctx, h))) return rc;
    h[i].rref = ctx->inst[i].rref;
    HIPC(hipMemcpy(ctx->st + i, &h[i], sizeof(InstState), hipMemcpyHostToDevice));
    return NF_OK;
}

int nf_eval_residuals(nf_ctx *ctx, double *out) {
    if (ctx && ctx->poly) return fail(ctx, NF_E_ARG, POLY_UNSUPPORTED);
    if (!ctx || !out) return NF_E_ARG;
    return eval_residuals(ctx, out);
}

int nf_inst_features(nf_ctx *ctx, int32_t i, const double *cp, float *x) {
    if (!valid_inst(ctx, i) || !cp || !x) return NF_E_ARG;
    int rc;
    if ((rc = features(ctx, i, cp))) return rc;
    const Inst &I = ctx->inst[i];
    std::vector<float> fm((size_t)6 * I.N);
    HIPC(hipStreamSynchronize(ctx->stream));
    HIPC(hipMemcpy(fm.data(), ctx->feat, fm.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (uint32_t c = 0; c < I.N; ++c)
        for (int f = 0; f < 6; ++f) x[6 * (size_t)I.perm[c] + f] = fm[(size_t)f * I.N + c];
    return NF_OK;
}

int nf_inst_warm_start(nf_ctx *ctx, int32_t i, nf_mlp *net, const double *cp, int32_t *fell_back) {
    if (!valid_inst(ctx, i) || !net || !cp) return NF_E_ARG;
    int rc;
    if ((rc = features(ctx, i, cp))) return rc;
    Inst &I = ctx->inst[i];
    HIPC(nf::mlp::predict_device(net, ctx->feat, I.N, ctx->vpred, ctx->stream));
    if ((rc = clear_flags(ctx))) return rc;
    hipLaunchKernelGGL(k_warm_state, dim3((I.N + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, ctx->stream,
                       (const float *)ctx->vpred, I.N, ctx->Ua + (size_t)5 * I.cell0, ctx->st, i);
    HIPC(hipGetLastError());
    std::vector<InstState> h;
    if ((rc = read_states(ctx, h))) return rc;
    const bool bad = h[i].bad_state > 0;
    if (fell_back) *fell_back = bad ? 1 : 0;
    if (bad) return set_freestream(ctx, i);
    // U_old = U: the first RK3 step must not scale a warm start
    if ((rc = after_state_set(ctx, i, false))) return rc;
    I.warm = true;
    return NF_OK;
}

int nf_set_precision(nf_ctx *ctx, int32_t precision) {
    if (ctx && ctx->poly) return fail(ctx, NF_E_ARG, POLY_UNSUPPORTED);
    if (!ctx || (precision != NF_PREC_FP64 && precision != NF_PREC_MIXED)) return NF_E_ARG;
    const bool mixed = precision == NF_PREC_MIXED;
    if (mixed == ctx->mixed) return NF_OK;
    if (mixed && ctx->prm.inv_flux_scheme == NF_FLUX_HLLC)
        return fail(ctx, NF_E_ARG, "the HLLC flux is implemented in FP64 only");
    int rc;
    const size_t NC = ctx->Ncell;
    if (mixed && !ctx->V32) {
        if ((rc = dalloc(ctx, &ctx->V32, 5 * NC))) return rc;
        if ((rc = dalloc(ctx, &ctx->FV32, 20 * NC))) return rc;
        if ((rc = dalloc(ctx, &ctx->Gc32, 5 * ((size_t)ctx->nbf + TILE - 1) / TILE * TILE))) return rc;
        if ((rc = dalloc(ctx, &ctx->Gf32, 5 * ((size_t)ctx->nbf + TILE - 1) / TILE * TILE))) return rc;
    }
    if (!mixed && !ctx->V) {
        if ((rc = dalloc(ctx, &ctx->V, 5 * NC))) return rc;
        if ((rc = dalloc(ctx, &ctx->FV, 20 * NC))) return rc;
        if ((rc = dalloc(ctx, &ctx->Gc, 5 * ((size_t)ctx->nbf + TILE - 1) / TILE * TILE))) return rc;
        if ((rc = dalloc(ctx, &ctx->Gf, 5 * ((size_t)ctx->nbf + TILE - 1) / TILE * TILE))) return rc;
    }
    HIPC(hipStreamSynchronize(ctx->stream));
    ctx->mixed = mixed;
    ctx->prm.precision = precision;
    if (ctx->graph) {  // captured with the other precision's kernels and buffers
        (void)hipGraphExecDestroy(ctx->graph);
        ctx->graph = nullptr;
    }
    // primvars of the current states in the new precision; dt from them;
    // every instance runs again (converged flags cleared)
    if ((rc = clear_flags(ctx))) return rc;
    if ((rc = prim_all(ctx))) return rc;
    for (Inst &I : ctx->inst) I.dt_valid = false;
    ctx->aux_valid = false;
    HIPC(hipStreamSynchronize(ctx->stream));
    return NF_OK;
}

int nf_set_launch(nf_ctx *ctx, int32_t block, int32_t xcd_remap) {
    if (ctx && ctx->poly) return fail(ctx, NF_E_ARG, POLY_UNSUPPORTED);
    if (!ctx || (block != 128 && block != 256)) return NF_E_ARG;
    ctx->block = block;
    ctx->xcd_remap = xcd_remap ? 1 : 0;
    return NF_OK;
}

// Times every launch configuration (sweep block 256/128 x XCD remap on/off)
// over `steps` graph-launched steps of the current state, keeps the fastest,
// and restores the state, run counters and dt afterwards.  Results do not
// depend on the configuration.
int nf_autotune(nf_ctx *ctx, int32_t steps, nf_tune_result *out) {
    if (ctx && ctx->poly) return fail(ctx, NF_E_ARG, POLY_UNSUPPORTED);
    if (!ctx || steps < 1) return NF_E_ARG;
    int rc;
    Snapshot snap;
    if ((rc = save(ctx, snap))) return rc;
    static const int cand[4][2] = {{256, 1}, {256, 0}, {128, 1}, {128, 0}};
    nf_tune_result r{};
    r.ms_per_step = 1e300;
    r.block = ctx->block;
    r.xcd_remap = ctx->xcd_remap;
    for (int q = 0; q < 4; ++q) {
        ctx->block = cand[q][0];
        ctx->xcd_remap = cand[q][1];
        nf_run_opts o{steps, 0.0, steps, 1};
        std::vector<nf_run_stats> s(ctx->n_inst);
        // the first run builds the graph and warms up; the second is timed
        for (int pass = 0; pass < 2; ++pass) {
            if ((rc = restore(ctx, snap))) return rc;
            if ((rc = run_loop(ctx, &o, s.data()))) {
                (void)restore(ctx, snap);
                return rc;
            }
        }
        const double ms = s[0].elapsed_ms / steps;
        r.tried_ms[q] = ms;
        if (ms < r.ms_per_step) {
            r.ms_per_step = ms;
            r.block = cand[q][0];
            r.xcd_remap = cand[q][1];
        }
    }
    r.n_tried = 4;
    ctx->block = r.block;
    ctx->xcd_remap = r.xcd_remap;
    if ((rc = restore(ctx, snap))) return rc;
    if (out) *out = r;
    return NF_OK;
}

}  // extern "C"
