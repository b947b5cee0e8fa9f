// Mixed-element FP64 step on the GPU (poly.hpp).  Per-cell slot layout:
// slot s of cell c is its s-th face in increasing reference face index, and
// every per-slot array is [slot][...][N] so a wave's 64 cells read each
// slot field with one coalesced load.
//   w      [6][N]     NONE, or bit 31 = the cell is the face's j side,
//                     bit 30 = boundary, bits 0-29 = neighbour cell or ghost id
//   ns     [6][N]     interior: the neighbour's slot of the same face;
//                     boundary: the boundary type
//   geo    [6][6][N]  S of the face (as stored, i -> j), then r of this side
//   U, U_old, R  [5][N];  V [5][NT] (cells, then ghosts);  FV [6][5][N]
//   reconstructed face values of each slot;  grad [15][N], psi [5][N]
// Kernels per stage: prim (U -> V), cell ghosts, sweep A (gradient, min/max,
// limiter, face values; second order only), sweep B (flux balance and the
// stage update).  Each follows the CPU solver's loops (src/Solver.cpp:54-230,
// include/Numerics.hpp:86-292) with the same operations in the same order;
// gradient, limiter and R are stored only when they are to be exported.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <vector>

#include "fp64_math.hpp"
#include "physics.hpp"
#include "poly.hpp"

namespace nf {

namespace {
constexpr int PB = 256;
constexpr uint32_t NONE = 0xFFFFFFFFu, SIDE_J = 0x80000000u;
using V3 = Vec3T<double>;

struct PolyDev {
    uint32_t N, NT, G;
    const uint32_t *w;
    const uint8_t *ns;
    const double *geo, *vol, *dS;
    // boundary faces, for the cell ghosts: cell i, ghost j, S, type
    const uint32_t *bi, *bj;
    const double *bS;
    const uint8_t *bbc;
    const int *stop;  // device-resident run: every kernel returns once it is set (null: no check)
};
constexpr uint32_t BND = 0x40000000u, IDX = 0x3FFFFFFFu;

__device__ __forceinline__ bool stopped(const int *stop) { return stop && *stop; }

__device__ __forceinline__ double geo(const PolyDev &m, int s, int q, uint32_t c) {
    return m.geo[((size_t)s * 6 + q) * m.N + c];
}

__global__ __launch_bounds__(PB) void k_poly_prim(const double *U, double *V, uint32_t N, uint32_t NT,
                                                  const int *stop) {
    const uint32_t c = blockIdx.x * PB + threadIdx.x;
    if (c >= N || stopped(stop)) return;
    double u[5], v[5];
    for (int k = 0; k < 5; ++k) u[k] = U[k * (size_t)N + c];
    cons_to_prim_d(u, v);
    for (int k = 0; k < 5; ++k) V[k * (size_t)NT + c] = v[k];
}

// set_constant_ghost_values (src/Solver.cpp:300-330)
__global__ __launch_bounds__(PB) void k_poly_ghost(PolyDev m, double *V, FarField<double> fs) {
    const uint32_t b = blockIdx.x * PB + threadIdx.x;
    if (b >= m.G || stopped(m.stop)) return;
    const uint32_t i = m.bi[b], j = m.bj[b];
    double vd[5], vg[5];
    for (int k = 0; k < 5; ++k) vd[k] = V[k * (size_t)m.NT + i];
    const V3 S = {m.bS[b], m.bS[m.G + b], m.bS[2 * (size_t)m.G + b]};
    ghost_value<double>((int)m.bbc[b], vd, S, fs, vg);
    for (int k = 0; k < 5; ++k) V[k * (size_t)m.NT + j] = vg[k];
}

// Green-Gauss (Numerics.hpp:86-128), max/min (:253-292), Barth (:148-250)
// and the reconstructed face values (Solver.hpp:81-112) of one cell.  The
// gradient terms divide by the cell volume through one refined reciprocal
// (rdiv: the quotient of a plain division), the Barth limiter takes two
// divisions per variable over the largest positive and most negative face
// Delta_2 (barth_psi_hilo), and the face-centroid vectors are read once.
__global__ __launch_bounds__(PB) void k_poly_sweepA(PolyDev m, const double *V, int limiter, double *FV,
                                                    double *grad, double *psi) {
    const uint32_t c = logical_block(blockIdx.x, gridDim.x, 1) * PB + threadIdx.x;  // XCD-contiguous
    if (c >= m.N || stopped(m.stop)) return;
    const size_t NT = m.NT, N = m.N;
    double vc[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) vc[k] = V[k * NT + c];
    const Rcp ivol = rcp_of(m.vol[c]);
    uint32_t w[6];
    int nfc = 0;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        w[s] = m.w[s * N + c];
        if (w[s] != NONE) nfc = s + 1;
    }
    double g[15];
#pragma unroll
    for (int q = 0; q < 15; ++q) g[q] = 0.0;
    double vmax[5], vmin[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) vmax[k] = -DBL_MAX, vmin[k] = DBL_MAX;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        // loads before the slot test (absent slots read the cell's own
        // entries, in bounds), so that later faces' loads need not wait
        const uint32_t n = s < nfc ? (w[s] & IDX) : c;
        const double Sx = geo(m, s, 0, c), Sy = geo(m, s, 1, c), Sz = geo(m, s, 2, c);
        double vnb[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) vnb[k] = V[k * NT + n];
        if (s >= nfc) break;
        const bool sj = (w[s] & SIDE_J) != 0;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const double vn = vnb[k];
            const double vi = sj ? vn : vc[k], vj = sj ? vc[k] : vn;  // the face's i and j sides
            const double uf = 0.5 * (vi + vj);
            const double t0 = rdiv(uf * Sx, ivol), t1 = rdiv(uf * Sy, ivol), t2 = rdiv(uf * Sz, ivol);
            if (!sj) {
                g[k] += t0;
                g[k + 5] += t1;
                g[k + 10] += t2;
            } else {
                g[k] -= t0;
                g[k + 5] -= t1;
                g[k + 10] -= t2;
            }
            vmax[k] = smax(vmax[k], smax(vi, vj));
            vmin[k] = smin(vmin[k], smin(vi, vj));
        }
    }
    V3d r[6];
#pragma unroll
    for (int s = 0; s < 6; ++s)
        r[s] = s < nfc ? V3d{geo(m, s, 3, c), geo(m, s, 4, c), geo(m, s, 5, c)} : V3d{0.0, 0.0, 0.0};
    double ps[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        ps[k] = 1.0;  // SolverData.cpp:17 (no limiter)
        if (limiter) {
            double hi = 0.0, lo = 0.0;
#pragma unroll
            for (int s = 0; s < 6; ++s) {
                if (s >= nfc) break;
                const double D = (g[k] * r[s].x + g[k + 5] * r[s].y) + g[k + 10] * r[s].z;
                hi = dmax(hi, D);
                lo = dmin(lo, D);
            }
            ps[k] = barth_psi_hilo(hi, lo, vmax[k], vmin[k], vc[k]);
        }
    }
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        if (s >= nfc) break;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const double D = (g[k] * r[s].x + g[k + 5] * r[s].y) + g[k + 10] * r[s].z;
            FV[((size_t)s * 5 + k) * N + c] = vc[k] + ps[k] * D;
        }
    }
    if (grad) {
        for (int q = 0; q < 15; ++q) grad[q * N + c] = g[q];
        for (int k = 0; k < 5; ++k) psi[k * N + c] = ps[k];
    }
}

struct StageArgs {
    const double *V, *FV;  // FV == nullptr: first order (face values = cell values)
    double *U;
    const double *Uo;
    double *R;     // stored when not null
    double *Rrho;  // the density flux balance, stored when not null (residual norm)
    int *bad;
    int flux;    // 0 Rusanov, 1 HLLC
    int kind;    // 0: U + dt/vol R, 1: RK3 stage 2, 2: RK3 stage 3, -1: no update
    double dt;
    const double *dtp;  // the step's dt in device memory (device-resident run), or null: dt
    const double *dtc;  // per-cell steps (local time stepping), or null
    FarField<double> fs;
};

// inviscid_fluxes (src/Solver.cpp:144-230) gathered per cell, the validity
// check (src/SolverData.cpp:66-79) and the stage update (src/Solver.cpp:83-133)
__global__ __launch_bounds__(PB) void k_poly_sweepB(PolyDev m, StageArgs a) {
    const uint32_t c = logical_block(blockIdx.x, gridDim.x, 1) * PB + threadIdx.x;  // XCD-contiguous
    if (c >= m.N || stopped(m.stop)) return;
    const size_t NT = m.NT, N = m.N;
    double R[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    // slot words first, then every face's loads without waiting on them;
    // absent slots and boundary faces read the cell's own entries (in
    // bounds) and are masked afterwards
    uint32_t w[6], qs[6];
    int nfc = 0;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        w[s] = m.w[s * N + c];
        qs[s] = m.ns[s * N + c];
    }
#pragma unroll
    for (int s = 0; s < 6; ++s)
        if (w[s] != NONE) nfc = s + 1;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const bool inner = s < nfc && !(w[s] & BND);
        const uint32_t n = inner ? (w[s] & IDX) : c, q = inner ? qs[s] : (uint32_t)s;
        const V3 S = {geo(m, s, 0, c), geo(m, s, 1, c), geo(m, s, 2, c)};
        double own[5], oth[5], UO[5], UN[5], F[5];
        for (int k = 0; k < 5; ++k) own[k] = a.FV ? a.FV[((size_t)s * 5 + k) * N + c] : a.V[k * NT + c];
        for (int k = 0; k < 5; ++k) oth[k] = a.FV ? a.FV[((size_t)q * 5 + k) * N + n] : a.V[k * NT + n];
        if (s >= nfc) break;
        if (w[s] & BND)  // the cell is the face's i side; ghost of its face value
            ghost_d((int)qs[s], own, S, a.fs, oth);
        const bool sj = (w[s] & SIDE_J) != 0;
        prim_to_cons(own, UO);
        prim_to_cons(oth, UN);
        if (a.flux == 1) {
            double ul[5], ur[5];  // the face's i and j sides
#pragma unroll
            for (int k = 0; k < 5; ++k) ul[k] = sj ? UN[k] : UO[k], ur[k] = sj ? UO[k] : UN[k];
            hllc_d(ul, ur, S, F);
        } else {
            rusanov_side(UO, UN, S, sj ? 0x80000000u : 0u, F);  // = rusanov(ul, ur, S)
        }
        if (sj) {
            for (int k = 0; k < 5; ++k) R[k] += F[k];
        } else {
            for (int k = 0; k < 5; ++k) R[k] -= F[k];
        }
    }
    bool ok = true;
    for (int k = 0; k < 5; ++k) {
        if (a.R) a.R[k * N + c] = R[k];
        ok = ok && std::isfinite(R[k]);
    }
    if (a.Rrho) a.Rrho[c] = R[0];
    if (!ok && a.bad) atomicAdd(a.bad, 1);
    if (a.kind < 0) return;
    const double vol = m.vol[c], dt = a.dtc ? a.dtc[c] : (a.dtp ? *a.dtp : a.dt);
    for (int k = 0; k < 5; ++k) {
        const size_t q = k * N + c;
        const double u = a.U[q];
        if (a.kind == 0) a.U[q] = u + dt / vol * R[k];
        else if (a.kind == 1) a.U[q] = 3.0 / 4.0 * a.Uo[q] + 1.0 / 4.0 * u + 1.0 / 4.0 * dt / vol * R[k];
        else a.U[q] = 1.0 / 3.0 * a.Uo[q] + 2.0 / 3.0 * u + 2.0 / 3.0 * dt / vol * R[k];
    }
}

// the stage update alone, after the host has checked the flux balance
__global__ __launch_bounds__(PB) void k_poly_update(PolyDev m, double *U, const double *Uo, const double *R,
                                                    int kind, double dt_all, const double *dtc) {
    const uint32_t c = blockIdx.x * PB + threadIdx.x;
    if (c >= m.N) return;
    const double vol = m.vol[c], dt = dtc ? dtc[c] : dt_all;
    for (int k = 0; k < 5; ++k) {
        const size_t q = k * (size_t)m.N + c;
        const double u = U[q];
        if (kind == 0) U[q] = u + dt / vol * R[q];
        else if (kind == 1) U[q] = 3.0 / 4.0 * Uo[q] + 1.0 / 4.0 * u + 1.0 / 4.0 * dt / vol * R[q];
        else U[q] = 1.0 / 3.0 * Uo[q] + 2.0 / 3.0 * u + 2.0 / 3.0 * dt / vol * R[q];
    }
}

// calc_timestep (src/Solver.cpp:232-274): per-block minimum of the cell steps
__global__ __launch_bounds__(PB) void k_poly_dt(PolyDev m, const double *V, double CFL, double *part,
                                                double *dt_cell) {
    __shared__ double sh[PB];
    if (stopped(m.stop)) return;  // the same for every thread of the block
    const uint32_t c = blockIdx.x * PB + threadIdx.x;
    double t = DBL_MAX;
    if (c < m.N) {
        double v[5];
        for (int k = 0; k < 5; ++k) v[k] = V[k * (size_t)m.NT + c];
        const double cs = sound_speed_prim(v);
        const double sx = (std::fabs(v[1]) + cs) * m.dS[c];
        const double sy = (std::fabs(v[2]) + cs) * m.dS[(size_t)m.N + c];
        const double sz = (std::fabs(v[3]) + cs) * m.dS[2 * (size_t)m.N + c];
        const double x = CFL * m.vol[c] / ((sx + sy) + sz);
        if (dt_cell) dt_cell[c] = x;  // local time stepping: each cell's own CFL step
        t = smin(t, x);
    }
    sh[threadIdx.x] = t;
    __syncthreads();
    for (int h = PB / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] = smin(sh[threadIdx.x], sh[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// per-block sums of x^2 (the residual norm), fixed order
__global__ __launch_bounds__(PB) void k_poly_sumsq(const double *x, uint32_t n, double *part, const int *stop) {
    __shared__ double sh[PB];
    if (stopped(stop)) return;  // the same for every thread of the block
    const uint32_t c = blockIdx.x * PB + threadIdx.x;
    sh[threadIdx.x] = c < n ? x[c] * x[c] : 0.0;
    __syncthreads();
    for (int h = PB / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Device-resident run, one block each.  flags: [0] stop, [1] converged,
// [2] invalid dt, [3] steps completed (the index of the step in progress), so
// a graph of several steps can be replayed unchanged.
// The step's dt: minimum of the block partials (calc_timestep).
__global__ __launch_bounds__(PB) void k_poly_step_dt(const double *part, uint32_t nb, int *flags, double *dtdev,
                                                     double *dts) {
    __shared__ double sh[PB];
    if (flags[0]) return;
    const int k = flags[3];  // this step's index: the steps completed so far
    double t = DBL_MAX;
    for (uint32_t q = threadIdx.x; q < nb; q += PB) t = smin(t, part[q]);
    sh[threadIdx.x] = t;
    __syncthreads();
    for (int h = PB / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] = smin(sh[threadIdx.x], sh[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double dt = sh[0];
        dts[k] = dt;
        if (!((dt - dt == 0.0) && dt > 0.0)) {
            flags[2] = 1;
            flags[0] = 1;
        } else {
            dtdev[0] = dt;
        }
    }
}

// The step's residual: sum of the block partials in a fixed order (each
// thread a strided sum, then a tree), its square root, and the drop test.
// runp = {drop, ref}: ref is the first residual since the state was set, or
// negative when this run's first step is that residual (the host loop's
// hist[0]; a first residual of 0 never converges, as there).  Read from
// device memory so a captured graph serves any drop and reference.
__global__ __launch_bounds__(PB) void k_poly_step_res(const double *part, uint32_t nb, int *flags, double *hres,
                                                      const double *runp) {
    __shared__ double sh[PB];
    if (flags[0]) return;
    const int k = flags[3];
    double v = 0.0;
    for (uint32_t q = threadIdx.x; q < nb; q += PB) v += part[q];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int h = PB / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double r = sqrt(sh[0]);
        hres[k] = r;
        flags[3] = k + 1;
        const double drop = runp[0], ref = runp[1];
        const double r0 = ref >= 0.0 ? ref : hres[0];
        if (drop > 0.0 && r0 > 0.0 && r / r0 <= drop) {
            flags[1] = 1;
            flags[0] = 1;
        }
    }
}

__global__ __launch_bounds__(PB) void k_poly_copy(double *dst, const double *src, size_t n, const int *stop) {
    const size_t q = (size_t)blockIdx.x * PB + threadIdx.x;
    if (q < n && !stopped(stop)) dst[q] = src[q];
}

// reference-order AoS [N][cols] <-> [cols][N] in device cell order
__global__ __launch_bounds__(PB) void k_poly_aos_in(const double *aos, const uint32_t *perm, double *soa, uint32_t N,
                                                    int cols) {
    const uint32_t d = blockIdx.x * PB + threadIdx.x;
    if (d >= N) return;
    const size_t r = (size_t)cols * perm[d];
    for (int k = 0; k < cols; ++k) soa[k * (size_t)N + d] = aos[r + k];
}
__global__ __launch_bounds__(PB) void k_poly_aos_out(const double *soa, const uint32_t *perm, double *aos, uint32_t N,
                                                     int cols) {
    const uint32_t d = blockIdx.x * PB + threadIdx.x;
    if (d >= N) return;
    const size_t r = (size_t)cols * perm[d];
    for (int k = 0; k < cols; ++k) aos[r + k] = soa[k * (size_t)N + d];
}
__global__ __launch_bounds__(PB) void k_poly_fill(double *p, double v, size_t n) {
    const size_t q = (size_t)blockIdx.x * PB + threadIdx.x;
    if (q < n) p[q] = v;
}
}  // namespace

struct PolyPath {
    hipStream_t st = nullptr;
    nf_params prm{};
    FarField<double> fs{};
    PolyDev m{};
    nf_fv_sizes sizes{};
    double *U = nullptr, *Uo = nullptr, *V = nullptr, *Ve = nullptr, *FV = nullptr;
    double *grad = nullptr, *psi = nullptr, *R = nullptr, *part = nullptr;
    double *Rrho = nullptr;
    double *dtc = nullptr;  // [N] per-cell steps (local time stepping)
    int *flags = nullptr;   // device-resident run flags
    double *dtdev = nullptr, *dts = nullptr, *hres = nullptr;  // its dt, and per step dt and residual
    int run_cap = 0;
    double *runp = nullptr;  // device {drop, ref} of the current run
    hipGraphExec_t graph = nullptr;  // check_every steps of nf_run, captured once per K and buffers
    int graph_K = 0;
    int *bad = nullptr;
    bool aux_valid = false;  // grad / psi / R hold nf_eval_flux_balance of the current state
    // run state since the last set_state / set_freestream
    // kernel timing (nf_time_kernels): HIP events around each launch
    bool timing = false;
    std::vector<std::pair<const char *, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::map<std::string, std::pair<double, int>> ktime;
    std::vector<uint32_t> perm;  // device cell -> reference cell
    uint32_t *perm_dev = nullptr;  // the same on the device (made on first state transfer)
    double *stage = nullptr;       // [15][N] AoS staging of state and field transfers
    std::vector<double> hist;    // residual norm of each step (flux balance of U^n)
    double time = 0.0;
    bool converged = false;
    std::vector<void *> owned;
    ~PolyPath() {
        if (st) (void)hipStreamSynchronize(st);
        if (graph) (void)hipGraphExecDestroy(graph);
        for (void *p : owned) (void)hipFree(p);
    }
};

namespace {
#define PHIP(expr)                                                                  \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess) {                                                     \
            err = std::string(#expr) + ": " + hipGetErrorString(e_);                \
            return NF_E_HIP;                                                        \
        }                                                                           \
    } while (0)

template <typename T> int alloc(PolyPath *P, T **p, size_t n, std::string &err) {
    PHIP(hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)));
    P->owned.push_back(*p);
    return NF_OK;
}
template <typename T> int upload(T *dst, const std::vector<T> &src, std::string &err) {
    PHIP(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return NF_OK;
}
dim3 grid(size_t n) { return dim3((unsigned)((n + PB - 1) / PB)); }

// AoS [N][cols] in reference cell order <-> SoA [cols][N] in device order
// Reference-order AoS [N][cols] crosses as is; the device reorders it into
// [cols][N] in device cell order (and back), the host never touches it.
int ensure_stage(PolyPath *P, std::string &err) {
    if (P->stage) return NF_OK;
    int rc;
    if ((rc = alloc(P, &P->perm_dev, P->m.N, err)) || (rc = alloc(P, &P->stage, 15 * (size_t)P->m.N, err))) return rc;
    PHIP(hipMemcpy(P->perm_dev, P->perm.data(), (size_t)P->m.N * sizeof(uint32_t), hipMemcpyHostToDevice));
    return NF_OK;
}
int put_soa(PolyPath *P, double *dst, const double *aos, int cols, std::string &err) {
    const uint32_t N = P->m.N;
    int rc;
    if ((rc = ensure_stage(P, err))) return rc;
    PHIP(hipMemcpyAsync(P->stage, aos, (size_t)cols * N * 8, hipMemcpyHostToDevice, P->st));
    hipLaunchKernelGGL(k_poly_aos_in, grid(N), dim3(PB), 0, P->st, (const double *)P->stage,
                       (const uint32_t *)P->perm_dev, dst, N, cols);
    PHIP(hipGetLastError());
    PHIP(hipStreamSynchronize(P->st));
    return NF_OK;
}
int get_soa(PolyPath *P, const double *src, int cols, double *aos, std::string &err) {
    const uint32_t N = P->m.N;
    int rc;
    if ((rc = ensure_stage(P, err))) return rc;
    hipLaunchKernelGGL(k_poly_aos_out, grid(N), dim3(PB), 0, P->st, src, (const uint32_t *)P->perm_dev, P->stage,
                       N, cols);
    PHIP(hipGetLastError());
    PHIP(hipMemcpyAsync(aos, P->stage, (size_t)cols * N * 8, hipMemcpyDeviceToHost, P->st));
    PHIP(hipStreamSynchronize(P->st));
    return NF_OK;
}

int tstart(PolyPath *P, hipEvent_t *e, std::string &err) {
    if (!P->timing) return NF_OK;
    PHIP(hipEventCreate(e));
    PHIP(hipEventRecord(*e, P->st));
    return NF_OK;
}
int tstop(PolyPath *P, const char *name, hipEvent_t e0, std::string &err) {
    if (!P->timing) return NF_OK;
    hipEvent_t e1;
    PHIP(hipEventCreate(&e1));
    PHIP(hipEventRecord(e1, P->st));
    P->pending.push_back({name, {e0, e1}});
    return NF_OK;
}
int tflush(PolyPath *P, std::string &err) {
    PHIP(hipStreamSynchronize(P->st));
    for (auto &q : P->pending) {
        float ms = 0.0f;
        PHIP(hipEventElapsedTime(&ms, q.second.first, q.second.second));
        auto &slot = P->ktime[q.first];
        slot.first += ms;
        slot.second += 1;
        (void)hipEventDestroy(q.second.first);
        (void)hipEventDestroy(q.second.second);
    }
    P->pending.clear();
    return NF_OK;
}

// evaluate_flux_balance of U into R (and the stage update when kind >= 0):
// primvars into V, ghosts, sweep A for second order, sweep B
// lts_dt: also store each cell's CFL step of these primvars (first stage of a
// locally time-stepped step); the update then uses the per-cell steps
// stop, dtp: device-resident run (kernels return once *stop is set; the dt is
// read from *dtp)
int flux_balance(PolyPath *P, double *V, int kind, double dt, bool aux, bool res, bool lts_dt, std::string &err,
                 const int *stop = nullptr, const double *dtp = nullptr) {
    PolyDev m = P->m;
    m.stop = stop;
    const bool check = P->prm.check_physical_validity != 0;
    int rc;
    hipEvent_t e = nullptr;
    if ((rc = tstart(P, &e, err))) return rc;
    hipLaunchKernelGGL(k_poly_prim, grid(m.N), dim3(PB), 0, P->st, (const double *)P->U, V, m.N, m.NT, stop);
    if ((rc = tstop(P, "poly_prim", e, err)) || (rc = tstart(P, &e, err))) return rc;
    if (m.G) hipLaunchKernelGGL(k_poly_ghost, grid(m.G), dim3(PB), 0, P->st, m, V, P->fs);
    if ((rc = tstop(P, "poly_ghost", e, err))) return rc;
    const bool lts = P->prm.local_time_step != 0 && kind >= 0;
    if (lts_dt)
        hipLaunchKernelGGL(k_poly_dt, grid(m.N), dim3(PB), 0, P->st, m, (const double *)V, P->prm.CFL, P->part,
                           P->dtc);
    const bool second = P->prm.spatial_order == NF_SECOND_ORDER;
    if (second) {
        if ((rc = tstart(P, &e, err))) return rc;
        hipLaunchKernelGGL(k_poly_sweepA, grid(m.N), dim3(PB), 0, P->st, m, (const double *)V,
                           P->prm.limiter != NF_LIMITER_NONE ? 1 : 0, P->FV, aux ? P->grad : nullptr,
                           aux ? P->psi : nullptr);
        if ((rc = tstop(P, "poly_sweepA", e, err))) return rc;
    }
    StageArgs a;
    a.V = V;
    a.FV = second ? P->FV : nullptr;
    a.U = P->U;
    a.Uo = P->Uo;
    a.R = (aux || check) ? P->R : nullptr;
    a.Rrho = res ? P->Rrho : nullptr;
    a.bad = check ? P->bad : nullptr;
    a.flux = P->prm.inv_flux_scheme == NF_FLUX_HLLC ? 1 : 0;
    a.kind = check ? -1 : kind;
    a.dt = dt;
    a.dtp = dtp;
    a.dtc = lts ? P->dtc : nullptr;
    a.fs = P->fs;
    if (check) PHIP(hipMemsetAsync(P->bad, 0, sizeof(int), P->st));
    if ((rc = tstart(P, &e, err))) return rc;
    hipLaunchKernelGGL(k_poly_sweepB, grid(m.N), dim3(PB), 0, P->st, m, a);
    PHIP(hipGetLastError());
    if ((rc = tstop(P, "poly_sweepB", e, err))) return rc;
    if (check) {
        int nbad = 0;
        PHIP(hipMemcpyAsync(&nbad, P->bad, sizeof(int), hipMemcpyDeviceToHost, P->st));
        PHIP(hipStreamSynchronize(P->st));
        if (nbad > 0) {
            err = std::to_string(nbad) + " cells with unphysical values detected in the flux_balance field of the "
                                         "Euler solver";
            return NF_E_NONPHYSICAL;
        }
        if (kind >= 0) {
            hipLaunchKernelGGL(k_poly_update, grid(m.N), dim3(PB), 0, P->st, m, P->U, (const double *)P->Uo,
                               (const double *)P->R, kind, dt, (const double *)a.dtc);
            PHIP(hipGetLastError());
        }
    }
    return NF_OK;
}
}  // namespace

int poly_create(const PrimalMesh &pm, const FVGrid &g, const nf_params &p, hipStream_t st, PolyPath **out,
                std::string &err) {
    (void)pm;
    *out = nullptr;
    if (p.precision != NF_PREC_FP64) {
        err = "a mixed-element mesh runs in FP64 only";
        return NF_E_ARG;
    }
    std::unique_ptr<PolyPath> hold(new PolyPath);
    PolyPath *P = hold.get();
    P->st = st;
    P->prm = p;
    P->fs = make_farfield<double>(p.primvars_inf);
    const uint32_t N = g.N, NT = g.NT, Fint = g.Fint, F = g.F;
    P->sizes = {N, NT, Fint, F, (uint32_t)g.pfirst.size()};
    if (NT - 1 > IDX) {
        err = "mesh too large for the mixed-element path";
        return NF_E_MESH;
    }
    // device numbering (cell_order); each cell keeps its faces in increasing
    // reference face index, so every sum keeps the reference's order
    if (p.cell_order == NF_ORDER_RCM) P->perm = rcm_order(g);
    else if (p.cell_order == NF_ORDER_MORTON) P->perm = morton_order(g);
    else {
        P->perm.resize(N);
        for (uint32_t c = 0; c < N; ++c) P->perm[c] = c;
    }
    std::vector<uint32_t> iperm(N);
    for (uint32_t d = 0; d < N; ++d) iperm[P->perm[d]] = d;
    // slots: the faces of each cell in increasing face index
    std::vector<uint32_t> w((size_t)6 * N, NONE);
    std::vector<uint8_t> ns((size_t)6 * N, 0), cnt(N, 0);
    std::vector<double> geo((size_t)36 * N, 0.0);
    const uint32_t G = F - Fint;
    std::vector<uint32_t> bi(G), bj(G);
    std::vector<double> bS((size_t)3 * G);
    std::vector<uint8_t> bbc(G);
    for (size_t q = 0; q < g.pfirst.size(); ++q)
        for (uint32_t f = g.pfirst[q]; f < g.pfirst[q] + g.pn[q]; ++f) bbc[f - Fint] = (uint8_t)g.pbc[q];
    auto put_geo = [&](uint32_t c, int s, uint32_t f, const std::vector<double> &r) {
        for (int d = 0; d < 3; ++d) {
            geo[((size_t)s * 6 + d) * N + c] = g.S[3 * (size_t)f + d];
            geo[((size_t)s * 6 + 3 + d) * N + c] = r[3 * (size_t)f + d];
        }
    };
    for (uint32_t f = 0; f < F; ++f) {
        const uint32_t i = iperm[g.ij[2 * (size_t)f]], jr = g.ij[2 * (size_t)f + 1];
        const uint32_t j = jr < N ? iperm[jr] : jr;  // ghosts keep their reference id
        if (cnt[i] >= 6 || (j < N && cnt[j] >= 6)) {
            err = "cell with more than six faces";
            return NF_E_MESH;
        }
        const int si = cnt[i]++;
        if (j < N) {
            const int sj = cnt[j]++;
            w[(size_t)si * N + i] = j;
            ns[(size_t)si * N + i] = (uint8_t)sj;
            w[(size_t)sj * N + j] = i | SIDE_J;
            ns[(size_t)sj * N + j] = (uint8_t)si;
            put_geo(j, sj, f, g.rj);
        } else {
            const uint32_t b = f - Fint;
            w[(size_t)si * N + i] = j | BND;
            ns[(size_t)si * N + i] = bbc[b];
            bi[b] = i;
            bj[b] = j;
            for (int d = 0; d < 3; ++d) bS[(size_t)d * G + b] = g.S[3 * (size_t)f + d];
        }
        put_geo(i, si, f, g.ri);
    }
    std::vector<double> vol(N), dS((size_t)3 * N);
    for (uint32_t c = 0; c < N; ++c) {
        vol[c] = g.vol[P->perm[c]];
        for (int d = 0; d < 3; ++d) dS[d * (size_t)N + c] = g.dS[3 * (size_t)P->perm[c] + d];
    }
    uint32_t *dw, *dbi, *dbj;
    uint8_t *dns, *dbbc;
    double *dgeo, *dvol, *ddS, *dbS;
    int rc;
    if ((rc = alloc(P, &dw, w.size(), err)) || (rc = alloc(P, &dns, ns.size(), err)) ||
        (rc = alloc(P, &dgeo, geo.size(), err)) || (rc = alloc(P, &dvol, N, err)) ||
        (rc = alloc(P, &ddS, 3 * (size_t)N, err)) || (rc = alloc(P, &dbi, G, err)) || (rc = alloc(P, &dbj, G, err)) ||
        (rc = alloc(P, &dbS, 3 * (size_t)G, err)) || (rc = alloc(P, &dbbc, G, err)))
        return rc;
    if ((rc = upload(dw, w, err)) || (rc = upload(dns, ns, err)) || (rc = upload(dgeo, geo, err)) ||
        (rc = upload(dvol, vol, err)) || (rc = upload(ddS, dS, err)) || (rc = upload(dbi, bi, err)) ||
        (rc = upload(dbj, bj, err)) || (rc = upload(dbS, bS, err)) || (rc = upload(dbbc, bbc, err)))
        return rc;
    P->m = {N, NT, G, dw, dns, dgeo, dvol, ddS, dbi, dbj, dbS, dbbc};
    if ((rc = alloc(P, &P->U, 5 * (size_t)N, err)) || (rc = alloc(P, &P->Uo, 5 * (size_t)N, err)) ||
        (rc = alloc(P, &P->V, 5 * (size_t)NT, err)) || (rc = alloc(P, &P->Ve, 5 * (size_t)NT, err)) ||
        (rc = alloc(P, &P->FV, 30 * (size_t)N, err)) || (rc = alloc(P, &P->grad, 15 * (size_t)N, err)) ||
        (rc = alloc(P, &P->psi, 5 * (size_t)N, err)) || (rc = alloc(P, &P->R, 5 * (size_t)N, err)) ||
        (rc = alloc(P, &P->part, (N + PB - 1) / PB, err)) || (rc = alloc(P, &P->bad, 1, err)) ||
        (rc = alloc(P, &P->Rrho, N, err)) || (rc = alloc(P, &P->dtc, N, err)))
        return rc;
    if ((rc = poly_set_freestream(P, err))) return rc;
    *out = hold.release();
    return NF_OK;
}

void poly_destroy(PolyPath *P) { delete P; }

nf_fv_sizes poly_sizes(const PolyPath *P) { return P->sizes; }

// set_freestream_values (src/SolverData.cpp:50-58)
static void reset_run(PolyPath *P) {
    P->hist.clear();
    P->time = 0.0;
    P->converged = false;
    P->aux_valid = false;
}

int poly_set_freestream(PolyPath *P, std::string &err) {
    const uint32_t N = P->m.N;
    reset_run(P);
    double U[5];
    prim_to_cons(P->prm.primvars_inf, U);
    for (int k = 0; k < 5; ++k) {
        hipLaunchKernelGGL(k_poly_fill, grid(N), dim3(PB), 0, P->st, P->U + k * (size_t)N, U[k], (size_t)N);
        hipLaunchKernelGGL(k_poly_fill, grid(N), dim3(PB), 0, P->st, P->V + k * (size_t)P->m.NT,
                           P->prm.primvars_inf[k], (size_t)N);
    }
    if (P->prm.rk3_zero_old_first_step)
        hipLaunchKernelGGL(k_poly_fill, grid(5 * (size_t)N), dim3(PB), 0, P->st, P->Uo, 0.0, 5 * (size_t)N);
    else
        PHIP(hipMemcpyAsync(P->Uo, P->U, 5 * (size_t)N * 8, hipMemcpyDeviceToDevice, P->st));
    PHIP(hipGetLastError());
    PHIP(hipStreamSynchronize(P->st));
    return NF_OK;
}

// U as given, primvars = prim(U); U_old = zero (the reference's first RK3
// step) unless given, and a given U_old must equal U
int poly_set_state(PolyPath *P, const double *U, const double *U_old, std::string &err) {
    const uint32_t N = P->m.N;
    bool zero = P->prm.rk3_zero_old_first_step != 0;
    if (U_old) {
        if (std::memcmp(U, U_old, (size_t)5 * N * sizeof(double)) != 0) {
            err = "U_old must be NULL (zero) or equal to U";
            return NF_E_ARG;
        }
        zero = false;
    }
    int rc;
    reset_run(P);
    if ((rc = put_soa(P, P->U, U, 5, err))) return rc;
    hipLaunchKernelGGL(k_poly_prim, grid(N), dim3(PB), 0, P->st, (const double *)P->U, P->V, N, P->m.NT,
                       (const int *)nullptr);
    if (zero)
        hipLaunchKernelGGL(k_poly_fill, grid(5 * (size_t)N), dim3(PB), 0, P->st, P->Uo, 0.0, 5 * (size_t)N);
    else
        PHIP(hipMemcpyAsync(P->Uo, P->U, 5 * (size_t)N * 8, hipMemcpyDeviceToDevice, P->st));
    PHIP(hipGetLastError());
    PHIP(hipStreamSynchronize(P->st));
    return NF_OK;
}

int poly_calc_timestep(PolyPath *P, double *dt_out, std::string &err) {
    const uint32_t nb = (P->m.N + PB - 1) / PB;
    hipLaunchKernelGGL(k_poly_dt, dim3(nb), dim3(PB), 0, P->st, P->m, (const double *)P->V, P->prm.CFL, P->part,
                       (double *)nullptr);
    PHIP(hipGetLastError());
    std::vector<double> part(nb);
    PHIP(hipMemcpyAsync(part.data(), P->part, nb * 8, hipMemcpyDeviceToHost, P->st));
    PHIP(hipStreamSynchronize(P->st));
    double dt = DBL_MAX;
    for (double x : part) dt = smin(dt, x);
    if (dt_out) *dt_out = dt;
    if (!((dt - dt == 0.0) && dt > 0.0)) {
        char buf[96];
        std::snprintf(buf, sizeof buf, "Invalid dt calculated (dt = %f)", dt);
        err = buf;
        return NF_E_INVALID_DT;
    }
    return NF_OK;
}

// Solver::step (src/Solver.cpp:34-52, 83-133)
int poly_step(PolyPath *P, double dt, std::string &err) {
    int rc;
    P->aux_valid = false;
    const bool rk3 = P->prm.time_scheme == NF_TVD_RK3, lts = P->prm.local_time_step != 0;
    // the residual is the flux balance of U^n (first stage), as in the batched path
    if ((rc = flux_balance(P, P->V, 0, dt, false, true, lts, err))) return rc;
    if (rk3) {
        if ((rc = flux_balance(P, P->V, 1, dt, false, false, false, err))) return rc;
        if ((rc = flux_balance(P, P->V, 2, dt, false, false, false, err))) return rc;
    }
    PHIP(hipMemcpyAsync(P->Uo, P->U, 5 * (size_t)P->m.N * 8, hipMemcpyDeviceToDevice, P->st));
    // residual norm ||R_rho||_2 of the first stage
    const uint32_t nb = (P->m.N + PB - 1) / PB;
    hipLaunchKernelGGL(k_poly_sumsq, dim3(nb), dim3(PB), 0, P->st, (const double *)P->Rrho, P->m.N, P->part,
                       (const int *)nullptr);
    PHIP(hipGetLastError());
    std::vector<double> part(nb);
    PHIP(hipMemcpyAsync(part.data(), P->part, nb * 8, hipMemcpyDeviceToHost, P->st));
    PHIP(hipStreamSynchronize(P->st));
    // the summation order of k_poly_step_res, so both loops give the same bits
    std::vector<double> acc(PB, 0.0);
    for (uint32_t q = 0; q < nb; ++q) acc[q % PB] += part[q];
    for (int h = PB / 2; h > 0; h >>= 1)
        for (int t = 0; t < h; ++t) acc[t] += acc[t + h];
    P->hist.push_back(std::sqrt(acc[0]));
    P->time += dt;
    return NF_OK;
}

// the flux balance of the current state, once, with its own primvar buffer:
// the primvars a following calc_timestep reads are left as the last step set them
int poly_eval_flux_balance(PolyPath *P, std::string &err) {
    int rc;
    if ((rc = flux_balance(P, P->Ve, -1, 0.0, true, false, false, err))) return rc;
    PHIP(hipStreamSynchronize(P->st));
    P->aux_valid = true;
    return NF_OK;
}

int poly_get_state(PolyPath *P, double *U, std::string &err) { return get_soa(P, P->U, 5, U, err); }

// The Driver loop (calc_timestep, step) for n steps, stopping after the step
// whose residual norm is at most drop times the first step's; a run that has
// reached its drop stays stopped until the state is set again.
static int poly_run_host(PolyPath *P, const nf_run_opts *o, nf_run_stats *stats, std::string &err) {
    const double drop = o->residual_drop;
    hipEvent_t e0, e1;
    PHIP(hipEventCreate(&e0));
    PHIP(hipEventCreate(&e1));
    PHIP(hipEventRecord(e0, P->st));
    int done = 0, rc = NF_OK;
    double dt = 0.0;
    if (!(drop > 0.0 && P->converged)) {
        while (done < o->n_steps) {
            if ((rc = poly_calc_timestep(P, &dt, err)) || (rc = poly_step(P, dt, err))) break;
            ++done;
            if (drop > 0.0 && P->hist[0] > 0.0 && P->hist.back() / P->hist[0] <= drop) {
                P->converged = true;
                break;
            }
        }
    }
    float ms = 0.0f;
    (void)hipEventRecord(e1, P->st);
    (void)hipEventSynchronize(e1);
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (stats) {
        stats->steps_done = done;
        stats->converged = P->converged ? 1 : 0;
        stats->time = P->time;
        stats->last_dt = dt;
        stats->res_first = P->hist.empty() ? 0.0 : P->hist.front();
        stats->res_last = P->hist.empty() ? 0.0 : P->hist.back();
        stats->elapsed_ms = ms;
        stats->fell_back = 0;
        stats->pad = 0;
    }
    return rc;
}

// HIP-event durations of each kernel over n steps of the Driver loop
int poly_time_kernels(PolyPath *P, int32_t n_steps, nf_kernel_time *out, int32_t max, int32_t *n_out,
                      std::string &err) {
    if (n_steps < 1 || !out) return NF_E_ARG;
    P->ktime.clear();
    P->timing = true;
    int rc = NF_OK;
    for (int q = 0; q < n_steps && rc == NF_OK; ++q) {
        double dt;
        if ((rc = poly_calc_timestep(P, &dt, err)) == NF_OK) rc = poly_step(P, dt, err);
        const int f = tflush(P, err);
        if (rc == NF_OK) rc = f;
    }
    P->timing = false;
    if (rc) return rc;
    int n = 0;
    for (auto &kv : P->ktime) {
        if (n >= max) break;
        std::snprintf(out[n].name, sizeof out[n].name, "%s", kv.first.c_str());
        out[n].avg_ms = kv.second.first / kv.second.second;
        out[n].calls = kv.second.second;
        ++n;
    }
    if (n_out) *n_out = n;
    return NF_OK;
}

// One step of the device-resident run: the same kernels and values as
// calc_timestep + step, with dt, residual and the stop test on the device.
static int enqueue_run_step(PolyPath *P, std::string &err) {
    PolyDev m = P->m;
    m.stop = P->flags;
    const uint32_t nb = (m.N + PB - 1) / PB;
    hipLaunchKernelGGL(k_poly_dt, dim3(nb), dim3(PB), 0, P->st, m, (const double *)P->V, P->prm.CFL, P->part,
                       (double *)nullptr);
    hipLaunchKernelGGL(k_poly_step_dt, dim3(1), dim3(PB), 0, P->st, (const double *)P->part, nb, P->flags,
                       P->dtdev, P->dts);
    PHIP(hipGetLastError());
    const bool rk3 = P->prm.time_scheme == NF_TVD_RK3, lts = P->prm.local_time_step != 0;
    int rc;
    if ((rc = flux_balance(P, P->V, 0, 0.0, false, true, lts, err, P->flags, P->dtdev))) return rc;
    if (rk3) {
        if ((rc = flux_balance(P, P->V, 1, 0.0, false, false, false, err, P->flags, P->dtdev))) return rc;
        if ((rc = flux_balance(P, P->V, 2, 0.0, false, false, false, err, P->flags, P->dtdev))) return rc;
    }
    hipLaunchKernelGGL(k_poly_copy, grid(5 * (size_t)m.N), dim3(PB), 0, P->st, P->Uo, (const double *)P->U,
                       5 * (size_t)m.N, (const int *)P->flags);
    hipLaunchKernelGGL(k_poly_sumsq, dim3(nb), dim3(PB), 0, P->st, (const double *)P->Rrho, m.N, P->part,
                       (const int *)P->flags);
    hipLaunchKernelGGL(k_poly_step_res, dim3(1), dim3(PB), 0, P->st, (const double *)P->part, nb, P->flags,
                       P->hres, (const double *)P->runp);
    PHIP(hipGetLastError());
    return NF_OK;
}

// nf_run: check_every steps, captured once as a HIP graph (use_graph) or
// launched eagerly, run between host synchronisations; the stop flag ends the
// remaining ones on the device, so the run stops after exactly the step that
// reached the drop, as the host loop does.  With the flux-balance check on (a
// host read per stage) the host loop runs instead.
int poly_run(PolyPath *P, const nf_run_opts *o, nf_run_stats *stats, std::string &err) {
    if (!o || o->n_steps < 0) return NF_E_ARG;
    if (P->prm.check_physical_validity) return poly_run_host(P, o, stats, err);
    const double drop = o->residual_drop;
    int rc;
    P->aux_valid = false;
    if (!P->flags && ((rc = alloc(P, &P->flags, 4, err)) || (rc = alloc(P, &P->dtdev, 1, err)) ||
                      (rc = alloc(P, &P->runp, 2, err))))
        return rc;
    if (o->n_steps > P->run_cap) {  // per-step history buffers, grown (the graph refers to them)
        const int cap = std::max(o->n_steps, 1024);
        PHIP(hipStreamSynchronize(P->st));
        if (P->graph) {
            (void)hipGraphExecDestroy(P->graph);
            P->graph = nullptr;
        }
        for (double *old : {P->dts, P->hres}) {
            if (!old) continue;
            P->owned.erase(std::find(P->owned.begin(), P->owned.end(), (void *)old));
            PHIP(hipFree(old));
        }
        P->dts = P->hres = nullptr;
        if ((rc = alloc(P, &P->dts, cap, err)) || (rc = alloc(P, &P->hres, cap, err))) return rc;
        P->run_cap = cap;
    }
    PHIP(hipMemsetAsync(P->flags, 0, 4 * sizeof(int), P->st));
    // this run's drop and reference (negative: the run's own first residual)
    const double runp[2] = {drop, P->hist.empty() ? -1.0 : P->hist.front()};
    PHIP(hipMemcpyAsync(P->runp, runp, sizeof runp, hipMemcpyHostToDevice, P->st));
    hipEvent_t e0, e1;
    PHIP(hipEventCreate(&e0));
    PHIP(hipEventCreate(&e1));
    PHIP(hipEventRecord(e0, P->st));
    const int K = std::max(1, o->check_every);
    int flags[4] = {0, 0, 0, 0}, queued = 0;
    if (!(drop > 0.0 && P->converged)) {
        while (queued < o->n_steps && !flags[0]) {
            const int chunk = std::min(K, o->n_steps - queued);
            if (o->use_graph && chunk == K && K > 1) {
                if (!P->graph || P->graph_K != K) {
                    if (P->graph) (void)hipGraphExecDestroy(P->graph);
                    P->graph = nullptr;
                    hipGraph_t gr;
                    PHIP(hipStreamBeginCapture(P->st, hipStreamCaptureModeThreadLocal));
                    for (int q = 0; q < K; ++q)
                        if ((rc = enqueue_run_step(P, err))) {
                            (void)hipStreamEndCapture(P->st, &gr);
                            return rc;
                        }
                    PHIP(hipStreamEndCapture(P->st, &gr));
                    const hipError_t ie = hipGraphInstantiate(&P->graph, gr, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(gr);
                    PHIP(ie);
                    P->graph_K = K;
                }
                PHIP(hipGraphLaunch(P->graph, P->st));
            } else {
                for (int q = 0; q < chunk; ++q)
                    if ((rc = enqueue_run_step(P, err))) return rc;
            }
            queued += chunk;
            PHIP(hipMemcpyAsync(flags, P->flags, sizeof flags, hipMemcpyDeviceToHost, P->st));
            PHIP(hipStreamSynchronize(P->st));
        }
    }
    float ms = 0.0f;
    (void)hipEventRecord(e1, P->st);
    (void)hipEventSynchronize(e1);
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const int done = flags[3];
    std::vector<double> dts(done + 1), res(done);
    if (done + flags[2] > 0) PHIP(hipMemcpy(dts.data(), P->dts, (done + flags[2]) * 8, hipMemcpyDeviceToHost));
    if (done > 0) PHIP(hipMemcpy(res.data(), P->hres, done * 8, hipMemcpyDeviceToHost));
    for (int k = 0; k < done; ++k) {
        P->time += dts[k];
        P->hist.push_back(res[k]);
    }
    if (flags[1]) P->converged = true;
    if (stats) {
        stats->steps_done = done;
        stats->converged = P->converged ? 1 : 0;
        stats->time = P->time;
        stats->last_dt = done > 0 ? dts[done - 1] : 0.0;
        stats->res_first = P->hist.empty() ? 0.0 : P->hist.front();
        stats->res_last = P->hist.empty() ? 0.0 : P->hist.back();
        stats->elapsed_ms = ms;
        stats->fell_back = 0;
        stats->pad = 0;
    }
    if (flags[2]) {
        char buf[96];
        std::snprintf(buf, sizeof buf, "Invalid dt calculated (dt = %f)", dts[done]);
        err = buf;
        return NF_E_INVALID_DT;
    }
    return NF_OK;
}

int poly_get_residuals(PolyPath *P, double *out, int32_t max, int32_t *n_out) {
    const int32_t n = (int32_t)std::min<size_t>(P->hist.size(), (size_t)std::max(max, 0));
    std::copy(P->hist.begin(), P->hist.begin() + n, out);
    if (n_out) *n_out = n;
    return NF_OK;
}

int poly_get_field(PolyPath *P, int32_t which, double *out, std::string &err) {
    switch (which) {
    case NF_FIELD_PRIM:  // prim(U) into Ve with row stride N for the export
        hipLaunchKernelGGL(k_poly_prim, grid(P->m.N), dim3(PB), 0, P->st, (const double *)P->U, P->Ve, P->m.N,
                           P->m.N, (const int *)nullptr);
        PHIP(hipGetLastError());
        return get_soa(P, P->Ve, 5, out, err);
    case NF_FIELD_STATE_OLD:
        return get_soa(P, P->Uo, 5, out, err);
    case NF_FIELD_LIMITER:
    case NF_FIELD_GRADIENT:
    case NF_FIELD_RESIDUAL:
        if (!P->aux_valid) {
            err = "call nf_eval_flux_balance first";
            return NF_E_ARG;
        }
        if (which == NF_FIELD_LIMITER) {
            if (P->prm.spatial_order != NF_SECOND_ORDER) {  // psi keeps its initial 1 (SolverData.cpp:17)
                for (size_t q = 0; q < 5 * (size_t)P->m.N; ++q) out[q] = 1.0;
                return NF_OK;
            }
            return get_soa(P, P->psi, 5, out, err);
        }
        if (which == NF_FIELD_GRADIENT) {
            if (P->prm.spatial_order != NF_SECOND_ORDER) {
                std::memset(out, 0, 15 * (size_t)P->m.N * 8);
                return NF_OK;
            }
            return get_soa(P, P->grad, 15, out, err);
        }
        return get_soa(P, P->R, 5, out, err);
    default:
        err = "unknown field";
        return NF_E_ARG;
    }
}

}  // namespace nf
