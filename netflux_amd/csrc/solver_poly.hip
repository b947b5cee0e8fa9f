// Mixed-element FP64 step on the GPU (poly.hpp).  Layouts, all in the
// reference's cell and face numbering:
//   U, U_old, R      [5][N]     conservative state, old state, flux balance
//   V                [5][NT]    primvars of the cells, then of the ghosts
//   grad             [15][N]    Green-Gauss gradient, index k + 5 d
//   psi              [5][N]     Barth limiter
//   VL / VR          [5][F] / [5][Fint]  face values reconstructed from
//                               the cell i / the cell j side of each face
//   cf               [6][N]     faces of each cell in increasing face index,
//                               bit 31 = the cell is the face's j side
// Kernels per stage: prim (U -> V), cell ghosts, sweep A (gradient, min/max,
// limiter, face values; second order only), sweep B (flux balance and the
// stage update).  Each follows the CPU solver's loops (src/Solver.cpp:54-230,
// include/Numerics.hpp:86-292) with the same operations in the same order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "physics.hpp"
#include "poly.hpp"

namespace nf {

namespace {
constexpr int PB = 256;
constexpr uint32_t NONE = 0xFFFFFFFFu, SIDE_J = 0x80000000u;
using V3 = Vec3T<double>;

struct PolyDev {
    uint32_t N, NT, Fint, F;
    const uint32_t *cf, *fi, *fj;
    const double *S, *ri, *rj, *vol, *dS;
    const uint8_t *bbc;
};

__device__ __forceinline__ V3 ld3(const double *a, uint32_t stride, uint32_t i) {
    return {a[i], a[stride + i], a[2 * (size_t)stride + i]};
}

__global__ __launch_bounds__(PB) void k_poly_prim(const double *U, double *V, uint32_t N, uint32_t NT) {
    const uint32_t c = blockIdx.x * PB + threadIdx.x;
    if (c >= N) return;
    double u[5], v[5];
    for (int k = 0; k < 5; ++k) u[k] = U[k * (size_t)N + c];
    cons_to_prim(u, v);
    for (int k = 0; k < 5; ++k) V[k * (size_t)NT + c] = v[k];
}

// set_constant_ghost_values (src/Solver.cpp:300-330)
__global__ __launch_bounds__(PB) void k_poly_ghost(PolyDev m, double *V, FarField<double> fs) {
    const uint32_t b = blockIdx.x * PB + threadIdx.x;
    if (b >= m.F - m.Fint) return;
    const uint32_t f = m.Fint + b, i = m.fi[f], j = m.fj[f];
    double vd[5], vg[5];
    for (int k = 0; k < 5; ++k) vd[k] = V[k * (size_t)m.NT + i];
    ghost_value<double>((int)m.bbc[b], vd, ld3(m.S, m.F, f), fs, vg);
    for (int k = 0; k < 5; ++k) V[k * (size_t)m.NT + j] = vg[k];
}

// Green-Gauss (Numerics.hpp:86-128), max/min (:253-292), Barth (:148-250)
// and the reconstructed face values (Solver.hpp:81-112) of one cell
__global__ __launch_bounds__(PB) void k_poly_sweepA(PolyDev m, const double *V, int limiter, double *grad,
                                                    double *psi, double *VL, double *VR) {
    const uint32_t c = blockIdx.x * PB + threadIdx.x;
    if (c >= m.N) return;
    const size_t NT = m.NT;
    double vc[5];
    for (int k = 0; k < 5; ++k) vc[k] = V[k * NT + c];
    const double vol = m.vol[c];
    uint32_t w[6];
    int nfc = 0;
    for (int s = 0; s < 6; ++s) {
        w[s] = m.cf[s * (size_t)m.N + c];
        if (w[s] != NONE) nfc = s + 1;
    }
    double g[15];
    for (int q = 0; q < 15; ++q) g[q] = 0.0;
    double vmax[5], vmin[5];
    for (int k = 0; k < 5; ++k) vmax[k] = -DBL_MAX, vmin[k] = DBL_MAX;
    for (int s = 0; s < nfc; ++s) {
        const uint32_t f = w[s] & ~SIDE_J;
        const bool sj = (w[s] & SIDE_J) != 0;
        const uint32_t i = m.fi[f], j = m.fj[f];
        const V3 S = ld3(m.S, m.F, f);
        for (int k = 0; k < 5; ++k) {
            const double vi = V[k * NT + i], vj = V[k * NT + j];
            const double uf = 0.5 * (vi + vj);
            const double t0 = uf * S.x, t1 = uf * S.y, t2 = uf * S.z;
            if (!sj) {
                g[k] += t0 / vol;
                g[k + 5] += t1 / vol;
                g[k + 10] += t2 / vol;
            } else {
                g[k] -= t0 / vol;
                g[k + 5] -= t1 / vol;
                g[k + 10] -= t2 / vol;
            }
            vmax[k] = smax(vmax[k], smax(vi, vj));
            vmin[k] = smin(vmin[k], smin(vi, vj));
        }
    }
    double ps[5];
    if (limiter) {
        for (int k = 0; k < 5; ++k) ps[k] = DBL_MAX;
        for (int s = 0; s < nfc; ++s) {
            const uint32_t f = w[s] & ~SIDE_J;
            const V3 r = ld3((w[s] & SIDE_J) ? m.rj : m.ri, m.F, f);
            for (int k = 0; k < 5; ++k) {
                double D = (g[k] * r.x + g[k + 5] * r.y) + g[k + 10] * r.z;
                D = (double)isign(D) * (std::fabs(D) + DBL_EPSILON);
                double fcon;
                if (D > 0.0) fcon = smin(1.0, (vmax[k] - vc[k]) / D);
                else if (D < 0.0) fcon = smin(1.0, (vmin[k] - vc[k]) / D);
                else fcon = 1.0;
                ps[k] = smin(ps[k], fcon);
            }
        }
    } else {
        for (int k = 0; k < 5; ++k) ps[k] = 1.0;  // SolverData.cpp:17
    }
    for (int s = 0; s < nfc; ++s) {
        const uint32_t f = w[s] & ~SIDE_J;
        const bool sj = (w[s] & SIDE_J) != 0;
        const V3 r = ld3(sj ? m.rj : m.ri, m.F, f);
        for (int k = 0; k < 5; ++k) {
            const double D = (g[k] * r.x + g[k + 5] * r.y) + g[k + 10] * r.z;
            const double vf = vc[k] + ps[k] * D;
            if (sj) VR[k * (size_t)m.Fint + f] = vf;
            else VL[k * (size_t)m.F + f] = vf;
        }
    }
    for (int q = 0; q < 15; ++q) grad[q * (size_t)m.N + c] = g[q];
    for (int k = 0; k < 5; ++k) psi[k * (size_t)m.N + c] = ps[k];
}

struct StageArgs {
    const double *V, *VL, *VR;  // VL == nullptr: first order (face values = cell values)
    double *U;
    const double *Uo;
    double *R;
    int *bad;
    int flux;    // 0 Rusanov, 1 HLLC
    int kind;    // 0: U + dt/vol R, 1: RK3 stage 2, 2: RK3 stage 3, -1: no update
    double dt;
    FarField<double> fs;
};

// inviscid_fluxes (src/Solver.cpp:144-230) gathered per cell, the validity
// check (src/SolverData.cpp:66-79) and the stage update (src/Solver.cpp:83-133)
__global__ __launch_bounds__(PB) void k_poly_sweepB(PolyDev m, StageArgs a) {
    const uint32_t c = blockIdx.x * PB + threadIdx.x;
    if (c >= m.N) return;
    const size_t NT = m.NT;
    double R[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int s = 0; s < 6; ++s) {
        const uint32_t w = m.cf[s * (size_t)m.N + c];
        if (w == NONE) break;
        const uint32_t f = w & ~SIDE_J;
        const V3 S = ld3(m.S, m.F, f);
        double vl[5], vr[5], ul[5], ur[5], F[5];
        const uint32_t i = m.fi[f], j = m.fj[f];
        for (int k = 0; k < 5; ++k) vl[k] = a.VL ? a.VL[k * (size_t)m.F + f] : a.V[k * NT + i];
        if (f < m.Fint) {
            for (int k = 0; k < 5; ++k) vr[k] = a.VL ? a.VR[k * (size_t)m.Fint + f] : a.V[k * NT + j];
        } else {
            ghost_value<double>((int)m.bbc[f - m.Fint], vl, S, a.fs, vr);
        }
        prim_to_cons(vl, ul);
        prim_to_cons(vr, ur);
        if (a.flux == 1) hllc_d(ul, ur, S, F);
        else rusanov<double>(ul, ur, S, F);
        if (w & SIDE_J) {
            for (int k = 0; k < 5; ++k) R[k] += F[k];
        } else {
            for (int k = 0; k < 5; ++k) R[k] -= F[k];
        }
    }
    bool ok = true;
    for (int k = 0; k < 5; ++k) {
        a.R[k * (size_t)m.N + c] = R[k];
        ok = ok && std::isfinite(R[k]);
    }
    if (!ok && a.bad) atomicAdd(a.bad, 1);
    if (a.kind < 0) return;
    const double vol = m.vol[c];
    for (int k = 0; k < 5; ++k) {
        const size_t q = k * (size_t)m.N + c;
        const double u = a.U[q];
        if (a.kind == 0) a.U[q] = u + a.dt / vol * R[k];
        else if (a.kind == 1) a.U[q] = 3.0 / 4.0 * a.Uo[q] + 1.0 / 4.0 * u + 1.0 / 4.0 * a.dt / vol * R[k];
        else a.U[q] = 1.0 / 3.0 * a.Uo[q] + 2.0 / 3.0 * u + 2.0 / 3.0 * a.dt / vol * R[k];
    }
}

// the stage update alone, after the host has checked the flux balance
__global__ __launch_bounds__(PB) void k_poly_update(PolyDev m, double *U, const double *Uo, const double *R,
                                                    int kind, double dt) {
    const uint32_t c = blockIdx.x * PB + threadIdx.x;
    if (c >= m.N) return;
    const double vol = m.vol[c];
    for (int k = 0; k < 5; ++k) {
        const size_t q = k * (size_t)m.N + c;
        const double u = U[q];
        if (kind == 0) U[q] = u + dt / vol * R[q];
        else if (kind == 1) U[q] = 3.0 / 4.0 * Uo[q] + 1.0 / 4.0 * u + 1.0 / 4.0 * dt / vol * R[q];
        else U[q] = 1.0 / 3.0 * Uo[q] + 2.0 / 3.0 * u + 2.0 / 3.0 * dt / vol * R[q];
    }
}

// calc_timestep (src/Solver.cpp:232-274): per-block minimum of the cell steps
__global__ __launch_bounds__(PB) void k_poly_dt(PolyDev m, const double *V, double CFL, double *part) {
    __shared__ double sh[PB];
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
    double *U = nullptr, *Uo = nullptr, *V = nullptr, *Ve = nullptr, *VL = nullptr, *VR = nullptr;
    double *grad = nullptr, *psi = nullptr, *R = nullptr, *part = nullptr;
    int *bad = nullptr;
    bool aux_valid = false;
    std::vector<void *> owned;
    ~PolyPath() {
        if (st) (void)hipStreamSynchronize(st);
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

// AoS [N][cols] <-> SoA [cols][N]
int put_soa(PolyPath *P, double *dst, const double *aos, int cols, std::string &err) {
    const uint32_t N = P->m.N;
    std::vector<double> soa((size_t)cols * N);
    for (uint32_t c = 0; c < N; ++c)
        for (int k = 0; k < cols; ++k) soa[k * (size_t)N + c] = aos[(size_t)cols * c + k];
    PHIP(hipMemcpyAsync(dst, soa.data(), soa.size() * 8, hipMemcpyHostToDevice, P->st));
    PHIP(hipStreamSynchronize(P->st));
    return NF_OK;
}
int get_soa(PolyPath *P, const double *src, int cols, double *aos, std::string &err) {
    const uint32_t N = P->m.N;
    std::vector<double> soa((size_t)cols * N);
    PHIP(hipStreamSynchronize(P->st));
    PHIP(hipMemcpy(soa.data(), src, soa.size() * 8, hipMemcpyDeviceToHost));
    for (uint32_t c = 0; c < N; ++c)
        for (int k = 0; k < cols; ++k) aos[(size_t)cols * c + k] = soa[k * (size_t)N + c];
    return NF_OK;
}

// evaluate_flux_balance of U into R (and the stage update when kind >= 0):
// primvars into V, ghosts, sweep A for second order, sweep B
int flux_balance(PolyPath *P, double *V, int kind, double dt, std::string &err) {
    const PolyDev &m = P->m;
    const uint32_t G = m.F - m.Fint;
    const bool check = P->prm.check_physical_validity != 0;
    hipLaunchKernelGGL(k_poly_prim, grid(m.N), dim3(PB), 0, P->st, (const double *)P->U, V, m.N, m.NT);
    if (G) hipLaunchKernelGGL(k_poly_ghost, grid(G), dim3(PB), 0, P->st, m, V, P->fs);
    const bool second = P->prm.spatial_order == NF_SECOND_ORDER;
    if (second)
        hipLaunchKernelGGL(k_poly_sweepA, grid(m.N), dim3(PB), 0, P->st, m, (const double *)V,
                           P->prm.limiter != NF_LIMITER_NONE ? 1 : 0, P->grad, P->psi, P->VL, P->VR);
    StageArgs a;
    a.V = V;
    a.VL = second ? P->VL : nullptr;
    a.VR = second ? P->VR : nullptr;
    a.U = P->U;
    a.Uo = P->Uo;
    a.R = P->R;
    a.bad = check ? P->bad : nullptr;
    a.flux = P->prm.inv_flux_scheme == NF_FLUX_HLLC ? 1 : 0;
    a.kind = check ? -1 : kind;
    a.dt = dt;
    a.fs = P->fs;
    if (check) PHIP(hipMemsetAsync(P->bad, 0, sizeof(int), P->st));
    hipLaunchKernelGGL(k_poly_sweepB, grid(m.N), dim3(PB), 0, P->st, m, a);
    PHIP(hipGetLastError());
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
                               (const double *)P->R, kind, dt);
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
    if (p.local_time_step) {
        err = "a mixed-element mesh takes the global time step only";
        return NF_E_ARG;
    }
    std::unique_ptr<PolyPath> hold(new PolyPath);
    PolyPath *P = hold.get();
    P->st = st;
    P->prm = p;
    P->fs = make_farfield<double>(p.primvars_inf);
    const uint32_t N = g.N, NT = g.NT, Fint = g.Fint, F = g.F;
    P->sizes = {N, NT, Fint, F, (uint32_t)g.pfirst.size()};
    // faces of each cell in increasing face index
    std::vector<uint32_t> cf((size_t)6 * N, NONE);
    std::vector<uint8_t> cnt(N, 0);
    std::vector<uint32_t> fi(F), fj(F);
    for (uint32_t f = 0; f < F; ++f) {
        const uint32_t i = g.ij[2 * (size_t)f], j = g.ij[2 * (size_t)f + 1];
        fi[f] = i;
        fj[f] = j;
        if (cnt[i] >= 6) {
            err = "cell with more than six faces";
            return NF_E_MESH;
        }
        cf[cnt[i]++ * (size_t)N + i] = f;
        if (j < N) {
            if (cnt[j] >= 6) {
                err = "cell with more than six faces";
                return NF_E_MESH;
            }
            cf[cnt[j]++ * (size_t)N + j] = f | SIDE_J;
        }
    }
    auto soa3 = [&](const std::vector<double> &a, uint32_t n) {
        std::vector<double> o((size_t)3 * n);
        for (uint32_t q = 0; q < n; ++q)
            for (int d = 0; d < 3; ++d) o[d * (size_t)n + q] = a[3 * (size_t)q + d];
        return o;
    };
    std::vector<uint8_t> bbc(F - Fint);
    for (size_t q = 0; q < g.pfirst.size(); ++q)
        for (uint32_t f = g.pfirst[q]; f < g.pfirst[q] + g.pn[q]; ++f) bbc[f - Fint] = (uint8_t)g.pbc[q];
    std::vector<double> vol(g.vol.begin(), g.vol.begin() + N);
    std::vector<double> dS((size_t)3 * N);
    for (uint32_t c = 0; c < N; ++c)
        for (int d = 0; d < 3; ++d) dS[d * (size_t)N + c] = g.dS[3 * (size_t)c + d];
    uint32_t *dcf, *dfi, *dfj;
    double *dSv, *dri, *drj, *dvol, *ddS;
    uint8_t *dbbc;
    int rc;
    if ((rc = alloc(P, &dcf, cf.size(), err)) || (rc = alloc(P, &dfi, F, err)) || (rc = alloc(P, &dfj, F, err)) ||
        (rc = alloc(P, &dSv, 3 * (size_t)F, err)) || (rc = alloc(P, &dri, 3 * (size_t)F, err)) ||
        (rc = alloc(P, &drj, 3 * (size_t)F, err)) || (rc = alloc(P, &dvol, N, err)) ||
        (rc = alloc(P, &ddS, 3 * (size_t)N, err)) || (rc = alloc(P, &dbbc, bbc.size(), err)))
        return rc;
    if ((rc = upload(dcf, cf, err)) || (rc = upload(dfi, fi, err)) || (rc = upload(dfj, fj, err)) ||
        (rc = upload(dSv, soa3(g.S, F), err)) || (rc = upload(dri, soa3(g.ri, F), err)) ||
        (rc = upload(drj, soa3(g.rj, F), err)) || (rc = upload(dvol, vol, err)) || (rc = upload(ddS, dS, err)) ||
        (rc = upload(dbbc, bbc, err)))
        return rc;
    P->m = {N, NT, Fint, F, dcf, dfi, dfj, dSv, dri, drj, dvol, ddS, dbbc};
    if ((rc = alloc(P, &P->U, 5 * (size_t)N, err)) || (rc = alloc(P, &P->Uo, 5 * (size_t)N, err)) ||
        (rc = alloc(P, &P->V, 5 * (size_t)NT, err)) || (rc = alloc(P, &P->Ve, 5 * (size_t)NT, err)) ||
        (rc = alloc(P, &P->VL, 5 * (size_t)F, err)) || (rc = alloc(P, &P->VR, 5 * (size_t)Fint, err)) ||
        (rc = alloc(P, &P->grad, 15 * (size_t)N, err)) || (rc = alloc(P, &P->psi, 5 * (size_t)N, err)) ||
        (rc = alloc(P, &P->R, 5 * (size_t)N, err)) || (rc = alloc(P, &P->part, (N + PB - 1) / PB, err)) ||
        (rc = alloc(P, &P->bad, 1, err)))
        return rc;
    if ((rc = poly_set_freestream(P, err))) return rc;
    *out = hold.release();
    return NF_OK;
}

void poly_destroy(PolyPath *P) { delete P; }

nf_fv_sizes poly_sizes(const PolyPath *P) { return P->sizes; }

// set_freestream_values (src/SolverData.cpp:50-58)
int poly_set_freestream(PolyPath *P, std::string &err) {
    const uint32_t N = P->m.N;
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
    if ((rc = put_soa(P, P->U, U, 5, err))) return rc;
    hipLaunchKernelGGL(k_poly_prim, grid(N), dim3(PB), 0, P->st, (const double *)P->U, P->V, N, P->m.NT);
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
    hipLaunchKernelGGL(k_poly_dt, dim3(nb), dim3(PB), 0, P->st, P->m, (const double *)P->V, P->prm.CFL, P->part);
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
    if ((rc = flux_balance(P, P->V, 0, dt, err))) return rc;
    if (P->prm.time_scheme == NF_TVD_RK3) {
        if ((rc = flux_balance(P, P->V, 1, dt, err))) return rc;
        if ((rc = flux_balance(P, P->V, 2, dt, err))) return rc;
    }
    PHIP(hipMemcpyAsync(P->Uo, P->U, 5 * (size_t)P->m.N * 8, hipMemcpyDeviceToDevice, P->st));
    PHIP(hipStreamSynchronize(P->st));
    return NF_OK;
}

// the flux balance of the current state, once, with its own primvar buffer:
// the primvars a following calc_timestep reads are left as the last step set them
int poly_eval_flux_balance(PolyPath *P, std::string &err) {
    int rc;
    if ((rc = flux_balance(P, P->Ve, -1, 0.0, err))) return rc;
    PHIP(hipStreamSynchronize(P->st));
    P->aux_valid = true;
    return NF_OK;
}

int poly_get_state(PolyPath *P, double *U, std::string &err) { return get_soa(P, P->U, 5, U, err); }

int poly_get_field(PolyPath *P, int32_t which, double *out, std::string &err) {
    switch (which) {
    case NF_FIELD_PRIM:  // prim(U) into Ve with row stride N for the export
        hipLaunchKernelGGL(k_poly_prim, grid(P->m.N), dim3(PB), 0, P->st, (const double *)P->U, P->Ve, P->m.N,
                           P->m.N);
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
