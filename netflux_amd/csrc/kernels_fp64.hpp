// FP64 stage kernels of the explicit Euler step.
//
// One thread per cell, gathering from the cell's four faces in increasing
// reference face index, so every accumulation (Green-Gauss gradient, flux
// balance, Delta S) adds the same terms in the same order as the reference's
// face-ordered scatter loops; no atomics, no colouring.  Geometry (volume,
// face normals, centroid-to-face vectors) is recomputed from the four nodes
// with the reference formulas instead of being streamed from memory.
//
//   sweep A (per stage): V -> Green-Gauss gradient, neighbour min/max, Barth
//            limiter, limited face values of the cell's four faces
//            (+ Delta S and the CFL time step of the cell when requested).
//   sweep B (per stage): face values -> Rusanov fluxes of the four faces
//            (wall/farfield ghosts from the reconstructed state), flux
//            balance, RK stage combination, prim(U) of the new state.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "physics.hpp"

namespace nf {

constexpr int BLOCK = 256;

struct DevMesh {
    uint32_t N;
    const double *nodes;    // [n_nodes][4] x,y,z,pad
    const uint32_t *cnodes; // [4][N]
    const uint32_t *nbr;    // [4][N]
    const uint32_t *fcode;  // [N]
};

using V3d = Vec3T<double>;

__device__ __forceinline__ V3d vsub(V3d a, V3d b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3d vadd(V3d a, V3d b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3d vscl(double s, V3d a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ V3d vcross(V3d a, V3d b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3d vsel(uint32_t q, const V3d *P) {
    V3d r = P[0];
    r = (q == 1) ? P[1] : r;
    r = (q == 2) ? P[2] : r;
    r = (q == 3) ? P[3] : r;
    return r;
}

// Logical block index: blocks are dispatched round robin over the 8 XCDs;
// give each XCD a contiguous range of cells so that neighbour gathers hit the
// L2 of the XCD that produced/consumed the same region.
__device__ __forceinline__ uint32_t logical_block(uint32_t b, uint32_t nb, int xcd_remap) {
    if (!xcd_remap) return b;
    const uint32_t x = b & 7u, i = b >> 3, q = nb >> 3, r = nb & 7u;
    return (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

struct CellGeo {
    V3d P[4];
    V3d cen;
    double vol;
};

__device__ __forceinline__ void load_cell(const DevMesh &m, uint32_t c, CellGeo &g) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t n = m.cnodes[(size_t)q * m.N + c];
        const double2 a = *reinterpret_cast<const double2 *>(m.nodes + 4 * (size_t)n);
        const double b = m.nodes[4 * (size_t)n + 2];
        g.P[q] = {a.x, a.y, b};
    }
    // tetrahedron_calc_volume / _centroid (src/geometry/GeometryUtilities.cpp:289-316)
    const V3d ab = vsub(g.P[1], g.P[0]), ac = vsub(g.P[2], g.P[0]), ad = vsub(g.P[3], g.P[0]);
    g.vol = dot3(vcross(ab, ac), ad) / 6;
    g.cen = vscl(1.0 / 4, vadd(vadd(vadd(g.P[0], g.P[1]), g.P[2]), g.P[3]));
}

// Face normal oriented i->j and face centroid (triangle formulas of
// src/geometry/GeometryUtilities.cpp:318-329, flip of src/geometry/
// FV_Grid.cpp:247-250) from the face element's node order.
__device__ __forceinline__ void face_geo(const CellGeo &g, uint32_t byte, V3d &S, V3d &fc) {
    const V3d a = vsel(byte & 3u, g.P), b = vsel((byte >> 2) & 3u, g.P), c = vsel((byte >> 4) & 3u, g.P);
    S = vscl(0.5, vcross(vsub(b, a), vsub(c, a)));
    if (byte & 0x40u) S = vscl(-1, S);
    const V3d s = vadd(vadd(a, b), c);
    fc = {s.x / 3, s.y / 3, s.z / 3};
}

// wave + block min / sum; result valid in thread 0
template <bool MIN>
__device__ __forceinline__ double block_reduce(double v, double *sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o, 64);
        v = MIN ? ((w < v) ? w : v) : v + w;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        v = sh[0];
        for (int q = 1; q < BLOCK / 64; ++q) v = MIN ? ((sh[q] < v) ? sh[q] : v) : v + sh[q];
    }
    return v;
}

struct SweepAArgs {
    const double *V;   // [5][N]
    double *FV;        // [4][5][N]
    double *dt_part;   // per-block CFL dt minimum, or null
    double CFL;
    FarField<double> fs;
    const int *stop;
    int xcd_remap;
    double *grad_aux;  // [15][N] column-major 5x3 per cell, or null
    double *psi_aux;   // [5][N] or null
};

// ORDER: 0 first order, 1 second order without limiter, 2 second order Barth
template <int ORDER>
__global__ __launch_bounds__(BLOCK) void k_sweepA(DevMesh m, SweepAArgs a) {
    __shared__ double sh[BLOCK / 64];
    if (*a.stop) return;
    const uint32_t N = m.N;
    const uint32_t c = logical_block(blockIdx.x, gridDim.x, a.xcd_remap) * BLOCK + threadIdx.x;
    const bool live = c < N;
    double dtc = DBL_MAX;
    if (live) {
        double Vc[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) Vc[k] = a.V[(size_t)k * N + c];
        CellGeo g;
        load_cell(m, c, g);
        const uint32_t code = m.fcode[c];
        double grad[5][3];
        double vmax[5], vmin[5];
        V3d r[4];
        double dS[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            grad[k][0] = grad[k][1] = grad[k][2] = 0.0;
            vmax[k] = -DBL_MAX;
            vmin[k] = DBL_MAX;
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const uint32_t byte = (code >> (8 * s)) & 0xFFu;
            const bool owner = byte & 0x80u;
            V3d S, fc;
            face_geo(g, byte, S, fc);
            r[s] = vsub(fc, g.cen);
            const uint32_t w = m.nbr[(size_t)s * N + c];
            const bool bnd = w & 0x80000000u;
            if (!bnd) {
                dS[0] += 0.5 * sabs(S.x);
                dS[1] += 0.5 * sabs(S.y);
                dS[2] += 0.5 * sabs(S.z);
            }
            if (ORDER == 0) continue;
            double Vn[5];
            if (bnd) {
                ghost_value<double>((int)(w & 0xFFu), Vc, S, a.fs, Vn);
            } else {
                const uint32_t n = w & 0x1FFFFFFFu;
#pragma unroll
                for (int k = 0; k < 5; ++k) Vn[k] = a.V[(size_t)k * N + n];
            }
            // Green-Gauss (include/Numerics.hpp:112-127); i/j as in the reference
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const double vi = owner ? Vc[k] : Vn[k], vj = owner ? Vn[k] : Vc[k];
                const double uf = 0.5 * (vi + vj);
                const double t0 = uf * S.x / g.vol, t1 = uf * S.y / g.vol, t2 = uf * S.z / g.vol;
                if (owner) {
                    grad[k][0] = grad[k][0] + t0;
                    grad[k][1] = grad[k][1] + t1;
                    grad[k][2] = grad[k][2] + t2;
                } else {
                    grad[k][0] = grad[k][0] - t0;
                    grad[k][1] = grad[k][1] - t1;
                    grad[k][2] = grad[k][2] - t2;
                }
                if (ORDER == 2) {
                    // neighbour min/max (include/Numerics.hpp:276-291)
                    vmax[k] = smax(vmax[k], smax(vi, vj));
                    vmin[k] = smin(vmin[k], smin(vi, vj));
                }
            }
        }
        if (ORDER == 0) {
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int k = 0; k < 5; ++k) a.FV[((size_t)s * 5 + k) * N + c] = Vc[k];
        } else {
            double delta[4][5];
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int k = 0; k < 5; ++k)
                    delta[s][k] = (grad[k][0] * r[s].x + grad[k][1] * r[s].y) + grad[k][2] * r[s].z;
            double psi[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                if (ORDER == 2) {
                    // Barth limiter (include/Numerics.hpp:179-236), faces in order
                    double p = DBL_MAX;
#pragma unroll
                    for (int s = 0; s < 4; ++s) p = smin(p, barth_phi(delta[s][k], vmax[k], vmin[k], Vc[k]));
                    psi[k] = p;
                } else {
                    psi[k] = 1.0;
                }
            }
            // limited reconstruction (include/Numerics.hpp:135-145)
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int k = 0; k < 5; ++k)
                    a.FV[((size_t)s * 5 + k) * N + c] = Vc[k] + psi[k] * delta[s][k];
            if (a.grad_aux) {
#pragma unroll
                for (int k = 0; k < 5; ++k)
#pragma unroll
                    for (int d = 0; d < 3; ++d) a.grad_aux[(size_t)(k + 5 * d) * N + c] = grad[k][d];
#pragma unroll
                for (int k = 0; k < 5; ++k) a.psi_aux[(size_t)k * N + c] = psi[k];
            }
        }
        if (a.dt_part) {
            // EulerSolver::calc_timestep (src/Solver.cpp:254-266)
            const double cs = sound_speed_prim(Vc);
            const double sx = (sabs(Vc[1]) + cs) * dS[0];
            const double sy = (sabs(Vc[2]) + cs) * dS[1];
            const double sz = (sabs(Vc[3]) + cs) * dS[2];
            const double t = a.CFL * g.vol / ((sx + sy) + sz);
            dtc = (t < DBL_MAX) ? t : DBL_MAX;  // NaN/+inf never win std::min
        }
    }
    if (a.dt_part) {
        const double v = block_reduce<true>(dtc, sh);
        if (threadIdx.x == 0) a.dt_part[blockIdx.x] = v;
    }
}

struct SweepBArgs {
    const double *FV;  // [4][5][N]
    const double *U;   // [5][N] state entering the stage
    const double *Uo;  // [5][N] U_old, or null for zero
    double *Uout;      // [5][N] (may alias U)
    double *Vout;      // [5][N] prim(Uout)
    const double *dt;  // device scalar
    int kind;          // 1: U + dt/vol R, 2: RK3 stage 2, 3: RK3 stage 3
    double *res_part;  // per-block sum of R_rho^2, or null
    int *bad_flux;     // count of cells with non-finite R (if check)
    int check;
    int *bad_state;    // count of cells with rho<=0, p<=0 or NaN in Uout
    FarField<double> fs;
    const int *stop;
    int xcd_remap;
    double *R_aux;     // [5][N] or null: store R, skip the update
};

__global__ __launch_bounds__(BLOCK) void k_sweepB(DevMesh m, SweepBArgs a) {
    __shared__ double sh[BLOCK / 64];
    if (*a.stop) return;
    const uint32_t N = m.N;
    const uint32_t c = logical_block(blockIdx.x, gridDim.x, a.xcd_remap) * BLOCK + threadIdx.x;
    const bool live = c < N;
    double r2 = 0.0;
    if (live) {
        CellGeo g;
        load_cell(m, c, g);
        const uint32_t code = m.fcode[c];
        double R[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const uint32_t byte = (code >> (8 * s)) & 0xFFu;
            const bool owner = byte & 0x80u;
            V3d S, fc;
            face_geo(g, byte, S, fc);
            const uint32_t w = m.nbr[(size_t)s * N + c];
            double Vo[5], Vn[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) Vo[k] = a.FV[((size_t)s * 5 + k) * N + c];
            if (w & 0x80000000u) {
                ghost_value<double>((int)(w & 0xFFu), Vo, S, a.fs, Vn);
            } else {
                const uint32_t n = w & 0x1FFFFFFFu, sn = (w >> 29) & 3u;
#pragma unroll
                for (int k = 0; k < 5; ++k) Vn[k] = a.FV[((size_t)sn * 5 + k) * N + n];
            }
            double UL[5], UR[5], F[5];
            // interior faces (src/Solver.cpp:171-192), boundary (:197-229)
            if (owner) {
                prim_to_cons(Vo, UL);
                prim_to_cons(Vn, UR);
            } else {
                prim_to_cons(Vn, UL);
                prim_to_cons(Vo, UR);
            }
            rusanov(UL, UR, S, F);
#pragma unroll
            for (int k = 0; k < 5; ++k) R[k] = owner ? R[k] - F[k] : R[k] + F[k];
        }
        if (a.check) {
            bool ok = true;
#pragma unroll
            for (int k = 0; k < 5; ++k) ok = ok && (R[k] - R[k] == 0.0);
            if (!ok) atomicAdd(a.bad_flux, 1);
        }
        r2 = R[0] * R[0];
        if (a.R_aux) {
#pragma unroll
            for (int k = 0; k < 5; ++k) a.R_aux[(size_t)k * N + c] = R[k];
        } else {
            const double dt = *a.dt;
            double Un[5], Vn[5];
            // Solver::TVD_RK3 / explicit_euler (src/Solver.cpp:98-132)
            if (a.kind == 1) {
                const double f = dt / g.vol;
#pragma unroll
                for (int k = 0; k < 5; ++k) Un[k] = a.U[(size_t)k * N + c] + f * R[k];
            } else if (a.kind == 2) {
                const double f = 1.0 / 4.0 * dt / g.vol;
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    const double uo = a.Uo ? a.Uo[(size_t)k * N + c] : 0.0;
                    Un[k] = 3.0 / 4.0 * uo + 1.0 / 4.0 * a.U[(size_t)k * N + c] + f * R[k];
                }
            } else {
                const double f = 2.0 / 3.0 * dt / g.vol;
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    const double uo = a.Uo ? a.Uo[(size_t)k * N + c] : 0.0;
                    Un[k] = 1.0 / 3.0 * uo + 2.0 / 3.0 * a.U[(size_t)k * N + c] + f * R[k];
                }
            }
            cons_to_prim(Un, Vn);
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                a.Uout[(size_t)k * N + c] = Un[k];
                a.Vout[(size_t)k * N + c] = Vn[k];
            }
            if (!(Vn[0] > 0.0) || !(Vn[4] > 0.0) || !(Un[4] > 0.0)) atomicAdd(a.bad_state, 1);
        }
    }
    if (a.res_part) {
        const double v = block_reduce<false>(r2, sh);
        if (threadIdx.x == 0) a.res_part[blockIdx.x] = v;
    }
}

}  // namespace nf
