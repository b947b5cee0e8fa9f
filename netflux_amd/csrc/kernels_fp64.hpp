// FP64 stage kernels of the explicit Euler step.
//
// Mapping: one tetrahedron per lane quad, one face per lane (lane = 4*cell +
// slot).  Slot s of a cell is its s-th face in increasing reference face
// index, and every per-cell accumulation is formed by adding the four lane
// contributions of the quad in slot order, ((((0 + t0) + t1) + t2) + t3),
// with DPP quad permutes -- the same terms in the same order as the
// reference's face-ordered scatter loops, so FP64 results are bit-identical
// to a serial evaluation, with no atomics and no colouring.  Geometry
// (volume, centroid, face normal, centroid-to-face vector) is recomputed from
// the four nodes with the reference formulas instead of streamed.
//
//   sweep A (per stage): V -> Green-Gauss gradient, neighbour min/max, Barth
//            limiter, limited face value of the lane's face
//            (+ Delta S and the CFL time step when requested).
//   sweep B (per stage): face values -> Rusanov flux of the lane's face
//            (wall/farfield ghost from the reconstructed state), flux balance,
//            RK stage combination, prim(U) of the new state.
//
// Layouts (device cell order): V, U: [N][5]; face values FV: [N][4][5];
// cnodes, nbr: [N][4]; fcode: [N]; nodes: [n_nodes][4] (x, y, z, pad).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "physics.hpp"

namespace nf {

constexpr int BLOCK = 256;            // threads
constexpr int CELLS = BLOCK / 4;      // cells per block

struct DevMesh {
    uint32_t N;
    double three, six;  // 3.0 and 6.0, divisors of the centroid / volume formulas
    const double *nodes;
    const uint32_t *cnodes;
    const uint32_t *nbr;
    const uint32_t *fcode;
};

using V3d = Vec3T<double>;

__device__ __forceinline__ V3d vsub(V3d a, V3d b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3d vadd(V3d a, V3d b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3d vscl(double s, V3d a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ V3d vcross(V3d a, V3d b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// node q of the cell, selected per scalar component so the choice stays in
// registers (v_cndmask) instead of becoming an indexed scratch access
__device__ __forceinline__ double dsel(uint32_t q, double a0, double a1, double a2, double a3) {
    double r = a0;
    r = (q == 1u) ? a1 : r;
    r = (q == 2u) ? a2 : r;
    r = (q == 3u) ? a3 : r;
    return r;
}
__device__ __forceinline__ V3d vsel(uint32_t q, const V3d &p0, const V3d &p1, const V3d &p2, const V3d &p3) {
    return {dsel(q, p0.x, p1.x, p2.x, p3.x), dsel(q, p0.y, p1.y, p2.y, p3.y), dsel(q, p0.z, p1.z, p2.z, p3.z)};
}

// value held by lane Q of this lane's quad (DPP quad_perm [Q,Q,Q,Q])
template <int Q>
__device__ __forceinline__ double qget(double v) {
    constexpr int ctrl = Q | (Q << 2) | (Q << 4) | (Q << 6);
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), ctrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ V3d qget3(V3d v, int q) {
    switch (q) {
    case 0: return {qget<0>(v.x), qget<0>(v.y), qget<0>(v.z)};
    case 1: return {qget<1>(v.x), qget<1>(v.y), qget<1>(v.z)};
    case 2: return {qget<2>(v.x), qget<2>(v.y), qget<2>(v.z)};
    default: return {qget<3>(v.x), qget<3>(v.y), qget<3>(v.z)};
    }
}
// ((((0 + x0) + x1) + x2) + x3) over the quad, identical in all four lanes
__device__ __forceinline__ double qsum(double x) {
    const double a = qget<0>(x), b = qget<1>(x), c = qget<2>(x), d = qget<3>(x);
    return (((0.0 + a) + b) + c) + d;
}
__device__ __forceinline__ double qmax(double init, double x) {
    const double a = qget<0>(x), b = qget<1>(x), c = qget<2>(x), d = qget<3>(x);
    return smax(smax(smax(smax(init, a), b), c), d);
}
__device__ __forceinline__ double qmin(double init, double x) {
    const double a = qget<0>(x), b = qget<1>(x), c = qget<2>(x), d = qget<3>(x);
    return smin(smin(smin(smin(init, a), b), c), d);
}

// IEEE maxNum/minNum of two doubles (v_max_f64 / v_min_f64).  Over finite
// values the maximum/minimum of a set does not depend on the order of
// evaluation, so the quad reductions below may use a butterfly; only the sign
// of a zero result can differ from a sequential std::max / std::min.
__device__ __forceinline__ double dmax(double a, double b) {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ double dmin(double a, double b) {
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
template <int CTRL>
__device__ __forceinline__ double qperm(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
// quad_perm [1,0,3,2] then [2,3,0,1]
__device__ __forceinline__ double qmax_all(double x) {
    x = dmax(x, qperm<0xB1>(x));
    return dmax(x, qperm<0x4E>(x));
}
__device__ __forceinline__ double qmin_all(double x) {
    x = dmin(x, qperm<0xB1>(x));
    return dmin(x, qperm<0x4E>(x));
}

// Division by a divisor used several times.  The f64 divide of the target is
// the sequence  div_scale, rcp, fma, fma, fma, fma (refined reciprocal y),
// mul q = a*y, fma r = a - b*q, div_fmas q + r*y, div_fixup.  Outside extreme
// exponents (|a| < 2^-969, denormal b or quotient) div_scale leaves both
// operands unchanged and div_fixup only restores the sign of a zero quotient,
// so refining y once per divisor and applying the last three steps gives the
// same quotient as a / b.
struct Rcp {
    double b, y;
};
__device__ __forceinline__ Rcp rcp_of(double b) {
    const double y0 = __builtin_amdgcn_rcp(b);
    const double e0 = __builtin_fma(-b, y0, 1.0);
    const double y1 = __builtin_fma(y0, e0, y0);
    const double e1 = __builtin_fma(-b, y1, 1.0);
    return {b, __builtin_fma(y1, e1, y1)};
}
__device__ __forceinline__ double rdiv(double a, const Rcp &r) {
    const double q = a * r.y;
    const double e = __builtin_fma(-r.b, q, a);
    return __builtin_fma(e, r.y, q);
}

// EulerEqs::cons_to_prim with one reciprocal of rho
__device__ __forceinline__ void cons_to_prim_d(const double *U, double *V) {
    const Rcp ir = rcp_of(U[0]);
    V[0] = U[0];
    V[1] = rdiv(U[1], ir);
    V[2] = rdiv(U[2], ir);
    V[3] = rdiv(U[3], ir);
    V[4] = Consts<double>::GM1 * (U[4] - rdiv(0.5, ir) * ((U[1] * U[1] + U[2] * U[2]) + U[3] * U[3]));
}

// NumericalFlux::rusanov (src/Numerics.cpp:10-18) with the EulerEqs helpers
// inlined and one reciprocal per repeated divisor (|S|, rho_L, rho_R)
__device__ __forceinline__ void rusanov_d(const double *UL, const double *UR, Vec3T<double> S, double *F) {
    const double z = (S.x * S.x + S.y * S.y) + S.z * S.z;
    const double area = sqrt(z);
    const Rcp ia = rcp_of(area);
    Vec3T<double> n = S;
    if (z > 0.0) n = {rdiv(S.x, ia), rdiv(S.y, ia), rdiv(S.z, ia)};
    const Rcp iL = rcp_of(UL[0]), iR = rcp_of(UR[0]);
    const double G = Consts<double>::GAMMA, GM1 = Consts<double>::GM1;
    const double pL = GM1 * (UL[4] - rdiv(0.5, iL) * ((UL[1] * UL[1] + UL[2] * UL[2]) + UL[3] * UL[3]));
    const double pR = GM1 * (UR[4] - rdiv(0.5, iR) * ((UR[1] * UR[1] + UR[2] * UR[2]) + UR[3] * UR[3]));
    const double vnL = rdiv((UL[1] * n.x + UL[2] * n.y) + UL[3] * n.z, iL);
    const double vnR = rdiv((UR[1] * n.x + UR[2] * n.y) + UR[3] * n.z, iR);
    const double cL = sqrt(rdiv(G * pL, iL));
    const double cR = sqrt(rdiv(G * pR, iR));
    const double sL = fabs(vnL) + cL, sR = fabs(vnR) + cR;
    const double sm = smax(sR, sL);
    const double FR[5] = {vnR * UR[0], vnR * UR[1] + pR * n.x, vnR * UR[2] + pR * n.y,
                          vnR * UR[3] + pR * n.z, vnR * (UR[4] + pR)};
    const double FL[5] = {vnL * UL[0], vnL * UL[1] + pL * n.x, vnL * UL[2] + pL * n.y,
                          vnL * UL[3] + pL * n.z, vnL * (UL[4] + pL)};
    const double h = area * 0.5;
#pragma unroll
    for (int k = 0; k < 5; ++k) F[k] = h * ((FR[k] + FL[k]) - sm * (UR[k] - UL[k]));
}

// BC_SlipWall / BC_NoSlipWall (src/Numerics.cpp:59-81); farfield via physics.hpp
__device__ __forceinline__ void ghost_d(int bc, const double *Vd, Vec3T<double> S, const FarField<double> &fs,
                                        double *Vg) {
    if (bc == BC_SLIP) {
        const double z = (S.x * S.x + S.y * S.y) + S.z * S.z;
        const Rcp ia = rcp_of(sqrt(z));
        Vec3T<double> n = S;
        if (z > 0.0) n = {rdiv(S.x, ia), rdiv(S.y, ia), rdiv(S.z, ia)};
        const double vn = (Vd[1] * n.x + Vd[2] * n.y) + Vd[3] * n.z;
        Vg[0] = Vd[0];
        Vg[1] = Vd[1] - 2.0 * vn * n.x;
        Vg[2] = Vd[2] - 2.0 * vn * n.y;
        Vg[3] = Vd[3] - 2.0 * vn * n.z;
        Vg[4] = Vd[4];
    } else {
        ghost_value<double>(bc, Vd, S, fs, Vg);
    }
}

// Barth face contribution (include/Numerics.hpp:190-206); min(1, x) as
// minNum, which also returns 1 for a NaN quotient like std::min(1.0, x)
__device__ __forceinline__ double barth_phi_d(double delta, double vmax, double vmin, double v) {
    const double d2 = double(isign(delta)) * (fabs(delta) + DBL_EPSILON);
    if (d2 > 0.0) return dmin(1.0, (vmax - v) / d2);
    if (d2 < 0.0) return dmin(1.0, (vmin - v) / d2);
    return 1.0;
}

// Logical block index: blocks are dispatched round robin over the 8 XCDs;
// give each XCD a contiguous range of cells so that neighbour gathers hit the
// L2 of the XCD that touched the same region.
__device__ __forceinline__ uint32_t logical_block(uint32_t b, uint32_t nb, int xcd_remap) {
    if (!xcd_remap) return b;
    const uint32_t x = b & 7u, i = b >> 3, q = nb >> 3, r = nb & 7u;
    return (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

struct LaneGeo {
    V3d p0, p1, p2, p3;
    V3d cen;
    double vol;
    Rcp ivol;
};

// Each lane loads node `slot` of its cell; the quad shares the four nodes.
__device__ __forceinline__ void load_cell(const DevMesh &m, uint32_t c, uint32_t s, LaneGeo &g) {
    const uint32_t n = m.cnodes[4 * (size_t)c + s];
    const double2 a = *reinterpret_cast<const double2 *>(m.nodes + 4 * (size_t)n);
    const double b = m.nodes[4 * (size_t)n + 2];
    const V3d mine = {a.x, a.y, b};
    g.p0 = qget3(mine, 0);
    g.p1 = qget3(mine, 1);
    g.p2 = qget3(mine, 2);
    g.p3 = qget3(mine, 3);
    // tetrahedron_calc_volume / _centroid (src/geometry/GeometryUtilities.cpp:289-316)
    const V3d ab = vsub(g.p1, g.p0), ac = vsub(g.p2, g.p0), ad = vsub(g.p3, g.p0);
    g.vol = rdiv(dot3(vcross(ab, ac), ad), rcp_of(m.six));
    g.cen = vscl(1.0 / 4, vadd(vadd(vadd(g.p0, g.p1), g.p2), g.p3));
    g.ivol = rcp_of(g.vol);
}

// Face normal oriented i->j and face centroid (triangle formulas of
// src/geometry/GeometryUtilities.cpp:318-329, flip of src/geometry/
// FV_Grid.cpp:247-250) from the face element's node order.
__device__ __forceinline__ void face_geo(const DevMesh &m, const LaneGeo &g, uint32_t byte, V3d &S, V3d &fc) {
    const V3d a = vsel(byte & 3u, g.p0, g.p1, g.p2, g.p3);
    const V3d b = vsel((byte >> 2) & 3u, g.p0, g.p1, g.p2, g.p3);
    const V3d c = vsel((byte >> 4) & 3u, g.p0, g.p1, g.p2, g.p3);
    S = vscl(0.5, vcross(vsub(b, a), vsub(c, a)));
    if (byte & 0x40u) S = vscl(-1, S);
    const V3d s = vadd(vadd(a, b), c);
    const Rcp i3 = rcp_of(m.three);
    fc = {rdiv(s.x, i3), rdiv(s.y, i3), rdiv(s.z, i3)};
}

// wave + block min / sum over all threads; result valid in thread 0
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
    const double *V;   // [N][5]
    double *FV;        // [N][4][5]
    double *dt_part;   // per-block CFL dt minimum, or null
    double CFL;
    FarField<double> fs;
    const int *stop;
    int xcd_remap;
    double *grad_aux;  // [N][15] column-major 5x3 per cell, or null
    double *psi_aux;   // [N][5] or null
};

// ORDER: 0 first order, 1 second order without limiter, 2 second order Barth
template <int ORDER>
__global__ __launch_bounds__(BLOCK) void k_sweepA(DevMesh m, SweepAArgs a) {
    __shared__ double sh[BLOCK / 64];
    if (*a.stop) return;
    const uint32_t N = m.N;
    const uint32_t s = threadIdx.x & 3u;
    const uint32_t c0 = logical_block(blockIdx.x, gridDim.x, a.xcd_remap) * CELLS + (threadIdx.x >> 2);
    const bool live = c0 < N;
    const uint32_t c = live ? c0 : N - 1;  // dead quads mirror the last cell, store nothing

    double Vc[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) Vc[k] = a.V[5 * (size_t)c + k];
    LaneGeo g;
    load_cell(m, c, s, g);
    const uint32_t byte = (m.fcode[c] >> (8 * s)) & 0xFFu;
    const bool owner = byte & 0x80u;
    V3d S, fc;
    face_geo(m, g, byte, S, fc);
    const V3d r = vsub(fc, g.cen);
    const uint32_t w = m.nbr[4 * (size_t)c + s];
    const bool bnd = w & 0x80000000u;
    double dtc = DBL_MAX;

    if (ORDER == 0) {
        if (live) {
#pragma unroll
            for (int k = 0; k < 5; ++k) a.FV[(4 * (size_t)c + s) * 5 + k] = Vc[k];
        }
    } else {
        double Vn[5];
        if (bnd) {
            ghost_d((int)(w & 0xFFu), Vc, S, a.fs, Vn);
        } else {
            const uint32_t n = w & 0x1FFFFFFFu;
#pragma unroll
            for (int k = 0; k < 5; ++k) Vn[k] = a.V[5 * (size_t)n + k];
        }
        // One variable at a time: gradient row k, its neighbour min/max, the
        // Barth factor and the face value only depend on variable k, so the
        // five passes are kept apart (sched_barrier) to bound register use.
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            // Green-Gauss (include/Numerics.hpp:112-127): i/j as in the
            // reference; cell i adds tmp/vol, cell j subtracts it
            // (x - t == x + (-t) exactly)
            const double vi = owner ? Vc[k] : Vn[k], vj = owner ? Vn[k] : Vc[k];
            const double uf = 0.5 * (vi + vj);
            const double t0 = rdiv(uf * S.x, g.ivol), t1 = rdiv(uf * S.y, g.ivol), t2 = rdiv(uf * S.z, g.ivol);
            const double g0 = qsum(owner ? t0 : -t0);
            const double g1 = qsum(owner ? t1 : -t1);
            const double g2 = qsum(owner ? t2 : -t2);
            const double delta = (g0 * r.x + g1 * r.y) + g2 * r.z;
            double psi = 1.0;
            if (ORDER == 2) {
                // neighbour min/max (include/Numerics.hpp:276-291) and the Barth
                // limiter (include/Numerics.hpp:179-236), faces in order
                const double vmax = dmax(-DBL_MAX, qmax_all(dmax(vi, vj)));
                const double vmin = dmin(DBL_MAX, qmin_all(dmin(vi, vj)));
                psi = dmin(DBL_MAX, qmin_all(barth_phi_d(delta, vmax, vmin, Vc[k])));
            }
            if (live) {
                // limited reconstruction (include/Numerics.hpp:135-145)
                a.FV[(4 * (size_t)c + s) * 5 + k] = Vc[k] + psi * delta;
                if (a.grad_aux && s == 0) {
                    a.grad_aux[15 * (size_t)c + k] = g0;
                    a.grad_aux[15 * (size_t)c + k + 5] = g1;
                    a.grad_aux[15 * (size_t)c + k + 10] = g2;
                    a.psi_aux[5 * (size_t)c + k] = psi;
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    if (a.dt_part) {
        // Delta S over interior faces (src/Solver.cpp:287-297) and the CFL
        // step of the cell (src/Solver.cpp:254-266)
        const double h0 = bnd ? 0.0 : 0.5 * sabs(S.x), h1 = bnd ? 0.0 : 0.5 * sabs(S.y),
                     h2 = bnd ? 0.0 : 0.5 * sabs(S.z);
        const double d0 = qsum(h0), d1 = qsum(h1), d2 = qsum(h2);
        const double cs = sound_speed_prim(Vc);
        const double sx = (sabs(Vc[1]) + cs) * d0;
        const double sy = (sabs(Vc[2]) + cs) * d1;
        const double sz = (sabs(Vc[3]) + cs) * d2;
        const double t = a.CFL * g.vol / ((sx + sy) + sz);
        dtc = (live && t < DBL_MAX) ? t : DBL_MAX;  // NaN/+inf never win std::min
        const double v = block_reduce<true>(dtc, sh);
        if (threadIdx.x == 0) a.dt_part[blockIdx.x] = v;
    }
}

struct SweepBArgs {
    const double *FV;  // [N][4][5]
    const double *U;   // [N][5] state entering the stage
    const double *Uo;  // [N][5] U_old, or null for zero
    double *Uout;      // [N][5] (may alias U)
    double *Vout;      // [N][5] prim(Uout)
    const double *dt;  // device scalar
    int kind;          // 1: U + dt/vol R, 2: RK3 stage 2, 3: RK3 stage 3
    double *res_part;  // per-block sum of R_rho^2, or null
    int *bad_flux;     // count of cells with non-finite R (if check)
    int check;
    int *bad_state;    // count of cells with rho<=0, p<=0 or NaN in Uout
    FarField<double> fs;
    const int *stop;
    int xcd_remap;
    double *R_aux;     // [N][5] or null: store R, skip the update
};

__global__ __launch_bounds__(BLOCK) void k_sweepB(DevMesh m, SweepBArgs a) {
    __shared__ double sh[BLOCK / 64];
    if (*a.stop) return;
    const uint32_t N = m.N;
    const uint32_t s = threadIdx.x & 3u;
    const uint32_t c0 = logical_block(blockIdx.x, gridDim.x, a.xcd_remap) * CELLS + (threadIdx.x >> 2);
    const bool live = c0 < N;
    const uint32_t c = live ? c0 : N - 1;

    LaneGeo g;
    load_cell(m, c, s, g);
    const uint32_t byte = (m.fcode[c] >> (8 * s)) & 0xFFu;
    const bool owner = byte & 0x80u;
    V3d S, fc;
    face_geo(m, g, byte, S, fc);
    const uint32_t w = m.nbr[4 * (size_t)c + s];
    double Vo[5], Vn[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) Vo[k] = a.FV[(4 * (size_t)c + s) * 5 + k];
    if (w & 0x80000000u) {
        ghost_d((int)(w & 0xFFu), Vo, S, a.fs, Vn);
    } else {
        const uint32_t n = w & 0x1FFFFFFFu, sn = (w >> 29) & 3u;
#pragma unroll
        for (int k = 0; k < 5; ++k) Vn[k] = a.FV[(4 * (size_t)n + sn) * 5 + k];
    }
    // interior faces (src/Solver.cpp:171-192), boundary faces (:197-229)
    double UL[5], UR[5], F[5];
    prim_to_cons(owner ? Vo : Vn, UL);
    prim_to_cons(owner ? Vn : Vo, UR);
    rusanov_d(UL, UR, S, F);
    double R[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) R[k] = qsum(owner ? -F[k] : F[k]);  // R_i -= F, R_j += F

    if (live && s == 0 && a.check) {
        bool ok = true;
#pragma unroll
        for (int k = 0; k < 5; ++k) ok = ok && (R[k] - R[k] == 0.0);
        if (!ok) atomicAdd(a.bad_flux, 1);
    }
    if (a.R_aux) {
        if (live && s == 0) {
#pragma unroll
            for (int k = 0; k < 5; ++k) a.R_aux[5 * (size_t)c + k] = R[k];
        }
    } else {
        const double dt = *a.dt;
        double Un[5], Vnew[5];
        // Solver::TVD_RK3 / explicit_euler (src/Solver.cpp:98-132)
        if (a.kind == 1) {
            const double f = rdiv(dt, g.ivol);
#pragma unroll
            for (int k = 0; k < 5; ++k) Un[k] = a.U[5 * (size_t)c + k] + f * R[k];
        } else if (a.kind == 2) {
            const double f = rdiv(1.0 / 4.0 * dt, g.ivol);
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const double uo = a.Uo ? a.Uo[5 * (size_t)c + k] : 0.0;
                Un[k] = 3.0 / 4.0 * uo + 1.0 / 4.0 * a.U[5 * (size_t)c + k] + f * R[k];
            }
        } else {
            const double f = rdiv(2.0 / 3.0 * dt, g.ivol);
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const double uo = a.Uo ? a.Uo[5 * (size_t)c + k] : 0.0;
                Un[k] = 1.0 / 3.0 * uo + 2.0 / 3.0 * a.U[5 * (size_t)c + k] + f * R[k];
            }
        }
        cons_to_prim_d(Un, Vnew);
        if (live) {
            // lane s stores component s, lane 0 also component 4
            double us = Un[0], vs = Vnew[0];
            us = (s == 1) ? Un[1] : us;
            vs = (s == 1) ? Vnew[1] : vs;
            us = (s == 2) ? Un[2] : us;
            vs = (s == 2) ? Vnew[2] : vs;
            us = (s == 3) ? Un[3] : us;
            vs = (s == 3) ? Vnew[3] : vs;
            a.Uout[5 * (size_t)c + s] = us;
            a.Vout[5 * (size_t)c + s] = vs;
            if (s == 0) {
                a.Uout[5 * (size_t)c + 4] = Un[4];
                a.Vout[5 * (size_t)c + 4] = Vnew[4];
                if (!(Vnew[0] > 0.0) || !(Vnew[4] > 0.0) || !(Un[4] > 0.0)) atomicAdd(a.bad_state, 1);
            }
        }
    }
    if (a.res_part) {
        const double r2 = (live && s == 0) ? R[0] * R[0] : 0.0;
        const double v = block_reduce<false>(r2, sh);
        if (threadIdx.x == 0) a.res_part[blockIdx.x] = v;
    }
}

}  // namespace nf
