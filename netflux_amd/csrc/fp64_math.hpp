// Device helpers shared by the tetrahedral and the mixed-element FP64
// kernels: the XCD-aware block order, vector helpers, maxNum/minNum, exact
// sign flips, division by a reused divisor, cons->prim, the Rusanov flux,
// boundary ghosts and the Barth limiter, each giving the bits of the
// reference's plain FP64 formulas.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "physics.hpp"

namespace nf {

// Logical block index: blocks are dispatched round robin over the 8 XCDs;
// give each XCD a contiguous range of cells so that neighbour gathers hit the
// L2 of the XCD that touched the same region.
__device__ __forceinline__ uint32_t logical_block(uint32_t b, uint32_t nb, int xcd_remap) {
    if (!xcd_remap) return b;
    const uint32_t x = b & 7u, i = b >> 3, q = nb >> 3, r = nb & 7u;
    return (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

using V3d = Vec3T<double>;

__device__ __forceinline__ V3d vsub(V3d a, V3d b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3d vadd(V3d a, V3d b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3d vscl(double s, V3d a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ V3d vcross(V3d a, V3d b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// IEEE maxNum/minNum (v_max_f64 / v_min_f64): over finite values the
// maximum/minimum of a set does not depend on evaluation order; only the sign
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

// x with its sign bit xor-ed with m (m = 0 or 0x80000000 on the high word):
// exact negation without a select.  x - t == x + (-t) and s*(-x) == -(s*x)
// hold exactly in IEEE arithmetic, which lets the owner/neighbour side of a
// face be handled by sign flips.
__device__ __forceinline__ double flip(double x, uint32_t m) {
    return __hiloint2double(__double2hiint(x) ^ (int)m, __double2loint(x));
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
__device__ __forceinline__ void rusanov_d(const double *UL, const double *UR, V3d S, double *F) {
    const double z = (S.x * S.x + S.y * S.y) + S.z * S.z;
    const double area = sqrt(z);
    const Rcp ia = rcp_of(area);
    V3d n = S;
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

// The same flux seen from one cell: UO is the cell's own face state, UN the
// neighbour's, mneg = 0 if the cell is the reference owner i (UL = UO) and
// the sign bit otherwise (UL = UN).  F(UL,UR) = h((FR+FL) - s(UR-UL)); with
// UR-UL = -(UO-UN) for the owner this is h((FR+FL) + s(UO-UN)), and
// h((FR+FL) - s(UO-UN)) otherwise -- bit-identical to rusanov_d(UL, UR).
__device__ __forceinline__ void rusanov_side(const double *UO, const double *UN, V3d S, uint32_t mneg, double *F) {
    const double z = (S.x * S.x + S.y * S.y) + S.z * S.z;
    const double area = sqrt(z);
    const Rcp ia = rcp_of(area);
    V3d n = S;
    if (z > 0.0) n = {rdiv(S.x, ia), rdiv(S.y, ia), rdiv(S.z, ia)};
    const Rcp iO = rcp_of(UO[0]), iN = rcp_of(UN[0]);
    const double G = Consts<double>::GAMMA, GM1 = Consts<double>::GM1;
    const double pO = GM1 * (UO[4] - rdiv(0.5, iO) * ((UO[1] * UO[1] + UO[2] * UO[2]) + UO[3] * UO[3]));
    const double pN = GM1 * (UN[4] - rdiv(0.5, iN) * ((UN[1] * UN[1] + UN[2] * UN[2]) + UN[3] * UN[3]));
    const double vnO = rdiv((UO[1] * n.x + UO[2] * n.y) + UO[3] * n.z, iO);
    const double vnN = rdiv((UN[1] * n.x + UN[2] * n.y) + UN[3] * n.z, iN);
    const double cO = sqrt(rdiv(G * pO, iO));
    const double cN = sqrt(rdiv(G * pN, iN));
    const double sm = dmax(fabs(vnO) + cO, fabs(vnN) + cN);
    const double FO[5] = {vnO * UO[0], vnO * UO[1] + pO * n.x, vnO * UO[2] + pO * n.y,
                          vnO * UO[3] + pO * n.z, vnO * (UO[4] + pO)};
    const double FN[5] = {vnN * UN[0], vnN * UN[1] + pN * n.x, vnN * UN[2] + pN * n.y,
                          vnN * UN[3] + pN * n.z, vnN * (UN[4] + pN)};
    const double h = area * 0.5;
#pragma unroll
    for (int k = 0; k < 5; ++k) F[k] = h * ((FO[k] + FN[k]) + flip(sm * (UO[k] - UN[k]), mneg));
}

// HLLC flux: hllc_d in physics.hpp

// BoundaryCondition::calc_ghost_val (src/Numerics.cpp:59-171); slip wall with
// one reciprocal of |S|, the others as in physics.hpp
__device__ __forceinline__ void ghost_d(int bc, const double *Vd, V3d S, const FarField<double> &fs, double *Vg) {
    if (bc == BC_SLIP) {
        const double z = (S.x * S.x + S.y * S.y) + S.z * S.z;
        const Rcp ia = rcp_of(sqrt(z));
        V3d n = S;
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

// Barth limiter of one variable over the cell's faces (four for a tetrahedron)
// (include/Numerics.hpp:179-236) with two divisions instead of one per face.  Faces
// with Delta_2 > 0 all divide vmax - v >= 0 by a positive Delta_2, so the
// smallest quotient is the one with the largest Delta_2, and because
// correctly rounded division is monotone, min_s round(A / d_s) =
// round(A / max_s d_s); likewise vmin - v <= 0 over the most negative
// Delta_2.  min(1, .) and the minimum over faces commute with this, a NaN
// Delta_2 falls in neither group (contribution 1, as in the reference) and
// Delta_2 == 0 contributes 1.  Only the sign of a zero limiter can differ
// from a face-by-face minimum, and it cannot reach a nonzero result.
// Delta_2 = sign(d) (|d| + eps) is positive exactly when d is, and rounding
// of x + eps is monotone in x, so the largest positive Delta_2 is
// max(d) + eps and the most negative is min(d) - eps (= -(|min d| + eps));
// maxNum/minNum against 0 drop NaN and the values of the other sign.
// A division whose quotient is at least 1 is skipped: for b > 0, a >= b
// gives a / b >= 1 and, 1 being representable, round(a / b) >= 1, so
// min(1, .) is 1 either way (for b < 0 the same holds with a <= b).  NaN
// operands fail both comparisons and give 1, as minNum(1, NaN) does.  The
// division then runs only in waves with a lane that limits.
// hi = maxNum(0, d_s), lo = minNum(0, d_s) over the cell's faces.
__device__ __forceinline__ double barth_psi_hilo(double hi, double lo, double vmax, double vmin, double v) {
    const double ap = vmax - v, bp = hi + DBL_EPSILON, am = vmin - v, bm = lo - DBL_EPSILON;
    double pp = 1.0, pm = 1.0;
    if (hi > 0.0 && ap < bp) pp = dmin(1.0, ap / bp);
    if (lo < 0.0 && am > bm) pm = dmin(1.0, am / bm);
    return dmin(pp, pm);
}
__device__ __forceinline__ double barth_psi_d(const double *delta, double vmax, double vmin, double v) {
    double hi = 0.0, lo = 0.0;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        hi = dmax(hi, delta[s]);
        lo = dmin(lo, delta[s]);
    }
    return barth_psi_hilo(hi, lo, vmax, vmin, v);
}

}  // namespace nf
