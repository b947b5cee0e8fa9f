// Euler physics shared by host and device code.
//
// Every expression is written in the association order in which the
// reference evaluates it (EulerEqs, include/SolverData.hpp:166-262;
// NumericalFlux::rusanov, src/Numerics.cpp:10-18; BCs, src/Numerics.cpp:
// 59-171), so that with floating-point contraction disabled the FP64 results
// are bit-identical to a serial evaluation of the reference formulas.
// Three-term sums are evaluated left to right: (a + b) + c.
#pragma once
#include <cfloat>
#include <cmath>

#if defined(__HIPCC__)
#define NF_HD __host__ __device__ __forceinline__
#else
#define NF_HD inline
#endif

namespace nf {

template <typename T> struct Consts;
template <> struct Consts<double> {
    static constexpr double GAMMA = 1.4;                   // include/Utilities.hpp:103
    static constexpr double GM1 = GAMMA - 1;               // include/SolverData.hpp:130
    static constexpr double GM1_INV = 1 / GM1;             // include/SolverData.hpp:131
    static constexpr double EPS = DBL_EPSILON;             // include/Numerics.hpp:171
    static constexpr double BIG = DBL_MAX;                 // include/Numerics.hpp:179,267
};
template <> struct Consts<float> {
    static constexpr float GAMMA = 1.4f;
    static constexpr float GM1 = GAMMA - 1;
    static constexpr float GM1_INV = 1 / GM1;
    static constexpr float EPS = FLT_EPSILON;
    static constexpr float BIG = FLT_MAX;
};

template <typename T> struct Vec3T {
    T x, y, z;
};

// std::max / std::min semantics, including which operand wins on ties
template <typename T> NF_HD T smax(T a, T b) { return (a < b) ? b : a; }
template <typename T> NF_HD T smin(T a, T b) { return (b < a) ? b : a; }
template <typename T> NF_HD T sabs(T a) { return std::fabs(a); }
// sign() of include/Includes.hpp:100-104
template <typename T> NF_HD int isign(T v) { return (v > T(0)) ? 1 : ((v < T(0)) ? -1 : 0); }

template <typename T> NF_HD T dot3(Vec3T<T> a, Vec3T<T> b) {
    return (a.x * b.x + a.y * b.y) + a.z * b.z;
}

// EulerEqs::pressure
template <typename T> NF_HD T pressure(const T *U) {
    return Consts<T>::GM1 * (U[4] - T(0.5) / U[0] * ((U[1] * U[1] + U[2] * U[2]) + U[3] * U[3]));
}
// EulerEqs::prim_to_cons
template <typename T> NF_HD void prim_to_cons(const T *V, T *U) {
    U[0] = V[0];
    U[1] = V[0] * V[1];
    U[2] = V[0] * V[2];
    U[3] = V[0] * V[3];
    U[4] = Consts<T>::GM1_INV * V[4] + T(0.5) * V[0] * ((V[1] * V[1] + V[2] * V[2]) + V[3] * V[3]);
}
// EulerEqs::cons_to_prim
template <typename T> NF_HD void cons_to_prim(const T *U, T *V) {
    V[0] = U[0];
    V[1] = U[1] / U[0];
    V[2] = U[2] / U[0];
    V[3] = U[3] / U[0];
    V[4] = pressure(U);
}
// EulerEqs::sound_speed_primitive
template <typename T> NF_HD T sound_speed_prim(const T *V) {
    return std::sqrt(Consts<T>::GAMMA * V[4] / V[0]);
}
// EulerEqs::entropy_primitive
template <typename T> NF_HD T entropy_prim(const T *V) {
    return V[4] * std::pow(V[0], -Consts<T>::GAMMA);
}

// Eigen normalized(): n = S / sqrt(|S|^2) when |S|^2 > 0; area = sqrt(|S|^2)
template <typename T> NF_HD void normal_area(Vec3T<T> S, Vec3T<T> &n, T &area) {
    T z = (S.x * S.x + S.y * S.y) + S.z * S.z;
    T r = std::sqrt(z);
    if (z > T(0)) n = {S.x / r, S.y / r, S.z / r};
    else n = S;
    area = r;
}

// NumericalFlux::rusanov with the EulerEqs helpers inlined
template <typename T> NF_HD void rusanov(const T *UL, const T *UR, Vec3T<T> S, T *F) {
    Vec3T<T> n;
    T area;
    normal_area(S, n, area);
    const T pL = pressure(UL), pR = pressure(UR);
    const T vnL = ((UL[1] * n.x + UL[2] * n.y) + UL[3] * n.z) / UL[0];
    const T vnR = ((UR[1] * n.x + UR[2] * n.y) + UR[3] * n.z) / UR[0];
    const T cL = std::sqrt(Consts<T>::GAMMA * pL / UL[0]);
    const T cR = std::sqrt(Consts<T>::GAMMA * pR / UR[0]);
    const T sL = sabs(vnL) + cL, sR = sabs(vnR) + cR;
    const T s = smax(sR, sL);
    const T FR[5] = {vnR * UR[0], vnR * UR[1] + pR * n.x, vnR * UR[2] + pR * n.y,
                     vnR * UR[3] + pR * n.z, vnR * (UR[4] + pR)};
    const T FL[5] = {vnL * UL[0], vnL * UL[1] + pL * n.x, vnL * UL[2] + pL * n.y,
                     vnL * UL[3] + pL * n.z, vnL * (UL[4] + pL)};
    const T h = area * T(0.5);
#pragma unroll
    for (int k = 0; k < 5; ++k) F[k] = h * ((FR[k] + FL[k]) - s * (UR[k] - UL[k]));
}

// Freestream constants of BC_FarField (src/Numerics.cpp:83-91)
template <typename T> struct FarField {
    T rho, vx, vy, vz, p, c, entropy;
};
template <typename T> inline FarField<T> make_farfield(const double *Vinf) {
    T V[5] = {T(Vinf[0]), T(Vinf[1]), T(Vinf[2]), T(Vinf[3]), T(Vinf[4])};
    FarField<T> f;
    f.rho = V[0];
    f.vx = V[1];
    f.vy = V[2];
    f.vz = V[3];
    f.p = V[4];
    f.c = std::sqrt(Consts<T>::GAMMA * V[4] / V[0]);
    f.entropy = V[4] * std::pow(V[0], -Consts<T>::GAMMA);
    return f;
}

enum : int { BC_NOSLIP = 0, BC_SLIP = 1, BC_FARFIELD = 2 };

// BoundaryCondition::calc_ghost_val for the three boundary types
template <typename T>
NF_HD void ghost_value(int bc, const T *Vd, Vec3T<T> S, const FarField<T> &fs, T *Vg) {
    if (bc == BC_SLIP) {
        Vec3T<T> n;
        T area;
        normal_area(S, n, area);
        const T vn = (Vd[1] * n.x + Vd[2] * n.y) + Vd[3] * n.z;
        Vg[0] = Vd[0];
        Vg[1] = Vd[1] - T(2) * vn * n.x;
        Vg[2] = Vd[2] - T(2) * vn * n.y;
        Vg[3] = Vd[3] - T(2) * vn * n.z;
        Vg[4] = Vd[4];
    } else if (bc == BC_NOSLIP) {
        Vg[0] = Vd[0];
        Vg[1] = -Vd[1];
        Vg[2] = -Vd[2];
        Vg[3] = -Vd[3];
        Vg[4] = Vd[4];
    } else {
        Vec3T<T> n;
        T area;
        normal_area(S, n, area);
        const Vec3T<T> vd = {Vd[1], Vd[2], Vd[3]};
        const Vec3T<T> vf = {fs.vx, fs.vy, fs.vz};
        const T vn_fs = dot3(n, vf);
        const T vn_d = dot3(n, vd);
        const T c_d = std::sqrt(Consts<T>::GAMMA * Vd[4] / Vd[0]);
        const T GM1 = Consts<T>::GM1;
        T Rp, Rm;
        if (vn_fs + fs.c > T(0)) Rp = vn_d + T(2) * c_d / GM1;
        else Rp = vn_fs + T(2) * fs.c / GM1;
        if (vn_fs - fs.c > T(0)) Rm = vn_d - T(2) * c_d / GM1;
        else Rm = vn_fs - T(2) * fs.c / GM1;
        const T vn_b = T(0.5) * (Rp + Rm);
        const T c_b = GM1 / T(4) * (Rp - Rm);
        Vec3T<T> vb;
        T s_b;
        if (vn_fs > T(0)) {
            const T d = vn_b - vn_d;
            vb = {vd.x + d * n.x, vd.y + d * n.y, vd.z + d * n.z};
            s_b = Vd[4] * std::pow(Vd[0], -Consts<T>::GAMMA);
        } else {
            const T d = vn_fs - vn_b;  // sign as in src/Numerics.cpp:146
            vb = {vf.x + d * n.x, vf.y + d * n.y, vf.z + d * n.z};
            s_b = fs.entropy;
        }
        const T rho_b = std::pow(Consts<T>::GAMMA * s_b / (c_b * c_b), Consts<T>::GM1_INV);
        const T p_b = c_b * c_b * rho_b / Consts<T>::GAMMA;
        Vg[0] = rho_b;
        Vg[1] = vb.x;
        Vg[2] = vb.y;
        Vg[3] = vb.z;
        Vg[4] = p_b;
    }
}

#ifdef __HIPCC__
// HLLC flux (Toro, "Riemann Solvers and Numerical Methods for Fluid
// Dynamics", 3rd ed., 2009, section 10.4; the reference's HLLC is an
// assert(false) stub, src/Numerics.cpp:20-23): Davis wave speeds, contact
// speed (10.37), star states (10.39), F*_K = F_K + S_K (U*_K - U_K) (10.38),
// scaled by |S|.  The oracle's hllc in the same operations and order, with
// plain (correctly rounded) divisions, so both give the same bits; UL is the
// state of the reference's cell i, UR of cell j (or the ghost), S = S_ij.
__device__ __forceinline__ double pressure_d(const double *U) {
    return Consts<double>::GM1 * (U[4] - 0.5 / U[0] * ((U[1] * U[1] + U[2] * U[2]) + U[3] * U[3]));
}
__device__ __forceinline__ void inviscid_flux_d(const double *U, Vec3T<double> n, double *F) {
    const double rho = U[0], p = pressure_d(U);
    const double vn = ((U[1] * n.x + U[2] * n.y) + U[3] * n.z) / rho;
    F[0] = vn * rho;
    F[1] = vn * U[1] + p * n.x;
    F[2] = vn * U[2] + p * n.y;
    F[3] = vn * U[3] + p * n.z;
    F[4] = vn * (U[4] + p);
}
__device__ __forceinline__ void hllc_d(const double *UL, const double *UR, Vec3T<double> S, double *Flux) {
    const double z = (S.x * S.x + S.y * S.y) + S.z * S.z;
    Vec3T<double> n = S;
    if (z > 0.0) {
        const double r = sqrt(z);
        n = {S.x / r, S.y / r, S.z / r};
    }
    const double area = sqrt(z);
    const double G = Consts<double>::GAMMA;
    const double pL = pressure_d(UL), pR = pressure_d(UR);
    const double vL = ((UL[1] * n.x + UL[2] * n.y) + UL[3] * n.z) / UL[0];
    const double vR = ((UR[1] * n.x + UR[2] * n.y) + UR[3] * n.z) / UR[0];
    const double cL = sqrt(G * pL / UL[0]), cR = sqrt(G * pR / UR[0]);
    const double aL = vL - cL, bL = vR - cR, aR = vL + cL, bR = vR + cR;
    const double sL = (bL < aL) ? bL : aL;  // std::min / std::max
    const double sR = (aR < bR) ? bR : aR;
    const double mL = UL[0] * (sL - vL), mR = UR[0] * (sR - vR);
    const double sM = ((pR - pL) + (mL * vL - mR * vR)) / (mL - mR);
    double F[5];
    if (0.0 <= sL) {
        inviscid_flux_d(UL, n, F);
    } else if (sR <= 0.0) {
        inviscid_flux_d(UR, n, F);
    } else {
        const bool left = 0.0 <= sM;
        double U[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) U[k] = left ? UL[k] : UR[k];
        const double sK = left ? sL : sR, vK = left ? vL : vR, pK = left ? pL : pR, mK = left ? mL : mR;
        double FK[5];
        inviscid_flux_d(U, n, FK);
        const double rs = mK / (sK - sM);  // star density
        const double dv = sM - vK;
        const double Us[5] = {rs, rs * (U[1] / U[0] + dv * n.x), rs * (U[2] / U[0] + dv * n.y),
                              rs * (U[3] / U[0] + dv * n.z), rs * (U[4] / U[0] + dv * (sM + pK / mK))};
#pragma unroll
        for (int k = 0; k < 5; ++k) F[k] = FK[k] + sK * (Us[k] - U[k]);
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) Flux[k] = area * F[k];
}

#endif

}  // namespace nf
