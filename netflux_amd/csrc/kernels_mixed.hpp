// Mixed-precision stage kernels: FP32 gradient, limiter, reconstruction and
// Rusanov flux on FP32 copies of the primitive and face states; FP64 state,
// flux-balance accumulation, RK combination, time step and residual norms.
//
// Same mapping, layouts and face order as the FP64 kernels
// (kernels_fp64.hpp), with FP32 V [tile][5][64], FV [tile][4][5][64] and
// ghost states.  Face geometry is formed in FP64 from the nodes and rounded
// once: S from the face's own node order, so the two cells of a face hold
// bit-identical FP32 normals, and each face flux is evaluated by both cells
// with identical FP32 arithmetic -- the FP64 flux balance conserves exactly.
// This header is compiled with floating-point contraction enabled.
#pragma once
#include "kernels_fp64.hpp"

namespace nf {
namespace mx {

// waves per SIMD the mixed sweep B is compiled for (0: the compiler's choice)
#ifndef NF_SWEEPB_MX_WAVES
#define NF_SWEEPB_MX_WAVES 0
#endif
#if NF_SWEEPB_MX_WAVES
#define NF_SWEEPB_MX_ATTR __attribute__((amdgpu_waves_per_eu(NF_SWEEPB_MX_WAVES, NF_SWEEPB_MX_WAVES)))
#else
#define NF_SWEEPB_MX_ATTR
#endif

// kernel arguments
struct GhostArgsMx {
    const uint32_t *bface;
    const uint8_t *bbc;
    const uint16_t *binst;
    uint32_t nb;
    const float *src;          // V32 [tile][5][64] (cell) or FV32 (face)
    float *out;                // [nb/64][5][64], tiled like V
    const FarField<float> *fs; // [instances]
};

struct SweepAArgsMx {
    const float *V;
    const float *Gc;
    float *FV;
    double *dt_part;
    double *dt_cell;
    double CFL;
    int xcd_remap;
    double *grad_aux, *psi_aux;
};

struct SweepBArgsMx {
    const float *FV;
    const float *Gf;
    const double *U;
    const double *Uo;
    double *Uout;
    float *Vout;
    const double *dt_cell;
    int kind;
    double *res_part;
    int check;
    int xcd_remap;
    double *R_aux;
    double *Rst, *fst;  // checked stage: R and f stored for the apply pass, no update
};

#ifdef NF_MIXED_KERNELS
using V3f = Vec3T<float>;

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ float fflip(float x, uint32_t m) { return __int_as_float(__float_as_int(x) ^ (int)m); }

__device__ __forceinline__ V3f to_f(V3d v) { return {(float)v.x, (float)v.y, (float)v.z}; }

// S (FP32, from FP64 edge vectors of the face's own node order) and, if r is
// given, the FP32 centroid-to-face vector
__device__ __forceinline__ void face_geo_mx(const NodeView &nv, const CellGeo &g, uint32_t byte, V3f &S, V3f *r) {
    const V3d a = nv_node(nv, byte & 3u), b = nv_node(nv, (byte >> 2) & 3u), c = nv_node(nv, (byte >> 4) & 3u);
    const V3f e1 = to_f(vsub(b, a)), e2 = to_f(vsub(c, a));
    S = {0.5f * (e1.y * e2.z - e1.z * e2.y), 0.5f * (e1.z * e2.x - e1.x * e2.z), 0.5f * (e1.x * e2.y - e1.y * e2.x)};
    if (byte & 0x40u) S = {-S.x, -S.y, -S.z};
    if (r) {
        const V3d s = vadd(vadd(a, b), c);
        *r = to_f(vsub({s.x / 3, s.y / 3, s.z / 3}, g.cen));
    }
}

__device__ __forceinline__ void prim_to_cons_f(const float *V, float *U) {
    U[0] = V[0];
    U[1] = V[0] * V[1];
    U[2] = V[0] * V[2];
    U[3] = V[0] * V[3];
    U[4] = Consts<float>::GM1_INV * V[4] + 0.5f * V[0] * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3]);
}

// Rusanov flux seen from one cell (see rusanov_side): F for UL = own if the
// cell is the owner (mneg = 0), UL = neighbour otherwise
__device__ __forceinline__ void rusanov_side_f(const float *UO, const float *UN, V3f S, uint32_t mneg, float *F) {
    // No contraction here: the two cells of a face evaluate this with UO and
    // UN swapped, and a fused FO + FN would round the owner's product on one
    // side and the neighbour's on the other.  Unfused, every operation is
    // either symmetric in (UO, UN) or exactly negated, so both cells get the
    // same flux bit for bit and interior faces conserve exactly.
#pragma clang fp contract(off)
    const float z = S.x * S.x + S.y * S.y + S.z * S.z;
    const float area = fsqrt(z);
    const float ia = frcp(area);
    const V3f n = {S.x * ia, S.y * ia, S.z * ia};
    const float iO = frcp(UO[0]), iN = frcp(UN[0]);
    const float G = Consts<float>::GAMMA, GM1 = Consts<float>::GM1;
    const float pO = GM1 * (UO[4] - 0.5f * iO * (UO[1] * UO[1] + UO[2] * UO[2] + UO[3] * UO[3]));
    const float pN = GM1 * (UN[4] - 0.5f * iN * (UN[1] * UN[1] + UN[2] * UN[2] + UN[3] * UN[3]));
    const float vnO = (UO[1] * n.x + UO[2] * n.y + UO[3] * n.z) * iO;
    const float vnN = (UN[1] * n.x + UN[2] * n.y + UN[3] * n.z) * iN;
    const float sm = fmaxf(fabsf(vnO) + fsqrt(G * pO * iO), fabsf(vnN) + fsqrt(G * pN * iN));
    const float FO[5] = {vnO * UO[0], vnO * UO[1] + pO * n.x, vnO * UO[2] + pO * n.y, vnO * UO[3] + pO * n.z,
                         vnO * (UO[4] + pO)};
    const float FN[5] = {vnN * UN[0], vnN * UN[1] + pN * n.x, vnN * UN[2] + pN * n.y, vnN * UN[3] + pN * n.z,
                         vnN * (UN[4] + pN)};
    const float h = 0.5f * area;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const float d = sm * (UO[k] - UN[k]);
        const float c = FO[k] + FN[k];
        F[k] = h * (c + fflip(d, mneg));
    }
}

// Barth limiter of one variable over the four faces with one division per
// sign group of Delta_2 (the FP64 kernels' barth_psi_d in FP32: the face
// with the largest |Delta_2| of a group gives the group's smallest quotient,
// and correctly rounded division is monotone in the divisor); a division
// whose quotient is at least 1 is skipped, as in barth_psi_d
__device__ __forceinline__ float barth_psi_f(const float *delta, float vmax, float vmin, float v) {
    float hi = 0.0f, lo = 0.0f;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        hi = fmaxf(hi, delta[s]);
        lo = fminf(lo, delta[s]);
    }
    const float ap = vmax - v, bp = hi + FLT_EPSILON, am = vmin - v, bm = lo - FLT_EPSILON;
    float pp = 1.0f, pm = 1.0f;
    if (hi > 0.0f && ap < bp) pp = fminf(1.0f, ap / bp);
    if (lo < 0.0f && am > bm) pm = fminf(1.0f, am / bm);
    return fminf(pp, pm);
}

// ---------------------------------------------------------------------------

template <bool FACE>
__global__ __launch_bounds__(BLOCK) void k_ghost_mx(DevMesh m, GhostArgsMx a) {
    __shared__ double lds[12 * BLOCK];
    const uint32_t b0 = blockIdx.x * BLOCK + threadIdx.x;
    const bool live = b0 < a.nb;
    const uint32_t b = live ? b0 : a.nb - 1;
    const uint32_t inst = a.binst[b];
    const uint32_t w = a.bface[b], c = w & 0x1FFFFFFFu, s = w >> 29;
    CellGeo g;
    load_cell(m, c, g);
    const NodeView nv = stage_nodes<BLOCK>(lds, g);
    V3f S;
    face_geo_mx(nv, g, (m.fcode[c] >> (8 * s)) & 0xFFu, S, nullptr);
    float Vd[5], Vg[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) Vd[k] = FACE ? a.src[fix(c, s, k)] : a.src[tix(c, k)];
    ghost_value<float>((int)a.bbc[b], Vd, S, a.fs[inst], Vg);
    if (live && !m.st[inst].stop) {
#pragma unroll
        for (int k = 0; k < 5; ++k) a.out[tix(b, k)] = Vg[k];
    }
}

template <int ORDER, int B>
__global__ __launch_bounds__(B) void k_sweepA_mx(DevMesh m, SweepAArgsMx a) {
    __shared__ double sh[B / 64];
    __shared__ double lds[12 * B];
    __shared__ float ldsv[5 * B];  // [wave][5][64] own states
    const LaneCell L = lane_cell<B>(m, a.xcd_remap);
    if (m.st[L.inst].stop) return;
    const uint32_t c = L.c;
    const bool live = L.live;
    const uint32_t p = L.lb * B + threadIdx.x;  // device position (padding included)

    const uint32_t code = m.fcode[c];
    const uint4 w4 = *reinterpret_cast<const uint4 *>(m.nbr + 4 * (size_t)c);
    const uint32_t w[4] = {w4.x, w4.y, w4.z, w4.w};
    const NodeTableLoad ntl = node_table_issue(m, p);
    __builtin_amdgcn_sched_barrier(0);
    float Vc[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) Vc[k] = a.V[tix(c, k)];
    // neighbours in the same block come from an LDS copy of the block's
    // states; the gathers are exec-masked to the other lanes (as in the FP64
    // sweep)
    float Vn[4][5];
    bool loc[4];
#pragma unroll
    for (int s = 0; s < 4; ++s)
        loc[s] = !(w[s] & 0x80000000u) && ((w[s] & 0x1FFFFFFFu) / B) == (p / B);
    if (ORDER != 0) {
#pragma unroll
        for (int s = 0; s < 4; ++s)
            if (!loc[s]) {
                const uint32_t idx = w[s] & 0x1FFFFFFFu;
                const float *src = ((w[s] & 0x80000000u) ? a.Gc : a.V) + tix(idx, 0);
#pragma unroll
                for (int k = 0; k < 5; ++k)
                    Vn[s][k] = *(const __attribute__((address_space(1))) float *)(src + k * TILE);
            }
    }
    float *vt = ldsv + (threadIdx.x >> 6) * 5 * TILE;
    if (ORDER != 0) {
#pragma unroll
        for (int k = 0; k < 5; ++k) vt[k * TILE + (p & 63u)] = Vc[k];
    }
    CellGeo g;
    // (its wave-level fences also order the state writes above before the
    // reads below)
    const NodeView nv = node_table_finish<B>(m, p, ntl, lds, g);
    if (ORDER != 0) __syncthreads();  // every wave's states are in LDS
    cell_geo(m, g);
    V3f S[4], r[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) face_geo_mx(nv, g, (code >> (8 * s)) & 0xFFu, S[s], &r[s]);
    const float ivol = (float)(1.0 / g.vol);
    __builtin_amdgcn_sched_barrier(0);
    if (ORDER != 0) {
#pragma unroll
        for (int s = 0; s < 4; ++s)
            if (loc[s]) {
                const uint32_t idx = w[s] & 0x1FFFFFFFu, j = idx & 63u;
                const float *src = ldsv + ((idx >> 6) - (p & ~(uint32_t)(B - 1)) / TILE) * 5 * TILE;
#pragma unroll
                for (int k = 0; k < 5; ++k) Vn[s][k] = src[k * TILE + j];
            }
    }

    if (ORDER == 0) {
        if (live) {
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int k = 0; k < 5; ++k) a.FV[fix(c, s, k)] = Vc[k];
        }
    } else {
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            float g0 = 0.f, g1 = 0.f, g2 = 0.f, vmax = Vc[k], vmin = Vc[k];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const uint32_t mneg = ((code >> (8 * s + 7)) & 1u) ? 0u : 0x80000000u;
                const float uf = fflip(0.5f * (Vc[k] + Vn[s][k]) * ivol, mneg);
                g0 += uf * S[s].x;
                g1 += uf * S[s].y;
                g2 += uf * S[s].z;
                if (ORDER == 2) {
                    vmax = fmaxf(vmax, Vn[s][k]);
                    vmin = fminf(vmin, Vn[s][k]);
                }
            }
            float delta[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) delta[s] = g0 * r[s].x + g1 * r[s].y + g2 * r[s].z;
            const float psi = (ORDER == 2) ? barth_psi_f(delta, vmax, vmin, Vc[k]) : 1.0f;
            if (live) {
#pragma unroll
                for (int s = 0; s < 4; ++s) a.FV[fix(c, s, k)] = Vc[k] + psi * delta[s];
                if (a.grad_aux) {
                    a.grad_aux[15 * (size_t)c + k] = g0;
                    a.grad_aux[15 * (size_t)c + k + 5] = g1;
                    a.grad_aux[15 * (size_t)c + k + 10] = g2;
                    a.psi_aux[5 * (size_t)c + k] = psi;
                }
            }
        }
    }
    if (a.dt_part) {
        const V3d Sd[4] = {{S[0].x, S[0].y, S[0].z}, {S[1].x, S[1].y, S[1].z}, {S[2].x, S[2].y, S[2].z},
                           {S[3].x, S[3].y, S[3].z}};
        const double Vd[5] = {Vc[0], Vc[1], Vc[2], Vc[3], Vc[4]};
        const double t = cell_dt(Vd, Sd, w, g.vol, a.CFL);
        const double dtc = (live && t < DBL_MAX) ? t : DBL_MAX;
        if (a.dt_cell && live) a.dt_cell[c] = t;
        const double v = block_reduce<true, B>(dtc, sh);
        if (threadIdx.x == 0) a.dt_part[L.lb] = v;
    }
}

template <int B>
__global__ __launch_bounds__(B) NF_SWEEPB_MX_ATTR void k_sweepB_mx(DevMesh m, SweepBArgsMx a) {
    __shared__ double sh[B / 64];
    __shared__ double lds[12 * B];
    const LaneCell L = lane_cell<B>(m, a.xcd_remap);
    InstState &st = m.st[L.inst];
    if (st.stop) return;
    const uint32_t c = L.c;
    const bool live = L.live;

    const uint32_t code = m.fcode[c];
    const uint4 w4 = *reinterpret_cast<const uint4 *>(m.nbr + 4 * (size_t)c);
    const uint32_t w[4] = {w4.x, w4.y, w4.z, w4.w};
    const NodeTableLoad ntl = node_table_issue(m, L.lb * B + threadIdx.x);
    __builtin_amdgcn_sched_barrier(0);
    // the stage state first, outside the stage-kind branches (stage 1 and a
    // zero U_old read U twice: same lines, no branch)
    const bool zero_old = st.uold_zero;
    double Up[5], Uop[5];
    {
        const double *uo = (a.kind == 1 || zero_old) ? a.U : a.Uo;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            Up[k] = a.U[tix(c, k)];
            Uop[k] = uo[tix(c, k)];
        }
    }
    // face values of both sides, one face at a time, the next face's loads
    // issued before the current face is computed (as the FP64 sweep)
    float Vo4[4][5], Vn4[4][5];
    auto load_face = [&](int s) {
        const uint32_t idx = w[s] & 0x1FFFFFFFu, sn = (w[s] >> 29) & 3u;
        const bool bnd = w[s] & 0x80000000u;
        const float *src = bnd ? a.Gf + tix(idx, 0) : a.FV + fix(idx, sn, 0);
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            Vo4[s][k] = a.FV[fix(c, s, k)];
            Vn4[s][k] = src[k * TILE];
        }
    };
    load_face(0);
    CellGeo g;
    const NodeView nv = node_table_finish<B>(m, L.lb * B + threadIdx.x, ntl, lds, g);
    cell_geo(m, g);
    double R[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (s + 1 < 4) load_face(s + 1);
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t byte = (code >> (8 * s)) & 0xFFu;
        V3f S;
        face_geo_mx(nv, g, byte, S, nullptr);
        float UO[5], UN[5], F[5];
        prim_to_cons_f(Vo4[s], UO);
        prim_to_cons_f(Vn4[s], UN);
        const uint32_t mneg = (byte & 0x80u) ? 0u : 0x80000000u;
        rusanov_side_f(UO, UN, S, mneg, F);
#pragma unroll
        for (int k = 0; k < 5; ++k) R[k] += (double)fflip(F[k], mneg ^ 0x80000000u);
        // the face's flux is complete here: without this the compiler sinks
        // every face's arithmetic below the last face's loads, and all four
        // faces' values are live at once
        asm volatile("" : "+v"(R[0]), "+v"(R[1]), "+v"(R[2]), "+v"(R[3]), "+v"(R[4]) : : "memory");
    }
    if (live && a.check) {
        bool ok = true;
#pragma unroll
        for (int k = 0; k < 5; ++k) ok = ok && (R[k] - R[k] == 0.0);
        if (!ok) atomicAdd(&st.bad_flux, 1);
    }
    if (a.R_aux) {
        if (live) {
#pragma unroll
            for (int k = 0; k < 5; ++k) a.R_aux[5 * (size_t)c + k] = R[k];
        }
    } else {
        const double dt = a.dt_cell ? a.dt_cell[c] : st.dt;
        const double f = (a.kind == 1 ? dt : (a.kind == 2 ? 1.0 / 4.0 * dt : 2.0 / 3.0 * dt)) / g.vol;
        if (a.Rst) {
            if (live) {
#pragma unroll
                for (int k = 0; k < 5; ++k) a.Rst[tix(c, k)] = R[k];
                a.fst[c] = f;
            }
        } else {
            double Un[5], Vnew[5];
            stage_combine(a.kind, zero_old, f, Up, Uop, R, Un);
            cons_to_prim_d(Un, Vnew);
            if (live) {
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    a.Uout[tix(c, k)] = Un[k];
                    a.Vout[tix(c, k)] = (float)Vnew[k];
                }
                if (!(Vnew[0] > 0.0) || !(Vnew[4] > 0.0) || !(Un[4] > 0.0)) atomicAdd(&st.bad_state, 1);
            }
        }
    }
    if (a.res_part) {
        const double r2 = live ? R[0] * R[0] : 0.0;
        const double v = block_reduce<false, B>(r2, sh);
        if (threadIdx.x == 0) a.res_part[L.lb] = v;
    }
}

__global__ __launch_bounds__(BLOCK) void k_apply_mx(DevMesh m, ApplyArgs a) { apply_stage_cell<float>(m, a); }

__global__ __launch_bounds__(BLOCK) void k_prim_mx(const double *U, float *V, uint32_t N) {
    const uint32_t c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= N) return;
    double u[5], v[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) u[k] = U[tix(c, k)];
    cons_to_prim_d(u, v);
#pragma unroll
    for (int k = 0; k < 5; ++k) V[tix(c, k)] = (float)v[k];
}

#endif  // NF_MIXED_KERNELS

// host-side launchers (compiled in solver_mixed.hip)
hipError_t launch_ghost_mx(bool face, dim3 g, hipStream_t st, const DevMesh &m, const GhostArgsMx &a);
hipError_t launch_sweepA_mx(int order, int block, dim3 g, hipStream_t st, const DevMesh &m, const SweepAArgsMx &a);
hipError_t launch_sweepB_mx(int block, dim3 g, hipStream_t st, const DevMesh &m, const SweepBArgsMx &a);
hipError_t launch_prim_mx(dim3 g, hipStream_t st, const double *U, float *V, uint32_t N);
hipError_t launch_apply_mx(dim3 g, hipStream_t st, const DevMesh &m, const ApplyArgs &a);

}  // namespace mx
}  // namespace nf
