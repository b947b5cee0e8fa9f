// FP64 stage kernels of the explicit Euler step.
//
// Mapping: one tetrahedron per lane.  The lane visits its four faces in slot
// order -- slot s is the cell's s-th face in increasing reference face index
// -- so every per-cell accumulation (Green-Gauss gradient, flux balance,
// Delta S) adds the same terms in the same order as the reference's
// face-ordered scatter loops, and FP64 results are identical to a serial
// evaluation of the reference formulas, with no atomics and no colouring.
// Geometry (volume, centroid, face normals, centroid-to-face vectors) is
// recomputed from the four nodes with the reference formulas instead of
// being streamed from memory.
//
//   sweep A (per stage): V -> Green-Gauss gradient, neighbour min/max, Barth
//            limiter, limited face values of the four faces
//            (+ Delta S and the CFL time step when requested).
//   sweep B (per stage): face values -> Rusanov fluxes of the four faces,
//            flux balance, RK stage combination, prim(U) of the new state.
//   Boundary faces read their ghost state from small per-boundary-face
//   buffers filled by k_ghost<false> (ghost of the cell-centre state, for the
//   gradient and min/max) and k_ghost<true> (ghost of the reconstructed face
//   state, for the boundary flux), so the sweeps carry no boundary code.
//
// Several meshes (instances) can share one launch: their cells are
// concatenated, each instance padded to whole blocks of 256 cells, and every
// block looks up its instance for the time step, the first-step flag, the
// stop flag and the error counters (InstState).
//
// Layouts: cell fields are tiled by 64 cells (one wavefront):
//   V, U  : [tile][5][64]          FV : [tile][4 slots][5][64]
// so a wave's own-cell loads and stores are fully coalesced; cnodes and nbr
// are [N][4] u32, fcode [N], nodes [n_nodes][4] (x, y, z, pad).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fp64_math.hpp"
#include "physics.hpp"

namespace nf {

// sweep A reads neighbours in the lane's own 64-cell tile (1) or in its
// block (2, with block barriers) from LDS
#ifndef NF_A_LDS
#define NF_A_LDS 2
#endif
// sweep A writes its face values with 16-byte stores (needs NF_A_LDS)
#ifndef NF_A_STX4
#define NF_A_STX4 1
#endif
// waves per SIMD the sweeps are compiled for (0: the compiler's choice)
#ifndef NF_SWEEP_WAVES
#define NF_SWEEP_WAVES 0
#endif
#if NF_SWEEP_WAVES
#define NF_SWEEP_ATTR __attribute__((amdgpu_waves_per_eu(NF_SWEEP_WAVES, NF_SWEEP_WAVES)))
#else
#define NF_SWEEP_ATTR
#endif
#ifndef NF_SWEEPB_WAVES
#define NF_SWEEPB_WAVES 0
#endif
#if NF_SWEEPB_WAVES
#define NF_SWEEPB_ATTR __attribute__((amdgpu_waves_per_eu(NF_SWEEPB_WAVES, NF_SWEEPB_WAVES)))
#else
#define NF_SWEEPB_ATTR NF_SWEEP_ATTR
#endif

constexpr int BLOCK = 256;  // threads of the auxiliary kernels; instance padding unit
constexpr int TILE = 64;    // cells per layout tile

// load through a global-address-space pointer (keeps the compiler from
// falling back to flat loads, which also count against the LDS counter)
__device__ __forceinline__ double ldg(const double *p) {
    return *(const __attribute__((address_space(1))) double *)p;
}

__device__ __forceinline__ size_t tix(uint32_t c, int k) {  // [tile][5][64]
    return ((size_t)(c >> 6) * 5 + k) * TILE + (c & 63u);
}
__device__ __forceinline__ size_t fix(uint32_t c, uint32_t s, int k) {  // [tile][4][5][64]
    return (((size_t)(c >> 6) * 4 + s) * 5 + k) * TILE + (c & 63u);
}

// Per-instance run state, one record per mesh of the context.
struct InstState {
    double dt;         // time step of the current step
    double time;       // sum of completed steps' dt
    double rref;       // residual the drop is measured against (0: the first step's)
    int nstep;         // completed steps
    int err;           // NF_E_INVALID_DT once raised
    int stop;          // nonzero: kernels skip this instance
    int converged;     // residual drop reached
    int bad_flux;      // cells with non-finite flux balance (checked runs)
    int bad_state;     // cells with rho <= 0, p <= 0 or NaN after an update
    int uold_zero;     // U_old is all zeros (before the first step)
    int pad;
};

struct DevMesh {
    double three, six;       // 3.0 and 6.0, divisors of the centroid / volume formulas
    const double *nodes;
    const uint32_t *cnodes;
    // per 64-cell group: its distinct nodes' coordinates, SoA, entries
    // [grp_off[g], grp_off[g+1]); cell_ln: the group-local entry of each
    // cell's four nodes (4 x u16)
    const uint32_t *grp_off;
    const double *gx, *gy, *gz;
    const uint2 *cell_ln;
    const uint32_t *nbr;
    const uint32_t *fcode;
    const uint32_t *blk_inst;  // [blocks] instance of each 256-cell block (32-bit: scalar loads)
    const uint32_t *inst_end;  // [instances] one past the instance's last device cell
    InstState *st;             // [instances]
};

// The lane's cell, its instance, and whether it is a real cell.  Dead lanes
// of the last block of an instance mirror the instance's last cell and store
// nothing.
struct LaneCell {
    uint32_t c, inst, lb;  // cell, instance, logical block
    bool live;
};
template <int B>
__device__ __forceinline__ LaneCell lane_cell(const DevMesh &m, int xcd_remap) {
    const uint32_t lb = logical_block(blockIdx.x, gridDim.x, xcd_remap);
    const uint32_t c0 = lb * B + threadIdx.x;
    const uint32_t inst = m.blk_inst[(lb * B) / BLOCK];
    const uint32_t end = m.inst_end[inst];
    return {c0 < end ? c0 : end - 1, inst, lb, c0 < end};
}

struct CellGeo {
    V3d p0, p1, p2, p3;
    V3d cen;
    double vol;
    Rcp ivol;
};

// Node loads of cell c from the node array (boundary-face kernels).
__device__ __forceinline__ void load_nodes(const DevMesh &m, uint32_t c, CellGeo &g) {
    const uint4 q = *reinterpret_cast<const uint4 *>(m.cnodes + 4 * (size_t)c);
    auto node = [&](uint32_t n) {
        const double2 a = *reinterpret_cast<const double2 *>(m.nodes + 4 * (size_t)n);
        return V3d{a.x, a.y, m.nodes[4 * (size_t)n + 2]};
    };
    g.p0 = node(q.x);
    g.p1 = node(q.y);
    g.p2 = node(q.z);
    g.p3 = node(q.w);
}

// volume, centroid and 1/volume of loaded nodes
__device__ __forceinline__ void cell_geo(const DevMesh &m, CellGeo &g) {
    // tetrahedron_calc_volume / _centroid (src/geometry/GeometryUtilities.cpp:289-316)
    const V3d ab = vsub(g.p1, g.p0), ac = vsub(g.p2, g.p0), ad = vsub(g.p3, g.p0);
    g.vol = rdiv(dot3(vcross(ab, ac), ad), rcp_of(m.six));
    g.cen = vscl(1.0 / 4, vadd(vadd(vadd(g.p0, g.p1), g.p2), g.p3));
    g.ivol = rcp_of(g.vol);
}

__device__ __forceinline__ void load_cell(const DevMesh &m, uint32_t c, CellGeo &g) {
    load_nodes(m, c, g);
    cell_geo(m, g);
}

// Node coordinates in LDS, rows x, y, z of `stride` entries each; l[q] is
// the entry of the lane's q-th cell node.  Selecting a face's nodes by index
// from LDS replaces register selects.
struct NodeView {
    const double *lds;
    uint32_t l[4];
    uint32_t stride;
};
__device__ __forceinline__ V3d nv_node(const NodeView &v, uint32_t q) {
    const uint32_t i = q == 0 ? v.l[0] : (q == 1 ? v.l[1] : (q == 2 ? v.l[2] : v.l[3]));
    return {v.lds[i], v.lds[v.stride + i], v.lds[2 * v.stride + i]};
}

// Per-lane staging (boundary-face kernels): lane t writes its four nodes at
// entries q*B + t; only its own entries are read, so no barrier is needed.
template <int B>
__device__ __forceinline__ NodeView stage_nodes(double *lds, const CellGeo &g) {
    const uint32_t t = threadIdx.x;
    const V3d P[4] = {g.p0, g.p1, g.p2, g.p3};
    NodeView v{lds, {t, B + t, 2 * B + t, 3 * B + t}, 4 * B};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        lds[q * B + t] = P[q].x;
        lds[4 * B + q * B + t] = P[q].y;
        lds[8 * B + q * B + t] = P[q].z;
    }
    return v;
}

// Per-wave node table (sweeps): the wave's 64 cells (one group) load their
// distinct nodes once, coalesced, into the wave's LDS slice (at most 256
// entries); each lane then takes its cell's nodes from there.  The loads are
// issued here, before the sweep's state gathers, so that waiting for them
// does not wait for the gathers (memory counters retire in order).  p is the
// lane's device cell position (padding included).  Only lanes of the same
// wave exchange data through the slice, so no block barrier is needed.
struct NodeTableLoad {
    double x, y, z;
    uint32_t n;
    uint2 q;
};
__device__ __forceinline__ NodeTableLoad node_table_issue(const DevMesh &m, uint32_t p) {
    const uint32_t gi = p >> 6, lane = p & 63u;
    const uint32_t o0 = m.grp_off[gi], n = m.grp_off[gi + 1] - o0;
    NodeTableLoad t;
    t.n = n;
    t.q = m.cell_ln[p];
    // up to 256 nodes: four per lane at most; the first 64 here, the rest
    // (rare: only very irregular orderings) in node_table_finish
    const uint32_t i = lane < n ? lane : 0;
    t.x = m.gx[o0 + i];
    t.y = m.gy[o0 + i];
    t.z = m.gz[o0 + i];
    return t;
}
template <int B>
__device__ __forceinline__ NodeView node_table_finish(const DevMesh &m, uint32_t p, const NodeTableLoad &t,
                                                      double *lds, CellGeo &g) {
    const uint32_t lane = p & 63u, wv = threadIdx.x >> 6;
    double *w = lds + wv * 768;  // [3][256] per wave
    if (lane < t.n) {
        w[lane] = t.x;
        w[256 + lane] = t.y;
        w[512 + lane] = t.z;
    }
    if (t.n > 64) {  // uniform per wave
        const uint32_t o0 = m.grp_off[p >> 6];
        for (uint32_t i = lane + 64; i < t.n; i += 64) {
            w[i] = m.gx[o0 + i];
            w[256 + i] = m.gy[o0 + i];
            w[512 + i] = m.gz[o0 + i];
        }
    }
    // LDS operations of one wave complete in order; the fences keep the
    // compiler from moving the reads above the writes
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    NodeView v{w, {t.q.x & 0xFFFFu, t.q.x >> 16, t.q.y & 0xFFFFu, t.q.y >> 16}, 256};
    g.p0 = nv_node(v, 0);
    g.p1 = nv_node(v, 1);
    g.p2 = nv_node(v, 2);
    g.p3 = nv_node(v, 3);
    return v;
}

// Face normal oriented i->j (triangle formula of src/geometry/
// GeometryUtilities.cpp:318-324, flip of src/geometry/FV_Grid.cpp:247-250)
// and centroid-to-face vector (triangle_calc_centroid, :326-329) from the
// face element's node order.
__device__ __forceinline__ void face_geo(const DevMesh &m, const CellGeo &g, const NodeView &nv, uint32_t byte,
                                         V3d &S, V3d *r) {
    const V3d a = nv_node(nv, byte & 3u), b = nv_node(nv, (byte >> 2) & 3u), c = nv_node(nv, (byte >> 4) & 3u);
    S = vscl(0.5, vcross(vsub(b, a), vsub(c, a)));
    if (byte & 0x40u) S = vscl(-1, S);
    if (r) {
        const V3d s = vadd(vadd(a, b), c);
        const Rcp i3 = rcp_of(m.three);
        const V3d fc = {rdiv(s.x, i3), rdiv(s.y, i3), rdiv(s.z, i3)};
        *r = vsub(fc, g.cen);
    }
}

// wave + block min / sum over all threads; result valid in thread 0.  The min
// skips NaN inputs, as the reference's running std::min(delta_time, x) does
// (x = NaN compares false and leaves delta_time), whatever lane holds them.
template <bool MIN, int B>
__device__ __forceinline__ double block_reduce(double v, double *sh) {
    if (MIN && !(v == v)) v = DBL_MAX;
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
        for (int q = 1; q < B / 64; ++q) v = MIN ? ((sh[q] < v) ? sh[q] : v) : v + sh[q];
    }
    return v;
}

// CFL step of a cell (src/Solver.cpp:254-266) with Delta S over its interior
// faces (src/Solver.cpp:287-297), in slot order
__device__ __forceinline__ double cell_dt(const double *Vc, const V3d *S, const uint32_t *w, double vol, double CFL) {
    double d0 = 0.0, d1 = 0.0, d2 = 0.0;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (!(w[s] & 0x80000000u)) {
            d0 += 0.5 * fabs(S[s].x);
            d1 += 0.5 * fabs(S[s].y);
            d2 += 0.5 * fabs(S[s].z);
        }
    }
    const double cs = sound_speed_prim(Vc);
    const double sx = (fabs(Vc[1]) + cs) * d0;
    const double sy = (fabs(Vc[2]) + cs) * d1;
    const double sz = (fabs(Vc[3]) + cs) * d2;
    return CFL * vol / ((sx + sy) + sz);
}

// Stage combination of Solver::TVD_RK3 / explicit_euler (src/Solver.cpp:
// 98-132) with f = (stage weight x dt) / volume: kind 1 U + f R, kind 2
// 3/4 U_old + 1/4 U + f R, kind 3 1/3 U_old + 2/3 U + f R (U_old = 0 on the
// first step of a zero-U_old run).  Shared by the fused sweeps and the checked
// stage update, so both give the same bits in a translation unit.
__device__ __forceinline__ void stage_combine(int kind, bool zero_old, double f, const double *Up,
                                              const double *Uop, const double *R, double *Un) {
    if (kind == 1) {
#pragma unroll
        for (int k = 0; k < 5; ++k) Un[k] = Up[k] + f * R[k];
    } else if (kind == 2) {
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const double uo = zero_old ? 0.0 : Uop[k];
            Un[k] = 3.0 / 4.0 * uo + 1.0 / 4.0 * Up[k] + f * R[k];
        }
    } else {
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const double uo = zero_old ? 0.0 : Uop[k];
            Un[k] = 1.0 / 3.0 * uo + 2.0 / 3.0 * Up[k] + f * R[k];
        }
    }
}

// Checked stage (check_physical_validity): the sweep stores R [tile][5][64]
// and f [cell] instead of updating; this pass updates only if no cell of the
// instance had a non-finite flux balance.  Otherwise the instance stops with
// U untouched and bad_flux holding that one evaluation's count, as the
// reference throws at the end of evaluate_flux_balance before any update
// (src/Solver.cpp:80, src/SolverData.cpp:66-79).  One lane per device cell,
// BLOCK-cell blocks (the instance padding unit).
struct ApplyArgs {
    const double *U, *Uo;
    double *Uout;
    void *Vout;  // prim(Uout): double (FP64 path, may be null) or float (mixed path)
    const double *Rst, *fst;
    int kind;
};

template <typename VT>
__device__ __forceinline__ void apply_stage_cell(const DevMesh &m, const ApplyArgs &a) {
    const uint32_t inst = m.blk_inst[blockIdx.x];
    InstState &st = m.st[inst];
    if (st.stop) return;
    if (st.bad_flux > 0) {
        if (threadIdx.x == 0) st.stop = 1;
        return;
    }
    const uint32_t c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= m.inst_end[inst]) return;
    const bool zero_old = st.uold_zero;
    const double *uo = (a.kind == 1 || zero_old) ? a.U : a.Uo;
    double Up[5], Uop[5], R[5], Un[5], Vnew[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        Up[k] = a.U[tix(c, k)];
        Uop[k] = uo[tix(c, k)];
        R[k] = a.Rst[tix(c, k)];
    }
    stage_combine(a.kind, zero_old, a.fst[c], Up, Uop, R, Un);
    cons_to_prim_d(Un, Vnew);
#pragma unroll
    for (int k = 0; k < 5; ++k) a.Uout[tix(c, k)] = Un[k];
    if (a.Vout) {
#pragma unroll
        for (int k = 0; k < 5; ++k) static_cast<VT *>(a.Vout)[tix(c, k)] = (VT)Vnew[k];
    }
    if (!(Vnew[0] > 0.0) || !(Vnew[4] > 0.0) || !(Un[4] > 0.0)) atomicAdd(&st.bad_state, 1);
}

#ifndef NF_DEVICE_HELPERS_ONLY
// ---------------------------------------------------------------------------
// boundary ghost states.  bface[b] = device cell | slot << 29.
struct GhostArgs {
    const uint32_t *bface;
    const uint8_t *bbc;          // boundary type
    const uint16_t *binst;       // instance
    uint32_t nb;                 // boundary faces
    const double *src;           // V or U [tile][5][64] (cell) or FV [tile][4][5][64] (face)
    double *out;                 // [nb/64][5][64], tiled like V
    const FarField<double> *fs;  // [instances]
    int src_is_u;                // cell ghosts: src is U, the primvars are prim(U)
    const double *bS;            // [3][nb] S_ij of each boundary face (cell -> ghost)
};

// FACE = false: ghost of the cell-centre state (set_constant_ghost_values,
// src/Solver.cpp:300-330); FACE = true: ghost of the reconstructed face state
// (src/Solver.cpp:210-212)
// The face vector comes precomputed from the host face graph: the same
// cross product and sign the sweeps evaluate from the nodes, so the same bits.
template <bool FACE>
__global__ __launch_bounds__(BLOCK) void k_ghost(DevMesh m, GhostArgs a) {
    const uint32_t b0 = blockIdx.x * BLOCK + threadIdx.x;
    const bool live = b0 < a.nb;
    const uint32_t b = live ? b0 : a.nb - 1;
    const uint32_t inst = a.binst[b];
    const uint32_t w = a.bface[b], c = w & 0x1FFFFFFFu, s = w >> 29;
    const V3d S = {a.bS[b], a.bS[a.nb + b], a.bS[2 * (size_t)a.nb + b]};
    double Vd[5], Vg[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) Vd[k] = FACE ? a.src[fix(c, s, k)] : a.src[tix(c, k)];
    if (!FACE && a.src_is_u) {
        const double Ud[5] = {Vd[0], Vd[1], Vd[2], Vd[3], Vd[4]};
        cons_to_prim_d(Ud, Vd);
    }
    ghost_d((int)a.bbc[b], Vd, S, a.fs[inst], Vg);
    if (live && !m.st[inst].stop) {
#pragma unroll
        for (int k = 0; k < 5; ++k) a.out[tix(b, k)] = Vg[k];
    }
}

// ---------------------------------------------------------------------------
struct SweepAArgs {
    const double *V;   // [tile][5][64] primvars (FROM_U = false) or the state U (FROM_U)
    const double *Gc;  // [nb/64][5][64] ghost of the cell-centre state
    double *FV;        // [tile][4][5][64]
    double *dt_part;   // per-block CFL dt minimum, or null
    double *dt_cell;   // [N] per-cell CFL step (local time stepping), or null
    double CFL;
    int xcd_remap;
    double *grad_aux;  // [N][15] column-major 5x3 per cell, or null
    double *psi_aux;   // [N][5] or null
};

// One 64-cell tile of sweep A: its loads (tileA_issue) and the rest
// (tileA_run).
//
// ORDER: 0 first order, 1 second order without limiter, 2 second order Barth.
// FROM_U: a.V holds the stage's state U and the sweep evaluates the
// primvars itself (cons_to_prim of the own cell and of the neighbours it
// gathers; neighbours in the tile come from the LDS copy of the own values),
// so that sweep B need not store them.
struct TileA {
    uint32_t c, p;  // cell (clamped for padding lanes), device position
    bool live;
    uint32_t code, w[4];
    NodeTableLoad ntl;
    double Vc[5], Vn[4][5];
    const double *nsrc[4];
    bool loc[4];
};

template <int ORDER, int B, bool FROM_U>
__device__ __forceinline__ void tileA_issue(const DevMesh &m, const SweepAArgs &a, TileA &t) {
    const uint32_t c = t.c, p = t.p;
    // all loads first: own state, neighbour states (or ghost states), nodes
    t.code = m.fcode[c];
    const uint4 w4 = *reinterpret_cast<const uint4 *>(m.nbr + 4 * (size_t)c);
    t.w[0] = w4.x, t.w[1] = w4.y, t.w[2] = w4.z, t.w[3] = w4.w;
    t.ntl = node_table_issue(m, p);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < 5; ++k) t.Vc[k] = a.V[tix(c, k)];  // (U if FROM_U: converted in tileA_run)
    // neighbour (or ghost) states; ghosts are tiled like V, so every source
    // has the same stride and the five loads share one address.  A neighbour
    // in the lane's own block (83 % of interior faces in Morton order at
    // 256 cells; 71 % within one 64-cell tile) is read from the block's copy
    // of its states in LDS instead (NF_A_LDS): the gathers are exec-masked
    // to the other lanes, which is what saves time (the texture path's work
    // grows with the lanes that access memory).  Every cell of the block is
    // a lane of it and live, since neighbours are real cells.  The gathers go first: the waits for the node and state loads
    // then cover them too (masked loads cannot be counted), which measured
    // faster than issuing them after the LDS copies (1.04 vs 0.93 ms).
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const uint32_t idx = t.w[s] & 0x1FFFFFFFu;
        t.nsrc[s] = ((t.w[s] & 0x80000000u) ? a.Gc : a.V) + tix(idx, 0);
        t.loc[s] = NF_A_LDS && !(t.w[s] & 0x80000000u) &&
                   (NF_A_LDS == 2 ? (idx / B) == (p / B) : (idx >> 6) == (p >> 6));
    }
    if (ORDER != 0) {
#pragma unroll
        for (int s = 0; s < 4; ++s)
            if (!t.loc[s]) {
#pragma unroll
                for (int k = 0; k < 5; ++k) t.Vn[s][k] = ldg(t.nsrc[s] + k * TILE);
            }
    }
}

// the tile's gradient, limiter and face values; returns the cell's CFL step
// (DBL_MAX for padding lanes or without a.dt_part)
template <int ORDER, int B, bool FROM_U>
__device__ __forceinline__ double tileA_run(const DevMesh &m, const SweepAArgs &a, TileA &t, double *lds,
                                            double *ldsv) {
    const uint32_t c = t.c, p = t.p, code = t.code;
    const bool live = t.live;
    const uint32_t *w = t.w;
    const bool *loc = t.loc;
    double *Vc = t.Vc;
    double(*Vn)[5] = t.Vn;
    if (FROM_U) {
        const double Uc[5] = {Vc[0], Vc[1], Vc[2], Vc[3], Vc[4]};
        cons_to_prim_d(Uc, Vc);
    }
    double *vt = ldsv + (NF_A_LDS ? (threadIdx.x >> 6) * 5 * TILE : 0);
    if (NF_A_LDS && ORDER != 0) {
#pragma unroll
        for (int k = 0; k < 5; ++k) vt[k * TILE + (p & 63u)] = Vc[k];
    }
    CellGeo g;
    // (its wave-level fences also order the state writes above before the
    // reads below)
    const NodeView nv = node_table_finish<B>(m, p, t.ntl, lds, g);
    if (NF_A_LDS == 2 && ORDER != 0) __syncthreads();  // every wave's states are in LDS
    cell_geo(m, g);
    V3d S[4], r[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) face_geo(m, g, nv, (code >> (8 * s)) & 0xFFu, S[s], &r[s]);
    __builtin_amdgcn_sched_barrier(0);
    if (FROM_U && ORDER != 0) {
        // gathered neighbour states to primvars (ghost values already are)
#pragma unroll
        for (int s = 0; s < 4; ++s)
            if (!loc[s] && !(w[s] & 0x80000000u)) {
                const double Un[5] = {Vn[s][0], Vn[s][1], Vn[s][2], Vn[s][3], Vn[s][4]};
                cons_to_prim_d(Un, Vn[s]);
            }
    }
    if (NF_A_LDS && ORDER != 0) {
#pragma unroll
        for (int s = 0; s < 4; ++s)
            if (loc[s]) {
                const uint32_t idx = w[s] & 0x1FFFFFFFu, j = idx & 63u;
                const double *src = NF_A_LDS == 2 ? ldsv + ((idx >> 6) - (p & ~(uint32_t)(B - 1)) / TILE) * 5 * TILE : vt;
#pragma unroll
                for (int k = 0; k < 5; ++k) Vn[s][k] = src[k * TILE + j];
            }
        if (NF_A_LDS == 2) __syncthreads();  // before the slices are reused for the stores
    }

    if (ORDER == 0) {
        if (live) {
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int k = 0; k < 5; ++k) a.FV[fix(c, s, k)] = Vc[k];
        }
    } else {
        // face normals as the cell adds them: -S where the cell is the
        // reference's neighbour j (cell j subtracts tmp/vol); in place, as
        // the time step below uses |S| only
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const uint32_t mneg = ((code >> (8 * s + 7)) & 1u) ? 0u : 0x80000000u;
            S[s] = {flip(S[s].x, mneg), flip(S[s].y, mneg), flip(S[s].z, mneg)};
        }
        // One variable at a time: gradient row k, its neighbour min/max, the
        // Barth factor and the face values depend on variable k only.
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            // Green-Gauss (include/Numerics.hpp:112-127): cell i adds tmp/vol,
            // cell j subtracts it, faces in reference order
            double g0 = 0.0, g1 = 0.0, g2 = 0.0;
            double vmax = Vc[k], vmin = Vc[k];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                // 0.5 (V_i + V_j) is symmetric in i, j; the non-owner subtracts
                // (through the flipped S: the sign carries through the
                // product and the division exactly)
                const double uf = 0.5 * (Vc[k] + Vn[s][k]);
                g0 = g0 + rdiv(uf * S[s].x, g.ivol);
                g1 = g1 + rdiv(uf * S[s].y, g.ivol);
                g2 = g2 + rdiv(uf * S[s].z, g.ivol);
                if (ORDER == 2) {
                    // neighbour min/max (include/Numerics.hpp:276-291): the
                    // maximum of {v, v_j} over the faces, one v_j at a time
                    vmax = dmax(vmax, Vn[s][k]);
                    vmin = dmin(vmin, Vn[s][k]);
                }
            }
            double delta[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) delta[s] = (g0 * r[s].x + g1 * r[s].y) + g2 * r[s].z;
            const double psi = (ORDER == 2) ? barth_psi_d(delta, vmax, vmin, Vc[k]) : 1.0;
            // limited reconstruction (include/Numerics.hpp:135-145)
            if (NF_A_LDS && NF_A_STX4) {
                // The wave's four rows (s, k) of the tile are written with
                // 16-byte stores: the values pass through the LDS slice (free
                // once the neighbour states are read), and lane l stores the
                // values of lanes 2(l%32), 2(l%32)+1 of slots l/32 and l/32+2.
                // Padding lanes store too: no real cell reads their entries.
                const uint32_t lane = p & 63u;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
                for (int s = 0; s < 4; ++s) vt[s * TILE + lane] = Vc[k] + psi * delta[s];
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const double2 v0 = reinterpret_cast<const double2 *>(vt)[lane];
                const double2 v1 = reinterpret_cast<const double2 *>(vt)[lane + TILE];
                double *row = a.FV + fix(p & ~63u, lane >> 5, k) + 2 * (lane & 31u);
                *reinterpret_cast<double2 *>(row) = v0;
                *reinterpret_cast<double2 *>(row + 2 * 5 * TILE) = v1;
            } else if (live) {
#pragma unroll
                for (int s = 0; s < 4; ++s) a.FV[fix(c, s, k)] = Vc[k] + psi * delta[s];
            }
            if (live) {
                if (a.grad_aux) {
                    a.grad_aux[15 * (size_t)c + k] = g0;
                    a.grad_aux[15 * (size_t)c + k + 5] = g1;
                    a.grad_aux[15 * (size_t)c + k + 10] = g2;
                    a.psi_aux[5 * (size_t)c + k] = psi;
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    double dtc = DBL_MAX;
    if (a.dt_part) {
        const double tc = cell_dt(Vc, S, w, g.vol, a.CFL);
        dtc = (live && tc < DBL_MAX) ? tc : DBL_MAX;  // NaN/+inf never win std::min
        if (a.dt_cell && live) a.dt_cell[c] = tc;
    }
    return dtc;
}

template <int ORDER, int B, bool FROM_U>
__global__ __launch_bounds__(B) NF_SWEEP_ATTR void k_sweepA(DevMesh m, SweepAArgs a) {
    __shared__ double sh[B / 64];
    __shared__ double lds[12 * B];
    __shared__ __align__(16) double ldsv[NF_A_LDS ? 5 * B : 2];  // [wave][5][64] own states
    const LaneCell L = lane_cell<B>(m, a.xcd_remap);
    if (m.st[L.inst].stop) return;
    TileA t;
    t.c = L.c;
    t.live = L.live;
    t.p = L.lb * B + threadIdx.x;
    tileA_issue<ORDER, B, FROM_U>(m, a, t);
    __builtin_amdgcn_sched_barrier(0);
    const double dtc = tileA_run<ORDER, B, FROM_U>(m, a, t, lds, ldsv);
    if (a.dt_part) {
        const double v = block_reduce<true, B>(dtc, sh);
        if (threadIdx.x == 0) a.dt_part[L.lb] = v;
    }
}

// ---------------------------------------------------------------------------
struct SweepBArgs {
    const double *FV;       // [tile][4][5][64]
    const double *Gf;       // [nb/64][5][64] ghost of the reconstructed face state
    const double *U;        // [tile][5][64] state entering the stage
    const double *Uo;       // U_old (U^n)
    double *Uout;           // (may alias U)
    double *Vout;           // prim(Uout), or null (the next sweep A evaluates it)
    const double *dt_cell;  // per-cell step (local time stepping), or null
    int kind;               // 1: U + dt/vol R, 2: RK3 stage 2, 3: RK3 stage 3
    double *res_part;       // per-block sum of R_rho^2, or null
    int check;
    int xcd_remap;
    double *R_aux;          // [N][5] or null: store R, skip the update
    double *Rst, *fst;      // checked stage: R [tile][5][64] and f [cell] for k_apply, no update
};

// FLUX: 0 Rusanov, 1 HLLC
template <int B, int FLUX>
__global__ __launch_bounds__(B) NF_SWEEPB_ATTR void k_sweepB(DevMesh m, SweepBArgs a) {
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
    CellGeo g;
    NodeView nv;
    double R[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const bool zero_old = st.uold_zero;
    double Up[5], Uop[5];
    {
        // loaded first, outside the stage-kind branches (stage 1 and a zero
        // U_old read U twice: same lines, no branch)
        const double *uo = (a.kind == 1 || zero_old) ? a.U : a.Uo;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            Up[k] = a.U[tix(c, k)];
            Uop[k] = uo[tix(c, k)];
        }
    }
    // face values of both sides, one face at a time, the next face's loads
    // issued before the current face is computed
    double Vo4[4][5], Vn4[4][5];
    auto load_face = [&](int s) {
        const uint32_t idx = w[s] & 0x1FFFFFFFu, sn = (w[s] >> 29) & 3u;
        const bool bnd = w[s] & 0x80000000u;
        const double *src = bnd ? a.Gf + tix(idx, 0) : a.FV + fix(idx, sn, 0);
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            Vo4[s][k] = a.FV[fix(c, s, k)];
            Vn4[s][k] = src[k * TILE];
        }
    };
    load_face(0);
    nv = node_table_finish<B>(m, L.lb * B + threadIdx.x, ntl, lds, g);
    cell_geo(m, g);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (s + 1 < 4) load_face(s + 1);
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t byte = (code >> (8 * s)) & 0xFFu;
        V3d S;
        face_geo(m, g, nv, byte, S, nullptr);
        // interior faces (src/Solver.cpp:171-192), boundary faces (:197-229)
        double UO[5], UN[5], F[5];
        prim_to_cons(Vo4[s], UO);
        prim_to_cons(Vn4[s], UN);
        const uint32_t mneg = (byte & 0x80u) ? 0u : 0x80000000u;
        if (FLUX == 1) {
            // F(U_i side, U_j side, S_ij) on both cells of the face
            const bool own = byte & 0x80u;
            double UL[5], UR[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                UL[k] = own ? UO[k] : UN[k];
                UR[k] = own ? UN[k] : UO[k];
            }
            hllc_d(UL, UR, S, F);
        } else {
            rusanov_side(UO, UN, S, mneg, F);
        }
#pragma unroll
        for (int k = 0; k < 5; ++k) R[k] = R[k] + flip(F[k], mneg ^ 0x80000000u);  // R_i -= F, R_j += F
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
        // Solver::TVD_RK3 / explicit_euler (src/Solver.cpp:98-132)
        const double f = rdiv(a.kind == 1 ? dt : (a.kind == 2 ? 1.0 / 4.0 * dt : 2.0 / 3.0 * dt), g.ivol);
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
                for (int k = 0; k < 5; ++k) a.Uout[tix(c, k)] = Un[k];
                if (a.Vout) {
#pragma unroll
                    for (int k = 0; k < 5; ++k) a.Vout[tix(c, k)] = Vnew[k];
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

__global__ __launch_bounds__(BLOCK) void k_apply(DevMesh m, ApplyArgs a) { apply_stage_cell<double>(m, a); }

// prim of a tiled state: V = prim(U)
__global__ __launch_bounds__(BLOCK) void k_prim(const double *U, double *V, uint32_t N) {
    const uint32_t c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= N) return;
    double u[5], v[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) u[k] = U[tix(c, k)];
    cons_to_prim_d(u, v);
#pragma unroll
    for (int k = 0; k < 5; ++k) V[tix(c, k)] = v[k];
}

// State across the boundary: reference-order AoS [N][5] (copied to the
// device as is) to the instance's device tiles, padding cells zero; perm is
// the instance's device cell -> reference cell map.
__global__ __launch_bounds__(BLOCK) void k_state_in(const double *aos, const uint32_t *perm, double *dev,
                                                    uint32_t N, uint32_t Npad) {
    const uint32_t p = blockIdx.x * BLOCK + threadIdx.x;
    if (p >= Npad) return;
    const size_t r = p < N ? 5 * (size_t)perm[p] : 0;
#pragma unroll
    for (int k = 0; k < 5; ++k) dev[tix(p, k)] = p < N ? aos[r + k] : 0.0;
}
// and back: device tiles to reference-order AoS
__global__ __launch_bounds__(BLOCK) void k_state_out(const double *dev, const uint32_t *perm, double *aos,
                                                     uint32_t N) {
    const uint32_t p = blockIdx.x * BLOCK + threadIdx.x;
    if (p >= N) return;
    const size_t r = 5 * (size_t)perm[p];
#pragma unroll
    for (int k = 0; k < 5; ++k) aos[r + k] = dev[tix(p, k)];
}
#endif  // NF_DEVICE_HELPERS_ONLY

}  // namespace nf
