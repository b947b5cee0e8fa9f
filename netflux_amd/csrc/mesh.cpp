// Host mesh pipeline (see mesh.hpp).  Compiled with -ffp-contract=off: the
// geometry must be evaluated exactly in the stated operation order.
#include "mesh.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace nf {

namespace {

// Local node triples of the four faces of a tetrahedron
// (src/geometry/GeometryUtilities.cpp:71-105).
const int TET_FACE[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

struct V3 {
    double x, y, z;
};
inline V3 ld(const double *p) { return {p[0], p[1], p[2]}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 scl(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

// tetrahedron_calc_volume / _centroid (GeometryUtilities.cpp:289-316)
inline double tet_volume(V3 n0, V3 n1, V3 n2, V3 n3) {
    V3 ab = sub(n1, n0), ac = sub(n2, n0), ad = sub(n3, n0);
    return dot(cross(ab, ac), ad) / 6;
}
inline V3 tet_centroid(V3 n0, V3 n1, V3 n2, V3 n3) {
    return scl(1.0 / 4, add(add(add(n0, n1), n2), n3));
}
// triangle_calc_face_normal / _centroid (GeometryUtilities.cpp:318-329)
inline V3 tri_normal(V3 n0, V3 n1, V3 n2) {
    V3 ab = sub(n1, n0), ac = sub(n2, n0);
    return scl(0.5, cross(ab, ac));
}
inline V3 tri_centroid(V3 n0, V3 n1, V3 n2) {
    V3 s = add(add(n0, n1), n2);
    return {s.x / 3, s.y / 3, s.z / 3};
}

inline uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline double uniform01(uint64_t &s) {
    return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
}

inline void sort3(uint32_t &a, uint32_t &b, uint32_t &c) {
    if (a > b) std::swap(a, b);
    if (b > c) std::swap(b, c);
    if (a > b) std::swap(a, b);
}

}  // namespace

// ---------------------------------------------------------------------------
void build_box(uint32_t nx, uint32_t ny, uint32_t nz, double lx, double ly,
               double lz, double jitter, uint64_t seed, double bump_h,
               double bump_sigma, const int32_t bc[6], PrimalMesh &m) {
    if (nx == 0 || ny == 0 || nz == 0) throw MeshError{"empty box"};
    const uint64_t NX = nx + 1, NY = ny + 1, NZ = nz + 1;
    const uint64_t nn = NX * NY * NZ;
    const uint64_t nc = 6ull * nx * ny * nz;
    if (nn >= (1ull << 32) || nc >= (1ull << 29)) throw MeshError{"box too large"};
    m.nodes.assign(3 * nn, 0.0);
    const double hx = lx / nx, hy = ly / ny, hz = lz / nz;
    uint64_t rng = seed;
    for (uint64_t k = 0; k < NZ; ++k)
        for (uint64_t j = 0; j < NY; ++j)
            for (uint64_t i = 0; i < NX; ++i) {
                uint64_t n = i + NX * (j + NY * k);
                double x = lx * (double)i / nx, y = ly * (double)j / ny,
                       z = lz * (double)k / nz;
                bool interior = i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz;
                if (interior && jitter > 0.0) {
                    x += (2.0 * uniform01(rng) - 1.0) * jitter * hx;
                    y += (2.0 * uniform01(rng) - 1.0) * jitter * hy;
                    z += (2.0 * uniform01(rng) - 1.0) * jitter * hz;
                }
                if (bump_h > 0.0) {
                    double d = x - 0.5 * lx;
                    double yb = bump_h * std::exp(-d * d / (2.0 * bump_sigma * bump_sigma));
                    y = yb + y * (1.0 - yb / ly);
                }
                m.nodes[3 * n + 0] = x;
                m.nodes[3 * n + 1] = y;
                m.nodes[3 * n + 2] = z;
            }
    // Six tetrahedra per hex around the (0,0,0)-(1,1,1) diagonal, one per
    // axis permutation; odd permutations get nodes 1,2 swapped so every
    // volume is positive in VTK order.
    static const int PERM[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2},
                                   {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    static const bool ODD[6] = {false, true, true, false, false, true};
    m.cells.resize(4 * nc);
    uint64_t c = 0;
    for (uint64_t k = 0; k < nz; ++k)
        for (uint64_t j = 0; j < ny; ++j)
            for (uint64_t i = 0; i < nx; ++i) {
                for (int p = 0; p < 6; ++p) {
                    uint64_t d[3] = {0, 0, 0};
                    uint32_t v[4];
                    auto node = [&](const uint64_t *o) {
                        return (uint32_t)((i + o[0]) + NX * ((j + o[1]) + NY * (k + o[2])));
                    };
                    v[0] = node(d);
                    d[PERM[p][0]] = 1;
                    v[1] = node(d);
                    d[PERM[p][1]] = 1;
                    v[2] = node(d);
                    d[PERM[p][2]] = 1;
                    v[3] = node(d);
                    if (ODD[p]) std::swap(v[1], v[2]);
                    std::memcpy(&m.cells[4 * c], v, sizeof v);
                    ++c;
                }
            }
    // Boundary triangles: tet faces lying in one of the six box planes, in
    // (cell, local face) order, node order of the tet face.
    std::vector<std::vector<uint32_t>> tri(6);
    auto logical = [&](uint32_t n, uint64_t &i, uint64_t &j, uint64_t &k) {
        i = n % NX;
        j = (n / NX) % NY;
        k = n / (NX * NY);
    };
    for (uint64_t t = 0; t < nc; ++t) {
        for (int f = 0; f < 4; ++f) {
            uint32_t a = m.cells[4 * t + TET_FACE[f][0]], b = m.cells[4 * t + TET_FACE[f][1]],
                     cc = m.cells[4 * t + TET_FACE[f][2]];
            uint64_t I[3], J[3], K[3];
            logical(a, I[0], J[0], K[0]);
            logical(b, I[1], J[1], K[1]);
            logical(cc, I[2], J[2], K[2]);
            int side = -1;
            if (I[0] == 0 && I[1] == 0 && I[2] == 0) side = 0;
            else if (I[0] == nx && I[1] == nx && I[2] == nx) side = 1;
            else if (J[0] == 0 && J[1] == 0 && J[2] == 0) side = 2;
            else if (J[0] == ny && J[1] == ny && J[2] == ny) side = 3;
            else if (K[0] == 0 && K[1] == 0 && K[2] == 0) side = 4;
            else if (K[0] == nz && K[1] == nz && K[2] == nz) side = 5;
            if (side < 0) continue;
            tri[side].push_back(a);
            tri[side].push_back(b);
            tri[side].push_back(cc);
        }
    }
    m.patch_n.resize(6);
    m.patch_bc.resize(6);
    m.patch_tri.clear();
    for (int s = 0; s < 6; ++s) {
        m.patch_n[s] = (uint32_t)(tri[s].size() / 3);
        m.patch_bc[s] = bc[s];
        m.patch_tri.insert(m.patch_tri.end(), tri[s].begin(), tri[s].end());
    }
}

// ---------------------------------------------------------------------------
void build_fv_grid(const PrimalMesh &m, FVGrid &g) {
    if (m.mixed()) {
        build_fv_grid_mixed(m, g);
        return;
    }
    const uint32_t N = m.n_cells(), nn = m.n_nodes();
    const uint32_t P = (uint32_t)m.patch_n.size();
    if (N == 0) throw MeshError{"mesh has no cells"};
    if (m.patch_bc.size() != P) throw MeshError{"patch_bc size mismatch"};
    uint64_t n_btri = 0;
    for (uint32_t p = 0; p < P; ++p) n_btri += m.patch_n[p];
    if (m.patch_tri.size() != 3 * n_btri) throw MeshError{"patch face array size mismatch"};
    for (uint64_t q = 0; q < m.cells.size(); ++q)
        if (m.cells[q] >= nn) throw MeshError{"cell node index out of range"};
    for (uint64_t q = 0; q < m.patch_tri.size(); ++q)
        if (m.patch_tri[q] >= nn) throw MeshError{"patch node index out of range"};

    // Bucket every tet face by its smallest node.
    struct Ent {
        uint32_t b, c, code;  // code = cell*4 + k; bit31 of c: used
    };
    std::vector<uint32_t> start((size_t)nn + 1, 0);
    for (uint32_t t = 0; t < N; ++t)
        for (int k = 0; k < 4; ++k) {
            uint32_t a = m.cells[4 * (size_t)t + TET_FACE[k][0]],
                     b = m.cells[4 * (size_t)t + TET_FACE[k][1]],
                     c = m.cells[4 * (size_t)t + TET_FACE[k][2]];
            sort3(a, b, c);
            start[a + 1]++;
        }
    for (uint32_t n = 0; n < nn; ++n) start[n + 1] += start[n];
    std::vector<Ent> ent((size_t)4 * N);
    {
        std::vector<uint32_t> fill(start.begin(), start.end() - 1);
        for (uint32_t t = 0; t < N; ++t)
            for (int k = 0; k < 4; ++k) {
                uint32_t a = m.cells[4 * (size_t)t + TET_FACE[k][0]],
                         b = m.cells[4 * (size_t)t + TET_FACE[k][1]],
                         c = m.cells[4 * (size_t)t + TET_FACE[k][2]];
                sort3(a, b, c);
                ent[fill[a]++] = {b, c, t * 4u + (uint32_t)k};
            }
    }
    // Pair faces inside each bucket.  state: 0 single, 1 paired, 2 consumed
    std::vector<uint8_t> state(ent.size(), 0);
    struct IFace {
        uint32_t i, j, ki;
    };
    std::vector<IFace> ifaces;
    ifaces.reserve((size_t)2 * N + 16);
    for (uint32_t a = 0; a < nn; ++a) {
        uint32_t s0 = start[a], s1 = start[a + 1];
        for (uint32_t p = s0; p < s1; ++p) {
            if (state[p]) continue;
            for (uint32_t q = p + 1; q < s1; ++q) {
                if (state[q] || ent[q].b != ent[p].b || ent[q].c != ent[p].c) continue;
                uint32_t tp = ent[p].code >> 2, tq = ent[q].code >> 2;
                if (tp == tq) throw MeshError{"degenerate tetrahedron"};
                state[p] = state[q] = 1;
                if (tp < tq) ifaces.push_back({tp, tq, ent[p].code & 3u});
                else ifaces.push_back({tq, tp, ent[q].code & 3u});
                for (uint32_t r = q + 1; r < s1; ++r)
                    if (ent[r].b == ent[p].b && ent[r].c == ent[p].c)
                        throw MeshError{"face shared by more than two cells"};
                break;
            }
        }
    }
    // Sort interior faces by (i, j): counting sort on i, insertion on j.
    const uint32_t Fint = (uint32_t)ifaces.size();
    std::vector<IFace> sorted(Fint);
    {
        std::vector<uint32_t> cnt((size_t)N + 1, 0);
        for (auto &f : ifaces) cnt[f.i + 1]++;
        for (uint32_t i = 0; i < N; ++i) cnt[i + 1] += cnt[i];
        std::vector<uint32_t> pos(cnt.begin(), cnt.end() - 1);
        for (auto &f : ifaces) sorted[pos[f.i]++] = f;
        for (uint32_t i = 0; i < N; ++i) {
            uint32_t b = cnt[i], e = cnt[i + 1];
            for (uint32_t x = b + 1; x < e; ++x) {
                IFace v = sorted[x];
                uint32_t y = x;
                while (y > b && sorted[y - 1].j > v.j) {
                    sorted[y] = sorted[y - 1];
                    --y;
                }
                sorted[y] = v;
            }
        }
    }
    ifaces.clear();
    ifaces.shrink_to_fit();

    // Boundary faces patch by patch; ghost ids in patch element order.
    const uint32_t G = (uint32_t)n_btri;
    g.N = N;
    g.NT = N + G;
    g.Fint = Fint;
    g.F = Fint + G;
    const uint32_t F = g.F;
    g.ij.resize((size_t)2 * F);
    g.fnodes.resize((size_t)3 * F);
    for (uint32_t f = 0; f < Fint; ++f) {
        const IFace &x = sorted[f];
        g.ij[2 * (size_t)f] = x.i;
        g.ij[2 * (size_t)f + 1] = x.j;
        for (int q = 0; q < 3; ++q)
            g.fnodes[3 * (size_t)f + q] = m.cells[4 * (size_t)x.i + TET_FACE[x.ki][q]];
    }
    sorted.clear();
    sorted.shrink_to_fit();
    g.pfirst.resize(P);
    g.pn.resize(P);
    g.pbc = m.patch_bc;
    uint32_t jg = N, f = Fint;
    uint64_t tri_off = 0;
    for (uint32_t p = 0; p < P; ++p) {
        g.pfirst[p] = f;
        g.pn[p] = m.patch_n[p];
        struct BF {
            uint32_t i, j;
            uint64_t tri;
        };
        std::vector<BF> bf(m.patch_n[p]);
        for (uint32_t e = 0; e < m.patch_n[p]; ++e, ++tri_off) {
            uint32_t a = m.patch_tri[3 * tri_off], b = m.patch_tri[3 * tri_off + 1],
                     c = m.patch_tri[3 * tri_off + 2];
            sort3(a, b, c);
            int64_t hit = -1;
            for (uint32_t q = start[a]; q < start[a + 1]; ++q)
                if (ent[q].b == b && ent[q].c == c) {
                    hit = q;
                    break;
                }
            if (hit < 0) throw MeshError{"patch face matches no cell face"};
            if (state[hit] != 0) throw MeshError{"patch face is interior or listed twice"};
            state[hit] = 2;
            bf[e] = {ent[hit].code >> 2, jg++, tri_off};
        }
        std::stable_sort(bf.begin(), bf.end(),
                         [](const BF &x, const BF &y) { return x.i < y.i; });
        for (auto &x : bf) {
            g.ij[2 * (size_t)f] = x.i;
            g.ij[2 * (size_t)f + 1] = x.j;
            for (int q = 0; q < 3; ++q) g.fnodes[3 * (size_t)f + q] = m.patch_tri[3 * x.tri + q];
            ++f;
        }
    }
    for (size_t q = 0; q < state.size(); ++q)
        if (state[q] == 0) throw MeshError{"boundary face of cell " + std::to_string(ent[q].code >> 2) +
                                           " is not covered by any patch"};
    ent.clear();
    ent.shrink_to_fit();

    // Geometry (FV_Grid::calc_geometry_properties, src/geometry/FV_Grid.cpp:154-205)
    g.vol.assign(g.NT, 0.0);
    g.cen.assign((size_t)3 * g.NT, 0.0);
    const double *X = m.nodes.data();
    for (uint32_t i = 0; i < N; ++i) {
        const uint32_t *v = &m.cells[4 * (size_t)i];
        V3 n0 = ld(X + 3 * (size_t)v[0]), n1 = ld(X + 3 * (size_t)v[1]),
           n2 = ld(X + 3 * (size_t)v[2]), n3 = ld(X + 3 * (size_t)v[3]);
        g.vol[i] = tet_volume(n0, n1, n2, n3);
        V3 c = tet_centroid(n0, n1, n2, n3);
        g.cen[3 * (size_t)i] = c.x;
        g.cen[3 * (size_t)i + 1] = c.y;
        g.cen[3 * (size_t)i + 2] = c.z;
    }
    g.S.resize((size_t)3 * F);
    g.ri.resize((size_t)3 * F);
    g.rj.resize((size_t)3 * F);
    g.flip.assign(F, 0);
    for (uint32_t q = 0; q < F; ++q) {
        uint32_t i = g.ij[2 * (size_t)q], j = g.ij[2 * (size_t)q + 1];
        const uint32_t *fn = &g.fnodes[3 * (size_t)q];
        V3 a = ld(X + 3 * (size_t)fn[0]), b = ld(X + 3 * (size_t)fn[1]), c = ld(X + 3 * (size_t)fn[2]);
        V3 fc = tri_centroid(a, b, c);
        V3 ci = ld(&g.cen[3 * (size_t)i]);
        if (j >= N) {
            g.vol[j] = 0.0;
            V3 cg = sub(scl(2, fc), ci);
            g.cen[3 * (size_t)j] = cg.x;
            g.cen[3 * (size_t)j + 1] = cg.y;
            g.cen[3 * (size_t)j + 2] = cg.z;
        }
        V3 cj = ld(&g.cen[3 * (size_t)j]);
        V3 S = tri_normal(a, b, c);
        double nd = dot(S, sub(cj, ci));
        if (nd == 0.0) throw MeshError{"face normal orthogonal to the cell-centre line"};
        if (nd < 0) {
            S = scl(-1, S);
            g.flip[q] = 1;
        }
        V3 r1 = sub(fc, ci), r2 = sub(fc, cj);
        double *Sp = &g.S[3 * (size_t)q], *r1p = &g.ri[3 * (size_t)q], *r2p = &g.rj[3 * (size_t)q];
        Sp[0] = S.x; Sp[1] = S.y; Sp[2] = S.z;
        r1p[0] = r1.x; r1p[1] = r1.y; r1p[2] = r1.z;
        r2p[0] = r2.x; r2p[1] = r2.y; r2p[2] = r2.z;
    }
    // Delta S over interior faces only (src/Solver.cpp:276-298)
    g.dS.assign((size_t)3 * g.NT, 0.0);
    for (uint32_t q = 0; q < Fint; ++q) {
        uint32_t i = g.ij[2 * (size_t)q], j = g.ij[2 * (size_t)q + 1];
        for (int d = 0; d < 3; ++d) {
            double t = 0.5 * std::fabs(g.S[3 * (size_t)q + d]);
            g.dS[3 * (size_t)i + d] += t;
            g.dS[3 * (size_t)j + d] += t;
        }
    }
}

// ---------------------------------------------------------------------------
std::vector<uint32_t> rcm_order(const FVGrid &g) {
    const uint32_t N = g.N;
    std::vector<uint32_t> deg(N, 0), off((size_t)N + 1, 0);
    for (uint32_t f = 0; f < g.Fint; ++f) {
        deg[g.ij[2 * (size_t)f]]++;
        deg[g.ij[2 * (size_t)f + 1]]++;
    }
    for (uint32_t i = 0; i < N; ++i) off[i + 1] = off[i] + deg[i];
    std::vector<uint32_t> adj(off[N]), pos(off.begin(), off.end() - 1);
    for (uint32_t f = 0; f < g.Fint; ++f) {
        uint32_t i = g.ij[2 * (size_t)f], j = g.ij[2 * (size_t)f + 1];
        adj[pos[i]++] = j;
        adj[pos[j]++] = i;
    }
    std::vector<uint32_t> order;
    order.reserve(N);
    std::vector<uint8_t> seen(N, 0);
    std::vector<uint32_t> level(N, 0);
    auto bfs_last = [&](uint32_t s) {
        // returns the last vertex reached by a BFS from s (pseudo-peripheral)
        std::vector<uint32_t> q{s};
        std::vector<uint8_t> vis(N, 0);
        vis[s] = 1;
        for (size_t h = 0; h < q.size(); ++h)
            for (uint32_t e = off[q[h]]; e < off[q[h] + 1]; ++e)
                if (!vis[adj[e]]) {
                    vis[adj[e]] = 1;
                    q.push_back(adj[e]);
                }
        return q.back();
    };
    for (uint32_t s0 = 0; s0 < N; ++s0) {
        if (seen[s0]) continue;
        uint32_t s = bfs_last(s0);
        size_t head = order.size();
        order.push_back(s);
        seen[s] = 1;
        std::vector<uint32_t> nb;
        for (size_t h = head; h < order.size(); ++h) {
            uint32_t v = order[h];
            nb.clear();
            for (uint32_t e = off[v]; e < off[v + 1]; ++e)
                if (!seen[adj[e]]) nb.push_back(adj[e]);
            std::sort(nb.begin(), nb.end(), [&](uint32_t a, uint32_t b) {
                return deg[a] != deg[b] ? deg[a] < deg[b] : a < b;
            });
            for (uint32_t w : nb) {
                seen[w] = 1;
                order.push_back(w);
            }
        }
    }
    std::reverse(order.begin(), order.end());
    return order;  // order[device] = reference cell
}

// Cells sorted by the Morton (Z-order) code of their centroid on a 2^21 grid
// over the bounding box: neighbours end up close in memory at every scale,
// which keeps face gathers inside the caches.
std::vector<uint32_t> morton_order(const FVGrid &g) {
    const uint32_t N = g.N;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (uint32_t c = 0; c < N; ++c)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], g.cen[3 * (size_t)c + d]);
            hi[d] = std::max(hi[d], g.cen[3 * (size_t)c + d]);
        }
    auto spread = [](uint64_t x) {
        x &= 0x1FFFFF;
        x = (x | x << 32) & 0x1F00000000FFFFull;
        x = (x | x << 16) & 0x1F0000FF0000FFull;
        x = (x | x << 8) & 0x100F00F00F00F00Full;
        x = (x | x << 4) & 0x10C30C30C30C30C3ull;
        x = (x | x << 2) & 0x1249249249249249ull;
        return x;
    };
    std::vector<std::pair<uint64_t, uint32_t>> key(N);
    for (uint32_t c = 0; c < N; ++c) {
        uint64_t q[3];
        for (int d = 0; d < 3; ++d) {
            double ext = hi[d] - lo[d];
            double t = ext > 0 ? (g.cen[3 * (size_t)c + d] - lo[d]) / ext : 0.0;
            q[d] = (uint64_t)std::min(2097151.0, std::max(0.0, t * 2097151.0));
        }
        key[c] = {spread(q[0]) | spread(q[1]) << 1 | spread(q[2]) << 2, c};
    }
    std::sort(key.begin(), key.end());
    std::vector<uint32_t> order(N);
    for (uint32_t d = 0; d < N; ++d) order[d] = key[d].second;
    return order;
}

void build_cell_layout(const PrimalMesh &m, const FVGrid &g, int order, CellLayout &L) {
    const uint32_t N = g.N;
    if (N > NBR_INDEX_MASK) throw MeshError{"too many cells for 29-bit neighbour index"};
    L.N = N;
    if (order == 1) L.perm = rcm_order(g);
    else if (order == 2) L.perm = morton_order(g);
    else {
        L.perm.resize(N);
        std::iota(L.perm.begin(), L.perm.end(), 0u);
    }
    L.iperm.assign(N, 0);
    for (uint32_t d = 0; d < N; ++d) L.iperm[L.perm[d]] = d;
    // faces of each reference cell in increasing face index
    std::vector<uint32_t> cf((size_t)4 * N), cnt(N, 0);
    auto push = [&](uint32_t c, uint32_t f) {
        if (cnt[c] >= 4) throw MeshError{"cell with more than four faces"};
        cf[4 * (size_t)c + cnt[c]++] = f;
    };
    for (uint32_t f = 0; f < g.F; ++f) {
        push(g.ij[2 * (size_t)f], f);
        if (g.ij[2 * (size_t)f + 1] < N) push(g.ij[2 * (size_t)f + 1], f);
    }
    for (uint32_t c = 0; c < N; ++c)
        if (cnt[c] != 4) throw MeshError{"cell " + std::to_string(c) + " is not closed by four faces"};
    std::vector<int32_t> face_bc(g.F - g.Fint, 0);
    for (size_t p = 0; p < g.pfirst.size(); ++p)
        for (uint32_t q = 0; q < g.pn[p]; ++q) face_bc[g.pfirst[p] - g.Fint + q] = g.pbc[p];
    auto slot_of = [&](uint32_t c, uint32_t f) -> uint32_t {
        for (uint32_t s = 0; s < 4; ++s)
            if (cf[4 * (size_t)c + s] == f) return s;
        throw MeshError{"internal: face not found in cell"};
    };
    L.cnodes.assign((size_t)4 * N, 0);
    L.nbr.assign((size_t)4 * N, 0);
    L.ghost.assign((size_t)4 * N, 0xFFFFFFFFu);
    L.fcode.assign(N, 0);
    L.bface.clear();
    L.bbc.clear();
    L.bfidx.clear();
    for (uint32_t d = 0; d < N; ++d) {
        const uint32_t r = L.perm[d];
        const uint32_t *v = &m.cells[4 * (size_t)r];
        for (int q = 0; q < 4; ++q) L.cnodes[4 * (size_t)d + q] = v[q];
        uint32_t code = 0;
        for (uint32_t s = 0; s < 4; ++s) {
            const uint32_t f = cf[4 * (size_t)r + s];
            const uint32_t i = g.ij[2 * (size_t)f], j = g.ij[2 * (size_t)f + 1];
            uint32_t byte = 0;
            for (int q = 0; q < 3; ++q) {
                uint32_t node = g.fnodes[3 * (size_t)f + q], loc = 4;
                for (uint32_t t = 0; t < 4; ++t)
                    if (v[t] == node) loc = t;
                if (loc == 4) throw MeshError{"internal: face node not in cell"};
                byte |= loc << (2 * q);
            }
            if (g.flip[f]) byte |= FC_FLIP;
            if (i == r) byte |= FC_OWNER;
            code |= byte << (8 * s);
            uint32_t w;
            if (j >= N) {
                w = NBR_BOUNDARY | (uint32_t)L.bface.size();
                L.bface.push_back(d | (s << 29));
                L.bbc.push_back((uint8_t)face_bc[f - g.Fint]);
                L.bfidx.push_back(f);
                L.ghost[4 * (size_t)d + s] = j;
            } else {
                uint32_t other = (i == r) ? j : i;
                w = L.iperm[other] | (slot_of(other, f) << 29);
            }
            L.nbr[4 * (size_t)d + s] = w;
        }
        L.fcode[d] = code;
    }
}

}  // namespace nf
