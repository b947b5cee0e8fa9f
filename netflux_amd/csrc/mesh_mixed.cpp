// Mixed-element meshes (tetrahedra, hexahedra, wedges, pyramids; triangle
// and quadrilateral boundary faces): element face tables, element geometry
// and the finite-volume face graph, following the reference's
// GeometryUtilities (src/geometry/GeometryUtilities.cpp:48-441) and FV_Grid
// (src/geometry/FV_Grid.cpp:21-267) with the same operations in the same
// order, so volumes, centroids and face vectors carry the reference's bits.
// CPU only.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mesh.hpp"

namespace nf {

namespace {
struct V3 {
    double x, y, z;
};
inline V3 ld(const double *p) { return {p[0], p[1], p[2]}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }

// GeometryUtilities.cpp:307-329
V3 tri_normal(V3 n0, V3 n1, V3 n2) { return 0.5 * cross(n1 - n0, n2 - n0); }
V3 tri_centroid(V3 n0, V3 n1, V3 n2) { return ((n0 + n1) + n2) / 3; }
// :411-441: two triangles (n0 n1 n2) and (n0 n2 n3)
V3 quad_normal(V3 n0, V3 n1, V3 n2, V3 n3) { return tri_normal(n0, n1, n2) + tri_normal(n0, n2, n3); }
V3 quad_centroid(V3 n0, V3 n1, V3 n2, V3 n3) {
    const double a = norm(tri_normal(n0, n1, n2)), b = norm(tri_normal(n0, n2, n3));
    const V3 ca = tri_centroid(n0, n1, n2), cb = tri_centroid(n0, n2, n3);
    const double s = a + b;
    return (a * ca + b * cb) / s;
}
// :283-305
void tet_props(V3 n0, V3 n1, V3 n2, V3 n3, double &vol, V3 &cen) {
    vol = dot(cross(n1 - n0, n2 - n0), n3 - n0) / 6;
    cen = 0.25 * (((n0 + n1) + n2) + n3);
}
// :331-342: base (n0 n1 n2 n3), apex n4
void pyramid_props(V3 n0, V3 n1, V3 n2, V3 n3, V3 n4, double &vol, V3 &cen) {
    const V3 cb = quad_centroid(n0, n1, n2, n3);
    const V3 d = cb - n4;
    const V3 Sb = quad_normal(n0, n3, n2, n1);
    cen = 0.75 * cb + 0.25 * n4;
    vol = dot(d, Sb) / 3;
}
// volume-weighted sum of sub-element centroids (:343-409)
struct Acc {
    double vol = 0;
    V3 cen{0, 0, 0};
    void add(double v, V3 c) {
        vol += v;
        cen = cen + v * c;
    }
};

const uint8_t TET_F[4][4] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
const uint8_t HEX_F[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4}, {3, 2, 6, 7}, {3, 0, 4, 7}, {1, 2, 6, 5}};
const uint8_t PYR_F[5][4] = {{0, 1, 2, 3}, {0, 1, 4}, {1, 2, 4}, {2, 3, 4}, {3, 0, 4}};
const uint8_t WEDGE_F[5][4] = {{0, 1, 2}, {4, 3, 5}, {1, 0, 3, 4}, {0, 2, 5, 3}, {4, 5, 2, 1}};

void cell_props(int type, const double *X, const uint32_t *v, double &vol, V3 &cen) {
    auto P = [&](int k) { return ld(X + 3 * (size_t)v[k]); };
    if (type == ELEM_TET) {
        tet_props(P(0), P(1), P(2), P(3), vol, cen);
    } else if (type == ELEM_PYRAMID) {
        pyramid_props(P(0), P(1), P(2), P(3), P(4), vol, cen);
    } else if (type == ELEM_HEX) {
        const V3 n0 = P(0), n1 = P(1), n2 = P(2), n3 = P(3), n4 = P(4), n5 = P(5), n6 = P(6), n7 = P(7);
        const V3 g = ((((((n0 + n1) + n2) + n3) + n4) + n5) + n6 + n7) / 8;
        Acc a;
        double pv;
        V3 pc;
        pyramid_props(n0, n1, n2, n3, g, pv, pc), a.add(pv, pc);
        pyramid_props(n4, n7, n6, n5, g, pv, pc), a.add(pv, pc);
        pyramid_props(n4, n0, n3, n7, g, pv, pc), a.add(pv, pc);
        pyramid_props(n2, n1, n5, n6, g, pv, pc), a.add(pv, pc);
        pyramid_props(n1, n0, n4, n5, g, pv, pc), a.add(pv, pc);
        pyramid_props(n3, n2, n6, n7, g, pv, pc), a.add(pv, pc);
        vol = a.vol;
        cen = a.cen / a.vol;
    } else {  // wedge
        const V3 n0 = P(0), n1 = P(1), n2 = P(2), n3 = P(3), n4 = P(4), n5 = P(5);
        const V3 g = ((((n0 + n1) + n2) + n3) + n4 + n5) / 6;
        Acc a;
        double pv;
        V3 pc;
        tet_props(n0, n2, n1, g, pv, pc), a.add(pv, pc);
        tet_props(n3, n4, n5, g, pv, pc), a.add(pv, pc);
        pyramid_props(n0, n3, n5, n2, g, pv, pc), a.add(pv, pc);
        pyramid_props(n0, n1, n4, n3, g, pv, pc), a.add(pv, pc);
        pyramid_props(n4, n1, n2, n5, g, pv, pc), a.add(pv, pc);
        vol = a.vol;
        cen = a.cen / a.vol;
    }
}

void face_props(int n, const double *X, const uint32_t *f, V3 &S, V3 &c) {
    const V3 a = ld(X + 3 * (size_t)f[0]), b = ld(X + 3 * (size_t)f[1]), d = ld(X + 3 * (size_t)f[2]);
    if (n == 3) {
        S = tri_normal(a, b, d);
        c = tri_centroid(a, b, d);
    } else {
        const V3 e = ld(X + 3 * (size_t)f[3]);
        S = quad_normal(a, b, d, e);
        c = quad_centroid(a, b, d, e);
    }
}

// a face as the reference's map key sees it: node count, then sorted nodes
struct Key {
    uint32_t n, s[4];
    bool operator<(const Key &o) const {
        if (n != o.n) return n < o.n;
        for (uint32_t q = 0; q < n; ++q)
            if (s[q] != o.s[q]) return s[q] < o.s[q];
        return false;
    }
    bool operator==(const Key &o) const { return !(*this < o) && !(o < *this); }
};
Key make_key(int n, const uint32_t *f) {
    Key k{(uint32_t)n, {f[0], f[1], f[2], n == 4 ? f[3] : 0u}};
    std::sort(k.s, k.s + n);
    return k;
}
}  // namespace

int elem_n_nodes(int t) {
    switch (t) {
        case ELEM_TRI: return 3;
        case ELEM_QUAD: return 4;
        case ELEM_TET: return 4;
        case ELEM_HEX: return 8;
        case ELEM_WEDGE: return 6;
        case ELEM_PYRAMID: return 5;
        default: return 0;
    }
}
int elem_n_faces(int t) {
    switch (t) {
        case ELEM_TET: return 4;
        case ELEM_HEX: return 6;
        case ELEM_WEDGE: return 5;
        case ELEM_PYRAMID: return 5;
        default: return 0;
    }
}
int elem_face(int t, const uint32_t *ve, int k, uint32_t fe[4]) {
    const uint8_t *row;
    int n;
    if (t == ELEM_TET) row = TET_F[k], n = 3;
    else if (t == ELEM_HEX) row = HEX_F[k], n = 4;
    else if (t == ELEM_PYRAMID) row = PYR_F[k], n = k == 0 ? 4 : 3;
    else row = WEDGE_F[k], n = k < 2 ? 3 : 4;
    for (int q = 0; q < n; ++q) fe[q] = ve[row[q]];
    return n;
}

void primal_from_desc(const nf_mesh_desc &d, PrimalMesh &m) {
    if (!d.nodes || !d.cells || d.n_cells == 0) throw MeshError("mesh has no cells");
    if (d.n_patches && (!d.patch_n_faces || !d.patch_bc)) throw MeshError("missing patch arrays");
    m = PrimalMesh{};
    m.nodes.assign(d.nodes, d.nodes + 3 * (size_t)d.n_nodes);
    m.patch_n.assign(d.patch_n_faces, d.patch_n_faces + d.n_patches);
    m.patch_bc.assign(d.patch_bc, d.patch_bc + d.n_patches);
    size_t nf = 0;
    for (uint32_t p = 0; p < d.n_patches; ++p) nf += d.patch_n_faces[p];
    if (nf && !d.patch_faces) throw MeshError("missing patch faces");
    const bool typed = d.cell_types || d.cell_offsets || d.patch_face_types || d.patch_face_offsets;
    if (!typed) {
        m.cells.assign(d.cells, d.cells + 4 * (size_t)d.n_cells);
        m.patch_tri.assign(d.patch_faces, d.patch_faces + 3 * nf);
        return;
    }
    if (!d.cell_types || !d.cell_offsets || (nf && (!d.patch_face_types || !d.patch_face_offsets)))
        throw MeshError("element types need cell_types, cell_offsets, patch_face_types and patch_face_offsets");
    bool all_tet = true;
    if (d.cell_offsets[0] != 0) throw MeshError("cell_offsets[0] must be 0");
    for (uint32_t i = 0; i < d.n_cells; ++i) {
        const int t = d.cell_types[i];
        if (elem_n_faces(t) == 0) throw MeshError("cell " + std::to_string(i) + ": not a volume element type");
        if (d.cell_offsets[i + 1] - d.cell_offsets[i] != (uint32_t)elem_n_nodes(t) ||
            d.cell_offsets[i + 1] < d.cell_offsets[i])
            throw MeshError("cell " + std::to_string(i) + ": node count does not match its type");
        all_tet = all_tet && t == ELEM_TET;
    }
    if (nf && d.patch_face_offsets[0] != 0) throw MeshError("patch_face_offsets[0] must be 0");
    for (size_t e = 0; e < nf; ++e) {
        const int t = d.patch_face_types[e];
        if (t != ELEM_TRI && t != ELEM_QUAD) throw MeshError("patch face " + std::to_string(e) + ": not a face type");
        if (d.patch_face_offsets[e + 1] - d.patch_face_offsets[e] != (uint32_t)elem_n_nodes(t) ||
            d.patch_face_offsets[e + 1] < d.patch_face_offsets[e])
            throw MeshError("patch face " + std::to_string(e) + ": node count does not match its type");
        all_tet = all_tet && t == ELEM_TRI;
    }
    const size_t nc = d.cell_offsets[d.n_cells], nfn = nf ? d.patch_face_offsets[nf] : 0;
    m.cells.assign(d.cells, d.cells + nc);
    m.patch_tri.assign(d.patch_faces, d.patch_faces + nfn);
    if (all_tet) return;  // same arrays as the tetrahedral layout
    m.ctype.assign(d.cell_types, d.cell_types + d.n_cells);
    m.coff.assign(d.cell_offsets, d.cell_offsets + d.n_cells + 1);
    m.ftype.assign(d.patch_face_types, d.patch_face_types + nf);
    m.foff.assign(d.patch_face_offsets, d.patch_face_offsets + nf + 1);
}

void build_fv_grid_mixed(const PrimalMesh &m, FVGrid &g) {
    const uint32_t N = m.n_cells(), nn = m.n_nodes();
    const uint32_t P = (uint32_t)m.patch_n.size();
    if (m.patch_bc.size() != P) throw MeshError{"patch_bc size mismatch"};
    size_t nb = 0;
    for (uint32_t p = 0; p < P; ++p) nb += m.patch_n[p];
    if (m.ftype.size() != nb || m.foff.size() != nb + 1 || m.coff.size() != (size_t)N + 1)
        throw MeshError{"element offset arrays do not match the counts"};
    for (uint32_t v : m.cells)
        if (v >= nn) throw MeshError{"cell node index out of range"};
    for (uint32_t v : m.patch_tri)
        if (v >= nn) throw MeshError{"patch node index out of range"};

    // every cell face with its key; the reference's std::map becomes a sort
    struct Ent {
        Key k;
        uint32_t cell, face;
    };
    std::vector<Ent> ent;
    ent.reserve((size_t)6 * N);
    for (uint32_t c = 0; c < N; ++c) {
        const int t = m.ctype[c];
        for (int k = 0; k < elem_n_faces(t); ++k) {
            uint32_t fe[4];
            const int n = elem_face(t, &m.cells[m.coff[c]], k, fe);
            ent.push_back({make_key(n, fe), c, (uint32_t)k});
        }
    }
    std::stable_sort(ent.begin(), ent.end(), [](const Ent &a, const Ent &b) { return a.k < b.k; });
    // key order = the reference map's order; pairs are interior faces with
    // i = the first cell that listed the face (FV_Grid.cpp:38-85)
    struct Face {
        uint32_t i, j, n, nodes[4];
    };
    std::vector<Face> faces;
    std::vector<uint8_t> used(ent.size(), 0);  // 1 = interior, 2 = boundary
    for (size_t a = 0; a < ent.size();) {
        size_t b = a + 1;
        while (b < ent.size() && ent[b].k == ent[a].k) ++b;
        if (b - a > 2) throw MeshError{"face shared by more than two cells"};
        if (b - a == 2) {
            const Ent &x = ent[a], &y = ent[a + 1];
            if (x.cell == y.cell) throw MeshError{"degenerate element: cell " + std::to_string(x.cell)};
            Face f{x.cell, y.cell, x.k.n, {0, 0, 0, 0}};
            elem_face(m.ctype[x.cell], &m.cells[m.coff[x.cell]], (int)x.face, f.nodes);
            faces.push_back(f);
            used[a] = used[a + 1] = 1;
        }
        a = b;
    }
    const uint32_t Fint = (uint32_t)faces.size();
    std::stable_sort(faces.begin(), faces.end(),
                     [](const Face &x, const Face &y) { return x.i != y.i ? x.i < y.i : x.j < y.j; });
    // boundary faces per patch, ghosts in patch element order (:93-115)
    g.pfirst.resize(P);
    g.pn.resize(P);
    g.pbc = m.patch_bc;
    uint32_t jg = N;
    size_t e = 0;
    for (uint32_t p = 0; p < P; ++p) {
        g.pfirst[p] = (uint32_t)faces.size();
        g.pn[p] = m.patch_n[p];
        const size_t b0 = faces.size();
        for (uint32_t q = 0; q < m.patch_n[p]; ++q, ++e) {
            const int n = elem_n_nodes(m.ftype[e]);
            const uint32_t *fn = &m.patch_tri[m.foff[e]];
            const Key k = make_key(n, fn);
            auto it = std::lower_bound(ent.begin(), ent.end(), k, [](const Ent &x, const Key &y) { return x.k < y; });
            if (it == ent.end() || !(it->k == k)) throw MeshError{"patch face matches no cell face"};
            const size_t h = (size_t)(it - ent.begin());
            if (used[h]) throw MeshError{"patch face is interior or listed twice"};
            used[h] = 2;
            Face f{it->cell, jg++, (uint32_t)n, {fn[0], fn[1], fn[2], n == 4 ? fn[3] : 0u}};
            faces.push_back(f);
        }
        std::stable_sort(faces.begin() + b0, faces.end(), [](const Face &x, const Face &y) { return x.i < y.i; });
    }
    for (size_t h = 0; h < ent.size(); ++h)
        if (!used[h]) throw MeshError{"boundary face of cell " + std::to_string(ent[h].cell) +
                                      " is not covered by any patch"};

    g.N = N;
    g.NT = jg;
    g.Fint = Fint;
    g.F = (uint32_t)faces.size();
    g.fnodes.clear();
    g.flip.clear();
    // geometry (FV_Grid::calc_geometry_properties, :154-205)
    const double *X = m.nodes.data();
    g.vol.assign(g.NT, 0.0);
    g.cen.assign((size_t)3 * g.NT, 0.0);
    for (uint32_t c = 0; c < N; ++c) {
        double v;
        V3 cc;
        cell_props(m.ctype[c], X, &m.cells[m.coff[c]], v, cc);
        g.vol[c] = v;
        g.cen[3 * (size_t)c] = cc.x;
        g.cen[3 * (size_t)c + 1] = cc.y;
        g.cen[3 * (size_t)c + 2] = cc.z;
    }
    g.ij.resize((size_t)2 * g.F);
    g.S.resize((size_t)3 * g.F);
    g.ri.resize((size_t)3 * g.F);
    g.rj.resize((size_t)3 * g.F);
    for (uint32_t q = 0; q < g.F; ++q) {
        const Face &f = faces[q];
        g.ij[2 * (size_t)q] = f.i;
        g.ij[2 * (size_t)q + 1] = f.j;
        V3 S, fc;
        face_props((int)f.n, X, f.nodes, S, fc);
        const V3 ci = ld(&g.cen[3 * (size_t)f.i]);
        if (f.j >= N) {  // calc_ghost_centroid
            const V3 cg = 2 * fc - ci;
            g.cen[3 * (size_t)f.j] = cg.x;
            g.cen[3 * (size_t)f.j + 1] = cg.y;
            g.cen[3 * (size_t)f.j + 2] = cg.z;
        }
        const V3 cj = ld(&g.cen[3 * (size_t)f.j]);
        const double nd = dot(S, cj - ci);  // calc_face_properties
        if (nd == 0.0) throw MeshError{"face normal orthogonal to the cell-centre line"};
        if (nd < 0) S = -1 * S;
        const V3 r1 = fc - ci, r2 = fc - cj;
        double *Sp = &g.S[3 * (size_t)q], *a = &g.ri[3 * (size_t)q], *b = &g.rj[3 * (size_t)q];
        Sp[0] = S.x, Sp[1] = S.y, Sp[2] = S.z;
        a[0] = r1.x, a[1] = r1.y, a[2] = r1.z;
        b[0] = r2.x, b[1] = r2.y, b[2] = r2.z;
    }
    g.dS.assign((size_t)3 * g.NT, 0.0);  // interior faces only (src/Solver.cpp:276-298)
    for (uint32_t q = 0; q < Fint; ++q) {
        const uint32_t i = g.ij[2 * (size_t)q], j = g.ij[2 * (size_t)q + 1];
        for (int d = 0; d < 3; ++d) {
            const double t = 0.5 * std::fabs(g.S[3 * (size_t)q + d]);
            g.dS[3 * (size_t)i + d] += t;
            g.dS[3 * (size_t)j + d] += t;
        }
    }
}

}  // namespace nf
