// SU2 mesh files: reader with the reference's conventions
// (PrimalGrid::read_su2_mesh, src/geometry/PrimalGrid.cpp:140-244) and a
// writer for round trips and generated test cases.  CPU only.
//
//   NDIME= 3
//   NELEM= n            then n lines: type nodes... [index]   (Salome node order)
//   NPOIN= n            then n lines: x y z [index]
//   NMARK= m            then per marker: MARKER_TAG= name, MARKER_ELEMS= k,
//                       k lines: type nodes...
//
// Volume elements: tetrahedra (10), hexahedra (12), wedges (13), pyramids
// (14); marker elements: triangles (5), quadrilaterals (9).  Elements arrive
// in Salome order and are stored in VTK order (salome_to_vtk_connectivity,
// src/geometry/GeometryUtilities.cpp:443-466: tetrahedron nodes 1<->2,
// hexahedron 1<->3 and 5<->7, pyramid 1<->3); marker elements are stored as
// read.  A file of only tetrahedra and triangles gives the tetrahedral layout.
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/netflux.h"
#include "mesh.hpp"

namespace nf {

namespace {

// next meaningful token stream line: skips blank lines and '%' comments
bool next_line(std::istream &is, std::istringstream &ls) {
    std::string line;
    while (std::getline(is, line)) {
        const size_t p = line.find_first_not_of(" \t\r");
        if (p == std::string::npos || line[p] == '%') continue;
        ls.clear();
        ls.str(line);
        return true;
    }
    return false;
}

template <typename T>
T keyword(std::istream &is, const char *key) {
    std::istringstream ls;
    if (!next_line(is, ls)) throw MeshError(std::string("missing ") + key);
    std::string k;
    ls >> k;
    // "KEY= value" or "KEY=value"
    std::string rest;
    if (k.size() > std::strlen(key) && k.compare(0, std::strlen(key), key) == 0) {
        rest = k.substr(std::strlen(key));
    } else {
        if (k != key) throw MeshError("symbol '" + k + "' parsed instead of correct symbol '" + key + "'");
        ls >> rest;
    }
    std::istringstream rs(rest);
    T v{};
    if (!(rs >> v)) throw MeshError(std::string("bad value after ") + key);
    return v;
}
}  // namespace

// Salome <-> VTK node order (the swaps are their own inverse)
static void salome_vtk(int t, uint32_t *n) {
    if (t == ELEM_TET) std::swap(n[1], n[2]);
    else if (t == ELEM_HEX) std::swap(n[1], n[3]), std::swap(n[5], n[7]);
    else if (t == ELEM_PYRAMID) std::swap(n[1], n[3]);
}

void read_su2(const std::string &path, PrimalMesh &m, std::vector<std::string> &names) {
    std::ifstream is(path);
    if (!is) throw MeshError("Couldn't open the mesh file: " + path);
    if (keyword<int>(is, "NDIME=") != 3) throw MeshError("Only three dimensions (NDIME=3) is accepted");
    const long ne = keyword<long>(is, "NELEM=");
    if (ne <= 0) throw MeshError("NELEM must be positive");
    m = PrimalMesh{};
    m.coff.push_back(0);
    std::istringstream ls;
    bool all_tet = true;
    for (long e = 0; e < ne; ++e) {
        if (!next_line(is, ls)) throw MeshError("unexpected end of element list");
        int t = -1;
        ls >> t;
        const int nn = elem_n_nodes(t);
        if (nn == 0)
            throw MeshError("Illegal su2 volume element identifier (" + std::to_string(t) +
                            ") detected in su2 mesh file");
        if (elem_n_faces(t) == 0)
            throw MeshError("Face element of type " + std::to_string(t) + " detected in su2 mesh file in the volume mesh");
        uint32_t n[8];
        for (int q = 0; q < nn; ++q)
            if (!(ls >> n[q])) throw MeshError("bad element line");
        salome_vtk(t, n);
        m.cells.insert(m.cells.end(), n, n + nn);
        m.ctype.push_back((uint8_t)t);
        m.coff.push_back((uint32_t)m.cells.size());
        all_tet = all_tet && t == ELEM_TET;
    }
    const long np = keyword<long>(is, "NPOIN=");
    if (np <= 0) throw MeshError("NPOIN must be positive");
    m.nodes.resize(3 * (size_t)np);
    for (long i = 0; i < np; ++i) {
        if (!next_line(is, ls)) throw MeshError("unexpected end of point list");
        if (!(ls >> m.nodes[3 * i] >> m.nodes[3 * i + 1] >> m.nodes[3 * i + 2])) throw MeshError("bad point line");
    }
    for (uint32_t v : m.cells)
        if (v >= (uint64_t)np) throw MeshError("element node index out of range");
    const long nm = keyword<long>(is, "NMARK=");
    names.clear();
    m.foff.push_back(0);
    for (long q = 0; q < nm; ++q) {
        std::istringstream tl;
        if (!next_line(is, tl)) throw MeshError("missing MARKER_TAG=");
        std::string k, name;
        tl >> k;
        if (k.compare(0, 11, "MARKER_TAG=") != 0)
            throw MeshError("symbol '" + k + "' parsed instead of correct symbol 'MARKER_TAG='");
        name = k.size() > 11 ? k.substr(11) : "";
        if (name.empty()) tl >> name;
        const long nf = keyword<long>(is, "MARKER_ELEMS=");
        for (long f = 0; f < nf; ++f) {
            if (!next_line(is, ls)) throw MeshError("unexpected end of marker " + name);
            int t = -1;
            ls >> t;
            const int nn = elem_n_nodes(t);
            if (nn == 0)
                throw MeshError("Illegal su2 element identifier (" + std::to_string(t) + ") detected in su2 mesh file");
            if (elem_n_faces(t) != 0)
                throw MeshError("Volume element of type " + std::to_string(t) +
                                " detected in su2 mesh file among marker elements (marker " + name + ")");
            uint32_t n[4];
            for (int r = 0; r < nn; ++r)
                if (!(ls >> n[r])) throw MeshError("bad marker element line");
            for (int r = 0; r < nn; ++r)
                if (n[r] >= (uint64_t)np) throw MeshError("marker node index out of range");
            m.patch_tri.insert(m.patch_tri.end(), n, n + nn);
            m.ftype.push_back((uint8_t)t);
            m.foff.push_back((uint32_t)m.patch_tri.size());
            all_tet = all_tet && t == ELEM_TRI;
        }
        m.patch_n.push_back((uint32_t)nf);
        names.push_back(name);
    }
    if (all_tet) m.ctype.clear(), m.coff.clear(), m.ftype.clear(), m.foff.clear();
}

void write_su2(const std::string &path, const PrimalMesh &m, const std::vector<std::string> &names) {
    std::FILE *f = std::fopen(path.c_str(), "w");
    if (!f) throw MeshError("Couldn't open " + path + " for writing");
    std::fprintf(f, "NDIME= 3\nNELEM= %u\n", m.n_cells());
    for (uint32_t e = 0; e < m.n_cells(); ++e) {
        const int t = m.mixed() ? (int)m.ctype[e] : (int)ELEM_TET;
        const uint32_t *src = m.mixed() ? &m.cells[m.coff[e]] : &m.cells[4 * (size_t)e];
        uint32_t n[8];
        const int nn = elem_n_nodes(t);
        std::copy(src, src + nn, n);
        salome_vtk(t, n);  // VTK -> Salome
        std::fprintf(f, "%d", t);
        for (int q = 0; q < nn; ++q) std::fprintf(f, " %u", n[q]);
        std::fprintf(f, " %u\n", e);
    }
    std::fprintf(f, "NPOIN= %u\n", m.n_nodes());
    for (uint32_t i = 0; i < m.n_nodes(); ++i)
        std::fprintf(f, "%.17g %.17g %.17g %u\n", m.nodes[3 * (size_t)i], m.nodes[3 * (size_t)i + 1],
                     m.nodes[3 * (size_t)i + 2], i);
    std::fprintf(f, "NMARK= %zu\n", m.patch_n.size());
    size_t t = 0;
    for (size_t q = 0; q < m.patch_n.size(); ++q) {
        std::fprintf(f, "MARKER_TAG= %s\nMARKER_ELEMS= %u\n", names[q].c_str(), m.patch_n[q]);
        for (uint32_t k = 0; k < m.patch_n[q]; ++k, ++t) {
            const int ty = m.mixed() ? (int)m.ftype[t] : (int)ELEM_TRI;
            const uint32_t *src = m.mixed() ? &m.patch_tri[m.foff[t]] : &m.patch_tri[3 * t];
            std::fprintf(f, "%d", ty);
            for (int r = 0; r < elem_n_nodes(ty); ++r) std::fprintf(f, " %u", src[r]);
            std::fprintf(f, "\n");
        }
    }
    if (std::fclose(f) != 0) throw MeshError("write error on " + path);
}

}  // namespace nf
