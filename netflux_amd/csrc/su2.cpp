// SU2 mesh files for the tetrahedral path: reader with the reference's
// conventions (PrimalGrid::read_su2_mesh, src/geometry/PrimalGrid.cpp:140-244)
// and a writer for round trips and generated test cases.  CPU only.
//
//   NDIME= 3
//   NELEM= n            then n lines: 10 a b c d [index]   (Salome node order)
//   NPOIN= n            then n lines: x y z [index]
//   NMARK= m            then per marker: MARKER_TAG= name, MARKER_ELEMS= k,
//                       k lines: 5 a b c
//
// Tetrahedra arrive in Salome order and are stored in VTK order by swapping
// nodes 1 and 2 (salome_to_vtk_connectivity, src/geometry/
// GeometryUtilities.cpp:443-466); triangles are stored as read.  Only
// tetrahedra and triangles are accepted: the device path is tetrahedral.
#include <cstdio>
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
constexpr int VTK_TRIANGLE = 5, VTK_TETRA = 10;

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

void read_su2(const std::string &path, PrimalMesh &m, std::vector<std::string> &names) {
    std::ifstream is(path);
    if (!is) throw MeshError("Couldn't open the mesh file: " + path);
    if (keyword<int>(is, "NDIME=") != 3) throw MeshError("Only three dimensions (NDIME=3) is accepted");
    const long ne = keyword<long>(is, "NELEM=");
    if (ne <= 0) throw MeshError("NELEM must be positive");
    m.cells.resize(4 * (size_t)ne);
    std::istringstream ls;
    for (long e = 0; e < ne; ++e) {
        if (!next_line(is, ls)) throw MeshError("unexpected end of element list");
        int t = 0;
        ls >> t;
        if (t != VTK_TETRA)
            throw MeshError("element type " + std::to_string(t) +
                            " in the volume mesh: only tetrahedra (10) are supported");
        uint32_t n[4];
        if (!(ls >> n[0] >> n[1] >> n[2] >> n[3])) throw MeshError("bad tetrahedron line");
        // Salome -> VTK
        m.cells[4 * e] = n[0];
        m.cells[4 * e + 1] = n[2];
        m.cells[4 * e + 2] = n[1];
        m.cells[4 * e + 3] = n[3];
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
    m.patch_n.clear();
    m.patch_tri.clear();
    names.clear();
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
            int t = 0;
            ls >> t;
            if (t != VTK_TRIANGLE)
                throw MeshError("marker element type " + std::to_string(t) + " in marker " + name +
                                ": only triangles (5) are supported");
            uint32_t a, b, c;
            if (!(ls >> a >> b >> c)) throw MeshError("bad marker element line");
            if (a >= (uint64_t)np || b >= (uint64_t)np || c >= (uint64_t)np)
                throw MeshError("marker node index out of range");
            m.patch_tri.insert(m.patch_tri.end(), {a, b, c});
        }
        m.patch_n.push_back((uint32_t)nf);
        names.push_back(name);
    }
}

void write_su2(const std::string &path, const PrimalMesh &m, const std::vector<std::string> &names) {
    std::FILE *f = std::fopen(path.c_str(), "w");
    if (!f) throw MeshError("Couldn't open " + path + " for writing");
    std::fprintf(f, "NDIME= 3\nNELEM= %u\n", m.n_cells());
    for (uint32_t e = 0; e < m.n_cells(); ++e) {
        const uint32_t *n = &m.cells[4 * (size_t)e];
        std::fprintf(f, "%d %u %u %u %u %u\n", VTK_TETRA, n[0], n[2], n[1], n[3], e);  // VTK -> Salome
    }
    std::fprintf(f, "NPOIN= %u\n", m.n_nodes());
    for (uint32_t i = 0; i < m.n_nodes(); ++i)
        std::fprintf(f, "%.17g %.17g %.17g %u\n", m.nodes[3 * (size_t)i], m.nodes[3 * (size_t)i + 1],
                     m.nodes[3 * (size_t)i + 2], i);
    std::fprintf(f, "NMARK= %zu\n", m.patch_n.size());
    size_t t = 0;
    for (size_t q = 0; q < m.patch_n.size(); ++q) {
        std::fprintf(f, "MARKER_TAG= %s\nMARKER_ELEMS= %u\n", names[q].c_str(), m.patch_n[q]);
        for (uint32_t k = 0; k < m.patch_n[q]; ++k, ++t)
            std::fprintf(f, "%d %u %u %u\n", VTK_TRIANGLE, m.patch_tri[3 * t], m.patch_tri[3 * t + 1],
                         m.patch_tri[3 * t + 2]);
    }
    if (std::fclose(f) != 0) throw MeshError("write error on " + path);
}

}  // namespace nf
