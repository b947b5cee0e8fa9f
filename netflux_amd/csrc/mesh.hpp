// Host mesh pipeline: primal tetrahedral mesh -> finite-volume face graph in
// the reference's face order -> per-cell device layout.
//
// The face graph follows FV_Grid (src/geometry/FV_Grid.cpp:21-267): interior
// faces sorted by (i,j) with i<j, then one block per patch in patch order,
// ghost cells numbered sequentially per patch; the geometry uses the
// reference's tet/triangle formulas (src/geometry/GeometryUtilities.cpp:
// 289-329) evaluated in the same operation order, so every face vector is
// bit-identical to what the reference computes from the same nodes.
// Matching uses buckets keyed by the smallest node of a face instead of the
// reference's std::map (O(F) instead of O(F log F)).
#pragma once
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/netflux.h"

namespace nf {

struct PrimalMesh {
    std::vector<double> nodes;        // [n_nodes][3]
    std::vector<uint32_t> cells;      // [n_cells][4] VTK order (mixed: element nodes, see coff)
    std::vector<uint32_t> patch_n;    // [P]
    std::vector<uint32_t> patch_tri;  // [sum][3] (mixed: face nodes, see foff)
    std::vector<int32_t> patch_bc;    // [P]
    // Mixed-element meshes only; empty when every cell is a tetrahedron and
    // every patch face a triangle.  ctype = VTK type, coff/foff = offsets
    // into cells/patch_tri ([n+1]).
    std::vector<uint8_t> ctype, ftype;
    std::vector<uint32_t> coff, foff;
    bool mixed() const { return !ctype.empty(); }
    uint32_t n_nodes() const { return (uint32_t)(nodes.size() / 3); }
    uint32_t n_cells() const { return mixed() ? (uint32_t)ctype.size() : (uint32_t)(cells.size() / 4); }
};

// Element tables of the reference (include/geometry/GeometryUtilities.hpp:
// 8-32, src/geometry/GeometryUtilities.cpp:48-240), VTK numbering.
enum : int { ELEM_TRI = 5, ELEM_QUAD = 9, ELEM_TET = 10, ELEM_HEX = 12, ELEM_WEDGE = 13, ELEM_PYRAMID = 14 };
int elem_n_nodes(int type);  // 0 for an unknown type
int elem_n_faces(int type);  // volume elements
// face k of a volume element: its node count (3 or 4) and nodes
int elem_face(int type, const uint32_t *ve, int k, uint32_t fe[4]);

struct FVGrid {
    uint32_t N = 0, NT = 0, Fint = 0, F = 0;
    std::vector<uint32_t> ij;         // [F][2]
    std::vector<uint32_t> fnodes;     // [F][3] face element node order
    std::vector<double> S, ri, rj;    // [F][3]
    std::vector<double> vol;          // [NT]
    std::vector<double> cen;          // [NT][3]
    std::vector<uint32_t> pfirst, pn; // [P]
    std::vector<int32_t> pbc;         // [P]
    std::vector<uint8_t> flip;        // [F] 1 if S was negated
    std::vector<double> dS;           // [NT][3] interior-face half |S| sums
};

// Device layout, one record per device cell.
// Slot s of a cell holds its s-th face in increasing reference face index,
// so gathers add contributions in the reference's scatter order.
struct CellLayout {
    uint32_t N = 0;
    std::vector<uint32_t> perm;   // device cell -> reference cell
    std::vector<uint32_t> iperm;  // reference cell -> device cell
    std::vector<uint32_t> cnodes; // [N][4] tet node ids (VTK order)
    std::vector<uint32_t> nbr;    // [N][4] neighbour words (see below)
    std::vector<uint32_t> fcode;  // [N] 4 x 8-bit face codes
    std::vector<uint32_t> ghost;  // [N][4] reference ghost id (boundary slots)
    std::vector<uint32_t> bface;  // [nb] device cell | slot << 29, per boundary face
    std::vector<uint8_t> bbc;     // [nb] boundary type of that face
    std::vector<uint32_t> bfidx;  // [nb] its reference face index
};

// neighbour word: bit31 = boundary.  interior: bits 29-30 neighbour slot,
// bits 0-28 neighbour device cell.  boundary: bits 0-28 boundary face index
// (position in bface/bbc, numbered in device cell then slot order).
constexpr uint32_t NBR_BOUNDARY = 0x80000000u;
constexpr uint32_t NBR_INDEX_MASK = 0x1FFFFFFFu;
// face code byte: bits 0-1,2-3,4-5 local index of face nodes a,b,c;
// bit 6 = S negated; bit 7 = this cell is the reference owner i.
constexpr uint32_t FC_FLIP = 0x40u;
constexpr uint32_t FC_OWNER = 0x80u;

struct MeshError {
    std::string msg;
    MeshError(std::string m = std::string()) : msg(std::move(m)) {}
};

// SU2 files (su2.cpp): tetrahedra + triangle markers; marker names in file order
void read_su2(const std::string &path, PrimalMesh &m, std::vector<std::string> &names);
void write_su2(const std::string &path, const PrimalMesh &m, const std::vector<std::string> &names);

void build_box(uint32_t nx, uint32_t ny, uint32_t nz, double lx, double ly,
               double lz, double jitter, uint64_t seed, double bump_h,
               double bump_sigma, const int32_t bc[6], PrimalMesh &out);

// Throws MeshError on inconsistent input.  A mixed-element mesh fills
// everything but fnodes and flip (which only the tetrahedral layout uses).
void build_fv_grid(const PrimalMesh &m, FVGrid &g);
void build_fv_grid_mixed(const PrimalMesh &m, FVGrid &g);

// Primal mesh from the C descriptor (either layout); a mixed description of
// an all-tetrahedron mesh is stored in the tetrahedral layout.
void primal_from_desc(const nf_mesh_desc &d, PrimalMesh &m);

// order: 0 = input order, 1 = reverse Cuthill-McKee, 2 = Morton (Z-order of centroids).
void build_cell_layout(const PrimalMesh &m, const FVGrid &g, int order,
                       CellLayout &L);

std::vector<uint32_t> rcm_order(const FVGrid &g);
std::vector<uint32_t> morton_order(const FVGrid &g);

}  // namespace nf
