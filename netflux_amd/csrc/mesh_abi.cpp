// C ABI of the host mesh utilities (include/netflux.h): generators and the
// finite-volume face graph export.  CPU only.
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/netflux.h"
#include "mesh.hpp"

struct nf_mesh {
    nf::PrimalMesh pm;
    nf::FVGrid fv;
    bool fv_built = false;
    std::vector<std::string> names;  // patch names (SU2 marker tags)
};

namespace {
thread_local std::string g_mesh_err;

int ensure_fv(nf_mesh *m) {
    if (m->fv_built) return NF_OK;
    try {
        nf::build_fv_grid(m->pm, m->fv);
    } catch (const nf::MeshError &e) {
        g_mesh_err = e.msg;
        return NF_E_MESH;
    }
    m->fv_built = true;
    return NF_OK;
}
}  // namespace

extern "C" {

int nf_mesh_box(uint32_t nx, uint32_t ny, uint32_t nz, double lx, double ly, double lz,
                double jitter, uint64_t seed, double bump_h, double bump_sigma,
                const int32_t patch_bc[6], nf_mesh **out) {
    if (!out || !patch_bc) return NF_E_ARG;
    *out = nullptr;
    if (!(lx > 0 && ly > 0 && lz > 0) || !(jitter >= 0 && jitter < 0.5)) return NF_E_ARG;
    if (bump_h > 0 && !(bump_sigma > 0)) return NF_E_ARG;
    for (int s = 0; s < 6; ++s)
        if (patch_bc[s] < 0 || patch_bc[s] > 2) return NF_E_ARG;
    nf_mesh *m = new (std::nothrow) nf_mesh;
    if (!m) return NF_E_ARG;
    try {
        nf::build_box(nx, ny, nz, lx, ly, lz, jitter, seed, bump_h, bump_sigma, patch_bc, m->pm);
    } catch (const nf::MeshError &) {
        delete m;
        return NF_E_ARG;
    }
    *out = m;
    return NF_OK;
}

int nf_mesh_from_desc(const nf_mesh_desc *d, nf_mesh **out) {
    if (!d || !out) return NF_E_ARG;
    *out = nullptr;
    nf_mesh *m = new (std::nothrow) nf_mesh;
    if (!m) return NF_E_ARG;
    try {
        nf::primal_from_desc(*d, m->pm);
    } catch (const nf::MeshError &e) {
        g_mesh_err = e.msg;
        delete m;
        return NF_E_MESH;
    }
    *out = m;
    return NF_OK;
}

int nf_mesh_get_desc(const nf_mesh *m, nf_mesh_desc *d) {
    if (!m || !d) return NF_E_ARG;
    std::memset(d, 0, sizeof *d);
    d->n_nodes = m->pm.n_nodes();
    d->nodes = m->pm.nodes.data();
    d->n_cells = m->pm.n_cells();
    d->cells = m->pm.cells.data();
    d->n_patches = (uint32_t)m->pm.patch_n.size();
    d->patch_n_faces = m->pm.patch_n.data();
    d->patch_faces = m->pm.patch_tri.data();
    d->patch_bc = m->pm.patch_bc.data();
    if (m->pm.mixed()) {
        d->cell_types = m->pm.ctype.data();
        d->cell_offsets = m->pm.coff.data();
        d->patch_face_types = m->pm.ftype.data();
        d->patch_face_offsets = m->pm.foff.data();
    }
    return NF_OK;
}

void nf_mesh_destroy(nf_mesh *m) { delete m; }

const char *nf_mesh_last_error(void) { return g_mesh_err.c_str(); }

int nf_mesh_read_su2(const char *path, int32_t n_names, const char *const *names, const int32_t *bcs,
                     nf_mesh **out) {
    if (!path || !out || n_names < 0 || (n_names > 0 && (!names || !bcs))) return NF_E_ARG;
    *out = nullptr;
    nf_mesh *m = new (std::nothrow) nf_mesh;
    if (!m) return NF_E_ARG;
    try {
        nf::read_su2(path, m->pm, m->names);
        // boundary type of each marker from the patch list of the input file
        // (PrimalGrid::read_su2_mesh rejects a marker the input does not name)
        for (const std::string &name : m->names) {
            int32_t bc = -1;
            for (int32_t i = 0; i < n_names; ++i)
                if (name == names[i]) bc = bcs[i];
            if (bc < 0) throw nf::MeshError("Patch with name '" + name + "' is not named in input file");
            if (bc > 2) throw nf::MeshError("bad boundary type for patch '" + name + "'");
            m->pm.patch_bc.push_back(bc);
        }
    } catch (const nf::MeshError &e) {
        g_mesh_err = e.msg;
        delete m;
        return NF_E_MESH;
    }
    *out = m;
    return NF_OK;
}

int nf_mesh_write_su2(const nf_mesh *m, const char *path, int32_t n_names, const char *const *names) {
    if (!m || !path || n_names < 0 || (n_names > 0 && !names)) return NF_E_ARG;
    std::vector<std::string> tags;
    for (size_t q = 0; q < m->pm.patch_n.size(); ++q) {
        if ((int32_t)q < n_names && names[q]) tags.push_back(names[q]);
        else if (q < m->names.size()) tags.push_back(m->names[q]);
        else tags.push_back("patch_" + std::to_string(q));
        if (tags.back().empty() || tags.back().find_first_of(" \t\n") != std::string::npos) return NF_E_ARG;
    }
    try {
        nf::write_su2(path, m->pm, tags);
    } catch (const nf::MeshError &e) {
        g_mesh_err = e.msg;
        return NF_E_MESH;
    }
    return NF_OK;
}

int nf_mesh_patch_name(const nf_mesh *m, int32_t i, char *buf, int32_t len) {
    if (!m || !buf || len < 1 || i < 0 || i >= (int32_t)m->pm.patch_n.size()) return NF_E_ARG;
    const std::string s = i < (int32_t)m->names.size() ? m->names[i] : "patch_" + std::to_string(i);
    std::snprintf(buf, (size_t)len, "%s", s.c_str());
    return NF_OK;
}

int nf_fv_sizes_of(const nf_mesh *mc, nf_fv_sizes *s) {
    if (!mc || !s) return NF_E_ARG;
    nf_mesh *m = const_cast<nf_mesh *>(mc);
    int rc = ensure_fv(m);
    if (rc) return rc;
    const nf::FVGrid &g = m->fv;
    *s = {g.N, g.NT, g.Fint, g.F, (uint32_t)g.pfirst.size()};
    return NF_OK;
}

int nf_fv_export(const nf_mesh *mc, uint32_t *cell_ij, double *S, double *r_i, double *r_j,
                 double *volume, double *centroid, uint32_t *patch_first, uint32_t *patch_nfaces,
                 double *delta_S) {
    if (!mc) return NF_E_ARG;
    nf_mesh *m = const_cast<nf_mesh *>(mc);
    int rc = ensure_fv(m);
    if (rc) return rc;
    const nf::FVGrid &g = m->fv;
    auto cp = [](void *dst, const void *src, size_t bytes) {
        if (dst) std::memcpy(dst, src, bytes);
    };
    cp(cell_ij, g.ij.data(), g.ij.size() * 4);
    cp(S, g.S.data(), g.S.size() * 8);
    cp(r_i, g.ri.data(), g.ri.size() * 8);
    cp(r_j, g.rj.data(), g.rj.size() * 8);
    cp(volume, g.vol.data(), g.vol.size() * 8);
    cp(centroid, g.cen.data(), g.cen.size() * 8);
    cp(patch_first, g.pfirst.data(), g.pfirst.size() * 4);
    cp(patch_nfaces, g.pn.data(), g.pn.size() * 4);
    cp(delta_S, g.dS.data(), g.dS.size() * 8);
    return NF_OK;
}

}  // extern "C"
