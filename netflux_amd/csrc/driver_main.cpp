// netflux_driver: runs NetFlux's time loop on the GPU through the C++ host
// classes (netflux_solver.hpp), on a generated box mesh or an SU2 file.
//
//   netflux_driver <n | su2:path> <steps> [sod|pulse] [state_out.bin]
//
// sod:   unit cube n^3 hexes x 6 tets, slip walls, Sod tube along x
// pulse: same mesh jittered by 15%, farfield, Mach 0.5 + Gaussian pulse
// su2:path reads that mesh (any element types; every marker a slip wall)
// and applies the Sod state by cell centroid.
// Prints one JSON line (cells, steps, time, sum of rho*vol); with a fourth
// argument writes U (AoS [N][5] doubles, reference cell order) to that file.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "netflux_solver.hpp"

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <n | su2:path> <steps> [sod|pulse] [state_out.bin]\n", argv[0]);
        return 2;
    }
    const uint32_t n = (uint32_t)std::atoi(argv[1]);
    const int steps = std::atoi(argv[2]);
    const std::string cas = argc > 3 ? argv[3] : "sod";
    const bool pulse = cas == "pulse";
    const int32_t bc_slip[6] = {NF_BC_SLIP_WALL, NF_BC_SLIP_WALL, NF_BC_SLIP_WALL,
                                NF_BC_SLIP_WALL, NF_BC_SLIP_WALL, NF_BC_SLIP_WALL};
    const int32_t bc_ff[6] = {NF_BC_FARFIELD, NF_BC_FARFIELD, NF_BC_FARFIELD,
                              NF_BC_FARFIELD, NF_BC_FARFIELD, NF_BC_FARFIELD};
    nf_mesh *mesh = nullptr;
    const std::string arg1 = argv[1];
    if (arg1.compare(0, 4, "su2:") == 0) {
        // marker tags as they appear in the file, each a slip wall
        std::ifstream is(arg1.substr(4));
        std::vector<std::string> tags;
        for (std::string tok; is >> tok;) {
            if (tok.compare(0, 11, "MARKER_TAG=") != 0) continue;
            std::string name = tok.substr(11);
            if (name.empty()) is >> name;
            tags.push_back(name);
        }
        std::vector<const char *> names;
        for (const std::string &t : tags) names.push_back(t.c_str());
        const std::vector<int32_t> bcs(tags.size(), NF_BC_SLIP_WALL);
        if (nf_mesh_read_su2(arg1.c_str() + 4, (int32_t)names.size(), names.data(), bcs.data(), &mesh)) {
            std::fprintf(stderr, "%s\n", nf_mesh_last_error());
            return 1;
        }
    } else if (nf_mesh_box(n, n, n, 1, 1, 1, pulse ? 0.15 : 0.0, 20260101ull, 0, 0.2, pulse ? bc_ff : bc_slip,
                           &mesh)) {
        std::fprintf(stderr, "mesh generation failed\n");
        return 1;
    }
    nf_mesh_desc desc;
    nf_mesh_get_desc(mesh, &desc);
    nf_fv_sizes sz;
    nf_fv_sizes_of(mesh, &sz);
    std::vector<double> vol(sz.n_total_cells), cen(3 * (size_t)sz.n_total_cells);
    nf_fv_export(mesh, nullptr, nullptr, nullptr, nullptr, vol.data(), cen.data(), nullptr, nullptr, nullptr);

    netflux::Config cfg;
    cfg.n_timesteps = (size_t)steps;
    const double g = 1.4;
    if (pulse) {
        const double Vinf[5] = {1.0, 0.5, 0.0, 0.0, 1.0 / g};
        std::memcpy(cfg.primvars_inf, Vinf, sizeof Vinf);
    } else {
        const double Vinf[5] = {1.0, 0.0, 0.0, 0.0, 1.0};
        std::memcpy(cfg.primvars_inf, Vinf, sizeof Vinf);
    }
    const uint32_t N = desc.n_cells;
    std::vector<double> U(5 * (size_t)N);
    for (uint32_t c = 0; c < N; ++c) {
        double V[5];
        const double *x = &cen[3 * (size_t)c];
        if (pulse) {
            const double dx = x[0] - 0.5, dy = x[1] - 0.5, dz = x[2] - 0.5;
            V[0] = 1.0 + 0.1 * std::exp(-(dx * dx + dy * dy + dz * dz) / (2 * 0.1 * 0.1));
            V[1] = 0.5;
            V[2] = V[3] = 0.0;
            V[4] = 1.0 / g;
        } else {
            const bool left = x[0] < 0.5;
            V[0] = left ? 1.0 : 0.125;
            V[1] = V[2] = V[3] = 0.0;
            V[4] = left ? 1.0 : 0.1;
        }
        double *u = &U[5 * (size_t)c];
        u[0] = V[0];
        u[1] = V[0] * V[1];
        u[2] = V[0] * V[2];
        u[3] = V[0] * V[3];
        u[4] = (1 / (g - 1)) * V[4] + 0.5 * V[0] * ((V[1] * V[1] + V[2] * V[2]) + V[3] * V[3]);
    }
    try {
        netflux::Driver driver(cfg, desc);
        driver.solver().set_state(U);
        driver.solve();
        const netflux::VecField Uf = driver.solver().get_solver_data().get_solution();
        double mass = 0.0;
        for (uint32_t c = 0; c < N; ++c) mass += Uf(c, 0) * vol[c];
        std::printf("{\"cells\": %u, \"steps\": %zu, \"time\": %.17g, \"mass\": %.17g}\n", N,
                    cfg.get_timestep(), cfg.get_time(), mass);
        if (argc > 4) {
            FILE *f = std::fopen(argv[4], "wb");
            if (!f || std::fwrite(Uf.data.data(), sizeof(double), Uf.data.size(), f) != Uf.data.size()) {
                std::fprintf(stderr, "cannot write %s\n", argv[4]);
                return 1;
            }
            std::fclose(f);
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        nf_mesh_destroy(mesh);
        return 1;
    }
    nf_mesh_destroy(mesh);
    return 0;
}
