// Mixed-element (polyhedral) path of the FP64 step: hexahedra, wedges,
// pyramids and tetrahedra in one mesh, with triangle and quadrilateral
// faces.  Geometry comes precomputed from the host face graph (S, r_i, r_j,
// volumes, Delta S in the reference's face order); one thread per cell
// gathers its faces in increasing face index, so every sum runs in the
// reference's scatter order and the result is bit-identical to the CPU
// solver.  Single mesh, FP64, global time step.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/netflux.h"
#include "mesh.hpp"

namespace nf {

struct PolyPath;

// NF_OK or an NF_E_* status with the message in err
int poly_create(const PrimalMesh &pm, const FVGrid &g, const nf_params &p, hipStream_t st, PolyPath **out,
                std::string &err);
void poly_destroy(PolyPath *P);
nf_fv_sizes poly_sizes(const PolyPath *P);
int poly_set_freestream(PolyPath *P, std::string &err);
int poly_set_state(PolyPath *P, const double *U, const double *U_old, std::string &err);
int poly_calc_timestep(PolyPath *P, double *dt, std::string &err);
int poly_step(PolyPath *P, double dt, std::string &err);
int poly_eval_flux_balance(PolyPath *P, std::string &err);
int poly_get_state(PolyPath *P, double *U, std::string &err);
int poly_run(PolyPath *P, const nf_run_opts *o, nf_run_stats *stats, std::string &err);
int poly_get_residuals(PolyPath *P, double *out, int32_t max, int32_t *n_out);
int poly_time_kernels(PolyPath *P, int32_t n_steps, nf_kernel_time *out, int32_t max, int32_t *n_out,
                      std::string &err);
int poly_get_field(PolyPath *P, int32_t which, double *out, std::string &err);

}  // namespace nf
