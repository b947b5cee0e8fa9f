// Mixed-precision kernels (kernels_mixed.hpp) in their own translation unit,
// compiled with floating-point contraction enabled; the FP64 kernels live in
// solver.hip, compiled without it.
#define NF_DEVICE_HELPERS_ONLY
#define NF_MIXED_KERNELS
#include "kernels_mixed.hpp"

namespace nf {
namespace mx {

hipError_t launch_ghost_mx(bool face, dim3 g, hipStream_t st, const DevMesh &m, const GhostArgsMx &a) {
    if (face) hipLaunchKernelGGL(k_ghost_mx<true>, g, dim3(BLOCK), 0, st, m, a);
    else hipLaunchKernelGGL(k_ghost_mx<false>, g, dim3(BLOCK), 0, st, m, a);
    return hipGetLastError();
}

hipError_t launch_sweepA_mx(int order, dim3 g, hipStream_t st, const DevMesh &m, const SweepAArgsMx &a) {
    if (order == 0) hipLaunchKernelGGL(k_sweepA_mx<0>, g, dim3(BLOCK), 0, st, m, a);
    else if (order == 1) hipLaunchKernelGGL(k_sweepA_mx<1>, g, dim3(BLOCK), 0, st, m, a);
    else hipLaunchKernelGGL(k_sweepA_mx<2>, g, dim3(BLOCK), 0, st, m, a);
    return hipGetLastError();
}

hipError_t launch_sweepB_mx(dim3 g, hipStream_t st, const DevMesh &m, const SweepBArgsMx &a) {
    hipLaunchKernelGGL(k_sweepB_mx, g, dim3(BLOCK), 0, st, m, a);
    return hipGetLastError();
}

hipError_t launch_prim_mx(dim3 g, hipStream_t st, const double *U, float *V, uint32_t N) {
    hipLaunchKernelGGL(k_prim_mx, g, dim3(BLOCK), 0, st, U, V, N);
    return hipGetLastError();
}

}  // namespace mx
}  // namespace nf
