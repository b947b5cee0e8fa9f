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

template <int B>
static void sweepA_b(int order, dim3 g, hipStream_t st, const DevMesh &m, const SweepAArgsMx &a) {
    if (order == 0) hipLaunchKernelGGL((k_sweepA_mx<0, B>), g, dim3(B), 0, st, m, a);
    else if (order == 1) hipLaunchKernelGGL((k_sweepA_mx<1, B>), g, dim3(B), 0, st, m, a);
    else hipLaunchKernelGGL((k_sweepA_mx<2, B>), g, dim3(B), 0, st, m, a);
}

hipError_t launch_sweepA_mx(int order, int block, dim3 g, hipStream_t st, const DevMesh &m, const SweepAArgsMx &a) {
    if (block == 128) sweepA_b<128>(order, g, st, m, a);
    else sweepA_b<256>(order, g, st, m, a);
    return hipGetLastError();
}

hipError_t launch_sweepB_mx(int block, dim3 g, hipStream_t st, const DevMesh &m, const SweepBArgsMx &a) {
    if (block == 128) hipLaunchKernelGGL(k_sweepB_mx<128>, g, dim3(128), 0, st, m, a);
    else hipLaunchKernelGGL(k_sweepB_mx<256>, g, dim3(256), 0, st, m, a);
    return hipGetLastError();
}

hipError_t launch_prim_mx(dim3 g, hipStream_t st, const double *U, float *V, uint32_t N) {
    hipLaunchKernelGGL(k_prim_mx, g, dim3(BLOCK), 0, st, U, V, N);
    return hipGetLastError();
}

hipError_t launch_apply_mx(dim3 g, hipStream_t st, const DevMesh &m, const ApplyArgs &a) {
    hipLaunchKernelGGL(k_apply_mx, g, dim3(BLOCK), 0, st, m, a);
    return hipGetLastError();
}

}  // namespace mx
}  // namespace nf
