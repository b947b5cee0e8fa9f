// Internal interface of the warm-start MLP (mlp.hip) used by the solver
// context: device-side prediction on features already in device memory.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

struct nf_mlp;

namespace nf {
namespace mlp {

constexpr int IN = 6, HID = 128, OUT = 5;

// raw features x [6][n] (feature-major, FP32, device) -> raw outputs
// y [5][n] (FP32, device), on `stream`; returns a HIP status
hipError_t predict_device(nf_mlp *m, const float *x, int64_t n, float *y, hipStream_t stream);

}  // namespace mlp
}  // namespace nf
