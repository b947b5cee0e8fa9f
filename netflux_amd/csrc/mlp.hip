// Warm-start network: MLP 6 -> 128 -> 128 -> 128 -> 5 with GELU, BF16
// weights and activations as MFMA operands, FP32 accumulation, FP32 master
// weights and Adam state.  Inputs are per-cell features (Mach, bump height,
// centroid x, y, z, wall distance); outputs are the primitive variables,
// both normalised per component.
//
// All products -- forward Z = W X, backward dX = W^T dZ and dW = dZ X^T --
// go through one BF16 MFMA GEMM kernel (v_mfma_f32_16x16x32_bf16) with
// arbitrary operand strides, so the transposed operands of the backward pass
// need no copies.  Activations are feature-major ([feature][sample]) so the
// sample dimension is the contiguous one.  Weight gradients reduce over the
// batch with a fixed split-K partition and a fixed-order sum (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/netflux.h"
#include "mlp.hpp"

using nf::mlp::HID;
using nf::mlp::IN;
using nf::mlp::OUT;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// flattened parameter layout: W1 [HID][IN], b1, W2 [HID][HID], b2, W3, b3,
// W4 [OUT][HID], b4 (row-major [out][in], as a linear layer stores them)
constexpr int64_t OFF_W1 = 0, OFF_B1 = OFF_W1 + HID * IN;
constexpr int64_t OFF_W2 = OFF_B1 + HID, OFF_B2 = OFF_W2 + HID * HID;
constexpr int64_t OFF_W3 = OFF_B2 + HID, OFF_B3 = OFF_W3 + HID * HID;
constexpr int64_t OFF_W4 = OFF_B3 + HID, OFF_B4 = OFF_W4 + OUT * HID;
constexpr int64_t NPARAM = OFF_B4 + OUT;

// ---------------------------------------------------------------------------
// GEMM  C(m, n) = sum_k A(m, k) B(k, n),  A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn], C(m,n) = C[z*sc_split + m*scm + n*scn], where
// blockIdx.z selects the K range [z*ksplit, (z+1)*ksplit).
// Tile 64x64 per 256-thread block, K step 64; each wave computes 32x32 as
// 2x2 MFMA 16x16x32 tiles.  Operand tiles are staged in LDS as [row][k]
// (k contiguous) so each lane's MFMA fragment -- 8 consecutive k of one row
// -- is a single 16-byte LDS read.
constexpr int TM = 64, TN = 64, TK = 64, LDK = TK + 8;

struct Gemm {
    const __bf16 *A;
    int64_t sam, sak;
    const __bf16 *B;
    int64_t sbk, sbn;
    float *C;
    int64_t scm, scn, sc_split;
    int M, N, K, ksplit;
};

// tile S[r][k] = P[(r0 + r)*s_row + (k0 + k)*s_k] for r < 64, k < TK, zero
// outside r0 + r < R, k0 + k < kend
__device__ __forceinline__ void load_tile(const __bf16 *P, int64_t s_row, int64_t s_k, int r0, int k0, int R,
                                          int kend, __bf16 (*S)[LDK]) {
    const int t = threadIdx.x;
    const bool aligned = (((uintptr_t)P) & 15) == 0;
    const bool vec = s_k == 1 && (s_row & 7) == 0 && aligned;
    const bool vec_rows = !vec && s_row == 1 && (s_k & 7) == 0 && aligned;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (vec_rows) {
            // rows contiguous: 16-byte loads of 8 rows at one k (a wave reads
            // 128 contiguous bytes per k), transposed into the [row][k] tile
            const int rr = (t & 7) * 8, kk = (t >> 3) + 32 * h;
            const int gr = r0 + rr, gk = k0 + kk;
            bf16x8 v;
            if (gk < kend && gr + 8 <= R) {
                v = *reinterpret_cast<const bf16x8 *>(P + gr + gk * s_k);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = (gk < kend && gr + e < R) ? P[(gr + e) + gk * s_k] : (__bf16)0.0f;
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) S[rr + e][kk] = v[e];
        } else if (vec) {
            // thread -> (row, 8 consecutive k): 16-byte loads along k
            const int r = (t >> 3) + 32 * h, kk = (t & 7) * 8;
            const int gr = r0 + r, gk = k0 + kk;
            bf16x8 v;
            if (gr < R && gk + 8 <= kend) {
                v = *reinterpret_cast<const bf16x8 *>(P + gr * s_row + gk);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    v[e] = (gr < R && gk + e < kend) ? P[gr * s_row + (gk + e)] : (__bf16)0.0f;
            }
            *reinterpret_cast<bf16x8 *>(&S[r][kk]) = v;
        } else {
            // thread -> (row, k block): consecutive threads take consecutive
            // rows, which are contiguous when s_row == 1
            const int r = t & 63, kk = (t >> 6) * 8 + 32 * h;
            const int gr = r0 + r;
            bf16x8 v;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int gk = k0 + kk + e;
                v[e] = (gr < R && gk < kend) ? P[gr * s_row + gk * s_k] : (__bf16)0.0f;
            }
            *reinterpret_cast<bf16x8 *>(&S[r][kk]) = v;
        }
    }
}

__global__ __launch_bounds__(256) void k_gemm(Gemm g) {
    __shared__ __attribute__((aligned(16))) __bf16 As[TM][LDK];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[TN][LDK];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w >> 1, wn = w & 1;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int kb = blockIdx.z * g.ksplit, ke = min(g.K, kb + g.ksplit);
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = kb; k0 < ke; k0 += TK) {
        load_tile(g.A, g.sam, g.sak, m0, k0, g.M, ke, As);
        load_tile(g.B, g.sbn, g.sbk, n0, k0, g.N, ke, Bs);
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < TK; ks += 32) {
            bf16x8 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                a[i] = *reinterpret_cast<const bf16x8 *>(&As[wm * 32 + i * 16 + (lane & 15)][ks + 8 * (lane >> 4)]);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                b[j] = *reinterpret_cast<const bf16x8 *>(&Bs[wn * 32 + j * 16 + (lane & 15)][ks + 8 * (lane >> 4)]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // D fragment: lane holds rows 4*(lane/16) + r of column lane%16
    float *C = g.C + blockIdx.z * g.sc_split;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + r;
                const int n = n0 + wn * 32 + j * 16 + (lane & 15);
                if (m < g.M && n < g.N) C[m * g.scm + n * g.scn] = acc[i][j][r];
            }
}

// ---------------------------------------------------------------------------
// elementwise kernels over feature-major [rows][ld] arrays
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// Z += b (per row); A = bf16(gelu(Z))
__global__ void k_bias_gelu(float *Z, const float *b, __bf16 *A, int rows, int n) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (j >= n || r >= rows) return;
    const size_t i = (size_t)r * n + j;
    const float z = Z[i] + b[r];
    Z[i] = z;
    A[i] = (__bf16)gelu(z);
}

// Y += b; for training: D = Y - T, dY = 2 D / (OUT n) (FP32 and BF16), and
// per-block sums of D^2 for the loss
__global__ void k_out_loss(float *Y, const float *b, const float *T, float *dY, __bf16 *dYb, float *loss_part,
                           int n, float inv_count) {
    __shared__ float sh[4];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    float s = 0.f;
    if (j < n) {
        for (int o = 0; o < OUT; ++o) {
            const size_t i = (size_t)o * n + j;
            const float y = Y[i] + b[o];
            Y[i] = y;
            if (T) {
                const float d = y - T[i];
                s += d * d;
                const float g = 2.0f * d * inv_count;
                dY[i] = g;
                dYb[i] = (__bf16)g;
            }
        }
    }
    if (!T) return;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) loss_part[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// dZ = dA * gelu'(Z), FP32 (bias gradient) and BF16 (next product)
__global__ void k_gelu_back(float *dA, const float *Z, __bf16 *dZb, int rows, int n) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (j >= n || r >= rows) return;
    const size_t i = (size_t)r * n + j;
    const float d = dA[i] * gelu_d(Z[i]);
    dA[i] = d;
    dZb[i] = (__bf16)d;
}

// out[r] = sum_j X[r][j] in a fixed order (one block per row)
__global__ __launch_bounds__(256) void k_rowsum(const float *X, int n, float *out) {
    __shared__ float sh[4];
    const int r = blockIdx.x;
    float s = 0.f;
    for (int j = threadIdx.x; j < n; j += 256) s += X[(size_t)r * n + j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[r] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// out[i] = sum over splits s (in order) of part[s*stride + i]
__global__ void k_sum_splits(const float *part, int splits, int64_t stride, int64_t n, float *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int q = 0; q < splits; ++q) s += part[q * stride + i];
    out[i] = s;
}

// raw feature-major inputs -> normalised BF16 operand; idx (optional) picks
// the samples of a minibatch; targets normalised alongside
__global__ void k_gather(const float *x, const float *y, int64_t n_all, const int64_t *idx, int64_t first, int n,
                         const float *in_shift, const float *in_scale, const float *out_shift,
                         const float *out_iscale, __bf16 *X0, float *T) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int64_t s = idx ? idx[first + j] : first + j;
    for (int f = 0; f < IN; ++f) X0[(size_t)f * n + j] = (__bf16)((x[f * n_all + s] - in_shift[f]) * in_scale[f]);
    if (T)
        for (int o = 0; o < OUT; ++o) T[(size_t)o * n + j] = (y[o * n_all + s] - out_shift[o]) * out_iscale[o];
}

// y_raw[o][first + j] = Y[o][j] * out_scale + out_shift
__global__ void k_denorm(const float *Y, int n, const float *out_shift, const float *out_scale, float *y,
                         int64_t n_all, int64_t first) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    for (int o = 0; o < OUT; ++o) y[o * n_all + first + j] = Y[(size_t)o * n + j] * out_scale[o] + out_shift[o];
}
// the same normalisation from / de-normalisation to sample-major [n][w]
// arrays (the layout of nf_mlp_predict's host arrays)
__global__ void k_gather_aos(const float *x, int n, const float *in_shift, const float *in_scale, __bf16 *X0) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    for (int f = 0; f < IN; ++f) X0[(size_t)f * n + j] = (__bf16)((x[(size_t)j * IN + f] - in_shift[f]) * in_scale[f]);
}
__global__ void k_denorm_aos(const float *Y, int n, const float *out_shift, const float *out_scale, float *y) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    for (int o = 0; o < OUT; ++o) y[(size_t)j * OUT + o] = Y[(size_t)o * n + j] * out_scale[o] + out_shift[o];
}

// Adam step on the FP32 masters (bias-corrected step size lr_t), BF16 copy
__global__ void k_adam(float *p, const float *g, float *m, float *v, __bf16 *pb, int64_t n, float lr_t, float b1,
                       float b2, float eps) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr_t * mi / (sqrtf(vi) + eps);
    p[i] = pi;
    pb[i] = (__bf16)pi;
}

__global__ void k_to_bf16(const float *p, __bf16 *pb, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pb[i] = (__bf16)p[i];
}

}  // namespace

// ---------------------------------------------------------------------------
struct nf_mlp {
    std::string err;
    hipStream_t stream = nullptr;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // FP32 [NPARAM]
    __bf16 *pb = nullptr;                                          // BF16 copy
    float *norm = nullptr;  // device: in_shift[6], in_scale[6], out_shift[5], out_scale[5], out_iscale[5]
    float hnorm[27] = {};
    bool norm_set = false;  // false: training sets it from the data
    int64_t adam_t = 0;
    // workspace for batches of up to `cap` samples
    int cap = 0;
    __bf16 *X0 = nullptr, *A[3] = {nullptr, nullptr, nullptr}, *dZb = nullptr, *dYb = nullptr;
    float *Z[3] = {nullptr, nullptr, nullptr}, *Y = nullptr, *T = nullptr, *dY = nullptr, *dA = nullptr;
    float *part = nullptr, *loss_part = nullptr;
    std::vector<void *> ws;
    int device = 0;  // device current at nf_mlp_create, holding the parameters
    // feature-major device copies of the caller's host arrays (inputs [6][n],
    // targets or outputs [5][n], minibatch order), kept and grown as needed
    float *dx = nullptr, *dy = nullptr;
    int64_t *didx = nullptr;
    int64_t data_cap = 0;
};

namespace {

constexpr int KSPLIT = 512;  // batch samples per split of the weight-gradient sums

int fail(nf_mlp *m, int code, const std::string &msg) {
    if (m) m->err = msg;
    return code;
}
#define HIPM(expr)                                                                       \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(mm, NF_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

hipError_t gemm(hipStream_t s, const __bf16 *A, int64_t sam, int64_t sak, const __bf16 *B, int64_t sbk, int64_t sbn,
                float *C, int64_t scm, int64_t scn, int M, int N, int K, int ksplit = 0, int64_t sc_split = 0) {
    if (ksplit <= 0) ksplit = (K + TK - 1) / TK * TK;
    const Gemm g{A, sam, sak, B, sbk, sbn, C, scm, scn, sc_split, M, N, K, ksplit};
    const dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM, (K + ksplit - 1) / ksplit);
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, s, g);
    return hipGetLastError();
}

int ensure_ws(nf_mlp *mm, int n) {
    if (n <= mm->cap) return NF_OK;
    for (void *q : mm->ws) (void)hipFree(q);
    mm->ws.clear();
    const int cap = (n + 255) / 256 * 256;
    auto al = [&](auto **ptr, size_t count) -> int {
        void *q = nullptr;
        HIPM(hipMalloc(&q, count * sizeof(**ptr)));
        mm->ws.push_back(q);
        *ptr = static_cast<std::remove_reference_t<decltype(*ptr)>>(q);
        return NF_OK;
    };
    int rc;
    if ((rc = al(&mm->X0, (size_t)IN * cap))) return rc;
    for (int l = 0; l < 3; ++l) {
        if ((rc = al(&mm->A[l], (size_t)HID * cap))) return rc;
        if ((rc = al(&mm->Z[l], (size_t)HID * cap))) return rc;
    }
    if ((rc = al(&mm->dZb, (size_t)HID * cap))) return rc;
    if ((rc = al(&mm->dA, (size_t)HID * cap))) return rc;
    if ((rc = al(&mm->Y, (size_t)OUT * cap))) return rc;
    if ((rc = al(&mm->T, (size_t)OUT * cap))) return rc;
    if ((rc = al(&mm->dY, (size_t)OUT * cap))) return rc;
    if ((rc = al(&mm->dYb, (size_t)OUT * cap))) return rc;
    if ((rc = al(&mm->part, (size_t)((cap + KSPLIT - 1) / KSPLIT) * HID * HID))) return rc;
    if ((rc = al(&mm->loss_part, (size_t)cap / 256 + 1))) return rc;
    mm->cap = cap;
    return NF_OK;
}

const float *d_in_shift(const nf_mlp *m) { return m->norm; }
const float *d_in_scale(const nf_mlp *m) { return m->norm + 6; }
const float *d_out_shift(const nf_mlp *m) { return m->norm + 12; }
const float *d_out_scale(const nf_mlp *m) { return m->norm + 17; }
const float *d_out_iscale(const nf_mlp *m) { return m->norm + 22; }

// activation buffers of one forward pass, feature-major [rows][n]
struct FwdBufs {
    const __bf16 *X0;
    float *Z[3];
    __bf16 *A[3];
    float *Y;
};

// forward of the n samples in X0 (normalised); Y = normalised outputs
hipError_t forward_on(const nf_mlp *m, const FwdBufs &b, int n, hipStream_t s) {
    hipError_t e;
    const __bf16 *in = b.X0;
    int k = IN;
    const int64_t offW[3] = {OFF_W1, OFF_W2, OFF_W3}, offB[3] = {OFF_B1, OFF_B2, OFF_B3};
    for (int l = 0; l < 3; ++l) {
        if ((e = gemm(s, m->pb + offW[l], k, 1, in, n, 1, b.Z[l], n, 1, HID, n, k))) return e;
        hipLaunchKernelGGL(k_bias_gelu, dim3((n + 255) / 256, HID), dim3(256), 0, s, b.Z[l],
                           (const float *)(m->p + offB[l]), b.A[l], HID, n);
        if ((e = hipGetLastError())) return e;
        in = b.A[l];
        k = HID;
    }
    return gemm(s, m->pb + OFF_W4, HID, 1, in, n, 1, b.Y, n, 1, OUT, n, HID);
}

// forward over the network's own workspace (training keeps every layer's
// pre-activations and activations for the backward pass)
hipError_t forward(nf_mlp *m, int n, hipStream_t s) {
    const FwdBufs b{m->X0, {m->Z[0], m->Z[1], m->Z[2]}, {m->A[0], m->A[1], m->A[2]}, m->Y};
    return forward_on(m, b, n, s);
}

// loss and gradient (into m->g) of the n samples in X0 / T
hipError_t backward(nf_mlp *m, int n, hipStream_t s) {
    hipError_t e;
    const float inv_count = 1.0f / ((float)OUT * (float)n);
    const int nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_out_loss, dim3(nb), dim3(256), 0, s, m->Y, (const float *)(m->p + OFF_B4),
                       (const float *)m->T, m->dY, m->dYb, m->loss_part, n, inv_count);
    if ((e = hipGetLastError())) return e;
    const int splits = (n + KSPLIT - 1) / KSPLIT;
    // weight gradient dW = dZ X^T over the batch, split over samples
    auto dW = [&](const __bf16 *dZ, int rows, const __bf16 *X, int cols, int64_t off) -> hipError_t {
        hipError_t e2;
        if ((e2 = gemm(s, dZ, n, 1, X, 1, n, m->part, cols, 1, rows, cols, n, KSPLIT, (int64_t)rows * cols)))
            return e2;
        const int64_t cnt = (int64_t)rows * cols;
        hipLaunchKernelGGL(k_sum_splits, dim3((cnt + 255) / 256), dim3(256), 0, s, (const float *)m->part, splits,
                           cnt, cnt, m->g + off);
        return hipGetLastError();
    };
    // output layer
    if ((e = dW(m->dYb, OUT, m->A[2], HID, OFF_W4))) return e;
    hipLaunchKernelGGL(k_rowsum, dim3(OUT), dim3(256), 0, s, (const float *)m->dY, n, m->g + OFF_B4);
    if ((e = hipGetLastError())) return e;
    // dA3 = W4^T dY
    if ((e = gemm(s, m->pb + OFF_W4, 1, HID, m->dYb, n, 1, m->dA, n, 1, HID, n, OUT))) return e;
    const int64_t offW[3] = {OFF_W1, OFF_W2, OFF_W3}, offB[3] = {OFF_B1, OFF_B2, OFF_B3};
    for (int l = 2; l >= 0; --l) {
        hipLaunchKernelGGL(k_gelu_back, dim3(nb, HID), dim3(256), 0, s, m->dA, (const float *)m->Z[l], m->dZb, HID,
                           n);
        if ((e = hipGetLastError())) return e;
        hipLaunchKernelGGL(k_rowsum, dim3(HID), dim3(256), 0, s, (const float *)m->dA, n, m->g + offB[l]);
        if ((e = hipGetLastError())) return e;
        const __bf16 *X = l ? m->A[l - 1] : m->X0;
        const int kin = l ? HID : IN;
        if ((e = dW(m->dZb, HID, X, kin, offW[l]))) return e;
        if (l > 0 && (e = gemm(s, m->pb + offW[l], 1, kin, m->dZb, n, 1, m->dA, n, 1, kin, n, HID))) return e;
    }
    return hipSuccess;
}

float sum_loss(nf_mlp *m, int n) {
    const int nb = (n + 255) / 256;
    std::vector<float> h(nb);
    if (hipMemcpy(h.data(), m->loss_part, nb * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return NAN;
    double s = 0;
    for (float x : h) s += x;
    return (float)(s / ((double)OUT * n));
}

int upload_norm(nf_mlp *mm) {
    float h[27];
    std::memcpy(h, mm->hnorm, 22 * sizeof(float));
    for (int o = 0; o < OUT; ++o) h[22 + o] = 1.0f / mm->hnorm[17 + o];
    HIPM(hipMemcpy(mm->norm, h, sizeof h, hipMemcpyHostToDevice));
    return NF_OK;
}

// room for n samples in the data buffers (allocates only when n grows)
int ensure_data(nf_mlp *mm, int64_t n) {
    if (n <= mm->data_cap) return NF_OK;
    HIPM(hipStreamSynchronize(mm->stream));
    for (void *q : {(void *)mm->dx, (void *)mm->dy, (void *)mm->didx})
        if (q) HIPM(hipFree(q));
    mm->dx = mm->dy = nullptr;
    mm->didx = nullptr;
    mm->data_cap = 0;
    const int64_t cap = (n + 255) / 256 * 256;
    HIPM(hipMalloc(&mm->dx, (size_t)IN * cap * sizeof(float)));
    HIPM(hipMalloc(&mm->dy, (size_t)OUT * cap * sizeof(float)));
    HIPM(hipMalloc(&mm->didx, (size_t)cap * sizeof(int64_t)));
    mm->data_cap = cap;
    return NF_OK;
}

// a host AoS [n][w] array into a feature-major device buffer [w][n]
int upload_fm(nf_mlp *mm, const float *aos, int64_t n, int w, float *dev) {
    std::vector<float> fm((size_t)w * n);
    for (int64_t j = 0; j < n; ++j)
        for (int f = 0; f < w; ++f) fm[(size_t)f * n + j] = aos[(size_t)j * w + f];
    HIPM(hipMemcpy(dev, fm.data(), fm.size() * sizeof(float), hipMemcpyHostToDevice));
    return NF_OK;
}

}  // namespace

namespace nf {
namespace mlp {

hipError_t predict_device(nf_mlp *m, const float *x, int64_t n, float *y, hipStream_t s) {
    constexpr int CH = 1 << 18;
    if (ensure_ws(m, (int)std::min<int64_t>(n, CH))) return hipErrorOutOfMemory;
    hipError_t e;
    for (int64_t first = 0; first < n; first += CH) {
        const int c = (int)std::min<int64_t>(CH, n - first);
        hipLaunchKernelGGL(k_gather, dim3((c + 255) / 256), dim3(256), 0, s, x, (const float *)nullptr, n,
                           (const int64_t *)nullptr, first, c, d_in_shift(m), d_in_scale(m), d_out_shift(m),
                           d_out_iscale(m), m->X0, (float *)nullptr);
        if ((e = hipGetLastError())) return e;
        if ((e = forward(m, c, s))) return e;
        hipLaunchKernelGGL(k_out_loss, dim3((c + 255) / 256), dim3(256), 0, s, m->Y, (const float *)(m->p + OFF_B4),
                           (const float *)nullptr, (float *)nullptr, (__bf16 *)nullptr, (float *)nullptr, c, 0.f);
        if ((e = hipGetLastError())) return e;
        hipLaunchKernelGGL(k_denorm, dim3((c + 255) / 256), dim3(256), 0, s, (const float *)m->Y, c, d_out_shift(m),
                           d_out_scale(m), y, n, first);
        if ((e = hipGetLastError())) return e;
    }
    return hipSuccess;
}

}  // namespace mlp
}  // namespace nf

// ============================================================================
// ---------------------------------------------------------------------------
// Inference that allocates nothing, for callers that may not call hipMalloc
// (code under HIP graph capture, serving loops).  The activations of one call
// -- normalised inputs X0 (BF16 [6][n]), one layer's pre-activations Z (FP32
// [128][n]), two BF16 activation buffers [128][n] used in turn, and the
// outputs Y (FP32 [5][n]) -- live in `scratch`, a file-scope device array
// sized for NOALLOC_MAX samples (about 17 MB).  Each GPU the library runs on
// holds its own copy until the process ends.  Host code never uses the
// array's name as an address: nf_mlp_create looks up the copy of the current
// device once (hipGetSymbolAddress) and the kernels take that address.  Every
// buffer is written in full before it is read, so nothing needs clearing
// between calls; calls that overlap on one GPU (two streams or threads)
// share the array and race.
namespace {

constexpr int64_t NOALLOC_MAX = 16384;  // samples per call
constexpr size_t al256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t NA_X0 = 0;
constexpr size_t NA_Z = al256(NA_X0 + sizeof(__bf16) * IN * NOALLOC_MAX);
constexpr size_t NA_A0 = al256(NA_Z + sizeof(float) * HID * NOALLOC_MAX);
constexpr size_t NA_A1 = al256(NA_A0 + sizeof(__bf16) * HID * NOALLOC_MAX);
constexpr size_t NA_Y = al256(NA_A1 + sizeof(__bf16) * HID * NOALLOC_MAX);
constexpr size_t NA_BYTES = al256(NA_Y + sizeof(float) * OUT * NOALLOC_MAX);

__device__ __align__(256) unsigned char scratch[NA_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
unsigned char *scratch_ptr[MAX_DEVICES] = {};  // per device: its copy of `scratch`

// once per device, with that device current (called by nf_mlp_create)
int scratch_setup(nf_mlp *mm, int dev) {
    if (dev < 0 || dev >= MAX_DEVICES) return fail(mm, NF_E_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (scratch_ptr[dev]) return NF_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(mm, NF_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    scratch_ptr[dev] = static_cast<unsigned char *>(p);
    return NF_OK;
}

}  // namespace

extern "C" {

int64_t nf_mlp_n_params(void) { return NPARAM; }

int nf_mlp_predict_noalloc(nf_mlp *mm, const float *x, int64_t n, float *y, void *stream) {
    if (!mm || !x || !y || n < 0) return NF_E_ARG;
    if (n > NOALLOC_MAX)
        return fail(mm, NF_E_ARG,
                    "nf_mlp_predict_noalloc: " + std::to_string(n) + " samples exceed the maximum of " +
                        std::to_string(NOALLOC_MAX) + " per call");
    if (n == 0) return NF_OK;
    int dev = -1;
    HIPM(hipGetDevice(&dev));
    if (dev != mm->device)
        return fail(mm, NF_E_ARG, "nf_mlp_predict_noalloc: the network's device must be current");
    unsigned char *base;
    {
        std::lock_guard<std::mutex> lk(scratch_mu);
        base = scratch_ptr[dev];
    }
    if (!base) return fail(mm, NF_E_HIP, "nf_mlp_predict_noalloc: scratch not set up on this device");
    const int c = (int)n;
    __bf16 *X0 = reinterpret_cast<__bf16 *>(base + NA_X0);
    float *Z = reinterpret_cast<float *>(base + NA_Z);
    __bf16 *A0 = reinterpret_cast<__bf16 *>(base + NA_A0), *A1 = reinterpret_cast<__bf16 *>(base + NA_A1);
    float *Y = reinterpret_cast<float *>(base + NA_Y);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : mm->stream;
    const dim3 g((c + 255) / 256), b(256);
    hipLaunchKernelGGL(k_gather_aos, g, b, 0, s, x, c, d_in_shift(mm), d_in_scale(mm), X0);
    HIPM(hipGetLastError());
    // layer l reads A[l-1] and writes A[l]: A0, A1, A0
    const FwdBufs fb{X0, {Z, Z, Z}, {A0, A1, A0}, Y};
    HIPM(forward_on(mm, fb, c, s));
    hipLaunchKernelGGL(k_out_loss, g, b, 0, s, Y, (const float *)(mm->p + OFF_B4), (const float *)nullptr,
                       (float *)nullptr, (__bf16 *)nullptr, (float *)nullptr, c, 0.f);
    HIPM(hipGetLastError());
    hipLaunchKernelGGL(k_denorm_aos, g, b, 0, s, (const float *)Y, c, d_out_shift(mm), d_out_scale(mm), y);
    HIPM(hipGetLastError());
    return NF_OK;
}

int nf_mlp_create(uint64_t seed, nf_mlp **out) {
    if (!out) return NF_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return NF_E_NOGPU;
    std::unique_ptr<nf_mlp> hold(new nf_mlp);
    nf_mlp *mm = hold.get();
    HIPM(hipGetDevice(&mm->device));
    int rc0 = scratch_setup(mm, mm->device);
    if (rc0) return rc0;
    HIPM(hipStreamCreateWithFlags(&mm->stream, hipStreamNonBlocking));
    HIPM(hipMalloc(&mm->p, NPARAM * sizeof(float)));
    HIPM(hipMalloc(&mm->g, NPARAM * sizeof(float)));
    HIPM(hipMalloc(&mm->m, NPARAM * sizeof(float)));
    HIPM(hipMalloc(&mm->v, NPARAM * sizeof(float)));
    HIPM(hipMalloc(&mm->pb, NPARAM * sizeof(__bf16)));
    HIPM(hipMalloc(&mm->norm, 27 * sizeof(float)));
    // uniform(+-1/sqrt(fan_in)) weights and biases, a linear layer's default
    std::mt19937_64 rng(seed);
    std::vector<float> h(NPARAM);
    auto fill = [&](int64_t off, int64_t cnt, int fan_in) {
        std::uniform_real_distribution<float> U(-1.0f / std::sqrt((float)fan_in), 1.0f / std::sqrt((float)fan_in));
        for (int64_t i = 0; i < cnt; ++i) h[off + i] = U(rng);
    };
    fill(OFF_W1, HID * IN + HID, IN);
    fill(OFF_W2, HID * HID + HID, HID);
    fill(OFF_W3, HID * HID + HID, HID);
    fill(OFF_W4, OUT * HID + OUT, HID);
    int rc = nf_mlp_set_params(mm, h.data());
    if (rc) return rc;
    for (int f = 0; f < 6; ++f) mm->hnorm[6 + f] = 1.0f;
    for (int o = 0; o < 5; ++o) mm->hnorm[17 + o] = 1.0f;
    if ((rc = upload_norm(mm))) return rc;
    *out = hold.release();
    return NF_OK;
}

int nf_mlp_destroy(nf_mlp *mm) {
    if (!mm) return NF_OK;
    if (mm->stream) (void)hipStreamSynchronize(mm->stream);
    for (void *q : mm->ws) (void)hipFree(q);
    for (void *q : {(void *)mm->p, (void *)mm->g, (void *)mm->m, (void *)mm->v, (void *)mm->pb, (void *)mm->norm,
                    (void *)mm->dx, (void *)mm->dy, (void *)mm->didx})
        if (q) (void)hipFree(q);
    if (mm->stream) (void)hipStreamDestroy(mm->stream);
    delete mm;
    return NF_OK;
}

const char *nf_mlp_last_error(const nf_mlp *m) { return m ? m->err.c_str() : "null network"; }

int nf_mlp_get_params(nf_mlp *mm, float *params) {
    if (!mm || !params) return NF_E_ARG;
    HIPM(hipStreamSynchronize(mm->stream));
    HIPM(hipMemcpy(params, mm->p, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    return NF_OK;
}

int nf_mlp_set_params(nf_mlp *mm, const float *params) {
    if (!mm || !params) return NF_E_ARG;
    HIPM(hipStreamSynchronize(mm->stream));
    HIPM(hipMemcpy(mm->p, params, NPARAM * sizeof(float), hipMemcpyHostToDevice));
    HIPM(hipMemset(mm->m, 0, NPARAM * sizeof(float)));
    HIPM(hipMemset(mm->v, 0, NPARAM * sizeof(float)));
    mm->adam_t = 0;
    hipLaunchKernelGGL(k_to_bf16, dim3((NPARAM + 255) / 256), dim3(256), 0, mm->stream, (const float *)mm->p, mm->pb,
                       NPARAM);
    HIPM(hipGetLastError());
    HIPM(hipStreamSynchronize(mm->stream));
    return NF_OK;
}

int nf_mlp_set_norm(nf_mlp *mm, const float *in_shift, const float *in_scale, const float *out_shift,
                    const float *out_scale) {
    if (!mm || !in_shift || !in_scale || !out_shift || !out_scale) return NF_E_ARG;
    for (int o = 0; o < OUT; ++o)
        if (!(out_scale[o] != 0.0f)) return NF_E_ARG;
    std::memcpy(mm->hnorm, in_shift, 6 * sizeof(float));
    std::memcpy(mm->hnorm + 6, in_scale, 6 * sizeof(float));
    std::memcpy(mm->hnorm + 12, out_shift, 5 * sizeof(float));
    std::memcpy(mm->hnorm + 17, out_scale, 5 * sizeof(float));
    mm->norm_set = true;
    return upload_norm(mm);
}

int nf_mlp_get_norm(nf_mlp *mm, float *in_shift, float *in_scale, float *out_shift, float *out_scale) {
    if (!mm || !in_shift || !in_scale || !out_shift || !out_scale) return NF_E_ARG;
    std::memcpy(in_shift, mm->hnorm, 6 * sizeof(float));
    std::memcpy(in_scale, mm->hnorm + 6, 6 * sizeof(float));
    std::memcpy(out_shift, mm->hnorm + 12, 5 * sizeof(float));
    std::memcpy(out_scale, mm->hnorm + 17, 5 * sizeof(float));
    return NF_OK;
}

int nf_mlp_predict(nf_mlp *mm, const float *x, int64_t n, float *y) {
    if (!mm || !x || !y || n < 0) return NF_E_ARG;
    if (n == 0) return NF_OK;
    int rc;
    if ((rc = ensure_data(mm, n)) || (rc = upload_fm(mm, x, n, IN, mm->dx))) return rc;
    const hipError_t e = nf::mlp::predict_device(mm, mm->dx, n, mm->dy, mm->stream);
    if (e != hipSuccess) return fail(mm, NF_E_HIP, std::string("predict: ") + hipGetErrorString(e));
    std::vector<float> fm((size_t)OUT * n);
    HIPM(hipStreamSynchronize(mm->stream));
    HIPM(hipMemcpy(fm.data(), mm->dy, fm.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (int64_t j = 0; j < n; ++j)
        for (int o = 0; o < OUT; ++o) y[(size_t)j * OUT + o] = fm[(size_t)o * n + j];
    return NF_OK;
}

// loss and gradient of one batch; x [n][6], y [n][5] raw host arrays
int nf_mlp_loss_grad(nf_mlp *mm, const float *x, const float *y, int64_t n, float *loss, float *grad) {
    if (!mm || !x || !y || n < 1 || n > (1 << 20)) return NF_E_ARG;
    int rc;
    if ((rc = ensure_ws(mm, (int)n)) || (rc = ensure_data(mm, n))) return rc;
    if ((rc = upload_fm(mm, x, n, IN, mm->dx)) || (rc = upload_fm(mm, y, n, OUT, mm->dy))) return rc;
    const int c = (int)n;
    hipLaunchKernelGGL(k_gather, dim3((c + 255) / 256), dim3(256), 0, mm->stream, (const float *)mm->dx,
                       (const float *)mm->dy, n, (const int64_t *)nullptr, (int64_t)0, c, d_in_shift(mm),
                       d_in_scale(mm), d_out_shift(mm), d_out_iscale(mm), mm->X0, mm->T);
    hipError_t e = hipGetLastError();
    if (!e) e = forward(mm, c, mm->stream);
    if (!e) e = backward(mm, c, mm->stream);
    if (!e) e = hipStreamSynchronize(mm->stream);
    if (e != hipSuccess) return fail(mm, NF_E_HIP, std::string("loss_grad: ") + hipGetErrorString(e));
    if (loss) *loss = sum_loss(mm, c);
    if (grad) HIPM(hipMemcpy(grad, mm->g, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    return NF_OK;
}

// Adam over shuffled minibatches of the whole set; x [n][6], y [n][5] raw
// host arrays.  Normalisation: if never set, mean/std of the data.
int nf_mlp_train(nf_mlp *mm, const float *x, const float *y, int64_t n, const nf_train_opts *o,
                 nf_train_stats *st) {
    if (!mm || !x || !y || !o || o->batch < 256 || o->batch % 256 || o->epochs < 0 || n < 256) return NF_E_ARG;
    int rc;
    const int B = (int)std::min<int64_t>(o->batch, n / 256 * 256);
    if ((rc = ensure_ws(mm, B))) return rc;
    if (!mm->norm_set) {
        // per-component mean and standard deviation of the training data
        for (int f = 0; f < IN + OUT; ++f) {
            double s = 0, s2 = 0;
            for (int64_t j = 0; j < n; ++j) {
                const double v = f < IN ? x[j * IN + f] : y[j * OUT + (f - IN)];
                s += v;
                s2 += v * v;
            }
            const double mean = s / n, sd = std::sqrt(std::max(0.0, s2 / n - mean * mean));
            if (f < IN) {
                mm->hnorm[f] = (float)mean;
                mm->hnorm[6 + f] = sd > 0 ? (float)(1.0 / sd) : 1.0f;
            } else {
                mm->hnorm[12 + f - IN] = (float)mean;
                mm->hnorm[17 + f - IN] = sd > 0 ? (float)sd : 1.0f;
            }
        }
        if ((rc = upload_norm(mm))) return rc;
        mm->norm_set = true;
    }
    if ((rc = ensure_data(mm, n))) return rc;
    if ((rc = upload_fm(mm, x, n, IN, mm->dx)) || (rc = upload_fm(mm, y, n, OUT, mm->dy))) return rc;
    const float *dx = mm->dx, *dy = mm->dy;
    int64_t *didx = mm->didx;
    std::vector<int64_t> idx(n);
    for (int64_t j = 0; j < n; ++j) idx[j] = j;
    std::mt19937_64 rng(o->seed);
    hipEvent_t e0, e1;
    HIPM(hipEventCreate(&e0));
    HIPM(hipEventCreate(&e1));
    HIPM(hipEventRecord(e0, mm->stream));
    const int64_t per_epoch = n / B;
    float first = NAN, last = NAN;
    int64_t steps = 0;
    hipError_t e = hipSuccess;
    for (int ep = 0; ep < o->epochs && !e; ++ep) {
        std::shuffle(idx.begin(), idx.end(), rng);
        e = hipStreamSynchronize(mm->stream);  // the previous epoch still reads didx
        if (!e) e = hipMemcpy(didx, idx.data(), n * sizeof(int64_t), hipMemcpyHostToDevice);
        const float lr = o->lr * std::pow(o->lr_decay > 0 ? o->lr_decay : 1.0f, (float)ep);
        double ep_loss = 0;
        for (int64_t q = 0; q < per_epoch && !e; ++q) {
            hipLaunchKernelGGL(k_gather, dim3(B / 256), dim3(256), 0, mm->stream, (const float *)dx, (const float *)dy,
                               n, (const int64_t *)didx, q * B, B, d_in_shift(mm), d_in_scale(mm), d_out_shift(mm),
                               d_out_iscale(mm), mm->X0, mm->T);
            e = hipGetLastError();
            if (!e) e = forward(mm, B, mm->stream);
            if (!e) e = backward(mm, B, mm->stream);
            if (e) break;
            ++mm->adam_t;
            const double bc1 = 1.0 - std::pow((double)o->beta1, (double)mm->adam_t);
            const double bc2 = 1.0 - std::pow((double)o->beta2, (double)mm->adam_t);
            const float lr_t = (float)(lr * std::sqrt(bc2) / bc1);
            hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, mm->stream, mm->p,
                               (const float *)mm->g, mm->m, mm->v, mm->pb, NPARAM, lr_t, o->beta1, o->beta2, o->eps);
            e = hipGetLastError();
            ++steps;
            if (q == 0 || q == per_epoch - 1) {
                if (!e) e = hipStreamSynchronize(mm->stream);
                const float l = e ? NAN : sum_loss(mm, B);
                if (ep == 0 && q == 0) first = l;
                ep_loss = l;
            }
        }
        last = (float)ep_loss;
    }
    if (!e) e = hipEventRecord(e1, mm->stream);
    if (!e) e = hipEventSynchronize(e1);
    float ms = 0;
    if (!e) (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return fail(mm, NF_E_HIP, std::string("train: ") + hipGetErrorString(e));
    if (st) {
        st->loss_first = first;
        st->loss_last = last;
        st->steps = steps;
        st->elapsed_ms = ms;
    }
    return NF_OK;
}

}  // extern "C"
