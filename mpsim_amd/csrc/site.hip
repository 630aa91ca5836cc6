// One-site gate and elementwise scale kernels (HBM-bound).
//
//   one-site   out[l, o, r] = sum_p G[o, p] a[l, p, r]
//              (mpsim/core.py:753-826, the tensordot of core.py:825)
//   scale      x <- f * x or f * conj(x)
//              (MPS.renormalize core.py:591-594, MPS.dagger core.py:502-505)
#include <hip/hip_runtime.h>

#include "common.h"
#include "launch.h"

constexpr int SITE_MAXD = 16;

// Each thread owns one (l, r) column of a site tensor and reads its d
// physical entries (stride chi_r: consecutive threads take consecutive r,
// so every load and store instruction is coalesced).
template <int D>
MPSX_DEV void one_site_body(const mpsx_site_op &op, const float2 *g, int64_t e) {
    const int64_t chi_r = op.chi_r;
    const int64_t l = e / chi_r, r = e - l * chi_r;
    const float2 *src = reinterpret_cast<const float2 *>(op.a) + l * D * chi_r + r;
    float2 *dst = reinterpret_cast<float2 *>(op.out) + l * D * chi_r + r;
    float2 v[D];
#pragma unroll
    for (int p = 0; p < D; ++p) v[p] = src[p * chi_r];
#pragma unroll
    for (int o = 0; o < D; ++o) {
        float2 acc = make_float2(0.f, 0.f);
#pragma unroll
        for (int p = 0; p < D; ++p) acc = cfma(g[o * D + p], v[p], acc);
        dst[o * chi_r] = acc;
    }
}

MPSX_DEV void one_site_generic(const mpsx_site_op &op, const float2 *g, int64_t e) {
    const int d = op.d;
    const int64_t chi_r = op.chi_r;
    const int64_t l = e / chi_r, r = e - l * chi_r;
    const float2 *src = reinterpret_cast<const float2 *>(op.a) + l * d * chi_r + r;
    float2 *dst = reinterpret_cast<float2 *>(op.out) + l * d * chi_r + r;
    float2 v[SITE_MAXD];
    for (int p = 0; p < d; ++p) v[p] = src[p * chi_r];
    for (int o = 0; o < d; ++o) {
        float2 acc = make_float2(0.f, 0.f);
        for (int p = 0; p < d; ++p) acc = cfma(g[o * d + p], v[p], acc);
        dst[o * chi_r] = acc;
    }
}

// One kernel per qudit dimension 2..4 (ops of other dimensions return at
// once), launched for the dimensions a call has.
template <int D>
__global__ void __launch_bounds__(256) k_one_site(const mpsx_site_op *ops) {
    __shared__ float2 g[D * D];
    const mpsx_site_op op = ops[blockIdx.y];
    if (op.d != D) return;
    const float2 *G = reinterpret_cast<const float2 *>(op.gate);
    for (int i = threadIdx.x; i < D * D; i += blockDim.x) g[i] = G[i];
    __syncthreads();
    const int64_t pairs = op.chi_l * op.chi_r;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < pairs;
         e += (int64_t)gridDim.x * blockDim.x)
        one_site_body<D>(op, g, e);
}

// Every other dimension (1, and 5..16). Kept apart from the unrolled kernels
// because its runtime-indexed input array lives in scratch memory, which
// every launch of a kernel that has it must reserve.
__global__ void __launch_bounds__(256) k_one_site_gen(const mpsx_site_op *ops) {
    __shared__ float2 g[SITE_MAXD * SITE_MAXD];
    const mpsx_site_op op = ops[blockIdx.y];
    const int d = op.d;
    if (d >= 2 && d <= 4) return;
    const float2 *G = reinterpret_cast<const float2 *>(op.gate);
    for (int i = threadIdx.x; i < d * d; i += blockDim.x) g[i] = G[i];
    __syncthreads();
    const int64_t pairs = op.chi_l * op.chi_r;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < pairs;
         e += (int64_t)gridDim.x * blockDim.x)
        one_site_generic(op, g, e);
}

__global__ void __launch_bounds__(256) k_scale(const mpsx_scale_op *ops) {
    const mpsx_scale_op op = ops[blockIdx.y];
    float2 *x = reinterpret_cast<float2 *>(op.x);
    const float2 f = make_float2(op.re, op.im);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < op.n;
         e += (int64_t)gridDim.x * blockDim.x) {
        float2 v = x[e];
        if (op.conj) v.y = -v.y;
        x[e] = cmul(f, v);
    }
}

static unsigned blocks_for(int64_t n, int64_t per_block, unsigned cap) {
    int64_t b = (n + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

hipError_t launch_one_site(hipStream_t s, const mpsx_site_op *d, int count, int64_t max_pairs,
                           uint32_t dmask) {
    dim3 grid(blocks_for(max_pairs, 256, 8192), (unsigned)count);
    if (dmask & (1u << 2)) hipLaunchKernelGGL(k_one_site<2>, grid, dim3(256), 0, s, d);
    if (dmask & (1u << 3)) hipLaunchKernelGGL(k_one_site<3>, grid, dim3(256), 0, s, d);
    if (dmask & (1u << 4)) hipLaunchKernelGGL(k_one_site<4>, grid, dim3(256), 0, s, d);
    if (dmask & ~((1u << 2) | (1u << 3) | (1u << 4)))
        hipLaunchKernelGGL(k_one_site_gen, grid, dim3(256), 0, s, d);
    return hipGetLastError();
}

hipError_t launch_scale(hipStream_t s, const mpsx_scale_op *d, int count, int64_t max_n) {
    dim3 grid(blocks_for(max_n, 256 * 4, 8192), (unsigned)count);
    hipLaunchKernelGGL(k_scale, grid, dim3(256), 0, s, d);
    return hipGetLastError();
}
