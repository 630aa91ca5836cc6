// Grouped complex64 GEMM for the read-out contractions.
//
// C = alpha * op(A) op(B) (+ C), op = identity or conjugate-transpose.
// Used for the left-to-right chain of MPS.wavefunction (core.py:483-500) and
// the two contractions per site of the transfer sweep in
// MPS.inner_product (core.py:553-559). 32x32 output tiles, 16-deep LDS
// stages, 2x2 outputs per thread, FP32 accumulation.
#include <hip/hip_runtime.h>

#include "common.h"
#include "launch.h"

constexpr int GT = 32;  // output tile edge
constexpr int GK = 16;  // k per stage

// Grid: x = (m tile, n tile) folded as m_tile * gn + n_tile (so m may need
// more than 65535 tiles, as the last steps of a wavefunction do), y = op.
__global__ void __launch_bounds__(256) k_gemm(const mpsx_gemm_op *ops, int64_t gn) {
    __shared__ float2 As[GT][GK + 1];
    __shared__ float2 Bs[GK][GT + 1];
    const mpsx_gemm_op op = ops[blockIdx.y];
    const int64_t tile = blockIdx.x;
    const int64_t m0 = (tile / gn) * GT, n0 = (tile % gn) * GT;
    if (m0 >= op.m || n0 >= op.n) return;
    const bool ta = (op.flags & 1) != 0, tb = (op.flags & 2) != 0, accum = (op.flags & 4) != 0;
    const float2 *A = reinterpret_cast<const float2 *>(op.A);
    const float2 *B = reinterpret_cast<const float2 *>(op.B);
    float2 *C = reinterpret_cast<float2 *>(op.C);
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    float2 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = make_float2(0.f, 0.f);

    for (int64_t k0 = 0; k0 < op.k; k0 += GK) {
        for (int e = tid; e < GT * GK; e += 256) {
            int mi, kk;
            if (!ta) { mi = e / GK; kk = e % GK; } else { kk = e / GT; mi = e % GT; }
            const int64_t gm = m0 + mi, gk = k0 + kk;
            float2 v = make_float2(0.f, 0.f);
            if (gm < op.m && gk < op.k) v = ta ? cconj(A[gk * op.lda + gm]) : A[gm * op.lda + gk];
            As[mi][kk] = v;
        }
        for (int e = tid; e < GK * GT; e += 256) {
            int ni, kk;
            if (!tb) { kk = e / GT; ni = e % GT; } else { ni = e / GK; kk = e % GK; }
            const int64_t gn = n0 + ni, gk = k0 + kk;
            float2 v = make_float2(0.f, 0.f);
            if (gn < op.n && gk < op.k) v = tb ? cconj(B[gn * op.ldb + gk]) : B[gk * op.ldb + gn];
            Bs[kk][ni] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GK; ++kk) {
            const float2 a0 = As[ty][kk], a1 = As[ty + 16][kk];
            const float2 b0 = Bs[kk][tx], b1 = Bs[kk][tx + 16];
            acc[0][0] = cfma(a0, b0, acc[0][0]);
            acc[0][1] = cfma(a0, b1, acc[0][1]);
            acc[1][0] = cfma(a1, b0, acc[1][0]);
            acc[1][1] = cfma(a1, b1, acc[1][1]);
        }
        __syncthreads();
    }
    const float2 alpha = make_float2(op.alpha_re, op.alpha_im);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int64_t gm = m0 + ty + 16 * i, gn = n0 + tx + 16 * j;
            if (gm < op.m && gn < op.n) {
                float2 v = cmul(alpha, acc[i][j]);
                float2 *dst = C + gm * op.ldc + gn;
                if (accum) {
                    const float2 old = *dst;
                    v.x += old.x;
                    v.y += old.y;
                }
                *dst = v;
            }
        }
}

hipError_t launch_gemm(hipStream_t s, const mpsx_gemm_op *d, int count, int64_t max_m,
                       int64_t max_n) {
    const int64_t gm = (max_m + GT - 1) / GT, gn = (max_n + GT - 1) / GT;
    if (gm < 1 || gn < 1) return hipSuccess;
    if (gm * gn > 2147483647LL || count > 65535) return hipErrorInvalidValue;
    dim3 grid((unsigned)(gm * gn), (unsigned)count);
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, s, d, gn);
    return hipGetLastError();
}
