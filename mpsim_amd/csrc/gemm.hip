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

// Tall, narrow products (no conjugation, k <= ROWS_KMAX, n <= 256): the
// late steps of the wavefunction chain, m up to d^(n-1) rows with k and n of
// a few. A workgroup takes a block of rows_per_block(k, n) consecutive rows:
// it copies their k entries of A into LDS with 16-byte loads (one
// contiguous, coalesced block when the rows are packed, lda = k), then
// thread t computes column j = t % n of rows t / n, t / n + 256 / n, ...
// from LDS (row-mates read the same entries, a broadcast) with column j of
// B in registers; a wave's stores cover consecutive outputs. Bound by HBM
// for small k and n.
constexpr int ROWS_KMAX = 32;
constexpr int ROWS_LDS_ELEMS = 8192;  // complex entries of A staged per workgroup (64 KiB)

MPSX_HD int rows_per_block(int km, int n) {
    const int per = 256 / n;
    int passes = ROWS_LDS_ELEMS / (km * per);
    passes = passes < 1 ? 1 : (passes > 8 ? 8 : passes);
    return per * passes;
}

// KM: the op's k rounded up to a power of two (the unrolled length).
template <int KM>
__global__ void __launch_bounds__(256) k_gemm_rows(const mpsx_gemm_op *ops) {
    extern __shared__ float2 As[];  // [rows of the block][KM]
    const mpsx_gemm_op op = ops[blockIdx.y];
    if (op.m <= 0 || op.n <= 0) return;
    const int n = (int)op.n, k = (int)op.k, per = 256 / n;
    const int rb = rows_per_block(KM, n);
    const int64_t rbase = (int64_t)blockIdx.x * rb;
    if (rbase >= op.m) return;
    const int nrows = (int)((op.m - rbase) < rb ? (op.m - rbase) : rb);
    const float2 *A = reinterpret_cast<const float2 *>(op.A);
    const float2 *B = reinterpret_cast<const float2 *>(op.B);
    float2 *C = reinterpret_cast<float2 *>(op.C);
    const int t = threadIdx.x;
    // stage A rows [rbase, rbase + nrows) as As[row][kk], kk < KM (zero-padded)
    if (op.lda == k && (k & 1) == 0) {
        const float4 *src = reinterpret_cast<const float4 *>(A + rbase * k);
        const int pairs = nrows * k / 2;
        for (int e = t; e < pairs; e += 256) {
            const float4 v = src[e];
            const int c = 2 * e, row = c / k, kk = c % k;
            As[row * KM + kk] = make_float2(v.x, v.y);
            As[row * KM + kk + 1] = make_float2(v.z, v.w);
        }
        if (k < KM)
            for (int e = t; e < nrows * (KM - k); e += 256)
                As[(e / (KM - k)) * KM + k + e % (KM - k)] = make_float2(0.f, 0.f);
    } else {
        for (int e = t; e < nrows * KM; e += 256) {
            const int row = e / KM, kk = e % KM;
            As[e] = kk < k ? A[(rbase + row) * op.lda + kk] : make_float2(0.f, 0.f);
        }
    }
    __syncthreads();
    if (t >= per * n) return;
    const int j = t % n;
    float2 bcol[KM];
#pragma unroll
    for (int kk = 0; kk < KM; ++kk)
        bcol[kk] = kk < k ? B[(int64_t)kk * op.ldb + j] : make_float2(0.f, 0.f);
    const float2 alpha = make_float2(op.alpha_re, op.alpha_im);
    const bool accum = (op.flags & 4) != 0;
    for (int row = t / n; row < nrows; row += per) {
        float2 acc = make_float2(0.f, 0.f);
#pragma unroll
        for (int kk = 0; kk < KM; ++kk) acc = cfma(As[row * KM + kk], bcol[kk], acc);
        float2 v = cmul(alpha, acc);
        float2 *dst = C + (rbase + row) * op.ldc + j;
        if (accum) {
            const float2 old = *dst;
            v.x += old.x;
            v.y += old.y;
        }
        *dst = v;
    }
}

bool gemm_rows_fit(const mpsx_gemm_op &o) {
    return (o.flags & 3) == 0 && o.k <= ROWS_KMAX && o.n >= 1 && o.n <= 256;
}

hipError_t launch_gemm_rows(hipStream_t s, const mpsx_gemm_op *h, const mpsx_gemm_op *d,
                            int count) {
    int64_t kmax = 1;
    for (int i = 0; i < count; ++i) kmax = kmax > h[i].k ? kmax : h[i].k;
    const int km = kmax <= 2 ? 2 : kmax <= 4 ? 4 : kmax <= 8 ? 8 : kmax <= 16 ? 16 : 32;
    int64_t blocks = 1;
    int rb_max = 1;
    for (int i = 0; i < count; ++i) {
        if (h[i].m <= 0 || h[i].n <= 0) continue;
        const int rb = rows_per_block(km, (int)h[i].n);
        rb_max = rb_max > rb ? rb_max : rb;
        blocks = blocks > (h[i].m + rb - 1) / rb ? blocks : (h[i].m + rb - 1) / rb;
    }
    if (blocks > 2147483647LL || count > 65535) return hipErrorInvalidValue;
    const dim3 g((unsigned)blocks, (unsigned)count), blk(256);
    const size_t lds = sizeof(float2) * (size_t)rb_max * km;
    if (km == 2)
        hipLaunchKernelGGL(k_gemm_rows<2>, g, blk, lds, s, d);
    else if (km == 4)
        hipLaunchKernelGGL(k_gemm_rows<4>, g, blk, lds, s, d);
    else if (km == 8)
        hipLaunchKernelGGL(k_gemm_rows<8>, g, blk, lds, s, d);
    else if (km == 16)
        hipLaunchKernelGGL(k_gemm_rows<16>, g, blk, lds, s, d);
    else
        hipLaunchKernelGGL(k_gemm_rows<32>, g, blk, lds, s, d);
    return hipGetLastError();
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
