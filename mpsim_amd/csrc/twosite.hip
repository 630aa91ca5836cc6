// Two-site update kernels for gfx950.
//
//   theta      fused  theta = G . (A . B)        (mpsim/core.py:1045-1068)
//   jacobi     blocked one-sided Jacobi SVD of theta's rows
//              (replaces np.linalg.svd inside TensorNetwork's
//               split_node_full_svd, called at core.py:1132-1137)
//   finalize   singular values, stable descending order, truncation to k,
//              absorb of S (core.py:1139-1145), discarded weight.
//
// Working matrix. Let X be theta (absorb right) or theta^T (absorb left):
// C rows of length R. The Jacobi iteration applies unitary 2b x 2b pivot
// transforms Q^H to pairs of row blocks of Z = [X | Y], Y starting as the
// identity, until the rows of X are mutually orthogonal. Then X = W Xt with
// W = Y^H unitary and the rows of Xt orthogonal, so
//   absorb right:  theta   = W Xt     -> U = W,  S V^H = Xt
//   absorb left:   theta^T = W Xt     -> U S = Xt^T,  V^H = W^T
// and the factor that carries S is read straight from Xt: nothing is ever
// divided by a singular value, and the unscaled factor W is exactly unitary
// also for the zero singular values the reference keeps.
//
// Precision: theta is formed in complex FP32 (from the complex64 site
// tensors). The working matrix Z is complex FP64 by default: Gram matrices
// and rotations run on the FP64 matrix cores, and each update is then as
// accurate as a complex64 LAPACK SVD. Over the ~10 sweeps of an update
// every row is rewritten a few hundred times, and an FP32 Z accumulates
// about 1e-6 relative error per update from those stores and from rounding
// the rotations; the FP32 variant (v_mfma_f32_16x16x4f32, Gram partials in
// FP32 per wave, FP64 across waves) remains selectable as the faster mode.
// The pivot eigensolve (cyclic Jacobi on the 2b x 2b Hermitian Gram), the
// convergence test and the row norms / discarded weight are FP64 in both.
#include <hip/hip_runtime.h>

#include <cstddef>

#include "common.h"
#include "launch.h"

#define OP_FIELD(f) offsetof(mpsx_two_site_op, f)

MPSX_DEV const mpsx_two_site_op &op_at(const mpsx_two_site_op *ops, int i) { return ops[i]; }

template <class T>
MPSX_DEV typename Cx<T>::type *z_ptr(char *ws, const mpsx_plan &p) {
    return reinterpret_cast<typename Cx<T>::type *>(ws + z_off(p));
}
MPSX_DEV double *sig_ptr(char *ws, const mpsx_plan &p) {
    return reinterpret_cast<double *>(ws + sig_off(p));
}
MPSX_DEV int *perm_ptr(char *ws, const mpsx_plan &p) {
    return reinterpret_cast<int *>(ws + perm_off(p));
}
MPSX_DEV ProblemCtl *ctl_ptr(char *ws, const mpsx_plan &p) {
    return reinterpret_cast<ProblemCtl *>(ws + p.ws_off);
}

// Block-wide FP64 sum (256 threads).
MPSX_DEV double block_sum_256(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) s = red[0] + red[1] + red[2] + red[3];
    return s;  // valid in thread 0
}

// ---------------------------------------------------------------- init
// One workgroup per padded row of every problem: Y <- identity, padded X
// rows <- 0. Row 0 of each problem also resets its control block.
template <class T>
__global__ void __launch_bounds__(256)
k_init_rows(const mpsx_two_site_op *ops, int count, char *ws) {
    using Z2 = typename Cx<T>::type;
    if (blockIdx.x == 0 && threadIdx.x < 128)
        reinterpret_cast<TwoSiteHeader *>(ws)->active[threadIdx.x & 63] = 0;
    const int pi = find_record(reinterpret_cast<const char *>(ops), count, sizeof(mpsx_two_site_op),
                               OP_FIELD(plan.row_base), blockIdx.x);
    const mpsx_plan p = ops[pi].plan;
    const int row = (int)(blockIdx.x - p.row_base);
    if (row >= p.Cp) return;
    Z2 *z = z_ptr<T>(ws, p) + (int64_t)row * p.LD;
    if (row >= p.C)
        for (int c = threadIdx.x; c < p.R; c += 256) z[c] = Cx<T>::make(0, 0);
    for (int c = threadIdx.x; c < p.Cp; c += 256)
        z[p.R + c] = Cx<T>::make(c == row ? 1 : 0, 0);
    if (row == 0 && threadIdx.x == 0) {
        ProblemCtl *ctl = ctl_ptr(ws, p);
        ctl->fro2 = 0.0;
        ctl->null2 = 0.0;
        ctl->counter = 0;
        ctl->conv = (p.C == 0 || p.R == 0) ? 1 : 0;
        ctl->sweeps = 0;
    }
}

// ---------------------------------------------------------------- theta
// One workgroup per 32x32 (l, r) tile; thread (ty, tx) owns l in
// {ty, ty+16}, r in {tx, tx+16} and all D*D (p1, p2) partial products.
template <int D, class T>
__global__ void __launch_bounds__(256)
k_theta(const mpsx_two_site_op *ops, int count, char *ws) {
    constexpr int TL = THETA_TL, TR = THETA_TR, KT = THETA_KT;
    __shared__ float2 As[TL * D][KT + 1];
    __shared__ float2 Bs[KT * D][TR + 1];
    __shared__ float2 Gs[D * D * D * D];

    const int pi = find_record(reinterpret_cast<const char *>(ops), count, sizeof(mpsx_two_site_op),
                               OP_FIELD(plan.tile_base), blockIdx.x);
    const mpsx_two_site_op &op = ops[pi];
    if (op.d != D || op.gate == nullptr) return;
    const int chi_l = op.chi_l, chi_m = op.chi_m, chi_r = op.chi_r;
    const mpsx_plan p = op.plan;
    const int tiles_r = (chi_r + TR - 1) / TR;
    const int t = (int)(blockIdx.x - p.tile_base);
    const int l0 = (t / tiles_r) * TL, r0 = (t % tiles_r) * TR;
    const float2 *A = reinterpret_cast<const float2 *>(op.a);
    const float2 *B = reinterpret_cast<const float2 *>(op.b);
    const float2 *G = reinterpret_cast<const float2 *>(op.gate);
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;

    for (int i = tid; i < D * D * D * D; i += 256) Gs[i] = G[i];

    float2 acc[2][2][D][D];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int i = 0; i < D; ++i)
#pragma unroll
                for (int j = 0; j < D; ++j) acc[a][b][i][j] = make_float2(0.f, 0.f);

    for (int m0 = 0; m0 < chi_m; m0 += KT) {
        for (int e = tid; e < TL * D * KT; e += 256) {
            const int km = e % KT, row = e / KT;  // row = il*D + p1
            const int l = l0 + row / D, m = m0 + km;
            float2 v = make_float2(0.f, 0.f);
            if (l < chi_l && m < chi_m) v = A[((int64_t)l0 * D + row) * chi_m + m];
            As[row][km] = v;
        }
        for (int e = tid; e < KT * D * TR; e += 256) {
            const int ir = e % TR, row = e / TR;  // row = km*D + p2
            const int m = m0 + row / D, r = r0 + ir;
            float2 v = make_float2(0.f, 0.f);
            if (m < chi_m && r < chi_r) v = B[((int64_t)m0 * D + row) * chi_r + r];
            Bs[row][ir] = v;
        }
        __syncthreads();
#pragma unroll 4
        for (int km = 0; km < KT; ++km) {
            float2 av[2][D], bv[2][D];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int i = 0; i < D; ++i) av[a][i] = As[(ty + 16 * a) * D + i][km];
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int j = 0; j < D; ++j) bv[b][j] = Bs[km * D + j][tx + 16 * b];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int i = 0; i < D; ++i)
#pragma unroll
                        for (int j = 0; j < D; ++j)
                            acc[a][b][i][j] = cfma(av[a][i], bv[b][j], acc[a][b][i][j]);
        }
        __syncthreads();
    }
    __syncthreads();

    // Epilogue: theta[l,o1,o2,r] = sum G[(o1,o2),(p1,p2)] T[l,p1,p2,r]
    typename Cx<T>::type *Z = z_ptr<T>(ws, p);
    const int LD = p.LD;
    const bool transposed = (op.flags & 1) != 0;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int l = l0 + ty + 16 * a;
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int r = r0 + tx + 16 * b;
            if (l >= chi_l || r >= chi_r) continue;
#pragma unroll
            for (int o1 = 0; o1 < D; ++o1)
#pragma unroll
                for (int o2 = 0; o2 < D; ++o2) {
                    float2 v = make_float2(0.f, 0.f);
#pragma unroll
                    for (int i = 0; i < D; ++i)
#pragma unroll
                        for (int j = 0; j < D; ++j)
                            v = cfma(Gs[(o1 * D + o2) * D * D + i * D + j], acc[a][b][i][j], v);
                    const int64_t row = (int64_t)l * D + o1;  // theta row (l, o1)
                    const int64_t col = (int64_t)o2 * chi_r + r;  // theta col (o2, r)
                    if (!transposed)
                        Z[row * LD + col] = Cx<T>::make(v.x, v.y);
                    else
                        Z[col * LD + row] = Cx<T>::make(v.x, v.y);
                }
        }
    }
}

// theta for a qudit dimension D in 5..16 (the kernels above are specialised
// for D <= 4). One workgroup per (l, tile of tr r values), tr =
// theta_gen_tr(D): first T[p1 p2][r] = sum_m A[l, p1, m] B[m, p2, r] for the
// tile into LDS (FP32 accumulation, as k_theta), then
// theta[(l, o1), (o2, r)] = sum_{p1 p2} G[(o1 o2), (p1 p2)] T[p1 p2][r].
// The flops are those of the specialised kernel (d^2 chi_l chi_m chi_r for
// T, d^4 chi_l chi_r for the gate); a wave's lanes take consecutive r, so
// the reads of B and the writes of theta are coalesced and A is read once
// per wave.
template <class T>
__global__ void __launch_bounds__(256)
k_theta_gen(const mpsx_two_site_op *ops, int count, char *ws) {
    __shared__ float2 Ts[4096];
    const int pi = find_record(reinterpret_cast<const char *>(ops), count, sizeof(mpsx_two_site_op),
                               OP_FIELD(plan.tile_base), blockIdx.x);
    const mpsx_two_site_op &op = ops[pi];
    const int D = op.d;
    if (op.gate == nullptr || D <= 4 || D > THETA_GEN_DMAX) return;
    const int chi_l = op.chi_l, chi_m = op.chi_m, chi_r = op.chi_r;
    const mpsx_plan p = op.plan;
    const int tr = theta_gen_tr(D), tiles_r = (chi_r + tr - 1) / tr;
    const int64_t t = (int64_t)blockIdx.x - p.tile_base;
    if (t >= (int64_t)chi_l * tiles_r) return;
    const int l = (int)(t / tiles_r), r0 = (int)(t % tiles_r) * tr;
    const float2 *A = reinterpret_cast<const float2 *>(op.a);
    const float2 *B = reinterpret_cast<const float2 *>(op.b);
    const float2 *G = reinterpret_cast<const float2 *>(op.gate);
    const int n = D * D * tr;  // <= 4096
    for (int e = threadIdx.x; e < n; e += 256) {
        const int rr = e % tr, pp = e / tr, p1 = pp / D, p2 = pp % D, r = r0 + rr;
        float2 acc = make_float2(0.f, 0.f);
        if (r < chi_r) {
            const float2 *a = A + ((int64_t)l * D + p1) * chi_m;
            const float2 *bcol = B + (int64_t)p2 * chi_r + r;
            for (int m = 0; m < chi_m; ++m) acc = cfma(a[m], bcol[(int64_t)m * D * chi_r], acc);
        }
        Ts[e] = acc;
    }
    __syncthreads();
    typename Cx<T>::type *Z = z_ptr<T>(ws, p);
    const int LD = p.LD;
    const bool transposed = (op.flags & 1) != 0;
    const int DD = D * D;
    for (int e = threadIdx.x; e < n; e += 256) {
        const int rr = e % tr, oo = e / tr, o1 = oo / D, o2 = oo % D, r = r0 + rr;
        if (r >= chi_r) continue;
        const float2 *g = G + (int64_t)oo * DD;
        float2 v = make_float2(0.f, 0.f);
        for (int pp = 0; pp < DD; ++pp) v = cfma(g[pp], Ts[pp * tr + rr], v);
        const int64_t row = (int64_t)l * D + o1;      // theta row (l, o1)
        const int64_t col = (int64_t)o2 * chi_r + r;  // theta col (o2, r)
        if (!transposed)
            Z[row * LD + col] = Cx<T>::make(v.x, v.y);
        else
            Z[col * LD + row] = Cx<T>::make(v.x, v.y);
    }
}

// Plain split: one workgroup per row m of the M x N input matrix, copied
// into X (or X^T for absorb-left problems).
template <class T>
__global__ void __launch_bounds__(256)
k_load(const mpsx_two_site_op *ops, int count, char *ws) {
    const int pi = find_record(reinterpret_cast<const char *>(ops), count, sizeof(mpsx_two_site_op),
                               OP_FIELD(plan.tile_base), blockIdx.x);
    const mpsx_two_site_op &op = ops[pi];
    if (op.gate != nullptr) return;
    const mpsx_plan p = op.plan;
    const int m = (int)(blockIdx.x - p.tile_base);
    if (m >= op.chi_l) return;
    const int N = op.chi_r, LD = p.LD;
    const float2 *src = reinterpret_cast<const float2 *>(op.a) + (int64_t)m * N;
    typename Cx<T>::type *Z = z_ptr<T>(ws, p);
    if (!(op.flags & 1))
        for (int n = threadIdx.x; n < N; n += 256) Z[(int64_t)m * LD + n] = Cx<T>::make(src[n].x, src[n].y);
    else
        for (int n = threadIdx.x; n < N; n += 256) Z[(int64_t)n * LD + m] = Cx<T>::make(src[n].x, src[n].y);
}

// ---------------------------------------------------------------- norms
// One workgroup per padded row: sig2[row] = sum_c |X[row, c]|^2 (FP64).
template <class T>
__global__ void __launch_bounds__(256)
k_rownorm(const mpsx_two_site_op *ops, int count, char *ws) {
    __shared__ double red[4];
    const int pi = find_record(reinterpret_cast<const char *>(ops), count, sizeof(mpsx_two_site_op),
                               OP_FIELD(plan.row_base), blockIdx.x);
    const mpsx_plan p = ops[pi].plan;
    const int row = (int)(blockIdx.x - p.row_base);
    if (row >= p.Cp) return;
    const typename Cx<T>::type *z = z_ptr<T>(ws, p) + (int64_t)row * p.LD;
    double s = 0.0;
    if (row < p.C)
        for (int c = threadIdx.x; c < p.R; c += 256) {
            const auto v = z[c];
            s += (double)v.x * v.x + (double)v.y * v.y;
        }
    s = block_sum_256(s, red);
    if (threadIdx.x == 0) sig_ptr(ws, p)[row] = s;
}

// One workgroup per problem: Frobenius norm and the null-row threshold.
__global__ void __launch_bounds__(256)
k_fro(const mpsx_two_site_op *ops, int count, char *ws, double null_rel) {
    __shared__ double red[4];
    const int pi = blockIdx.x;
    if (pi >= count) return;
    const mpsx_plan p = ops[pi].plan;
    const double *sig = sig_ptr(ws, p);
    double s = 0.0;
    for (int i = threadIdx.x; i < p.C; i += 256) s += sig[i];
    s = block_sum_256(s, red);
    if (threadIdx.x == 0) {
        ProblemCtl *ctl = ctl_ptr(ws, p);
        ctl->fro2 = s;
        ctl->null2 = null_rel * null_rel * s;
    }
}

// ---------------------------------------------------------------- Jacobi
// One workgroup (4 waves) per (problem, pivot pair) of round `round`; the
// pivot is B2 = 2b rows of Z. Both GEMM phases run on the FP32 matrix cores
// (v_mfma_f32_16x16x4f32) over the real embedding of the complex rows:
//   S = [Re Xp; Im Xp]  (NS = 2*B2 real rows)
//   Gram:   G = S S^T,  H = (G11 + G22) + i (G21 - G12)
//   apply:  [Re Zp'; Im Zp'] = T [Re Zp; Im Zp],  T = [[Ar, -Ai], [Ai, Ar]],
//           A = Q^H
// MFMA 16x16x4 operand layout (wave64): lane l holds A[l%16][l/16],
// B[l/16][l%16]; result register r of lane l is D[4*(l/16)+r][l%16].
typedef float mfma_f4 __attribute__((ext_vector_type(4)));

MPSX_DEV mfma_f4 mfma16(float a, float b, mfma_f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Rotation parameters of the pivot eigensolve. They sit on the serial path
// of every step, and the correctly rounded FP64 sqrt, division and rsqrt are
// long instruction sequences, so both start from the hardware estimates
// (v_rsq_f64, v_rcp_f64) refined by Newton steps:
//  * w = sign(tau) / (|tau| + sqrt(tau^2 + g2)) only sets the angle; an
//    error there leaves a slightly smaller off-diagonal that later sweeps
//    remove, so one step each suffices;
//  * cs = 1 / sqrt(1 + w^2 g2) must make [[cs, s], [-conj(s), cs]] with
//    s = w cs h unitary to FP64 rounding (cs^2 (1 + w^2 g2) = 1), so it gets
//    two steps (each squares the relative error).
// Inputs are positive and normal (g2 > 0 is required by the caller).
MPSX_DEV double rotation_w(double tau, double g2) {
    const double x = tau * tau + g2;
    double r = __builtin_amdgcn_rsq(x);
    r = r * fma(-0.5 * x * r, r, 1.5);
    const double den = fabs(tau) + x * r;
    double inv = __builtin_amdgcn_rcp(den);
    inv = inv * fma(-den, inv, 2.0);
    return tau >= 0.0 ? inv : -inv;
}
MPSX_DEV double unit_cos(double y) {
    double r = __builtin_amdgcn_rsq(y);
    r = r * fma(-0.5 * y * r, r, 1.5);
    r = r * fma(-0.5 * y * r, r, 1.5);
    return r;
}

// Convergence test and eigensolve of one pivot (shared by all round
// kernels). H: the pivot's Hermitian Gram (FP64, LDS); Q (LDS, identity on
// entry) accumulates the rotations. A pair rotates when |h_ij|^2 exceeds
// both tol^2 h_ii h_jj and abs2 (an absolute floor, 0 for Grams formed from
// the rows; the Gram method, whose H carries FP64 rounding from every
// round, sets it above that noise; inner steps use 1e-4 abs2). Returns false
// in every thread when no pair of the pivot exceeds the tolerance.
template <int B2, int NT = 256>
MPSX_DEV bool pivot_solve(double2 (*H)[B2 + 1], double2 (*Q)[B2 + 1], double *prm_c,
                          double2 *prm_s, int &s_any, ProblemCtl *ctl, double null2, double tol,
                          double tol_inner, int inner_sweeps, double abs2) {
    constexpr int b = B2 / 2, NW = NT / 64;  // NT threads, NW waves
    const int tid = threadIdx.x;
    // ---- convergence test on every off-diagonal pair of the pivot (the
    // blocks' own rows included: a block is only orthogonal internally once
    // it has been part of a rotated pivot)
    int rot = 0;
    for (int e = tid; e < B2 * B2; e += NT) {
        const int i = e / B2, j = e % B2;
        if (i >= j) continue;
        const double hi = H[i][i].x, hj = H[j][j].x;
        const double2 h = H[i][j];
        const double g2 = h.x * h.x + h.y * h.y;
        if (hi > null2 && hj > null2 && g2 > tol * tol * hi * hj && g2 > abs2) rot = 1;
    }
    if (!__syncthreads_or(rot)) return false;
    if (tid == 0) atomicAdd(&ctl->counter, 1);

    // ---- pivot eigensolve: cyclic Jacobi on H (FP64), Q accumulates.
    // The index pairs of every step come from a table built once (pt[step][k]
    // = (p, q) of pair k), so a step carries no index arithmetic. Work items
    // of a step: the b*b 2x2 blocks (k1, k2) of H, B <- J_k1^H B J_k2, taken
    // by threads 0 .. b*b-1 (every entry of H lies in exactly one block, so
    // a step reads and writes disjoint entries and needs no mirror), and the
    // NQ = B2*b (row, pair) entries of Q, Q <- Q J, dealt out per wave in
    // NQI full slots. An item whose rotations are all the identity (s == 0
    // exactly) is skipped.
    constexpr int NB = b * b, NQ = B2 * b;
    constexpr int QPW = NQ / NW, NQI = (QPW + 63) / 64;
    static_assert(NB <= NT && NQ % NW == 0, "pivot too large for one workgroup");
    __shared__ unsigned char pt[B2 - 1][b][2];
    for (int e = tid; e < (B2 - 1) * b; e += NT) {
        int pp, qq;
        inner_pair(e % b, e / b, b, &pp, &qq);
        pt[e / b][e % b][0] = (unsigned char)pp;
        pt[e / b][e % b][1] = (unsigned char)qq;
    }
    const int lane64 = tid & 63, wv = tid >> 6;
    const int hk1 = tid < NB ? tid / b : -1, hk2 = tid % b;
    int qrow[NQI], qk[NQI];
#pragma unroll
    for (int t = 0; t < NQI; ++t) {
        const int r = t * 64 + lane64;
        const int j = wv * QPW + r;
        const bool ok = r < QPW;
        qrow[t] = ok ? j / b : -1;
        qk[t] = ok ? j % b : 0;
    }
    auto nz = [](double2 v) { return v.x != 0.0 || v.y != 0.0; };
    const double abs2_inner = 1e-4 * abs2;
    // step_any[step & 1]: some pair of the step rotates. The parameter lanes
    // and lane 0, which clears the other entry for the next step, are all in
    // wave 0, so their LDS writes stay in program order.
    __shared__ int step_any[2];
    for (int sw = 0; sw < inner_sweeps; ++sw) {
        __syncthreads();  // every wave has read s_any of the previous sweep (and pt is built)
        if (tid == 0) {
            s_any = 0;
            step_any[0] = 0;
        }
        __syncthreads();
        for (int step = 0; step < B2 - 1; ++step) {
            if (tid < b) {
                const int pp = pt[step][tid][0], qq = pt[step][tid][1];
                const double a = H[pp][pp].x, c = H[qq][qq].x;
                const double2 h = H[pp][qq];
                const double g2 = h.x * h.x + h.y * h.y;
                double cs = 1.0;
                double2 s = make_double2(0.0, 0.0);
                if (a > null2 && c > null2 && g2 > tol_inner * tol_inner * a * c && g2 > 0.0 &&
                    g2 > abs2_inner) {
                    // tan of the rotation angle t = w |h| with
                    // w = sign(tau) / (|tau| + sqrt(tau^2 + |h|^2)), tau = (c - a)/2;
                    // s = sn e^{i phi} = (t cs / |h|) h = w cs h.
                    const double tau = 0.5 * (c - a);
                    const double w = rotation_w(tau, g2);
                    cs = unit_cos(1.0 + w * w * g2);
                    const double wc = w * cs;
                    s = make_double2(wc * h.x, wc * h.y);
                    s_any = 1;
                    step_any[step & 1] = 1;
                }
                prm_c[tid] = cs;
                prm_s[tid] = s;
            }
            __syncthreads();
            const bool any = step_any[step & 1] != 0;
            if (tid == 0) step_any[(step + 1) & 1] = 0;
            if (!any) continue;  // nothing to rotate in this step: H and Q unchanged
            // H <- J^H H J on the 2x2 block (k1, k2), Q <- Q J, with
            // J = [[c, s], [-conj(s), c]] for each index pair (p, q).
            if (hk1 >= 0) {
                const int k1 = hk1, k2 = hk2;
                const double2 s1 = prm_s[k1], s2 = prm_s[k2];
                if (nz(s1) || nz(s2)) {
                    const int p1 = pt[step][k1][0], q1 = pt[step][k1][1];
                    const int p2 = pt[step][k2][0], q2 = pt[step][k2][1];
                    const double c1 = prm_c[k1], c2 = prm_c[k2];
                    const double2 h00 = H[p1][p2], h01 = H[p1][q2], h10 = H[q1][p2], h11 = H[q1][q2];
                    const double2 ms2 = make_double2(-s2.x, s2.y);  // -conj(s2)
                    const double2 m00 = zadd(zscale(h00, c2), zmul(h01, ms2));
                    const double2 m01 = zadd(zmul(h00, s2), zscale(h01, c2));
                    const double2 m10 = zadd(zscale(h10, c2), zmul(h11, ms2));
                    const double2 m11 = zadd(zmul(h10, s2), zscale(h11, c2));
                    const double2 ns1 = make_double2(-s1.x, -s1.y);  // -s1
                    const double2 cs1 = zconj(s1);
                    H[p1][p2] = zadd(zscale(m00, c1), zmul(ns1, m10));
                    H[p1][q2] = zadd(zscale(m01, c1), zmul(ns1, m11));
                    H[q1][p2] = zadd(zmul(cs1, m00), zscale(m10, c1));
                    H[q1][q2] = zadd(zmul(cs1, m01), zscale(m11, c1));
                }
            }
#pragma unroll
            for (int t = 0; t < NQI; ++t) {
                if (qrow[t] < 0) continue;
                const int row = qrow[t], k = qk[t];
                const double2 s = prm_s[k];
                if (!nz(s)) continue;
                const int pp = pt[step][k][0], qq = pt[step][k][1];
                const double c = prm_c[k];
                const double2 qp = Q[row][pp], qv = Q[row][qq];
                const double2 ms = make_double2(-s.x, s.y);  // -conj(s)
                Q[row][pp] = zadd(zscale(qp, c), zmul(qv, ms));
                Q[row][qq] = zadd(zmul(qp, s), zscale(qv, c));
            }
            __syncthreads();
        }
        if (!s_any) break;
    }

    return true;
}

template <int B2>
__global__ void __launch_bounds__(256)
k_jacobi_round32(const mpsx_two_site_op *ops, const int *owner, char *ws, int round, double tol,
               double tol_inner, int inner_sweeps) {
    constexpr int b = B2 / 2;
    constexpr int NS = 2 * B2;          // stacked real rows
    constexpr int NT = NS / 16;         // 16-row tiles per side
    constexpr int RG = B2 / 16;         // 16-row groups of complex pivot rows

    // gp: FP32 partial Grams of waves (0+2) and (1+3); Q reuses the space
    // once H has been formed.
    __shared__ union {
        float gp[2][NS][NS];
        double2 q[B2][B2 + 1];
    } sm;
    __shared__ double2 H[B2][B2 + 1];
    __shared__ int rows[B2];
    __shared__ double prm_c[b];
    __shared__ double2 prm_s[b];
    __shared__ int s_any;
    auto &Q = sm.q;

    const int pi = owner[blockIdx.x];
    const mpsx_plan p = ops[pi].plan;
    if (p.b != b || p.method != 0) return;
    if (round >= p.nb - 1) return;
    ProblemCtl *ctl = ctl_ptr(ws, p);
    if (ctl->conv) return;
    const double null2 = ctl->null2;
    const int pair = (int)(blockIdx.x - p.pair_base);
    int I, J;
    rr_pair(pair, round, p.nb, &I, &J);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    auto row_of = [&](int i) { return i < b ? I * b + i : J * b + (i - b); };
    if (tid < B2) rows[tid] = row_of(tid);
    float2 *Z = z_ptr<float>(ws, p);
    const int LD = p.LD, R = p.R;

    // ---- Gram G = S S^T on the matrix cores, S = [Re Xp; Im Xp]. Wave w
    // takes the 16-column blocks w, w+4, ... of X; lane (l16, lq) loads four
    // consecutive columns of each of its complex rows g*16 + l16 and feeds
    // column c0 + 4 lq + s at k-step s (the sum over columns is
    // order-free; A and B operands use the same order). Operands go from
    // global memory straight to registers: no staging, no barriers.
    mfma_f4 acc[NT][NT];
#pragma unroll
    for (int u = 0; u < NT; ++u)
#pragma unroll
        for (int v = 0; v < NT; ++v) acc[u][v] = mfma_f4{0.f, 0.f, 0.f, 0.f};
    int64_t goff[RG];
#pragma unroll
    for (int g = 0; g < RG; ++g) goff[g] = (int64_t)row_of(g * 16 + l16) * LD;
    for (int c0 = wave * 16; c0 < R; c0 += 64) {
        float2 v[RG][4];
        const int cq = c0 + 4 * lq;
        if (cq + 4 <= R) {  // 32 contiguous bytes per row: two 16-byte loads
#pragma unroll
            for (int g = 0; g < RG; ++g) {
                const float4 *src = reinterpret_cast<const float4 *>(Z + goff[g] + cq);
                const float4 lo = src[0], hi = src[1];
                v[g][0] = make_float2(lo.x, lo.y);
                v[g][1] = make_float2(lo.z, lo.w);
                v[g][2] = make_float2(hi.x, hi.y);
                v[g][3] = make_float2(hi.z, hi.w);
            }
        } else {
#pragma unroll
            for (int g = 0; g < RG; ++g)
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int c = cq + k;
                    v[g][k] = c < R ? Z[goff[g] + c] : make_float2(0.f, 0.f);
                }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float op[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) op[t] = t < NT / 2 ? v[t][k].x : v[t - NT / 2][k].y;
#pragma unroll
            for (int u = 0; u < NT; ++u)
#pragma unroll
                for (int w = 0; w < NT; ++w) acc[u][w] = mfma16(op[u], op[w], acc[u][w]);
        }
    }
    // deterministic reduction of the four wave partials: (w2 + w0), (w3 + w1)
    // in FP32, then the two halves in FP64 while forming H.
    if (wave >= 2) {
#pragma unroll
        for (int u = 0; u < NT; ++u)
#pragma unroll
            for (int w = 0; w < NT; ++w)
#pragma unroll
                for (int r = 0; r < 4; ++r) sm.gp[wave - 2][u * 16 + 4 * lq + r][w * 16 + l16] = acc[u][w][r];
    }
    __syncthreads();
    if (wave < 2) {
#pragma unroll
        for (int u = 0; u < NT; ++u)
#pragma unroll
            for (int w = 0; w < NT; ++w)
#pragma unroll
                for (int r = 0; r < 4; ++r) sm.gp[wave][u * 16 + 4 * lq + r][w * 16 + l16] += acc[u][w][r];
    }
    __syncthreads();
    double2 hv[(B2 * B2 + 255) / 256];
#pragma unroll
    for (int t = 0; t < (B2 * B2 + 255) / 256; ++t) {
        const int e = tid + 256 * t;
        if (e < B2 * B2) {
            const int i = e / B2, j = e % B2;
            const int u = i <= j ? i : j, v = i <= j ? j : i;  // upper triangle, mirrored
            auto G = [&](int r, int c) { return (double)sm.gp[0][r][c] + (double)sm.gp[1][r][c]; };
            const double re = G(u, v) + G(B2 + u, B2 + v);
            double im = G(B2 + u, v) - G(u, B2 + v);
            if (i > j) im = -im;
            if (i == j) im = 0.0;
            hv[t] = make_double2(re, im);
        }
    }
    __syncthreads();  // gp is dead from here on; Q reuses it
#pragma unroll
    for (int t = 0; t < (B2 * B2 + 255) / 256; ++t) {
        const int e = tid + 256 * t;
        if (e < B2 * B2) {
            const int i = e / B2, j = e % B2;
            H[i][j] = hv[t];
            Q[i][j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
        }
    }
    __syncthreads();

    if (!pivot_solve<B2>(H, Q, prm_c, prm_s, s_any, ctl, null2, tol, tol_inner, inner_sweeps, 0.0))
        return;

    // ---- apply on the matrix cores. T[i][j] (i, j < NS) from Q:
    //   T[i][j] = Re Q[j][i], T[i][B2+j] = Im Q[j][i],
    //   T[B2+i][j] = -Im Q[j][i], T[B2+i][B2+j] = Re Q[j][i]   (i, j < B2)
    // Lane l of tile row tm, k-step s needs T[16 tm + l%16][4 s + l/16].
    float tA[NT][NS / 4];
#pragma unroll
    for (int tm = 0; tm < NT; ++tm)
#pragma unroll
        for (int s = 0; s < NS / 4; ++s) {
            const int i = tm * 16 + l16, j = 4 * s + lq;
            const int ii = i % B2, jj = j % B2;
            const double2 q = Q[jj][ii];
            float v;
            if (i < B2)
                v = (float)(j < B2 ? q.x : q.y);
            else
                v = (float)(j < B2 ? -q.y : q.x);
            tA[tm][s] = v;
        }
    const int W = p.R + p.Cp;
    int64_t roff[B2 / 4];
#pragma unroll
    for (int s = 0; s < B2 / 4; ++s) roff[s] = (int64_t)rows[4 * s + lq] * LD;
    // lane's complex inputs of the column block: rows (4 s + lq), s < B2/4;
    // the next block's inputs are loaded before this block's MFMAs.
    float2 zin[B2 / 4], znx[B2 / 4];
    {
        const int c = wave * 16 + l16;
#pragma unroll
        for (int s = 0; s < B2 / 4; ++s) zin[s] = c < W ? Z[roff[s] + c] : make_float2(0.f, 0.f);
    }
    for (int c0 = wave * 16; c0 < W; c0 += 64) {
        const int c = c0 + l16;
        const bool ok = c < W;
        const int cn = c + 64;
#pragma unroll
        for (int s = 0; s < B2 / 4; ++s) znx[s] = cn < W ? Z[roff[s] + cn] : make_float2(0.f, 0.f);
        mfma_f4 acc[NT];
#pragma unroll
        for (int tm = 0; tm < NT; ++tm) acc[tm] = mfma_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < NS / 4; ++s) {
            const float bv = s < B2 / 4 ? zin[s].x : zin[s - B2 / 4].y;
#pragma unroll
            for (int tm = 0; tm < NT; ++tm) acc[tm] = mfma16(tA[tm][s], bv, acc[tm]);
        }
        if (ok) {
#pragma unroll
            for (int tm = 0; tm < NT / 2; ++tm)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = tm * 16 + 4 * lq + r;  // complex row in the pivot
                    Z[(int64_t)rows[row] * LD + c] = make_float2(acc[tm][r], acc[tm + NT / 2][r]);
                }
        }
#pragma unroll
        for (int s = 0; s < B2 / 4; ++s) zin[s] = znx[s];
    }
}

// The same round with an FP64 working matrix, on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64: operands laid out as in the FP32 16x16x4
// instruction, lane l holding A[l%16][l/16] and B[l/16][l%16]; but result
// register r of lane l is D[4*r + l/16][l%16]). Both products work on the
// real and imaginary parts directly (no real embedding, which would double
// the accumulators and rotation operands):
//   Gram:  Re H = Xr Xr^T + Xi Xi^T,   Im H = Xi Xr^T - Xr Xi^T
//   apply: Re Z' = Re(Q)^T Zr + Im(Q)^T Zi,  Im Z' = -Im(Q)^T Zr + Re(Q)^T Zi
// FP64 matrix-core time dominates the Gram, so it computes only the tiles on
// and above the diagonal of the Hermitian H, each from three real products
// instead of four (Gauss's form), and so does the apply, which reads its
// rotation operands from LDS at every column block (held in registers, the
// three operand arrays would halve the occupancy, which measured slower).
// Each wave accumulates a partial Gram over its columns; the four partials
// are summed in a fixed order, (w2 + w0) + (w3 + w1), through LDS.

template <int B2>
__global__ void __launch_bounds__(256)
k_jacobi_round64(const mpsx_two_site_op *ops, const int *owner, char *ws, int round, double tol,
                 double tol_inner, int inner_sweeps) {
    constexpr int b = B2 / 2;
    constexpr int RG = B2 / 16;         // 16-row groups of complex pivot rows
    constexpr int HT = RG;              // 16-row tiles of the complex Gram

    // hp: partial complex Grams of waves (0+2) and (1+3); Q reuses the space
    // once H has been formed.
    __shared__ union {
        double2 hp[2][B2][B2];
        double2 q[B2][B2 + 1];
    } sm;
    __shared__ double2 H[B2][B2 + 1];
    __shared__ int rows[B2];
    __shared__ double prm_c[b];
    __shared__ double2 prm_s[b];
    __shared__ int s_any;
    auto &Q = sm.q;

    const int pi = owner[blockIdx.x];
    const mpsx_plan p = ops[pi].plan;
    if (p.b != b || p.method != 0) return;
    if (round >= p.nb - 1) return;
    ProblemCtl *ctl = ctl_ptr(ws, p);
    if (ctl->conv) return;
    const double null2 = ctl->null2;
    const int pair = (int)(blockIdx.x - p.pair_base);
    int I, J;
    rr_pair(pair, round, p.nb, &I, &J);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    auto row_of = [&](int i) { return i < b ? I * b + i : J * b + (i - b); };
    if (tid < B2) rows[tid] = row_of(tid);
    double2 *Z = z_ptr<double>(ws, p);
    const int LD = p.LD, R = p.R;

    // ---- Gram on the matrix cores. Wave w takes the 16-column blocks
    // w, w+4, ... and lane (l16, lq) feeds column c0 + 4 lq + s of complex
    // row u*16 + l16 at k-step s (as A operand for tile row u and as B
    // operand for tile column u).
    // upper tiles u <= w only (H is Hermitian), three real products each:
    //   P1 = Xr Xr^T, P2 = Xi Xi^T, P3 = (Xr + Xi)(Xr - Xi)^T,
    //   Re H = P1 + P2, Im H = P3 - P1 + P2
    constexpr int NTL = HT * (HT + 1) / 2;
    mfma_d4 p1[NTL], p2[NTL], p3[NTL];
#pragma unroll
    for (int t = 0; t < NTL; ++t) {
        p1[t] = mfma_d4{0.0, 0.0, 0.0, 0.0};
        p2[t] = mfma_d4{0.0, 0.0, 0.0, 0.0};
        p3[t] = mfma_d4{0.0, 0.0, 0.0, 0.0};
    }
    int64_t goff[RG];
#pragma unroll
    for (int g = 0; g < RG; ++g) goff[g] = (int64_t)row_of(g * 16 + l16) * LD;
    // the next column block's operands are loaded before this block's MFMAs
    auto load_block = [&](int c0, double2 (&v)[RG][4]) {
        const int cq = c0 + 4 * lq;
        if (cq + 4 <= R) {
#pragma unroll
            for (int g = 0; g < RG; ++g)
#pragma unroll
                for (int k = 0; k < 4; ++k) v[g][k] = Z[goff[g] + cq + k];
        } else {
#pragma unroll
            for (int g = 0; g < RG; ++g)
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int c = cq + k;
                    v[g][k] = c < R ? Z[goff[g] + c] : make_double2(0.0, 0.0);
                }
        }
    };
    double2 v[RG][4], vn[RG][4];
    if (wave * 16 < R) load_block(wave * 16, v);
    for (int c0 = wave * 16; c0 < R; c0 += 64) {
        if (c0 + 64 < R) load_block(c0 + 64, vn);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            int t = 0;
#pragma unroll
            for (int u = 0; u < HT; ++u)
#pragma unroll
                for (int w = u; w < HT; ++w, ++t) {
                    const double ar = v[u][k].x, ai = v[u][k].y, br = v[w][k].x, bi = v[w][k].y;
                    p1[t] = mfma16d(ar, br, p1[t]);
                    p2[t] = mfma16d(ai, bi, p2[t]);
                    p3[t] = mfma16d(ar + ai, br - bi, p3[t]);
                }
        }
#pragma unroll
        for (int g = 0; g < RG; ++g)
#pragma unroll
            for (int k = 0; k < 4; ++k) v[g][k] = vn[g][k];
    }
    // this wave's partial H[16 u + 4 r + lq][16 w + l16] for the upper tiles
    double2 ph[NTL][4];
#pragma unroll
    for (int t = 0; t < NTL; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            ph[t][r] = make_double2(p1[t][r] + p2[t][r], p3[t][r] - p1[t][r] + p2[t][r]);
    if (wave >= 2) {
        int t = 0;
#pragma unroll
        for (int u = 0; u < HT; ++u)
#pragma unroll
            for (int w = u; w < HT; ++w, ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) sm.hp[wave - 2][u * 16 + 4 * r + lq][w * 16 + l16] = ph[t][r];
    }
    __syncthreads();
    if (wave < 2) {
        int t = 0;
#pragma unroll
        for (int u = 0; u < HT; ++u)
#pragma unroll
            for (int w = u; w < HT; ++w, ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    double2 &d = sm.hp[wave][u * 16 + 4 * r + lq][w * 16 + l16];
                    d = make_double2(d.x + ph[t][r].x, d.y + ph[t][r].y);
                }
    }
    __syncthreads();
    double2 hv[(B2 * B2 + 255) / 256];
#pragma unroll
    for (int t = 0; t < (B2 * B2 + 255) / 256; ++t) {
        const int e = tid + 256 * t;
        if (e < B2 * B2) {
            const int i = e / B2, j = e % B2;
            const int u = i <= j ? i : j, v = i <= j ? j : i;  // upper triangle, mirrored
            const double2 a0 = sm.hp[0][u][v], a1 = sm.hp[1][u][v];
            double im = a0.y + a1.y;
            if (i > j) im = -im;
            if (i == j) im = 0.0;
            hv[t] = make_double2(a0.x + a1.x, im);
        }
    }
    __syncthreads();  // hp is dead from here on; Q reuses it
#pragma unroll
    for (int t = 0; t < (B2 * B2 + 255) / 256; ++t) {
        const int e = tid + 256 * t;
        if (e < B2 * B2) {
            const int i = e / B2, j = e % B2;
            H[i][j] = hv[t];
            Q[i][j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
        }
    }
    __syncthreads();

    if (!pivot_solve<B2>(H, Q, prm_c, prm_s, s_any, ctl, null2, tol, tol_inner, inner_sweeps, 0.0))
        return;

    // ---- apply on the matrix cores: Z' = Q^H Z for the pivot rows. Lane
    // (l16, lq) of output tile tm at k-step s takes Q[4 s + lq][16 tm + l16]
    // (A operand) and input row 4 s + lq at its column (B operand); three
    // real products per step: P1 = Re(Q)^T Zr, P2 = Im(Q)^T Zi,
    // P3 = (Re Q - Im Q)^T (Zr + Zi);  Re Z' = P1 + P2, Im Z' = P3 - P1 + P2.
    const int W = p.R + p.Cp;
    int64_t roff[B2 / 4];
#pragma unroll
    for (int s = 0; s < B2 / 4; ++s) roff[s] = (int64_t)rows[4 * s + lq] * LD;
    double2 zin[B2 / 4], znx[B2 / 4];
    {
        const int c = wave * 16 + l16;
#pragma unroll
        for (int s = 0; s < B2 / 4; ++s) zin[s] = c < W ? Z[roff[s] + c] : make_double2(0.0, 0.0);
    }
    for (int c0 = wave * 16; c0 < W; c0 += 64) {
        const int c = c0 + l16;
        const bool ok = c < W;
        const int cn = c + 64;
#pragma unroll
        for (int s = 0; s < B2 / 4; ++s) znx[s] = cn < W ? Z[roff[s] + cn] : make_double2(0.0, 0.0);
        // compiler fence: the rotation operands are read from LDS inside the
        // loop (held in registers they would halve the occupancy)
        asm volatile("" ::: "memory");
        mfma_d4 a1[HT], a2[HT], a3[HT];
#pragma unroll
        for (int tm = 0; tm < HT; ++tm) {
            a1[tm] = mfma_d4{0.0, 0.0, 0.0, 0.0};
            a2[tm] = mfma_d4{0.0, 0.0, 0.0, 0.0};
            a3[tm] = mfma_d4{0.0, 0.0, 0.0, 0.0};
        }
#pragma unroll
        for (int s = 0; s < B2 / 4; ++s) {
            const double xr = zin[s].x, xi = zin[s].y, xs = zin[s].x + zin[s].y;
#pragma unroll
            for (int tm = 0; tm < HT; ++tm) {
                const double2 q = Q[4 * s + lq][16 * tm + l16];
                a1[tm] = mfma16d(q.x, xr, a1[tm]);
                a2[tm] = mfma16d(q.y, xi, a2[tm]);
                a3[tm] = mfma16d(q.x - q.y, xs, a3[tm]);
            }
        }
        if (ok) {
#pragma unroll
            for (int tm = 0; tm < HT; ++tm)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = tm * 16 + 4 * r + lq;  // complex row in the pivot
                    Z[(int64_t)rows[row] * LD + c] =
                        make_double2(a1[tm][r] + a2[tm][r], a3[tm][r] - a1[tm][r] + a2[tm][r]);
                }
        }
#pragma unroll
        for (int s = 0; s < B2 / 4; ++s) zin[s] = znx[s];
    }
}

// ---------------------------------------------------------------- Gram method
// Two-sided Jacobi on the Hermitian Gram H = X X^H (FP64), for problems
// planned with method 1 (FP64, 32-row pivots). It is the one-sided
// iteration above written on H: a round diagonalises the same pivot Grams
// with the same rotations J = diag(Q_P) and sets H <- J^H H J and
// Y <- J^H Y. The one-sided round forms each pivot's Gram from its rows,
// so it reads the pivot rows of [X | Y] twice and writes them once per
// round; here H is formed once, a round reads and writes the upper block
// tiles of H and the rows of Y, and X is rotated once at the end,
// Xt = Y X. A round has two parts: the eigensolves (one workgroup per
// pivot: convergence test and eigensolve of the pivot block of H,
// latency-bound) and the update (one workgroup per pair of pivots and per
// 64 columns of Y: matrix-core products, memory-bound). Rounds are
// pipelined: the eigensolve workgroup of a pivot of round r + 1 first
// updates (round r) the one tile that holds its cross block, and the rest of
// round r's update runs in the same launch (k_hround), so the memory-bound
// work of a round overlaps the latency of the next and the chain of a sweep
// is one launch per round. A sweep is k_hsolve(0), then k_hround(r) for
// r = 0 .. R-2, then k_hupdate(R-1, whole).
//
// H is stored as Cp x Cp (row stride Cp) with the 16x16 block tiles
// (bi <= bj) valid; of the diagonal block tiles only the upper triangles are
// kept current. Element (i, j) of the lower triangle is conj(H[j][i]).
MPSX_DEV double2 *h_ptr(char *ws, const mpsx_plan &p) {
    return reinterpret_cast<double2 *>(ws + h_off(p));
}
// Q and flags of round `round` (two sets, so that the update launches of a
// round may still read its rotations while the next round's eigensolves
// write theirs).
MPSX_DEV double2 *qmat_ptr(char *ws, const mpsx_plan &p, int round) {
    return reinterpret_cast<double2 *>(ws + qmat_off(p)) + (int64_t)(round & 1) * (p.nb / 2) * 1024;
}
MPSX_DEV int *qflag_ptr(char *ws, const mpsx_plan &p, int round) {
    return reinterpret_cast<int *>(ws + qflag_off(p)) + (int64_t)(round & 1) * (p.nb / 2);
}
// The FP32 H of a mixed problem's first phase, and the FP64 X1 = Y X that
// the same region receives at the switch to FP64.
MPSX_DEV float2 *h32_ptr(char *ws, const mpsx_plan &p) {
    return reinterpret_cast<float2 *>(ws + t_off(p));
}
MPSX_DEV double2 *t64_ptr(char *ws, const mpsx_plan &p) {
    return reinterpret_cast<double2 *>(ws + t_off(p));
}
// The copy of X (Cp x R, row stride R) that the end's X <- Y X reads: made
// in the H region after the sweeps, or, QR-preconditioned, the X0 saved
// before the QR.
MPSX_DEV const double2 *xc_ptr(char *ws, const mpsx_plan &p) {
    return p.method == 3 ? reinterpret_cast<const double2 *>(ws + qx0_off(p)) : h_ptr(ws, p);
}
// Problems the Gram round kernels work on: in the FP32 phase (F32H) the
// mixed ones (method 2), in the FP64 rounds every Gram problem (methods 1
// and 2; a mixed problem there has passed its switch to FP64).
template <bool F32H>
MPSX_DEV bool gram_on(const mpsx_plan &p) {
    return F32H ? p.method == 2 : p.method >= 1;
}
// Element access to H (Cp x Cp, row stride Cp) in the phase's precision;
// values travel as double2 either way.
template <bool F32H>
MPSX_DEV double2 h_load(char *ws, const mpsx_plan &p, int64_t idx) {
    if (F32H) {
        const float2 v = h32_ptr(ws, p)[idx];
        return make_double2(v.x, v.y);
    }
    return h_ptr(ws, p)[idx];
}
template <bool F32H>
MPSX_DEV void h_store(char *ws, const mpsx_plan &p, int64_t idx, double2 v) {
    if (F32H)
        h32_ptr(ws, p)[idx] = make_float2((float)v.x, (float)v.y);
    else
        h_ptr(ws, p)[idx] = v;
}

// Block paired with block `item` in round r of the circle method over n
// blocks (the inverse of rr_item, then the slot opposite).
MPSX_DEV int rr_partner(int item, int r, int n) {
    const int pos = item == 0 ? 0 : ((item - 1 - r) % (n - 1) + (n - 1)) % (n - 1) + 1;
    return rr_item(n - 1 - pos, r, n);
}
// Pivot (slot pair) that holds block `item` in round r.
MPSX_DEV int rr_pivot(int item, int r, int n) {
    const int pos = item == 0 ? 0 : ((item - 1 - r) % (n - 1) + (n - 1)) % (n - 1) + 1;
    return pos < n - 1 - pos ? pos : n - 1 - pos;
}

// The Gram method's products run on one of two kernels by problem size:
// the direct ones below (k_hgram: one 16x16 block tile of H per workgroup;
// k_hxgemm: one 32x32 tile), whose many small workgroups fill the GPU for
// problems of up to kTiledGemmRows padded rows, and k_hgemm (64x64 tiles,
// operands staged through LDS), which re-reads the operands four times less
// and wins on the large ones (theta of 3072 rows: H in 1.7 s instead of 4.6 s
// over a 48-qutrit circuit).
// The choice goes by the Jacobi's rows, 16 nb: for a tall QR-preconditioned
// problem they are fewer than the working matrix's Cp rows, and the
// launchers size their grids from nb too.
constexpr int kTiledGemmRows = 512;
MPSX_DEV bool tiled_gemm(const mpsx_plan &p) { return 16 * p.nb > kTiledGemmRows; }

// One workgroup per upper 16x16 block tile (bi <= bj) of H = X X^H. x: the
// problem's pivot slot j (nb/2 per problem), y in [0, nb]; the tile is
// number j (nb + 1) + y of the nb (nb + 1) / 2 upper tiles in row-major
// order. Waves take 16-column blocks w, w + 4, ...; three real products per
// tile as in the round kernel's Gram; partials summed in a fixed order.
// STAGE 0: X is the working matrix's X part (row stride LD); H goes to the
// H region (method 1) or, rounded, to H32 (method 2). STAGE 1, the switch
// of a mixed problem to FP64: X is X1 = Y X in region T (row stride R), H
// goes to the H region, and the problem's convergence state is reset for
// the FP64 sweeps.
template <int STAGE>
__global__ void __launch_bounds__(256)
k_hgram(const mpsx_two_site_op *ops, const int *owner, char *ws) {
    __shared__ double2 part[4][16][17];
    const int pi = owner[blockIdx.x];
    const mpsx_plan p = ops[pi].plan;
    if ((STAGE == 0 ? p.method < 1 : p.method != 2) || tiled_gemm(p)) return;
    if (STAGE == 1 && blockIdx.x == p.pair_base && blockIdx.y == 0 && threadIdx.x == 0) {
        ProblemCtl *ctl = ctl_ptr(ws, p);
        ctl->conv = 0;
        ctl->counter = 0;
    }
    const int nb = p.nb;
    if ((int)blockIdx.y > nb) return;
    int rem = (int)(blockIdx.x - p.pair_base) * (nb + 1) + (int)blockIdx.y;
    int bi = 0;
    while (rem >= nb - bi) {
        rem -= nb - bi;
        ++bi;
    }
    const int bj = bi + rem;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const double2 *Z = STAGE == 0 ? z_ptr<double>(ws, p) : t64_ptr(ws, p);
    const int LD = STAGE == 0 ? p.LD : p.R, R = p.R;
    const int64_t ra = (int64_t)(bi * 16 + l16) * LD, rb = (int64_t)(bj * 16 + l16) * LD;
    mfma_d4 p1 = {0.0, 0.0, 0.0, 0.0}, p2 = p1, p3 = p1;
    // A QR-preconditioned problem's X is R, upper trapezoidal (row i is zero
    // left of column i): the 16-column blocks left of row block bj hold zeros
    // in the tile's second operand and add exact zeros, so they are skipped
    // (the sums are bitwise those of the full loop).
    int cs = wave * 16;
    if (STAGE == 0 && p.method == 3 && 16 * bj >= cs + 16) cs += 64 * ((16 * bj - 16 - cs) / 64 + 1);
    for (int c0 = cs; c0 < R; c0 += 64) {
        const int cq = c0 + 4 * lq;
        double2 va[4], vb[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int c = cq + k;
            va[k] = c < R ? Z[ra + c] : make_double2(0.0, 0.0);
            vb[k] = c < R ? Z[rb + c] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            p1 = mfma16d(va[k].x, vb[k].x, p1);
            p2 = mfma16d(va[k].y, vb[k].y, p2);
            p3 = mfma16d(va[k].x + va[k].y, vb[k].x - vb[k].y, p3);
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
        part[wave][4 * r + lq][l16] = make_double2(p1[r] + p2[r], p3[r] - p1[r] + p2[r]);
    __syncthreads();
    const int i = tid >> 4, j = tid & 15;
    const double2 a = part[0][i][j], b = part[1][i][j], c = part[2][i][j], d = part[3][i][j];
    double2 v = make_double2((a.x + c.x) + (b.x + d.x), (a.y + c.y) + (b.y + d.y));
    if (bi == bj && i == j) v.y = 0.0;
    const int64_t idx = (int64_t)(bi * 16 + i) * p.Cp + bj * 16 + j;
    if (STAGE == 0 && p.method == 2)
        h_store<true>(ws, p, idx, v);
    else
        h_ptr(ws, p)[idx] = v;
}

// Workgroup size of the eigensolve kernel (eight waves measured slower than
// four: the per-step cost is latency, not issue).
constexpr int HSOLVE_THREADS = 256;

constexpr int HSOLVE_NE = 32 * 32 / HSOLVE_THREADS;  // pivot block elements per thread
static_assert((32 * 32) % HSOLVE_THREADS == 0, "pivot block loads per thread");

// This thread's elements (i, j), i <= j, of the 32x32 pivot block of blocks
// I < J: element e = tid + 256 t of the block in row-major order lies in a
// stored tile, (I, I), (I, J) or (J, J), read row-contiguously.
template <bool F32H>
MPSX_DEV void hsolve_load(char *ws, const mpsx_plan &p, int I, int J, double2 (&hv)[HSOLVE_NE]) {
    constexpr int B2 = 32, b = 16;
    const int Cp = p.Cp;
    auto grow = [&](int i) { return i < b ? I * b + i : J * b + (i - b); };
#pragma unroll
    for (int t = 0; t < HSOLVE_NE; ++t) {
        const int e = threadIdx.x + HSOLVE_THREADS * t, i = e / B2, j = e % B2;
        if (i <= j) hv[t] = h_load<F32H>(ws, p, (int64_t)grow(i) * Cp + grow(j));
    }
}

// The eigensolve of pivot `pair` (blocks I < J) of round `round` from the
// loaded upper triangle hv of its 32x32 block of H: runs the pivot
// eigensolve and writes Q and the rotated block (Q^H H_PP Q, which the
// eigensolve leaves in LDS) back; the pivot's flag tells the update whether
// Q is the identity. The absolute floor abs_rel * ||theta||_F^2 sits above
// the FP64 rounding H gathers over the rounds (pairs below it would
// otherwise keep rotating on noise; their residual coupling changes the
// truncated factors only to second order).
template <bool F32H>
MPSX_DEV void hsolve_core(const mpsx_plan &p, char *ws, int round, int pair, int I, int J,
                          const double2 (&hv)[HSOLVE_NE], double tol, double tol_inner, int inner_sweeps,
                          double abs_rel, double2 (*H)[33], double2 (*Q)[33], double *prm_c,
                          double2 *prm_s, int &s_any) {
    constexpr int B2 = 32, b = 16, NT = HSOLVE_THREADS, NE = HSOLVE_NE;
    ProblemCtl *ctl = ctl_ptr(ws, p);
    auto grow = [&](int i) { return i < b ? I * b + i : J * b + (i - b); };
    const int Cp = p.Cp, tid = threadIdx.x;
    // mirrored into the lower triangle
#pragma unroll
    for (int t = 0; t < NE; ++t) {
        const int e = tid + NT * t, i = e / B2, j = e % B2;
        Q[i][j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
        if (i > j) continue;
        double2 h = hv[t];
        if (i == j) h.y = 0.0;
        H[i][j] = h;
        H[j][i] = make_double2(h.x, -h.y);
    }
    __syncthreads();
    const double abs_t = abs_rel * ctl->fro2;
    int *flag = qflag_ptr(ws, p, round);
    if (!pivot_solve<B2, NT>(H, Q, prm_c, prm_s, s_any, ctl, ctl->null2, tol, tol_inner, inner_sweeps,
                         abs_t * abs_t)) {
        if (tid == 0) flag[pair] = 0;
        return;
    }
    double2 *Qg = qmat_ptr(ws, p, round) + (int64_t)pair * B2 * B2;
    for (int e = tid; e < B2 * B2; e += NT) {
        const int i = e / B2, j = e % B2;
        Qg[e] = Q[i][j];
        if (i <= j) h_store<F32H>(ws, p, (int64_t)grow(i) * Cp + grow(j), H[i][j]);
    }
    if (tid == 0) flag[pair] = 1;
}

// Round `round`, eigensolves: one workgroup per pivot. Loads the pivot's
// 32x32 block of H (all loads in flight together with the convergence flag:
// the addresses are valid either way) and solves it.
template <bool F32H>
MPSX_DEV void hsolve_body(const mpsx_two_site_op *ops, const int *owner, char *ws, int round,
                          double tol, double tol_inner, int inner_sweeps, double abs_rel,
                          double2 (*H)[33], double2 (*Q)[33], double *prm_c, double2 *prm_s,
                          int &s_any) {
    const int pi = owner[blockIdx.x];
    const mpsx_plan p = ops[pi].plan;
    if (!gram_on<F32H>(p) || round >= p.nb - 1) return;
    const int pair = (int)(blockIdx.x - p.pair_base);
    int I, J;
    rr_pair(pair, round, p.nb, &I, &J);
    double2 hv[HSOLVE_NE];
    hsolve_load<F32H>(ws, p, I, J, hv);
    if (ctl_ptr(ws, p)->conv) return;
    hsolve_core<F32H>(p, ws, round, pair, I, J, hv, tol, tol_inner, inner_sweeps, abs_rel, H, Q, prm_c,
                      prm_s, s_any);
}

template <bool F32H>
__global__ void __launch_bounds__(HSOLVE_THREADS)
k_hsolve(const mpsx_two_site_op *ops, const int *owner, char *ws, int round, double tol,
         double tol_inner, int inner_sweeps, double abs_rel) {
    __shared__ double2 H[32][33];
    __shared__ double2 Q[32][33];
    __shared__ double prm_c[16];
    __shared__ double2 prm_s[16];
    __shared__ int s_any;
    hsolve_body<F32H>(ops, owner, ws, round, tol, tol_inner, inner_sweeps, abs_rel, H, Q, prm_c, prm_s,
                      s_any);
}

// Result of an H tile update: the LDS array holding it.
using HTile = double2[33];

// T <- Q_P^H T Q_P2 in place for the 32x32 tile of H between pivots P != P2
// of round `round` (row blocks I, J of P; column blocks I2, J2 of P2; a
// factor whose pivot did not rotate is the identity and skipped). FP64
// 16x16x4 MFMAs, three real products per complex product; wave w owns output
// tile (w / 2, w % 2). The tile passes through LDS (Ts: T, then the result;
// Ws: W = Q_P^H T). Returns the array holding the result, or nullptr when
// nothing was written (problem converged, or neither pivot rotated). Uniform
// over the workgroup.
MPSX_DEV const HTile *htile_update(const mpsx_plan &p, char *ws, int round, int P, int P2,
                                   double2 (*Ts)[33], double2 (*Ws)[33]) {
    constexpr int b = 16;
    const int *convp = &ctl_ptr(ws, p)->conv;  // read with the data loads below
    const int *flag = qflag_ptr(ws, p, round);
    const double2 *Qall = qmat_ptr(ws, p, round);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    int I, J, I2, J2;
    rr_pair(P, round, p.nb, &I, &J);
    rr_pair(P2, round, p.nb, &I2, &J2);
    double2 *Hg = h_ptr(ws, p);
    const int Cp = p.Cp;
    // The tile's four 16x16 block tiles (row block a of P: I or J, column
    // block c of P2: I2 or J2) are stored as H[bi][bj] when bi < bj, else
    // as the conjugate transpose in H[bj][bi]; both are read and written
    // row-contiguously (256 B per block row) and transposed in LDS. The
    // loads are issued together with the convergence and rotation flags
    // (valid addresses either way).
    const int ii = tid >> 4, jj = tid & 15;  // one element of each block tile per thread
    int64_t gaddr[4];
    bool lower[4];
    double2 tv[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int bi = (t >> 1) ? J : I, bj = (t & 1) ? J2 : I2;
        lower[t] = bi > bj;
        gaddr[t] = lower[t] ? (int64_t)(bj * b + ii) * Cp + bi * b + jj
                            : (int64_t)(bi * b + ii) * Cp + bj * b + jj;
        tv[t] = Hg[gaddr[t]];
    }
    const int fa = flag[P], fb = flag[P2];
    const int tm = wave >> 1, tn = wave & 1;
    // this lane's rotation operands of both products, loaded up front
    const double2 *Qa = Qall + (int64_t)P * 1024, *Qb = Qall + (int64_t)P2 * 1024;
    double2 qa[8], qb[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int k = 4 * s + lq;
        qa[s] = Qa[k * 32 + 16 * tm + l16];  // conj: A = Q_P^H[16 tm + l16][k]
        qb[s] = Qb[k * 32 + 16 * tn + l16];  // B = Q_P2[k][16 tn + l16]
    }
    if (*convp || (!fa && !fb)) return nullptr;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const double2 v = tv[t];
        const int r0 = 16 * (t >> 1), c0 = 16 * (t & 1);
        if (lower[t])
            Ts[r0 + jj][c0 + ii] = make_double2(v.x, -v.y);
        else
            Ts[r0 + ii][c0 + jj] = v;
    }
    __syncthreads();
    // W = Q_P^H T: register r of lane (l16, lq) is W[16 tm + 4 r + lq][16 tn + l16]
    if (fa) {
        mfma_d4 a1 = {0.0, 0.0, 0.0, 0.0}, a2 = a1, a3 = a1;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int k = 4 * s + lq;
            const double2 q = qa[s];                   // conj: A = Q^H[16 tm + l16][k]
            const double2 t = Ts[k][16 * tn + l16];    // B = T[k][16 tn + l16]
            a1 = mfma16d(q.x, t.x, a1);
            a2 = mfma16d(q.y, t.y, a2);
            a3 = mfma16d(q.x - q.y, t.x + t.y, a3);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
            Ws[16 * tm + 4 * r + lq][16 * tn + l16] =
                make_double2(a1[r] + a2[r], a3[r] - a1[r] + a2[r]);
        __syncthreads();
    }
    if (fb) {
        const double2(*Wsrc)[33] = fa ? Ws : Ts;
        mfma_d4 c1 = {0.0, 0.0, 0.0, 0.0}, c2 = c1, c3 = c1;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int k = 4 * s + lq;
            const double2 wa = Wsrc[16 * tm + l16][k];  // A = W[16 tm + l16][k]
            const double2 qv = qb[s];                   // B = Q_P2[k][16 tn + l16]
            c1 = mfma16d(wa.x, qv.x, c1);
            c2 = mfma16d(wa.y, qv.y, c2);
            c3 = mfma16d(wa.x + wa.y, qv.x + qv.y, c3);
        }
        __syncthreads();  // every wave has read Ts (as W) before it is overwritten
#pragma unroll
        for (int r = 0; r < 4; ++r)
            Ts[16 * tm + 4 * r + lq][16 * tn + l16] =
                make_double2(c1[r] - c2[r], c3[r] - c1[r] - c2[r]);
        __syncthreads();
    }
    const HTile *Out = fb ? Ts : Ws;  // fa only: the result is W
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int r0 = 16 * (t >> 1), c0 = 16 * (t & 1);
        if (lower[t]) {
            const double2 v = Out[r0 + jj][c0 + ii];
            Hg[gaddr[t]] = make_double2(v.x, -v.y);
        } else {
            Hg[gaddr[t]] = Out[r0 + ii][c0 + jj];
        }
    }
    return Out;
}

// The same tile update for the FP32 phase of a mixed problem: the tile is
// read from and written to H32 and the products run on the FP32 matrix
// cores (v_mfma_f32_16x16x4f32: operands laid out as in the FP64
// instruction; result register r of lane l is D[4*(l/16) + r][l%16]), with
// the FP64 rotations rounded to FP32 as they are loaded. Ts and Ws are
// float2 views of the workgroup's LDS tiles.
using HTile32 = float2[33];
MPSX_DEV const HTile32 *htile_update32(const mpsx_plan &p, char *ws, int round, int P, int P2,
                                       float2 (*Ts)[33], float2 (*Ws)[33]) {
    constexpr int b = 16;
    const int *convp = &ctl_ptr(ws, p)->conv;
    const int *flag = qflag_ptr(ws, p, round);
    const double2 *Qall = qmat_ptr(ws, p, round);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    int I, J, I2, J2;
    rr_pair(P, round, p.nb, &I, &J);
    rr_pair(P2, round, p.nb, &I2, &J2);
    float2 *Hg = h32_ptr(ws, p);
    const int Cp = p.Cp;
    const int ii = tid >> 4, jj = tid & 15;
    int64_t gaddr[4];
    bool lower[4];
    float2 tv[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int bi = (t >> 1) ? J : I, bj = (t & 1) ? J2 : I2;
        lower[t] = bi > bj;
        gaddr[t] = lower[t] ? (int64_t)(bj * b + ii) * Cp + bi * b + jj
                            : (int64_t)(bi * b + ii) * Cp + bj * b + jj;
        tv[t] = Hg[gaddr[t]];
    }
    const int fa = flag[P], fb = flag[P2];
    const int tm = wave >> 1, tn = wave & 1;
    const double2 *Qa = Qall + (int64_t)P * 1024, *Qb = Qall + (int64_t)P2 * 1024;
    float2 qa[8], qb[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int k = 4 * s + lq;
        const double2 a = Qa[k * 32 + 16 * tm + l16];  // conj: A = Q_P^H[16 tm + l16][k]
        const double2 c = Qb[k * 32 + 16 * tn + l16];  // B = Q_P2[k][16 tn + l16]
        qa[s] = make_float2((float)a.x, (float)a.y);
        qb[s] = make_float2((float)c.x, (float)c.y);
    }
    if (*convp || (!fa && !fb)) return nullptr;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const float2 v = tv[t];
        const int r0 = 16 * (t >> 1), c0 = 16 * (t & 1);
        if (lower[t])
            Ts[r0 + jj][c0 + ii] = make_float2(v.x, -v.y);
        else
            Ts[r0 + ii][c0 + jj] = v;
    }
    __syncthreads();
    // W = Q_P^H T: register r of lane (l16, lq) is W[16 tm + 4 lq + r][16 tn + l16]
    if (fa) {
        mfma_f4 a1 = {0.f, 0.f, 0.f, 0.f}, a2 = a1, a3 = a1;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int k = 4 * s + lq;
            const float2 q = qa[s];
            const float2 t = Ts[k][16 * tn + l16];
            a1 = mfma16(q.x, t.x, a1);
            a2 = mfma16(q.y, t.y, a2);
            a3 = mfma16(q.x - q.y, t.x + t.y, a3);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
            Ws[16 * tm + 4 * lq + r][16 * tn + l16] = make_float2(a1[r] + a2[r], a3[r] - a1[r] + a2[r]);
        __syncthreads();
    }
    if (fb) {
        const float2(*Wsrc)[33] = fa ? Ws : Ts;
        mfma_f4 c1 = {0.f, 0.f, 0.f, 0.f}, c2 = c1, c3 = c1;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int k = 4 * s + lq;
            const float2 wa = Wsrc[16 * tm + l16][k];
            const float2 qv = qb[s];
            c1 = mfma16(wa.x, qv.x, c1);
            c2 = mfma16(wa.y, qv.y, c2);
            c3 = mfma16(wa.x + wa.y, qv.x + qv.y, c3);
        }
        __syncthreads();  // every wave has read Ts (as W) before it is overwritten
#pragma unroll
        for (int r = 0; r < 4; ++r)
            Ts[16 * tm + 4 * lq + r][16 * tn + l16] = make_float2(c1[r] - c2[r], c3[r] - c1[r] - c2[r]);
        __syncthreads();
    }
    const HTile32 *Out = fb ? Ts : Ws;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int r0 = 16 * (t >> 1), c0 = 16 * (t & 1);
        if (lower[t]) {
            const float2 v = Out[r0 + jj][c0 + ii];
            Hg[gaddr[t]] = make_float2(v.x, -v.y);
        } else {
            Hg[gaddr[t]] = Out[r0 + ii][c0 + jj];
        }
    }
    return Out;
}

// Round `round`, update (`part` 0: all of it; 2: all but the tiles that hold
// the cross block of a pivot of round + 1, which the workgroups of round
// + 1's eigensolves update themselves, see hnext_body). x: pivot P of a
// problem; y < n/2 (n = nb/2 pivots): the 32x32 tile of H between P and
// P2 = (P + y + 1) mod n, each unordered pair once; y >= n/2: columns
// 64 (y - n/2) .. +63 of the Y rows of P, Y <- Q_P^H Y (FP64 MFMAs, wave w
// owns 16 columns of all 32 rows).
template <bool F32H>
MPSX_DEV void hupdate_body(const mpsx_two_site_op *ops, const int *owner, char *ws, int round,
                           int part, int y, double2 (*Ts)[33], double2 (*Ws)[33]) {
    constexpr int b = 16;
    const int pi = owner[blockIdx.x];
    const mpsx_plan p = ops[pi].plan;
    if (!gram_on<F32H>(p) || round >= p.nb - 1) return;
    const int n = p.nb / 2, nt = n / 2, nb = p.nb;
    const int P = (int)(blockIdx.x - p.pair_base);
    int I, J;
    rr_pair(P, round, nb, &I, &J);
    if (y < nt) {
        const int dd = y + 1;
        if (2 * dd == n && P >= nt) return;
        const int P2 = (P + dd) % n;
        if (part == 2) {  // a tile holding a pivot pair of the next round is updated with its eigensolve
            int I2, J2;
            rr_pair(P2, round, nb, &I2, &J2);
            const int pi_ = rr_partner(I, round + 1, nb), pj_ = rr_partner(J, round + 1, nb);
            if (pi_ == I2 || pi_ == J2 || pj_ == I2 || pj_ == J2) return;
        }
        if (F32H)
            htile_update32(p, ws, round, P, P2, reinterpret_cast<float2(*)[33]>(Ts),
                           reinterpret_cast<float2(*)[33]>(Ws));
        else
            htile_update(p, ws, round, P, P2, Ts, Ws);
        return;
    }
    // Y rows of pivot P, 16 columns per wave (FP64 in both phases)
    const int c0 = (y - nt) * 64;
    const int *flag = qflag_ptr(ws, p, round);
    if (c0 >= p.Cp || ctl_ptr(ws, p)->conv || !flag[P]) return;
    const double2 *Qa = qmat_ptr(ws, p, round) + (int64_t)P * 1024;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    double2 *Z = z_ptr<double>(ws, p);
    const int LD = p.LD, R = p.R;
    const int c = c0 + 16 * wave + l16;
    const bool ok = c < p.Cp;
    int64_t roff[8];
    double2 v[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int i = 4 * s + lq;
        roff[s] = (int64_t)(i < b ? I * b + i : J * b + (i - b)) * LD + R;
        v[s] = ok ? Z[roff[s] + c] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int tm = 0; tm < 2; ++tm) {
        mfma_d4 a1 = {0.0, 0.0, 0.0, 0.0}, a2 = a1, a3 = a1;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const double2 q = Qa[(4 * s + lq) * 32 + 16 * tm + l16];  // conj: A = Q^H[16 tm + l16][k]
            a1 = mfma16d(q.x, v[s].x, a1);
            a2 = mfma16d(q.y, v[s].y, a2);
            a3 = mfma16d(q.x - q.y, v[s].x + v[s].y, a3);
        }
        if (ok) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 16 * tm + 4 * r + lq;
                const int64_t ro = (int64_t)(i < b ? I * b + i : J * b + (i - b)) * LD + R;
                Z[ro + c] = make_double2(a1[r] + a2[r], a3[r] - a1[r] + a2[r]);
            }
        }
    }
}

template <bool F32H>
__global__ void __launch_bounds__(256, 4)
k_hupdate(const mpsx_two_site_op *ops, const int *owner, char *ws, int round, int part) {
    __shared__ double2 Ts[32][33];
    __shared__ double2 Ws[32][33];
    hupdate_body<F32H>(ops, owner, ws, round, part, (int)blockIdx.y, Ts, Ws);
}

// The eigensolve of pivot Pn of round + 1, preceded by the round `round`
// update of the one H tile that holds its cross block: blocks In < Jn of Pn
// were held by pivots pa and pb of this round, and the tile between them is
// updated here (when pa != pb; when pa == pb that pivot's eigensolve has
// written the block). For nb >= 6 no such tile holds the cross blocks of two
// pivots of round + 1 (the circle method sends the two blocks of a pivot to
// pivots two slots apart), so this workgroup is the only one in the launch
// that reads or writes the tile; problems with nb = 4 use the one-sided
// method. The diagonal block tiles (In, In) and (Jn, Jn) were written by this
// round's eigensolves, in the launch before; the cross block is taken from
// the updated tile in LDS. The tile is updated in the
// orientation part 0 uses (P2 = P + d mod n, 1 <= d <= n/2), so that its
// arithmetic does not depend on which launch updates it.
template <bool F32H>
MPSX_DEV void hnext_body(const mpsx_two_site_op *ops, const int *owner, char *ws, int round, double tol,
                         double tol_inner, int inner_sweeps, double abs_rel, double2 (*A)[33],
                         double2 (*B)[33], double *prm_c, double2 *prm_s, int &s_any) {
    constexpr int B2 = 32, b = 16;
    const int pi = owner[blockIdx.x];
    const mpsx_plan p = ops[pi].plan;
    if (!gram_on<F32H>(p) || round >= p.nb - 1) return;
    const int nb = p.nb, n = nb / 2, nt = n / 2;
    const int Pn = (int)(blockIdx.x - p.pair_base);
    int In, Jn;
    rr_pair(Pn, round + 1, nb, &In, &Jn);
    const bool solve = round + 1 < nb - 1;
    const int pa = rr_pivot(In, round, nb), pb = rr_pivot(Jn, round, nb);
    const HTile *Out = nullptr;
    const HTile32 *Out32 = nullptr;
    int P = pa, P2 = pb;
    if (pa != pb) {
        const int d = (pb - pa + n) % n;
        const bool keep = 2 * d < n || (2 * d == n && pa < nt);
        P = keep ? pa : pb;
        P2 = keep ? pb : pa;
        if (F32H)
            Out32 = htile_update32(p, ws, round, P, P2, reinterpret_cast<float2(*)[33]>(A),
                                   reinterpret_cast<float2(*)[33]>(B));
        else
            Out = htile_update(p, ws, round, P, P2, A, B);
    }
    if (!solve) return;
    // loaded after the update (registers), the cross block replaced from LDS
    double2 hv[HSOLVE_NE];
    hsolve_load<F32H>(ws, p, In, Jn, hv);
    auto tile = [&](int r, int c) {
        if (F32H) {
            const float2 v = Out32[r][c];
            return make_double2(v.x, v.y);
        }
        return Out[r][c];
    };
    if (Out || Out32) {
        int I, J, I2, J2;
        rr_pair(P, round, nb, &I, &J);
        rr_pair(P2, round, nb, &I2, &J2);
        const bool in_p = In == I || In == J;  // In a block of P (rows of Out), Jn of P2
#pragma unroll
        for (int t = 0; t < HSOLVE_NE; ++t) {
            const int e = threadIdx.x + HSOLVE_THREADS * t, i = e / B2, j = e % B2;
            if (i >= b || j < b) continue;
            const int jj = j - b;
            if (in_p) {
                hv[t] = tile((In == J ? b : 0) + i, (Jn == J2 ? b : 0) + jj);
            } else {  // H[In][Jn] = H[Jn][In]^H
                const double2 v = tile((Jn == J ? b : 0) + jj, (In == J2 ? b : 0) + i);
                hv[t] = make_double2(v.x, -v.y);
            }
        }
    }
    if (ctl_ptr(ws, p)->conv) return;
    __syncthreads();  // the tile in LDS has been read before the eigensolve reuses A and B
    hsolve_core<F32H>(p, ws, round + 1, Pn, In, Jn, hv, tol, tol_inner, inner_sweeps, abs_rel, A, B, prm_c,
                      prm_s, s_any);
}

// Rounds r + 1 and r in one launch: y = 0 runs the eigensolves of round
// r + 1 (each after the update of the tile holding its cross block), y >= 1
// the rest of the update of round r (the other tiles, and the rows of Y).
// The two touch disjoint tiles of H and different sets of Q and flags, so
// they overlap freely. LDS is shared between the two roles. Four waves per
// SIMD (at most 128 registers a lane, accumulators included) match the four
// workgroups per CU that the LDS allows.
template <bool F32H>
__global__ void __launch_bounds__(256, 4)
k_hround(const mpsx_two_site_op *ops, const int *owner, char *ws, int round, double tol,
         double tol_inner, int inner_sweeps, double abs_rel) {
    __shared__ double2 A[32][33];
    __shared__ double2 B[32][33];
    __shared__ double prm_c[16];
    __shared__ double2 prm_s[16];
    __shared__ int s_any;
    if (blockIdx.y == 0)
        hnext_body<F32H>(ops, owner, ws, round, tol, tol_inner, inner_sweeps, abs_rel, A, B, prm_c, prm_s,
                         s_any);
    else
        hupdate_body<F32H>(ops, owner, ws, round, 2, (int)blockIdx.y - 1, A, B);
}

// End of the Gram method, 1: copy X (Cp x R) into the H region (row stride
// R), one workgroup per padded row.
__global__ void __launch_bounds__(256)
k_hxcopy(const mpsx_two_site_op *ops, int count, char *ws) {
    const int pi = find_record(reinterpret_cast<const char *>(ops), count, sizeof(mpsx_two_site_op),
                               OP_FIELD(plan.row_base), blockIdx.x);
    const mpsx_plan p = ops[pi].plan;
    if (p.method < 1 || p.method == 3) return;  // method 3 kept X in X0
    const int row = (int)(blockIdx.x - p.row_base);
    if (row >= p.Cp) return;
    const double2 *z = z_ptr<double>(ws, p) + (int64_t)row * p.LD;
    double2 *x = h_ptr(ws, p) + (int64_t)row * p.R;
    for (int c = threadIdx.x; c < p.R; c += 256) x[c] = z[c];
}

// End of the Gram method, 2: X <- Y Xc (Xc the copy), so that the rows of X
// are those the one-sided iteration would have produced. x: pivot slot j
// = rows 32 j .. 32 j + 31 (Cp = 32 nb/2), y: 32 columns; wave w computes
// the 16x16 output tile (w / 2, w % 2) over K = Cp with three real products.
// MODE 0: that product (every Gram problem). MODE 1, the switch of a mixed
// problem to FP64: X1 = Y X into region T (row stride R); X stays.
template <int MODE>
__global__ void __launch_bounds__(256)
k_hxgemm(const mpsx_two_site_op *ops, const int *owner, char *ws) {
    const int pi = owner[blockIdx.x];
    const mpsx_plan p = ops[pi].plan;
    if ((MODE == 0 ? p.method < 1 : p.method != 2) || tiled_gemm(p)) return;
    const int R = p.R, LD = p.LD, Cp = p.Cp;
    const int col0 = (int)blockIdx.y * 32;
    if (col0 >= R) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int tm = wave >> 1, tn = wave & 1;
    const int row0 = (int)(blockIdx.x - p.pair_base) * 32 + 16 * tm;
    const int col = col0 + 16 * tn + l16;
    const bool ok = col < R;
    double2 *Z = z_ptr<double>(ws, p);
    const double2 *Xc = MODE == 0 ? xc_ptr(ws, p) : Z;  // B = X[k][col]
    const int ldx = MODE == 0 ? R : LD;
    double2 *Out = MODE == 0 ? Z : t64_ptr(ws, p);
    const int ldo = MODE == 0 ? LD : R;
    const double2 *yrow = Z + (int64_t)(row0 + l16) * LD + R;  // A = Y[row0 + l16][k]
    mfma_d4 a1 = {0.0, 0.0, 0.0, 0.0}, a2 = a1, a3 = a1;
    for (int k0 = 0; k0 < Cp; k0 += 16) {
        double2 ya[4], xb[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int k = k0 + 4 * s + lq;
            ya[s] = yrow[k];
            xb[s] = ok ? Xc[(int64_t)k * ldx + col] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            a1 = mfma16d(ya[s].x, xb[s].x, a1);
            a2 = mfma16d(ya[s].y, xb[s].y, a2);
            a3 = mfma16d(ya[s].x + ya[s].y, xb[s].x + xb[s].y, a3);
        }
    }
    if (ok) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
            Out[(int64_t)(row0 + 4 * r + lq) * ldo + col] =
                make_double2(a1[r] - a2[r], a3[r] - a1[r] - a2[r]);
    }
}

// The Gram method's products, one 64x64 output tile per workgroup, FP64
// matrix cores (three real products per complex product), K staged through
// LDS 16 at a time (each staged operand is read once per tile instead of
// once per 16x16 output block). x: pivot slot j of the problem, the output
// rows 64 j .. 64 j + 63 (slots beyond Cp / 64 exit); y: 64 output columns.
// Wave w computes the 32x32 quadrant (w / 2, w % 2) as 2 x 2 MFMA tiles.
//   MODE 0: H = X X^H (upper 64x64 block tiles; H region, or rounded to
//           H32 for a mixed problem), X the working matrix's X part.
//   MODE 1: the switch of a mixed problem to FP64: H = X1 X1^H with X1 in
//           region T (row stride R); resets the problem's convergence state.
//   MODE 2: the end: X <- Y Xc (Xc the copy of X in the H region, row
//           stride R), so the rows of X are those the one-sided iteration
//           would have produced.
//   MODE 3: the switch of a mixed problem: X1 = Y X into region T.
template <int MODE>
__global__ void __launch_bounds__(256)
k_hgemm(const mpsx_two_site_op *ops, const int *owner, char *ws) {
    constexpr int TM = 64, TK = 16;
    __shared__ double2 As[TM][TK + 1];  // A[i][k]
    __shared__ double2 Bs[TK][TM + 1];  // B[k][j]
    const int pi = owner[blockIdx.x];
    const mpsx_plan p = ops[pi].plan;
    const bool mixed_only = MODE == 1 || MODE == 3;
    if ((mixed_only ? p.method != 2 : p.method < 1) || !tiled_gemm(p)) return;
    if (MODE == 1 && blockIdx.x == p.pair_base && blockIdx.y == 0 && threadIdx.x == 0) {
        ProblemCtl *ctl = ctl_ptr(ws, p);
        ctl->conv = 0;
        ctl->counter = 0;
    }
    const bool gram = MODE <= 1;
    const int Cp = p.Cp, R = p.R, LD = p.LD;
    const int i0 = (int)(blockIdx.x - p.pair_base) * TM, j0 = (int)blockIdx.y * TM;
    const int ncols = gram ? Cp : R, K = gram ? R : Cp;
    if (i0 >= Cp || j0 >= ncols || (gram && j0 < i0)) return;
    double2 *Z = z_ptr<double>(ws, p);
    // row-major sources: A rows (length K), and for the Gram the rows of X
    // that form B = X^H; for MODES 2, 3 B = X[k][j]
    const double2 *Asrc = gram ? (MODE == 0 ? Z : t64_ptr(ws, p)) : Z + R;
    const int lda = gram ? (MODE == 0 ? LD : R) : LD;
    const double2 *Bsrc = gram ? Asrc : (MODE == 2 ? xc_ptr(ws, p) : Z);
    const int ldb = gram ? lda : (MODE == 2 ? R : LD);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int wm = wave >> 1, wn = wave & 1;
    mfma_d4 acc[2][2][3];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[a][b][c] = mfma_d4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += TK) {
        // stage A (64 x 16) and B (16 x 64): four complex entries per thread each
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int e = tid + 256 * t;
            const int ai = e >> 4, ak = e & 15;  // A[i][k], row-contiguous in k
            const int gk = k0 + ak;
            As[ai][ak] = (i0 + ai < Cp && gk < K) ? Asrc[(int64_t)(i0 + ai) * lda + gk]
                                                  : make_double2(0.0, 0.0);
            if (gram) {  // B[k][j] = conj(X[j0 + j][k]): read X rows contiguously in k
                const int bj = e >> 4, bk = e & 15;
                const double2 v = (j0 + bj < ncols && k0 + bk < K)
                                      ? Bsrc[(int64_t)(j0 + bj) * ldb + k0 + bk]
                                      : make_double2(0.0, 0.0);
                Bs[bk][bj] = make_double2(v.x, -v.y);
            } else {     // B[k][j], row-contiguous in j
                const int bk = e >> 6, bj = e & 63;
                Bs[bk][bj] = (k0 + bk < K && j0 + bj < ncols)
                                 ? Bsrc[(int64_t)(k0 + bk) * ldb + j0 + bj]
                                 : make_double2(0.0, 0.0);
            }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < TK / 4; ++s) {
            const int kk = 4 * s + lq;
            double2 av[2], bv[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) av[a] = As[32 * wm + 16 * a + l16][kk];
#pragma unroll
            for (int b = 0; b < 2; ++b) bv[b] = Bs[kk][32 * wn + 16 * b + l16];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    acc[a][b][0] = mfma16d(av[a].x, bv[b].x, acc[a][b][0]);
                    acc[a][b][1] = mfma16d(av[a].y, bv[b].y, acc[a][b][1]);
                    acc[a][b][2] = mfma16d(av[a].x + av[a].y, bv[b].x + bv[b].y, acc[a][b][2]);
                }
        }
        __syncthreads();
    }
    // C = (p1 - p2) + i (p3 - p1 - p2); register r of lane (l16, lq) is
    // C[16 a + 4 r + lq][16 b + l16] of the wave's quadrant
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + 32 * wm + 16 * a + 4 * r + lq;
                const int j = j0 + 32 * wn + 16 * b + l16;
                if (i >= Cp || j >= ncols) continue;
                double2 v = make_double2(acc[a][b][0][r] - acc[a][b][1][r],
                                         acc[a][b][2][r] - acc[a][b][0][r] - acc[a][b][1][r]);
                if (gram) {
                    if (i == j) v.y = 0.0;
                    const int64_t idx = (int64_t)i * Cp + j;
                    if (MODE == 0 && p.method == 2)
                        h_store<true>(ws, p, idx, v);
                    else
                        h_ptr(ws, p)[idx] = v;
                } else if (MODE == 2) {
                    Z[(int64_t)i * LD + j] = v;
                } else {
                    t64_ptr(ws, p)[(int64_t)i * R + j] = v;
                }
            }
}

// Round workgroup -> problem table (one workgroup per problem), so that the
// round kernels need no search.
__global__ void k_owner(const mpsx_two_site_op *ops, int count, int *owner) {
    const int pi = blockIdx.x;
    if (pi >= count) return;
    const mpsx_plan p = ops[pi].plan;
    for (int j = threadIdx.x; j < p.nb / 2; j += blockDim.x) owner[p.pair_base + j] = pi;
}

// After each sweep: a problem whose sweep rotated nothing is converged.
// `sweep` counts the sweeps of both phases of the call (it indexes the
// header's active counters, reused modulo 64: the next one is cleared
// here). In the FP32 phase (fp32_phase) only the mixed Gram problems swept.
__global__ void k_sweep_end(const mpsx_two_site_op *ops, int count, char *ws, int sweep,
                            int fp32_phase) {
    const int pi = blockIdx.x * blockDim.x + threadIdx.x;
    if (pi == 0) reinterpret_cast<TwoSiteHeader *>(ws)->active[(sweep + 1) & 63] = 0;
    if (pi >= count) return;
    const mpsx_plan p = ops[pi].plan;
    if (fp32_phase && p.method != 2) return;
    ProblemCtl *ctl = ctl_ptr(ws, p);
    if (ctl->conv) return;
    if (ctl->counter == 0) {
        ctl->conv = 1;
    } else {
        ctl->sweeps += 1;
        ctl->counter = 0;
        atomicAdd(&reinterpret_cast<TwoSiteHeader *>(ws)->active[sweep & 63], 1);
    }
}

// ---------------------------------------------------------------- finalize
// One workgroup per problem: stable descending order of the row norms
// (ties keep the lower row index), singular values, discarded weight, and
// a count of problems left unconverged (when `unconverged` is not null).
constexpr int SORT_LDS = 4096;
__global__ void __launch_bounds__(256)
k_sort(const mpsx_two_site_op *ops, int count, char *ws, int *unconverged) {
    __shared__ double sg[SORT_LDS];
    __shared__ double red[4];
    const int pi = blockIdx.x;
    if (pi >= count) return;
    const mpsx_two_site_op &op = ops[pi];
    const mpsx_plan p = op.plan;
    const int C = p.C, k = op.k;
    const double *sig = sig_ptr(ws, p);
    int *perm = perm_ptr(ws, p);
    const bool in_lds = C <= SORT_LDS;
    if (in_lds)
        for (int i = threadIdx.x; i < C; i += 256) sg[i] = sig[i];
    __syncthreads();
    double disc = 0.0;
    for (int i = threadIdx.x; i < C; i += 256) {
        const double si = in_lds ? sg[i] : sig[i];
        int rank = 0;
        for (int j = 0; j < C; ++j) {
            const double sj = in_lds ? sg[j] : sig[j];
            rank += (sj > si) || (sj == si && j < i);
        }
        perm[rank] = i;
        if (rank >= k) disc += si;
    }
    disc = block_sum_256(disc, red);
    __syncthreads();
    if (threadIdx.x == 0 && op.discarded) *op.discarded = disc;
    if (threadIdx.x == 0 && unconverged && p.nb >= 2 && !ctl_ptr(ws, p)->conv)
        atomicAdd(unconverged, 1);
    if (op.s_out)
        for (int j = threadIdx.x; j < k; j += 256) {
            const double s2 = in_lds ? sg[perm[j]] : sig[perm[j]];
            op.s_out[j] = (float)sqrt(s2);
        }
}

// Gather of the kept rows into the two factors (S absorbed as the rows of X
// already carry it). The d*chi_l x k factor a_out takes one entry from each
// kept row of Z per output row, i.e. a transpose: 32x32 tiles are staged
// through LDS so that both the reads (32 consecutive columns of a row of Z)
// and the writes (32 consecutive entries of a row of a_out) are coalesced.
// The k x N factor b_out is a row copy, one workgroup per row.
//   absorb right: a_out[m][j] = conj(Y[perm[j]][m]),  b_out[j][n] = X[perm[j]][n]
//   absorb left:  a_out[m][j] = X[perm[j]][m],        b_out[j][n] = conj(Y[perm[j]][n])
template <class T>
__global__ void __launch_bounds__(256)
k_gather(const mpsx_two_site_op *ops, int count, char *ws) {
    using Z2 = typename Cx<T>::type;
    __shared__ Z2 tile[32][33];
    const int pi = find_record(reinterpret_cast<const char *>(ops), count, sizeof(mpsx_two_site_op),
                               OP_FIELD(plan.out_base), blockIdx.x);
    const mpsx_two_site_op &op = ops[pi];
    const mpsx_plan p = op.plan;
    const int M = op.d * op.chi_l, N = op.d * op.chi_r, k = op.k;
    const int64_t u = (int64_t)blockIdx.x - p.out_base;
    if (u >= gather_units(M, k)) return;
    const Z2 *Z = z_ptr<T>(ws, p);
    const int *perm = perm_ptr(ws, p);
    const int LD = p.LD, R = p.R;
    const bool left_absorb = (op.flags & 1) != 0;
    // complex64 outputs; the conjugate of Y is taken for the unscaled factor
    auto f2 = [](Z2 v, bool conj) { return make_float2((float)v.x, conj ? -(float)v.y : (float)v.y); };
    const int64_t ta = ((int64_t)(M + 31) / 32) * ((k + 31) / 32);
    if (u < ta) {
        const int kt = (k + 31) / 32;
        const int m0 = (int)(u / kt) * 32, j0 = (int)(u % kt) * 32;
        const int col0 = left_absorb ? 0 : R;
        for (int e = threadIdx.x; e < 32 * 32; e += 256) {
            const int jj = e >> 5, mi = e & 31;
            const int j = j0 + jj, m = m0 + mi;
            if (j < k && m < M) tile[jj][mi] = Z[(int64_t)perm[j] * LD + col0 + m];
        }
        __syncthreads();
        float2 *out = reinterpret_cast<float2 *>(op.a_out);
        for (int e = threadIdx.x; e < 32 * 32; e += 256) {
            const int mi = e >> 5, jj = e & 31;
            const int j = j0 + jj, m = m0 + mi;
            if (j < k && m < M) out[(int64_t)m * k + j] = f2(tile[jj][mi], !left_absorb);
        }
    } else {
        const int j = (int)(u - ta);
        float2 *out = reinterpret_cast<float2 *>(op.b_out) + (int64_t)j * N;
        const int64_t src = (int64_t)perm[j] * LD + (left_absorb ? R : 0);
        for (int n = threadIdx.x; n < N; n += 256) out[n] = f2(Z[src + n], left_absorb);
    }
}

// ---------------------------------------------------------------- launchers
template <class T>
static void launch_init_rows_t(hipStream_t s, const mpsx_two_site_op *d, int count, char *ws,
                               int64_t rows) {
    hipLaunchKernelGGL(k_init_rows<T>, dim3((unsigned)rows), dim3(256), 0, s, d, count, ws);
}
hipError_t launch_init_rows(hipStream_t s, bool f32, const mpsx_two_site_op *d, int count, char *ws,
                            int64_t rows) {
    if (f32)
        launch_init_rows_t<float>(s, d, count, ws, rows);
    else
        launch_init_rows_t<double>(s, d, count, ws, rows);
    return hipGetLastError();
}

template <class T>
static hipError_t launch_theta_t(hipStream_t s, int D, const mpsx_two_site_op *d, int count,
                                 char *ws, int64_t tiles) {
    const dim3 g((unsigned)tiles), blk(256);
    switch (D) {
        case 1: hipLaunchKernelGGL(k_load<T>, g, blk, 0, s, d, count, ws); break;
        case 2: hipLaunchKernelGGL((k_theta<2, T>), g, blk, 0, s, d, count, ws); break;
        case 3: hipLaunchKernelGGL((k_theta<3, T>), g, blk, 0, s, d, count, ws); break;
        case 4: hipLaunchKernelGGL((k_theta<4, T>), g, blk, 0, s, d, count, ws); break;
        default: return hipErrorInvalidValue;
    }
    return hipSuccess;
}
hipError_t launch_theta(hipStream_t s, bool f32, int D, const mpsx_two_site_op *d, int count,
                        char *ws, int64_t tiles) {
    const hipError_t e = f32 ? launch_theta_t<float>(s, D, d, count, ws, tiles)
                             : launch_theta_t<double>(s, D, d, count, ws, tiles);
    return e != hipSuccess ? e : hipGetLastError();
}

hipError_t launch_theta_gen(hipStream_t s, bool f32, const mpsx_two_site_op *d, int count, char *ws,
                            int64_t tiles) {
    if (f32)
        hipLaunchKernelGGL(k_theta_gen<float>, dim3((unsigned)tiles), dim3(256), 0, s, d, count, ws);
    else
        hipLaunchKernelGGL(k_theta_gen<double>, dim3((unsigned)tiles), dim3(256), 0, s, d, count, ws);
    return hipGetLastError();
}

hipError_t launch_rownorm(hipStream_t s, bool f32, const mpsx_two_site_op *d, int count, char *ws,
                          int64_t rows) {
    if (f32)
        hipLaunchKernelGGL(k_rownorm<float>, dim3((unsigned)rows), dim3(256), 0, s, d, count, ws);
    else
        hipLaunchKernelGGL(k_rownorm<double>, dim3((unsigned)rows), dim3(256), 0, s, d, count, ws);
    return hipGetLastError();
}

hipError_t launch_fro(hipStream_t s, const mpsx_two_site_op *d, int count, char *ws, double null_rel) {
    hipLaunchKernelGGL(k_fro, dim3((unsigned)count), dim3(256), 0, s, d, count, ws, null_rel);
    return hipGetLastError();
}

hipError_t launch_owner(hipStream_t s, const mpsx_two_site_op *d, int count, int *owner) {
    hipLaunchKernelGGL(k_owner, dim3((unsigned)count), dim3(256), 0, s, d, count, owner);
    return hipGetLastError();
}

hipError_t launch_jacobi_round(hipStream_t s, bool f32, int b, const mpsx_two_site_op *d,
                               const int *owner, char *ws, int64_t pairs, int round, double tol,
                               double tol_inner, int inner_sweeps) {
    const dim3 g((unsigned)pairs), blk(256);
    if (b != 8 && b != 16) return hipErrorInvalidValue;
    if (f32) {
        if (b == 8)
            hipLaunchKernelGGL(k_jacobi_round32<16>, g, blk, 0, s, d, owner, ws, round, tol, tol_inner,
                               inner_sweeps);
        else
            hipLaunchKernelGGL(k_jacobi_round32<32>, g, blk, 0, s, d, owner, ws, round, tol, tol_inner,
                               inner_sweeps);
    } else {
        if (b == 8)
            hipLaunchKernelGGL(k_jacobi_round64<16>, g, blk, 0, s, d, owner, ws, round, tol, tol_inner,
                               inner_sweeps);
        else
            hipLaunchKernelGGL(k_jacobi_round64<32>, g, blk, 0, s, d, owner, ws, round, tol, tol_inner,
                               inner_sweeps);
    }
    return hipGetLastError();
}

hipError_t launch_sweep_end(hipStream_t s, const mpsx_two_site_op *d, int count, char *ws, int sweep,
                            bool fp32_phase) {
    hipLaunchKernelGGL(k_sweep_end, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, d, count, ws,
                       sweep, fp32_phase ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_sort(hipStream_t s, const mpsx_two_site_op *d, int count, char *ws,
                       int *unconverged) {
    hipLaunchKernelGGL(k_sort, dim3((unsigned)count), dim3(256), 0, s, d, count, ws, unconverged);
    return hipGetLastError();
}

hipError_t launch_gather(hipStream_t s, bool f32, const mpsx_two_site_op *d, int count, char *ws,
                         int64_t outs) {
    if (f32)
        hipLaunchKernelGGL(k_gather<float>, dim3((unsigned)outs), dim3(256), 0, s, d, count, ws);
    else
        hipLaunchKernelGGL(k_gather<double>, dim3((unsigned)outs), dim3(256), 0, s, d, count, ws);
    return hipGetLastError();
}

hipError_t launch_hgram(hipStream_t s, const mpsx_two_site_op *d, const int *owner, char *ws,
                        int64_t pairs, int max_nb, int stage) {
    const dim3 small((unsigned)pairs, (unsigned)(max_nb + 1));
    const bool large = 16 * max_nb > kTiledGemmRows;
    const dim3 tiled((unsigned)pairs, (unsigned)((16 * max_nb + 63) / 64));  // 64-column tiles of Cp
    if (stage == 0) {
        hipLaunchKernelGGL(k_hgram<0>, small, dim3(256), 0, s, d, owner, ws);
        if (large) hipLaunchKernelGGL(k_hgemm<0>, tiled, dim3(256), 0, s, d, owner, ws);
    } else {
        hipLaunchKernelGGL(k_hgram<1>, small, dim3(256), 0, s, d, owner, ws);
        if (large) hipLaunchKernelGGL(k_hgemm<1>, tiled, dim3(256), 0, s, d, owner, ws);
    }
    return hipGetLastError();
}

hipError_t launch_hsolve(hipStream_t s, bool f32h, const mpsx_two_site_op *d, const int *owner, char *ws,
                         int64_t pairs, int round, double tol, double tol_inner, int inner_sweeps,
                         double abs_rel) {
    const dim3 g((unsigned)pairs), blk(HSOLVE_THREADS);
    if (f32h)
        hipLaunchKernelGGL(k_hsolve<true>, g, blk, 0, s, d, owner, ws, round, tol, tol_inner, inner_sweeps,
                           abs_rel);
    else
        hipLaunchKernelGGL(k_hsolve<false>, g, blk, 0, s, d, owner, ws, round, tol, tol_inner, inner_sweeps,
                           abs_rel);
    return hipGetLastError();
}

hipError_t launch_hround(hipStream_t s, bool f32h, const mpsx_two_site_op *d, const int *owner, char *ws,
                         int64_t pairs, int ytiles, int round, double tol, double tol_inner,
                         int inner_sweeps, double abs_rel) {
    static_assert(HSOLVE_THREADS == 256, "k_hround runs both roles with one workgroup size");
    const dim3 g((unsigned)pairs, (unsigned)(ytiles + 1)), blk(256);
    if (f32h)
        hipLaunchKernelGGL(k_hround<true>, g, blk, 0, s, d, owner, ws, round, tol, tol_inner, inner_sweeps,
                           abs_rel);
    else
        hipLaunchKernelGGL(k_hround<false>, g, blk, 0, s, d, owner, ws, round, tol, tol_inner, inner_sweeps,
                           abs_rel);
    return hipGetLastError();
}

hipError_t launch_hupdate(hipStream_t s, bool f32h, const mpsx_two_site_op *d, const int *owner, char *ws,
                          int64_t pairs, int ytiles, int round, int part) {
    const dim3 g((unsigned)pairs, (unsigned)ytiles), blk(256);
    if (f32h)
        hipLaunchKernelGGL(k_hupdate<true>, g, blk, 0, s, d, owner, ws, round, part);
    else
        hipLaunchKernelGGL(k_hupdate<false>, g, blk, 0, s, d, owner, ws, round, part);
    return hipGetLastError();
}

hipError_t launch_hxcopy(hipStream_t s, const mpsx_two_site_op *d, int count, char *ws, int64_t rows) {
    hipLaunchKernelGGL(k_hxcopy, dim3((unsigned)rows), dim3(256), 0, s, d, count, ws);
    return hipGetLastError();
}

hipError_t launch_hxgemm(hipStream_t s, const mpsx_two_site_op *d, const int *owner, char *ws,
                         int64_t pairs, int ctiles, int mode, int max_nb) {
    // ctiles counts 32-column tiles of R; k_hgemm takes 64 columns per workgroup
    const dim3 small((unsigned)pairs, (unsigned)ctiles), tiled((unsigned)pairs, (unsigned)((ctiles + 1) / 2));
    const bool large = 16 * max_nb > kTiledGemmRows;
    if (mode == 0) {
        hipLaunchKernelGGL(k_hxgemm<0>, small, dim3(256), 0, s, d, owner, ws);
        if (large) hipLaunchKernelGGL(k_hgemm<2>, tiled, dim3(256), 0, s, d, owner, ws);
    } else {
        hipLaunchKernelGGL(k_hxgemm<1>, small, dim3(256), 0, s, d, owner, ws);
        if (large) hipLaunchKernelGGL(k_hgemm<3>, tiled, dim3(256), 0, s, d, owner, ws);
    }
    return hipGetLastError();
}
