// QR preconditioning of the Gram-method Jacobi (plan.method 3), gfx950.
//
// The Gram method orthogonalises the C rows of X (C x R, FP64 complex) by
// two-sided Jacobi on H = X X^H, accumulating the rotations in Y (starting
// as the identity) and finishing with X <- Y X. Preconditioned, it first
// takes a Householder QR of X with its columns in a fixed order P (by
// decreasing norm), X P = Q R, applied to the whole working row [X P | Y]:
//   [X P | I]  ->  [R | Q^H]
// so that the Jacobi starts from H = R R^H = Q^H (X X^H) Q and Y = Q^H, and
// the rotations W it finds give Y = W Q^H. The end is unchanged except that
// X <- Y X0 reads the saved, unpermuted X0: W Q^H X0 = W R P^T has
// orthogonal rows, the absorbed factor, and U = Y^H = Q W^H is exactly
// unitary (products of unitary reflectors and rotations), so nothing is
// divided by a singular value. P only decides which Q the QR finds; it is
// never undone. R R^H is nearer diagonal than X X^H when the spectrum
// decays (the Drmac-Veselic preconditioner of one-sided Jacobi), which
// saves sweeps.
//
// Householder reflector of x (column below and at the pivot row q):
//   H = I - tau v v^H,  v_q = 1,  H x = beta e_q,
//   beta = -e^{i arg x_q} ||x||,  tau = 1 + |x_q| / ||x|| (real, H Hermitian
//   and unitary),  v = x / (x_q - beta) below q;  tau = 0 when x is zero
//   below q.
// Sixteen reflectors form a panel, Q_p = H_0 ... H_15 = I - V T V^H with T
// upper triangular (T[j][j] = tau_j, T[0:j, j] = -tau_j T[0:j, 0:j] V^H v_j),
// and the columns right of the panel take Q_p^H A = A - V (T^H (V^H A)).
#include <hip/hip_runtime.h>

#include <cstddef>

#include "common.h"
#include "launch.h"

#define OP_FIELD(f) offsetof(mpsx_two_site_op, f)

namespace {

MPSX_DEV double2 *zq_ptr(char *ws, const mpsx_plan &p) {
    return reinterpret_cast<double2 *>(ws + z_off(p));
}
MPSX_DEV double2 *qv_ptr(char *ws, const mpsx_plan &p, int panel) {
    return reinterpret_cast<double2 *>(ws + qv_off(p)) + (int64_t)(panel % QR_RING) * 16 * p.Cp;
}
MPSX_DEV double2 *qt_ptr(char *ws, const mpsx_plan &p, int panel) {
    return reinterpret_cast<double2 *>(ws + qt_off(p)) + (panel % QR_RING) * 256;
}
MPSX_DEV int *qorder_ptr(char *ws, const mpsx_plan &p) {
    return reinterpret_cast<int *>(ws + qorder_off(p));
}
// Component-wise select by value (a conditional on two double2 lvalues
// selects an address and keeps the register arrays in memory).
MPSX_DEV double2 zsel(bool c, double2 x, double2 y) {
    return make_double2(c ? x.x : y.x, c ? x.y : y.y);
}
// With 8 chunks a wave (1024 rows) the strip takes 128 registers: keep the
// scheduler from hoisting every chunk's loads ahead of the first, which
// would take as many registers again (the form still spills a few hundred
// bytes a lane in its once-per-launch code).
template <int CH>
MPSX_DEV void chunk_fence() {
    if constexpr (CH > 4) __builtin_amdgcn_sched_barrier(0);
}
// Sum over the 64 lanes of a wave, the same in every lane (fixed order):
// DPP swaps within quads, then the half-row and row mirrors, then two
// cross-row exchanges.
template <int CTRL>
MPSX_DEV double dpp_d(double x) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
MPSX_DEV double wave_sum(double x) {
    x += dpp_d<0xB1>(x);   // quad_perm [1,0,3,2]
    x += dpp_d<0x4E>(x);   // quad_perm [2,3,0,1]
    x += dpp_d<0x141>(x);  // row_half_mirror
    x += dpp_d<0x140>(x);  // row_mirror
    x += __shfl_xor(x, 16, 64);
    x += __shfl_xor(x, 32, 64);
    return x;
}
// Element (rr, c) of the LDS copy of a 512 x 16 strip: rows of 16, the
// column XOR-swizzled by the row so that a column read by consecutive rows
// spreads over the banks.
MPSX_DEV int sidx(int rr, int c) { return rr * 16 + (c ^ (rr & 15)); }
MPSX_DEV double2 shfl_xor_z(double2 v, int m) {
    return make_double2(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64));
}


// T of the panel just factored, row k by thread k < 16, in registers with
// the loops unrolled (the fastest of the forms measured: a loop over LDS
// takes one round trip per term): T[k][k] = tau_k, T[k][j] = -tau_j
// sum_{m<j} T[k][m] Ys[j][m] for k < j < ncols, zero elsewhere.
MPSX_DEV void build_t(char *ws, const mpsx_plan &p, int panel, int ncols, const double2 (*Ys)[17],
                      const double *taus) {
    if (threadIdx.x >= 16) return;
    const int k = threadIdx.x;
    double2 tr[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        double2 sum = make_double2(0.0, 0.0);
#pragma unroll
        for (int m = 0; m < j; ++m) sum = zadd(sum, zmul(tr[m], Ys[j][m]));
        const double tj = j < ncols ? taus[j] : 0.0;
        tr[j] = zsel(j == k, make_double2(tj, 0.0),
                     zsel(j > k && j < ncols, zscale(sum, -tj), make_double2(0.0, 0.0)));
    }
    double2 *T = qt_ptr(ws, p, panel + 1) + k * 16;
#pragma unroll
    for (int j = 0; j < 16; ++j) T[j] = tr[j];
}

// Factorisation of panel + 1 by the workgroup of strip 0 (the panel's own
// columns, already updated by panel `panel` in `a`), for problems of at most
// 512 rows. The strip goes through LDS into a column layout: wave w owns
// columns 2w and 2w + 1 whole, lane l their rows rf + l + 64 t, so every
// reduction over rows stays inside a wave (DPP, no barrier). Step j: the
// owner of column j forms the reflector (norm below q, x_q, beta, tau, sc)
// and publishes w_j (x_j below q, 1 / sc at q) in one of two LDS buffers
// and its scalars; one barrier; every wave takes g'_k = w_j^H a_k for its
// columns and updates those right of j (a_k -= tau |sc|^2 g'_k w_j), or
// records v_k^H v_j for T (k < j). The buffers alternate, so the owner of
// step j + 1 may publish while slower waves still read step j's.
MPSX_DEV void factor_512(const mpsx_two_site_op *ops, char *ws, int panel, const double2 (&a)[4][4],
                         double2 *big, double2 (*wbuf)[512], double2 (*Ys)[17], double *taus,
                         double2 *scs, double2 *betas) {
    const mpsx_plan p = ops[blockIdx.x].plan;
    const int C = p.C, LD = p.LD, ncol = p.R + p.C;
    const int rb = panel >= 0 ? 16 * panel : 0, rf = 16 * (panel + 1);
    const int ncols = min(16, qr_reflectors(p) - rf);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, l16 = lane & 15, lq = lane >> 4;
    double2 *Z = zq_ptr(ws, p);
    __syncthreads();  // the update phase is done with the shared buffer
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) big[sidx(16 * (w + 8 * i) + 4 * r + lq, l16)] = a[i][r];
    __syncthreads();
    double2 x[2][8];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int rr = rf - rb + lane + 64 * t;
            x[c][t] = rr < 512 ? big[sidx(rr, 2 * w + c)] : make_double2(0.0, 0.0);
        }
    for (int j = 0; j < ncols; ++j) {
        const int q = rf + j, buf = j & 1;
        if ((j >> 1) == w) {
            const bool hi = (j & 1) != 0;
            int d0 = rf + lane - q;  // row - q of element t is d0 + 64 t
            asm volatile("" : "+v"(d0));
            double s = 0.0;
            double2 xq = make_double2(0.0, 0.0);
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int dr = d0 + 64 * t;
                const double2 xv = zsel(hi, x[1][t], x[0][t]);
                const double2 xb = zsel(dr > 0, xv, make_double2(0.0, 0.0));
                s = fma(xb.x, xb.x, fma(xb.y, xb.y, s));
                xq = zsel(dr == 0, xv, xq);
            }
            const double s2 = wave_sum(s);
            const double2 al = make_double2(wave_sum(xq.x), wave_sum(xq.y));
            double tau = 0.0;
            double2 beta = al, sc = make_double2(1.0, 0.0), isc = sc;
            if (s2 > 0.0) {  // two long operations deep: the square roots, then the reciprocals
                const double a2 = al.x * al.x + al.y * al.y;
                const double aa = sqrt(a2), xn = sqrt(a2 + s2);
                const double ia = 1.0 / aa, ix = 1.0 / xn, inv = 1.0 / (aa + xn);
                const double2 u = aa > 0.0 ? make_double2(al.x * ia, al.y * ia) : make_double2(1.0, 0.0);
                beta = make_double2(-u.x * xn, -u.y * xn);
                tau = 1.0 + aa * ix;
                sc = make_double2(u.x * inv, -u.y * inv);              // conj(u) / (|x_q| + ||x||)
                isc = make_double2(u.x * (aa + xn), u.y * (aa + xn));  // 1 / sc
            }
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int dr = d0 + 64 * t;
                const double2 xv = zsel(hi, x[1][t], x[0][t]);
                wbuf[buf][lane + 64 * t] = zsel(dr > 0, xv, zsel(dr == 0, isc, make_double2(0.0, 0.0)));
            }
            if (lane == 0) {
                taus[j] = tau;
                scs[j] = sc;
                betas[j] = beta;
            }
        }
        __syncthreads();
        const double tau = taus[j];
        const double2 sc = scs[j];
        double2 wv[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) wv[t] = wbuf[buf][lane + 64 * t];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int k = 2 * w + c;
            double2 g = make_double2(0.0, 0.0);
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                g.x = fma(wv[t].x, x[c][t].x, fma(wv[t].y, x[c][t].y, g.x));
                g.y = fma(wv[t].x, x[c][t].y, fma(-wv[t].y, x[c][t].x, g.y));
            }
            g = make_double2(wave_sum(g.x), wave_sum(g.y));
            if (k > j) {
                const double2 cg = zscale(g, tau * (sc.x * sc.x + sc.y * sc.y));
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    x[c][t].x = fma(-wv[t].x, cg.x, fma(wv[t].y, cg.y, x[c][t].x));
                    x[c][t].y = fma(-wv[t].x, cg.y, fma(-wv[t].y, cg.x, x[c][t].y));
                }
            } else if (k < j && lane == 0) {
                Ys[j][k] = zmul(zconj(scs[k]), zmul(sc, zconj(g)));
            }
        }
    }
    __syncthreads();
    // reflected columns: beta on the diagonal and zero below in R, v = sc x
    // below (1 at the diagonal) in V; the other columns as they are
    double2 *V = qv_ptr(ws, p, panel + 1);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int k = 2 * w + c, qk = rf + k;
        const bool refl = k < ncols;
        const double2 sck = scs[k & 15], bk = betas[k & 15];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int row = rf + lane + 64 * t, rr = row - rb;
            if (rr >= 512) continue;
            const double2 xv = x[c][t];
            big[sidx(rr, k)] = zsel(refl && row > qk, make_double2(0.0, 0.0), zsel(refl && row == qk, bk, xv));
            if (row < C)
                V[(int64_t)row * 16 + k] =
                    refl ? zsel(row > qk, zmul(xv, sck), make_double2(row == qk ? 1.0 : 0.0, 0.0))
                         : make_double2(0.0, 0.0);
        }
    }
    __syncthreads();
    // the strip to Z (rows rb .. C-1, 16 consecutive columns a row)
    for (int e = tid; e < (C - rb) * 16; e += 512) {
        const int rr = e >> 4, c = e & 15;
        if (rf + c < ncol) Z[(int64_t)(rb + rr) * LD + rf + c] = big[sidx(rr, c)];
    }
    build_t(ws, p, panel, ncols, Ys, taus);
}

}  // namespace

namespace {
constexpr int QR_NORM_ROWS = 32;
constexpr int QR_NORM_COLS = 256;
// Cp is a multiple of 32, so a problem has at most Cp / 32 parts
static_assert((QR_MAX_CP / QR_NORM_ROWS) * QR_MAX_R * 8 <= QR_RING * 16 * 16 * QR_MAX_CP,
              "the column norm parts fit the V ring");
MPSX_DEV double *qnorm_ptr(char *ws, const mpsx_plan &p) {
    return reinterpret_cast<double *>(ws + qv_off(p));
}
}  // namespace

// Squared column norms of X in parts: workgroup (problem, y, z) sums the
// rows 32 y .. 32 y + 31 of the columns 256 z .. 256 z + 255 in row order
// into part[y][column] (FP64), so that many workgroups share the problem's
// read of X. part lies in the V ring, unused until the first panel is
// factored.
__global__ void __launch_bounds__(QR_NORM_COLS) k_qr_colnorm(const mpsx_two_site_op *ops, char *ws) {
    const mpsx_plan p = ops[blockIdx.x].plan;
    if (p.method != 3) return;
    const int C = p.C, R = p.R;
    const int r0 = (int)blockIdx.y * QR_NORM_ROWS, c = (int)blockIdx.z * QR_NORM_COLS + (int)threadIdx.x;
    if (r0 >= C || c >= R) return;
    const int r1 = r0 + QR_NORM_ROWS < C ? r0 + QR_NORM_ROWS : C;
    const double2 *z = zq_ptr(ws, p) + c;
    double s = 0.0;
#pragma unroll 8
    for (int r = r0; r < r1; ++r) {
        const double2 v = z[(int64_t)r * p.LD];
        s = fma(v.x, v.x, fma(v.y, v.y, s));
    }
    qnorm_ptr(ws, p)[(int64_t)blockIdx.y * R + c] = s;
}

// Column order of X, one workgroup per problem: the squared column norms
// (the parts summed in row order), then the rank of each column under (norm
// descending, index ascending); order[rank] = column. A NaN norm ranks
// last, so the order is a permutation whatever the data.
__global__ void __launch_bounds__(1024) k_qr_order(const mpsx_two_site_op *ops, char *ws) {
    __shared__ double nrm[QR_MAX_R];
    const mpsx_plan p = ops[blockIdx.x].plan;
    if (p.method != 3) return;
    const int R = p.R, parts = (p.C + QR_NORM_ROWS - 1) / QR_NORM_ROWS;
    const int tid = threadIdx.x;
    const double *part = qnorm_ptr(ws, p);
    for (int c = tid; c < R; c += 1024) {
        double v = 0.0;
        for (int y = 0; y < parts; ++y) v += part[(int64_t)y * R + c];
        nrm[c] = v == v ? v : -1.0;
    }
    __syncthreads();
    int *order = qorder_ptr(ws, p);
    for (int c = tid; c < R; c += 1024) {
        const double v = nrm[c];
        int rank = 0;
        // unrolled so that the broadcast LDS reads overlap
#pragma unroll 16
        for (int o = 0; o < R; ++o) {
            const double w = nrm[o];
            rank += (w > v) || (w == v && o < c);
        }
        order[rank] = c;
    }
}

// One workgroup per padded row: X0[row] <- X[row] (zero for padded rows),
// then X[row][j] <- X0[row][order[j]].
__global__ void __launch_bounds__(256) k_qr_permute(const mpsx_two_site_op *ops, int count, char *ws) {
    __shared__ double2 xs[QR_MAX_R];
    const int pi = find_record(reinterpret_cast<const char *>(ops), count, sizeof(mpsx_two_site_op),
                               OP_FIELD(plan.row_base), blockIdx.x);
    const mpsx_plan p = ops[pi].plan;
    if (p.method != 3) return;
    const int row = (int)(blockIdx.x - p.row_base);
    if (row >= p.Cp) return;
    const int R = p.R;
    double2 *x0 = reinterpret_cast<double2 *>(ws + qx0_off(p)) + (int64_t)row * R;
    if (row >= p.C) {
        for (int c = threadIdx.x; c < R; c += 256) x0[c] = make_double2(0.0, 0.0);
        return;
    }
    double2 *z = zq_ptr(ws, p) + (int64_t)row * p.LD;
    for (int c = threadIdx.x; c < R; c += 256) {
        const double2 v = z[c];
        xs[c] = v;
        x0[c] = v;
    }
    __syncthreads();
    const int *order = qorder_ptr(ws, p);
    for (int j = threadIdx.x; j < R; j += 256) {
        int src = order[j];
        if ((unsigned)src >= (unsigned)R) src = j;
        z[j] = xs[src];
    }
}

// One step of the blocked QR. x: problem; y: 16-column strip. Launch
// `panel` >= 0 applies panel `panel`'s block reflector (V, T from ring slot
// panel mod QR_RING) to the rows panel*16 .. C-1 of the strip's columns
// 16 (panel + 1 + y) .. +15 (the X columns right of the panel, then the Y
// columns; every column of [X | Y] right of the panel up to R + C - 1; the
// strips of Y columns alone take QR_DEFER panels at a time, below);
// strip 0 (the next panel's columns) then factors panel + 1 (factor_512 up
// to 512 rows, else in the register layout) and writes its V and T to ring
// slot panel + 1. Launch -1 only factors panel 0. Eight waves; wave w holds
// the 16-row chunks w, w + 8, ... of the strip, chunk i as the four
// registers a[i][r] = (row 16 (w + 8 i) + 4 r + lq, column l16) of lane
// (l16, lq), the matrix-core output layout.
template <int CH>
__global__ void __launch_bounds__(512) k_qr_step(const mpsx_two_site_op *ops, char *ws, int panel) {
    // The update phase's arrays and, for panels of at most 512 rows, the
    // factorisation's LDS copy of the strip share one buffer.
    constexpr int kUpd = 11 * 16 * 17;
    constexpr int kStrip = CH == 4 ? 512 * 16 : 0;
    __shared__ double2 big[kStrip > kUpd ? kStrip : kUpd];
    auto Wp = reinterpret_cast<double2 (*)[16][17]>(big);             // per-wave partial V^H A
    auto Wt = reinterpret_cast<double2 (*)[17]>(big + 8 * 16 * 17);   // V^H A
    auto Ws = reinterpret_cast<double2 (*)[17]>(big + 9 * 16 * 17);   // T^H V^H A
    auto Ts = reinterpret_cast<double2 (*)[17]>(big + 10 * 16 * 17);  // T of the panel applied
    __shared__ double2 Ys[16][17];     // Ys[j][k] = v_k^H v_j (k < j), for T
    // 512-row panels: w_j published by the wave that owns column j, two buffers
    __shared__ double2 wbuf[CH == 4 ? 2 : 1][CH == 4 ? 512 : 1];
    __shared__ double taus[16];
    __shared__ double2 scs[16], betas[16];
    __shared__ double red[8];
    __shared__ double2 redz[8][16];
    __shared__ double2 alpha_s;
    __shared__ double2 vcol[8][CH == 4 ? 1 : 16 * CH];  // 1024-row panels: v, per wave for its rows

    const mpsx_plan p = ops[blockIdx.x].plan;
    if (p.method != 3) return;
    // each problem always runs the same variant (4 chunks up to 512 rows, 8
    // above), so its result does not depend on what else is in the batch
    if ((CH == 4) != (p.C <= 512)) return;
    const int C = p.C, LD = p.LD, ncol = p.R + p.C;
    const int np = qr_panels(p);
    if (panel >= np) return;
    const int c0 = 16 * (panel + 1 + (int)blockIdx.y);
    if (c0 >= ncol) return;
    const bool update = panel >= 0;
    const bool factor = blockIdx.y == 0 && panel + 1 < np;
    if (!update && !factor) return;
    // A strip of Y columns only (c0 >= R) is read by no later panel: it takes
    // the reflectors of QR_DEFER panels at a time, at every QR_DEFER-th
    // launch and the last, from panel k0 on (one pass over it in HBM instead
    // of one per panel). Strips with X columns, the factored one included,
    // take every panel at once.
    const bool ystrip = c0 >= p.R;
    const int k0 = ystrip ? panel - panel % QR_DEFER : panel;
    if (update && ystrip && panel % QR_DEFER != QR_DEFER - 1 && panel != np - 1) return;
    const int rb = update ? 16 * k0 : 0;  // first row held
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, l16 = lane & 15, lq = lane >> 4;
    const int col = c0 + l16;
    const bool col_ok = col < ncol;
    double2 *Z = zq_ptr(ws, p);

    double2 a[CH][4];
#pragma unroll
    for (int i = 0; i < CH; ++i) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = rb + 16 * (w + 8 * i) + 4 * r + lq;
            a[i][r] = (row < C && col_ok) ? Z[(int64_t)row * LD + col] : make_double2(0.0, 0.0);
        }
        chunk_fence<CH>();
    }

    for (int k = k0; update && k <= panel; ++k) {
        const double2 *V = qv_ptr(ws, p, k);
        const double2 *T = qt_ptr(ws, p, k);
        const int vr0 = 16 * k;  // rows of panel k's reflectors (the ring slot's rows above are stale)
        // W = V^H A over this wave's rows: A operand conj(V[row][refl]),
        // B operand the strip; three real products per complex product
        mfma_d4 m1 = {0.0, 0.0, 0.0, 0.0}, m2 = m1, m3 = m1;
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            const int base = rb + 16 * (w + 8 * i);
            if (base >= C) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = base + 4 * r + lq;
                const double2 v = (row >= vr0 && row < C) ? V[(int64_t)row * 16 + l16] : make_double2(0.0, 0.0);
                m1 = mfma16d(v.x, a[i][r].x, m1);
                m2 = mfma16d(v.y, a[i][r].y, m2);
                m3 = mfma16d(v.x - v.y, a[i][r].x + a[i][r].y, m3);
            }
            chunk_fence<CH>();
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
            Wp[w][4 * r + lq][l16] = make_double2(m1[r] + m2[r], m3[r] - m1[r] + m2[r]);
        if (tid < 256) Ts[tid >> 4][tid & 15] = T[tid];
        __syncthreads();
        if (tid < 256) {
            const int i = tid >> 4, c = tid & 15;
            double2 s = make_double2(0.0, 0.0);
#pragma unroll
            for (int v = 0; v < 8; ++v) s = zadd(s, Wp[v][i][c]);
            Wt[i][c] = s;
        }
        __syncthreads();
        if (tid < 256) {  // Ws = T^H Wt
            const int i = tid >> 4, c = tid & 15;
            double2 s = make_double2(0.0, 0.0);
            for (int k = 0; k <= i; ++k) {
                const double2 t = Ts[k][i], x = Wt[k][c];
                s.x += t.x * x.x + t.y * x.y;
                s.y += t.x * x.y - t.y * x.x;
            }
            Ws[i][c] = s;
        }
        __syncthreads();
        // A -= V Ws: A operand V[row][refl], B operand Ws[refl][column]
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            const int base = rb + 16 * (w + 8 * i);
            if (base >= C) continue;
            mfma_d4 n1 = {0.0, 0.0, 0.0, 0.0}, n2 = n1, n3 = n1;
            const int vrow = base + l16;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const double2 v = (vrow >= vr0 && vrow < C) ? V[(int64_t)vrow * 16 + 4 * t + lq]
                                                            : make_double2(0.0, 0.0);
                const double2 x = Ws[4 * t + lq][l16];
                n1 = mfma16d(v.x, x.x, n1);
                n2 = mfma16d(v.y, x.y, n2);
                n3 = mfma16d(v.x + v.y, x.x + x.y, n3);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                a[i][r].x -= n1[r] - n2[r];
                a[i][r].y -= n3[r] - n1[r] - n2[r];
            }
            chunk_fence<CH>();
        }
    }

    if constexpr (CH == 4) {
        if (factor) {
            factor_512(ops, ws, panel, a, big, wbuf, Ys, taus, scs, betas);
            return;
        }
    }

    int ncols = 0, rf = 0;
    if (factor) {
        rf = 16 * (panel + 1);
        ncols = min(16, qr_reflectors(p) - rf);
        // Step j reflects column j (pivot row q = rf + j). The columns keep
        // their unscaled x (below q) in registers until the end; sc_j, beta_j
        // and tau_j go to LDS. Every lane reads w_j (x_j below q, 1 / sc_j at
        // q, zero above) from the wave's LDS column, so v_j = sc_j w_j and
        //   g'_k = w_j^H a_k,  a_k -= tau_j |sc_j|^2 g'_k w_j  (k > j),
        //   v_m^H v_j = conj(sc_m) sc_j conj(g'_m)  (m < j: a_m holds x_m).
        for (int j = 0; j < ncols; ++j) {
            const int q = rf + j;
            // Lane j of each row: x_j below q (zero elsewhere) into the
            // wave's LDS column, its squared norm, and x_q. Row - q of element
            // (i, r) is d0 + 128 i + 4 r; d0 passes through an empty asm
            // statement, so that the compares are against constants and no
            // predicate stays live from one pass to the next.
            double s = 0.0;
            if (l16 == j) {
                int d0 = rb + 16 * w + lq - q;
                asm volatile("" : "+v"(d0));
#pragma unroll
                for (int i = 0; i < CH; ++i)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int dr = d0 + 128 * i + 4 * r;
                        const double2 x = zsel(dr > 0, a[i][r], make_double2(0.0, 0.0));
                        vcol[w][16 * i + 4 * r + lq] = x;
                        s = fma(x.x, x.x, fma(x.y, x.y, s));
                        if (dr == 0) alpha_s = a[i][r];
                    }
            }
            s += __shfl_xor(s, 16, 64);
            s += __shfl_xor(s, 32, 64);
            if (lq == 0 && l16 == j) red[w] = s;
            __syncthreads();
            const double s2 = ((red[0] + red[1]) + (red[2] + red[3])) + ((red[4] + red[5]) + (red[6] + red[7]));
            const double2 al = alpha_s;
            double tau = 0.0;
            double2 beta = al, sc = make_double2(1.0, 0.0), isc = sc;
            if (s2 > 0.0) {  // two long operations deep: the square roots, then the reciprocals
                const double a2 = al.x * al.x + al.y * al.y;
                const double aa = sqrt(a2), xn = sqrt(a2 + s2);
                const double ia = 1.0 / aa, ix = 1.0 / xn, inv = 1.0 / (aa + xn);
                const double2 u = aa > 0.0 ? make_double2(al.x * ia, al.y * ia) : make_double2(1.0, 0.0);
                beta = make_double2(-u.x * xn, -u.y * xn);
                tau = 1.0 + aa * ix;
                sc = make_double2(u.x * inv, -u.y * inv);          // conj(u) / (|x_q| + ||x||)
                isc = make_double2(u.x * (aa + xn), u.y * (aa + xn));  // 1 / sc
            }
            if (tid == 0) {
                taus[j] = tau;
                scs[j] = sc;
                betas[j] = beta;
            }
            // w_j = x_j with 1 / sc at the pivot row: that entry of the LDS
            // column, written by the lane that wrote its zero (in order)
            if (l16 == j) {
                const int dq = q - rb, qw = (dq >> 4) & 7;
                if (qw == w && ((dq & 3) == lq)) vcol[w][16 * (dq >> 7) + 4 * ((dq >> 2) & 3) + lq] = isc;
            }
            __builtin_amdgcn_wave_barrier();
            // w in registers for 4 chunks a wave; for 8 (1024 rows) the strip
            // alone takes 128 registers, and w is read from LDS at each use
            double2 ww[CH <= 4 ? CH : 1][4];
            if constexpr (CH <= 4) {
#pragma unroll
                for (int i = 0; i < CH; ++i)
#pragma unroll
                    for (int r = 0; r < 4; ++r) ww[i][r] = vcol[w][16 * i + 4 * r + lq];
            }
            auto wget = [&](int i, int r) -> double2 {
                if constexpr (CH <= 4)
                    return ww[i][r];
                else
                    return vcol[w][16 * i + 4 * r + lq];
            };
            double2 g = make_double2(0.0, 0.0);
#pragma unroll
            for (int i = 0; i < CH; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const double2 x = wget(i, r);
                    g.x = fma(x.x, a[i][r].x, fma(x.y, a[i][r].y, g.x));
                    g.y = fma(x.x, a[i][r].y, fma(-x.y, a[i][r].x, g.y));
                }
            g = zadd(g, shfl_xor_z(g, 16));
            g = zadd(g, shfl_xor_z(g, 32));
            if (lq == 0) redz[w][l16] = g;
            __syncthreads();
            g = make_double2(0.0, 0.0);
#pragma unroll
            for (int v = 0; v < 8; ++v) g = zadd(g, redz[v][l16]);
            if (tid < 16 && tid < j)
                Ys[j][tid] = zmul(zconj(scs[tid]), zmul(sc, zconj(g)));
            const double cs = l16 > j ? tau * (sc.x * sc.x + sc.y * sc.y) : 0.0;
            const double2 cg = zscale(g, cs);
#pragma unroll
            for (int i = 0; i < CH; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const double2 x = wget(i, r);
                    a[i][r].x = fma(-x.x, cg.x, fma(x.y, cg.y, a[i][r].x));
                    a[i][r].y = fma(-x.x, cg.y, fma(-x.y, cg.x, a[i][r].y));
                }
        }
        __syncthreads();
        // V rows rf .. C-1 of the panel
        double2 *V = qv_ptr(ws, p, panel + 1);
#pragma unroll
        for (int i = 0; i < CH; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = rb + 16 * (w + 8 * i) + 4 * r + lq;
                if (row < rf || row >= C) continue;
                const int q = rf + l16;
                double2 v = make_double2(0.0, 0.0);
                if (l16 < ncols) v = zsel(row > q, zmul(a[i][r], scs[l16]), make_double2(row == q ? 1.0 : 0.0, 0.0));
                V[(int64_t)row * 16 + l16] = v;
            }
    }

    if (col_ok) {
        // the row stride passes through an empty asm statement, so that the
        // addresses are formed here rather than kept live from the loads
        int ld = LD;
        asm volatile("" : "+v"(ld));
        double2 *zc = Z + col;
#pragma unroll
        for (int i = 0; i < CH; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = rb + 16 * (w + 8 * i) + 4 * r + lq;
                if (row >= C) continue;
                // reflected columns: beta on the diagonal, zero below
                const bool refl = factor && l16 < ncols;
                const double2 out = zsel(refl && row > rf + l16, make_double2(0.0, 0.0),
                                         zsel(refl && row == rf + l16, betas[l16 & 15], a[i][r]));
                zc[(int64_t)row * ld] = out;
            }
    }
    // T after the strip is stored, so that its registers are free
    if (factor) {
        build_t(ws, p, panel, ncols, Ys, taus);
    }
}

hipError_t launch_qr_order(hipStream_t s, const mpsx_two_site_op *d, int count, char *ws, int max_c,
                           int max_r) {
    if (max_c > QR_MAX_CP || max_r > QR_MAX_R) return hipErrorInvalidValue;
    const dim3 parts((unsigned)count, (unsigned)((max_c + QR_NORM_ROWS - 1) / QR_NORM_ROWS),
                     (unsigned)((max_r + QR_NORM_COLS - 1) / QR_NORM_COLS));
    k_qr_colnorm<<<parts, QR_NORM_COLS, 0, s>>>(d, ws);
    if (hipError_t e = hipGetLastError()) return e;
    k_qr_order<<<count, 1024, 0, s>>>(d, ws);
    return hipGetLastError();
}

hipError_t launch_qr_permute(hipStream_t s, const mpsx_two_site_op *d, int count, char *ws,
                             int64_t rows) {
    k_qr_permute<<<(unsigned)rows, 256, 0, s>>>(d, count, ws);
    return hipGetLastError();
}

hipError_t launch_qr_step(hipStream_t s, const mpsx_two_site_op *d, int count, char *ws, int panel,
                          int strips, int max_c, bool small, bool large) {
    if (max_c > QR_MAX_CP) return hipErrorInvalidValue;
    const dim3 grid((unsigned)count, (unsigned)strips);
    if (small) {
        k_qr_step<4><<<grid, 512, 0, s>>>(d, ws, panel);
        if (hipError_t e = hipGetLastError()) return e;
    }
    if (large) k_qr_step<8><<<grid, 512, 0, s>>>(d, ws, panel);
    return hipGetLastError();
}
