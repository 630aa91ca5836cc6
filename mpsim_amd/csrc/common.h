// Shared device helpers and the library-private plan layout.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mpsx.h"

#define MPSX_DEV __device__ __forceinline__
#define MPSX_HD __host__ __device__ __forceinline__

// ---------------------------------------------------------------- complex
MPSX_DEV float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// c + a * b
MPSX_DEV float2 cfma(float2 a, float2 b, float2 c) {
    c.x = fmaf(a.x, b.x, c.x);
    c.x = fmaf(-a.y, b.y, c.x);
    c.y = fmaf(a.x, b.y, c.y);
    c.y = fmaf(a.y, b.x, c.y);
    return c;
}
// c + conj(a) * b
MPSX_DEV float2 cfma_conj(float2 a, float2 b, float2 c) {
    c.x = fmaf(a.x, b.x, c.x);
    c.x = fmaf(a.y, b.y, c.x);
    c.y = fmaf(a.x, b.y, c.y);
    c.y = fmaf(-a.y, b.x, c.y);
    return c;
}
MPSX_DEV float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

MPSX_DEV double2 zmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
MPSX_DEV double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
MPSX_DEV double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
MPSX_DEV double2 zscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }

// FP64 matrix-core product D = A B + C of 16x16x4 tiles. Lane l holds
// A[l % 16][l / 16] and B[l / 16][l % 16]; register r of lane l holds
// D[4 r + l / 16][l % 16].
typedef double mfma_d4 __attribute__((ext_vector_type(4)));
MPSX_DEV mfma_d4 mfma16d(double a, double b, mfma_d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Round-robin ("circle method") pairing of n (even) items: in round r
// (0 <= r < n-1) slot p (0 <= p < n/2) pairs item lo with item hi; every
// item appears once per round and every unordered pair meets exactly once
// over the n-1 rounds.
MPSX_HD int rr_item(int pos, int r, int n) {
    return pos == 0 ? 0 : ((pos - 1 + r) % (n - 1)) + 1;
}
MPSX_HD void rr_pair(int p, int r, int n, int *lo, int *hi) {
    int a = rr_item(p, r, n);
    int b = rr_item(n - 1 - p, r, n);
    *lo = a < b ? a : b;
    *hi = a < b ? b : a;
}

// Ordering of one inner sweep over a 2b-row pivot (rows 0..b-1 from block I,
// b..2b-1 from block J): steps 0..b-1 pair I row k with J row (k + step) mod b
// (the cross pairs); steps b..2b-2 pair the rows inside I and inside J by the
// circle method. Every pair of the pivot meets once per sweep; intra-block
// pairs, already orthogonal once the iteration settles, come last so their
// steps find nothing to rotate.
MPSX_HD void inner_pair(int k, int step, int b, int *p, int *q) {
    if (step < b) {
        *p = k;
        *q = b + (k + step) % b;
    } else if (k < b / 2) {
        rr_pair(k, step - b, b, p, q);
    } else {
        rr_pair(k - b / 2, step - b, b, p, q);
        *p += b;
        *q += b;
    }
}

// Binary search: index i with base[i] <= x < base[i+1] for a table of
// `count` records of byte size `stride` whose int64 field at `field_off`
// holds an exclusive prefix sum (records are sorted by construction).
MPSX_DEV int find_record(const char *table, int count, size_t stride, size_t field_off,
                         int64_t x) {
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        int64_t v = *reinterpret_cast<const int64_t *>(table + (size_t)mid * stride + field_off);
        if (v <= x)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Workspace header of a two-site batch (at ws + 0).
struct TwoSiteHeader {
    int32_t active[64];  // active[s]: problems still unconverged after sweep s
    int32_t pad[64];
};

// Per-problem control block inside the workspace.
struct ProblemCtl {
    double fro2;    // ||theta||_F^2
    double null2;   // rows with squared norm <= null2 are treated as zero
    int32_t counter;  // pivot pairs rotated in the current sweep
    int32_t conv;     // 1 once a full sweep rotated nothing
    int32_t sweeps;   // sweeps that did work
    int32_t pad;
};

// Byte layout of one problem's region starting at plan.ws_off:
//   ProblemCtl | sig2[Cp] (double) | perm[Cp] (int32) | Z[Cp][LD] (float2)
MPSX_HD int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }
MPSX_HD int64_t ctl_bytes() { return 64; }
MPSX_HD int64_t sig_off(const mpsx_plan &p) { return p.ws_off + ctl_bytes(); }
MPSX_HD int64_t perm_off(const mpsx_plan &p) { return sig_off(p) + align_up(8LL * p.Cp, 256); }
MPSX_HD int64_t z_off(const mpsx_plan &p) { return perm_off(p) + align_up(4LL * p.Cp, 256); }
// Problems of the Gram method (plan.method 1 or 2, b = 16, FP64) add after Z:
//   H[Cp][Cp] (double2, upper 16x16 block tiles valid; at the end the same
//   region holds a copy of X as Cp x R) | Q[2][nb/2][32][32] (double2, the
//   rotations of each pivot, for even and odd rounds) | flag[2][nb/2]
//   (int32, 1 when the pivot rotated in that round)
MPSX_HD int64_t h_elems(const mpsx_plan &p) {
    return (int64_t)p.Cp * (p.Cp > p.R ? p.Cp : p.R);
}
MPSX_HD int64_t h_off(const mpsx_plan &p) {
    return z_off(p) + align_up((int64_t)p.zbytes * p.Cp * p.LD, 256);
}
MPSX_HD int64_t qmat_off(const mpsx_plan &p) { return h_off(p) + align_up(16 * h_elems(p), 256); }
MPSX_HD int64_t qflag_off(const mpsx_plan &p) { return qmat_off(p) + 2 * 16LL * (p.nb / 2) * 1024; }
// Problems of the mixed Gram method (plan.method 2: an FP32 phase, then the
// FP64 rounds) add a region T of H's size after the flags. In the FP32 phase
// it holds H32 (Cp x Cp float2, row stride Cp, upper 16x16 block tiles
// valid); at the switch to FP64 it receives X1 = Y X (Cp x R double2, row
// stride R), from which H is formed again in FP64.
MPSX_HD int64_t t_off(const mpsx_plan &p) {
    return align_up(qflag_off(p) + 2 * 4LL * (p.nb / 2), 256);
}
// Problems of the QR-preconditioned Gram method (plan.method 3) add, at the
// same offset as region T:
//   X0[Cp][R] (double2, the unpermuted X, read by the final X <- Y X0;
//   padded rows zero) | V[QR_RING][Cp][16] (double2, the Householder vectors
//   of a 16-column panel, panel k in slot k mod QR_RING) | T[QR_RING][16][16]
//   (double2, the panels' triangular factors) | order[R] (int32, columns by
//   decreasing norm)
constexpr int QR_MAX_CP = 1024;  // rows one workgroup of the QR step holds (8 waves x 8 chunks of 16)
constexpr int QR_MAX_R = 4096;   // a row staged in LDS
// The columns of Y (Q^H being formed) take the panels' reflectors
// QR_DEFER at a time (no later panel reads them); the ring holds those
// panels and the one being factored meanwhile.
constexpr int QR_DEFER = 4;
constexpr int QR_RING = 8;
MPSX_HD int64_t qx0_off(const mpsx_plan &p) { return t_off(p); }
MPSX_HD int64_t qv_off(const mpsx_plan &p) { return qx0_off(p) + align_up(16LL * p.Cp * p.R, 256); }
MPSX_HD int64_t qt_off(const mpsx_plan &p) { return qv_off(p) + QR_RING * 16LL * 16 * p.Cp; }
MPSX_HD int64_t qorder_off(const mpsx_plan &p) { return qt_off(p) + QR_RING * 16LL * 256; }
// Householder reflectors of the QR: one per column of X up to the last row.
MPSX_HD int qr_reflectors(const mpsx_plan &p) { return p.C - 1 < p.R ? p.C - 1 : p.R; }
MPSX_HD int qr_panels(const mpsx_plan &p) { return (qr_reflectors(p) + 15) / 16; }
MPSX_HD int64_t problem_bytes(const mpsx_plan &p) {
    if (p.method == 0) return align_up(z_off(p) - p.ws_off + (int64_t)p.zbytes * p.Cp * p.LD, 256);
    if (p.method == 1) return align_up(qflag_off(p) - p.ws_off + 2 * 4LL * (p.nb / 2), 256);
    if (p.method == 3) return align_up(qorder_off(p) - p.ws_off + 4LL * p.R, 256);
    return align_up(t_off(p) - p.ws_off + 16 * h_elems(p), 256);
}

// Complex element of the working matrix: float2 (FP32 rotations) or
// double2 (FP64 rotations).
template <class T> struct Cx;
template <> struct Cx<float> {
    using type = float2;
    static MPSX_DEV float2 make(float x, float y) { return make_float2(x, y); }
};
template <> struct Cx<double> {
    using type = double2;
    static MPSX_DEV double2 make(double x, double y) { return make_double2(x, y); }
};

// Workgroups of the final gather of one problem: 32x32 tiles of the
// M x k unscaled-by-column factor (read as a transpose of rows of Z), then
// one per row of the k x N factor.
MPSX_HD int64_t gather_units(int64_t M, int64_t k) {
    return ((M + 31) / 32) * ((k + 31) / 32) + k;
}

// Tile sizes of the theta kernel.
constexpr int THETA_TL = 32;  // l values per tile
constexpr int THETA_TR = 32;  // r values per tile
constexpr int THETA_KT = 16;  // bond (m) values per LDS stage

// Qudit dimensions 5..16 take the generic theta kernel: one workgroup per
// (l, tile of r values); the tile's d^2 x tr contraction T = A B stays in
// LDS (at most 4096 complex64 entries) for the gate product.
constexpr int THETA_GEN_DMAX = 16;
MPSX_HD int theta_gen_tr(int d) { return d <= 8 ? 64 : 16; }
