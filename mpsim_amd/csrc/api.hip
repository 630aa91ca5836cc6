// C ABI of libmpsx: argument checks, launch planning and the two-site
// sweep driver. Kernels live in site.hip, gemm.hip and twosite.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "driver.h"
#include "launch.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(MPSX_EHIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                      \
    do {                                         \
        hipError_t e_ = (expr);                  \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

constexpr int64_t kHeaderBytes = 1024;
constexpr int kDefaultMaxSweeps = 30;
constexpr int kDefaultInnerSweeps = 1;
// Defaults by working precision. FP64 rotations converge to 1e-7 in the
// same number of sweeps as to 1e-5 (the iteration is quadratic at the end)
// and can follow rows down to 1e-9 of the norm; FP32 rotations stop at the
// FP32 noise floor.
constexpr double kDefaultTol64 = 1e-7, kDefaultTol32 = 1e-5;
constexpr double kNullRel64 = 1e-9, kNullRel32 = 1e-6;
constexpr int64_t kGridLimit = 2147483647LL;
// Absolute floor of the Gram method's rotation test, relative to
// ||theta||_F^2: above the FP64 rounding that H gathers over the rounds.
constexpr double kGramAbsRel = 1e-12;
// The FP32 phase of the mixed Gram method (opt-in: opts.fp32_sweeps > 0):
// relative tolerance and absolute floor (above the FP32 rounding H32
// gathers; pairs below it are left to the FP64 rounds).
constexpr double kGramTol32 = 1e-5, kGramAbsRel32 = 3e-8;

struct Resolved {
    int block, max_sweeps, check_every, inner_sweeps;
    double tol;
    int block_switch;
    int *unconverged;
    bool fp32;
    double null_rel;
    int zbytes;
    int method;
    int fp32_sweeps;  // sweep limit of the Gram method's FP32 phase; 0: no such phase
    int precondition;  // 0 none, 1 two-site problems of saturated bonds, 2 every Gram problem
};

Resolved resolve(const mpsx_svd_opts *o) {
    Resolved r{0, kDefaultMaxSweeps, 1, kDefaultInnerSweeps, 0.0, 64, nullptr, false, 0.0, 16,
               MPSX_SVD_AUTO, 0, 0};
    if (o) {
        if (o->block) r.block = o->block;
        if (o->max_sweeps) r.max_sweeps = o->max_sweeps;
        r.check_every = o->check_every;
        if (o->inner_sweeps) r.inner_sweeps = o->inner_sweeps;
        if (o->tol > 0) r.tol = o->tol;
        if (o->block_switch > 0) r.block_switch = o->block_switch;
        r.unconverged = o->unconverged;
        r.fp32 = o->rotation_fp32 != 0;
        r.method = o->method;
        r.fp32_sweeps = o->fp32_sweeps > 0 ? o->fp32_sweeps : 0;
        r.precondition = o->precondition;
    }
    if (r.tol <= 0) r.tol = r.fp32 ? kDefaultTol32 : kDefaultTol64;
    r.null_rel = r.fp32 ? kNullRel32 : kNullRel64;
    r.zbytes = r.fp32 ? 8 : 16;
    return r;
}

// ------------------------------------------------------------ profiler
struct ProfRec {
    const char *name;
    hipEvent_t a, b;
};
struct Profiler {
    std::mutex mu;
    bool on = false;
    std::vector<ProfRec> recs;
    std::vector<hipEvent_t> pool;
    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        // timing-only events: no system-scope fence (cache writeback and
        // invalidation) when they are recorded, which would otherwise stall
        // the stream between the kernels being timed
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
        return e;
    }
};
Profiler &prof() {
    static Profiler p;
    return p;
}

// Set while a driver call runs that must create no events (may_poll false).
thread_local bool g_no_profile = false;
struct NoProfile {
    bool prev;
    explicit NoProfile(bool on) : prev(g_no_profile) { g_no_profile = g_no_profile || on; }
    ~NoProfile() { g_no_profile = prev; }
};

// Launches `fn` and, when profiling is on, brackets it with events.
template <class F>
hipError_t timed(const char *name, hipStream_t s, F &&fn) {
    Profiler &p = prof();
    if (!p.on || g_no_profile) return fn();
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone)
        return fn();  // events inside a captured graph would not be replay-timed
    hipEvent_t a, b;
    {
        std::lock_guard<std::mutex> g(p.mu);
        a = p.get();
        b = p.get();
    }
    if (!a || !b) return fn();
    hipError_t e = hipEventRecord(a, s);
    if (e != hipSuccess) return e;
    e = fn();
    if (e != hipSuccess) return e;
    e = hipEventRecord(b, s);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> g(p.mu);
    p.recs.push_back({name, a, b});
    return hipSuccess;
}

// Workgroups of the theta phase: 32x32 (l, r) tiles of a two-site problem,
// or one per matrix row of a plain split.
int64_t theta_tiles(const mpsx_two_site_op &o) {
    if (o.gate == nullptr) return o.chi_l;
    if (o.d > 4) {  // generic kernel: one workgroup per (l, tile of r)
        const int tr = theta_gen_tr(o.d);
        return (int64_t)o.chi_l * (((int64_t)o.chi_r + tr - 1) / tr);
    }
    return ((int64_t)o.chi_l + THETA_TL - 1) / THETA_TL * (((int64_t)o.chi_r + THETA_TR - 1) / THETA_TR);
}

// Rows per Jacobi block: 2b-row pivots of 16 for small problems, 32 once
// there are more than `switch_rows` rows (the autotuner measures the
// crossover; 64 by default).
int auto_block(int C, int switch_rows) { return C <= switch_rows ? 8 : 16; }

// The Gram method serves FP64 problems with 32-row pivots and at least six
// 16-row blocks unless the one-sided method is asked for (with four blocks
// one H tile holds the cross blocks of both pivots of the next round, which
// the pipelined rounds do not allow): method 1, or method 2 with an FP32
// phase before the FP64 rounds when opts.fp32_sweeps asks for one.
int plan_method(const Resolved &r, int b, int nb) {
    if (r.fp32 || b != 16 || nb < 6 || r.method == MPSX_SVD_ROWS) return 0;
    return r.fp32_sweeps > 0 ? 2 : 1;
}

int *poll_buffer() {
    thread_local int *buf = nullptr;
    if (!buf) {
        void *p = nullptr;
        if (hipHostMalloc(&p, 256, hipHostMallocDefault) != hipSuccess) return nullptr;
        buf = static_cast<int *>(p);
    }
    return buf;
}

}  // namespace

extern "C" {

int mpsx_version(void) { return 100; }

const char *mpsx_last_error(void) { return g_err.c_str(); }

size_t mpsx_sizeof(int which) {
    switch (which) {
        case 0: return sizeof(mpsx_site_op);
        case 1: return sizeof(mpsx_two_site_op);
        case 2: return sizeof(mpsx_gemm_op);
        case 3: return sizeof(mpsx_scale_op);
        case 4: return sizeof(mpsx_svd_opts);
        case 5: return sizeof(mpsx_plan);
        default: return 0;
    }
}

int mpsx_device_info(int device, char *arch, int arch_len, int *num_cus) {
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    if (arch && arch_len > 0) {
        std::strncpy(arch, prop.gcnArchName, (size_t)arch_len - 1);
        arch[arch_len - 1] = '\0';
    }
    if (num_cus) *num_cus = prop.multiProcessorCount;
    return MPSX_OK;
}

// ---------------------------------------------------------------- one-site
int mpsx_apply_one_site(mpsx_stream stream, const mpsx_site_op *h, const mpsx_site_op *d,
                        int count) {
    if (count < 0) return fail(MPSX_EINVAL, "count < 0");
    if (count == 0) return MPSX_OK;
    if (!h || !d) return fail(MPSX_EINVAL, "null descriptor table");
    if (count > 65535) return fail(MPSX_EINVAL, "count %d > 65535", count);
    int64_t max_pairs = 0;
    uint32_t dmask = 0;  // bit d: some op has qudit dimension d
    for (int i = 0; i < count; ++i) {
        const mpsx_site_op &o = h[i];
        if (o.d < 1 || o.d > 16) return fail(MPSX_EINVAL, "op %d: d=%d outside [1,16]", i, o.d);
        if (o.chi_l < 0 || o.chi_r < 0) return fail(MPSX_EINVAL, "op %d: negative bond", i);
        const int64_t pairs = o.chi_l * o.chi_r;
        if (pairs > 0 && (!o.a || !o.out || !o.gate))
            return fail(MPSX_EINVAL, "op %d: null pointer", i);
        max_pairs = std::max(max_pairs, pairs);
        if (pairs > 0) dmask |= 1u << o.d;
    }
    if (max_pairs == 0) return MPSX_OK;
    HIP_TRY(timed("k_one_site", (hipStream_t)stream, [&] { return launch_one_site((hipStream_t)stream, d, count, max_pairs, dmask); }), "k_one_site");
    return MPSX_OK;
}

// ---------------------------------------------------------------- scale
int mpsx_scale(mpsx_stream stream, const mpsx_scale_op *h, const mpsx_scale_op *d, int count) {
    if (count < 0) return fail(MPSX_EINVAL, "count < 0");
    if (count == 0) return MPSX_OK;
    if (!h || !d) return fail(MPSX_EINVAL, "null descriptor table");
    if (count > 65535) return fail(MPSX_EINVAL, "count %d > 65535", count);
    int64_t max_n = 0;
    for (int i = 0; i < count; ++i) {
        if (h[i].n < 0) return fail(MPSX_EINVAL, "op %d: n < 0", i);
        if (h[i].n > 0 && !h[i].x) return fail(MPSX_EINVAL, "op %d: null pointer", i);
        max_n = std::max(max_n, h[i].n);
    }
    if (max_n == 0) return MPSX_OK;
    HIP_TRY(timed("k_scale", (hipStream_t)stream, [&] { return launch_scale((hipStream_t)stream, d, count, max_n); }), "k_scale");
    return MPSX_OK;
}

// ---------------------------------------------------------------- gemm
int mpsx_gemm(mpsx_stream stream, const mpsx_gemm_op *h, const mpsx_gemm_op *d, int count) {
    if (count < 0) return fail(MPSX_EINVAL, "count < 0");
    if (count == 0) return MPSX_OK;
    if (!h || !d) return fail(MPSX_EINVAL, "null descriptor table");
    if (count > 65535) return fail(MPSX_EINVAL, "count %d > 65535", count);
    int64_t max_m = 0, max_n = 0;
    for (int i = 0; i < count; ++i) {
        const mpsx_gemm_op &o = h[i];
        if (o.m < 0 || o.n < 0 || o.k < 0) return fail(MPSX_EINVAL, "op %d: negative size", i);
        const bool ta = o.flags & 1, tb = o.flags & 2;
        if (o.m > 0 && o.n > 0) {
            if (!o.C) return fail(MPSX_EINVAL, "op %d: null C", i);
            if (o.k > 0 && (!o.A || !o.B)) return fail(MPSX_EINVAL, "op %d: null A/B", i);
            if (o.ldc < o.n) return fail(MPSX_EINVAL, "op %d: ldc < n", i);
            if (o.k > 0 && o.lda < (ta ? o.m : o.k)) return fail(MPSX_EINVAL, "op %d: lda too small", i);
            if (o.k > 0 && o.ldb < (tb ? o.k : o.n)) return fail(MPSX_EINVAL, "op %d: ldb too small", i);
        }
        max_m = std::max(max_m, o.m);
        max_n = std::max(max_n, o.n);
    }
    if (max_m == 0 || max_n == 0) return MPSX_OK;
    bool rows = true;
    for (int i = 0; i < count && rows; ++i) rows = gemm_rows_fit(h[i]);
    if (rows) {
        HIP_TRY(timed("k_gemm_rows", (hipStream_t)stream, [&] { return launch_gemm_rows((hipStream_t)stream, h, d, count); }), "k_gemm_rows");
        return MPSX_OK;
    }
    if (((max_m + 31) / 32) * ((max_n + 31) / 32) > 2147483647LL)
        return fail(MPSX_EINVAL, "m x n too large for one launch");
    HIP_TRY(timed("k_gemm", (hipStream_t)stream, [&] { return launch_gemm((hipStream_t)stream, d, count, max_m, max_n); }), "k_gemm");
    return MPSX_OK;
}

// ---------------------------------------------------------------- two-site
}  // extern "C"

namespace {

// Validates op i and fills the shape fields of its plan (everything but the
// five prefix offsets) from (d, chi_l, chi_m, chi_r, k, flags) and the
// resolved options. The planner and the check in mpsx_two_site share it, so
// a table planned for other shapes, flags or options is caught before
// anything is launched.
int plan_shape(const mpsx_two_site_op &o, int i, const Resolved &r, mpsx_plan *p) {
    const bool split = o.gate == nullptr;
    if (split ? o.d != 1 : (o.d < 2 || o.d > THETA_GEN_DMAX))
        return fail(MPSX_EINVAL, "op %d: d=%d outside [2,%d] (or 1 for a plain split)", i, o.d,
                    THETA_GEN_DMAX);
    if (o.chi_l < 0 || o.chi_m < 0 || o.chi_r < 0)
        return fail(MPSX_EINVAL, "op %d: negative bond dimension", i);
    const int64_t M = (int64_t)o.d * o.chi_l, N = (int64_t)o.d * o.chi_r;
    if (o.k < 0 || o.k > std::min(M, N))
        return fail(MPSX_EINVAL, "op %d: k=%d outside [0, min(%lld,%lld)]", i, o.k, (long long)M,
                    (long long)N);
    if (M > 65536 || N > 65536) return fail(MPSX_EINVAL, "op %d: matrix too large", i);
    if (M * N > 0 && !split && o.chi_m > 0 && (!o.a || !o.b))
        return fail(MPSX_EINVAL, "op %d: null input", i);
    if (M * N > 0 && split && !o.a) return fail(MPSX_EINVAL, "op %d: null input", i);
    if (M * o.k > 0 && !o.a_out) return fail(MPSX_EINVAL, "op %d: null a_out", i);
    if (N * o.k > 0 && !o.b_out) return fail(MPSX_EINVAL, "op %d: null b_out", i);
    if (o.flags & ~1) return fail(MPSX_EINVAL, "op %d: unknown flag bits 0x%x", i, o.flags);
    const bool left_absorb = (o.flags & 1) != 0;
    std::memset(p, 0, sizeof *p);
    p->C = (int32_t)(left_absorb ? N : M);
    p->R = (int32_t)(left_absorb ? M : N);
    p->b = r.block ? r.block : auto_block(p->C, r.block_switch);
    if (p->C > 0) {
        int nb = (p->C + p->b - 1) / p->b;
        if (nb < 2) nb = 2;
        if (nb & 1) ++nb;
        p->nb = nb;
        p->Cp = nb * p->b;
    }
    p->LD = (int32_t)align_up((int64_t)p->R + p->Cp, 4);
    p->zbytes = r.zbytes;
    p->method = p->nb > 0 ? plan_method(r, p->b, p->nb) : 0;
    if (p->method == 1 && p->Cp <= QR_MAX_CP && p->R <= QR_MAX_R &&
        (r.precondition == 2 ||
         (r.precondition == 1 && !split && o.chi_m >= std::min(o.chi_l, o.chi_r))))
        p->method = 3;
    // Tall and preconditioned: R has no rows beyond R, so the Jacobi runs on
    // R's 16-row blocks (when that still makes at least six), while Z, Y and
    // the QR keep C's padded rows (Y is Cp wide, U has C rows)
    if (p->method == 3 && p->C > p->R) {
        int nb = (p->R + 15) / 16;
        if (nb & 1) ++nb;
        if (nb >= 6) p->nb = nb;
    }
    return MPSX_OK;
}

int check_options(const Resolved &r) {
    if (r.block != 0 && r.block != 8 && r.block != 16)
        return fail(MPSX_EINVAL, "block %d not in {0 (auto), 8, 16}", r.block);
    if (r.max_sweeps < 1 || r.max_sweeps > 64) return fail(MPSX_EINVAL, "max_sweeps outside [1,64]");
    if (r.method != MPSX_SVD_AUTO && r.method != MPSX_SVD_ROWS)
        return fail(MPSX_EINVAL, "method %d not in {0 (auto), 1 (rows)}", r.method);
    if (r.fp32_sweeps > 64) return fail(MPSX_EINVAL, "fp32_sweeps above 64");
    if (r.precondition < 0 || r.precondition > 2)
        return fail(MPSX_EINVAL, "precondition %d not in {0, 1, 2}", r.precondition);
    return MPSX_OK;
}

}  // namespace

extern "C" {

int mpsx_two_site_plan(mpsx_two_site_op *h, int count, const mpsx_svd_opts *opts,
                       size_t *bytes) {
    if (count < 0) return fail(MPSX_EINVAL, "count < 0");
    if (count > 0 && !h) return fail(MPSX_EINVAL, "null descriptor table");
    const Resolved r = resolve(opts);
    if (int e = check_options(r)) return e;
    int64_t off = kHeaderBytes, tiles = 0, pairs = 0, rows = 0, outs = 0;
    for (int i = 0; i < count; ++i) {
        mpsx_two_site_op &o = h[i];
        mpsx_plan p;
        if (int e = plan_shape(o, i, r, &p)) return e;
        p.ws_off = off;
        p.tile_base = tiles;
        p.pair_base = pairs;
        p.row_base = rows;
        p.out_base = outs;
        o.plan = p;
        off += problem_bytes(p);
        tiles += theta_tiles(o);
        pairs += p.nb / 2;
        rows += p.Cp;
        outs += gather_units((int64_t)o.d * o.chi_l, o.k);
    }
    if (tiles > kGridLimit || pairs > kGridLimit || rows > kGridLimit || outs > kGridLimit)
        return fail(MPSX_EINVAL, "batch too large for one launch");
    // round workgroup -> problem table after the problems
    off = align_up(off, 256) + align_up(4 * pairs, 256);
    if (bytes) *bytes = (size_t)off;
    g_err.clear();
    return MPSX_OK;
}

}  // extern "C"

namespace {

// The two-site driver behind mpsx_two_site. `may_poll` false: never read
// device memory back (no convergence poll; the sweep limit and the
// device-side early exit bound the work) and create no profiling events.
int two_site_run(hipStream_t stream, const mpsx_two_site_op *h, const mpsx_two_site_op *d,
                 int count, const Resolved &r, char *ws, size_t ws_bytes, int *sweeps_out,
                 bool may_poll, size_t *need_out = nullptr) {
    if (sweeps_out) *sweeps_out = 0;
    if (count < 0) return fail(MPSX_EINVAL, "count < 0");
    if (count == 0) return MPSX_OK;
    if (!h || !d) return fail(MPSX_EINVAL, "null descriptor table");
    if (!ws) return fail(MPSX_EINVAL, "null workspace");
    if (int e = check_options(r)) return e;
    NoProfile quiet(!may_poll);
    // Re-plan every op from its shape, flags and the options and compare with
    // the table's plan: a table that was not planned, was planned for other
    // options, or was edited after planning is rejected.
    int64_t need = kHeaderBytes, tiles = 0, pairs = 0, rows = 0, outs = 0;
    int rows_nb = 0, gram_nb = 0, gram_y = 0, gram_c = 0;
    bool has_d[THETA_GEN_DMAX + 1] = {};  // index 1: plain splits
    bool has_b[17] = {false};  // one-sided problems by block width
    bool has_gram = false, has_mixed = false;
    int qr_c = 0, qr_r = 0, qr_rc = 0, qr_np = 0;  // QR-preconditioned problems: max C, R, R + C, panels
    bool qr_small = false, qr_large = false;  // some of at most 512 rows, some above
    for (int i = 0; i < count; ++i) {
        const mpsx_two_site_op &o = h[i];
        const mpsx_plan &p = o.plan;
        mpsx_plan want;
        if (int e = plan_shape(o, i, r, &want)) return e;
        want.ws_off = need;
        want.tile_base = tiles;
        want.pair_base = pairs;
        want.row_base = rows;
        want.out_base = outs;
        if (std::memcmp(&want, &p, sizeof want) != 0)
            return fail(MPSX_EINVAL,
                        "op %d: plan missing or stale (call mpsx_two_site_plan with the same "
                        "shapes, flags and options)", i);
        need += problem_bytes(p);
        tiles += theta_tiles(o);
        pairs += p.nb / 2;
        rows += p.Cp;
        outs += gather_units((int64_t)o.d * o.chi_l, o.k);
        has_d[o.d] = true;
        if (p.nb > 0 && p.method == 0) {
            has_b[p.b] = true;
            rows_nb = std::max(rows_nb, p.nb);
        }
        if (p.nb > 0 && p.method >= 1) {
            has_gram = true;
            has_mixed = has_mixed || p.method == 2;
            gram_nb = std::max(gram_nb, p.nb);
            gram_y = std::max(gram_y, p.nb / 4 + (p.Cp + 63) / 64);
            gram_c = std::max(gram_c, (p.R + 31) / 32);
        }
        if (p.nb > 0 && p.method == 3) {
            qr_c = std::max(qr_c, (int)p.C);
            qr_r = std::max(qr_r, (int)p.R);
            qr_rc = std::max(qr_rc, (int)(p.R + p.C));
            qr_np = std::max(qr_np, qr_panels(p));
            (p.C <= 512 ? qr_small : qr_large) = true;
        }
    }
    if (tiles > kGridLimit || pairs > kGridLimit || rows > kGridLimit || outs > kGridLimit)
        return fail(MPSX_EINVAL, "batch too large for one launch");
    const int64_t owner_off = align_up(need, 256);
    need = owner_off + align_up(4 * pairs, 256);
    if (need_out) *need_out = (size_t)need;
    if ((int64_t)ws_bytes < need)
        return fail(MPSX_EWORKSPACE, "workspace %zu bytes < %lld needed", ws_bytes, (long long)need);
    int *owner = reinterpret_cast<int *>(ws + owner_off);

    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(stream, &cap), "hipStreamIsCapturing");
    const bool can_poll = may_poll && r.check_every > 0 && cap == hipStreamCaptureStatusNone;

    const bool f32 = r.fp32;
    if (rows > 0) HIP_TRY(timed("k_init_rows", stream, [&] { return launch_init_rows(stream, f32, d, count, ws, rows); }), "k_init_rows");
    for (int D = 1; D <= 4; ++D)
        if (has_d[D] && tiles > 0) HIP_TRY(timed("k_theta", stream, [&] { return launch_theta(stream, f32, D, d, count, ws, tiles); }), "k_theta");
    bool has_gen = false;
    for (int D = 5; D <= THETA_GEN_DMAX; ++D) has_gen = has_gen || has_d[D];
    if (has_gen && tiles > 0) HIP_TRY(timed("k_theta", stream, [&] { return launch_theta_gen(stream, f32, d, count, ws, tiles); }), "k_theta");
    if (rows > 0) {
        HIP_TRY(timed("k_rownorm", stream, [&] { return launch_rownorm(stream, f32, d, count, ws, rows); }), "k_rownorm");
        HIP_TRY(timed("k_fro", stream, [&] { return launch_fro(stream, d, count, ws, r.null_rel); }), "k_fro");
    }
    if (qr_c > 0) {
        // [X P | I] -> [R | Q^H] for the preconditioned problems, X0 saved
        HIP_TRY(timed("k_qr_order", stream, [&] { return launch_qr_order(stream, d, count, ws, qr_c, qr_r); }), "k_qr_order");
        HIP_TRY(timed("k_qr_permute", stream, [&] { return launch_qr_permute(stream, d, count, ws, rows); }),
                "k_qr_permute");
        for (int panel = -1; panel < qr_np; ++panel) {
            const int strips = panel < 0 ? 1 : (qr_rc - 16 * (panel + 1) + 15) / 16;
            if (strips > 0)
                HIP_TRY(timed("k_qr_step", stream,
                              [&] {
                                  return launch_qr_step(stream, d, count, ws, panel, strips, qr_c, qr_small,
                                                        qr_large);
                              }),
                        "k_qr_step");
        }
    }

    int sweeps = 0;
    bool converged = false;
    if (pairs > 0)
        HIP_TRY(timed("k_owner", stream, [&] { return launch_owner(stream, d, count, owner); }), "k_owner");
    if (has_gram)
        HIP_TRY(timed("k_hgram", stream, [&] { return launch_hgram(stream, d, owner, ws, pairs, gram_nb, 0); }), "k_hgram");
    // sweeps of both phases so far: indexes the header's active counters
    int sweep_idx = 0;
    int *pin = (pairs > 0 && can_poll) ? poll_buffer() : nullptr;
    auto poll_done = [&](int idx, bool *done) -> int {
        const TwoSiteHeader *hdr = reinterpret_cast<const TwoSiteHeader *>(ws);
        HIP_TRY(hipMemcpyAsync(pin, &hdr->active[idx & 63], sizeof(int), hipMemcpyDeviceToHost, stream),
                "hipMemcpyAsync");
        HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
        *done = *pin == 0;
        return MPSX_OK;
    };
    if (has_mixed) {
        // FP32 phase of the mixed Gram problems: the same pipelined rounds on
        // H32 (FP32 tiles on the FP32 matrix cores; FP64 eigensolves, Q and
        // Y), until a sweep rotates nothing or the phase's sweep limit; then
        // X1 = Y X and H = X1 X1^H in FP64 for the FP64 rounds below.
        const double ti = kGramTol32 * 0.01;
        for (int s = 0; s < r.fp32_sweeps; ++s) {
            HIP_TRY(timed("k_hsolve32", stream,
                          [&] {
                              return launch_hsolve(stream, true, d, owner, ws, pairs, 0, kGramTol32, ti,
                                                   r.inner_sweeps, kGramAbsRel32);
                          }),
                    "k_hsolve32");
            for (int round = 0; round < gram_nb - 2; ++round)
                HIP_TRY(timed("k_hround32", stream,
                              [&] {
                                  return launch_hround(stream, true, d, owner, ws, pairs, gram_y, round,
                                                       kGramTol32, ti, r.inner_sweeps, kGramAbsRel32);
                              }),
                        "k_hround32");
            HIP_TRY(timed("k_hupdate32", stream,
                          [&] { return launch_hupdate(stream, true, d, owner, ws, pairs, gram_y, gram_nb - 2, 0); }),
                    "k_hupdate32");
            HIP_TRY(timed("k_sweep_end", stream, [&] { return launch_sweep_end(stream, d, count, ws, sweep_idx, true); }), "k_sweep_end");
            ++sweep_idx;
            if (pin && ((s + 1) % r.check_every == 0 || s + 1 == r.fp32_sweeps)) {
                bool done = false;
                if (int e = poll_done(sweep_idx - 1, &done)) return e;
                if (done) break;
            }
        }
        HIP_TRY(timed("k_hxgemm", stream, [&] { return launch_hxgemm(stream, d, owner, ws, pairs, gram_c, 1, gram_nb); }), "k_hxgemm");
        HIP_TRY(timed("k_hgram", stream, [&] { return launch_hgram(stream, d, owner, ws, pairs, gram_nb, 1); }), "k_hgram");
    }
    if (pairs > 0) {
        for (int s = 0; s < r.max_sweeps; ++s) {
            for (int round = 0; round < rows_nb - 1; ++round)
                for (int b : {8, 16})
                    if (has_b[b])
                        HIP_TRY(timed(b == 8 ? (f32 ? "k_jacobi_round32<16>" : "k_jacobi_round64<16>")
                                             : (f32 ? "k_jacobi_round32<32>" : "k_jacobi_round64<32>"),
                                      stream,
                                      [&] {
                                          return launch_jacobi_round(stream, f32, b, d, owner, ws, pairs, round,
                                                                     r.tol, r.tol * 0.01, r.inner_sweeps);
                                      }),
                                "k_jacobi_round");
            // Gram rounds 0 .. R-1 (R = gram_nb - 1), pipelined on the stream:
            //   k_hsolve(0); then for r < R-1: k_hround(eigensolves of r+1,
            //   each after the tile holding its cross block, beside the rest
            //   of round r's update); finally k_hupdate(R-1, whole).
            if (has_gram) {
                const double ti = r.tol * 0.01;
                HIP_TRY(timed("k_hsolve", stream,
                              [&] {
                                  return launch_hsolve(stream, false, d, owner, ws, pairs, 0, r.tol, ti,
                                                       r.inner_sweeps, kGramAbsRel);
                              }),
                        "k_hsolve");
                for (int round = 0; round < gram_nb - 2; ++round) {
                    HIP_TRY(timed("k_hround", stream,
                                  [&] {
                                      return launch_hround(stream, false, d, owner, ws, pairs, gram_y, round, r.tol,
                                                           ti, r.inner_sweeps, kGramAbsRel);
                                  }),
                            "k_hround");
                }
                HIP_TRY(timed("k_hupdate", stream,
                              [&] {
                                  return launch_hupdate(stream, false, d, owner, ws, pairs, gram_y, gram_nb - 2, 0);
                              }),
                        "k_hupdate");
            }
            HIP_TRY(timed("k_sweep_end", stream, [&] { return launch_sweep_end(stream, d, count, ws, sweep_idx, false); }), "k_sweep_end");
            ++sweep_idx;
            ++sweeps;
            if (pin && ((s + 1) % r.check_every == 0 || s + 1 == r.max_sweeps)) {
                bool done = false;
                if (int e = poll_done(sweep_idx - 1, &done)) return e;
                if (done) {
                    converged = true;
                    break;
                }
            }
        }
    } else {
        converged = true;
    }

    if (has_gram) {
        HIP_TRY(timed("k_hxcopy", stream, [&] { return launch_hxcopy(stream, d, count, ws, rows); }), "k_hxcopy");
        HIP_TRY(timed("k_hxgemm", stream, [&] { return launch_hxgemm(stream, d, owner, ws, pairs, gram_c, 0, gram_nb); }), "k_hxgemm");
    }
    if (rows > 0) {
        HIP_TRY(timed("k_rownorm", stream, [&] { return launch_rownorm(stream, f32, d, count, ws, rows); }), "k_rownorm");
        HIP_TRY(timed("k_sort", stream, [&] { return launch_sort(stream, d, count, ws, r.unconverged); }), "k_sort");
    }
    if (outs > 0) HIP_TRY(timed("k_gather", stream, [&] { return launch_gather(stream, f32, d, count, ws, outs); }), "k_gather");
    if (sweeps_out) *sweeps_out = sweeps;
    if (can_poll && !converged)
        return fail(MPSX_ENOTCONV, "Jacobi SVD not converged after %d sweeps", sweeps);
    return MPSX_OK;
}

}  // namespace

// Library-internal entry of the driver for the other source files
// (driver.h); hidden from the shared library's exports.
int mpsx_internal_two_site(hipStream_t stream, const mpsx_two_site_op *h,
                           const mpsx_two_site_op *d, int count, const mpsx_svd_opts *opts,
                           char *ws, size_t ws_bytes, int *sweeps_out, bool may_poll,
                           size_t *need_out) {
    return two_site_run(stream, h, d, count, resolve(opts), ws, ws_bytes, sweeps_out, may_poll,
                        need_out);
}

int mpsx_internal_fail(int code, const char *msg) { return fail(code, "%s", msg); }

extern "C" {

int mpsx_two_site(mpsx_stream stream, const mpsx_two_site_op *h, const mpsx_two_site_op *d,
                  int count, const mpsx_svd_opts *opts, void *ws, size_t ws_bytes,
                  int *sweeps_out) {
    return two_site_run((hipStream_t)stream, h, d, count, resolve(opts), static_cast<char *>(ws),
                        ws_bytes, sweeps_out, true);
}


// ---------------------------------------------------------------- graphs
struct mpsx_graph_s {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

int mpsx_graph_begin(mpsx_stream stream) {
    if (!stream) return fail(MPSX_EINVAL, "the null stream cannot be captured");
    HIP_TRY(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal),
            "hipStreamBeginCapture");
    return MPSX_OK;
}

int mpsx_graph_end(mpsx_stream stream, mpsx_graph *out) {
    if (!out) return fail(MPSX_EINVAL, "null output");
    *out = nullptr;
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamEndCapture((hipStream_t)stream, &g), "hipStreamEndCapture");
    hipGraphExec_t e = nullptr;
    const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    if (err != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(MPSX_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(err));
    }
    *out = new mpsx_graph_s{g, e};
    return MPSX_OK;
}

int mpsx_graph_launch(mpsx_graph graph, mpsx_stream stream) {
    if (!graph) return fail(MPSX_EINVAL, "null graph");
    HIP_TRY(hipGraphLaunch(graph->exec, (hipStream_t)stream), "hipGraphLaunch");
    return MPSX_OK;
}

int mpsx_graph_destroy(mpsx_graph graph) {
    if (!graph) return MPSX_OK;
    HIP_TRY(hipGraphExecDestroy(graph->exec), "hipGraphExecDestroy");
    HIP_TRY(hipGraphDestroy(graph->graph), "hipGraphDestroy");
    delete graph;
    return MPSX_OK;
}

int mpsx_profile_enable(int on) {
    Profiler &p = prof();
    std::lock_guard<std::mutex> g(p.mu);
    p.on = on != 0;
    return MPSX_OK;
}

}  // extern "C"

namespace {
// Length of the union of intervals (sorted in place).
double union_length(std::vector<std::pair<double, double>> &sp) {
    std::sort(sp.begin(), sp.end());
    double busy = 0.0, lo = 0.0, hi = -1.0;
    for (const auto &iv : sp) {
        if (iv.first > hi) {
            if (hi > lo) busy += hi - lo;
            lo = iv.first;
            hi = iv.second;
        } else if (iv.second > hi) {
            hi = iv.second;
        }
    }
    if (hi > lo) busy += hi - lo;
    return busy;
}
}  // namespace

extern "C" {

int mpsx_profile_busy(const char *names, double *busy_ms) {
    if (!names || !busy_ms) return fail(MPSX_EINVAL, "null argument");
    std::vector<std::string> want;
    for (const char *a = names; *a;) {
        const char *b = std::strchr(a, ',');
        const size_t n = b ? (size_t)(b - a) : std::strlen(a);
        if (n) want.emplace_back(a, n);
        a += n + (b ? 1 : 0);
    }
    Profiler &p = prof();
    std::lock_guard<std::mutex> g(p.mu);
    std::vector<std::pair<double, double>> spans;
    for (const ProfRec &r : p.recs) {
        if (std::find(want.begin(), want.end(), std::string(r.name)) == want.end()) continue;
        HIP_TRY(hipEventSynchronize(r.b), "hipEventSynchronize");
        float ms = 0.f, t0 = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, r.a, r.b), "hipEventElapsedTime");
        HIP_TRY(hipEventElapsedTime(&t0, p.recs.front().a, r.a), "hipEventElapsedTime");
        spans.emplace_back((double)t0, (double)t0 + ms);
    }
    *busy_ms = union_length(spans);
    return MPSX_OK;
}

int mpsx_profile_read(char *names, int name_len, double *total_ms, double *busy_ms,
                      int64_t *launches, int cap, int reset) {
    Profiler &p = prof();
    std::lock_guard<std::mutex> g(p.mu);
    struct Agg {
        double total = 0.0;
        int64_t n = 0;
        std::vector<std::pair<double, double>> spans;  // relative to the first record
    };
    std::map<std::string, Agg> agg;
    for (const ProfRec &r : p.recs) {
        HIP_TRY(hipEventSynchronize(r.b), "hipEventSynchronize");
        float ms = 0.f, t0 = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, r.a, r.b), "hipEventElapsedTime");
        HIP_TRY(hipEventElapsedTime(&t0, p.recs.front().a, r.a), "hipEventElapsedTime");
        Agg &slot = agg[r.name];
        slot.total += ms;
        slot.n += 1;
        slot.spans.emplace_back((double)t0, (double)t0 + ms);
    }
    int i = 0;
    for (auto &kv : agg) {
        if (i < cap) {
            if (names && name_len > 0) {
                std::strncpy(names + (size_t)i * name_len, kv.first.c_str(), (size_t)name_len - 1);
                names[(size_t)i * name_len + name_len - 1] = '\0';
            }
            if (total_ms) total_ms[i] = kv.second.total;
            if (launches) launches[i] = kv.second.n;
            // union of the launch intervals (launches on several streams overlap)
            if (busy_ms) busy_ms[i] = union_length(kv.second.spans);
        }
        ++i;
    }
    if (reset) {
        for (const ProfRec &r : p.recs) {
            p.pool.push_back(r.a);
            p.pool.push_back(r.b);
        }
        p.recs.clear();
    }
    return i;
}

}  // extern "C"
