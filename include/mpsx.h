/*
 * mpsx.h - C ABI of the MI355X (gfx950) MPS gate-application library.
 *
 * The reference (grmlarose/mpsim) is pure Python with no FFI; its hot path
 * lives behind the Python object API of mpsim.MPS. Each entry point below
 * replaces one arithmetic step of that API; the Python package mpsim_amd
 * keeps the reference's names, arguments and exceptions and calls these
 * through ctypes. Citations are to the reference's files.
 *
 * Conventions
 *   - All tensors are complex64 (interleaved float re, im) in device memory,
 *     owned by the caller. A site tensor is [chi_l, d, chi_r] row-major;
 *     chain ends carry a dummy bond of size 1.
 *   - Every batched entry point takes the same descriptor array twice:
 *     `h` in host memory (read by the library to size launches) and `d` in
 *     device memory (read by the kernels). The two must hold identical
 *     bytes and `d` must stay valid until the enqueued work has finished.
 *   - `stream` is a hipStream_t (NULL = the default stream). Calls enqueue
 *     work and return; they never synchronise the stream unless stated.
 *   - The library allocates no device memory. Scratch space is passed in by
 *     the caller (`ws`, `ws_bytes`), sized by the matching *_plan call; the
 *     one exception is mpsx_two_site_static's fixed per-device array.
 *   - Return value: 0 on success, otherwise an MPSX_E* code; the message is
 *     available from mpsx_last_error() on the calling thread.
 */
#ifndef MPSX_H
#define MPSX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPSX_OK 0
#define MPSX_EINVAL 1      /* bad argument (shape, size, null pointer)      */
#define MPSX_EHIP 2        /* a HIP runtime call failed                     */
#define MPSX_EWORKSPACE 3  /* workspace too small                           */
#define MPSX_ENOTCONV 4    /* SVD hit the sweep limit before converging     */

typedef void *mpsx_stream; /* hipStream_t */

/* Library version (major*10000 + minor*100 + patch). */
int mpsx_version(void);

/* Message for the last failed call on this thread ("" if none). */
const char *mpsx_last_error(void);

/* gfx arch name of `device` (e.g. "gfx950") and its compute-unit count. */
int mpsx_device_info(int device, char *arch, int arch_len, int *num_cus);

/* sizeof of the ABI structs, so bindings can check their layouts:
 * 0 mpsx_site_op, 1 mpsx_two_site_op, 2 mpsx_gemm_op, 3 mpsx_scale_op,
 * 4 mpsx_svd_opts, 5 mpsx_plan. Returns 0 for an unknown index. */
size_t mpsx_sizeof(int which);

/* ------------------------------------------------------------------------
 * One-site gate.
 * Replaces MPS.apply_one_qudit_gate (mpsim/core.py:753-826): gate edge 1
 * contracts the site's physical leg, gate edge 0 becomes the new leg,
 *     out[l, o, r] = sum_p G[o, p] * a[l, p, r].
 * `out` may alias `a`.
 * ---------------------------------------------------------------------- */
typedef struct {
    const void *a;    /* [chi_l, d, chi_r]                       */
    void *out;        /* [chi_l, d, chi_r]                       */
    const void *gate; /* [d, d] row-major G[o, p]                */
    int64_t chi_l;
    int64_t chi_r;
    int32_t d;
    int32_t reserved;
} mpsx_site_op;

int mpsx_apply_one_site(mpsx_stream stream, const mpsx_site_op *h,
                        const mpsx_site_op *d, int count);

/* ------------------------------------------------------------------------
 * Two-site update on adjacent sites (i, i+1).
 * Replaces the arithmetic of MPS.apply_two_qudit_gate
 * (mpsim/core.py:1045-1152) and TensorNetwork 0.2.1's split_node_full_svd
 * (called at core.py:1132-1137):
 *   theta[l, o1, o2, r] = sum G[o1, o2, p1, p2] a[l, p1, m] b[m, p2, r]
 *   theta (rows (l,o1), cols (o2,r)) = U S V^H, singular values descending,
 *   ties in input order, truncated to k columns (zero values are kept).
 *   absorb right (flags bit0 = 0, keep_left_canonical=True):
 *       a_out = U[:, :k]            as [chi_l, d, k]
 *       b_out = S[:k] V^H[:k, :]    as [k, d, chi_r]
 *   absorb left  (flags bit0 = 1, keep_left_canonical=False):
 *       a_out = U[:, :k] S[:k],  b_out = V^H[:k, :]
 * The caller computes k = min(d*chi_l, d*chi_r, maxsvals) (core.py:1110-1130).
 * The SVD is a blocked one-sided Jacobi on the side that is not scaled by S,
 * so the scaled side never divides by a singular value. theta is formed in
 * complex FP32; by default the rotations are stored and applied in complex
 * FP64 (FP64 matrix cores), which keeps each truncated update as accurate
 * as a complex64 LAPACK SVD. With rotation_fp32 set they run in FP32
 * (faster, about 1e-6 relative error per update).
 *
 * Plain split: with gate == NULL and d == 1 the op factorises the M x N
 * matrix stored row-major at `a` (M = chi_l, N = chi_r; `b`, chi_m unused)
 * the same way. This replaces the other tn.split_node(_full_svd) calls of
 * the reference: MPS.from_wavefunction (core.py:314-319) and the
 * orthonormalisation after non-unitary gates (core.py:877-882, 924-929).
 * ---------------------------------------------------------------------- */
/* Launch plan of one problem. Filled by mpsx_two_site_plan; callers copy it
 * to the device descriptor unchanged and never edit it. */
typedef struct {
    int64_t ws_off;    /* byte offset of this problem's scratch in ws     */
    int64_t tile_base; /* exclusive prefix of theta tiles                 */
    int64_t pair_base; /* exclusive prefix of Jacobi pivot pairs          */
    int64_t row_base;  /* exclusive prefix of padded rows                 */
    int64_t out_base;  /* exclusive prefix of gather workgroups          */
    int32_t C;         /* rows orthogonalised (d*chi_l or d*chi_r)         */
    int32_t R;         /* row length (the other dimension)                */
    int32_t Cp;        /* C padded (to nb*b but for method 3 with C > R) */
    int32_t LD;        /* row stride of the working matrix (elements)     */
    int32_t nb;        /* number of row blocks the Jacobi pairs (even; for
                          method 3 with C > R, R's blocks: the rows of the
                          QR's R beyond R are zero)                      */
    int32_t b;         /* rows per block                                  */
    int32_t zbytes;    /* bytes per working-matrix element: 16 (complex
                          FP64 rotations) or 8 (complex FP32)             */
    int32_t method;    /* 0: one-sided Jacobi on the rows of [X | Y];
                          1: two-sided Jacobi on the Gram H = X X^H;
                          2: the same with an FP32 first phase;
                          3: method 1 after a Householder QR of X         */
} mpsx_plan;

typedef struct {
    const void *a;     /* [chi_l, d, chi_m]                              */
    const void *b;     /* [chi_m, d, chi_r]                              */
    const void *gate;  /* [d*d, d*d] row-major G[(o1,o2), (p1,p2)]       */
    void *a_out;       /* [chi_l, d, k]                                  */
    void *b_out;       /* [k, d, chi_r]                                  */
    float *s_out;      /* k singular values, or NULL                     */
    double *discarded; /* sum of squared discarded values, or NULL       */
    int32_t chi_l, chi_m, chi_r, d;  /* d: qudit dimension 2..16 (core.py:1013-1028
                          accepts any d matching the MPS), or 1 with gate
                          = NULL for a plain split of the chi_l x chi_r
                          matrix a                                        */
    int32_t k;
    int32_t flags;     /* bit0: absorb S into the left factor            */
    mpsx_plan plan;    /* library-owned, see mpsx_two_site_plan          */
} mpsx_two_site_op;

typedef struct {
    int32_t block;      /* Jacobi block width b (pivot is 2b x 2b); 0 = auto  */
    int32_t max_sweeps; /* outer sweep limit; 0 = default (30)              */
    int32_t check_every;/* host convergence poll period in sweeps; 0 = never
                           poll (graph-capture safe: runs max_sweeps with
                           device-side early exit)                         */
    int32_t inner_sweeps;/* pivot eigensolver sweep limit; 0 = default     */
    double tol;         /* relative orthogonality tolerance; 0 = default   */
    int32_t block_switch;/* with block = 0: problems with more than this many
                           rows use b = 16, the others b = 8; 0 = default 64 */
    int32_t rotation_fp32;/* 0: FP64 working matrix (default; tol default
                           1e-7); 1: FP32 working matrix (tol default 1e-5) */
    int32_t *unconverged;/* device int or NULL: incremented once per problem
                           whose last launched sweep still rotated (the
                           device-side status when check_every = 0); the
                           caller zeroes it                                 */
    int32_t method;     /* MPSX_SVD_AUTO (0): the Gram method for FP64
                           problems with b = 16 and nb >= 6 (more than 64
                           rows), rows otherwise;
                           MPSX_SVD_ROWS (1): the one-sided method for all */
    int32_t fp32_sweeps;/* > 0: the Gram method starts with an FP32 phase of
                           at most this many sweeps: its rounds keep H in
                           FP32 on the FP32 matrix cores (rotations, Y and
                           the convergence tests in FP64) until a sweep
                           rotates nothing, then H is formed again in FP64
                           from the rotated rows and the FP64 rounds finish
                           (max_sweeps bounds those). 0 (default) or < 0:
                           FP64 rounds only. The results agree with the FP64
                           rounds' to FP64 rounding of the converged rows. */
    int32_t precondition;/* Householder QR before the FP64 Gram method (no
                           counterpart in the reference, whose SVD is
                           LAPACK's): X P = Q R with the columns ordered by
                           decreasing norm, then the same Jacobi on
                           R R^H = Q^H X X^H Q starting from Y = Q^H. Fewer
                           sweeps where the spectrum of theta decays; more
                           where it is flat.
                           0 (default): none;
                           1: two-site problems whose middle bond is at least
                           the smaller outer bond (chi_m >= min(chi_l,
                           chi_r): theta built from an already truncated
                           bond);
                           2: every Gram problem (plain splits included).
                           Only problems of at most 1024 padded rows and
                           4096 columns; the others keep the plain Gram
                           method (plan.method 3 marks the preconditioned) */
} mpsx_svd_opts;

#define MPSX_SVD_AUTO 0
#define MPSX_SVD_ROWS 1

/* Validates the batch, fills every h[i].plan and returns the workspace bytes
 * mpsx_two_site needs. Call it before copying h to the device. */
int mpsx_two_site_plan(mpsx_two_site_op *h, int count,
                       const mpsx_svd_opts *opts, size_t *bytes);

/* Batched two-site update. Problems are independent (disjoint site pairs,
 * possibly of different circuits). `sweeps_out` (host, may be NULL)
 * receives the number of outer sweeps launched. */
int mpsx_two_site(mpsx_stream stream, const mpsx_two_site_op *h,
                  const mpsx_two_site_op *d, int count,
                  const mpsx_svd_opts *opts, void *ws, size_t ws_bytes,
                  int *sweeps_out);

/* ------------------------------------------------------------------------
 * mpsx_two_site on a fixed, library-owned device workspace.
 * No counterpart in the reference (which allocates through numpy on every
 * gate). For callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc. It
 * takes the same planned tables as mpsx_two_site but no workspace: the
 * library keeps one MPSX_STATIC_WS_BYTES (64 MiB) array per device for the
 * life of the process, looked up on the first call on a device (the
 * stream's device must be current). A batch whose planned workspace
 * (mpsx_two_site_plan's *bytes) exceeds it is rejected with MPSX_EWORKSPACE
 * before anything is enqueued; 64 MiB holds, e.g., five updates at
 * chi = 256 (theta 512 x 512) or one at chi = 512 (d = 2).
 * The call allocates nothing (no hipMalloc, hipHostMalloc, event or
 * library allocation) and never reads device memory back: it does not poll
 * convergence, it launches opts->max_sweeps sweeps (default 30) in which
 * converged problems exit at once, so the results are bitwise those of
 * mpsx_two_site; non-convergence is reported only through
 * opts->unconverged. All calls on one device share its array: calls that
 * overlap on one device (different streams or threads) race and corrupt
 * each other's results; serialise them.
 * ---------------------------------------------------------------------- */
#define MPSX_STATIC_WS_BYTES ((size_t)64 << 20)

int mpsx_two_site_static(mpsx_stream stream, const mpsx_two_site_op *h,
                         const mpsx_two_site_op *d, int count,
                         const mpsx_svd_opts *opts, int *sweeps_out);

/* ------------------------------------------------------------------------
 * Grouped complex GEMM used by the read-out contractions:
 *   C = alpha * op(A) * op(B) (+ C if accumulate),  op = identity, or
 *   conjugate-transpose when the corresponding flag bit is set.
 * Replaces the tn.contract_between chains of MPS.wavefunction
 * (core.py:483-500) and MPS.inner_product (core.py:507-561).
 * ---------------------------------------------------------------------- */
typedef struct {
    const void *A;
    const void *B;
    void *C;
    int64_t m, n, k;
    int64_t lda, ldb, ldc; /* row strides (elements) of the stored matrices */
    float alpha_re, alpha_im;
    int32_t flags;         /* bit0: op(A)=A^H, bit1: op(B)=B^H, bit2: C += */
    int32_t reserved;
} mpsx_gemm_op;

int mpsx_gemm(mpsx_stream stream, const mpsx_gemm_op *h, const mpsx_gemm_op *d,
              int count);

/* ------------------------------------------------------------------------
 * Scale every element of each buffer by a complex factor, optionally
 * conjugating. Replaces the tensor rescaling of MPS.renormalize
 * (core.py:591-594) and MPS.dagger (core.py:502-505).
 * ---------------------------------------------------------------------- */
typedef struct {
    void *x;
    int64_t n;          /* complex elements */
    float re, im;
    int32_t conj;
    int32_t reserved;
} mpsx_scale_op;

int mpsx_scale(mpsx_stream stream, const mpsx_scale_op *h,
               const mpsx_scale_op *d, int count);

/* ------------------------------------------------------------------------
 * Kernel timing. When enabled, every kernel the library launches is
 * bracketed by HIP events recorded on the launch stream. mpsx_profile_read
 * synchronises those events and returns, per kernel name, the summed
 * duration, the busy time (length of the union of its launch intervals,
 * smaller than the sum when launches on different streams overlap) and the
 * launch count since the last reset. No counterpart in the reference (which
 * has no timing hooks); used by the benchmark.
 * ---------------------------------------------------------------------- */
int mpsx_profile_enable(int on);
/* Writes up to `cap` entries; returns the number of distinct kernels. */
int mpsx_profile_read(char *names, int name_len, double *total_ms, double *busy_ms,
                      int64_t *launches, int cap, int reset);
/* Busy time (ms) of a group of kernels since the last reset: the length of
 * the union of the launch intervals of every kernel named in `names`
 * (comma-separated), e.g. the two launches of a Jacobi round. */
int mpsx_profile_busy(const char *names, double *busy_ms);

/* ------------------------------------------------------------------------
 * Graph capture of library calls. Every call above launches kernels only
 * (no allocation, no host synchronisation) when its stream is being
 * captured, with check_every = 0 for mpsx_two_site, so a fixed schedule of
 * calls on preallocated buffers can be recorded once and replayed as one
 * hipGraph. The descriptor tables and buffers the calls name are read at
 * replay time and must stay alive and unchanged (their contents may be
 * updated between replays). No counterpart in the reference, which applies
 * gates one by one from Python (core.py:1221-1247).
 * ---------------------------------------------------------------------- */
typedef struct mpsx_graph_s *mpsx_graph;

/* Starts capturing `stream` (not the null stream) in thread-local mode. */
int mpsx_graph_begin(mpsx_stream stream);
/* Ends the capture and instantiates it into *out. */
int mpsx_graph_end(mpsx_stream stream, mpsx_graph *out);
/* Replays the graph on `stream`. */
int mpsx_graph_launch(mpsx_graph graph, mpsx_stream stream);
int mpsx_graph_destroy(mpsx_graph graph);

#ifdef __cplusplus
}
#endif

#endif /* MPSX_H */
