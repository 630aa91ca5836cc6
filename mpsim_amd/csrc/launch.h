// Host-side launchers shared between the kernel files and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mpsx.h"

hipError_t launch_one_site(hipStream_t s, const mpsx_site_op *d, int count, int64_t max_pairs,
                           uint32_t dmask);
hipError_t launch_scale(hipStream_t s, const mpsx_scale_op *d, int count, int64_t max_n);
hipError_t launch_gemm(hipStream_t s, const mpsx_gemm_op *d, int count, int64_t max_m,
                       int64_t max_n);
// The tall-narrow variant: every op without conjugation, k <= 32, n <= 256.
bool gemm_rows_fit(const mpsx_gemm_op &o);
hipError_t launch_gemm_rows(hipStream_t s, const mpsx_gemm_op *h, const mpsx_gemm_op *d,
                            int count);

hipError_t launch_init_rows(hipStream_t s, bool f32, const mpsx_two_site_op *d, int count, char *ws,
                            int64_t rows);
hipError_t launch_theta(hipStream_t s, bool f32, int D, const mpsx_two_site_op *d, int count, char *ws,
                        int64_t tiles);
hipError_t launch_theta_gen(hipStream_t s, bool f32, const mpsx_two_site_op *d, int count, char *ws,
                            int64_t tiles);
hipError_t launch_rownorm(hipStream_t s, bool f32, const mpsx_two_site_op *d, int count, char *ws,
                          int64_t rows);
hipError_t launch_fro(hipStream_t s, const mpsx_two_site_op *d, int count, char *ws,
                      double null_rel);
hipError_t launch_owner(hipStream_t s, const mpsx_two_site_op *d, int count, int *owner);
hipError_t launch_jacobi_round(hipStream_t s, bool f32, int b, const mpsx_two_site_op *d, const int *owner,
                               char *ws, int64_t pairs, int round, double tol, double tol_inner,
                               int inner_sweeps);
hipError_t launch_sweep_end(hipStream_t s, const mpsx_two_site_op *d, int count, char *ws,
                            int sweep, bool fp32_phase);
hipError_t launch_sort(hipStream_t s, const mpsx_two_site_op *d, int count, char *ws,
                       int *unconverged);
hipError_t launch_gather(hipStream_t s, bool f32, const mpsx_two_site_op *d, int count, char *ws,
                         int64_t outs);
hipError_t launch_hgram(hipStream_t s, const mpsx_two_site_op *d, const int *owner, char *ws,
                        int64_t pairs, int max_nb, int stage);
hipError_t launch_hsolve(hipStream_t s, bool f32h, const mpsx_two_site_op *d, const int *owner, char *ws,
                         int64_t pairs, int round, double tol, double tol_inner, int inner_sweeps,
                         double abs_rel);
hipError_t launch_hupdate(hipStream_t s, bool f32h, const mpsx_two_site_op *d, const int *owner, char *ws,
                          int64_t pairs, int ytiles, int round, int part);
hipError_t launch_hround(hipStream_t s, bool f32h, const mpsx_two_site_op *d, const int *owner, char *ws,
                         int64_t pairs, int ytiles, int round, double tol, double tol_inner,
                         int inner_sweeps, double abs_rel);
hipError_t launch_hxcopy(hipStream_t s, const mpsx_two_site_op *d, int count, char *ws, int64_t rows);
hipError_t launch_hxgemm(hipStream_t s, const mpsx_two_site_op *d, const int *owner, char *ws,
                         int64_t pairs, int ctiles, int mode, int max_nb);
// QR preconditioning (plan.method 3): column order, X0 save and permute,
// and step `panel` (-1 factors panel 0) over `strips` 16-column strips, for
// the problems of at most 512 rows (`small`) and those above (`large`).
hipError_t launch_qr_order(hipStream_t s, const mpsx_two_site_op *d, int count, char *ws, int max_c,
                           int max_r);
hipError_t launch_qr_permute(hipStream_t s, const mpsx_two_site_op *d, int count, char *ws,
                             int64_t rows);
hipError_t launch_qr_step(hipStream_t s, const mpsx_two_site_op *d, int count, char *ws, int panel,
                          int strips, int max_c, bool small, bool large);
