// Library-internal entry points shared between the source files of libmpsx
// (not part of the C ABI, hidden from the exports).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "../../include/mpsx.h"

#define MPSX_INTERNAL __attribute__((visibility("hidden")))

// The two-site driver of mpsx_two_site (argument checks, plan check, launch
// sequence). may_poll false: no convergence poll (no device-to-host read)
// and no profiling events. need_out (may be NULL) receives the workspace
// bytes the batch needs once the table has passed its checks.
MPSX_INTERNAL int mpsx_internal_two_site(hipStream_t stream, const mpsx_two_site_op *h,
                                         const mpsx_two_site_op *d, int count,
                                         const mpsx_svd_opts *opts, char *ws, size_t ws_bytes,
                                         int *sweeps_out, bool may_poll, size_t *need_out);

// Records `msg` as this thread's mpsx_last_error() and returns `code`.
MPSX_INTERNAL int mpsx_internal_fail(int code, const char *msg);
