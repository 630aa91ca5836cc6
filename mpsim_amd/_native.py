"""ctypes binding of libmpsx.so (include/mpsx.h).

Descriptor tables are built as numpy structured arrays whose layout matches
the C structs byte for byte (checked against mpsx_sizeof at load time). Each
batched call passes the host table and a device copy of it.

There is no fallback: if the library or a GPU is missing, the calls raise.
"""
from __future__ import annotations

import ctypes
import os
import threading
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MPSX_LIB", os.path.join(HERE, "libmpsx.so"))

MPSX_OK, MPSX_EINVAL, MPSX_EHIP, MPSX_EWORKSPACE, MPSX_ENOTCONV = 0, 1, 2, 3, 4
MPSX_STATIC_WS_BYTES = 64 << 20  # fixed workspace of mpsx_two_site_static

SITE_OP = np.dtype([("a", "<u8"), ("out", "<u8"), ("gate", "<u8"), ("chi_l", "<i8"),
                    ("chi_r", "<i8"), ("d", "<i4"), ("reserved", "<i4")])
PLAN_FIELDS = [("ws_off", "<i8"), ("tile_base", "<i8"), ("pair_base", "<i8"),
               ("row_base", "<i8"), ("out_base", "<i8"), ("C", "<i4"), ("R", "<i4"),
               ("Cp", "<i4"), ("LD", "<i4"), ("nb", "<i4"), ("bw", "<i4"), ("zbytes", "<i4"),
               ("method", "<i4")]
PLAN = np.dtype(PLAN_FIELDS)
TWO_SITE_OP = np.dtype([("a", "<u8"), ("b", "<u8"), ("gate", "<u8"), ("a_out", "<u8"),
                        ("b_out", "<u8"), ("s_out", "<u8"), ("discarded", "<u8"),
                        ("chi_l", "<i4"), ("chi_m", "<i4"), ("chi_r", "<i4"), ("d", "<i4"),
                        ("k", "<i4"), ("flags", "<i4")] + PLAN_FIELDS)
GEMM_OP = np.dtype([("A", "<u8"), ("B", "<u8"), ("C", "<u8"), ("m", "<i8"), ("n", "<i8"),
                    ("k", "<i8"), ("lda", "<i8"), ("ldb", "<i8"), ("ldc", "<i8"),
                    ("alpha_re", "<f4"), ("alpha_im", "<f4"), ("flags", "<i4"),
                    ("reserved", "<i4")])
SCALE_OP = np.dtype([("x", "<u8"), ("n", "<i8"), ("re", "<f4"), ("im", "<f4"),
                     ("conj", "<i4"), ("reserved", "<i4")])

GEMM_CONJ_A, GEMM_CONJ_B, GEMM_ACCUM = 1, 2, 4


class SvdOpts(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int32), ("max_sweeps", ctypes.c_int32),
                ("check_every", ctypes.c_int32), ("inner_sweeps", ctypes.c_int32),
                ("tol", ctypes.c_double), ("block_switch", ctypes.c_int32),
                ("rotation_fp32", ctypes.c_int32), ("unconverged", ctypes.c_void_p),
                ("method", ctypes.c_int32), ("fp32_sweeps", ctypes.c_int32),
                ("precondition", ctypes.c_int32)]


class MpsxError(RuntimeError):
    """A libmpsx call failed; `code` is the MPSX_E* value."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"libmpsx error {code}: {msg}")
        self.code = code


class NotConvergedWarning(RuntimeWarning):
    pass


EXPORTS = ("mpsx_version", "mpsx_last_error", "mpsx_device_info", "mpsx_sizeof",
           "mpsx_apply_one_site", "mpsx_two_site_plan", "mpsx_two_site", "mpsx_two_site_static",
           "mpsx_gemm",
           "mpsx_scale", "mpsx_profile_enable", "mpsx_profile_read", "mpsx_profile_busy",
           "mpsx_graph_begin", "mpsx_graph_end", "mpsx_graph_launch", "mpsx_graph_destroy")

_lock = threading.Lock()
_lib: Optional[ctypes.CDLL] = None


def load(path: Optional[str] = None) -> ctypes.CDLL:
    """Loads libmpsx.so (once) and checks the struct layouts."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise MpsxError(-1, f"{p} not found; build it with `python -m mpsim_amd.build`")
        lib = ctypes.CDLL(p)
        vp, i32, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
        lib.mpsx_version.restype = i32
        lib.mpsx_last_error.restype = ctypes.c_char_p
        lib.mpsx_device_info.argtypes = [i32, ctypes.c_char_p, i32, ctypes.POINTER(i32)]
        lib.mpsx_sizeof.argtypes = [i32]
        lib.mpsx_sizeof.restype = sz
        lib.mpsx_apply_one_site.argtypes = [vp, vp, vp, i32]
        lib.mpsx_two_site_plan.argtypes = [vp, i32, ctypes.POINTER(SvdOpts), ctypes.POINTER(sz)]
        lib.mpsx_two_site.argtypes = [vp, vp, vp, i32, ctypes.POINTER(SvdOpts), vp, sz,
                                      ctypes.POINTER(i32)]
        lib.mpsx_two_site_static.argtypes = [vp, vp, vp, i32, ctypes.POINTER(SvdOpts),
                                             ctypes.POINTER(i32)]
        lib.mpsx_gemm.argtypes = [vp, vp, vp, i32]
        lib.mpsx_scale.argtypes = [vp, vp, vp, i32]
        lib.mpsx_profile_enable.argtypes = [i32]
        lib.mpsx_profile_read.argtypes = [ctypes.c_char_p, i32, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_int64), i32, i32]
        lib.mpsx_profile_busy.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)]
        lib.mpsx_graph_begin.argtypes = [vp]
        lib.mpsx_graph_end.argtypes = [vp, ctypes.POINTER(vp)]
        lib.mpsx_graph_launch.argtypes = [vp, vp]
        lib.mpsx_graph_destroy.argtypes = [vp]
        for name in ("mpsx_apply_one_site", "mpsx_two_site_plan", "mpsx_two_site",
                     "mpsx_two_site_static", "mpsx_gemm",
                     "mpsx_scale", "mpsx_device_info", "mpsx_profile_enable",
                     "mpsx_profile_read", "mpsx_profile_busy", "mpsx_graph_begin", "mpsx_graph_end",
                     "mpsx_graph_launch", "mpsx_graph_destroy"):
            getattr(lib, name).restype = i32
        sizes = [lib.mpsx_sizeof(i) for i in range(6)]
        want = [SITE_OP.itemsize, TWO_SITE_OP.itemsize, GEMM_OP.itemsize, SCALE_OP.itemsize,
                ctypes.sizeof(SvdOpts), PLAN.itemsize]
        if sizes != want:
            raise MpsxError(-1, f"ABI layout mismatch: library {sizes} vs binding {want}")
        if path is None:
            _lib = lib
        return lib


def check(code: int) -> None:
    if code != MPSX_OK:
        msg = load().mpsx_last_error().decode(errors="replace")
        raise MpsxError(code, msg)


# --------------------------------------------------------------------- torch
def _torch():
    import torch  # imported lazily: host-only tools do not need it
    return torch


def device(dev=None):
    torch = _torch()
    if not torch.cuda.is_available():
        raise MpsxError(-1, "no HIP device is visible (torch.cuda.is_available() is False)")
    if dev is None:
        return torch.device("cuda", torch.cuda.current_device())
    return torch.device(dev)


def stream_handle(dev) -> int:
    torch = _torch()
    return int(torch.cuda.current_stream(dev).cuda_stream)


def to_device_table(table: np.ndarray, dev):
    """Device copy of a descriptor table (kept alive by the caller)."""
    torch = _torch()
    raw = torch.from_numpy(np.ascontiguousarray(table).view(np.uint8))
    return raw.to(dev, non_blocking=False)


def ptr(t) -> int:
    """Device address of a contiguous tensor (0 for None or empty)."""
    if t is None or t.numel() == 0:
        return 0
    if not t.is_contiguous():
        raise ValueError("libmpsx needs contiguous tensors")
    return int(t.data_ptr())


# --------------------------------------------------------------------- timing
def profile_enable(on: bool) -> None:
    check(load().mpsx_profile_enable(1 if on else 0))


def profile_read_full(reset: bool = True) -> dict:
    """{kernel name: {"total_ms", "busy_ms", "launches"}} since the last
    reset; synchronises the recorded HIP events. busy_ms is the union of the
    kernel's launch intervals (launches on two streams may overlap)."""
    lib = load()
    cap, nlen = 64, 64
    names = ctypes.create_string_buffer(cap * nlen)
    tot = (ctypes.c_double * cap)()
    busy = (ctypes.c_double * cap)()
    cnt = (ctypes.c_int64 * cap)()
    n = lib.mpsx_profile_read(names, nlen, tot, busy, cnt, cap, 1 if reset else 0)
    if n < 0:
        check(n)
    out = {}
    for i in range(min(n, cap)):
        nm = names.raw[i * nlen:(i + 1) * nlen].split(b"\0", 1)[0].decode()
        out[nm] = {"total_ms": float(tot[i]), "busy_ms": float(busy[i]), "launches": int(cnt[i])}
    return out


def profile_busy(names) -> float:
    """Busy time (ms) of a group of kernels (union of all their launch
    intervals) since the last reset; does not reset."""
    lib = load()
    out = ctypes.c_double(0.0)
    check(lib.mpsx_profile_busy(",".join(names).encode(), ctypes.byref(out)))
    return float(out.value)


def profile_read(reset: bool = True) -> dict:
    """{kernel name: (total ms, launches)} since the last reset."""
    return {k: (v["total_ms"], v["launches"]) for k, v in profile_read_full(reset).items()}
