"""ctypes binding of libbfly_hip.so (the C ABI declared in include/bfly_hip.h).

The library is loaded from ascend-fft_amd/lib/ in this tree.  There is no
fallback: when the shared object is missing or fails to load, every entry
point raises, so nothing can silently run on another path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_ROOT = Path(__file__).resolve().parents[1]          # ascend-fft_amd/
LIB_DIR = PKG_ROOT / "lib"
HIP_LIB = LIB_DIR / "libbfly_hip.so"
STUB_LIB = LIB_DIR / "libbfly_stub.so"

OK = 0
ERR_NOT_LOADED = -1
ERR_TABLE = -2
ERR_ALLOC = -3
ERR_H2D = -4
ERR_LAUNCH = -5
ERR_SYNC = -6
ERR_D2H = -7
ERR_SHAPE = -8
ERR_SCRATCH = -9
NOALLOC_SCRATCH_BYTES = 64 << 20

VARIANT_AUTO = 0
VARIANT_PLAIN = 1
VARIANT_STREAM = 2
VARIANT_PLAIN_SWZ = 3
VARIANT_STREAM_SWZ = 4
VARIANT_STREAM_LT = 5
VARIANT_STREAM2 = 6
VARIANT_FOUR_STEP = 7
VARIANT_SPLIT_MFMA = 8
VARIANT_SPLIT_MFMA_STREAM = 9
VARIANT_LAST = 9

REAL_ADJOINT = 1  # bfly_real_ex flag: the transform's adjoint
MIX_FORWARD, MIX_GRAD_INPUT, MIX_GRAD_WEIGHT = 0, 1, 2  # bfly_mode_mix ops

KIND_C2C = 0
KIND_R2C = 1
KIND_C2R = 2

ERRORS = {
    ERR_NOT_LOADED: "kernel not available",
    ERR_TABLE: "table preparation failed",
    ERR_ALLOC: "device allocation failed",
    ERR_H2D: "host-to-device copy failed",
    ERR_LAUNCH: "kernel launch failed",
    ERR_SYNC: "stream synchronisation failed",
    ERR_D2H: "device-to-host copy failed",
    ERR_SHAPE: "shape / pointer / size guard",
    ERR_SCRATCH: "intermediate exceeds the static scratch of bfly_c2c_noalloc",
}

# Every symbol include/bfly_hip.h declares (tests check the export table).
SYMBOLS = (
    "bfly_rt_init", "bfly_rt_teardown", "bfly_rt_stream_create", "bfly_rt_stream_destroy", "bfly_rt_stream_sync",
    "bfly_rt_malloc", "bfly_rt_mem_info", "bfly_rt_free", "bfly_rt_memcpy_h2d", "bfly_rt_memcpy_d2h",
    "bfly_rt_memcpy_h2d_async", "bfly_rt_memcpy_d2h_async", "bfly_rt_malloc_host", "bfly_rt_free_host",
    "bfly_rt_event_create", "bfly_rt_event_destroy", "bfly_rt_event_record", "bfly_rt_event_elapsed",
    "bfly_graph_begin", "bfly_graph_end", "bfly_graph_launch", "bfly_graph_destroy",
    "bfly_c2c_launches", "bfly_c2c_traced", "bfly_c2c_columns_workspace_bytes", "bfly_c2c_columns", "bfly_c2c_rows",
    "bfly_fft2d_table_floats", "bfly_fft2d_table_fill",
    "bfly_fft2d_workspace_bytes", "bfly_fft2d", "bfly_rt_time_transform", "bfly_supported", "bfly_table_floats", "bfly_table_fill",
    "bfly_workspace_bytes", "bfly_c2c", "bfly_c2c_fwd", "bfly_c2c_inv", "bfly_c2c_scaled", "bfly_c2c_ex", "bfly_r2c", "bfly_c2r", "bfly_r2c_ex",
    "bfly_c2r_ex", "bfly_real_ex", "bfly_group_supported", "bfly_group_create", "bfly_group_launch",
    "bfly_group_query", "bfly_group_destroy", "bfly_mode_mix", "bfly_probe_copy", "bfly_probe_device",
    "bfly_noalloc_prepare", "bfly_c2c_noalloc", "bfly_host_pipelined",
)


class GroupRequest(C.Structure):
    """bfly_group_request of include/bfly_hip.h."""
    _fields_ = [("in_", C.c_void_p), ("out", C.c_void_p), ("kind", C.c_int32),
                ("sign", C.c_int32), ("n", C.c_uint32), ("batch", C.c_uint32)]


class GroupConfig(C.Structure):
    """bfly_group_config of include/bfly_hip.h."""
    _fields_ = [("workgroup", C.c_int32), ("schedule", C.c_int32), ("order", C.c_int32),
                ("dynamic", C.c_int32), ("grid", C.c_int32), ("twiddles", C.c_int32)]


class BflyError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: status {code} ({ERRORS.get(code, 'unknown')})")
        self.code = code


def _bind(lib):
    vp, u32, i32, sz = C.c_void_p, C.c_uint32, C.c_int, C.c_size_t
    sig = {
        "bfly_rt_init": (i32, [i32]),
        "bfly_rt_stream_create": (i32, [C.POINTER(vp)]),
        "bfly_rt_stream_destroy": (i32, [vp]),
        "bfly_rt_stream_sync": (i32, [vp]),
        "bfly_rt_malloc": (i32, [C.POINTER(vp), sz]),
        "bfly_rt_mem_info": (i32, [C.POINTER(sz), C.POINTER(sz)]),
        "bfly_rt_free": (i32, [vp]),
        "bfly_rt_memcpy_h2d": (i32, [vp, vp, sz, vp]),
        "bfly_rt_memcpy_d2h": (i32, [vp, vp, sz, vp]),
        "bfly_rt_memcpy_h2d_async": (i32, [vp, vp, sz, vp]),
        "bfly_rt_memcpy_d2h_async": (i32, [vp, vp, sz, vp]),
        "bfly_rt_malloc_host": (i32, [C.POINTER(vp), sz]),
        "bfly_rt_free_host": (i32, [vp]),
        "bfly_rt_event_create": (i32, [C.POINTER(vp)]),
        "bfly_rt_event_destroy": (i32, [vp]),
        "bfly_rt_event_record": (i32, [vp, vp]),
        "bfly_rt_event_elapsed": (i32, [vp, vp, C.POINTER(C.c_float)]),
        "bfly_graph_begin": (i32, [vp]),
        "bfly_graph_end": (i32, [vp, C.POINTER(vp)]),
        "bfly_graph_launch": (i32, [vp, vp]),
        "bfly_graph_destroy": (i32, [vp]),
        "bfly_c2c_launches": (i32, [u32, u32, i32]),
        "bfly_c2c_traced": (i32, [vp, vp, u32, u32, i32, i32, vp, vp, vp, C.POINTER(vp), i32]),
        "bfly_c2c_columns_workspace_bytes": (sz, [u32, u32, u32]),
        "bfly_c2c_columns": (i32, [vp, vp, u32, u32, u32, i32, vp, vp, vp]),
        "bfly_fft2d_table_floats": (sz, [u32]),
        "bfly_fft2d_table_fill": (i32, [u32, vp]),
        "bfly_fft2d_workspace_bytes": (sz, [u32, u32, u32]),
        "bfly_fft2d": (i32, [vp, vp, u32, u32, u32, i32, vp, vp, vp, vp]),
        "bfly_rt_time_transform": (i32, [i32, i32, vp, vp, u32, u32, vp, vp, i32, vp,
                                         C.POINTER(C.c_float)]),
        "bfly_supported": (i32, [i32, u32]),
        "bfly_table_floats": (sz, [i32, u32]),
        "bfly_table_fill": (i32, [i32, u32, vp]),
        "bfly_workspace_bytes": (sz, [i32, u32, u32]),
        "bfly_c2c": (i32, [vp, vp, u32, u32, i32, vp, vp, vp]),
        "bfly_c2c_ex": (i32, [vp, vp, u32, u32, i32, i32, vp, vp, vp]),
        "bfly_c2c_fwd": (i32, [vp, vp, u32, u32, vp, vp, vp]),
        "bfly_c2c_inv": (i32, [vp, vp, u32, u32, vp, vp, vp]),
        "bfly_c2c_scaled": (i32, [vp, vp, u32, u32, i32, C.c_float, vp, vp, vp]),
        "bfly_c2c_noalloc": (i32, [vp, vp, u32, u32, i32, vp, vp]),
        "bfly_host_pipelined": (i32, [i32, i32, i32, vp, vp, u32, u32, vp, vp, vp, vp, u32, vp]),
        "bfly_noalloc_prepare": (i32, []),
        "bfly_c2c_rows": (i32, [vp, C.c_uint64, vp, C.c_uint64, u32, u32, i32, vp, vp, vp]),
        "bfly_rt_teardown": (i32, []),
        "bfly_r2c": (i32, [vp, vp, u32, u32, vp, vp]),
        "bfly_c2r": (i32, [vp, vp, u32, u32, vp, vp]),
        "bfly_r2c_ex": (i32, [vp, vp, u32, u32, u32, u32, vp, vp]),
        "bfly_c2r_ex": (i32, [vp, vp, u32, u32, u32, u32, vp, vp]),
        "bfly_real_ex": (i32, [i32, vp, vp, u32, u32, u32, u32, i32, u32, vp, vp]),
        "bfly_mode_mix": (i32, [i32, vp, vp, vp, u32, u32, u32, u32, vp]),
        "bfly_group_supported": (i32, [i32, u32, i32]),
        "bfly_group_create": (i32, [C.POINTER(GroupRequest), i32, C.POINTER(GroupConfig),
                                    C.POINTER(vp)]),
        "bfly_group_launch": (i32, [vp, vp, vp]),
        "bfly_group_query": (i32, [vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(sz)]),
        "bfly_group_destroy": (i32, [vp]),
        "bfly_probe_copy": (i32, [vp, vp, sz, vp]),
        "bfly_probe_device": (i32, [i32, C.c_char_p, sz]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_hip = None


def hip():
    """The gfx950 library; raises when it is not built or cannot load."""
    global _hip
    if _hip is None:
        if not HIP_LIB.exists():
            raise RuntimeError(f"{HIP_LIB} is not built (run `make -C ascend-fft_amd`)")
        _hip = _bind(C.CDLL(str(HIP_LIB), mode=os.RTLD_LOCAL | getattr(os, "RTLD_NOW", 2)))
    return _hip


def stub():
    """The CPU stub of the same ABI (host-only tests)."""
    if not STUB_LIB.exists():
        raise RuntimeError(f"{STUB_LIB} is not built")
    lib = _bind(C.CDLL(str(STUB_LIB), mode=os.RTLD_LOCAL))
    lib.bfly_stub_reset.restype = None
    lib.bfly_stub_set_failures.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.bfly_stub_set_failures.restype = None
    lib.bfly_stub_counters.argtypes = [C.POINTER(C.c_int)]
    lib.bfly_stub_counters.restype = None
    return lib


def check(code: int, what: str) -> None:
    if code != OK:
        raise BflyError(code, what)
