"""CPU oracle for the FFT hot path -- test infrastructure, not product code.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker.  The shipped library (ascend-fft_amd/)
never imports, links or calls anything here.

`lib()` is our C restatement (oracle/bfly_oracle.c, each function citing the
reference file:line it follows).  `ref()` is the reference's own host code
compiled from its sources into oracle/_ref/libbfly_ref.so (None when that
build is absent, e.g. on a machine without the reference tree).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_PATH = HERE / "_ref" / "libbfly_ref.so"
REFERENCE_TREE = Path("/root/reference")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")

_lib = None
_ref = None

NUMERICAL_MODES = {
    "zero": 0, "impulse": 1, "constant": 2, "tone": 3, "alternating": 4,
    "random-seeded": 5, "high-dynamic-range": 6, "near-cancellation": 7,
}
APP_MODES = {"ofdm": 0, "radar": 1, "dl": 2}


def build(with_ref: bool = True) -> None:
    """Compile the C oracle (and, when the reference tree exists, oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)
    if with_ref and REFERENCE_TREE.exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


def _bind(lib):
    def fn(name, res, *args):
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = list(args)
    u32, i32, sz = C.c_uint32, C.c_int, C.c_size_t
    fn("orc_fft_f64_inplace", None, _f64p, u32, i32)
    fn("orc_c2c_f64", None, _f32p, _f64p, u32, u32, i32)
    fn("orc_fft2_f64", None, _f32p, _f64p, u32, u32, u32, i32)
    fn("orc_ref_fft_f32", None, _f32p, _f32p, u32)
    fn("orc_dft_direct_f64", None, _f32p, _f64p, u32, i32)
    fn("orc_ref_r2c_f32", None, _f32p, _f32p, u32)
    fn("orc_r2c_f64", None, _f32p, _f64p, u32, u32, u32, u32)
    fn("orc_ref_c2r_f32", None, _f32p, _f32p, u32)
    fn("orc_c2r_f64", None, _f32p, _f64p, u32, u32, u32, u32)
    fn("orc_max_rel_scaled", C.c_double, _f32p, _f32p, sz)
    fn("orc_rel_l2", C.c_double, _f32p, _f64p, sz)
    fn("orc_rel_l2_rows", C.c_double, _f32p, _f64p, u32, sz, sz, sz)
    fn("orc_pattern_at", C.c_float, sz)
    fn("orc_fill_pattern", None, _f32p, sz)
    fn("orc_sample_batches", u32, u32, u32, _u32p, u32)
    fn("orc_app_hash", u32, u32, u32)
    fn("orc_gen_numerical", None, _f32p, u32, u32, i32, u32)
    fn("orc_gen_app", None, _f32p, u32, u32, i32)
    fn("orc_gen_sin", None, _f32p, sz)
    fn("orc_gen_half", None, _f32p, u32, u32, u32)
    fn("orc_plane_k_for", u32, u32)
    fn("orc_fold_d_for", u32, u32, u32, u32)
    fn("orc_row_fft_ub_bytes", C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64)
    fn("orc_resolve_row_fft", i32, u32, u32, u32, u32, C.POINTER(u32), C.POINTER(u32))
    fn("orc_gen_twiddles", None, u32, _f32p, _f32p)
    fn("orc_gen_bit_reverse", None, u32, _u32p)
    fn("orc_gen_bit_reverse_plane", None, u32, _u32p)
    fn("orc_gen_transpose_plane", None, u32, _u32p)
    fn("orc_gen_interleave", None, u32, u32, _u32p)
    fn("orc_gen_r2c_twiddles", u32, u32, _f32p, _f32p)
    fn("orc_long_split", i32, u32, C.POINTER(u32), C.POINTER(u32))
    fn("orc_four_step_f64", None, _f32p, _f64p, u32, u32, u32, u32)
    return lib


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build(with_ref=False)
        _lib = _bind(C.CDLL(str(LIB_PATH)))
    return _lib


class RefMapping(C.Structure):
    _fields_ = [("n", C.c_uint32), ("batch", C.c_uint32), ("precision", C.c_int),
                ("direction", C.c_int), ("n1", C.c_uint32), ("n2", C.c_uint32),
                ("stage_partition", C.c_int), ("stage_space", C.c_int),
                ("stage_time", C.c_int), ("data_space", C.c_int), ("data_time", C.c_int),
                ("role0", C.c_int), ("role1", C.c_int), ("residence", C.c_int),
                ("boundary_layout", C.c_int), ("boundary_home", C.c_int),
                ("boundary_impl", C.c_int), ("pipeline_buffers", C.c_int),
                ("fold_d", C.c_uint32), ("plane_k", C.c_uint32),
                ("unit_ub_bytes", C.c_uint64)]


class RefLowering(C.Structure):
    _fields_ = [("supported", C.c_int), ("abstract_feasible", C.c_int),
                ("launches", C.c_int), ("gm_boundaries", C.c_int),
                ("host_assisted", C.c_int), ("on_chip", C.c_int),
                ("plan_ub_bytes", C.c_uint64), ("modeled_payload_gm", C.c_uint64),
                ("kinds", C.c_char * 256), ("reason", C.c_char * 512)]


def ref():
    """The reference's own code (oracle/_ref), or None when it was not built."""
    global _ref
    if _ref is None:
        if not REF_PATH.exists():
            if REFERENCE_TREE.exists():
                build(with_ref=True)
            if not REF_PATH.exists():
                return None
        r = C.CDLL(str(REF_PATH))

        def fn(name, res, *args):
            f = getattr(r, name)
            f.restype = res
            f.argtypes = list(args)
        u32 = C.c_uint32
        fn("ref_fft_f32", None, _f32p, _f32p, u32)
        fn("ref_r2c_f32", None, _f32p, _f32p, u32)
        fn("ref_c2r_f32", None, _f32p, _f32p, u32)
        fn("ref_max_rel_scaled", C.c_double, _f32p, _f32p, C.c_size_t)
        fn("ref_pattern_at", C.c_float, C.c_size_t)
        fn("ref_sample_batches", u32, u32, u32, _u32p, u32)
        fn("ref_plane_k_for", u32, u32)
        fn("ref_fold_d_for", u32, u32, u32, u32)
        fn("ref_row_fft_ub_bytes", C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64)
        fn("ref_resolve_row_fft", C.c_int, u32, u32, u32, u32, C.POINTER(u32), C.POINTER(u32))
        fn("ref_default_mapping", None, u32, u32, C.POINTER(RefMapping))
        fn("ref_query_lowering", None, C.POINTER(RefMapping), C.c_char_p, C.c_char_p,
           C.c_char_p, C.POINTER(RefLowering))
        _ref = r
    return _ref


# --------------------------------------------------------------- numpy API
def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def c2c_f64(x: np.ndarray, n: int, sign: int = -1) -> np.ndarray:
    """FP64 oracle C2C of interleaved fp32 rows [batch][2n] (sign -1 forward)."""
    x = _f32(x).reshape(-1, 2 * n)
    out = np.empty(x.shape, np.float64)
    lib().orc_c2c_f64(x, out, n, x.shape[0], sign)
    return out


def fft_f64(z: np.ndarray, n: int, sign: int = -1) -> np.ndarray:
    """FP64 oracle FFT of FP64 interleaved rows [rows][2n] (a copy)."""
    a = np.array(z, dtype=np.float64, copy=True).reshape(-1, 2 * n)
    for r in range(a.shape[0]):
        row = np.ascontiguousarray(a[r])
        lib().orc_fft_f64_inplace(row, n, sign)
        a[r] = row
    return a


def fft2_f64(x: np.ndarray, n1: int, n2: int, sign: int = -1) -> np.ndarray:
    """FP64 oracle 2-D C2C of interleaved fp32 arrays [batch][n1][2 n2]."""
    x = _f32(x).reshape(-1, n1 * 2 * n2)
    out = np.empty(x.shape, np.float64)
    lib().orc_fft2_f64(x, out, n1, n2, x.shape[0], sign)
    return out


def ref_fft_f32(x: np.ndarray, n: int) -> np.ndarray:
    x = _f32(x).reshape(-1, 2 * n)
    out = np.empty_like(x)
    for b in range(x.shape[0]):
        row = np.empty(2 * n, np.float32)
        lib().orc_ref_fft_f32(np.ascontiguousarray(x[b]), row, n)
        out[b] = row
    return out


def r2c_f64(x: np.ndarray, n: int) -> np.ndarray:
    x = _f32(x).reshape(-1, n)
    out = np.empty((x.shape[0], n + 2), np.float64)
    lib().orc_r2c_f64(x, out, n, x.shape[0], n, n + 2)
    return out


def c2r_f64(h: np.ndarray, n: int) -> np.ndarray:
    h = _f32(h).reshape(-1, n + 2)
    out = np.empty((h.shape[0], n), np.float64)
    lib().orc_c2r_f64(h, out, n, h.shape[0], n + 2, n)
    return out


def four_step_f64(x: np.ndarray, n: int, n1: int, n2: int) -> np.ndarray:
    x = _f32(x).reshape(-1, 2 * n)
    out = np.empty(x.shape, np.float64)
    lib().orc_four_step_f64(x, out, n, x.shape[0], n1, n2)
    return out


def rel_l2_rows(got: np.ndarray, ref: np.ndarray) -> float:
    """Worst per-row relative L2 error of fp32 rows against FP64 rows."""
    got = _f32(got).reshape(ref.shape[0], -1)
    ref = np.ascontiguousarray(ref, dtype=np.float64)
    L = ref.shape[1]
    return lib().orc_rel_l2_rows(got, ref.reshape(-1), ref.shape[0], L, got.shape[1], L)


def max_rel_scaled(got: np.ndarray, ref: np.ndarray) -> float:
    got, ref = _f32(got).reshape(-1), _f32(ref).reshape(-1)
    return lib().orc_max_rel_scaled(got, ref, got.size)


def uniform_input(seed: int, shape) -> np.ndarray:
    """Seeded U[-1,1) input: numpy default_rng(seed), drawn in FP64, cast to fp32."""
    return np.random.default_rng(seed).uniform(-1.0, 1.0, size=shape).astype(np.float32)


def numerical_input(n: int, batch: int, mode: str, seed: int = 0) -> np.ndarray:
    out = np.empty((batch, 2 * n), np.float32)
    lib().orc_gen_numerical(out, n, batch, NUMERICAL_MODES[mode], seed)
    return out


def app_input(n: int, batch: int, mode: str) -> np.ndarray:
    out = np.empty((batch, 2 * n), np.float32)
    lib().orc_gen_app(out, n, batch, APP_MODES[mode])
    return out


def sin_input(count: int) -> np.ndarray:
    """The `sin` / `deterministic-sin-cos` sequence over a flat float index."""
    out = np.empty(count, np.float32)
    lib().orc_gen_sin(out, count)
    return out


def half_input(n: int, batch: int) -> np.ndarray:
    out = np.zeros((batch, n + 2), np.float32)
    lib().orc_gen_half(out, n, batch, n + 2)
    return out


def pattern(count: int) -> np.ndarray:
    out = np.empty(count, np.float32)
    lib().orc_fill_pattern(out, count)
    return out


def long_split(n: int):
    a, b = C.c_uint32(0), C.c_uint32(0)
    ok = lib().orc_long_split(n, C.byref(a), C.byref(b))
    return (a.value, b.value) if ok else None


def cpu_threads() -> int:
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
