"""Host-side checks of the C ABI: both libraries load and export every
symbol include/bfly_hip.h declares; the host table functions produce the
FP64 twiddle layout the kernels expect; execute guards through the stub.
No device compute is called here."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import bfly
from bfly import _lib

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "bfly_hip.h"
KINDS = (bfly.KIND_C2C, bfly.KIND_R2C, bfly.KIND_C2R)


def declared():
    return sorted(set(re.findall(r"BFLY_API\s+(?:int|size_t)\s+(bfly_\w+)\s*\(", HEADER.read_text())))


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", str(path)], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_header_lists_the_binding_symbols():
    assert declared() == sorted(_lib.SYMBOLS)


@pytest.mark.parametrize("lib", [_lib.HIP_LIB, _lib.STUB_LIB])
def test_libraries_export_every_declared_symbol(lib):
    assert lib.exists(), f"{lib} not built"
    names = exported(lib)
    missing = [s for s in declared() if s not in names]
    assert not missing, missing
    # nothing but the C ABI (and the stub's controls) leaks out
    extra = {s for s in names if s.startswith("bfly") and s not in declared()}
    if lib == _lib.HIP_LIB:
        assert not extra, extra


def test_hip_library_loads_without_a_gpu():
    lib = bfly.hip()  # dlopen only; no device call
    assert lib.bfly_supported(bfly.KIND_C2C, 4096) == 1


def _stage_table(radices):
    """Closed form of the inter-pass table: pass s >= 1 holds
    W_{Ns R}^{r k} at [k R + r], Ns = R_0 ... R_{s-1}."""
    out, ns = [], 1
    for s, R in enumerate(radices):
        if s > 0:
            k = np.arange(ns)[:, None]
            r = np.arange(R)[None, :]
            out.append(np.exp(-2j * np.pi * (r * k % (ns * R)) / (ns * R)).reshape(-1))
        ns *= R
    return np.concatenate(out) if out else np.zeros(2, complex)


PLANS = {16: [16], 32: [16, 2], 64: [16, 4], 128: [16, 8], 256: [16, 16], 512: [16, 16, 2],
         1024: [16, 16, 4], 2048: [16, 16, 8], 4096: [16, 16, 16]}


@pytest.mark.parametrize("n", sorted(PLANS))
def test_c2c_table_is_fp64_twiddles(n):
    # [pass tables][W_n^j, j < 64 | W_n^{64 j}, j < n/64], each entry the FP64
    # value rounded once to fp32.
    lib = bfly.hip()
    nf = lib.bfly_table_floats(bfly.KIND_C2C, n)
    t = np.empty(nf, np.float32)
    assert lib.bfly_table_fill(bfly.KIND_C2C, n, t.ctypes.data) == 0
    stage = _stage_table(PLANS[n]) if len(PLANS[n]) > 1 else np.zeros(0, complex)
    lo = min(n, 64)
    two = np.concatenate([np.exp(-2j * np.pi * np.arange(lo) / n),
                          np.exp(-2j * np.pi * ((lo * np.arange(n // lo)) % n) / n)])
    want = np.concatenate([stage, two])
    assert nf == 2 * want.size
    assert np.array_equal(t[0::2], want.real.astype(np.float32))
    assert np.array_equal(t[1::2], want.imag.astype(np.float32))


@pytest.mark.parametrize("n", [8192, 16384, 32768, 65536])
def test_long_table_layout(n):
    lib = bfly.hip()
    t = np.empty(lib.bfly_table_floats(bfly.KIND_C2C, n), np.float32)
    assert lib.bfly_table_fill(bfly.KIND_C2C, n, t.ctypes.data) == 0
    n1 = 1 << ((n.bit_length() - 1) // 2)
    n2 = n // n1
    s1, s2 = _stage_table(PLANS[n1]), _stage_table(PLANS[n2])
    c = t[0::2] + 1j * t[1::2]
    assert np.array_equal(c[:s1.size], s1.astype(np.complex64))
    assert np.array_equal(c[s1.size:s1.size + s2.size], s2.astype(np.complex64))
    w = np.exp(-2j * np.pi * np.arange(n) / n).astype(np.complex64)
    o = s1.size + s2.size
    assert np.array_equal(c[o:o + n], w)
    # lengths a single workgroup holds append that plan's inter-pass table
    single = {8192: [16, 16, 16, 2], 16384: [16, 16, 16, 4]}.get(n)
    tail = _stage_table(single).astype(np.complex64) if single else np.zeros(0, np.complex64)
    assert np.array_equal(c[o + n:], tail)
    # the default path needs the n*batch workspace only where it runs four-step:
    # above 16384, and 16384 below the batch that fills the GPU one transform
    # per workgroup (256); 8192 always runs one workgroup per transform
    assert lib.bfly_workspace_bytes(bfly.KIND_C2C, n, 64) == (0 if n == 8192 else n * 64 * 8)
    if n <= 16384:
        assert lib.bfly_workspace_bytes(bfly.KIND_C2C, n, 256) == 0
        assert lib.bfly_workspace_bytes(bfly.KIND_C2C, n, 63) == (0 if n == 8192 else n * 63 * 8)


@pytest.mark.parametrize("kind", [bfly.KIND_R2C, bfly.KIND_C2R])
@pytest.mark.parametrize("n", [32, 128, 8192])
def test_real_table_layout(kind, n):
    lib = bfly.hip()
    t = np.empty(lib.bfly_table_floats(kind, n), np.float32)
    assert lib.bfly_table_fill(kind, n, t.ctypes.data) == 0
    s = _stage_table(PLANS[n // 2])
    c = t[0::2] + 1j * t[1::2]
    off = 0 if len(PLANS[n // 2]) == 1 else s.size
    w = np.exp(-2j * np.pi * np.arange(n // 2 + 1) / n).astype(np.complex64)
    assert np.array_equal(c[off:off + n // 2 + 1], w)


def test_supported_envelopes():
    lib = bfly.hip()
    for n in [1 << k for k in range(4, 17)]:
        assert lib.bfly_supported(bfly.KIND_C2C, n) == 1
    for n in (0, 1, 8, 100, 131072, 1 << 20):
        assert lib.bfly_supported(bfly.KIND_C2C, n) == 0
    for kind in (bfly.KIND_R2C, bfly.KIND_C2R):
        assert [lib.bfly_supported(kind, 1 << k) for k in range(3, 16)] == \
            [0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0]
    assert lib.bfly_supported(7, 64) == 0
    assert lib.bfly_table_floats(bfly.KIND_C2C, 100) == 0
    assert lib.bfly_table_fill(bfly.KIND_C2C, 100, None) == bfly.ERR_SHAPE


def test_execute_guards_through_stub():
    s = bfly.stub()
    s.bfly_stub_reset()
    buf = np.zeros(1 << 16, np.float32)
    p = buf.ctypes.data
    # the stub never dereferences: distinct, far-apart addresses stand for
    # the input, output, workspace and table buffers
    i, o, w, t = p, p + (1 << 30), p + (2 << 30), p + (3 << 30)
    assert s.bfly_c2c(i, o, 4096, 1, -1, t, None, None) == 0
    assert s.bfly_c2c(i, o, 4096, 0, -1, t, None, None) == bfly.ERR_SHAPE
    assert s.bfly_c2c(None, o, 4096, 1, -1, t, None, None) == bfly.ERR_SHAPE
    assert s.bfly_c2c(i, o, 4096, 1, 0, t, None, None) == bfly.ERR_SHAPE
    assert s.bfly_c2c(i, o, 4096, 1, -1, None, None, None) == _lib.ERR_TABLE
    assert s.bfly_c2c(i, o, 32768, 1, -1, t, None, None) == bfly.ERR_SHAPE  # four-step needs work
    assert s.bfly_c2c(i, o, 32768, 1, -1, t, w, None) == 0
    assert s.bfly_c2c(i, o, 8192, 64, -1, t, None, None) == 0  # one workgroup per transform
    assert s.bfly_c2c(i, o, 16384, 1, -1, t, None, None) == bfly.ERR_SHAPE  # small batch: four-step
    assert s.bfly_c2c_ex(i, o, 8192, 1, -1, _lib.VARIANT_PLAIN, t, None, None) == 0
    assert s.bfly_c2c_ex(i, o, 8192, 64, -1, _lib.VARIANT_FOUR_STEP, t, None, None) == bfly.ERR_SHAPE
    assert s.bfly_c2c_ex(i, o, 8192, 1, -1, _lib.VARIANT_LAST + 1, t, None, None) == bfly.ERR_SHAPE
    assert s.bfly_r2c(i, o, 16, 1, t, None) == bfly.ERR_SHAPE
    assert s.bfly_c2r(i, o, 16384, 1, t, None) == bfly.ERR_SHAPE
    assert s.bfly_c2c(i, p + (1 << 46), 4096, 0x7FFFFFFF, -1, t, None, None) == 0  # grid fits (64 TiB apart)
    assert s.bfly_real_ex(1, i, o, 1024, 1, 513, 1026, 0, _lib.REAL_ADJOINT, t, None) == 0
    assert s.bfly_real_ex(2, i, o, 1024, 1, 40, 80, _lib.VARIANT_STREAM_LT, _lib.REAL_ADJOINT, t, None) == 0
    assert s.bfly_real_ex(1, i, o, 1024, 1, 513, 1026, 0, 2, t, None) == bfly.ERR_SHAPE  # unknown flag
    assert s.bfly_real_ex(2, i, o, 1024, 1, 513, 1026, _lib.VARIANT_STREAM2, 0, t, None) == bfly.ERR_SHAPE
    # overlapping buffers are refused at the extents each call touches
    assert s.bfly_c2c(i, i, 4096, 1, -1, t, None, None) == bfly.ERR_SHAPE  # in place
    assert s.bfly_c2c(i, i + 8 * 4096 - 8, 4096, 1, -1, t, None, None) == bfly.ERR_SHAPE  # one complex shared
    assert s.bfly_c2c(i, i + 8 * 4096, 4096, 1, -1, t, None, None) == 0  # adjacent
    assert s.bfly_c2c(i, o, 32768, 1, -1, t, o + 8, None) == bfly.ERR_SHAPE  # workspace inside out
    assert s.bfly_c2c(i, o, 4096, 1, -1, t, o, None) == 0  # a workspace the call does not use
    assert s.bfly_r2c(i, i + 4 * 1024, 1024, 1, t, None) == 0  # row of n reals, then the spectrum
    assert s.bfly_r2c(i, i + 4 * 1022, 1024, 1, t, None) == bfly.ERR_SHAPE
    assert s.bfly_real_ex(2, i, i + 4 * (3 * 80 + 2 * 40), 1024, 4, 40, 80, 0, 0, t, None) == 0  # truncated
    assert s.bfly_real_ex(2, i, i + 4 * (3 * 80 + 2 * 40) - 4, 1024, 4, 40, 80, 0, 0, t, None) == bfly.ERR_SHAPE
    cnt = (C.c_int * 8)()
    s.bfly_stub_counters(cnt)
    assert cnt[3] == 11  # launches counted only for accepted calls


def test_scaled_c2c_guards_through_stub():
    s = bfly.stub()
    buf = np.zeros(1 << 16, np.float32)
    p = buf.ctypes.data
    i, o, w, t = p, p + (1 << 30), p + (2 << 30), p + (3 << 30)
    assert s.bfly_c2c_scaled(i, o, 4096, 3, 1, 1.0 / 4096, t, None, None) == 0
    assert s.bfly_c2c_scaled(i, o, 16, 1, -1, -2.5, t, None, None) == 0
    assert s.bfly_c2c_scaled(i, o, 4096, 3, 1, float("nan"), t, None, None) == bfly.ERR_SHAPE
    assert s.bfly_c2c_scaled(i, o, 4096, 3, 1, float("inf"), t, None, None) == bfly.ERR_SHAPE
    assert s.bfly_c2c_scaled(i, o, 4095, 3, 1, 1.0, t, None, None) == bfly.ERR_SHAPE
    assert s.bfly_c2c_scaled(i, o, 4096, 0, 1, 1.0, t, None, None) == bfly.ERR_SHAPE
    assert s.bfly_c2c_scaled(i, o, 4096, 3, 0, 1.0, t, None, None) == bfly.ERR_SHAPE
    assert s.bfly_c2c_scaled(i, o, 8192, 1, 1, 1.0, t, None, None) == 0  # LDS-twiddle kernel, no workspace
    assert s.bfly_c2c_scaled(i, o, 16384, 64, 1, 1.0, t, None, None) == bfly.ERR_SHAPE  # four-step: work
    assert s.bfly_c2c_scaled(i, o, 16384, 64, 1, 1.0, t, w, None) == 0
    assert s.bfly_c2c_scaled(i, o, 16384, 256, 1, 1.0, t, None, None) == 0
    assert s.bfly_c2c_scaled(i, o, 65536, 2, 1, 1.0, t, w, None) == 0
    assert s.bfly_c2c_scaled(i, i, 4096, 1, 1, 1.0, t, None, None) == bfly.ERR_SHAPE  # in place


def test_padded_rows_guards_through_stub():
    s = bfly.stub()
    buf = np.zeros(1 << 16, np.float32)
    p = buf.ctypes.data
    i, o, w, t = p, p + (1 << 30), p + (2 << 30), p + (3 << 30)
    assert s.bfly_c2c_rows(i, 4100, o, 4096, 4096, 16, -1, t, None, None) == 0
    assert s.bfly_c2c_rows(i, 20000, o, 16384, 16384, 3, 1, t, None, None) == 0
    assert s.bfly_c2c_rows(i, 1, o, 1, 64, 1, -1, t, None, None) == bfly.ERR_SHAPE  # one row still needs dist >= n
    assert s.bfly_c2c_rows(i, 4095, o, 4096, 4096, 16, -1, t, None, None) == bfly.ERR_SHAPE  # rows would overlap
    assert s.bfly_c2c_rows(i, 4096, o, 4095, 4096, 16, -1, t, None, None) == bfly.ERR_SHAPE
    assert s.bfly_c2c_rows(i, 48, o, 48, 48, 1, -1, t, None, None) == bfly.ERR_SHAPE  # not a power of two
    assert s.bfly_c2c_rows(i, 131072, o, 131072, 131072, 1, -1, t, w, None) == bfly.ERR_SHAPE  # past 65536
    assert s.bfly_c2c_rows(i + 4, 64, o, 64, 64, 1, -1, t, None, None) == bfly.ERR_SHAPE  # not 8-byte aligned
    assert s.bfly_c2c_rows(i, 64, o, 64, 64, 0, -1, t, None, None) == bfly.ERR_SHAPE
    assert s.bfly_c2c_rows(i, 64, o, 64, 64, 1, 0, t, None, None) == bfly.ERR_SHAPE
    assert s.bfly_c2c_rows(i, 64, o, 64, 64, 1, -1, None, None, None) == _lib.ERR_TABLE
    assert s.bfly_c2c_rows(i, 1 << 58, o, 64, 64, 2, -1, t, None, None) == bfly.ERR_SHAPE  # extent past 2^58
    # spans: 8 ((b-1) dist + n) bytes; n=64, b=4, dist=100 -> 2912
    assert s.bfly_c2c_rows(i, 100, i + 2912, 100, 64, 4, -1, t, None, None) == 0
    assert s.bfly_c2c_rows(i, 100, i + 2904, 100, 64, 4, -1, t, None, None) == bfly.ERR_SHAPE
    assert s.bfly_c2c_rows(i + 2912, 100, i, 100, 64, 4, -1, t, None, None) == 0
    # four-step lengths need a contiguous workspace of 8 n batch bytes
    assert s.bfly_c2c_rows(i, 70000, o, 65536, 65536, 2, -1, t, None, None) == bfly.ERR_SHAPE
    assert s.bfly_c2c_rows(i, 70000, o, 65536, 65536, 2, -1, t, w, None) == 0
    assert s.bfly_c2c_rows(i, 40000, o, 32768, 32768, 3, 1, t, w, None) == 0
    assert s.bfly_c2c_rows(i, 70000, o, 65536, 65536, 2, -1, t, o + 8 * 65536, None) == bfly.ERR_SHAPE  # in out
    assert s.bfly_c2c_rows(i, 70000, o, 65536, 65536, 2, -1, t, i + 8 * 70000 - 8, None) == bfly.ERR_SHAPE


def test_rows_and_scale_wrappers_refuse_bad_arguments_on_the_host():
    import torch
    from bfly import ops
    with pytest.raises(ValueError):  # host tensor: refused before any device call
        ops.c2c_rows(torch.zeros(2, 8, dtype=torch.complex64))
    with pytest.raises(ValueError):
        ops.c2c(torch.zeros(2, 8, dtype=torch.complex64), -1, norm="ortho", scale=0.5)
    assert ops.norm_scale(64, -1, None) is None and ops.norm_scale(64, 1, None) == 1.0 / 64
    assert ops.norm_scale(64, 1, "backward") == 1.0 / 64
    assert ops.norm_scale(64, -1, "forward") == 1.0 / 64
    assert ops.norm_scale(64, 1, "ortho") == ops.norm_scale(64, -1, "ortho") == 0.125


@pytest.mark.parametrize("norm", [None, "backward", "ortho", "forward"])
def test_norm_scale_follows_numpy(norm):
    from bfly import ops
    rng = np.random.default_rng(5)
    n = 32
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign, fn in ((-1, np.fft.fft), (1, np.fft.ifft)):
        sc = ops.norm_scale(n, sign, norm)
        raw = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * n  # unnormalised both ways
        np.testing.assert_allclose(raw * (1.0 if sc is None else sc), fn(x, norm=norm), rtol=1e-12, atol=1e-12)
    with pytest.raises(ValueError):
        ops.norm_scale(n, 1, "unitary")


def test_chain_launch_count_and_traced_guards_through_stub():
    s = bfly.stub()
    # one launch up to 4096, one workgroup per transform for 8192 / 16384
    # from 256 rows (any batch at 8192), the two-launch four-step otherwise
    for n, b, v, want in ((4096, 4096, 0, 1), (8192, 1, 0, 1), (16384, 256, 0, 1), (16384, 255, 0, 2),
                          (32768, 1, 0, 2), (65536, 256, 0, 2), (8192, 64, _lib.VARIANT_FOUR_STEP, 2),
                          (4096, 16, _lib.VARIANT_SPLIT_MFMA, 1)):
        assert s.bfly_c2c_launches(n, b, v) == want, (n, b, v)
    assert s.bfly_c2c_launches(100, 1, 0) == bfly.ERR_SHAPE
    assert s.bfly_c2c_launches(4096, 0, 0) == bfly.ERR_SHAPE
    assert s.bfly_c2c_launches(4096, 1, _lib.VARIANT_LAST + 1) == bfly.ERR_SHAPE
    s.bfly_stub_reset()
    buf = np.zeros(1 << 16, np.float32)
    p = buf.ctypes.data
    i, o, w, t = p, p + (1 << 30), p + (2 << 30), p + (3 << 30)
    ev = [C.c_void_p() for _ in range(3)]
    for e in ev:
        assert s.bfly_rt_event_create(C.byref(e)) == 0
    arr = (C.c_void_p * 3)(*[e.value for e in ev])
    assert s.bfly_c2c_traced(i, o, 4096, 2, -1, 0, t, None, None, arr, 2) == 0
    assert s.bfly_c2c_traced(i, o, 4096, 2, -1, 0, t, None, None, arr, 3) == bfly.ERR_SHAPE  # one launch
    assert s.bfly_c2c_traced(i, o, 65536, 2, -1, 0, t, w, None, arr, 3) == 0
    assert s.bfly_c2c_traced(i, o, 65536, 2, -1, 0, t, None, None, arr, 3) == bfly.ERR_SHAPE  # no workspace
    assert s.bfly_c2c_traced(i, i, 4096, 2, -1, 0, t, None, None, arr, 2) == bfly.ERR_SHAPE  # in place
    assert s.bfly_c2c_traced(i, o, 4096, 2, -1, 0, t, None, None, None, 2) == bfly.ERR_SHAPE
    ms = C.c_float()
    assert s.bfly_rt_event_elapsed(arr[0], arr[1], C.byref(ms)) == 0 and ms.value >= 0.0
    for e in ev:
        assert s.bfly_rt_event_destroy(e) == 0
    cnt = (C.c_int * 8)()
    s.bfly_stub_counters(cnt)
    assert cnt[3] == 2


@pytest.mark.parametrize("n1", [16, 64, 256, 512, 1024, 2048, 4096])
def test_fft2d_column_table_layout(n1):
    s = bfly.stub()
    nf = s.bfly_fft2d_table_floats(n1)
    tab = np.empty(nf, np.float32)
    assert s.bfly_fft2d_table_fill(n1, tab.ctypes.data) == 0
    c = tab.view(np.complex64)
    p = n1 if n1 <= 2048 else 64
    q = n1 // p
    # [local C2C table of p][local C2C table of q][W_n1^j, j<64 | W_n1^{64 j}] (128 entries)
    lp = s.bfly_table_floats(bfly.KIND_C2C, p) // 2
    tp = np.empty(2 * lp, np.float32)
    assert s.bfly_table_fill(bfly.KIND_C2C, p, tp.ctypes.data) == 0
    np.testing.assert_array_equal(c[:lp], tp.view(np.complex64))
    if q == 1:
        assert c.size == lp
        return
    lq = s.bfly_table_floats(bfly.KIND_C2C, q) // 2
    tq = np.empty(2 * lq, np.float32)
    assert s.bfly_table_fill(bfly.KIND_C2C, q, tq.ctypes.data) == 0
    np.testing.assert_array_equal(c[lp:lp + lq], tq.view(np.complex64))
    w = c[lp + lq:]
    assert w.size == 128
    j = np.arange(64)
    np.testing.assert_allclose(w[:64], np.exp(-2j * np.pi * j / n1), atol=1e-7)
    np.testing.assert_allclose(w[64:64 + n1 // 64], np.exp(-2j * np.pi * 64 * np.arange(n1 // 64) / n1), atol=1e-7)
    assert not w[64 + n1 // 64:].any()


def test_fft2d_guards_through_stub():
    s = bfly.stub()
    buf = np.zeros(16, np.float32)
    p = buf.ctypes.data
    i, o, w, t, tc = p, p + (1 << 30), p + (2 << 30), p + (3 << 30), p + (4 << 30)
    assert s.bfly_fft2d(i, o, 64, 128, 3, -1, t, tc, w, None) == 0
    assert s.bfly_fft2d(i, o, 4096, 4096, 1, 1, t, tc, w, None) == 0
    assert s.bfly_fft2d_workspace_bytes(64, 128, 3) == 8 * 64 * 128 * 3
    assert s.bfly_fft2d_workspace_bytes(8, 128, 3) == 0 and s.bfly_fft2d_table_floats(8192) == 0
    for n1, n2 in ((8, 64), (64, 8192), (48, 64), (64, 0)):
        assert s.bfly_fft2d(i, o, n1, n2, 1, -1, t, tc, w, None) == bfly.ERR_SHAPE
    assert s.bfly_fft2d(i, o, 64, 64, 0, -1, t, tc, w, None) == bfly.ERR_SHAPE
    assert s.bfly_fft2d(i, o, 64, 64, 1, -1, t, tc, None, None) == bfly.ERR_SHAPE
    assert s.bfly_fft2d(i, o, 64, 64, 1, -1, None, tc, w, None) == _lib.ERR_TABLE
    assert s.bfly_fft2d(i, i + 8 * 64 * 64 - 8, 64, 64, 1, -1, t, tc, w, None) == bfly.ERR_SHAPE
    assert s.bfly_fft2d(i, o, 64, 64, 1, -1, t, tc, o + 8, None) == bfly.ERR_SHAPE
    assert s.bfly_fft2d(i, i + 8 * 64 * 64, 64, 64, 1, -1, t, tc, w, None) == 0  # adjacent


def test_strided_columns_guards_through_stub():
    s = bfly.stub()
    buf = np.zeros(16, np.float32)
    p = buf.ctypes.data
    i, o, w, t = p, p + (1 << 30), p + (2 << 30), p + (3 << 30)
    assert s.bfly_c2c_columns(i, o, 1024, 100, 3, -1, t, None, None) == 0     # one pass: no workspace
    assert s.bfly_c2c_columns_workspace_bytes(1024, 100, 3) == 0
    assert s.bfly_c2c_columns_workspace_bytes(4096, 100, 3) == 8 * 4096 * 100 * 3
    assert s.bfly_c2c_columns(i, o, 4096, 100, 3, -1, t, None, None) == bfly.ERR_SHAPE
    assert s.bfly_c2c_columns(i, o, 4096, 100, 3, 1, t, w, None) == 0
    for n, ncols, b in ((8, 4, 1), (8192, 4, 1), (96, 4, 1), (64, 0, 1), (64, 4, 0)):
        assert s.bfly_c2c_columns(i, o, n, ncols, b, -1, t, w, None) == bfly.ERR_SHAPE, (n, ncols, b)
    assert s.bfly_c2c_columns(i, o, 64, 4, 1, 2, t, w, None) == bfly.ERR_SHAPE
    assert s.bfly_c2c_columns(i, o, 64, 4, 1, -1, None, w, None) == _lib.ERR_TABLE
    assert s.bfly_c2c_columns(i, i + 8, 64, 4, 1, -1, t, w, None) == bfly.ERR_SHAPE       # overlap
    assert s.bfly_c2c_columns(i, o, 4096, 4, 1, -1, t, o + 64, None) == bfly.ERR_SHAPE    # work inside out


def test_python_modules_import_and_variant_ids_match_header():
    import importlib
    for m in ("bfly.ops", "bfly.tune", "bfly.autograd", "bfly.spectral"):
        importlib.import_module(m)
    text = HEADER.read_text()
    for name in ("AUTO", "PLAIN", "STREAM", "PLAIN_SWZ", "STREAM_SWZ", "STREAM_LT", "STREAM2",
                 "FOUR_STEP", "SPLIT_MFMA", "SPLIT_MFMA_STREAM", "LAST"):
        m = re.search(rf"#define BFLY_VARIANT_{name} (\d+)", text)
        assert m and int(m.group(1)) == getattr(_lib, f"VARIANT_{name}"), name
    for name in ("OK", "ERR_NOT_LOADED", "ERR_TABLE", "ERR_ALLOC", "ERR_H2D", "ERR_LAUNCH",
                 "ERR_SYNC", "ERR_D2H", "ERR_SHAPE"):
        m = re.search(rf"#define BFLY_{name} \(?(-?\d+)\)?", text)
        assert m and int(m.group(1)) == getattr(_lib, name), name
    m = re.search(r"#define BFLY_REAL_ADJOINT (\d+)u", text)
    assert m and int(m.group(1)) == _lib.REAL_ADJOINT
    for name in ("FORWARD", "GRAD_INPUT", "GRAD_WEIGHT"):
        m = re.search(rf"#define BFLY_MIX_{name} (\d+)", text)
        assert m and int(m.group(1)) == getattr(_lib, f"MIX_{name}"), name


def test_host_pipelined_chunks_and_guards_through_stub():
    """bfly_host_pipelined on the CPU stub: the batch moves in chunks of
    chunk_rows transforms (upload, launch, download per chunk, in order),
    the whole batch is guarded as one execute call first, and an injected
    failure stops the call with the step's code."""
    s = bfly.stub()
    s.bfly_stub_reset()
    n, b = 64, 5
    host_in = np.arange(b * 2 * n, dtype=np.float32)
    host_out = np.zeros_like(host_in)
    dev_in = np.zeros_like(host_in)
    dev_out = np.full_like(host_in, 7.0)  # the stub's launch computes nothing
    tab = np.zeros(16, np.float32)
    hi, ho, di, do, tp = (a.ctypes.data for a in (host_in, host_out, dev_in, dev_out, tab))
    cnt = (C.c_int * 8)()

    def counters():
        s.bfly_stub_counters(cnt)
        return cnt[2], cnt[3]  # copies, launches

    c0, l0 = counters()
    assert s.bfly_host_pipelined(bfly.KIND_C2C, -1, 0, hi, ho, n, b, tp, di, do, None, 2, None) == 0
    c1, l1 = counters()
    assert (c1 - c0, l1 - l0) == (6, 3)  # three chunks of 2, 2, 1 transforms
    assert np.array_equal(dev_in, host_in) and np.all(host_out == 7.0)
    # the default chunk (about 8 MiB of input) holds this whole batch
    assert s.bfly_host_pipelined(bfly.KIND_C2C, 1, 0, hi, ho, n, b, tp, di, do, None, 0, None) == 0
    c2, l2 = counters()
    assert (c2 - c1, l2 - l1) == (2, 1)
    # the second chunk's upload fails: H2D, one launch done
    s.bfly_stub_set_failures(0, c2 + 3, 0)
    assert s.bfly_host_pipelined(bfly.KIND_C2C, -1, 0, hi, ho, n, b, tp, di, do, None, 2, None) == _lib.ERR_H2D
    s.bfly_stub_set_failures(0, 0, 0)
    c3, l3 = counters()
    assert (c3 - c2, l3 - l2) == (3, 1)
    # the first download fails: D2H
    s.bfly_stub_set_failures(0, c3 + 2, 0)
    assert s.bfly_host_pipelined(bfly.KIND_C2C, -1, 0, hi, ho, n, b, tp, di, do, None, 2, None) == _lib.ERR_D2H
    s.bfly_stub_set_failures(0, 0, 0)
    # real kinds: rows of n reals and n+2 floats
    r_in = np.zeros(b * n, np.float32); r_half = np.zeros(b * (n + 2), np.float32)
    d_in = np.zeros_like(r_in); d_half = np.zeros_like(r_half)
    args = (r_in.ctypes.data, r_half.ctypes.data, n, b, tp, d_in.ctypes.data, d_half.ctypes.data, None, 2, None)
    assert s.bfly_host_pipelined(bfly.KIND_R2C, 0, 0, *args) == 0  # the sign of a real kind is implied
    args = (r_half.ctypes.data, r_in.ctypes.data, n, b, tp, d_half.ctypes.data, d_in.ctypes.data, None, 2, None)
    assert s.bfly_host_pipelined(bfly.KIND_C2R, 0, 0, *args) == 0
    # guards (nothing is copied or launched)
    c4, l4 = counters()
    bad = [
        (7, -1, 0, hi, ho, n, b, tp, di, do, None, 2, None),                      # kind
        (bfly.KIND_C2C, 0, 0, hi, ho, n, b, tp, di, do, None, 2, None),           # sign
        (bfly.KIND_C2C, -1, 0, None, ho, n, b, tp, di, do, None, 2, None),        # host input
        (bfly.KIND_C2C, -1, 0, hi, ho, n, b, tp, None, do, None, 2, None),        # device output
        (bfly.KIND_C2C, -1, 0, hi, ho, 48, b, tp, di, do, None, 2, None),         # length
        (bfly.KIND_C2C, -1, 0, hi, ho, n, 0, tp, di, do, None, 2, None),          # batch
        (bfly.KIND_C2C, -1, _lib.VARIANT_LAST + 1, hi, ho, n, b, tp, di, do, None, 2, None),
        (bfly.KIND_R2C, 0, _lib.VARIANT_STREAM2, hi, ho, n, b, tp, di, do, None, 2, None),
        (bfly.KIND_C2C, -1, 0, hi, ho, n, b, tp, di, di, None, 2, None),          # device buffers overlap
        (bfly.KIND_C2C, -1, 0, hi, ho, 32768, 1, tp, di, do, None, 2, None),      # four-step needs work
    ]
    for a in bad:
        assert s.bfly_host_pipelined(*a) == bfly.ERR_SHAPE, a[:3]
    assert s.bfly_host_pipelined(bfly.KIND_C2C, -1, 0, hi, ho, n, b, None, di, do, None, 2, None) == _lib.ERR_TABLE
    assert counters() == (c4, l4)

