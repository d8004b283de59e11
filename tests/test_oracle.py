"""Pins the CPU oracle (oracle/bfly_oracle.c) before anything is checked against it.

Each restated function is checked against at least one thing that is not its
own formula:
  * the reference's own code compiled from its sources (oracle/_ref: reference.cpp
    and the header-only fft_k.hpp / long_fft_ub.h / descriptors.hpp, plain g++,
    no stand-in headers), bit for bit;
  * golden vectors written by that build and by the reference's Python input
    twin (tests/golden/reference_vectors.npz, tests/golden/make_goldens.py);
  * goldens the reference's tests hold (tests/test_limits.cpp,
    tests/test_runtime_guards.cpp, tests/test_descriptors.py);
  * mathematics: direct O(N^2) DFT, numpy pocketfft in FP64, closed forms
    (impulse, constant, tone), Parseval, linearity, round trips, the FFT the
    Generator tables drive, Marsaglia's xorshift32 known answer.
"""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

import oracle

GOLD = np.load(Path(__file__).resolve().parent / "golden" / "reference_vectors.npz")
REF = oracle.ref()
needs_ref = pytest.mark.skipif(REF is None, reason="oracle/_ref not built (no reference tree)")


def np_fft64(x, n, sign=-1):
    xc = x[:, 0::2].astype(np.float64) + 1j * x[:, 1::2].astype(np.float64)
    y = np.fft.fft(xc, axis=1) if sign < 0 else np.fft.ifft(xc, axis=1) * n
    out = np.empty((x.shape[0], 2 * n))
    out[:, 0::2], out[:, 1::2] = y.real, y.imag
    return out


# ------------------------------------------------------------ FP64 transforms
@pytest.mark.parametrize("n", [2, 4, 8, 16, 64, 1024, 4096, 65536])
def test_fft_f64_matches_pocketfft(n):
    x = oracle.uniform_input(3, (2, 2 * n))
    for sign in (-1, 1):
        got = oracle.c2c_f64(x, n, sign)
        want = np_fft64(x, n, sign)
        assert np.max(np.abs(got - want)) <= 1e-13 * max(1.0, np.max(np.abs(want))) * np.log2(n)


@pytest.mark.parametrize("n", [8, 16, 64, 256])
def test_fft_f64_matches_direct_dft(n):
    x = oracle.uniform_input(4, (1, 2 * n))
    direct = np.empty(2 * n)
    oracle.lib().orc_dft_direct_f64(np.ascontiguousarray(x[0]), direct, n, -1)
    fast = oracle.c2c_f64(x, n)[0]
    assert np.max(np.abs(fast - direct)) < 1e-12 * n


@pytest.mark.parametrize("n", [64, 4096])
def test_closed_forms(n):
    imp = oracle.numerical_input(n, 5, "impulse")
    y = oracle.c2c_f64(imp, n)
    for b in range(5):
        k = np.arange(n)
        want = np.exp(-2j * np.pi * k * (b % n) / n)
        assert np.allclose(y[b, 0::2] + 1j * y[b, 1::2], want, atol=1e-12)
    const = oracle.numerical_input(n, 1, "constant")
    y = oracle.c2c_f64(const, n)[0]
    assert np.isclose(y[0], 0.25 * n) and np.isclose(y[1], -0.5 * n)
    assert np.max(np.abs(y[2:])) < 1e-9
    tone = oracle.numerical_input(n, 3, "tone")
    y = oracle.c2c_f64(tone, n)
    for b in range(3):
        mag = np.hypot(y[b, 0::2], y[b, 1::2])
        assert np.argmax(mag) == 1 + b % 7 and abs(mag.max() - n) < 1e-6 * n


def test_parseval_and_linearity():
    n = 1024
    x = oracle.uniform_input(5, (2, 2 * n)).astype(np.float64)
    y = oracle.c2c_f64(x.astype(np.float32), n)
    e_t = np.sum(x.astype(np.float32).astype(np.float64) ** 2, axis=1)
    e_f = np.sum(y ** 2, axis=1) / n
    assert np.allclose(e_t, e_f, rtol=1e-12)
    # dyadic inputs so that a + b is exact in fp32
    a = (np.round(oracle.uniform_input(6, (1, 2 * n)) * 4096) / 4096).astype(np.float32)
    b = (np.round(oracle.uniform_input(7, (1, 2 * n)) * 4096) / 4096).astype(np.float32)
    lhs = oracle.c2c_f64(a + b, n)
    assert np.allclose(lhs, oracle.c2c_f64(a, n) + oracle.c2c_f64(b, n), atol=1e-9)


@pytest.mark.parametrize("n", [32, 128, 4096, 8192])
def test_real_transforms_against_pocketfft(n):
    x = oracle.uniform_input(8, (3, n))
    h = oracle.r2c_f64(x, n)
    want = np.fft.rfft(x.astype(np.float64), axis=1)
    assert np.allclose(h[:, 0::2], want.real, atol=1e-9) and np.allclose(h[:, 1::2], want.imag, atol=1e-9)
    hh = oracle.half_input(n, 3)
    got = oracle.c2r_f64(hh, n)
    ref = np.fft.irfft(hh[:, 0::2].astype(np.float64) + 1j * hh[:, 1::2].astype(np.float64), n=n, axis=1)
    assert np.allclose(got, ref, atol=1e-12)


def test_four_step_chain_equals_direct_fft():
    for n in (8192, 16384, 32768, 65536):
        n1, n2 = oracle.long_split(n)
        x = oracle.uniform_input(9, (1, 2 * n))
        four = oracle.four_step_f64(x, n, n1, n2)
        direct = oracle.c2c_f64(x, n)
        assert np.max(np.abs(four - direct)) < 1e-10 * np.max(np.abs(direct))


def test_long_split_rule():
    # the four-step splits of the long chain, src/host/fft_check.cpp:295-308
    assert oracle.long_split(8192) == (64, 128)
    assert oracle.long_split(16384) == (128, 128)
    assert oracle.long_split(32768) == (128, 256)
    assert oracle.long_split(65536) == (256, 256)
    assert oracle.long_split(4096 * 4096 * 2) is None


# ------------------------------------------------- bit-exact vs the reference
def _ref_rows(fn, x, out_len, n):
    out = np.empty((x.shape[0], out_len), np.float32)
    for b in range(x.shape[0]):
        row = np.empty(out_len, np.float32)
        fn(np.ascontiguousarray(x[b]), row, n)
        out[b] = row
    return out


@pytest.mark.parametrize("n", [64, 256, 1024, 4096])
def test_ref_fft_golden_vectors(n):
    x = GOLD[f"fft_in_{n}"]
    got = _ref_rows(oracle.lib().orc_ref_fft_f32, x, 2 * n, n)
    assert np.array_equal(got, GOLD[f"fft_out_{n}"])


@pytest.mark.parametrize("n", [128, 1024, 8192])
def test_ref_r2c_golden_vectors(n):
    got = _ref_rows(oracle.lib().orc_ref_r2c_f32, GOLD[f"r2c_in_{n}"], n + 2, n)
    assert np.array_equal(got, GOLD[f"r2c_out_{n}"])


@pytest.mark.parametrize("n", [64, 1024, 4096])
def test_ref_c2r_golden_vectors(n):
    got = _ref_rows(oracle.lib().orc_ref_c2r_f32, GOLD[f"c2r_in_{n}"], n, n)
    assert np.array_equal(got, GOLD[f"c2r_out_{n}"])


@needs_ref
@pytest.mark.parametrize("n", [2, 8, 64, 512, 2048, 16384])
def test_ref_transforms_bit_exact_live(n):
    x = oracle.uniform_input(21, (2, 2 * n))
    a = _ref_rows(oracle.lib().orc_ref_fft_f32, x, 2 * n, n)
    b = _ref_rows(REF.ref_fft_f32, x, 2 * n, n)
    assert np.array_equal(a, b)
    xr = oracle.uniform_input(22, (2, n))
    assert np.array_equal(_ref_rows(oracle.lib().orc_ref_r2c_f32, xr, n + 2, n),
                          _ref_rows(REF.ref_r2c_f32, xr, n + 2, n))
    if n >= 4:
        h = oracle.half_input(n, 2)
        assert np.array_equal(_ref_rows(oracle.lib().orc_ref_c2r_f32, h, n, n),
                              _ref_rows(REF.ref_c2r_f32, h, n, n))


def test_pattern_golden():
    got = oracle.pattern(4096)
    assert np.array_equal(got, GOLD["pattern_0_4096"])


def test_pattern_closed_form():
    """patternAt (src/framework/reference.cpp:10-12): the index is a size_t, so
    i * 2654435761 wraps modulo 2^64 (not 2^32); the residue mod 1000 is
    divided by 1000 in fp32 and 0.5 subtracted in fp32.  Python integers give
    the 64-bit product exactly; indices above 2^32 tell the two wraps apart."""
    idx = list(range(5000)) + [2**32 - 1, 2**32, 2**32 + 12345, 2**40 + 7, 2**63 + 99]
    want = [np.float32(np.float32(((i * 2654435761) % 2**64) % 1000) / np.float32(1000.0)) - np.float32(0.5)
            for i in idx]
    got = [oracle.lib().orc_pattern_at(i) for i in idx]
    assert np.array_equal(np.array(got, np.float32), np.array(want, np.float32))
    wrap32 = np.float32(((2**32 + 12345) * 2654435761 % 2**32) % 1000) / np.float32(1000) - np.float32(0.5)
    assert got[idx.index(2**32 + 12345)] != wrap32


@needs_ref
def test_pattern_live():
    for i in list(range(0, 70000, 7)) + [2**32 + 5, 2**40 + 7]:
        assert oracle.lib().orc_pattern_at(i) == REF.ref_pattern_at(i)


def test_integer_rules_golden():
    L = oracle.lib()
    ns = GOLD["rule_n"]
    assert [L.orc_plane_k_for(int(n)) for n in ns] == list(GOLD["plane_k"])
    for i, n in enumerate(ns):
        for j, b in enumerate(GOLD["fold_batches"]):
            assert L.orc_fold_d_for(int(n), int(b), 48) == GOLD["fold_d"][i, j], (n, b)


@needs_ref
def test_integer_rules_live():
    L = oracle.lib()
    for k in range(0, 20):
        n = 1 << k
        assert L.orc_plane_k_for(n) == REF.ref_plane_k_for(n)
        for b in (1, 2, 47, 96, 1000, 0xFFFFFFFF):
            for nblk in (1, 24, 48):
                assert L.orc_fold_d_for(n, b, nblk) == REF.ref_fold_d_for(n, b, nblk)
        for D in (1, 2, 4):
            for K in (8, 16, 32):
                assert L.orc_row_fft_ub_bytes(n, D, K) == REF.ref_row_fft_ub_bytes(n, D, K)


def test_row_fft_goldens_from_reference_tests():
    # tests/test_limits.cpp:114-180 of the reference.
    L = oracle.lib()
    d, k = C.c_uint32(), C.c_uint32()

    def rr(length, rows, dov=0, kov=0):
        ok = L.orc_resolve_row_fft(length, rows, dov, kov, C.byref(d), C.byref(k))
        return ok, d.value, k.value
    assert L.orc_row_fft_ub_bytes(4096, 1, 16) == 46 * 4096 + 12 * 4096 // 16 + 128 == 191616
    ok, D, K = rr(64, 128)
    assert (ok, D, K) == (1, 2, 8) and L.orc_row_fft_ub_bytes(64, D, K) == 5056
    ok, D, K = rr(128, 64)
    assert (ok, D, K) == (1, 1, 8) and L.orc_row_fft_ub_bytes(128, D, K) == 6208
    ok, D, K = rr(64, 524288)
    assert D == 4 and L.orc_row_fft_ub_bytes(64, D, K) == 8832
    ok, D, K = rr(4096, 4096)
    assert (D, K) == (1, 16)
    # Override (64, 128, D=4, K=32): test_limits.cpp:159,176-178 expects K to
    # clamp to 8, but descriptors.hpp:242-248 rejects it -- the source is the
    # contract, so the oracle reports illegal with K=32 kept.
    ok, D, K = rr(64, 128, 4, 32)
    assert (ok, D, K) == (0, 4, 32)


@needs_ref
def test_resolve_row_fft_live():
    L = oracle.lib()
    a, b, c, e = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
    for length in (64, 128, 256, 1024, 4096):
        for rows in (1, 3, 64, 4096, 524288):
            for dov in (0, 1, 3):
                for kov in (0, 8, 12, 16, 32):
                    x = L.orc_resolve_row_fft(length, rows, dov, kov, C.byref(a), C.byref(b))
                    y = REF.ref_resolve_row_fft(length, rows, dov, kov, C.byref(c), C.byref(e))
                    assert (x, a.value, b.value) == (y, c.value, e.value)


def _np_bitrev(n):
    """Bit reversal by reversing each index's binary string."""
    bits = n.bit_length() - 1
    if bits == 0:
        return np.zeros(1, np.int64)
    return np.array([int(format(i, f"0{bits}b")[::-1], 2) for i in range(n)], np.int64)


def _tables(n):
    L = oracle.lib()
    twr, twi = np.empty(n + 16, np.float32), np.empty(n + 16, np.float32)
    L.orc_gen_twiddles(n, twr, twi)
    out = {"twr": twr, "twi": twi}
    for name, fn in (("bitrev", L.orc_gen_bit_reverse), ("plane", L.orc_gen_bit_reverse_plane),
                     ("transpose", L.orc_gen_transpose_plane)):
        d = np.empty(n, np.uint32)
        fn(n, d)
        out[name] = d.astype(np.int64)
    return out


@pytest.mark.parametrize("n", [8, 16, 64, 256, 1024, 4096, 65536])
def test_generator_tables_drive_a_radix2_fft(n):
    """genTwiddles + genBitReverse (src/framework/butterfly.cpp:96-123) are the
    tables of a radix-2 DIT: gathering by the byte offsets / 8 and running the
    h = 1, 2, ..., n/2 butterfly sweep with stage slot s(h) = sum of align8(h')
    over h' < h gives the DFT (numpy pocketfft, FP64).  A wrong sign, slot
    alignment, stage order or permutation fails here."""
    t = _tables(n)
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    a = x[t["bitrev"] // 2].astype(np.complex128)
    slot, h = 0, 1
    while 2 * h <= n:
        w = t["twr"][slot:slot + h].astype(np.float64) + 1j * t["twi"][slot:slot + h].astype(np.float64)
        blk = a.reshape(-1, 2 * h)
        u, v = blk[:, :h], blk[:, h:] * w
        a = np.concatenate([u + v, u - v], axis=1).reshape(-1)
        slot += (h + 7) // 8 * 8
        h *= 2
    want = np.fft.fft(x)
    assert np.abs(a - want).max() <= 1e-6 * np.abs(want).max() * np.log2(n)
    assert slot <= n + 16 and not t["twr"][slot:].any() and not t["twi"][slot:].any()


@pytest.mark.parametrize("n", [8, 64, 256, 4096, 65536])
def test_generator_tables_closed_forms(n):
    """Every table against a numpy construction of its meaning, with K from the
    oracle's planeKFor (pinned below by the reference's goldens):
      twiddles: FP64 e^{-i pi j/h} rounded once to fp32, per align8(h) slot;
      bitrev:   2 * (index with its binary string reversed);
      plane:    the bit-reversed order viewed as [rows][K], transposed into K
                planes of rows, 8-byte offsets;
      transpose: [K][rows] -> [rows][K] index map, 4-byte offsets;
      interleave: gather of a planar [re plane | im plane] of D rows into
                interleaved [D][n][re, im], 4-byte offsets."""
    L = oracle.lib()
    t = _tables(n)
    K = L.orc_plane_k_for(n)
    rows = n // K
    slots = []
    h = 1
    while 2 * h <= n:
        slots += [np.exp(-1j * np.pi * np.arange(h) / h), np.zeros((h + 7) // 8 * 8 - h)]
        h *= 2
    want = np.concatenate(slots + [np.zeros(n + 16 - sum(len(s) for s in slots))])
    assert np.array_equal(t["twr"], want.real.astype(np.float32))
    assert np.array_equal(t["twi"], want.imag.astype(np.float32))
    br = _np_bitrev(n)
    assert np.array_equal(t["bitrev"], 2 * br)
    assert np.array_equal(t["plane"], 8 * br[np.arange(n).reshape(rows, K).T.ravel()])
    assert np.array_equal(t["transpose"], 4 * np.arange(n).reshape(K, rows).T.ravel())
    rng = np.random.default_rng(7)
    for D in (1, 2, 4):
        d = np.empty(2 * n * D, np.uint32)
        L.orc_gen_interleave(n, D, d)
        re = rng.normal(size=(D, n)).astype(np.float32)
        im = rng.normal(size=(D, n)).astype(np.float32)
        planar = np.concatenate([re.ravel(), im.ravel()]).view(np.uint8)
        gathered = np.stack([planar[d.astype(np.int64) + b] for b in range(4)], axis=1).copy().view(np.float32)
        assert np.array_equal(gathered.ravel(), np.stack([re, im], axis=-1).ravel())


def test_generator_permutation_properties():
    for n in (64, 1024, 4096):
        t = _tables(n)
        r = t["bitrev"] // 2
        assert sorted(r) == list(range(n)) and np.array_equal(r[r], np.arange(n))  # involution
        assert sorted(t["plane"] // 8) == list(range(n))
        assert sorted(t["transpose"] // 4) == list(range(n))


# ------------------------------------------------------------ error metrics
def test_max_rel_scaled_reference_values():
    # tests/test_runtime_guards.cpp:63-85 of the reference.
    f = lambda got, ref: oracle.lib().orc_max_rel_scaled(np.array(got, np.float32),
                                                         np.array(ref, np.float32), len(ref))
    assert f([1, 0], [0, 0]) == 1.0
    assert f([1e-6, 0], [0, 0]) > 0
    assert f([1e-6, 0], [1e-8, 0]) < 1e-4
    for v in (np.nan, np.inf, -np.inf):
        assert np.isinf(f([v, 0], [1e-8, 0]))
        assert np.isinf(f([0, 0], [v, 0]))
    assert f([101, 0], [100, 0]) == np.float64(np.float32(101) - np.float32(100)) / 100.0
    assert abs(f([101, 0], [100, 0]) - 0.01) < 1e-15


@needs_ref
def test_max_rel_scaled_live():
    rng = np.random.default_rng(0)
    for _ in range(20):
        a = rng.normal(size=33).astype(np.float32)
        b = (a + rng.normal(size=33) * 1e-3).astype(np.float32)
        assert oracle.lib().orc_max_rel_scaled(a, b, 33) == REF.ref_max_rel_scaled(a, b, 33)


def test_sample_batches():
    L = oracle.lib()
    out = np.zeros(8, np.uint32)
    assert L.orc_sample_batches(4096, 8, out, 8) == 8
    assert list(out) == [i * 4095 // 7 for i in range(8)]
    assert L.orc_sample_batches(3, 8, out, 8) == 3 and list(out[:3]) == [0, 1, 2]
    assert L.orc_sample_batches(0, 8, out, 8) == 0
    assert L.orc_sample_batches(10, 1, out, 8) == 1 and out[0] == 0


def test_rel_l2_metric():
    ref = np.array([[3.0, 4.0, 0.0, 0.0]])
    got = np.array([[3.0, 4.0, 0.0, 0.5]], np.float32)
    assert abs(oracle.rel_l2_rows(got, ref) - 0.1) < 1e-12
    got[0, 0] = np.nan
    assert np.isinf(oracle.rel_l2_rows(got, ref))


def test_app_hash_known_answer():
    """appHash (src/host/fft_check.cpp:63-68) is two steps of Marsaglia's
    xorshift32 (13, 17, 5) from the seed b*2654435761 + k*40503 mod 2^32.
    Marsaglia's published sequence from y = 2463534242 starts 723471715,
    2497366906 (Xorshift RNGs, J. Stat. Softw. 8(14), 2003); b = 296369474
    is the one b whose seed is 2463534242 at k = 0, so appHash(b, 0) is the
    sequence's second value.  The k multiplier is pinned by the reference
    twin's values below."""
    L = oracle.lib()
    assert (296369474 * 2654435761) % 2**32 == 2463534242
    assert L.orc_app_hash(296369474, 0) == 2497366906


def test_app_hash_matches_reference_twin():
    # scripts/bench_native_npu.py:27-37 (`_xh`), run by tests/golden/make_goldens.py
    L = oracle.lib()
    for i, b in enumerate(GOLD["hash_b"]):
        for j, k in enumerate(GOLD["hash_k"]):
            assert L.orc_app_hash(int(b), int(k)) == GOLD["hash_out"][i, j], (b, k)


@pytest.mark.parametrize("mode", ["ofdm", "radar", "dl"])
def test_app_inputs_match_reference_twin(mode):
    """ofdm / radar / dl (src/host/fft_check.cpp:118-159) against the vectors
    the reference's own Python twin wrote (scripts/bench_native_npu.py:40-84,
    stated bit-identical to fft_check): bit for bit for ofdm and dl (table
    lookups); radar sums FP64 cos/sin, so one fp32 ulp is allowed where the
    two libm evaluations round differently."""
    for n, b in ((64, 3), (256, 14), (2048, 5)):
        got = oracle.app_input(n, b, mode)
        want = GOLD[f"app_{mode}_{n}_{b}"]
        if mode == "radar":
            ulps = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
            assert ulps.max() <= 1 and (ulps == 0).mean() > 0.99, (n, b)
        else:
            assert np.array_equal(got, want), (n, b)


def test_app_input_properties():
    """Structure each application input must have, independent of the hash:
    ofdm guard bands and DC are zero, pilots every 12 bins are unit QPSK, the
    rest 16-QAM scaled 1/sqrt(10); dl is constant over blocks of 8 with a
    4-PSK phase; radar is bounded by the sum of its target amplitudes / 4."""
    n, b = 1024, 6
    x = oracle.app_input(n, b, "ofdm").reshape(b, n, 2)
    k = np.arange(n)
    guard = (k < n // 16) | (k >= n - n // 16)
    assert not x[:, guard].any()
    pil = x[:, (~guard) & (k % 12 == 0)]
    assert set(np.unique(np.abs(pil))) == {1.0}
    data = x[:, (~guard) & (k % 12 != 0)] * np.float32(np.sqrt(10))
    assert set(np.round(np.unique(data), 5)) == {-3.0, -1.0, 1.0, 3.0}
    d = oracle.app_input(n, b, "dl").reshape(b, n // 8, 8, 2)
    assert np.all(d == d[:, :, :1, :]) and np.all((d[..., 0] == 0) | (d[..., 1] == 0))
    r = oracle.app_input(n, b, "radar")
    assert np.abs(r).max() <= 0.25 * 1.875 + 1e-6
    h = oracle.half_input(64, 3)
    assert np.all(h[:, 65] == 0) and np.all(np.abs(h) <= 1)
    for mode in oracle.NUMERICAL_MODES:
        assert np.all(np.isfinite(oracle.numerical_input(64, 2, mode)))


def test_sin_input_sequence():
    # src/host/fft_check.cpp:398-399: sin(0.011 i) + 0.25 cos(0.037 i) over the
    # flat index, rounded once to fp32; 0.25 at i = 0, and numpy's FP64
    # evaluation agrees to the last fp32 bit (at most one ulp where the FP64
    # value sits on a rounding boundary)
    x = oracle.sin_input(100000)
    assert x[0] == np.float32(0.25)
    i = np.arange(100000, dtype=np.float64)
    want = (np.sin(0.011 * i) + 0.25 * np.cos(0.037 * i)).astype(np.float32)
    ulps = np.abs(x.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
    assert ulps.max() <= 1 and (ulps == 0).mean() > 0.999


def test_fft2_oracle_pinned():
    # against numpy's complex128 fft2 (pocketfft, an independent library), the
    # closed form of a shifted impulse, and the O(N^2) direct DFT on 8 x 16
    rng = np.random.default_rng(5)
    for n1, n2, b in ((16, 32, 2), (64, 16, 1), (8, 8, 3)):
        x = rng.uniform(-1, 1, (b, n1, 2 * n2)).astype(np.float32)
        got = oracle.fft2_f64(x, n1, n2).reshape(b, n1, n2, 2)
        z = x.reshape(b, n1, n2, 2).astype(np.float64)
        want = np.fft.fft2(z[..., 0] + 1j * z[..., 1])
        assert np.abs(got[..., 0] + 1j * got[..., 1] - want).max() <= 1e-12 * np.abs(want).max()
        inv = oracle.fft2_f64(x, n1, n2, sign=1).reshape(b, n1, n2, 2)
        want = np.fft.ifft2(z[..., 0] + 1j * z[..., 1]) * (n1 * n2)
        assert np.abs(inv[..., 0] + 1j * inv[..., 1] - want).max() <= 1e-12 * np.abs(want).max()
    n1, n2, a, c = 32, 64, 5, 9
    x = np.zeros((n1, 2 * n2), np.float32)
    x[a, 2 * c] = 1.0
    got = oracle.fft2_f64(x, n1, n2).reshape(n1, n2, 2)
    k1, k2 = np.meshgrid(np.arange(n1), np.arange(n2), indexing="ij")
    want = np.exp(-2j * np.pi * (a * k1 / n1 + c * k2 / n2))
    assert np.abs(got[..., 0] + 1j * got[..., 1] - want).max() < 1e-12
    n1, n2 = 8, 16
    x = rng.uniform(-1, 1, (n1, 2 * n2)).astype(np.float32)
    F1 = np.exp(-2j * np.pi * np.outer(np.arange(n1), np.arange(n1)) / n1)
    F2 = np.exp(-2j * np.pi * np.outer(np.arange(n2), np.arange(n2)) / n2)
    z = x.reshape(n1, n2, 2).astype(np.float64)
    want = F1 @ (z[..., 0] + 1j * z[..., 1]) @ F2
    got = oracle.fft2_f64(x, n1, n2).reshape(n1, n2, 2)
    assert np.abs(got[..., 0] + 1j * got[..., 1] - want).max() < 1e-11
