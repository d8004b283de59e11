"""GPU parity: libbfly_hip.so against the FP64 oracle on seeded inputs.

Gates (per transform): relative L2 error <= 1e-5 against FP64, and the
reference's maxRelScaled <= 1e-4 against its fp32-rounded FP64 result
(include/butterfly/reference.hpp:30-41, used by src/host/fft_check.cpp:1186-1248).
"""
import numpy as np
import pytest

import oracle
from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]

REL_L2 = 1e-5
MAX_REL_SCALED = 1e-4
GRID_N = [64, 128, 256, 512, 1024, 2048, 4096]
GRID_B = [1, 4, 16, 64, 256, 1024, 4096]


def _ops():
    from bfly import ops
    return ops


def run_c2c(x, n, sign=-1):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    y = _ops().c2c(t, sign, scale=1.0)      # the reference's unnormalised contract
    torch.cuda.synchronize()
    return y.cpu().numpy()


def gate_c2c(got, x, n, sign=-1):
    ref = oracle.c2c_f64(x, n, sign)
    e = oracle.rel_l2_rows(got, ref)
    assert e <= REL_L2, f"n={n} rel-L2 {e:.3e}"
    m = oracle.max_rel_scaled(got, ref.astype(np.float32))
    assert m <= MAX_REL_SCALED, f"n={n} maxRelScaled {m:.3e}"
    return e


def test_cfg1_c2c_n1024_b64_seed0():
    n, b = 1024, 64
    x = oracle.uniform_input(0, (b, 2 * n))
    got = run_c2c(x, n)
    e = gate_c2c(got, x, n)
    assert e < 1e-6  # an FP32 Stockham at n=1024 sits near 2e-7


@pytest.mark.parametrize("n", GRID_N)
def test_c2c_grid_seed1(n):
    # One seeded draw per (n, batch) point of the 49-point grid.
    for b in GRID_B:
        x = oracle.uniform_input(1, (b, 2 * n))
        got = run_c2c(x, n)
        gate_c2c(got, x, n)


VARIANTS = [1, 2, 3, 4, 5, 6]  # plain, stream, plain+swizzle, stream+swizzle, stream+LDS twiddles, stream x2


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("n", [16, 32, 64, 128, 256, 512, 1024, 2048, 4096])
def test_c2c_every_variant(n, variant):
    # Each kernel variant on ragged and multi-wave batches, both signs.
    import torch
    from bfly import _lib
    for b, sign in ((1, -1), (97, 1), (3000, -1)):
        x = oracle.uniform_input(13, (b, 2 * n))
        xt = torch.from_numpy(x).cuda()
        y = torch.empty_like(xt)
        tab = _ops().table(0, n, xt.device)
        rc = _lib.hip().bfly_c2c_ex(xt.data_ptr(), y.data_ptr(), n, b, sign, variant, tab.data_ptr(),
                                    None, torch.cuda.current_stream().cuda_stream)
        assert rc == 0
        torch.cuda.synchronize()
        gate_c2c(y.cpu().numpy(), x, n, sign)


def _run_variant(x, n, b, sign, variant):
    import torch
    from bfly import _lib
    xt = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    y = torch.empty_like(xt)
    tab = _ops().table(0, n, xt.device)
    rc = _lib.hip().bfly_c2c_ex(xt.data_ptr(), y.data_ptr(), n, b, sign, variant, tab.data_ptr(),
                                None, torch.cuda.current_stream().cuda_stream)
    assert rc == 0
    torch.cuda.synchronize()
    return y.cpu().numpy()


@pytest.mark.parametrize("n", [1024, 2048, 4096])
def test_split_mfma_variant_error_vs_fp64(n):
    # Radix-16 passes on the matrix cores in split FP16: same FP64 gates as
    # the FP32 paths; its error is stated against FP64 and stays within a
    # few times the FP32 Stockham's (about 1.3e-7 at these lengths).
    from bfly import _lib
    for b, sign in ((1, -1), (7, 1), (64, -1), (1000, 1), (3001, -1)):
        x = oracle.uniform_input(21, (b, 2 * n))
        got = _run_variant(x, n, b, sign, _lib.VARIANT_SPLIT_MFMA)
        e = gate_c2c(got, x, n, sign)
        assert e < 2e-6, f"n={n} b={b} rel-L2 {e:.3e}"
        # the resident-grid form runs the same per-transform code
        got_s = _run_variant(x, n, b, sign, _lib.VARIANT_SPLIT_MFMA_STREAM)
        assert np.array_equal(got_s, got), (n, b, sign)


@pytest.mark.timeout(300)
def test_cfg3_split_mfma_n4096_b4096_seed2():
    """The precision pattern's own workload: N=4096, batch=4096, U[-1,1)
    from seed 2, through both split-FP16 MFMA kernels, every row against the
    FP64 oracle.  Measured worst per-transform relative L2 error: about
    4.1e-7 (FP32 Stockham on the same rows: about 1.3e-7); the gate is the
    north-star 1e-5, and the two kernels give the same bits."""
    from bfly import _lib
    n, b = 4096, 4096
    x = oracle.uniform_input(2, (b, 2 * n))
    ref = oracle.c2c_f64(x, n)
    got = _run_variant(x, n, b, -1, _lib.VARIANT_SPLIT_MFMA)
    e = oracle.rel_l2_rows(got, ref)
    print(f"split MFMA 4096 x 4096 seed 2: worst rel-L2 {e:.3e}")
    assert e <= REL_L2 and e < 1e-6, e
    assert oracle.max_rel_scaled(got, ref.astype(np.float32)) <= MAX_REL_SCALED
    got_s = _run_variant(x, n, b, -1, _lib.VARIANT_SPLIT_MFMA_STREAM)
    assert np.array_equal(got_s, got)


@pytest.mark.parametrize("n", [1024, 4096])
@pytest.mark.parametrize("scale", [1e-30, 1e-6, 1.0, 1e6, 1e30])
def test_split_mfma_any_magnitude(n, scale):
    # Per-butterfly power-of-two scaling keeps both FP16 halves in range for
    # data far outside FP16's own range.
    from bfly import _lib
    b = 9
    x = (oracle.uniform_input(22, (b, 2 * n)).astype(np.float64) * scale).astype(np.float32)
    got = _run_variant(x, n, b, -1, _lib.VARIANT_SPLIT_MFMA)
    gate_c2c(got, x, n)


@pytest.mark.parametrize("mode", sorted(oracle.NUMERICAL_MODES))
def test_split_mfma_numerical_modes(mode):
    from bfly import _lib
    n, b = 2048, 3
    x = oracle.numerical_input(n, b, mode, seed=1)
    got = _run_variant(x, n, b, -1, _lib.VARIANT_SPLIT_MFMA)
    ref = oracle.c2c_f64(x, n).astype(np.float32)
    assert oracle.max_rel_scaled(got, ref) <= MAX_REL_SCALED


def test_split_mfma_rejects_other_lengths():
    import torch
    from bfly import _lib
    x = torch.zeros(2 * 512, device="cuda")
    tab = _ops().table(0, 512, x.device)
    assert _lib.hip().bfly_c2c_ex(x.data_ptr(), x.data_ptr(), 512, 1, -1, _lib.VARIANT_SPLIT_MFMA,
                                  tab.data_ptr(), None, 0) == _lib.ERR_SHAPE


@pytest.mark.parametrize("n", [32, 64, 128, 256, 1024, 4096, 8192])
def test_real_streaming_variant_equals_one_shot(n):
    # bfly_real_ex STREAM runs the same arithmetic as PLAIN on a resident
    # grid: outputs equal bit for bit (ragged and multi-round batches,
    # full and mode-truncated spectra), and both pass the FP64 gate.
    import torch
    from bfly import _lib
    lib = _lib.hip()
    s = torch.cuda.current_stream().cuda_stream
    for b, bins in ((1, n // 2 + 1), (333, n // 2 + 1), (2000, 7)):
        x = oracle.uniform_input(31, (b, n))
        xt = torch.from_numpy(x).cuda()
        tab = _ops().table(1, n, xt.device)
        outs = []
        for v in (_lib.VARIANT_PLAIN, _lib.VARIANT_STREAM):
            y = torch.full((b, 2 * bins), float("nan"), device="cuda")
            assert lib.bfly_real_ex(1, xt.data_ptr(), y.data_ptr(), n, b, bins, 2 * bins, v, 0,
                                    tab.data_ptr(), s) == 0
            outs.append(y)
        torch.cuda.synchronize()
        assert torch.equal(outs[0], outs[1]), (n, b, bins)
        ref = oracle.r2c_f64(x, n)[:, :2 * bins]
        assert oracle.rel_l2_rows(outs[1].cpu().numpy(), np.ascontiguousarray(ref)) <= REL_L2
        h = torch.zeros(b, n + 2, device="cuda")
        h[:, :2 * bins] = outs[1]
        h[:, 1] = 0
        if bins == n // 2 + 1:
            h[:, n + 1] = 0
        tab2 = _ops().table(2, n, xt.device)
        backs = []
        for v in (_lib.VARIANT_PLAIN, _lib.VARIANT_STREAM):
            z = torch.full((b, n), float("nan"), device="cuda")
            assert lib.bfly_real_ex(2, h.data_ptr(), z.data_ptr(), n, b, bins, n + 2, v, 0,
                                    tab2.data_ptr(), s) == 0
            backs.append(z)
        torch.cuda.synchronize()
        assert torch.equal(backs[0], backs[1]), (n, b, bins)
        assert oracle.rel_l2_rows(backs[1].cpu().numpy(), oracle.c2r_f64(h.cpu().numpy(), n)) <= REL_L2
    # ids other than AUTO / PLAIN / STREAM / STREAM_LT are refused
    assert lib.bfly_real_ex(1, xt.data_ptr(), xt.data_ptr(), n, 1, n // 2 + 1, n + 2,
                            _lib.VARIANT_STREAM2, 0, tab.data_ptr(), s) == _lib.ERR_SHAPE


@pytest.mark.parametrize("n", [256, 1024, 4096, 8192])
def test_real_lds_twiddle_variant_fp64(n):
    # bfly_real_ex STREAM_LT: twiddles from two-level LDS tables (different
    # rounding from the one-shot kernel); FP64 gates, round trip, truncation
    # (13 bins, and 3n/8 + 5 bins: from n = 2048 bins j and n/2 - j are
    # formed together, so a cut above n/4 keeps some mirrored bins).
    import torch
    from bfly import _lib
    lib = _lib.hip()
    s = torch.cuda.current_stream().cuda_stream
    for b, bins in ((1, n // 2 + 1), (777, n // 2 + 1), (3000, 13), (101, 3 * n // 8 + 5)):
        x = oracle.uniform_input(33, (b, n))
        xt = torch.from_numpy(x).cuda()
        y = torch.full((b, 2 * bins), float("nan"), device="cuda")
        assert lib.bfly_real_ex(1, xt.data_ptr(), y.data_ptr(), n, b, bins, 2 * bins, _lib.VARIANT_STREAM_LT, 0,
                                _ops().table(1, n, xt.device).data_ptr(), s) == 0
        torch.cuda.synchronize()
        ref = oracle.r2c_f64(x, n)[:, :2 * bins]
        assert oracle.rel_l2_rows(y.cpu().numpy(), np.ascontiguousarray(ref)) <= REL_L2
        h = torch.zeros(b, n + 2, device="cuda")
        h[:, :2 * bins] = y
        h[:, 1] = 0
        if bins == n // 2 + 1:
            h[:, n + 1] = 0
        z = torch.full((b, n), float("nan"), device="cuda")
        assert lib.bfly_real_ex(2, h.data_ptr(), z.data_ptr(), n, b, bins, n + 2, _lib.VARIANT_STREAM_LT, 0,
                                _ops().table(2, n, xt.device).data_ptr(), s) == 0
        torch.cuda.synchronize()
        assert oracle.rel_l2_rows(z.cpu().numpy(), oracle.c2r_f64(h.cpu().numpy(), n)) <= REL_L2
        if bins == n // 2 + 1:
            assert np.max(np.abs(z.cpu().numpy() - x)) < 1e-5


@pytest.mark.parametrize("n,b", [(16, 3), (256, 97), (4096, 40), (8192, 5), (16384, 64), (16384, 256),
                                 (32768, 3), (65536, 2)])
def test_c2c_scaled_equals_its_variant_times_scale_bit_for_bit(n, b):
    # bfly_c2c_scaled multiplies in the last store of the variant it runs
    # (LDS-twiddle streaming up to 8192, AUTO above: the one-workgroup
    # kernel, or the four-step chain at 16384 b=64 and 32768 / 65536), so
    # its output is fl(scale * that variant's output) bit for bit.
    import torch
    from bfly import _lib
    ops, lib = _ops(), _lib.hip()
    x = torch.from_numpy(oracle.uniform_input(11, (b, 2 * n))).cuda()
    tab = ops.table(_lib.KIND_C2C, n, x.device)
    work = ops.workspace(_lib.KIND_C2C, n, b, x.device)
    wp = work.data_ptr() if work is not None else None
    st = torch.cuda.current_stream().cuda_stream
    variant = _lib.VARIANT_STREAM_LT if n <= 8192 else _lib.VARIANT_AUTO
    for sign, scale in ((1, 1.0 / n), (-1, 1.0 / np.sqrt(n)), (1, -3.0), (-1, 1.0)):
        ref = torch.empty_like(x)
        got = torch.full_like(x, float("nan"))
        assert lib.bfly_c2c_ex(x.data_ptr(), ref.data_ptr(), n, b, sign, variant, tab.data_ptr(), wp, st) == 0
        assert lib.bfly_c2c_scaled(x.data_ptr(), got.data_ptr(), n, b, sign, scale, tab.data_ptr(), wp, st) == 0
        torch.cuda.synchronize()
        want = ref.cpu().numpy() * np.float32(scale)
        assert np.array_equal(got.cpu().numpy(), want), (n, b, sign, scale)
        if sign == 1 and scale == 1.0 / n:  # the normalised inverse against FP64
            gate_c2c(got.cpu().numpy() * n, x.cpu().numpy(), n, sign=+1)


@pytest.mark.parametrize("n,b,ipad,opad", [(16, 5, 1, 3), (256, 97, 8, 0), (4096, 40, 3, 64), (8192, 6, 1, 1),
                                           (16384, 3, 2, 0), (1024, 1, 5, 5), (32768, 3, 4, 2), (65536, 2, 1, 0)])
def test_c2c_rows_padded_equals_contiguous_bit_for_bit(n, b, ipad, opad):
    # bfly_c2c_rows over rows n + ipad apart into rows n + opad apart: the
    # LDS-twiddle streaming kernel (up to 16384) or the four-step chain
    # (32768 / 65536) with row distances, so it equals the contiguous call
    # of the same kernels bit for bit.  NaN in the input gaps would reach
    # the result if a gap were read; the output gaps keep their fill.
    import torch
    from bfly import _lib
    ops, lib = _ops(), _lib.hip()
    x = oracle.uniform_input(13, (b, 2 * n))
    xt = torch.from_numpy(x).cuda().view(torch.complex64)
    nan = complex(float("nan"), float("nan"))
    ibuf = torch.full((b, n + ipad), nan, dtype=torch.complex64, device="cuda")
    ibuf[:, :n] = xt
    tab = ops.table(_lib.KIND_C2C, n, xt.device)
    st = torch.cuda.current_stream().cuda_stream
    for sign in (-1, 1):
        obuf = torch.full((b, n + opad), complex(7.0, -7.0), dtype=torch.complex64, device="cuda")
        ops.c2c_rows(ibuf[:, :n], sign, out=obuf[:, :n])
        ref = torch.empty_like(xt)
        work = ops.workspace(_lib.KIND_C2C, n, b, xt.device) if n > 16384 else None
        variant = _lib.VARIANT_STREAM_LT if n <= 16384 else _lib.VARIANT_FOUR_STEP
        assert lib.bfly_c2c_ex(xt.data_ptr(), ref.data_ptr(), n, b, sign, variant, tab.data_ptr(),
                               work.data_ptr() if work is not None else None, st) == 0
        torch.cuda.synchronize()
        got = torch.view_as_real(obuf[:, :n]).cpu()
        assert torch.equal(got, torch.view_as_real(ref).cpu()), (n, b, sign)
        if opad:
            assert bool((obuf[:, n:] == complex(7.0, -7.0)).all())
        gate_c2c(got.numpy().reshape(b, 2 * n), x, n, sign=sign)
    # a new contiguous result when no `out` is given
    y = ops.c2c_rows(ibuf[:, :n])
    torch.cuda.synchronize()
    gate_c2c(torch.view_as_real(y).cpu().numpy().reshape(b, 2 * n), x, n)


@pytest.mark.parametrize("norm", [None, "backward", "ortho", "forward"])
@pytest.mark.parametrize("n,b", [(64, 7), (4096, 16), (16384, 4), (65536, 2)])
def test_c2c_norm_matches_numpy_and_round_trips(n, b, norm):
    # ops.c2c(norm=...) against numpy.fft's own normalisation of the FP64
    # transform, and forward then inverse under one convention returns x.
    import torch
    x = oracle.uniform_input(12, (b, 2 * n))
    xc = x.astype(np.float64).view(np.complex128)
    t = torch.from_numpy(x).cuda().view(torch.complex64)
    f = _ops().c2c(t, -1, norm=norm)
    back = _ops().c2c(f, +1, norm=norm)
    torch.cuda.synchronize()
    ref = np.fft.fft(xc, axis=-1, norm=norm)
    got = f.cpu().numpy().astype(np.complex128)
    e = np.sqrt(np.sum(np.abs(got - ref) ** 2, axis=-1) / np.sum(np.abs(ref) ** 2, axis=-1)).max()
    assert e <= REL_L2, f"n={n} norm={norm} rel-L2 {e:.3e}"
    assert np.max(np.abs(back.cpu().numpy() - xc)) < 1e-5


@pytest.mark.parametrize("n", [16, 32, 64, 1024, 4096])
def test_c2c_inverse_is_conjugate_transform(n):
    b = 5
    x = oracle.uniform_input(7, (b, 2 * n))
    got = run_c2c(x, n, sign=+1)
    gate_c2c(got, x, n, sign=+1)
    back = run_c2c(run_c2c(x, n), n, sign=+1) / n
    assert np.max(np.abs(back - x)) < 1e-5


@pytest.mark.parametrize("n", [8192, 16384, 32768, 65536])
def test_long_c2c_b64_seed2(n):
    b = 64
    x = oracle.uniform_input(2, (b, 2 * n))
    got = run_c2c(x, n)
    gate_c2c(got, x, n)


@pytest.mark.parametrize("n", [8192, 16384])
@pytest.mark.parametrize("b", [1, 5, 64, 700])
def test_single_workgroup_and_four_step_paths_agree(n, b):
    # n = 8192 / 16384 as one 512 / 1024-thread workgroup per transform
    # (BFLY_VARIANT_PLAIN, no workspace) and as the two-kernel four-step path
    # (BFLY_VARIANT_FOUR_STEP); AUTO picks per batch.  All pass the FP64 gates
    # on the same seeded input.
    import torch
    from bfly import _lib
    ops = _ops()
    x = oracle.uniform_input(2, (b, 2 * n))
    t = torch.from_numpy(x).cuda()
    tab = ops.table(_lib.KIND_C2C, n, t.device)
    work = torch.empty(n * b * 2, dtype=torch.float32, device="cuda")
    outs = []
    for variant, w in ((_lib.VARIANT_PLAIN, None), (_lib.VARIANT_FOUR_STEP, work.data_ptr()),
                       (_lib.VARIANT_STREAM_LT, None), (_lib.VARIANT_AUTO, work.data_ptr())):
        y = torch.empty_like(t)
        rc = _lib.hip().bfly_c2c_ex(t.data_ptr(), y.data_ptr(), n, b, -1, variant, tab.data_ptr(),
                                    w, torch.cuda.current_stream().cuda_stream)
        _lib.check(rc, "bfly_c2c_ex")
        torch.cuda.synchronize()
        outs.append(y.cpu().numpy())
        gate_c2c(outs[-1], x, n)
    # the four-step variant refuses to run without its workspace
    y = torch.empty_like(t)
    assert _lib.hip().bfly_c2c_ex(t.data_ptr(), y.data_ptr(), n, b, -1, _lib.VARIANT_FOUR_STEP,
                                  tab.data_ptr(), None, 0) == _lib.ERR_SHAPE


@pytest.mark.parametrize("sign", [-1, 1])
def test_streamed_long_column_pass_equals_the_one_shot_pass(sign):
    # At n = 65536 the four-step column pass streams over its column groups
    # once there is a group for every resident workgroup (batch >= 64 here);
    # smaller batches run the one-shot column kernel.  The per-column
    # arithmetic is the same: the batch of 64 in one call equals eight calls
    # of 8 transforms bit for bit, and passes the FP64 gates.
    import torch
    n, b = 65536, 64
    x = oracle.uniform_input(7, (b, 2 * n))
    got = run_c2c(x, n, sign)
    parts = np.concatenate([run_c2c(x[i:i + 8], n, sign) for i in range(0, b, 8)])
    assert np.array_equal(got, parts)
    rows = list(range(0, b, 9))
    gate_c2c(got[rows], x[rows], n, sign)
    torch.cuda.synchronize()


def test_long_c2c_matches_host_four_step_split():
    # The reference's host four-step chain (src/host/fft_check.cpp:904-963) in
    # FP64 with its own split (fft_check.cpp:295-308) agrees with the direct
    # FFT; the device result must agree with both.
    n, b = 8192, 3
    x = oracle.uniform_input(5, (b, 2 * n))
    n1, n2 = oracle.long_split(n)
    four = oracle.four_step_f64(x, n, n1, n2)
    direct = oracle.c2c_f64(x, n)
    assert np.max(np.abs(four - direct)) / np.max(np.abs(direct)) < 1e-12
    got = run_c2c(x, n)
    assert oracle.rel_l2_rows(got, four) <= REL_L2


@pytest.mark.parametrize("n", [64, 1024, 4096, 65536])
@pytest.mark.parametrize("mode", sorted(oracle.NUMERICAL_MODES))
def test_numerical_modes(n, mode):
    # src/host/fft_check.cpp:83-116 input modes, reference gate maxRelScaled.
    b = 3
    x = oracle.numerical_input(n, b, mode, seed=1)
    got = run_c2c(x, n)
    ref = oracle.c2c_f64(x, n).astype(np.float32)
    assert oracle.max_rel_scaled(got, ref) <= MAX_REL_SCALED
    if mode == "impulse":
        for row in range(b):
            k = np.arange(n)
            want = np.exp(-2j * np.pi * k * (row % n) / n)
            g = got[row, 0::2] + 1j * got[row, 1::2]
            assert np.max(np.abs(g - want)) < 1e-5
    if mode == "constant":
        g = got[:, 0::2] + 1j * got[:, 1::2]
        assert np.allclose(g[:, 0], n * (0.25 - 0.5j), rtol=1e-6)
        assert np.max(np.abs(g[:, 1:])) < 1e-6 * n
    if mode == "zero":
        assert not np.any(got)


@pytest.mark.parametrize("mode", ["ofdm", "radar", "dl"])
def test_app_inputs(mode):
    for n in (256, 4096):
        x = oracle.app_input(n, 5, mode)
        got = run_c2c(x, n)
        ref = oracle.c2c_f64(x, n).astype(np.float32)
        assert oracle.max_rel_scaled(got, ref) <= MAX_REL_SCALED


@pytest.mark.parametrize("b", [1, 2, 3, 47, 48, 49, 95, 97, 143, 193])
def test_ragged_batches_all_rows(b):
    # tests/test_limits.cpp:249-261: odd batches around workgroup tails, all rows.
    for n in (64, 1024):
        x = oracle.pattern(b * 2 * n).reshape(b, 2 * n)
        got = run_c2c(x, n)
        gate_c2c(got, x, n)


def test_pattern_input_reference_gate():
    # The reference's own smoke (tests/test_framework.cpp:109-137): patternAt
    # input, n=4096, batch=64, sampled rows against refFftF32.
    n, b = 4096, 64
    x = oracle.pattern(b * 2 * n).reshape(b, 2 * n)
    got = run_c2c(x, n)
    for row in (0, 9, 18, 27, 36, 45, 54, 63):
        ref = oracle.ref_fft_f32(x[row], n)[0]
        assert oracle.max_rel_scaled(got[row], ref) <= MAX_REL_SCALED


# ----------------------------------------------------------------- R2C / C2R
def run_r2c(x, n):
    import torch
    y = _ops().rfft(torch.from_numpy(np.ascontiguousarray(x)).cuda())
    torch.cuda.synchronize()
    return y.cpu().numpy()


def run_c2r(h, n):
    import torch
    y = _ops().irfft(torch.from_numpy(np.ascontiguousarray(h)).cuda(), n)
    torch.cuda.synchronize()
    return y.cpu().numpy()


@pytest.mark.parametrize("n", [128, 256, 512, 1024, 2048, 4096, 8192])
def test_r2c_grid_seed1(n):
    for b in GRID_B:
        x = oracle.uniform_input(1, (b, n))
        got = run_r2c(x, n)
        assert got.shape == (b, n + 2)
        ref = oracle.r2c_f64(x, n)
        e = oracle.rel_l2_rows(got, ref)
        assert e <= REL_L2, f"r2c n={n} b={b} rel-L2 {e:.3e}"
        assert oracle.max_rel_scaled(got, ref.astype(np.float32)) <= MAX_REL_SCALED


@pytest.mark.parametrize("n", [64, 128, 256, 512, 1024, 2048, 4096])
def test_c2r_grid_seed1(n):
    # Half spectra built as the R2C of seeded reals (the grid's C2R inputs).
    for b in GRID_B:
        x = oracle.uniform_input(1, (b, n))
        h = oracle.r2c_f64(x, n).astype(np.float32)
        got = run_c2r(h, n)
        assert got.shape == (b, n)
        ref = oracle.c2r_f64(h, n)
        e = oracle.rel_l2_rows(got, ref)
        assert e <= REL_L2, f"c2r n={n} b={b} rel-L2 {e:.3e}"
        assert np.max(np.abs(got - x)) < 1e-5  # round trip


@pytest.mark.parametrize("n", [32, 64, 128, 4096, 8192])
def test_c2r_arbitrary_half_spectrum(n):
    # src/host/fft_check.cpp:179-192 half spectra: DC imaginary part nonzero,
    # Nyquist imaginary 0; the reference's expansion ignores both imaginary
    # parts (src/framework/reference.cpp:64-79).
    b = 3
    h = oracle.half_input(n, b)
    got = run_c2r(h, n)
    ref = oracle.c2r_f64(h, n)
    assert oracle.rel_l2_rows(got, ref) <= REL_L2
    for row in range(b):
        r = np.empty(n, np.float32)
        oracle.lib().orc_ref_c2r_f32(np.ascontiguousarray(h[row]), r, n)
        assert oracle.max_rel_scaled(got[row], r) <= MAX_REL_SCALED


@pytest.mark.parametrize("n", [64, 128, 4096, 8192])
def test_real_boundary_round_trip(n):
    # tests/test_framework.cpp:16-87: R2C/C2R at the envelope ends, B=3,
    # patternAt input.
    b = 3
    x = oracle.pattern(b * n).reshape(b, n)
    if n >= 128:
        h = run_r2c(x, n)
        for row in range(b):
            r = np.empty(n + 2, np.float32)
            oracle.lib().orc_ref_r2c_f32(np.ascontiguousarray(x[row]), r, n)
            assert oracle.max_rel_scaled(h[row], r) <= MAX_REL_SCALED
    else:
        h = oracle.pattern(b * (n + 2)).reshape(b, n + 2).copy()
        h[:, 1] = 0
        h[:, n + 1] = 0
    back = run_c2r(h, n)
    for row in range(b):
        r = np.empty(n, np.float32)
        oracle.lib().orc_ref_c2r_f32(np.ascontiguousarray(h[row]), r, n)
        assert oracle.max_rel_scaled(back[row], r) <= MAX_REL_SCALED
        if n >= 128:
            assert oracle.max_rel_scaled(back[row], x[row]) <= MAX_REL_SCALED
