"""GPU: the 2-D C2C transform (bfly_fft2d; bfly.ops.fft2) against the FP64
oracle (relative L2 <= 1e-5 per 2-D transform) and torch.fft.fft2, for the
one-pass column lengths (n1 <= 256), the split ones (512..4096), both signs,
ragged batches, and the reference's declared 2-D shapes (fft2d 256..2048
squares, range-Doppler 1024x128 / 2048x256 / 4096x256)."""
import numpy as np
import pytest

import oracle
from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]

REL_L2 = 1e-5

SHAPES = [
    (16, 16, 3), (16, 4096, 2), (32, 64, 5), (64, 128, 4), (128, 16, 7), (256, 256, 4), (256, 4096, 1),
    (512, 512, 2), (512, 16, 3), (1024, 128, 4), (1024, 1024, 1), (2048, 256, 2), (2048, 2048, 1),
    (4096, 256, 1), (4096, 16, 2), (16, 1024, 9),
]


def _run(dev, x, sign):
    import torch
    from bfly import ops
    xt = torch.from_numpy(x).to(dev)
    b, n1, n2 = x.shape
    y = ops.fft2(xt, sign)
    torch.cuda.synchronize()
    return y.cpu().numpy()


@pytest.mark.parametrize("sign", [-1, 1])
@pytest.mark.parametrize("n1,n2,batch", SHAPES, ids=[f"{a}x{b}-b{c}" for a, b, c in SHAPES])
def test_fft2_matches_the_oracle(dev, n1, n2, batch, sign):
    print(f"fft2 n1={n1} n2={n2} batch={batch} sign={sign}", flush=True)
    rng = np.random.default_rng(n1 * 7919 + n2 + batch)
    z = rng.uniform(-1, 1, (batch, n1, n2, 2)).astype(np.float32)
    x = z[..., 0] + 1j * z[..., 1]
    y = _run(dev, x.astype(np.complex64), sign)
    want = oracle.fft2_f64(z.reshape(batch, n1, 2 * n2), n1, n2, sign)
    got = np.ascontiguousarray(y).view(np.float32).reshape(batch, -1)
    assert oracle.rel_l2_rows(got, want) <= REL_L2


def test_fft2_matches_torch_and_special_inputs(dev):
    import torch
    from bfly import ops
    g = torch.Generator(device="cpu").manual_seed(4)
    for n1, n2 in ((1024, 128), (256, 256), (4096, 16)):
        x = torch.complex(torch.rand(3, n1, n2, generator=g) * 2 - 1,
                          torch.rand(3, n1, n2, generator=g) * 2 - 1).to(dev)
        y = ops.fft2(x)
        ref = torch.fft.fft2(x)
        rel = (torch.linalg.vector_norm(y - ref) / torch.linalg.vector_norm(ref)).item()
        assert rel < 3e-6, (n1, n2, rel)
    # impulse at (a, c): e^{-2 pi i (a k1 / n1 + c k2 / n2)}; constant: n1 n2 at the origin
    n1, n2, a, c = 2048, 64, 77, 5
    x = torch.zeros(1, n1, n2, dtype=torch.complex64, device=dev)
    x[0, a, c] = 1
    y = ops.fft2(x)[0].cpu().numpy()
    k1, k2 = np.meshgrid(np.arange(n1), np.arange(n2), indexing="ij")
    assert np.abs(y - np.exp(-2j * np.pi * (a * k1 / n1 + c * k2 / n2))).max() < 2e-6
    y = ops.fft2(torch.ones(2, 512, 32, dtype=torch.complex64, device=dev)).cpu().numpy()
    assert y[:, 0, 0].real.tolist() == [512 * 32] * 2 and np.abs(y.reshape(2, -1)[:, 1:]).max() < 1e-2


def test_fft2_round_trip_and_guards(dev):
    import ctypes as C

    import torch
    from bfly import _lib, ops
    x = torch.complex(torch.randn(2, 512, 256), torch.randn(2, 512, 256)).to(dev)
    back = ops.fft2(ops.fft2(x), sign=1) / (512 * 256)
    assert (torch.linalg.vector_norm(back - x) / torch.linalg.vector_norm(x)).item() < 2e-6
    lib = _lib.hip()
    buf = torch.empty(1 << 20, dtype=torch.float32, device=dev)
    p = buf.data_ptr()
    t = ops.table(_lib.KIND_C2C, 64, dev).data_ptr()
    tc = ops.table2d_cols(64, dev).data_ptr()
    s = torch.cuda.current_stream(dev).cuda_stream
    # shapes outside [16, 4096] or not powers of two, overlapping buffers, bad sign
    for n1, n2 in ((8, 64), (64, 8192), (96, 64)):
        assert lib.bfly_fft2d(p, p + (1 << 19), n1, n2, 1, -1, t, tc, p + (2 << 19), s) == _lib.ERR_SHAPE
    assert lib.bfly_fft2d(p, p + 8, 64, 64, 1, -1, t, tc, p + (2 << 19), s) == _lib.ERR_SHAPE
    assert lib.bfly_fft2d(p, p + (1 << 19), 64, 64, 1, 0, t, tc, p + (2 << 19), s) == _lib.ERR_SHAPE
    assert lib.bfly_fft2d(p, p + (1 << 19), 64, 64, 1, -1, None, tc, p + (2 << 19), s) == _lib.ERR_TABLE


COLS = [(16, 1, 3), (64, 3, 5), (128, 1000, 1), (256, 100, 2), (512, 7, 4), (1024, 1000, 1), (2048, 64, 1),
        (4096, 33, 2), (4096, 1, 1)]


@pytest.mark.parametrize("sign", [-1, 1])
@pytest.mark.parametrize("n,ncols,batch", COLS, ids=[f"{a}x{b}-b{c}" for a, b, c in COLS])
def test_strided_columns_match_the_oracle(dev, n, ncols, batch, sign):
    import torch
    from bfly import ops
    print(f"columns n={n} ncols={ncols} batch={batch} sign={sign}", flush=True)
    rng = np.random.default_rng(n + 31 * ncols + batch)
    z = rng.uniform(-1, 1, (batch, n, ncols, 2)).astype(np.float32)
    x = torch.from_numpy(z[..., 0] + 1j * z[..., 1]).to(torch.complex64).to(dev)
    y = ops.fft_columns(x, sign)
    torch.cuda.synchronize()
    # each column as an interleaved row for the oracle
    rows = np.ascontiguousarray(z.transpose(0, 2, 1, 3)).reshape(batch * ncols, 2 * n)
    want = oracle.c2c_f64(rows, n, sign)
    got = np.ascontiguousarray(y.cpu().numpy().transpose(0, 2, 1)).view(np.float32).reshape(batch * ncols, 2 * n)
    assert oracle.rel_l2_rows(got, want) <= REL_L2
    if sign < 0:
        ref = torch.fft.fft(x, dim=-2)
        assert (torch.linalg.vector_norm(y - ref) / torch.linalg.vector_norm(ref)).item() < 3e-6


@pytest.mark.parametrize("sign", [-1, 1])
def test_streamed_column_pass_equals_the_one_shot_pass(dev, sign):
    """1024-point columns with a full column group for every resident
    workgroup run the streaming column kernel (the next group prefetched);
    fewer groups run the one-shot kernel.  The per-column arithmetic is the
    same, so strided columns transformed in one call (256 groups: streamed)
    equal the same columns transformed in 512-wide slices (64 groups each:
    one-shot) bit for bit; 2-D transforms whose column pass streams pass the
    FP64 gate."""
    import torch
    from bfly import ops
    g = torch.Generator(device="cpu").manual_seed(11 + sign)
    for n1, n2, batch in ((1024, 1024, 5), (1024, 128, 40)):
        print(f"fft2 streamed columns n1={n1} n2={n2} batch={batch} sign={sign}", flush=True)
        x = torch.complex(torch.rand(batch, n1, n2, generator=g) * 2 - 1,
                          torch.rand(batch, n1, n2, generator=g) * 2 - 1).to(dev)
        y = ops.fft2(x, sign)
        torch.cuda.synchronize()
        z = torch.view_as_real(x).cpu().numpy().reshape(batch, n1, 2 * n2)
        want = oracle.fft2_f64(z, n1, n2, sign)
        got = np.ascontiguousarray(torch.view_as_real(y).cpu().numpy()).reshape(batch, -1)
        assert oracle.rel_l2_rows(got, want) <= REL_L2
    n, ncols, batch = 1024, 2048, 2
    print(f"columns streamed n={n} ncols={ncols} batch={batch} sign={sign}", flush=True)
    x = torch.complex(torch.rand(batch, n, ncols, generator=g) * 2 - 1,
                      torch.rand(batch, n, ncols, generator=g) * 2 - 1).to(dev)
    y = ops.fft_columns(x, sign)
    # column blocks of 512 run with 32 groups each: the one-shot kernel
    parts = torch.cat([ops.fft_columns(x[..., j:j + 512].contiguous(), sign) for j in range(0, ncols, 512)], dim=-1)
    torch.cuda.synchronize()
    assert torch.equal(y, parts)
    xc = x.cpu().to(torch.complex128)
    ref = torch.fft.fft(xc, dim=-2) if sign < 0 else torch.fft.ifft(xc, dim=-2) * n  # unnormalised inverse
    rel = (torch.linalg.vector_norm(y.cpu().to(torch.complex128) - ref) / torch.linalg.vector_norm(ref)).item()
    assert rel <= REL_L2

