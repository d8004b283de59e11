"""GPU: the largest shapes of the reference's 32-bit offset guard
(batch * 2n <= 2^32 - 1, src/framework/butterfly.cpp:729) and of the long
path, checked through properties that do not depend on the size plus
sampled rows against the FP64 oracle:

* linearity: the sum over all rows of the outputs equals the transform of
  the sum of the inputs (both sums in FP64, the latter transformed by the
  FP64 oracle), relative error <= 1e-5;
* rows spread over the batch (first, last and evenly between, the
  reference's sampleBatches) at the usual gate, relative L2 <= 1e-5.

Inputs are drawn on the device (U[-1,1), seeded torch generator)."""
import numpy as np
import pytest

import oracle
from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]

REL = 1e-5


def _rows(batch, want=9):
    idx = np.empty(want, np.uint32)
    k = oracle.lib().orc_sample_batches(batch, want, idx, want)
    return [int(i) for i in idx[:k]]


def _fp64_row_sum(t, chunk=1 << 20):
    """Sum over dim 0 of a 2-D device tensor, in FP64, chunk by chunk."""
    import torch
    acc = torch.zeros(t.shape[1], dtype=torch.float64, device=t.device)
    for i in range(0, t.shape[0], chunk):
        acc += t[i:i + chunk].double().sum(0)
    return acc.cpu().numpy()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n,batch", [(64, (2**32 - 1) // 128), (4096, (2**32 - 1) // 8192),
                                     (65536, (2**32 - 1) // 131072)])
def test_c2c_at_the_offset_limit(dev, n, batch):
    import torch
    from bfly import ops
    print(f"c2c n={n} batch={batch}", flush=True)
    g = torch.Generator(device=dev).manual_seed(n)
    x = torch.rand(batch, 2 * n, generator=g, device=dev) * 2 - 1
    y = torch.empty_like(x)
    ops.c2c_raw(x, y, n, batch)
    torch.cuda.synchronize()
    want = oracle.fft_f64(_fp64_row_sum(x)[None, :], n)[0]
    assert _rel(_fp64_row_sum(y), want) <= REL
    rows = _rows(batch)
    xs, ys = x[rows].cpu().numpy(), y[rows].cpu().numpy()
    assert oracle.rel_l2_rows(ys, oracle.c2c_f64(xs, n)) <= REL
    del x, y
    torch.cuda.empty_cache()


@pytest.mark.parametrize("n,batch", [(8192, (2**32 - 1) // 16384), (4096, 1 << 18)])
def test_real_round_trip_at_scale(dev, n, batch):
    import torch
    from bfly import ops
    print(f"r2c/c2r n={n} batch={batch}", flush=True)
    g = torch.Generator(device=dev).manual_seed(n + 1)
    x = torch.rand(batch, n, generator=g, device=dev) * 2 - 1
    h = ops.rfft(x)
    torch.cuda.synchronize()
    # linearity of the R2C: sum of spectra = R2C of the summed rows (FP64)
    s = _fp64_row_sum(x)
    z = np.zeros(2 * n)
    z[0::2] = s
    want = oracle.fft_f64(z[None, :], n)[0][: n + 2]
    assert _rel(_fp64_row_sum(h), want) <= REL
    rows = _rows(batch)
    assert oracle.rel_l2_rows(h[rows].cpu().numpy(), oracle.r2c_f64(x[rows].cpu().numpy(), n)) <= REL
    back = ops.irfft(h, n)
    torch.cuda.synchronize()
    err = (back - x).abs().max().item()
    assert err < 2e-5, err
    del x, h, back
    torch.cuda.empty_cache()
