"""GPU: bfly_c2c_noalloc -- the C2C entry point that allocates nothing (its
four-step intermediate lives in a static device array of
BFLY_NOALLOC_SCRATCH_BYTES per GPU) -- gives bfly_c2c's results bit for bit,
through the C ABI, on calls in a row where a later call is smaller than an
earlier one (the array keeps the last call's contents), and refuses a shape
whose intermediate does not fit with BFLY_ERR_SCRATCH."""
import ctypes as C

import numpy as np
import pytest

import oracle
from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]


def _both(n, b, sign, seed):
    """(noalloc output, bfly_c2c output with its own workspace, input)."""
    import torch
    from bfly import _lib, ops
    lib = _lib.hip()
    x = torch.from_numpy(oracle.uniform_input(seed, (b, 2 * n))).cuda()
    tab = ops.table(_lib.KIND_C2C, n, x.device)
    s = torch.cuda.current_stream().cuda_stream
    y_na = torch.full_like(x, float("nan"))
    rc = lib.bfly_c2c_noalloc(x.data_ptr(), y_na.data_ptr(), n, b, sign, tab.data_ptr(), s)
    assert rc == _lib.OK, (n, b, rc)
    y_ref = torch.full_like(x, float("nan"))
    ws = lib.bfly_workspace_bytes(_lib.KIND_C2C, n, b)
    work = torch.empty(max(1, ws // 4), dtype=torch.float32, device=x.device) if ws else None
    rc = lib.bfly_c2c(x.data_ptr(), y_ref.data_ptr(), n, b, sign, tab.data_ptr(),
                      work.data_ptr() if work is not None else None, s)
    assert rc == _lib.OK
    torch.cuda.synchronize()
    return y_na.cpu().numpy(), y_ref.cpu().numpy(), x.cpu().numpy()


@pytest.mark.parametrize("shapes", [
    [(65536, 96, -1), (32768, 16, -1), (65536, 2, 1)],     # chain, smaller chain, smaller again
    [(16384, 200, -1), (16384, 7, 1), (4096, 33, -1)],      # four-step below 256 rows, then one pass
])
def test_noalloc_equals_the_allocating_call(shapes):
    for i, (n, b, sign) in enumerate(shapes):
        print(f"noalloc call {i}: n={n} batch={b} sign={sign}", flush=True)
        got, want, x = _both(n, b, sign, seed=40 + i)
        assert np.array_equal(got, want), (n, b, sign)
        assert oracle.rel_l2_rows(got, oracle.c2c_f64(x, n, sign)) <= 1e-5


def test_noalloc_refuses_what_does_not_fit():
    import torch
    from bfly import _lib, ops
    lib = _lib.hip()
    assert lib.bfly_noalloc_prepare() == _lib.OK
    n, b = 65536, _lib.NOALLOC_SCRATCH_BYTES // (8 * 65536) + 1        # one row too many
    x = torch.zeros(b, 2 * n, device="cuda")
    y = torch.empty_like(x)
    tab = ops.table(_lib.KIND_C2C, n, x.device)
    s = torch.cuda.current_stream().cuda_stream
    assert lib.bfly_c2c_noalloc(x.data_ptr(), y.data_ptr(), n, b, -1, tab.data_ptr(), s) == _lib.ERR_SCRATCH
    assert lib.bfly_c2c_noalloc(x.data_ptr(), y.data_ptr(), n, b - 1, -1, tab.data_ptr(), s) == _lib.OK
    assert lib.bfly_c2c_noalloc(x.data_ptr(), x.data_ptr(), n, 2, -1, tab.data_ptr(), s) == _lib.ERR_SHAPE
    assert lib.bfly_c2c_noalloc(x.data_ptr(), y.data_ptr(), 100, 2, -1, tab.data_ptr(), s) == _lib.ERR_SHAPE
    torch.cuda.synchronize()
