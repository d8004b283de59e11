"""GPU: bfly_host_pipelined (host buffers in, host buffers out, the batch
moved over the link in chunks whose uploads, transforms and downloads
overlap) against the device-pointer entry points on the whole batch.

The pipelined call runs one kernel variant on every chunk, so its result
equals bfly_c2c_ex / bfly_real_ex of that variant on the whole batch bit for
bit; the AUTO cases are held to the FP64 gates of the parity tests.  It
replaces the synchronous upload -> kernel -> download of the reference's
Plan::run / runR2C / runC2R (src/framework/butterfly.cpp:723-850).
"""
import numpy as np
import pytest

import oracle
from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]

REL_L2 = 1e-5
MAX_REL_SCALED = 1e-4


def _host(shape, pinned):
    import torch
    t = torch.empty(shape, dtype=torch.float32, pin_memory=pinned)
    return t, t.numpy()


def _pipelined(kind, sign, variant, x, n, b, row_out, chunk, pinned, work_rows=0, null_stream=False):
    import torch
    from bfly import _lib, ops
    lib = _lib.hip()
    dev = torch.device("cuda:0")
    tin, hin = _host(x.shape, pinned)
    hin[...] = x
    tout, hout = _host((b, row_out), pinned)
    hout[...] = np.nan
    din = torch.empty(x.size, dtype=torch.float32, device=dev)
    dout = torch.empty(b * row_out, dtype=torch.float32, device=dev)
    work = torch.empty(max(work_rows, 1) * 2 * n, dtype=torch.float32, device=dev) if work_rows else None
    tab = ops.table(kind, n, dev)
    torch.cuda.synchronize()
    rc = lib.bfly_host_pipelined(kind, sign, variant, tin.data_ptr(), tout.data_ptr(), n, b, tab.data_ptr(),
                                 din.data_ptr(), dout.data_ptr(), None if work is None else work.data_ptr(),
                                 chunk, None if null_stream else torch.cuda.current_stream().cuda_stream)
    assert rc == 0, f"bfly_host_pipelined kind={kind} n={n} b={b} chunk={chunk} rc={rc}"
    return hout.copy()


def _whole_c2c(x, n, b, sign, variant):
    import torch
    from bfly import _lib, ops
    xt = torch.from_numpy(x).cuda()
    y = torch.empty_like(xt)
    work = torch.empty(b * 2 * n, dtype=torch.float32, device=xt.device) if n > 4096 else None
    tab = ops.table(_lib.KIND_C2C, n, xt.device)
    rc = _lib.hip().bfly_c2c_ex(xt.data_ptr(), y.data_ptr(), n, b, sign, variant, tab.data_ptr(),
                                None if work is None else work.data_ptr(), torch.cuda.current_stream().cuda_stream)
    assert rc == 0
    torch.cuda.synchronize()
    return y.cpu().numpy()


def _whole_real(kind, x, n, b, variant):
    import torch
    from bfly import _lib, ops
    xt = torch.from_numpy(x).cuda()
    row_out = n + 2 if kind == _lib.KIND_R2C else n
    y = torch.empty((b, row_out), dtype=torch.float32, device=xt.device)
    tab = ops.table(kind, n, xt.device)
    rc = _lib.hip().bfly_real_ex(kind, xt.data_ptr(), y.data_ptr(), n, b, n // 2 + 1, n + 2, variant, 0,
                                 tab.data_ptr(), torch.cuda.current_stream().cuda_stream)
    assert rc == 0
    torch.cuda.synchronize()
    return y.cpu().numpy()


def _gate_c2c(got, x, n, sign):
    ref = oracle.c2c_f64(x, n, sign)
    assert oracle.rel_l2_rows(got, ref) <= REL_L2
    assert oracle.max_rel_scaled(got, ref.astype(np.float32)) <= MAX_REL_SCALED


# (n, batch, chunk rows, variant, sign): ragged last chunks, one chunk, the
# default chunk (1024 transforms of 1024 points: five chunks for 4100), the
# single-workgroup and four-step long kernels
C2C_CASES = [
    (1024, 100, 7, 5, -1),
    (1024, 4100, 0, 5, -1),
    (4096, 300, 0, 5, 1),
    (4096, 33, 64, 1, -1),
    (256, 4097, 1000, 2, -1),
    (16384, 70, 16, 7, -1),
    (16384, 40, 9, 5, 1),
    (65536, 5, 2, 7, -1),
    (32768, 6, 4, 7, 1),
]


@pytest.mark.parametrize("pinned", [True, False], ids=["pinned", "pageable"])
@pytest.mark.parametrize("n,b,chunk,variant,sign", C2C_CASES)
def test_pipelined_c2c_equals_the_whole_batch_call(n, b, chunk, variant, sign, pinned):
    from bfly import _lib
    print(f"pipelined c2c n={n} b={b} chunk={chunk} variant={variant} pinned={pinned}", flush=True)
    x = oracle.uniform_input(31, (b, 2 * n))
    work_rows = (chunk or b) if n > 4096 else 0
    got = _pipelined(_lib.KIND_C2C, sign, variant, x, n, b, 2 * n, chunk, pinned, work_rows)
    want = _whole_c2c(x, n, b, sign, variant)
    assert np.array_equal(got, want)
    if b * n <= (1 << 21):
        _gate_c2c(got, x, n, sign)


@pytest.mark.parametrize("n,b,chunk", [(4096, 512, 0), (16384, 64, 0), (2048, 9, 4)])
def test_pipelined_c2c_auto_meets_the_fp64_gates(n, b, chunk):
    from bfly import _lib
    print(f"pipelined c2c AUTO n={n} b={b} chunk={chunk}", flush=True)
    x = oracle.uniform_input(32, (b, 2 * n))
    work_rows = b if n > 4096 else 0
    got = _pipelined(_lib.KIND_C2C, -1, _lib.VARIANT_AUTO, x, n, b, 2 * n, chunk, True, work_rows)
    _gate_c2c(got, x, n, -1)


@pytest.mark.parametrize("pinned", [True, False], ids=["pinned", "pageable"])
@pytest.mark.parametrize("n,b,chunk,variant", [(4096, 50, 16, 5), (1024, 300, 0, 2), (256, 77, 10, 1)])
def test_pipelined_real_equals_the_whole_batch_call(n, b, chunk, variant, pinned):
    from bfly import _lib
    print(f"pipelined real n={n} b={b} chunk={chunk} variant={variant} pinned={pinned}", flush=True)
    x = oracle.uniform_input(33, (b, n))
    spec = _pipelined(_lib.KIND_R2C, 0, variant, x, n, b, n + 2, chunk, pinned)
    assert np.array_equal(spec, _whole_real(_lib.KIND_R2C, x, n, b, variant))
    ref = oracle.c2c_f64(np.stack([x, np.zeros_like(x)], -1).reshape(b, 2 * n), n, -1)[:, : n + 2]
    assert oracle.rel_l2_rows(spec, ref) <= REL_L2
    back = _pipelined(_lib.KIND_C2R, 0, variant, spec, n, b, n, chunk, pinned)
    assert np.array_equal(back, _whole_real(_lib.KIND_C2R, spec, n, b, variant))
    assert oracle.rel_l2_rows(back, x) <= REL_L2  # C2R scales 1/n: the round trip is the input


@pytest.mark.parametrize("chunk", [0, 24])
def test_pipelined_on_the_null_stream(chunk):
    # uploads on the legacy default stream, transforms and downloads on the
    # library's own streams: one chunk and seven chunks
    from bfly import _lib
    print(f"pipelined c2c on the null stream chunk={chunk}", flush=True)
    n, b = 2048, 150
    x = oracle.uniform_input(34, (b, 2 * n))
    got = _pipelined(_lib.KIND_C2C, -1, 5, x, n, b, 2 * n, chunk, True, null_stream=True)
    assert np.array_equal(got, _whole_c2c(x, n, b, -1, 5))
    _gate_c2c(got, x, n, -1)


def test_pipelined_guards_return_before_any_copy():
    # every buffer is sized for the call, so a guard that let one through
    # would still stay in bounds
    import torch
    from bfly import _lib, ops
    lib = _lib.hip()
    n, b = 1024, 4
    tab = ops.table(_lib.KIND_C2C, n, "cuda:0")
    d = torch.empty(2 * n * b, dtype=torch.float32, device="cuda:0")
    e = torch.empty_like(d)
    h = np.zeros(2 * n * b, np.float32)
    stream = torch.cuda.current_stream().cuda_stream
    assert lib.bfly_host_pipelined(_lib.KIND_C2C, 0, 0, h.ctypes.data, h.ctypes.data, n, b, tab.data_ptr(),
                                   d.data_ptr(), e.data_ptr(), None, 0, stream) == _lib.ERR_SHAPE
    assert lib.bfly_host_pipelined(_lib.KIND_C2C, -1, 0, h.ctypes.data, h.ctypes.data, n, b, tab.data_ptr(),
                                   d.data_ptr(), d.data_ptr(), None, 0, stream) == _lib.ERR_SHAPE


def test_pipelined_calls_from_four_threads_at_once():
    # four host threads (ctypes releases the GIL during the call) pipelining
    # their own buffers on one device: the library's streams are shared under
    # its lock, and every result equals the whole-batch call
    import threading
    import torch
    from bfly import _lib, ops
    lib = _lib.hip()
    n, b, chunk = 1024, 2048, 256
    dev = torch.device("cuda:0")
    tab = ops.table(_lib.KIND_C2C, n, dev)
    jobs = []
    for i in range(4):
        x = oracle.uniform_input(40 + i, (b, 2 * n))
        tin, hin = _host(x.shape, True)
        hin[...] = x
        tout, hout = _host(x.shape, True)
        din = torch.empty(x.size, dtype=torch.float32, device=dev)
        dout = torch.empty_like(din)
        jobs.append((x, tin, tout, hout, din, dout))
    torch.cuda.synchronize()
    rcs = [None] * 4

    def run(i):
        x, tin, tout, hout, din, dout = jobs[i]
        rcs[i] = lib.bfly_host_pipelined(_lib.KIND_C2C, -1, 5, tin.data_ptr(), tout.data_ptr(), n, b,
                                         tab.data_ptr(), din.data_ptr(), dout.data_ptr(), None, chunk, None)

    threads = [threading.Thread(target=run, args=(i,)) for i in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=60)
    assert rcs == [0, 0, 0, 0]
    for x, tin, tout, hout, din, dout in jobs:
        assert np.array_equal(hout, _whole_c2c(x, n, b, -1, 5))
