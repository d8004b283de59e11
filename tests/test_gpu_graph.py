"""GPU: hipGraph capture of library calls through the C ABI (bfly_graph_*).

One graph holds a C2C (streaming LDS-twiddle kernel), a four-step long C2C,
an R2C, a C2R, a 2-D C2C, a strided column transform, a normalised inverse
and padded-row C2Cs (one-pass and four-step).  It is replayed
several times with the program's own eager work between the replays (new
inputs written, the same calls run eagerly into separate outputs, full
synchronisations), and every replay must equal the eager results bit for
bit."""
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]


def test_graph_replays_equal_eager_calls(dev):
    import torch
    from bfly import _lib, ops
    from bfly.graph import Graph

    lib = _lib.hip()
    s = torch.cuda.Stream(device=dev)
    sid = s.cuda_stream
    g = torch.Generator(device=dev).manual_seed(9)

    def rnd(*shape):
        return torch.rand(*shape, generator=g, device=dev) * 2 - 1

    # inputs, graph outputs and eager outputs, all allocated before the capture
    x1 = rnd(64, 2 * 4096)
    xl = rnd(4, 2 * 65536)
    xr = rnd(32, 4096)
    hr = rnd(32, 4098)
    x2 = rnd(3, 512, 2 * 256)
    xc = rnd(2, 1024, 2 * 33)
    xp = rnd(16, 2 * (1024 + 8))  # padded rows: 1024 values every 1032
    xq = rnd(2, 2 * (32768 + 16))  # padded rows through the four-step chain
    ins = [x1, xl, xr, hr, x2, xc, xp, xq]
    outs = [torch.empty_like(x1), torch.empty_like(xl), torch.empty_like(hr), torch.empty_like(xr),
            torch.empty_like(x2), torch.empty_like(xc), torch.empty_like(x1), torch.empty(16, 2 * 1024, device=dev),
            torch.empty(2, 2 * 32768, device=dev)]
    eager = [torch.empty_like(o) for o in outs]
    t1 = ops.table(_lib.KIND_C2C, 4096, dev)
    tl = ops.table(_lib.KIND_C2C, 65536, dev)
    wl = ops.workspace(_lib.KIND_C2C, 65536, 4, dev)
    tr, tc2r = ops.table(_lib.KIND_R2C, 4096, dev), ops.table(_lib.KIND_C2R, 4096, dev)
    t2r, t2c = ops.table(_lib.KIND_C2C, 256, dev), ops.table2d_cols(512, dev)
    w2 = torch.empty(2 * x2.numel(), device=dev)
    tcc = ops.table2d_cols(1024, dev)
    wc = torch.empty(2 * xc.numel(), device=dev)
    tp = ops.table(_lib.KIND_C2C, 1024, dev)
    tq = ops.table(_lib.KIND_C2C, 32768, dev)
    wq = ops.workspace(_lib.KIND_C2C, 32768, 2, dev)

    def calls(o):
        rc = [lib.bfly_c2c(x1.data_ptr(), o[0].data_ptr(), 4096, 64, -1, t1.data_ptr(), None, sid),
              lib.bfly_c2c(xl.data_ptr(), o[1].data_ptr(), 65536, 4, -1, tl.data_ptr(), wl.data_ptr(), sid),
              lib.bfly_real_ex(1, xr.data_ptr(), o[2].data_ptr(), 4096, 32, 2049, 4098, 0, 0, tr.data_ptr(), sid),
              lib.bfly_real_ex(2, hr.data_ptr(), o[3].data_ptr(), 4096, 32, 2049, 4098, 0, 0, tc2r.data_ptr(), sid),
              lib.bfly_fft2d(x2.data_ptr(), o[4].data_ptr(), 512, 256, 3, -1, t2r.data_ptr(), t2c.data_ptr(),
                             w2.data_ptr(), sid),
              lib.bfly_c2c_columns(xc.data_ptr(), o[5].data_ptr(), 1024, 33, 2, 1, tcc.data_ptr(), wc.data_ptr(),
                                   sid),
              lib.bfly_c2c_scaled(x1.data_ptr(), o[6].data_ptr(), 4096, 64, 1, 1.0 / 4096, t1.data_ptr(), None, sid),
              lib.bfly_c2c_rows(xp.data_ptr(), 1032, o[7].data_ptr(), 1024, 1024, 16, -1, tp.data_ptr(), None,
                                sid),
              lib.bfly_c2c_rows(xq.data_ptr(), 32784, o[8].data_ptr(), 32768, 32768, 2, 1, tq.data_ptr(),
                                wq.data_ptr(), sid)]
        assert rc == [0] * 9, rc

    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        calls(eager)  # eager first: resident-grid queries and tables settle before the capture
    s.synchronize()
    graph = Graph()
    graph.capture(lambda: calls(outs), sid)
    try:
        for rep in range(4):
            with torch.cuda.stream(s):
                for t in ins:                      # new inputs (eager torch work)
                    t.copy_(rnd(*t.shape))
                calls(eager)                       # the same calls, eagerly, into separate outputs
                for o in outs:
                    o.fill_(float("nan"))
            s.synchronize()
            graph.replay(sid)
            s.synchronize()
            for k, (a, b) in enumerate(zip(outs, eager)):
                assert torch.equal(a, b), (rep, k)
    finally:
        graph.close()
    assert lib.bfly_graph_begin(None) == _lib.ERR_SHAPE


def test_forward_inverse_wrappers_and_teardown(dev):
    import torch
    from bfly import _lib, ops
    lib = _lib.hip()
    sid = torch.cuda.current_stream(dev).cuda_stream
    for n, b in ((1024, 7), (65536, 2)):
        x = torch.rand(b, 2 * n, device=dev) * 2 - 1
        t = ops.table(_lib.KIND_C2C, n, dev)
        w = ops.workspace(_lib.KIND_C2C, n, b, dev)
        wp = w.data_ptr() if w is not None else None
        outs = [torch.empty_like(x) for _ in range(4)]
        assert lib.bfly_c2c_fwd(x.data_ptr(), outs[0].data_ptr(), n, b, t.data_ptr(), wp, sid) == 0
        assert lib.bfly_c2c(x.data_ptr(), outs[1].data_ptr(), n, b, -1, t.data_ptr(), wp, sid) == 0
        assert lib.bfly_c2c_inv(x.data_ptr(), outs[2].data_ptr(), n, b, t.data_ptr(), wp, sid) == 0
        assert lib.bfly_c2c(x.data_ptr(), outs[3].data_ptr(), n, b, 1, t.data_ptr(), wp, sid) == 0
        torch.cuda.synchronize()
        assert torch.equal(outs[0], outs[1]) and torch.equal(outs[2], outs[3])
    assert lib.bfly_rt_teardown() == 0


def test_concurrent_host_threads_on_separate_streams(dev):
    """Four host threads, each with its own stream, call the library at once
    (ctypes releases the GIL during the calls): every result equals the same
    call made alone, bit for bit (the library's host state -- resident-grid
    caches, table layouts -- is shared and must stay consistent)."""
    import threading

    import torch
    from bfly import _lib, ops
    lib = _lib.hip()
    shapes = [(4096, 256), (1024, 1024), (65536, 8), (256, 4096)]
    g = torch.Generator(device=dev).manual_seed(11)
    xs = [torch.rand(b, 2 * n, generator=g, device=dev) * 2 - 1 for n, b in shapes]
    tabs = [ops.table(_lib.KIND_C2C, n, dev) for n, _ in shapes]
    works = [ops.workspace(_lib.KIND_C2C, n, b, dev) for n, b in shapes]
    alone = [torch.empty_like(x) for x in xs]
    sid = torch.cuda.current_stream(dev).cuda_stream
    for (n, b), x, y, t, w in zip(shapes, xs, alone, tabs, works):
        assert lib.bfly_c2c(x.data_ptr(), y.data_ptr(), n, b, -1, t.data_ptr(),
                            w.data_ptr() if w is not None else None, sid) == 0
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(device=dev) for _ in shapes]
    outs = [[torch.empty_like(x) for _ in range(8)] for x in xs]
    # each thread's own workspace (the four-step chain writes it), allocated
    # here and kept alive until every stream has finished
    wks = [torch.empty_like(w) if w is not None else None for w in works]
    errors = []

    def worker(k):
        n, b = shapes[k]
        s = streams[k].cuda_stream
        wk = wks[k]
        for o in outs[k]:
            rc = lib.bfly_c2c(xs[k].data_ptr(), o.data_ptr(), n, b, -1, tabs[k].data_ptr(),
                              wk.data_ptr() if wk is not None else None, s)
            if rc:
                errors.append((k, rc))

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(len(shapes))]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    torch.cuda.synchronize()
    assert not errors, errors
    for k in range(len(shapes)):
        for o in outs[k]:
            assert torch.equal(o, alone[k]), shapes[k]
