"""Host side of the grouped launch (bfly_group_*), through the CPU stub: the
request / configuration guards, the unit arithmetic and the uneven workload
generator.  No device compute."""
import ctypes as C

import numpy as np
import pytest

import bfly
from bfly import _lib, group

C2C, R2C, C2R = bfly.KIND_C2C, bfly.KIND_R2C, bfly.KIND_C2R
BUF = np.zeros(16, np.float32)
P = BUF.ctypes.data


def _reqs(specs):
    # the stub never dereferences request buffers: distinct, far-apart
    # addresses stand for each request's input and output
    arr = (_lib.GroupRequest * len(specs))()
    for i, (kind, n, b, *sign) in enumerate(specs):
        arr[i] = _lib.GroupRequest(P + (2 * i << 36), P + ((2 * i + 1) << 36), kind, sign[0] if sign else -1, n, b)
    return arr


def _create(lib, specs, cfg=None):
    h = C.c_void_p()
    arr = _reqs(specs)
    rc = lib.bfly_group_create(arr, len(specs), C.byref(cfg) if cfg is not None else None, C.byref(h))
    return rc, h


def _query(lib, h):
    u, g, b = C.c_uint32(), C.c_uint32(), C.c_size_t()
    assert lib.bfly_group_query(h, C.byref(u), C.byref(g), C.byref(b)) == 0
    return u.value, g.value


@pytest.mark.parametrize("lib_name", ["stub", "hip"])
def test_group_supported_envelope(lib_name):
    lib = getattr(bfly, lib_name)()
    for wg in (512, 1024):
        for k in range(3, 16):
            n = 1 << k
            c2c = 16 <= n <= (8192 if wg == 512 else 16384)
            real = 32 <= n <= 8192
            assert lib.bfly_group_supported(C2C, n, wg) == int(c2c), (wg, n)
            assert lib.bfly_group_supported(R2C, n, wg) == int(real), (wg, n)
            assert lib.bfly_group_supported(C2R, n, wg) == int(real), (wg, n)
    assert lib.bfly_group_supported(C2C, 4096, 256) == 0
    assert lib.bfly_group_supported(C2C, 1000, 512) == 0
    assert lib.bfly_group_supported(7, 64, 512) == 0


def test_group_unit_counts():
    # a unit is 16*wg points: wg/(n'/16) transforms, n' = n (C2C) or n/2
    s = bfly.stub()
    s.bfly_stub_reset()
    specs = [(C2C, 4096, 5), (R2C, 8192, 3), (C2C, 16, 1000), (C2R, 64, 7), (C2C, 8192, 2, 1)]
    rc, h = _create(s, specs)
    assert rc == 0
    units, grid = _query(s, h)
    # 4096: 2 per unit -> 3; R2C 8192: n'=4096 -> 2; C2C 16: 512 -> 2;
    # C2R 64: n'=32, T=2, 256 per unit -> 1; C2C 8192: T=512 -> 1 per unit -> 2
    assert units == 3 + 2 + 2 + 1 + 2
    assert grid == units
    assert s.bfly_group_launch(h, None, None) == 0
    assert s.bfly_group_destroy(h) == 0
    cfg = _lib.GroupConfig(1024, 0, 0, 0, 0)
    rc, h = _create(s, specs, cfg)
    assert rc == 0
    assert _query(s, h)[0] == 2 + 1 + 1 + 1 + 1
    assert s.bfly_group_destroy(h) == 0
    cfg = _lib.GroupConfig(512, 0, 1, 1, 4)
    rc, h = _create(s, [(C2C, 64, 100000)], cfg)
    assert rc == 0
    assert _query(s, h) == ((100000 + 127) // 128, 4)
    assert s.bfly_group_destroy(h) == 0
    # LDS twiddles change where the twiddles come from, not the unit list
    rc, h = _create(s, specs, _lib.GroupConfig(512, 0, 0, 0, 0, 1))
    assert rc == 0
    assert _query(s, h)[0] == 3 + 2 + 2 + 1 + 2
    assert s.bfly_group_destroy(h) == 0


@pytest.mark.parametrize("bad_cfg", [
    (256, 0, 0, 0, 0), (512, 2, 0, 0, 0), (512, 0, 4, 0, 0), (512, 0, -1, 0, 0),
    (512, 0, 0, 2, 0), (512, 0, 0, 0, -5), (512, 0, 0, 0, 0, 2), (512, 0, 0, 0, 0, -1),
    (512, 1, 0, 0, 0, 1)])  # LDS twiddles run radix schedule 0 only
def test_group_config_guards(bad_cfg):
    s = bfly.stub()
    rc, h = _create(s, [(C2C, 64, 1)], _lib.GroupConfig(*bad_cfg))
    assert rc == bfly.ERR_SHAPE and not h.value


@pytest.mark.parametrize("spec", [
    (C2C, 64, 0), (C2C, 100, 1), (C2C, 16384, 1), (R2C, 16, 1), (R2C, 16384, 1),
    (C2R, 16384, 1), (C2C, 64, 1, 0), (C2C, 64, 1, 2), (5, 64, 1)])
def test_group_request_guards(spec):
    s = bfly.stub()
    rc, h = _create(s, [(C2C, 64, 1), spec])
    assert rc == bfly.ERR_SHAPE and not h.value


def test_group_null_guards():
    s = bfly.stub()
    h = C.c_void_p()
    assert s.bfly_group_create(None, 1, None, C.byref(h)) == bfly.ERR_SHAPE
    assert s.bfly_group_create(_reqs([(C2C, 64, 1)]), 0, None, C.byref(h)) == bfly.ERR_SHAPE
    assert s.bfly_group_create(_reqs([(C2C, 64, 1)]), 1, None, None) == bfly.ERR_SHAPE
    arr = _reqs([(C2C, 64, 1)])
    arr[0].in_ = None
    assert s.bfly_group_create(arr, 1, None, C.byref(h)) == bfly.ERR_SHAPE
    # a request whose output overlaps its input
    arr = _reqs([(C2C, 64, 1), (R2C, 1024, 3)])
    arr[1].out = arr[1].in_ + 4 * 1024 * 3 - 4
    assert s.bfly_group_create(arr, 2, None, C.byref(h)) == bfly.ERR_SHAPE
    arr[1].out = arr[1].in_ + 4 * 1024 * 3
    assert s.bfly_group_create(arr, 2, None, C.byref(h)) == 0 and h.value
    assert s.bfly_group_destroy(h) == 0
    assert s.bfly_group_launch(None, None, None) == bfly.ERR_SHAPE
    assert s.bfly_group_destroy(None) == 0


def test_group_allocation_failure_leaves_no_handle():
    s = bfly.stub()
    s.bfly_stub_reset()
    s.bfly_stub_set_failures(1, 0, 0)
    rc, h = _create(s, [(C2C, 64, 1)])
    assert rc == _lib.ERR_ALLOC and not h.value
    s.bfly_stub_reset()


def test_uneven_workload_generator():
    specs = group.uneven_specs()
    assert len(specs) == 512
    kinds = np.array([k for k, _, _ in specs])
    ns = np.array([n for _, n, _ in specs])
    bs = np.array([b for _, _, b in specs])
    assert set(np.unique(kinds)) <= {C2C, R2C}
    assert set(np.unique(ns)) == {1 << k for k in range(6, 14)}
    assert bs.min() >= 1 and bs.max() <= 256
    # same seed, same draw
    assert group.uneven_specs() == specs
    # the draw is the documented one: per request kind, log2 n, batch
    rng = np.random.default_rng(3)
    k0 = C2C if rng.random() < 0.5 else R2C
    n0 = 1 << int(rng.integers(6, 14))
    b0 = int(rng.integers(1, 257))
    assert specs[0] == (k0, n0, b0)
    # roughly balanced kinds, every length present
    assert 200 < int((kinds == C2C).sum()) < 312
    s = bfly.stub()
    rc, h = _create(s, specs)
    assert rc == 0
    assert s.bfly_group_destroy(h) == 0


def test_pick_throughput_and_serving_objectives():
    # rows: (ms, max rel diff, p50 latency us)
    rows = [(0.40, 1e-7, 300.0), (0.41, 1e-7, 25.0), (0.39, 2e-5, 10.0), (0.43, 1e-7, 20.0), (0.405, 0.0, None)]
    # throughput: the fastest row inside the gate (row 2 is faster but fails it)
    assert group.pick(rows, gate=1e-5) == 0
    # serving: lowest median latency within 5 % of the fastest gated time
    # (0.42 us limit: row 3 at 0.43 is out, row 4 has no latency)
    assert group.pick(rows, gate=1e-5, latency_slack=0.05) == 1
    assert group.pick(rows, gate=1e-5, latency_slack=0.10) == 3
    assert group.pick(rows, gate=1e-5, latency_slack=0.0) == 0
    # ties on latency go to the faster row
    assert group.pick([(0.5, 0.0, 9.0), (0.45, 0.0, 9.0)], gate=1e-5, latency_slack=0.2) == 1
    # no latency measured anywhere: the fastest
    assert group.pick([(0.5, 0.0, None), (0.45, 0.0, None)], gate=1e-5, latency_slack=0.2) == 1
    with pytest.raises(RuntimeError):
        group.pick([(0.3, 1e-3, 5.0)], gate=1e-5)
    with pytest.raises(ValueError):
        group.pick(rows, gate=1e-5, latency_slack=-0.1)

