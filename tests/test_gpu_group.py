"""GPU: the grouped launch (one persistent kernel over many uneven requests).

With radix schedule 0 a grouped transform runs the same per-transform code
as its per-request launch, so outputs must equal the per-request PLAIN
launches (STREAM_LT launches with LDS twiddles) bit for bit, in every launch
configuration and under hipGraph replay.  Radix schedule 1 (remainder radix first) rounds differently and is
gated against the FP64 oracle (relative L2 <= 1e-5 per transform).
"""
import itertools

import numpy as np
import pytest

import oracle
from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]

REL_L2 = 1e-5


def _mixed_specs(max_c2c=8192):
    # every kind and length, ragged batches (partial units), both C2C signs
    specs, signs = [], []
    batches = [1, 3, 37, 5, 130, 2, 9, 64, 17, 1, 6]
    i = 0
    for k in range(4, 15):
        n = 1 << k
        if n <= max_c2c:
            specs.append((0, n, batches[i % len(batches)]))
            signs.append(-1 if i % 2 == 0 else 1)
            i += 1
        if 32 <= n <= 8192:
            specs.append((1, n, batches[i % len(batches)]))
            signs.append(-1)
            i += 1
            specs.append((2, n, batches[i % len(batches)]))
            signs.append(1)
            i += 1
    return specs, signs


def _fp64(r):
    x = r.x.cpu().numpy()
    if r.kind == 0:
        return oracle.c2c_f64(x.reshape(r.batch, 2 * r.n), r.n, r.sign)
    if r.kind == 1:
        return oracle.r2c_f64(x.reshape(r.batch, r.n), r.n)
    return oracle.c2r_f64(x.reshape(r.batch, r.n + 2), r.n)


def _gate_fp64(reqs):
    for r in reqs:
        got = r.y.cpu().numpy().reshape(r.batch, -1)
        e = oracle.rel_l2_rows(got, _fp64(r))
        assert e <= REL_L2, (r.kind, r.n, r.batch, e)


@pytest.fixture(scope="module")
def mixed():
    import torch
    from bfly import group
    specs, signs = _mixed_specs()
    reqs = group.make_requests(specs, seed=11, signs=signs)
    group.launch_per_request(reqs)
    torch.cuda.synchronize()
    return reqs, group.outputs(reqs)


@pytest.mark.parametrize("wg,order,dynamic", list(itertools.product((512, 1024), (0, 1, 2, 3), (0, 1))))
def test_group_equals_per_request_launches_bitwise(mixed, wg, order, dynamic):
    import torch
    from bfly import group
    reqs, ref = mixed
    g = group.Group(reqs, group.Config(workgroup=wg, order=order, dynamic=dynamic))
    for _ in range(2):  # the dynamic counter must come back to zero
        for r in reqs:
            r.y.fill_(float("nan"))
        g.launch()
        torch.cuda.synchronize()
        for r, want in zip(reqs, ref):
            assert torch.equal(r.y, want), (r.kind, r.n, r.batch, r.sign)
    g.close()


@pytest.fixture(scope="module")
def mixed_lt():
    # the same requests' per-request STREAM_LT launches (C2R / R2C below
    # n = 256 run their register-twiddle kernels, as the grouped LT path does)
    import torch
    from bfly import group
    specs, signs = _mixed_specs()
    reqs = group.make_requests(specs, seed=11, signs=signs)
    group.launch_per_request(reqs, variant=group.Config(twiddles=1).reference_variant())
    torch.cuda.synchronize()
    return reqs, group.outputs(reqs)


@pytest.mark.parametrize("wg,order,dynamic", list(itertools.product((512, 1024), (0, 3), (0, 1))))
def test_group_lds_twiddles_equal_stream_lt_launches_bitwise(mixed_lt, wg, order, dynamic):
    # twiddles = 1: each unit stages its plan's two-level tables in LDS and
    # runs the STREAM_LT kernels' per-transform code -- the same bits
    import torch
    from bfly import group
    reqs, ref = mixed_lt
    g = group.Group(reqs, group.Config(workgroup=wg, order=order, dynamic=dynamic, twiddles=1))
    for _ in range(2):
        for r in reqs:
            r.y.fill_(float("nan"))
        g.launch()
        torch.cuda.synchronize()
        for r, want in zip(reqs, ref):
            assert torch.equal(r.y, want), (r.kind, r.n, r.batch, r.sign)
    _gate_fp64(reqs)
    g.close()


def test_group_lds_twiddles_16384_and_graph_replays():
    import torch
    from bfly import group
    reqs = group.make_requests([(0, 16384, 3), (0, 8192, 2), (1, 8192, 3), (2, 4096, 5), (0, 16, 7)], seed=13,
                               signs=[-1, 1, -1, 1, 1])
    group.launch_per_request(reqs, variant=group.Config(twiddles=1).reference_variant())
    torch.cuda.synchronize()
    ref = group.outputs(reqs)
    _gate_fp64(reqs)
    g = group.Group(reqs, group.Config(workgroup=1024, twiddles=1, dynamic=1))
    graph = group.capture(g.launch)
    for rep in range(3):
        for r in reqs:
            r.y.fill_(float("nan"))
        torch.randn(1 << 20, device="cuda").mul_(2.0)  # eager work between replays
        graph.replay()
        torch.cuda.synchronize()
        for r, want in zip(reqs, ref):
            assert torch.equal(r.y, want), (rep, r.kind, r.n)
    del graph
    g.close()


@pytest.mark.parametrize("wg", [512, 1024])
def test_group_remainder_first_schedule_fp64(mixed, wg):
    import torch
    from bfly import group
    reqs, ref = mixed
    g = group.Group(reqs, group.Config(workgroup=wg, schedule=1))
    g.launch()
    torch.cuda.synchronize()
    _gate_fp64(reqs)
    assert group.max_rel_diff(group.outputs(reqs), ref) <= REL_L2
    g.close()


def test_group_c2c_16384_with_1024_thread_workgroups():
    import torch
    from bfly import group
    reqs = group.make_requests([(0, 16384, 3), (0, 8192, 2), (1, 8192, 3)], seed=12, signs=[-1, 1, -1])
    group.launch_per_request(reqs)
    torch.cuda.synchronize()
    ref = group.outputs(reqs)
    _gate_fp64(reqs)
    g = group.Group(reqs, group.Config(workgroup=1024))
    g.launch()
    torch.cuda.synchronize()
    for r, want in zip(reqs, ref):
        assert torch.equal(r.y, want)
    g.close()


@pytest.mark.parametrize("dynamic", [0, 1])
def test_group_graph_replays_match_eager(mixed, dynamic):
    # captured once, replayed several times with eager work in between; every
    # replay rewrites the outputs (poisoned before each replay) exactly
    import torch
    from bfly import group
    reqs, ref = mixed
    g = group.Group(reqs, group.Config(dynamic=dynamic, order=1))
    graph = group.capture(g.launch)
    for rep in range(3):
        for r in reqs:
            r.y.fill_(float("nan"))
        side = torch.randn(1 << 20, device="cuda")  # eager work between replays
        side.mul_(2.0)
        graph.replay()
        torch.cuda.synchronize()
        for r, want in zip(reqs, ref):
            assert torch.equal(r.y, want), (rep, r.kind, r.n)
        group.launch_per_request(reqs[:3])  # eager library launches between replays
        torch.cuda.synchronize()
    del graph
    g.close()


def test_uneven_workload_512_requests_seed3():
    # 512 requests: n log-uniform over 64..8192, batch U[1,256], C2C/R2C 50/50
    import torch
    from bfly import group
    specs = group.uneven_specs()
    reqs = group.make_requests(specs, seed=3)
    group.launch_per_request(reqs)
    torch.cuda.synchronize()
    ref = group.outputs(reqs)
    g = group.Group(reqs)
    for r in reqs:
        r.y.fill_(float("nan"))
    g.launch()
    torch.cuda.synchronize()
    for r, want in zip(reqs, ref):
        assert torch.equal(r.y, want), (r.kind, r.n, r.batch)
    # FP64 check of a sample of requests (one of each length and kind)
    seen, sample = set(), []
    for r in reqs:
        if (r.kind, r.n) not in seen:
            seen.add((r.kind, r.n))
            sample.append(r)
    _gate_fp64(sample)
    # per-request completion stamps: one per request, inside the launch
    lat = group.latencies_us(g)
    assert lat.shape == (512,) and np.all(lat > 0) and np.all(np.isfinite(lat))
    ms = group.time_ms(g.launch, reps=5)
    assert lat.max() < 20 * ms * 1e3
    g.close()


def test_least_work_first_order_lowers_median_latency():
    # order 3 puts the requests with the least total work first: on the
    # uneven workload their completion stamps come early, so the median
    # per-request latency drops well below that of most-work-first (order 1)
    from bfly import group
    reqs = group.make_requests(group.uneven_specs(), seed=3)
    p50 = {}
    for order in (1, 3):
        g = group.Group(reqs, group.Config(order=order, dynamic=1))
        g.launch()
        p50[order] = float(np.percentile(group.latencies_us(g), 50))
        g.close()
    assert p50[3] < 0.75 * p50[1], p50


def test_autotune_picks_a_gated_configuration():
    import torch
    from bfly import group
    specs = group.uneven_specs(count=64, seed=5)
    reqs = group.make_requests(specs, seed=5)
    res = group.autotune(reqs, space={"workgroup": (512, 1024), "schedule": (0, 1), "order": (0, 1),
                                      "dynamic": (0, 1), "grid": (0,), "twiddles": (0, 1)}, reps=3)
    assert len(res.table) == 24  # LDS twiddles with schedule 0 only
    assert all(err <= REL_L2 for _, _, err, _, _, _ in res.table)
    # schedule-0 candidates with device-memory twiddles are bit-exact with
    # the per-request PLAIN launches
    assert all(err == 0.0 for name, _, err, _, _, _ in res.table if "_sched0_" in name and name.endswith("_tw0"))
    assert res.best_ms == min(ms for _, ms, _, _, _, _ in res.table)
    torch.cuda.synchronize()


def test_autotune_serving_objective_trades_time_for_latency():
    """With latency_slack the kept configuration is within that fraction of
    the fastest gated time and has the lowest median request latency there;
    every gated row carries its measured median."""
    import torch
    from bfly import group
    specs = group.uneven_specs(count=64, seed=5)
    reqs = group.make_requests(specs, seed=5)
    slack = 0.05
    res = group.autotune(reqs, space={"workgroup": (512,), "schedule": (0,), "order": (0, 1, 3),
                                      "dynamic": (1,), "grid": (0,), "twiddles": (1,)}, reps=5,
                         latency_slack=slack)
    assert len(res.table) == 3
    gated = [row for row in res.table if row[2] <= REL_L2]
    assert gated and all(row[5] is not None and row[5] > 0 for row in gated)
    fastest = min(row[1] for row in gated)
    assert res.best_ms <= fastest * (1 + slack)
    near = [row for row in gated if row[1] <= fastest * (1 + slack)]
    assert res.best_p50_us == min(row[5] for row in near)
    assert res.best.name() in [row[0] for row in near]
    torch.cuda.synchronize()
