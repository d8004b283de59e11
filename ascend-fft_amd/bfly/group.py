"""Grouped launch: many uneven C2C / R2C / C2R requests in one persistent
kernel (bfly_group_* of include/bfly_hip.h), its hipGraph capture, and the
autotuner that picks the launch configuration.

A request is (kind, n, batch, sign) with device input and output tensors in
the kind's execute layout:
  C2C  float32 [batch, 2n] interleaved  -> float32 [batch, 2n]
  R2C  float32 [batch, n]               -> float32 [batch, n+2]
  C2R  float32 [batch, n+2]             -> float32 [batch, n]   (scaled 1/n)

The configuration axes the tuner searches are the ones the kernel exposes:
threads per workgroup (512 / 1024, which fixes the transforms per workgroup
at 512*16/n' or 1024*16/n'), radix schedule (radix 16 first or the
remainder radix first), unit order, static or atomic-counter scheduling,
the number of persistent workgroups and where the inter-pass twiddles come
from (device memory, or the two-level tables staged in LDS per unit).
"""
from __future__ import annotations

import ctypes as C
import itertools
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib

KIND_NAMES = {_lib.KIND_C2C: "c2c", _lib.KIND_R2C: "r2c", _lib.KIND_C2R: "c2r"}


@dataclass(frozen=True)
class Config:
    workgroup: int = 512
    schedule: int = 0
    order: int = 0
    dynamic: int = 0
    grid: int = 0
    twiddles: int = 0

    def as_c(self) -> _lib.GroupConfig:
        return _lib.GroupConfig(self.workgroup, self.schedule, self.order, self.dynamic, self.grid,
                                self.twiddles)

    def name(self) -> str:
        return (f"wg{self.workgroup}_sched{self.schedule}_order{self.order}_"
                f"{'dyn' if self.dynamic else 'static'}_grid{self.grid or 'res'}_tw{self.twiddles}")

    def valid(self) -> bool:
        """LDS twiddle tables run radix schedule 0 only."""
        return not (self.twiddles == 1 and self.schedule != 0)

    def reference_variant(self) -> int:
        """The per-request variant whose kernels compute the same bits."""
        return _lib.VARIANT_STREAM_LT if self.twiddles == 1 else _lib.VARIANT_PLAIN


@dataclass
class Request:
    kind: int
    n: int
    batch: int
    sign: int = -1
    x: object = None  # device input tensor
    y: object = None  # device output tensor

    def in_floats(self) -> int:
        return self.batch * (2 * self.n if self.kind == _lib.KIND_C2C else
                             self.n if self.kind == _lib.KIND_R2C else self.n + 2)

    def out_floats(self) -> int:
        return self.batch * (2 * self.n if self.kind == _lib.KIND_C2C else
                             self.n + 2 if self.kind == _lib.KIND_R2C else self.n)

    def algorithmic_bytes(self) -> int:
        """One read of the input plus one write of the output."""
        return 4 * (self.in_floats() + self.out_floats())


def uneven_specs(count: int = 512, seed: int = 3, n_min: int = 64, n_max: int = 8192,
                 batch_max: int = 256) -> list[tuple[int, int, int]]:
    """(kind, n, batch) of the uneven batched workload: n log-uniform over the
    powers of two n_min..n_max, batch uniform on 1..batch_max, C2C or R2C with
    probability 1/2 each; numpy default_rng(seed), drawn per request in the
    order kind, log2 n, batch."""
    rng = np.random.default_rng(seed)
    lo, hi = int(math.log2(n_min)), int(math.log2(n_max))
    out = []
    for _ in range(count):
        kind = _lib.KIND_C2C if rng.random() < 0.5 else _lib.KIND_R2C
        n = 1 << int(rng.integers(lo, hi + 1))
        b = int(rng.integers(1, batch_max + 1))
        out.append((kind, n, b))
    return out


def make_requests(specs, device="cuda", seed: int = 3, signs=None) -> list[Request]:
    """Allocates the requests' tensors with inputs uniform on [-1, 1) from a
    torch generator on the device (seeded once); C2R inputs get zero imaginary
    DC and Nyquist bins so every path reads the same Hermitian spectrum."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    reqs = []
    for i, (kind, n, b) in enumerate(specs):
        r = Request(kind, n, b, -1 if signs is None else signs[i])
        x = torch.rand(r.in_floats(), generator=g, device=device, dtype=torch.float32) * 2 - 1
        if kind == _lib.KIND_C2R:
            x = x.view(b, n + 2)
            x[:, 1] = 0
            x[:, n + 1] = 0
        r.x = x.reshape(-1).contiguous()
        r.y = torch.empty(r.out_floats(), device=device, dtype=torch.float32)
        reqs.append(r)
    return reqs


class Group:
    """A created group (request table + unit list + twiddles on the device)."""

    def __init__(self, reqs: list[Request], cfg: Config = Config()):
        self.reqs = reqs
        self.cfg = cfg
        arr = (_lib.GroupRequest * len(reqs))()
        for i, r in enumerate(reqs):
            arr[i] = _lib.GroupRequest(r.x.data_ptr(), r.y.data_ptr(), r.kind, r.sign, r.n, r.batch)
        h = C.c_void_p()
        ccfg = cfg.as_c()
        _lib.check(_lib.hip().bfly_group_create(arr, len(reqs), C.byref(ccfg), C.byref(h)),
                   f"bfly_group_create({cfg.name()})")
        self.handle = h
        units, grid, nbytes = C.c_uint32(), C.c_uint32(), C.c_size_t()
        _lib.check(_lib.hip().bfly_group_query(h, C.byref(units), C.byref(grid), C.byref(nbytes)),
                   "bfly_group_query")
        self.units, self.grid, self.device_bytes = units.value, grid.value, nbytes.value

    def launch(self, stamps=None) -> None:
        import torch
        s = torch.cuda.current_stream().cuda_stream
        rc = _lib.hip().bfly_group_launch(self.handle, stamps.data_ptr() if stamps is not None else None, s)
        _lib.check(rc, "bfly_group_launch")

    def close(self) -> None:
        if self.handle is not None and self.handle.value:
            _lib.check(_lib.hip().bfly_group_destroy(self.handle), "bfly_group_destroy")
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def algorithmic_bytes(self) -> int:
        return sum(r.algorithmic_bytes() for r in self.reqs)

    def transforms(self) -> int:
        return sum(r.batch for r in self.reqs)


def launch_per_request(reqs: list[Request], variant: int = _lib.VARIANT_PLAIN) -> None:
    """The same requests as one library call each (bfly_c2c_ex / bfly_real_ex
    of the given variant) on the current stream."""
    import torch
    from . import ops
    lib = _lib.hip()
    s = torch.cuda.current_stream().cuda_stream
    for r in reqs:
        tab = ops.table(r.kind, r.n, r.x.device)
        if r.kind == _lib.KIND_C2C:
            work = ops.workspace(r.kind, r.n, r.batch, r.x.device)
            rc = lib.bfly_c2c_ex(r.x.data_ptr(), r.y.data_ptr(), r.n, r.batch, r.sign, variant,
                                 tab.data_ptr(), work.data_ptr() if work is not None else None, s)
        else:
            rc = lib.bfly_real_ex(r.kind, r.x.data_ptr(), r.y.data_ptr(), r.n, r.batch, r.n // 2 + 1, r.n + 2,
                                  variant, 0, tab.data_ptr(), s)
        _lib.check(rc, f"per-request {KIND_NAMES[r.kind]} n={r.n} batch={r.batch}")


def capture(fn, warmup: int = 1):
    """hipGraph of fn() (torch.cuda.CUDAGraph is hipGraph on ROCm); the
    returned graph's replay() re-runs the captured launches."""
    import torch
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(warmup):
            fn()
    torch.cuda.current_stream().wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        fn()
    return g


def time_ms(fn, reps: int = 20, warmup: int = 3) -> float:
    """Mean ms per call of fn() over reps calls, HIP events on the current
    stream (the stream the library launches on)."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def outputs(reqs: list[Request]):
    return [r.y.clone() for r in reqs]


def max_rel_diff(a, b) -> float:
    """Largest per-request ||a-b|| / ||b|| (device tensors, computed in fp64)."""
    worst = 0.0
    for x, y in zip(a, b):
        d = (x.double() - y.double()).norm().item()
        ref = max(y.double().norm().item(), 1e-300)
        worst = max(worst, d / ref)
    return worst


def wall_clock_khz() -> float:
    import json
    from . import ops
    return float(json.loads(ops.device_facts()).get("wall_clock_khz", 0) or 0)


def latencies_us(group: Group) -> np.ndarray:
    """Per-request completion times (us after the earliest workgroup start)
    of one stamped launch."""
    import torch
    stamps = torch.zeros(len(group.reqs) + 1, dtype=torch.int64, device="cuda")
    group.launch(stamps)
    torch.cuda.synchronize()
    st = stamps.cpu().numpy().view(np.uint64)
    khz = wall_clock_khz()
    if khz <= 0:
        raise RuntimeError("device wall-clock rate unknown")
    return (st[1:].astype(np.float64) - float(st[0])) / khz * 1e3


def per_request_latencies_us(reqs: list[Request], variant: int = _lib.VARIANT_PLAIN) -> np.ndarray:
    """Completion time of each request (us after the first launch starts)
    when the requests are issued as back-to-back per-request launches on one
    stream: an event after every request."""
    import torch
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(len(reqs) + 1)]
    torch.cuda.synchronize()
    evs[0].record()
    for i, r in enumerate(reqs):
        launch_per_request([r], variant)
        evs[i + 1].record()
    evs[-1].synchronize()
    return np.array([evs[0].elapsed_time(e) * 1e3 for e in evs[1:]])


DEFAULT_SPACE = {
    "workgroup": (512, 1024),
    "schedule": (0, 1),
    "order": (0, 1, 2, 3),
    "dynamic": (0, 1),
    "grid": (0,),
    "twiddles": (0, 1),
}


@dataclass
class TuneResult:
    best: Config
    best_ms: float
    table: list = field(default_factory=list)  # (config name, ms, max rel diff, units, grid, p50 us or None)
    best_p50_us: float | None = None


def pick(rows: list, gate: float, latency_slack: float | None = None) -> int:
    """Index of the configuration to keep among `rows` of (ms, max rel diff,
    p50 latency us or None).  Throughput objective (latency_slack None): the
    fastest row within the accuracy gate.  Serving objective: among the
    gated rows at most (1 + latency_slack) times slower than the fastest,
    the one with the lowest median request latency (ties: the faster)."""
    ok = [i for i, (ms, err, _) in enumerate(rows) if err <= gate]
    if not ok:
        raise RuntimeError("no grouped configuration met the accuracy gate")
    fastest = min(ok, key=lambda i: rows[i][0])
    if latency_slack is None:
        return fastest
    if latency_slack < 0:
        raise ValueError("latency_slack must be >= 0")
    limit = rows[fastest][0] * (1.0 + latency_slack)
    near = [i for i in ok if rows[i][0] <= limit and rows[i][2] is not None]
    if not near:
        return fastest
    return min(near, key=lambda i: (rows[i][2], rows[i][0]))


def autotune(reqs: list[Request], space: dict | None = None, reps: int = 10,
             gate: float = 1e-5, reference=None, latency_slack: float | None = None) -> TuneResult:
    """Times every configuration of `space` on the same requests and keeps
    the fastest whose outputs stay within `gate` (relative L2 per request) of
    `reference` (default: per-request PLAIN launches).  Each candidate is
    timed as a captured hipGraph replay, the form a serving loop runs.  With
    `latency_slack` (e.g. 0.05) every gated candidate's per-request
    completion times are also measured (one stamped launch) and the kept one
    is the lowest median latency within that fraction of the fastest time
    (`pick`): a serving loop trades a few percent of throughput for its short
    requests finishing first."""
    import torch
    space = dict(DEFAULT_SPACE if space is None else space)
    if reference is None:
        launch_per_request(reqs)
        torch.cuda.synchronize()
        reference = outputs(reqs)
    table, cfgs = [], []
    keys = list(space)
    for vals in itertools.product(*(space[k] for k in keys)):
        cfg = Config(**dict(zip(keys, vals)))
        if not cfg.valid() or any(not _lib.hip().bfly_group_supported(r.kind, r.n, cfg.workgroup)
                                  for r in reqs):
            continue
        grp = Group(reqs, cfg)
        for r in reqs:
            r.y.zero_()
        grp.launch()
        torch.cuda.synchronize()
        err = max_rel_diff(outputs(reqs), reference)
        graph = capture(grp.launch)
        ms = time_ms(graph.replay, reps=reps)
        p50 = None
        if latency_slack is not None and err <= gate:
            p50 = float(np.percentile(latencies_us(grp), 50))
        table.append((cfg.name(), ms, err, grp.units, grp.grid, p50))
        cfgs.append(cfg)
        del graph
        grp.close()
    i = pick([(ms, err, p50) for _, ms, err, _, _, p50 in table], gate, latency_slack)
    return TuneResult(cfgs[i], table[i][1], table, table[i][5])
