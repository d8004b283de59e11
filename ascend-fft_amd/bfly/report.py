"""Reader and checker for the report that `bfly_check` prints.

The printed vocabulary is the reference driver's (src/host/fft_check.cpp:
1186-1315 and scripts/scopes.py:1-60 of the reference), so a collector that
reads one driver's lines reads the other's:

    scopes: plan_setup=<us> us first_use=<us> us host_end_to_end mean=<us>
            min=<us> us h2d=<us|NA> device_chain=<us> d2h=<us|NA> reps=<k>
    boundary_impl: <chain form>
    segments: <launch>=<us> us ... | NA

How the lines are read and what is checked is this library's own:

* every line body is read by one token reader: ``key=value`` pairs, the unit
  word ``us`` closing a value, and a bare word labelling the pairs up to the
  next unit (``host_end_to_end mean=.. min=.. us``);
* a report has one of two transfer modes: ``resident`` (the input already
  sits in device memory; h2d / d2h are NA) or ``staged`` (every execution
  uploads its input and downloads its output: the long chains);
* the checks are rows of a table (what, when, test, message) evaluated
  against the typed records.

Chain forms and the launches each one reports as segments, in launch order:
``single`` (one workgroup per transform: fft1), ``fused2`` (this library's
four-step: the column kernel fft1, the row kernel fft2), and the reference's
``fused`` (five launches) and ``separate`` (six) so that its lines are
understood too.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Tuple

CHAINS: Dict[str, Tuple[str, ...]] = {
    "single": ("fft1",),
    "fused2": ("fft1", "fft2"),
    "fused": ("transpose_in", "fft1", "transpose_boundary", "fft2", "transpose_out"),
    "separate": ("transpose_in", "fft1", "twiddle", "transpose_boundary", "fft2", "transpose_out"),
}
LAUNCH_NAMES = CHAINS["separate"]
MODES = ("resident", "staged")
SCOPE_KEYS = ("plan_setup", "first_use", "host_end_to_end.mean", "host_end_to_end.min",
              "h2d", "device_chain", "d2h", "reps")


class ReportError(ValueError):
    """A report line that does not follow the vocabulary."""


def read_pairs(body: str) -> Dict[str, Optional[str]]:
    """``key=value`` pairs of a line body as {qualified key: value text}.

    A bare word other than the unit ``us`` labels the pairs that follow it
    until the next ``us`` (their keys become ``label.key``); ``NA`` values
    are returned as None.  A key seen twice is an error.
    """
    pairs: Dict[str, Optional[str]] = {}
    label = ""
    for word in body.split():
        if word == "us":
            label = ""
            continue
        if "=" not in word:
            if label:
                raise ReportError(f"two labels in a row ({label!r}, {word!r}) in {body!r}")
            label = word
            continue
        key, _, value = word.partition("=")
        if not key or not value:
            raise ReportError(f"malformed pair {word!r} in {body!r}")
        key = f"{label}.{key}" if label else key
        if key in pairs:
            raise ReportError(f"{key} appears twice in {body!r}")
        pairs[key] = None if value == "NA" else value
    return pairs


def _number(key: str, text: Optional[str]) -> Optional[float]:
    if text is None:
        return None
    try:
        return float(text)
    except ValueError:
        raise ReportError(f"{key}={text!r} is not a number") from None


def _body(line: str, tag: str) -> str:
    if not isinstance(line, str) or not line.startswith(tag + ":"):
        raise ReportError(f"expected a {tag}: line, got {line!r}")
    return line[len(tag) + 1:]


@dataclass(frozen=True)
class Scopes:
    """The ``scopes:`` line; spans in microseconds, None where NA."""
    plan_setup: float
    first_use: float
    e2e_mean: float
    e2e_min: float
    h2d: Optional[float]
    device_chain: Optional[float]
    d2h: Optional[float]
    reps: int

    @classmethod
    def parse(cls, line: str) -> "Scopes":
        pairs = read_pairs(_body(line, "scopes"))
        extra = sorted(set(pairs) - set(SCOPE_KEYS))
        if extra:
            raise ReportError(f"scopes line has fields outside the vocabulary: {', '.join(extra)}")
        absent = [k for k in SCOPE_KEYS if k not in pairs]
        if absent:
            raise ReportError(f"scopes line lacks {', '.join(absent)}")
        reps = pairs["reps"]
        if reps is None or not reps.isdigit():
            raise ReportError(f"reps={reps!r} is not a count")
        num = {k: _number(k, pairs[k]) for k in SCOPE_KEYS if k != "reps"}
        return cls(num["plan_setup"], num["first_use"], num["host_end_to_end.mean"],
                   num["host_end_to_end.min"], num["h2d"], num["device_chain"], num["d2h"], int(reps))

    @property
    def mode(self) -> str:
        """``staged`` when the execution moved data across PCIe, else ``resident``."""
        return "resident" if self.h2d is None and self.d2h is None else "staged"


def parse_chain(line: str) -> str:
    """``boundary_impl: <form>`` -> the chain form (a key of CHAINS)."""
    form = _body(line, "boundary_impl").strip()
    if form not in CHAINS:
        raise ReportError(f"chain form {form!r} is none of {', '.join(CHAINS)}")
    return form


def parse_segments(line: str) -> Optional[Dict[str, float]]:
    """``segments:`` -> {launch: us}, or None for ``segments: NA``."""
    body = _body(line, "segments").strip()
    if body == "NA":
        return None
    pairs = read_pairs(body)
    if not pairs:
        raise ReportError(f"segments line names no launch: {line!r}")
    foreign = sorted(set(pairs) - set(LAUNCH_NAMES))
    if foreign:
        raise ReportError(f"segments line names launches no chain has: {', '.join(foreign)}")
    return {k: _number(k, v) for k, v in pairs.items()}


# --------------------------------------------------------------------- checks
@dataclass(frozen=True)
class Tolerance:
    """How far timing figures may disagree before a check fails."""
    wall_frac: float = 0.10      # wall clock may undershoot h2d + chain + d2h by this share ...
    wall_us: float = 200.0       # ... plus this many microseconds
    segments_frac: float = 0.01  # launch spans may miss the chain span by this share ...
    segments_us: float = 20.0    # ... or by this many microseconds, whichever is larger


@dataclass(frozen=True)
class Rule:
    what: str
    modes: Tuple[str, ...]
    broken: Callable[["Scopes", Tolerance], bool]
    message: Callable[["Scopes", Tolerance], str]


def _positive(v) -> bool:
    return v is not None and v > 0


def _wall_short(s: Scopes, t: Tolerance, wall: float) -> bool:
    spans = sum(v for v in (s.h2d, s.device_chain, s.d2h) if v is not None)
    return spans > wall * (1.0 + t.wall_frac) + t.wall_us


SCOPE_RULES: Tuple[Rule, ...] = (
    Rule("host timings", MODES,
         lambda s, t: not all(_positive(v) for v in (s.plan_setup, s.first_use, s.e2e_mean, s.e2e_min)),
         lambda s, t: (f"plan_setup / first_use / host_end_to_end must all be positive "
                       f"({s.plan_setup}, {s.first_use}, {s.e2e_mean}, {s.e2e_min})")),
    Rule("kernel span", MODES, lambda s, t: not _positive(s.device_chain),
         lambda s, t: f"device_chain is {s.device_chain}: every report needs a positive kernel span"),
    Rule("repetitions", MODES, lambda s, t: s.reps < 1, lambda s, t: f"reps is {s.reps}, needs >= 1"),
    Rule("resident transfers", ("resident",), lambda s, t: s.h2d is not None or s.d2h is not None,
         lambda s, t: f"a resident run moves no data over PCIe, yet h2d={s.h2d} d2h={s.d2h} (want NA)"),
    Rule("staged transfers", ("staged",), lambda s, t: not (_positive(s.h2d) and _positive(s.d2h)),
         lambda s, t: f"a staged run must time both copies, got h2d={s.h2d} d2h={s.d2h}"),
    Rule("wall clock covers the spans", ("staged",),
         lambda s, t: _wall_short(s, t, s.e2e_mean) or _wall_short(s, t, s.e2e_min),
         lambda s, t: (f"h2d + device_chain + d2h = {(s.h2d or 0) + (s.device_chain or 0) + (s.d2h or 0):.1f} us "
                       f"exceeds the host wall clock (mean {s.e2e_mean}, min {s.e2e_min}): "
                       f"copies were counted inside device_chain")),
)


def scope_problems(s: Scopes, mode: Optional[str] = None, tol: Tolerance = Tolerance()) -> List[str]:
    """Broken rules of a scopes record in `mode` (default: the record's own)."""
    mode = s.mode if mode is None else mode
    if mode not in MODES:
        return [f"mode {mode!r} is none of {', '.join(MODES)}"]
    return [r.message(s, tol) for r in SCOPE_RULES if mode in r.modes and r.broken(s, tol)]


def segment_problems(seg: Optional[Dict[str, float]], s: Scopes, chain: Optional[str],
                     mode: Optional[str] = None, tol: Tolerance = Tolerance()) -> List[str]:
    """Broken rules of a segments record: resident runs report NA; a staged
    run reports exactly its chain's launches, none negative, adding up to
    device_chain (the spans are events inside the same execution)."""
    mode = s.mode if mode is None else mode
    if mode not in MODES:
        return [f"mode {mode!r} is none of {', '.join(MODES)}"]
    if mode == "resident":
        return [] if seg is None else [f"a resident run has no chain launches to split, got {seg}"]
    if chain not in CHAINS:
        return [f"chain form {chain!r} is none of {', '.join(CHAINS)}"]
    want = CHAINS[chain]
    if seg is None:
        return [f"a staged {chain} run must time its launches {', '.join(want)}"]
    out = [f"launch {k} does not belong to the {chain} chain" for k in sorted(set(seg) - set(want))]
    out += [f"launch {k} of the {chain} chain is not timed" for k in want if k not in seg]
    out += [f"launch {k} has a negative span {seg[k]}" for k in want if k in seg and not seg[k] >= 0]
    total = sum(seg[k] for k in want if k in seg and seg[k] >= 0)
    chain_us = s.device_chain
    if chain_us and total > 0 and abs(total - chain_us) > max(tol.segments_frac * chain_us, tol.segments_us):
        out.append(f"launch spans add up to {total:.1f} us but device_chain is {chain_us:.1f} us")
    return out


# --------------------------------------------------------------------- report
@dataclass
class SeqStep:
    """One input of an AB_INPUT_SEQ run: its token, error, verdict and whether
    the output failed to follow the input (stale)."""
    token: str
    max_rel: float
    verdict: str
    stale: bool


@dataclass
class Report:
    scopes: Optional[Scopes] = None
    chain: Optional[str] = None
    segments: Optional[Dict[str, float]] = None
    seq: List[SeqStep] = field(default_factory=list)
    max_rel: Optional[float] = None
    launches: Optional[int] = None
    us_per_call: Optional[float] = None
    verdict: Optional[str] = None

    _SEQ = re.compile(r"seq\[\d+\]=(\S+) maxRel=(\S+) (PASS|FAIL)( STALE-OUTPUT)?$")

    @classmethod
    def parse(cls, text: str) -> "Report":
        rep = cls()
        for raw in text.splitlines():
            line = raw.strip()
            head = line.split(":", 1)[0] if ":" in line else ""
            if head == "scopes":
                rep.scopes = Scopes.parse(line)
            elif head == "boundary_impl":
                rep.chain = parse_chain(line)
            elif head == "segments":
                rep.segments = parse_segments(line)
            elif line.startswith("seq["):
                m = cls._SEQ.match(line)
                if m:
                    rep.seq.append(SeqStep(m.group(1), float(m.group(2)), m.group(3), bool(m.group(4))))
            elif line.startswith("n=") and "maxRel=" in line:
                fields = dict(w.split("=", 1) for w in line.split() if "=" in w)
                rep.max_rel = float(fields["maxRel"])
                rep.launches = int(fields["launches"]) if "launches" in fields else None
            elif re.match(r"bfly_(c2c|r2c|c2r): ", line):
                rep.us_per_call = float(line.split(":", 1)[1].split()[0])
            elif line in ("PASS", "FAIL"):
                rep.verdict = line
        return rep

    @property
    def mode(self) -> Optional[str]:
        return self.scopes.mode if self.scopes else None

    def problems(self, tol: Tolerance = Tolerance()) -> List[str]:
        """Every broken rule of the report, empty when it is well formed."""
        if self.scopes is None:
            return ["no scopes line"]
        out = scope_problems(self.scopes, tol=tol)
        chain = self.chain if self.mode == "staged" else None
        out += segment_problems(self.segments, self.scopes, chain, tol=tol)
        if self.verdict is None:
            out.append("no PASS / FAIL line")
        if self.max_rel is None or not math.isfinite(self.max_rel):
            out.append(f"maxRel {self.max_rel} is not a finite error")
        return out
