"""Parse and check the timing lines that `bfly_check` prints.

Same functions, field names and checks as the reference's
scripts/scopes.py:1-229 (parse_scopes, validate_scopes, parse_boundary_impl,
parse_segments, validate_segments, infer_mode), so a collector written
against the reference's driver reads this one unchanged.  The vocabulary:

  scopes: plan_setup=<us> us first_use=<us> us host_end_to_end mean=<us>
          min=<us> us h2d=<us|NA> device_chain=<us> d2h=<us|NA> reps=<k>

  h2d / d2h      the logical input upload / output download of one
                 execution; measured on long chains (C2C n > 4096), whose
                 every execution uploads its input and downloads its output,
                 and NA on the short path (input already resident)
  device_chain   the span of the execution's kernels, transfers excluded

Long chains also print `boundary_impl:` and `segments:`.  The reference's
device chain has six launches ("separate") or five ("fused"); this
library's has one fused launch ("single": n = 8192 / 16384 on one workgroup
per transform, segment fft1) or two ("fused2": the four-step column kernel
-- transpose in, N1-point FFTs, twiddle -- as fft1, and the row kernel --
N2-point FFTs, natural-order store -- as fft2).  All four forms are known
here, so both drivers' lines validate.
"""
import re

FIELDS = ("plan_setup", "first_use", "e2e_mean", "e2e_min",
          "h2d", "device_chain", "d2h", "reps")

SEGMENT_FIELDS = ("transpose_in", "fft1", "twiddle",
                  "transpose_boundary", "fft2", "transpose_out")

# segments each boundary implementation reports, in launch order
IMPL_SEGMENTS = {
    "separate": SEGMENT_FIELDS,
    "fused": tuple(f for f in SEGMENT_FIELDS if f != "twiddle"),
    "fused2": ("fft1", "fft2"),
    "single": ("fft1",),
}
IMPLS = tuple(IMPL_SEGMENTS)
FUSED_SEGMENT_FIELDS = IMPL_SEGMENTS["fused"]

_MODES = ("short", "long_host", "long_device")
_NUM = r"([0-9.eE+-]+)"


def parse_scopes(line):
    """One `scopes:` line -> dict of FIELDS (NA spans -> None).

    ValueError for another line, the retired `device_only` field, or a
    missing field.
    """
    if not isinstance(line, str) or not line.startswith("scopes:"):
        raise ValueError("not a scopes line: %r" % (line,))
    if "device_only" in line:
        raise ValueError("retired device_only scope; expect h2d/device_chain/d2h fields")
    fields = {}
    for key, pat in (("plan_setup", r"plan_setup=%s us" % _NUM),
                     ("first_use", r"first_use=%s us" % _NUM)):
        m = re.search(pat, line)
        if not m:
            raise ValueError("missing %s: %r" % (key, line))
        fields[key] = float(m.group(1))
    m = re.search(r"host_end_to_end mean=%s min=%s us" % (_NUM, _NUM), line)
    if not m:
        raise ValueError("missing host_end_to_end: %r" % (line,))
    fields["e2e_mean"], fields["e2e_min"] = float(m.group(1)), float(m.group(2))
    for key in ("h2d", "device_chain", "d2h"):
        m = re.search(r"\b%s=(NA|[0-9.eE+-]+(?: us)?)\b" % key, line)
        if not m:
            raise ValueError("missing %s: %r" % (key, line))
        fields[key] = None if m.group(1) == "NA" else float(m.group(1).split()[0])
    m = re.search(r"\breps=(\d+)\b", line)
    if not m:
        raise ValueError("missing reps: %r" % (line,))
    fields["reps"] = int(m.group(1))
    return fields


def validate_scopes(fields, mode, slack_frac=0.10, slack_us=200.0):
    """Problems (a list of strings; empty = valid) of parsed scopes for a
    mode in ("short", "long_host", "long_device").

    Every mode needs positive plan_setup / first_use / e2e and a measured
    device_chain; the short path reports h2d and d2h as NA; long chains
    measure both, and their wall-clock times must cover
    h2d + device_chain + d2h up to slack_frac * wall + slack_us.
    """
    if mode not in _MODES:
        return ["unknown mode %r (expected one of %r)" % (mode, _MODES)]
    problems = []
    for key in ("plan_setup", "first_use", "e2e_mean", "e2e_min"):
        v = fields.get(key)
        if v is None:
            problems.append("%s missing" % key)
        elif not v > 0:
            problems.append("%s must be > 0, got %r" % (key, v))
    chain = fields.get("device_chain")
    if chain is None:
        problems.append("device_chain must be measured in every mode")
    elif not chain > 0:
        problems.append("device_chain must be > 0, got %r" % (chain,))
    if fields.get("reps") is None or fields["reps"] < 1:
        problems.append("reps must be >= 1")
    if mode == "short":
        for key in ("h2d", "d2h"):
            if fields.get(key) is not None:
                problems.append("short path must report %s=NA, got %r" % (key, fields[key]))
        return problems
    for key in ("h2d", "d2h"):
        v = fields.get(key)
        if v is None:
            problems.append("%s must be measured on %s chains" % (key, mode))
        elif not v > 0:
            problems.append("%s must be > 0, got %r" % (key, v))
    budget = sum(v for v in (fields.get("h2d"), chain, fields.get("d2h")) if v is not None)
    for key in ("e2e_mean", "e2e_min"):
        wall = fields.get(key)
        if wall is not None and budget > wall + slack_frac * wall + slack_us:
            problems.append("%s=%r below h2d+device_chain+d2h=%r: transfers are not allowed "
                            "inside device_chain" % (key, wall, budget))
    return problems


def parse_boundary_impl(line):
    """`boundary_impl: <impl>` -> the impl name (one of IMPLS)."""
    if not isinstance(line, str) or not line.startswith("boundary_impl:"):
        raise ValueError("not a boundary_impl line: %r" % (line,))
    value = line[len("boundary_impl:"):].strip()
    if value not in IMPLS:
        raise ValueError("unknown boundary_impl %r (expected one of %r)" % (value, IMPLS))
    return value


def parse_segments(line):
    """`segments: ...` -> dict of per-launch spans in us, or None for NA."""
    if not isinstance(line, str) or not line.startswith("segments:"):
        raise ValueError("not a segments line: %r" % (line,))
    body = line[len("segments:"):].strip()
    if body == "NA":
        return None
    fields = {}
    for key in SEGMENT_FIELDS:
        m = re.search(r"\b%s=%s us\b" % (key, _NUM), body)
        if m:
            fields[key] = float(m.group(1))
    if not fields:
        raise ValueError("no segment fields: %r" % (line,))
    unknown = set(re.findall(r"([a-z0-9_]+)=", body)) - set(SEGMENT_FIELDS)
    if unknown:
        raise ValueError("unexpected segment field %r: %r" % (sorted(unknown)[0], line))
    return fields


def validate_segments(seg, fields, mode, impl="separate", rel_tol=0.01, abs_tol_us=20.0):
    """Problems of a parsed segments dict: NA off the device chain; on it,
    exactly the segments of `impl`, each >= 0, adding up to device_chain
    within max(rel_tol * chain, abs_tol_us) (the spans come from the same
    execution as device_chain)."""
    if mode not in _MODES:
        return ["unknown mode %r (expected one of %r)" % (mode, _MODES)]
    if impl not in IMPLS:
        return ["unknown impl %r (expected one of %r)" % (impl, IMPLS)]
    if mode != "long_device":
        return [] if seg is None else ["segments must be NA on %s, got %r" % (mode, seg)]
    expected = IMPL_SEGMENTS[impl]
    if seg is None:
        return ["long_device chain must report the %d %s segments" % (len(expected), impl)]
    problems = ["segment %s is not part of the %s chain" % (k, impl)
                for k in sorted(set(seg) - set(expected))]
    total = 0.0
    for key in expected:
        v = seg.get(key)
        if not isinstance(v, (int, float)):
            problems.append("segment %s missing" % key)
        elif not v >= 0:
            problems.append("segment %s must be >= 0, got %r" % (key, v))
        else:
            total += v
    chain = fields.get("device_chain")
    if chain is not None and total > 0.0 and abs(total - chain) > max(rel_tol * chain, abs_tol_us):
        problems.append("sum(segments)=%r != device_chain=%r: the segment events must telescope "
                        "inside the chain span" % (total, chain))
    return problems


def infer_mode(fields, dev_boundary=True):
    """Mode of a parsed scopes dict: short when h2d is NA, else the long
    chain on the device (bfly_check's only long chain) or on the host."""
    if fields.get("h2d") is None:
        return "short"
    return "long_device" if dev_boundary else "long_host"


def parse_report(text):
    """All of one bfly_check run's report lines: a dict with "scopes",
    "boundary_impl" (None when absent), "segments", "seq" (list of
    (token, maxRel, status, stale)), "max_rel", "launches", "us_per_call"
    and "verdict" ("PASS" / "FAIL")."""
    out = {"scopes": None, "boundary_impl": None, "segments": None, "seq": [],
           "max_rel": None, "launches": None, "us_per_call": None, "verdict": None}
    for line in text.splitlines():
        line = line.strip()
        if line.startswith("scopes:"):
            out["scopes"] = parse_scopes(line)
        elif line.startswith("boundary_impl:"):
            out["boundary_impl"] = parse_boundary_impl(line)
        elif line.startswith("segments:"):
            out["segments"] = parse_segments(line)
        elif line.startswith("seq["):
            m = re.match(r"seq\[\d+\]=(.*) maxRel=%s (PASS|FAIL)( STALE-OUTPUT)?$" % _NUM, line)
            if m:
                out["seq"].append((m.group(1), float(m.group(2)), m.group(3), bool(m.group(4))))
        elif line.startswith("n=") and "maxRel=" in line:
            out["max_rel"] = float(re.search(r"maxRel=%s" % _NUM, line).group(1))
            m = re.search(r"launches=(\d+)", line)
            out["launches"] = int(m.group(1)) if m else None
        elif re.match(r"bfly_(c2c|r2c|c2r): ", line):
            out["us_per_call"] = float(re.search(r": %s us/call" % _NUM, line).group(1))
        elif line in ("PASS", "FAIL"):
            out["verdict"] = line
    return out
