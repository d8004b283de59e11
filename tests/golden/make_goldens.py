"""Write tests/golden/reference_vectors.npz from the reference's own code.

* Runs the reference's host sources compiled into oracle/_ref/libbfly_ref.so
  (reference.cpp and the header-only fft_k.hpp / long_fft_ub.h, built with
  plain g++ against the reference's include/ tree) on small seeded inputs.
* Runs the reference's Python twin of the application input generators
  (scripts/bench_native_npu.py:27-84, `_xh` and `app_input`, stated there to
  be bit-identical to src/host/fft_check.cpp:63-159).  That script imports
  torch_npu at module level, so only those two numpy-only function
  definitions are compiled from its text, in memory; nothing of it is
  written into this repository.

Inputs and outputs are stored as data, so the oracle's restatement stays
pinned where the reference tree is absent.

    python tests/golden/make_goldens.py
"""
import ast
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
import oracle  # noqa: E402

OUT = Path(__file__).resolve().parent / "reference_vectors.npz"
TWIN = oracle.REFERENCE_TREE / "scripts" / "bench_native_npu.py"


def _load_twin():
    """`_xh` and `app_input` of the reference's Python input twin."""
    tree = ast.parse(TWIN.read_text())
    keep = [node for node in tree.body if isinstance(node, ast.FunctionDef) and node.name in ("_xh", "app_input")]
    if len(keep) != 2:
        raise SystemExit(f"{TWIN}: _xh / app_input not found")
    ns = {"np": np}
    exec(compile(ast.Module(body=keep, type_ignores=[]), str(TWIN), "exec"), ns)
    return ns["_xh"], ns["app_input"]


def main():
    r = oracle.ref()
    if r is None:
        raise SystemExit("oracle/_ref is not built (needs the reference tree)")
    data = {}
    # refFftF32 / refR2CF32 / refC2RF32 on seeded rows
    for n in (64, 256, 1024, 4096):
        x = oracle.uniform_input(11, (2, 2 * n))
        y = np.stack([np.empty(2 * n, np.float32) for _ in range(2)])
        for b in range(2):
            row = np.empty(2 * n, np.float32)
            r.ref_fft_f32(np.ascontiguousarray(x[b]), row, n)
            y[b] = row
        data[f"fft_in_{n}"] = x
        data[f"fft_out_{n}"] = y
    for n in (128, 1024, 8192):
        x = oracle.uniform_input(12, (2, n))
        h = np.empty((2, n + 2), np.float32)
        for b in range(2):
            row = np.empty(n + 2, np.float32)
            r.ref_r2c_f32(np.ascontiguousarray(x[b]), row, n)
            h[b] = row
        data[f"r2c_in_{n}"] = x
        data[f"r2c_out_{n}"] = h
    for n in (64, 1024, 4096):
        h = oracle.half_input(n, 2)
        y = np.empty((2, n), np.float32)
        for b in range(2):
            row = np.empty(n, np.float32)
            r.ref_c2r_f32(np.ascontiguousarray(h[b]), row, n)
            y[b] = row
        data[f"c2r_in_{n}"] = h
        data[f"c2r_out_{n}"] = y
    # integer rules
    ns = [1 << k for k in range(3, 17)]
    data["rule_n"] = np.array(ns, np.uint32)
    data["plane_k"] = np.array([r.ref_plane_k_for(n) for n in ns], np.uint32)
    batches = [1, 3, 47, 48, 49, 191, 192, 193, 4096, 0xFFFFFFFF]
    data["fold_batches"] = np.array(batches, np.uint32)
    data["fold_d"] = np.array([[r.ref_fold_d_for(n, b, 48) for b in batches] for n in ns], np.uint32)
    # application inputs and the hash, from the reference's Python twin
    xh, app = _load_twin()
    for mode in ("ofdm", "radar", "dl"):
        for n, b in ((64, 3), (256, 14), (2048, 5)):
            re, im = app(mode, n, b)
            data[f"app_{mode}_{n}_{b}"] = np.stack([re, im], axis=-1).reshape(b, 2 * n).astype(np.float32)
    bs = np.array([0, 1, 7, 4095, 296369474, 0xFFFFFFFF], np.uint32)[:, None]
    ks = np.array([0, 1, 2, 40503, 65535, 0x9E37, 0x5BD1, 0xFFFFFFFF], np.uint32)[None, :]
    data["hash_b"], data["hash_k"], data["hash_out"] = bs.ravel(), ks.ravel(), xh(bs, ks)
    # patternAt
    data["pattern_0_4096"] = np.array([r.ref_pattern_at(i) for i in range(4096)], np.float32)
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({len(data)} arrays)")


if __name__ == "__main__":
    main()
