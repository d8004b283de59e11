"""Write tests/golden/reference_vectors.npz from the reference's own code.

Runs the reference's host sources compiled into oracle/_ref/libbfly_ref.so
(reference.cpp, butterfly.cpp Generator tables, fft_k.hpp, long_fft_ub.h)
on small seeded inputs and stores inputs and outputs as data, so the
oracle's restatement stays pinned where the reference tree is absent.

    python tests/golden/make_goldens.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
import oracle  # noqa: E402

OUT = Path(__file__).resolve().parent / "reference_vectors.npz"


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
    # Generator tables
    for n in (64, 256, 4096):
        twr = np.empty(n + 16, np.float32)
        twi = np.empty(n + 16, np.float32)
        r.ref_gen_twiddles(n, twr, twi)
        data[f"gen_twr_{n}"], data[f"gen_twi_{n}"] = twr, twi
        for name, fn in (("bitrev", r.ref_gen_bit_reverse), ("plane", r.ref_gen_bit_reverse_plane),
                         ("transpose", r.ref_gen_transpose_plane)):
            d = np.empty(n, np.uint32)
            fn(n, d)
            data[f"gen_{name}_{n}"] = d
        for D in (1, 4):
            d = np.empty(2 * n * D, np.uint32)
            r.ref_gen_interleave(n, D, d)
            data[f"gen_interleave_{n}_{D}"] = d
    # patternAt
    data["pattern_0_4096"] = np.array([r.ref_pattern_at(i) for i in range(4096)], np.float32)
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({len(data)} arrays)")


if __name__ == "__main__":
    main()
