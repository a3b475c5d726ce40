"""Write the golden triplet-assembly fixtures from the reference's own code.

Runs ``wp_bsr_matrix_from_triplets_host`` from ``oracle/_ref/libwarp_ref.so``
(the reference's CPU native library, built from its unmodified sources by
``make -C oracle ref``) on seeded inputs and stores inputs + outputs as
``triplets_ref.npz``.  The reference host path sorts with std::sort, so only
the topology (offsets, columns, run ends) and, per run, the *set* of triplet
indices are reference-defined; the fixture stores the sorted set per run.

Usage: python tests/golden/make_golden.py
"""

from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

CASES = [
    # name, nrow, ncol, T, block shape, dtype, zero fraction, out-of-range fraction, count
    ("csr_8x6", 8, 6, 100, (), np.float32, 0.0, 0.0, None),
    ("bsr_3x2", 4, 9, 50, (3, 2), np.float32, 0.1, 0.0, None),
    ("bsr_3x3_oob", 37, 41, 2000, (3, 3), np.float32, 0.05, 0.1, None),
    ("bsr_3x3_count", 50, 50, 3000, (3, 3), np.float64, 0.02, 0.05, 2222),
    ("csr_dense_row", 3, 5000, 9000, (), np.float32, 0.0, 0.0, None),
    ("csr_many", 1000, 700, 20000, (), np.float64, 0.01, 0.02, None),
]


def make_case(rng, nrow, ncol, T, bshape, dtype, zero_frac, oob_frac):
    rows = rng.integers(0, nrow, size=T).astype(np.int32)
    cols = rng.integers(0, ncol, size=T).astype(np.int32)
    oob = rng.random(T) < oob_frac
    rows[oob & (rng.random(T) < 0.5)] = -1
    cols[oob & (rng.random(T) < 0.5)] = ncol + 3
    vals = rng.standard_normal(size=(T, *bshape)).astype(dtype)
    z = rng.random(T) < zero_frac
    vals[z] = 0.0
    # signed zeros count as zero
    if bshape:
        vals[z & (rng.random(T) < 0.5), 0, 0] = -0.0
    return rows, cols, vals


def main():
    if oracle.ref_lib() is None:
        oracle.build(ref=True)
    rng = np.random.default_rng(2024)
    out = {}
    for name, nrow, ncol, T, bshape, dtype, zf, of, count in CASES:
        rows, cols, vals = make_case(rng, nrow, ncol, T, bshape, dtype, zf, of)
        ref = oracle.ref_bsr_from_triplets(nrow, ncol, rows, cols, vals, prune=True, count=count)
        nnz = ref["nnz"]
        so = ref["summed_offsets"]
        runs = np.concatenate([np.sort(ref["summed_indices"][(0 if u == 0 else so[u - 1]) : so[u]]) for u in range(nnz)]) if nnz else np.zeros(0, np.int32)
        out[f"{name}/shape"] = np.array([nrow, ncol, T, -1 if count is None else count], np.int64)
        out[f"{name}/rows"] = rows
        out[f"{name}/cols"] = cols
        out[f"{name}/vals"] = vals
        out[f"{name}/offsets"] = ref["offsets"]
        out[f"{name}/columns"] = ref["columns"]
        out[f"{name}/summed_offsets"] = so
        out[f"{name}/run_indices_sorted"] = runs.astype(np.int32)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "triplets_ref.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(CASES)} cases)")


if __name__ == "__main__":
    main()
