"""Parity metrics between the GPU CSR and the oracle's CSR.

Tolerance (values are FP64): for every stored entry
    |K_gpu - K_ref| <= 1e-12 * |K_ref| + 1e-14 * max_j |K_ref[i, j]|
i.e. 1e-12 relative for every entry above 1% of its row's largest entry, and
below that an absolute floor of ~45 ulp of the row's largest entry.  Small
entries are differences of large terms (a beam's rotation-rotation entry is
EI/L - kGA L/4 before the frame transform; exactly-zero sums of plate
blocks), so their rounding error scales with the row, not with the entry:
the worst seen is 18 ulp of the row max on a 2.5e-4 * row-max beam entry.
The normwise max |dK| / max |K| is reported beside it.  The sparsity pattern
(row_ptr, col_idx) must be bit-identical.
"""

import numpy as np

REL = 1e-12
ROW_FLOOR = 1e-14


def row_max(row_ptr, val):
    n = len(row_ptr) - 1
    counts = np.diff(row_ptr)
    rows = np.repeat(np.arange(n), counts)
    rm = np.zeros(n)
    np.maximum.at(rm, rows, np.abs(val))
    return rm[rows]


def compare(gpu, ref, rel=REL, floor=ROW_FLOOR):
    rp, ci, v = gpu
    rp0, ci0, v0 = ref
    out = {"pattern_equal": bool(np.array_equal(rp, rp0) and np.array_equal(ci, ci0)), "nnz": int(len(v0))}
    if not out["pattern_equal"] or len(v0) == 0:
        out.update(max_ratio=np.inf if len(v0) else 0.0, max_rel=np.nan, normwise=np.nan)
        return out
    rm = row_max(rp0, v0)
    d = np.abs(v - v0)
    bound = rel * np.abs(v0) + floor * rm
    out["max_ratio"] = float(np.max(d / np.maximum(bound, 1e-300)))
    sig = np.abs(v0) > 1e-9 * rm
    out["max_rel"] = float(np.max(d[sig] / np.abs(v0[sig]))) if sig.any() else 0.0
    out["normwise"] = float(d.max() / np.abs(v0).max())
    return out
