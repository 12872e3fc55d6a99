"""Parity metrics between the GPU CSR and the oracle's CSR.

Tolerance (values are FP64): for every stored entry
    |K_gpu - K_ref| <= 1e-12 * |K_ref| + 2 n u M
with u = 2^-53 the unit roundoff, n = 32 and M the entry's term magnitude
from the oracle (OracleModel.csr_magnitude): the sum over the entry's element
contributions of the stiffness terms taken with the absolute value of every
factor, rotated as |T|^T |k| |T|.

Why this bound: the GPU and the oracle sum the same terms (products of
element constants, shape-function derivatives and frame entries) in different
orders and groupings.  For a sum of products evaluated in floating point,
|fl(s) - s| <= gamma_n * sum |terms|, gamma_n = n u / (1 - n u), where n
counts the roundings along one term's chain, the inputs' own (Higham,
Accuracy and Stability of Numerical Algorithms, 2nd ed., sec. 3.1).  Two such
evaluations differ by at most 2 gamma_n M.  The longest chains are a plate's
(local coordinates, Jacobian, its inverse, dh/dx, the MITC4 multipliers,
constitutive constants, the weight, up to four points and four elements):
fewer than 32 roundings.  So entries whose terms cancel (a beam's
rotation-rotation entry EI/L - kGA L/4 before the frame transform, sums of
plate blocks that vanish) get an allowance that scales with their own terms,
not with their row, and well-conditioned entries are held to 1e-12 relative
(2 n u M is then ~7e-15 relative).

Measured (all GPU parity cases): the worst entry uses 0.31 of its bound.
Cancelling entries stay within ~20 u M.  Well-conditioned beam entries
differ by up to ~2.8e3 u M (6e-13 relative, mixed bar model; 490 u M on the
64^3 frame): the beam frames come from cos, sin and acos, whose device and
host implementations may differ in the last bit, and the roll's acos near
+-1 magnifies that; the 1e-12 relative part of the bound holds them.

The sparsity pattern (row_ptr, col_idx) must be bit-identical.  Reported per
comparison: max_ratio (largest |dK| / bound, <= 1 passes), max_rel (largest
relative difference over entries above 1e-9 of their row maximum), u_of_M
(largest |dK| / (u M): the roundings actually used) and normwise
(max |dK| / max |K|).
"""

import numpy as np

REL = 1e-12
U = 2.0 ** -53
ROUNDINGS = 32  # n: roundings along one term's chain, per evaluation


def row_max(row_ptr, val):
    n = len(row_ptr) - 1
    counts = np.diff(row_ptr)
    rows = np.repeat(np.arange(n), counts)
    rm = np.zeros(n)
    np.maximum.at(rm, rows, np.abs(val))
    return rm[rows]


def oracle_reference(om):
    """(row_ptr, col_idx, val, magnitude) of the oracle's structural CSR."""
    rp, ci, v = om.csr(structural=True)
    return rp, ci, v, om.csr_magnitude()


def entry_bound(v0, mag, rel=REL):
    return rel * np.abs(v0) + 2.0 * ROUNDINGS * U * mag


def compare(gpu, ref, rel=REL):
    """gpu: (row_ptr, col_idx, val); ref: oracle_reference(...)."""
    rp, ci, v = gpu
    rp0, ci0, v0, mag = ref
    out = {"pattern_equal": bool(np.array_equal(rp, rp0) and np.array_equal(ci, ci0)), "nnz": int(len(v0))}
    if not out["pattern_equal"] or len(v0) == 0:
        out.update(max_ratio=np.inf if len(v0) else 0.0, max_rel=np.nan, normwise=np.nan)
        return out
    d = np.abs(v - v0)
    bound = entry_bound(v0, mag, rel)
    out["max_ratio"] = float(np.max(d / np.maximum(bound, 1e-300)))
    rm = row_max(rp0, v0)
    sig = np.abs(v0) > 1e-9 * rm
    out["max_rel"] = float(np.max(d[sig] / np.abs(v0[sig]))) if sig.any() else 0.0
    out["u_of_M"] = float(np.max(d / np.maximum(U * mag, 1e-300)))
    out["normwise"] = float(d.max() / np.abs(v0).max())
    return out
