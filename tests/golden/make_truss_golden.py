"""Writes truss_golden.json: the global stiffness of single 3-D bars and the
gradients of two scalars of it, in 50-digit arithmetic (mpmath), as fixtures
for the FP64 paths.

A restatement of the crate's truss written for this purpose (it shares no code
with the FP64 oracle or the product), from src/fem/structs/truss.rs:95-189 and
bar_2n_element_functions.rs:8-113: one point r = 0 with weight 2, J = L / 2,
B = [dh1/dr, 0, 0, dh2/dr, 0, 0] / J with dh/dr = -+ 1/2, the area at r
(A2 - A) / 2 r + A + (A2 - A) / 2 (A without A2), k = B^T B A(r) E det J alpha
= E A(0) / L on the axial dofs.  In global axes only the frame's first row
d / L enters (truss.rs:66-93, 191-214): K = E A(0) / L [e e^T, -e e^T;
-e e^T, e e^T] with e = d / L, whatever the frame's other rows.

Gradients: with G random on the 36 nonzeros, L = sum G_ij K_ij, and
L_u = -u^T K u for a random u, dL/d(E, A, A2) and dL/dX of both nodes by
central differences at 50 digits, steps 1e-20 relative: 30 significant digits.

Run from the repository root: python tests/golden/make_truss_golden.py
"""

import json
import math
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50

# node 1, node 2, (E, A, A2 or None)
CASES = [
    ((0.0, 0.0, 0.0), (1.5, 0.75, -0.5), (2.0e11, 0.01, 0.004)),
    ((0.25, -1.0, 2.0), (-0.875, 0.5, 1.125), (7.0e10, 0.0025, None)),
    ((3.0, 1.0, -2.0), (1.0, 2.5, 0.25), (2.1e11, 0.02, 0.03)),
]


def truss_k(x1, x2, E, A, A2):
    d = [mp.mpf(b) - mp.mpf(a) for a, b in zip(x1, x2)]
    L = mp.sqrt(sum(v * v for v in d))
    e = [v / L for v in d]
    A, E = mp.mpf(A), mp.mpf(E)
    r = mp.mpf(0)  # the point
    area = A if A2 is None else (mp.mpf(A2) - A) / 2 * r + A + (mp.mpf(A2) - A) / 2
    c = E * area / L
    K = mp.matrix(12, 12)
    for a in range(2):
        for b in range(2):
            sgn = 1 if a == b else -1
            for i in range(3):
                for j in range(3):
                    K[6 * a + i, 6 * b + j] = sgn * c * e[i] * e[j]
    return K


def gradients(x1, x2, props, weights):
    def Lw(x1_, x2_, p):
        K = truss_k(x1_, x2_, *p)
        return sum(w * K[i, j] for (i, j), w in weights.items())
    gp = []
    for q in range(3):
        if props[q] is None:
            gp.append(None)
            continue
        h = mp.mpf(props[q]) * mp.mpf("1e-20")
        up = [mp.mpf(v) if v is not None else None for v in props]
        dn = list(up)
        up[q] += h
        dn[q] -= h
        gp.append((Lw(x1, x2, up) - Lw(x1, x2, dn)) / (2 * h))
    scale = max(abs(mp.mpf(v)) for v in list(x1) + list(x2))
    h = scale * mp.mpf("1e-20")
    p = [mp.mpf(v) if v is not None else None for v in props]
    gx = []
    for node in range(2):
        row = []
        for c in range(3):
            xs = [[mp.mpf(v) for v in x1], [mp.mpf(v) for v in x2]]
            xs2 = [list(xs[0]), list(xs[1])]
            xs[node][c] += h
            xs2[node][c] -= h
            row.append((Lw(xs[0], xs[1], p) - Lw(xs2[0], xs2[1], p)) / (2 * h))
        gx.append(row)
    return gp, gx


def main():
    rng = np.random.default_rng(13)
    s30 = lambda v: None if v is None else mp.nstr(v, 30)
    out = {"description": "3-D bar stiffness and gradients, 50-digit evaluation (tests/golden/make_truss_golden.py)",
           "cases": []}
    for x1, x2, props in CASES:
        K = truss_k(x1, x2, *props)
        entries = [[i, j, s30(K[i, j])] for i in range(12) for j in range(12) if K[i, j] != 0]
        G = [float(v) for v in rng.normal(size=len(entries))]
        u = [float(v) for v in rng.normal(size=12)]
        gG = gradients(x1, x2, props, {(i, j): mp.mpf(g) for (i, j, _), g in zip(entries, G)})
        gu = gradients(x1, x2, props, {(i, j): -mp.mpf(u[i]) * mp.mpf(u[j]) for i, j, _ in entries})
        pack = lambda g: {"params": [s30(v) for v in g[0]], "xyz": [[s30(v) for v in r] for r in g[1]]}
        out["cases"].append({"xyz": [list(x1), list(x2)],
                             "props": [props[0], props[1], math.nan if props[2] is None else props[2]],
                             "entries": entries, "G": G, "grad_G": pack(gG), "u": u, "grad_u": pack(gu)})
        print("case", x1, x2, props, "nonzeros", len(entries))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "truss_golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", path)


if __name__ == "__main__":
    main()
