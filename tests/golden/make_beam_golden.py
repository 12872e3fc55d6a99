"""Writes beam_golden.json: the stiffness of single x-aligned beams and the
gradients of two scalars of it, in 50-digit arithmetic (mpmath), as fixtures
for the FP64 paths.

A restatement of the crate's beam element written for this purpose (it shares
no code with the FP64 oracle or the product), from src/fem/structs/beam.rs:
- one point r = 0 with weight 2; J = det J = L / 2 (bar_2n_element_functions.rs:
  8-113: x1 = -L/2, x2 = L/2), h = (1 -+ r) / 2, dh/dr = -+ 1/2;
- six rank-one terms (beam.rs:260-605): u with E A, v with G A k
  (B_v = dh/J at v minus h at theta_w), w with G A k (B_w = dh/J at w minus h
  at theta_v: the crate's sign), theta_u with G It, theta_v with E I22p,
  theta_w with E I11p, G = E / (2 (1 + nu)); each term c B^T B det J alpha.
The beams run along +x from the origin with local axis 1 = +z, I12 = 0 and
I11 > I22: the principal angle is 0 and the frame is the identity
(beam.rs:119-258), so the local matrix is also the global one.

Gradients: with G random on the nonzeros, L = sum G_ij K_ij, and
L_u = -u^T K u for a random u (the compliance sensitivities), dL/dp for
p = E, nu, A, I11, I22, It, k by central differences at 50 digits, steps 1e-20
relative: 30 significant digits.

Run from the repository root: python tests/golden/make_beam_golden.py
"""

import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50

# L, (E, nu, A, I11, I22, I12, It, k)
CASES = [
    (2.0, (2.0e11, 0.3, 0.01, 8.0e-5, 2.0e-5, 0.0, 1.0e-5, 5.0 / 6.0)),
    (0.625, (7.0e10, 0.33, 0.0025, 3.0e-6, 1.25e-6, 0.0, 2.5e-6, 0.85)),
    (5.5, (2.1e11, 0.25, 0.03, 6.0e-4, 1.0e-4, 0.0, 2.0e-4, 0.9)),
]
PARAMS = (0, 1, 2, 3, 4, 6, 7)  # E, nu, A, I11, I22, It, k in the 11-value layout


def beam_k(L, props):
    L = mp.mpf(L)
    E, nu, A, i11, i22, _, it, kf = (mp.mpf(v) for v in props)
    J = L / 2
    alpha = mp.mpf(2)
    h = mp.mpf(1) / 2
    d1, d2 = -1 / (2 * J), 1 / (2 * J)  # dh/dr / J at r = 0
    G = E / (2 * (1 + nu))

    def row(entries):
        b = [mp.mpf(0)] * 12
        for i, v in entries:
            b[i] += v
        return b

    terms = [
        (A * E, row([(0, d1), (6, d2)])),
        (G * A * kf, row([(1, d1), (7, d2), (5, -h), (11, -h)])),
        (G * A * kf, row([(2, d1), (8, d2), (4, -h), (10, -h)])),
        (G * it, row([(3, d1), (9, d2)])),
        (E * i22, row([(4, d1), (10, d2)])),
        (E * i11, row([(5, d1), (11, d2)])),
    ]
    K = mp.matrix(12, 12)
    for c, b in terms:
        for i in range(12):
            for j in range(12):
                K[i, j] += c * b[i] * b[j] * J * alpha
    return K


def gradients(L, props, weights):
    def Lw(p):
        K = beam_k(L, p)
        return sum(w * K[i, j] for (i, j), w in weights.items())
    out = []
    for q in PARAMS:
        hq = mp.mpf(props[q]) * mp.mpf("1e-20")
        up = [mp.mpf(v) for v in props]
        dn = [mp.mpf(v) for v in props]
        up[q] += hq
        dn[q] -= hq
        out.append((Lw(up) - Lw(dn)) / (2 * hq))
    return out


def main():
    rng = np.random.default_rng(11)
    out = {"description": "x-aligned beam stiffness and gradients, 50-digit evaluation "
                          "(tests/golden/make_beam_golden.py)", "params": list(PARAMS), "cases": []}
    s30 = lambda v: mp.nstr(v, 30)
    for L, props in CASES:
        K = beam_k(L, props)
        entries = [[i, j, s30(K[i, j])] for i in range(12) for j in range(12) if K[i, j] != 0]
        G = [float(v) for v in rng.normal(size=len(entries))]
        u = [float(v) for v in rng.normal(size=12)]
        gG = gradients(L, props, {(i, j): mp.mpf(g) for (i, j, _), g in zip(entries, G)})
        gu = gradients(L, props, {(i, j): -mp.mpf(u[i]) * mp.mpf(u[j]) for i, j, _ in entries})
        out["cases"].append({"xyz": [[0.0, 0.0, 0.0], [L, 0.0, 0.0]],
                             "props": list(props) + [0.0, 0.0, 1.0], "entries": entries,
                             "G": G, "grad_G": [s30(v) for v in gG], "u": u, "grad_u": [s30(v) for v in gu]})
        print("case L", L, "nonzeros", len(entries))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "beam_golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", path)


if __name__ == "__main__":
    main()
