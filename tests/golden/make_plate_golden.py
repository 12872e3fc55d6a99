"""Writes plate_mitc4_golden.json: the local stiffness of single MITC4 plates,
evaluated in 50-digit arithmetic (mpmath), as fixtures for the FP64 paths.

A restatement of the crate's plate element written for this purpose: it shares
no code with the FP64 oracle or the product.  It follows:
- shape functions and Jacobian (src/fem/quadrilateral_4n_element_functions.rs:
  252-338, 380-503, 505-653): node order (r, s) = 1 (+, +), 2 (-, +), 3 (-, -),
  4 (+, -), h_i = (1 + r r_i)(1 + s s_i) / 4, J = [[x_r, y_r], [x_s, y_s]],
  [dh/dx; dh/dy] = J^-1 [dh/dr; dh/ds];
- B matrices (src/fem/structs/plate.rs:160-274 membrane, 276-390 bending,
  392-511 MITC4 shear, with gamma multipliers sqrt(...) / (8 det J));
- constitutive matrices and the sum over points (plate.rs:513-672: membrane
  E t / (1 - nu^2), bending E t^3 / (12 (1 - nu^2)), both times
  [[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]]; shear E t k / (2 (1 + nu))
  times the weight again, then every part times det J and the weight);
- the 2 x 2 rule with the crate's point coordinate: sqrt of the f32 value of
  1/3 taken in f64 (plate.rs:1066-1091), weights 1; KROT6 = 1 added to each
  node's drilling diagonal (plate.rs:25, 714-721).

The quads lie in z = 0 with node 3 at the origin and (X4 - X3) x (X2 - X3)
along +z, so the element frame is the identity and the local matrix is also the
element's global one.  Inputs are dyadic (exact in binary); the FP64 inputs E,
nu, t, k are taken at their exact binary values.  Each nonzero entry is stored
to 30 significant digits.

Gradients, for the backward: with G random on the 212 nonzeros and
L = sum G_ij K_ij, and with L_u = -u^T K u for a random u (the compliance
sensitivities), dL/dp for p = E, nu, t, k and for the in-plane coordinates
x, y of the four nodes (the frame stays the identity under in-plane moves),
by central differences at 50 digits with steps 1e-20 relative (truncation
~1e-40, cancellation ~1e-30): 30 significant digits.

plate_tilted_golden.json holds the same quads rotated and translated into
general position (FP64 coordinates, taken exactly): the global matrix (all 576
entries) and the gradients with respect to E, nu, t, k and all twelve node
coordinates.

Run from the repository root: python tests/golden/make_plate_golden.py
"""

import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50

CASES = [
    # (x1, y1, x2, y2, x4, y4), E, nu, t, k  -- x3 = y3 = 0
    ((1.0, 1.0, 0.0, 1.0, 1.0, 0.0), 2.0e11, 0.3, 0.01, 5.0 / 6.0),  # unit square
]


def _random_cases(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        sx, sy = rng.uniform(0.5, 3.0, 2)
        j = rng.uniform(-0.25, 0.25, 6)
        xl = np.array([sx * (1 + j[0]), sy * (1 + j[1]), sx * j[2], sy * (1 + j[3]), sx * (1 + j[4]), sy * j[5]])
        xl = np.round(xl * 1024.0) / 1024.0  # dyadic
        x1, y1, x2, y2, x4, y4 = xl
        # convex and counterclockwise 1 -> 2 -> 3 -> 4 (positive cross products at every corner)
        p = [(x1, y1), (x2, y2), (0.0, 0.0), (x4, y4)]
        cross = [(p[(i + 1) % 4][0] - p[i][0]) * (p[(i + 2) % 4][1] - p[(i + 1) % 4][1])
                 - (p[(i + 1) % 4][1] - p[i][1]) * (p[(i + 2) % 4][0] - p[(i + 1) % 4][0]) for i in range(4)]
        if min(cross) <= 0.05 * sx * sy:
            continue
        E = float(rng.choice([70e9, 2.0e11, 2.1e11]))
        nu = float(rng.choice([0.25, 0.3, 0.33]))
        t = float(np.round(rng.uniform(0.002, 0.2), 4))
        out.append((tuple(float(v) for v in xl), E, nu, t, 5.0 / 6.0))
    return out


def gauss_point():
    """The crate's 2 x 2 point coordinate: f32(1/3) widened, then sqrt in f64."""
    return mp.mpf(float(np.sqrt(np.float64(np.float32(1.0 / 3.0)))))


def plate_k(xl, E, nu, t, kf, x3=0, y3=0):
    """Local stiffness; xl = (x1, y1, x2, y2, x4, y4) in the plane, node 3 at
    (x3, y3) (local coordinates are taken relative to node 3)."""
    x3, y3 = mp.mpf(x3), mp.mpf(y3)
    x1, y1, x2, y2, x4, y4 = (mp.mpf(v) - (x3 if i % 2 == 0 else y3) for i, v in enumerate(xl))
    x3 = y3 = mp.mpf(0)
    X, Y = [x1, x2, x3, x4], [y1, y2, y3, y4]
    E, nu, t, kf = mp.mpf(E), mp.mpf(nu), mp.mpf(t), mp.mpf(kf)
    g = gauss_point()
    one, two, four = mp.mpf(1), mp.mpf(2), mp.mpf(4)
    cmem = E * t / (one - nu ** 2)
    cbend = E * t ** 3 / (12 * (one - nu ** 2))
    cshear = E * t * kf / (two * (one + nu))
    C = mp.matrix([[1, nu, 0], [nu, 1, 0], [0, 0, (one - nu) / two]])
    ax, bx, cx = x1 - x2 - x3 + x4, x1 - x2 + x3 - x4, x1 + x2 - x3 - x4
    ay, by, cy = y1 - y2 - y3 + y4, y1 - y2 + y3 - y4, y1 + y2 - y3 - y4
    K = mp.matrix(24, 24)
    for r, s in ((g, g), (-g, g), (-g, -g), (g, -g)):
        alpha = one  # alpha_r * alpha_s
        dr = [(one + s) / four, -(one + s) / four, -(one - s) / four, (one - s) / four]
        ds = [(one + r) / four, (one - r) / four, -(one - r) / four, -(one + r) / four]
        xr = sum(dr[i] * X[i] for i in range(4))
        yr = sum(dr[i] * Y[i] for i in range(4))
        xs = sum(ds[i] * X[i] for i in range(4))
        ys = sum(ds[i] * Y[i] for i in range(4))
        det = xr * ys - yr * xs
        dx = [(ys * dr[i] - yr * ds[i]) / det for i in range(4)]
        dy = [(-xs * dr[i] + xr * ds[i]) / det for i in range(4)]
        bm = mp.matrix(3, 24)
        bb = mp.matrix(3, 24)
        for i in range(4):
            bm[0, 6 * i] = dx[i]
            bm[1, 6 * i + 1] = dy[i]
            bm[2, 6 * i] = dy[i]
            bm[2, 6 * i + 1] = dx[i]
            bb[0, 6 * i + 4] = -dx[i]
            bb[1, 6 * i + 3] = dy[i]
            bb[2, 6 * i + 3] = dx[i]
            bb[2, 6 * i + 4] = -dy[i]
        mr = mp.sqrt((cx + r * bx) ** 2 + (cy + r * by) ** 2) / (8 * det)
        ms = mp.sqrt((ax + s * bx) ** 2 + (ay + s * by) ** 2) / (8 * det)
        bs = mp.matrix(2, 24)
        # gamma_rz: edge 1-2 carries (1 + s), edge 4-3 carries (1 - s)
        for i, sign, f, ex, ey in ((0, 1, one + s, x1 - x2, y1 - y2), (1, -1, one + s, x1 - x2, y1 - y2),
                                   (2, -1, one - s, x4 - x3, y4 - y3), (3, 1, one - s, x4 - x3, y4 - y3)):
            bs[0, 6 * i + 2] = sign * f / two * mr
            bs[0, 6 * i + 3] = -f * ey / four * mr
            bs[0, 6 * i + 4] = f * ex / four * mr
        # gamma_sz: edge 1-4 carries (1 + r), edge 2-3 carries (1 - r)
        for i, sign, f, ex, ey in ((0, 1, one + r, x1 - x4, y1 - y4), (1, 1, one - r, x2 - x3, y2 - y3),
                                   (2, -1, one - r, x2 - x3, y2 - y3), (3, -1, one + r, x1 - x4, y1 - y4)):
            bs[1, 6 * i + 2] = sign * f / two * ms
            bs[1, 6 * i + 3] = -f * ey / four * ms
            bs[1, 6 * i + 4] = f * ex / four * ms
        K += (bm.T * (C * cmem) * bm) * (det * alpha)
        K += (bb.T * (C * cbend) * bb) * (det * alpha)
        K += (bs.T * (cshear * alpha) * bs) * (det * alpha)
    for i in range(4):
        K[6 * i + 5, 6 * i + 5] += 1
    return K


def gradients(xl, props, weights):
    """dL/d(E, nu, t, k) and dL/d(x_a, y_a) for a = 1..4 of L = sum w_ij K_ij,
    by central differences at 50 digits."""
    def L(xl_, props_, x3=0, y3=0):
        K = plate_k(xl_, *props_, x3=x3, y3=y3)
        return sum(w * K[i, j] for (i, j), w in weights.items())
    dp = []
    for q in range(4):
        h = mp.mpf(props[q]) * mp.mpf("1e-20")
        up = [mp.mpf(v) for v in props]
        dn = [mp.mpf(v) for v in props]
        up[q] += h
        dn[q] -= h
        dp.append((L(xl, up) - L(xl, dn)) / (2 * h))
    scale = max(abs(mp.mpf(v)) for v in xl)
    h = scale * mp.mpf("1e-20")
    dx = []
    for a in range(4):  # node a + 1; node 3 (a = 2) moves through x3, y3
        row = []
        for c in range(2):
            if a == 2:
                args = [(xl, props, h if c == 0 else 0, h if c == 1 else 0),
                        (xl, props, -h if c == 0 else 0, -h if c == 1 else 0)]
            else:
                k = {0: 0, 1: 2, 3: 4}[a] + c
                xu = [mp.mpf(v) for v in xl]
                xd = [mp.mpf(v) for v in xl]
                xu[k] += h
                xd[k] -= h
                args = [(xu, props, 0, 0), (xd, props, 0, 0)]
            row.append((L(*args[0]) - L(*args[1])) / (2 * h))
        dx.append(row)
    return dp, dx


def main():
    cases = CASES + _random_cases(20251203, 3)
    rng = np.random.default_rng(7)
    out = {"description": "MITC4 plate local stiffness, 50-digit evaluation (tests/golden/make_plate_golden.py)",
           "cases": []}
    for xl, E, nu, t, kf in cases:
        K = plate_k(xl, E, nu, t, kf)
        entries = [[i, j, mp.nstr(K[i, j], 30)] for i in range(24) for j in range(24) if K[i, j] != 0]
        x1, y1, x2, y2, x4, y4 = xl
        G = [float(v) for v in rng.normal(size=len(entries))]
        u = [float(v) for v in rng.normal(size=24)]
        props = (E, nu, t, kf)
        gp, gx = gradients(xl, props, {(i, j): mp.mpf(g) for (i, j, _), g in zip(entries, G)})
        up, ux = gradients(xl, props, {(i, j): -mp.mpf(u[i]) * mp.mpf(u[j]) for i, j, _ in entries})
        s30 = lambda v: mp.nstr(v, 30)
        out["cases"].append({"xyz": [[x1, y1, 0.0], [x2, y2, 0.0], [0.0, 0.0, 0.0], [x4, y4, 0.0]],
                             "props": [E, nu, t, kf], "entries": entries,
                             "G": G, "grad_G": {"params": [s30(v) for v in gp], "xy": [[s30(v) for v in r] for r in gx]},
                             "u": u, "grad_u": {"params": [s30(v) for v in up], "xy": [[s30(v) for v in r] for r in ux]}})
        print("case", xl, E, nu, t, "nonzeros", len(entries))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "plate_mitc4_golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", path)



# ----------------------------------------------------------------- tilted plates
# Quads in general position: the same element in global axes,
# K = T^T k(local coordinates) T with T = diag(R, R, ...), R's rows (e1, e2, e3),
# e3 along n = (X4 - X3) x (X2 - X3) and local coordinates the first two
# components of R (X_i - X3) (quadrilateral_4n_element_functions.rs:131-171,
# 340-378; plate.rs:726-1002).  The crate fixes e1 by its alignment rotation;
# K is the same for every e1 in the plane (the element is isotropic in its
# plane: an in-plane rotation of the frame rotates k and T together), so here
# e1 = (X4 - X3) projected on the plane, normalised.


def frame(X):
    X = [[mp.mpf(v) for v in p] for p in X]
    a = [X[3][k] - X[2][k] for k in range(3)]
    b = [X[1][k] - X[2][k] for k in range(3)]
    n = [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]
    nn = mp.sqrt(sum(v * v for v in n))
    e3 = [v / nn for v in n]
    d = sum(a[k] * e3[k] for k in range(3))
    p = [a[k] - d * e3[k] for k in range(3)]
    pn = mp.sqrt(sum(v * v for v in p))
    e1 = [v / pn for v in p]
    e2 = [e3[1] * e1[2] - e3[2] * e1[1], e3[2] * e1[0] - e3[0] * e1[2], e3[0] * e1[1] - e3[1] * e1[0]]
    return X, [e1, e2, e3]


def plate_k_global(X, props):
    X, R = frame(X)
    loc = []
    for i in (0, 1, 3):
        d = [X[i][k] - X[2][k] for k in range(3)]
        loc += [sum(R[0][k] * d[k] for k in range(3)), sum(R[1][k] * d[k] for k in range(3))]
    k = plate_k(loc, *props)
    T = mp.matrix(24, 24)
    for blk in range(8):
        for r in range(3):
            for c in range(3):
                T[3 * blk + r, 3 * blk + c] = R[r][c]
    return T.T * k * T


def tilted_gradients(X, props, weight_sets):
    """dL/d(E, nu, t, k) and dL/dX (4 x 3) for each weight set, by central
    differences at 50 digits (steps 1e-20 relative)."""
    def Ls(X_, p_):
        K = plate_k_global(X_, p_)
        return [sum(w * K[i, j] for (i, j), w in ws.items()) for ws in weight_sets]
    nset = len(weight_sets)
    gp = [[None] * 4 for _ in range(nset)]
    for q in range(4):
        h = mp.mpf(props[q]) * mp.mpf("1e-20")
        up = [mp.mpf(v) for v in props]
        dn = list(up)
        up[q] += h
        dn[q] -= h
        lu, ld = Ls(X, up), Ls(X, dn)
        for s in range(nset):
            gp[s][q] = (lu[s] - ld[s]) / (2 * h)
    scale = max(abs(mp.mpf(v)) for p in X for v in p)
    h = scale * mp.mpf("1e-20")
    gx = [[[None] * 3 for _ in range(4)] for _ in range(nset)]
    for a in range(4):
        for c in range(3):
            Xu = [[mp.mpf(v) for v in p] for p in X]
            Xd = [list(p) for p in Xu]
            Xu[a][c] += h
            Xd[a][c] -= h
            lu, ld = Ls(Xu, props), Ls(Xd, props)
            for s in range(nset):
                gx[s][a][c] = (lu[s] - ld[s]) / (2 * h)
    return gp, gx


def main_tilted():
    rng = np.random.default_rng(29)
    out = {"description": "MITC4 plates in general position: global stiffness and gradients, 50-digit "
                          "evaluation (tests/golden/make_plate_golden.py)", "cases": []}
    s30 = lambda v: mp.nstr(v, 30)
    for xl, E, nu, t, kf in CASES + _random_cases(20251203, 3):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        Q = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        flat = np.array([[xl[0], xl[1], 0.0], [xl[2], xl[3], 0.0], [0.0, 0.0, 0.0], [xl[4], xl[5], 0.0]])
        X = (flat @ Q.T + rng.uniform(-2, 2, 3)).tolist()  # the FP64 coordinates are the inputs
        props = (E, nu, t, kf)
        K = plate_k_global(X, props)
        entries = [[i, j, s30(K[i, j])] for i in range(24) for j in range(24) if K[i, j] != 0]
        G = [float(v) for v in rng.normal(size=len(entries))]
        u = [float(v) for v in rng.normal(size=24)]
        wG = {(i, j): mp.mpf(g) for (i, j, _), g in zip(entries, G)}
        wu = {(i, j): -mp.mpf(u[i]) * mp.mpf(u[j]) for i, j, _ in entries}
        gp, gx = tilted_gradients(X, props, [wG, wu])
        pack = lambda s: {"params": [s30(v) for v in gp[s]], "xyz": [[s30(v) for v in r] for r in gx[s]]}
        out["cases"].append({"xyz": X, "props": list(props), "entries": entries,
                             "G": G, "grad_G": pack(0), "u": u, "grad_u": pack(1)})
        print("tilted case", E, nu, t, "nonzeros", len(entries))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "plate_tilted_golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", path)


if __name__ == "__main__":
    main()
    main_tilted()
