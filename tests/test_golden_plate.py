"""The FP64 oracle's plate stiffness against the 50-digit MITC4 fixtures
(tests/golden/plate_mitc4_golden.json, written by make_plate_golden.py from
src/fem/structs/plate.rs:160-724 and quadrilateral_4n_element_functions.rs).

Each case is one flat quad with the identity frame, so the oracle's CSR of the
one-element model is the local matrix.  The fixtures are checked first for
what the mathematics fixes regardless of the restatement (symmetry, the rigid
translations and rotations, nullity 6 with the drilling term, positive
semi-definiteness, the gradients' homogeneity in E and translation
invariance); then the oracle must meet the parity bound of tests/parity.py
entry by entry with an identical sparsity pattern.  The same quads in general
position (plate_tilted_golden.json) must equal the flat fixtures rotated into
place, and the oracle must meet them too.  The device path is held to both in
test_gpu_golden_plate.py."""

import json
import os

import numpy as np
import pytest

from oracle.oracle import OracleModel
from tests.parity import compare

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "plate_mitc4_golden.json")


def cases():
    with open(FIXTURE) as f:
        return json.load(f)["cases"]


def golden_dense(case):
    K = np.zeros((24, 24))
    for i, j, v in case["entries"]:
        K[i, j] = float(v)
    return K


def one_plate_oracle(case):
    om = OracleModel(1e-4, 1e-12, 4)
    for n, p in enumerate(case["xyz"]):
        om.add_node(n + 1, *p)
    om.add_plate(1, 1, 2, 3, 4, *case["props"])
    return om


def golden_reference(case, om):
    """(row_ptr, col_idx, val, magnitude) of the fixture on the oracle's
    structural pattern, after checking that the two patterns agree."""
    rp, ci, _ = om.csr(structural=True)
    rows = np.repeat(np.arange(24), np.diff(rp))
    K = golden_dense(case)
    assert set(zip(rows.tolist(), ci.tolist())) == {(i, j) for i, j, _ in case["entries"]}
    return rp, ci, K[rows, ci], om.csr_magnitude()


@pytest.mark.parametrize("idx", range(4))
def test_fixture_invariants(idx):
    case = cases()[idx]
    K = golden_dense(case)
    scale = np.abs(K).max()
    assert np.abs(K - K.T).max() <= 1e-15 * scale
    xyz = np.array(case["xyz"])
    for d in range(3):  # rigid translations
        u = np.zeros(24)
        u[d::6] = 1.0
        assert np.abs(K @ u).max() <= 1e-12 * scale
    # rigid rotations (right-hand rule, small angle phi = 1): about z, u = -y,
    # v = x with the drilling rotation left at 0 (KROT6 resists it); about x,
    # w = y with theta_x = 1; about y, w = -x with theta_y = 1
    rz = np.zeros(24)
    rz[0::6], rz[1::6] = -xyz[:, 1], xyz[:, 0]
    assert np.abs(K @ rz).max() <= 1e-12 * scale
    rx = np.zeros(24)
    rx[2::6], rx[3::6] = xyz[:, 1], 1.0
    ry = np.zeros(24)
    ry[2::6], ry[4::6] = -xyz[:, 0], 1.0
    assert np.abs(K @ rx).max() <= 1e-12 * scale and np.abs(K @ ry).max() <= 1e-12 * scale
    # the drilling dofs are decoupled with stiffness KROT6 = 1: four eigenvalues
    # exactly 1; of the other 20, six vanish (to rounding: ~1e-16 of the
    # largest, 1e-5 here) and the rest are positive
    drill = [6 * a + 5 for a in range(4)]
    assert np.array_equal(K[drill][:, drill], np.eye(4))
    keep = [i for i in range(24) if i not in drill]
    assert not K[drill][:, keep].any()
    w = np.linalg.eigvalsh(K[np.ix_(keep, keep)])
    tiny = 1e-13 * np.abs(w).max()
    assert np.sum(np.abs(w) < tiny) == 6
    assert np.all(w > -tiny)
    assert len(case["entries"]) == 212


@pytest.mark.parametrize("idx", range(4))
def test_oracle_matches_golden(idx):
    case = cases()[idx]
    om = one_plate_oracle(case)
    ref = golden_reference(case, om)
    r = compare(om.csr(structural=True), ref)
    print("case", idx, r)
    assert r["pattern_equal"] and r["max_ratio"] <= 1.0, r
    assert r["normwise"] < 1e-14, r


def test_fixture_keeps_the_f32_point():
    """The fixtures use the crate's f32 Gauss point; with the exact 1/sqrt(3)
    the unit-square matrix moves by ~1e-8 relative, far above the parity
    bound, so a restatement with exact points would fail the tests above."""
    case = cases()[0]
    g_crate = float(np.sqrt(np.float64(np.float32(1.0 / 3.0))))
    rel = abs(g_crate - 1.0 / np.sqrt(3.0)) * np.sqrt(3.0)
    assert 1e-9 < rel < 1e-7
    K = golden_dense(case)
    om = one_plate_oracle(case)
    rp, ci, v = om.csr(structural=True)
    rows = np.repeat(np.arange(24), np.diff(rp))
    assert np.abs(v - K[rows, ci]).max() < 1e-3 * rel * np.abs(K).max()


def weights(case, key):
    """w_ij on the entries: G, or -u_i u_j for the compliance."""
    if key == "G":
        return np.array(case["G"])
    u = np.array(case["u"])
    return np.array([-u[i] * u[j] for i, j, _ in case["entries"]])


def grad_scale(case, key):
    """(sum |w_ij K_ij|, L = sum w_ij K_ij): the first is the natural size of
    L, hence of p dL/dp."""
    K = golden_dense(case)
    w = weights(case, key)
    vals = np.array([K[i, j] for i, j, _ in case["entries"]])
    return float(np.abs(w * vals).sum()), float(np.dot(w, vals))


@pytest.mark.parametrize("key", ["G", "u"])
@pytest.mark.parametrize("idx", range(4))
def test_fixture_gradients_self_consistent(idx, key):
    """K is E times a matrix free of E plus KROT6 = 1 on the drilling
    diagonals, so E dL/dE = L - sum of w over those four entries; K depends on
    positions only through differences, so the in-plane coordinate gradients
    sum to zero."""
    case = cases()[idx]
    g = case["grad_" + key]
    lab, L = grad_scale(case, key)
    w = weights(case, key)
    drill = sum(wk for wk, (i, j, _) in zip(w, case["entries"]) if i == j and i % 6 == 5)
    E = case["props"][0]
    assert abs(E * float(g["params"][0]) - (L - drill)) <= 1e-14 * lab
    xy = np.array([[float(v) for v in r] for r in g["xy"]])
    size = max(abs(v) for p in case["xyz"] for v in p)
    assert np.abs(xy.sum(axis=0)).max() <= 1e-14 * lab / size


# ----------------------------------------------------------------- tilted plates
TILTED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "plate_tilted_golden.json")


def tilted_cases():
    with open(TILTED) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("idx", range(4))
def test_tilted_fixture_is_the_rotated_flat_one(idx):
    """The tilted fixture (own frame, projection, T^T k T) equals the flat
    fixture rotated into place, T_Q K T_Q^T with Q recovered from the node
    coordinates; and its coordinate gradients sum to zero (translation)."""
    flat, tilt = cases()[idx], tilted_cases()[idx]
    Kf, Kt = golden_dense(flat), np.zeros((24, 24))
    for i, j, v in tilt["entries"]:
        Kt[i, j] = float(v)
    Xf, Xt = np.array(flat["xyz"]), np.array(tilt["xyz"])
    A = Xf[[0, 1, 3]] - Xf[2]
    B = Xt[[0, 1, 3]] - Xt[2]
    Q = np.linalg.lstsq(A, B, rcond=None)[0].T  # B = A Q^T; A has rank 2: fix the normal
    n = np.cross(Xt[3] - Xt[2], Xt[1] - Xt[2])
    Q[:, 2] = n / np.linalg.norm(n)
    T = np.kron(np.eye(8), Q)
    assert np.abs(Kt - T @ Kf @ T.T).max() <= 1e-12 * np.abs(Kf).max()
    assert np.abs(Kt - Kt.T).max() <= 1e-15 * np.abs(Kt).max()
    for key in ("G", "u"):
        gx = np.array([[float(v) for v in r] for r in tilt["grad_" + key]["xyz"]])
        assert np.abs(gx.sum(axis=0)).max() <= 1e-13 * np.abs(gx).max()


@pytest.mark.parametrize("idx", range(4))
def test_oracle_matches_tilted_golden(idx):
    case = tilted_cases()[idx]
    om = OracleModel(1e-4, 1e-12, 4)
    for n, p in enumerate(case["xyz"]):
        om.add_node(n + 1, *p)
    om.add_plate(1, 1, 2, 3, 4, *case["props"])
    rp, ci, v = om.csr(structural=True)
    rows = np.repeat(np.arange(24), np.diff(rp))
    K = np.zeros((24, 24))
    for i, j, val in case["entries"]:
        K[i, j] = float(val)
    assert set(zip(rows.tolist(), ci.tolist())) == {(i, j) for i, j, _ in case["entries"]}
    r = compare((rp, ci, v), (rp, ci, K[rows, ci], om.csr_magnitude()))
    print("tilted case", idx, r)
    assert r["pattern_equal"] and r["max_ratio"] <= 1.0, r
