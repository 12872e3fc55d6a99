"""The FP64 oracle's truss stiffness against the 50-digit fixtures of 3-D bars
(tests/golden/truss_golden.json, written by make_truss_golden.py from
src/fem/structs/truss.rs).  The fixtures are checked for what the mathematics
fixes (symmetry, rank 1, the translations and rotations as null space,
E dL/dE = L, coordinate gradients summing to zero), then the oracle must meet the parity bound of tests/parity.py with
an identical pattern.  test_gpu_golden_truss.py holds the device assembly and
its gradients to the same fixtures."""

import json
import os

import numpy as np
import pytest

from oracle.oracle import OracleModel
from tests.parity import compare

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "truss_golden.json")


def cases():
    with open(FIXTURE) as f:
        return json.load(f)["cases"]


def golden_dense(case):
    K = np.zeros((12, 12))
    for i, j, v in case["entries"]:
        K[i, j] = float(v)
    return K


def one_truss_oracle(case):
    om = OracleModel(1e-4, 1e-12, 2)
    for n, p in enumerate(case["xyz"]):
        om.add_node(n + 1, *p)
    E, A, A2 = case["props"]
    om.add_truss(1, 1, 2, E, A, None if np.isnan(A2) else A2)
    return om


def golden_reference(case, om):
    rp, ci, _ = om.csr(structural=True)
    rows = np.repeat(np.arange(12), np.diff(rp))
    assert set(zip(rows.tolist(), ci.tolist())) == {(i, j) for i, j, _ in case["entries"]}
    return rp, ci, golden_dense(case)[rows, ci], om.csr_magnitude()


def weights(case, key):
    if key == "G":
        return np.array(case["G"])
    u = np.array(case["u"])
    return np.array([-u[i] * u[j] for i, j, _ in case["entries"]])


def grad_scale(case, key):
    K = golden_dense(case)
    w = weights(case, key)
    vals = np.array([K[i, j] for i, j, _ in case["entries"]])
    return float(np.abs(w * vals).sum()), float(np.dot(w, vals))


@pytest.mark.parametrize("idx", range(3))
def test_fixture_invariants(idx):
    case = cases()[idx]
    K = golden_dense(case)
    scale = np.abs(K).max()
    assert np.abs(K - K.T).max() <= 1e-15 * scale
    w = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(w) > 1e-12 * np.abs(w).max()) == 1  # rank 1
    x = np.array(case["xyz"])
    for k in range(3):  # translations and rotations about the three axes
        t = np.zeros(12)
        t[k::6] = 1.0
        assert np.abs(K @ t).max() <= 1e-12 * scale
        axis = np.eye(3)[k]
        r = np.zeros(12)
        for a in range(2):
            r[6 * a:6 * a + 3] = np.cross(axis, x[a])
        assert np.abs(K @ r).max() <= 1e-12 * scale * np.abs(x).max()


@pytest.mark.parametrize("key", ["G", "u"])
@pytest.mark.parametrize("idx", range(3))
def test_fixture_gradient_identities(idx, key):
    case = cases()[idx]
    g = case["grad_" + key]
    lab, L = grad_scale(case, key)
    E = case["props"][0]
    assert abs(E * float(g["params"][0]) - L) <= 1e-14 * lab
    gx = np.array([[float(v) for v in r] for r in g["xyz"]])
    x = np.array(case["xyz"])
    size = np.abs(x).max()
    assert np.abs(gx.sum(axis=0)).max() <= 1e-14 * lab / size


@pytest.mark.parametrize("idx", range(3))
def test_oracle_matches_golden(idx):
    case = cases()[idx]
    om = one_truss_oracle(case)
    r = compare(om.csr(structural=True), golden_reference(case, om))
    print("truss case", idx, r)
    assert r["pattern_equal"] and r["max_ratio"] <= 1.0, r
    assert r["normwise"] < 1e-14, r
