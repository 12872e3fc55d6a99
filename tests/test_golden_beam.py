"""The FP64 oracle's beam stiffness against the 50-digit fixtures of x-aligned
beams (tests/golden/beam_golden.json, written by make_beam_golden.py from
src/fem/structs/beam.rs:260-605).  The fixtures are checked for what the
mathematics fixes (symmetry, nullity 6, E dL/dE = L since every term is
proportional to E at fixed nu), then the oracle must meet the parity bound of
tests/parity.py with an identical pattern.  test_gpu_golden_beam.py holds the
device assembly and its gradients to the same fixtures."""

import json
import os

import numpy as np
import pytest

from oracle.oracle import OracleModel
from tests.parity import compare

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "beam_golden.json")


def fixture():
    with open(FIXTURE) as f:
        return json.load(f)


def cases():
    return fixture()["cases"]


def golden_dense(case):
    K = np.zeros((12, 12))
    for i, j, v in case["entries"]:
        K[i, j] = float(v)
    return K


def one_beam_oracle(case):
    om = OracleModel(1e-4, 1e-12, 2)
    for n, p in enumerate(case["xyz"]):
        om.add_node(n + 1, *p)
    p = case["props"]
    om.add_beam(1, 1, 2, *p[:8], p[8:])
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
    for d in range(3):
        u = np.zeros(12)
        u[d::6] = 1.0
        assert np.abs(K @ u).max() <= 1e-12 * scale
    w = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(w) < 1e-12 * np.abs(w).max()) == 6
    assert np.all(w > -1e-12 * np.abs(w).max())


@pytest.mark.parametrize("key", ["G", "u"])
@pytest.mark.parametrize("idx", range(3))
def test_fixture_gradient_homogeneity(idx, key):
    case = cases()[idx]
    lab, L = grad_scale(case, key)
    E = case["props"][0]
    assert abs(E * float(case["grad_" + key][0]) - L) <= 1e-14 * lab


@pytest.mark.parametrize("idx", range(3))
def test_oracle_matches_golden(idx):
    case = cases()[idx]
    om = one_beam_oracle(case)
    r = compare(om.csr(structural=True), golden_reference(case, om))
    print("beam case", idx, r)
    assert r["pattern_equal"] and r["max_ratio"] <= 1.0, r
    assert r["normwise"] < 1e-14, r
