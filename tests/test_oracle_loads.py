"""Pins the oracle's distributed loads (methods_for_bc_data_handle.rs:58-174,
beam.rs:775-797, plate.rs:1145-1185) by closed forms: q L / 2 per beam node,
p A / 4 per node of a rectangle, and the total load and its first moment on
distorted quads (bilinear geometry: the total is exact under the 2 x 2
rule, the first moment up to the f32-rounded points).  CPU only."""

import numpy as np
import pytest

from oracle.oracle import OracleError, OracleModel


def test_line_load():
    m = OracleModel(1e-4, 1e-12, 2)
    m.add_node(1, 0.0, 0.0, 0.0)
    m.add_node(2, 1.0, 2.0, 2.0)  # L = 3
    m.add_beam(1, 1, 2, 2.1e11, 0.3, 1e-3, 2e-6, 1e-6, 0.0, 1e-6, 5 / 6, (0.0, 0.0, 1.0))
    m.add_uniformly_distributed_line_load(1, 2, -10.0)
    m.add_uniformly_distributed_line_load(1, 2, -10.0)  # loads add up
    f = m.forces(2).reshape(2, 6)
    assert f[0, 2] == pytest.approx(-30.0, rel=1e-15) and f[1, 2] == pytest.approx(-30.0, rel=1e-15)
    assert np.count_nonzero(f) == 2
    with pytest.raises(OracleError, match=r"^Beam element with number 7 does not exist!$"):
        m.add_uniformly_distributed_line_load(7, 2, 1.0)


def quad(xy, z=None):
    m = OracleModel(1e-4, 1e-12, 4)
    for i, (x, y) in enumerate(xy):
        m.add_node(i + 1, x, y, 0.0 if z is None else z[i])
    m.add_plate(1, 1, 2, 3, 4, 7e10, 0.3, 0.01, 5 / 6)
    return m


def test_surface_load_rectangle():
    m = quad([(2.0, 1.0), (0.0, 1.0), (0.0, 0.0), (2.0, 0.0)])
    m.add_uniformly_distributed_surface_load(1, 2, -100.0)
    f = m.forces(4).reshape(4, 6)
    assert np.allclose(f[:, 2], -50.0, rtol=1e-14, atol=0)
    with pytest.raises(OracleError, match=r"^Plate element with number 2 does not exist!$"):
        m.add_uniformly_distributed_surface_load(2, 2, 1.0)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_surface_load_distorted_total_and_moment(seed):
    rng = np.random.default_rng(seed)
    base = np.array([(1.0, 1.0), (-1.0, 1.0), (-1.0, -1.0), (1.0, -1.0)])
    xy = base + rng.uniform(-0.25, 0.25, size=(4, 2))
    m = quad([tuple(p) for p in xy])
    p = 3.5
    m.add_uniformly_distributed_surface_load(1, 0, p)
    f = m.forces(4).reshape(4, 6)[:, 0]
    x, y = xy[:, 0], xy[:, 1]
    area = 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))
    cx = np.dot(x + np.roll(x, -1), x * np.roll(y, -1) - np.roll(x, -1) * y) / (6 * area)
    assert f.sum() == pytest.approx(p * area, rel=1e-13)
    # the points are +-sqrt(f32(1/3)) (plate.rs:1066-1091), exact for the total but
    # not for the first moment's cubic terms: ~1e-8 relative, O(1) for a wrong weight
    assert (f * x).sum() == pytest.approx(p * area * cx, rel=1e-7)
