"""Pins the oracle's element post-processing (extract_element_analysis_result:
truss.rs:281-333, beam.rs:803-993, plate.rs:1196-1409) against the crate's
golden test, closed forms (bar extension, torsion, Timoshenko shear and
bending, plate patch tests), an energy identity that ties the beam forces to
the independently pinned stiffness, statics (dM/dx = V) and rigid-body
modes.  CPU only."""

import numpy as np
import pytest

from oracle.oracle import OracleModel

TRUSS, BEAM, PLATE = 0, 1, 2


def test_crate_golden_truss_force():
    """src/tests/fem/test_fem.rs:60-61: ForceR = 100 for u2x = 0.0015."""
    m = OracleModel(1e-4, 1e-12, 2)
    m.add_node(1, 0.0, 0.0, 0.0)
    m.add_node(2, 30.0, 0.0, 0.0)
    m.add_truss(1, 1, 2, 1e6, 2.0, None)
    u = np.zeros(12)
    u[6] = 0.0015
    assert m.element_forces(TRUSS, u)[0, 0] == pytest.approx(100.0, rel=1e-14)


def test_truss_axial_force_random_3d():
    rng = np.random.default_rng(1)
    for trial in range(20):
        p1, p2 = rng.normal(size=3), rng.normal(size=3)
        E, A = rng.uniform(1e9, 3e11), rng.uniform(1e-4, 1e-2)
        A2 = rng.uniform(1e-4, 1e-2) if trial % 2 else None
        m = OracleModel(1e-4, 1e-12, 2)
        m.add_node(1, *p1)
        m.add_node(2, *p2)
        m.add_truss(1, 1, 2, E, A, A2)
        u = rng.normal(size=12) * 1e-3
        d = p2 - p1
        L = np.linalg.norm(d)
        area = A if A2 is None else (A + A2) / 2
        want = E * area * (d @ (u[6:9] - u[0:3])) / L ** 2
        assert m.element_forces(TRUSS, u)[0, 0] == pytest.approx(want, rel=1e-10, abs=1e-12 * abs(E * area))


def beam_x(L=2.0, E=2.1e11, nu=0.3, A=3e-3, i11=4e-6, i22=2e-6, it=1e-6, k=5 / 6):
    m = OracleModel(1e-4, 1e-12, 2)
    m.add_node(1, 0.0, 0.0, 0.0)
    m.add_node(2, L, 0.0, 0.0)
    m.add_beam(1, 1, 2, E, nu, A, i11, i22, 0.0, it, k, (0.0, 0.0, 1.0))
    return m, dict(L=L, E=E, G=E / (2 * (1 + nu)), A=A, i11=i11, i22=i22, it=it, k=k)


def test_beam_closed_forms():
    """Beam along +x with local axis 1 = +z (identity frame): extension,
    twist, shear in v and w, and uniform curvature."""
    m, p = beam_x()
    L, E, G, A, k = p["L"], p["E"], p["G"], p["A"], p["k"]
    names = ["R", "S", "T", "MR", "MS1", "MS", "MS2", "MT1", "MT", "MT2"]

    def forces(u):
        return dict(zip(names, m.element_forces(BEAM, u)[0]))

    u = np.zeros(12)
    u[6] = 1e-3
    f = forces(u)
    assert f["R"] == pytest.approx(E * A * 1e-3 / L, rel=1e-13)
    u = np.zeros(12)
    u[9] = 2e-3  # twist
    assert forces(u)["MR"] == pytest.approx(G * p["it"] * 2e-3 / L, rel=1e-13)
    u = np.zeros(12)
    u[7] = 1e-3  # v2: 1-point Timoshenko shear strain (v2 - v1) / L - (tw1 + tw2) / 2
    f = forces(u)
    assert f["S"] == pytest.approx(G * A * k * 1e-3 / L, rel=1e-13)
    assert f["MT1"] - f["MT2"] == pytest.approx(L * f["S"], rel=1e-13)
    u = np.zeros(12)
    u[8] = 1e-3  # w2: (w2 - w1) / L - (tv1 + tv2) / 2 (beam.rs:346-399 sign)
    f = forces(u)
    assert f["T"] == pytest.approx(G * A * k * 1e-3 / L, rel=1e-13)
    assert f["MS1"] - f["MS2"] == pytest.approx(L * f["T"], rel=1e-13)
    kap = 1e-3  # uniform curvature about local z: v = kap x^2 / 2, theta_w = kap x
    u = np.zeros(12)
    u[7], u[11] = kap * L * L / 2, kap * L
    f = forces(u)
    assert f["MT"] == pytest.approx(E * p["i11"] * kap, rel=1e-12)
    assert abs(f["S"]) <= 1e-12 * G * A * k * kap * L


def test_beam_energy_identity_random_frames():
    """u^T K_e u = L sum_t F_t^2 / c_t with c = (EA, GAk, GAk, G It, E I22',
    E I11'): the forces agree with the stiffness in any frame."""
    rng = np.random.default_rng(2)
    for trial in range(10):
        p1, p2 = rng.normal(size=3), rng.normal(size=3)
        v1 = rng.normal(size=3)
        E, nu, A = 2.1e11, 0.3, rng.uniform(1e-3, 1e-2)
        i11, i22, it = rng.uniform(1e-6, 1e-5), rng.uniform(1e-7, 1e-6), rng.uniform(1e-7, 1e-6)
        k = 5 / 6
        m = OracleModel(1e-4, 1e-12, 2)
        m.add_node(1, *p1)
        m.add_node(2, *p2)
        m.add_beam(1, 1, 2, E, nu, A, i11, i22, 0.0, it, k, tuple(v1))
        K = m.element_matrix(BEAM, 1, global_frame=True)
        u = rng.normal(size=12) * 1e-3
        ue = np.concatenate([u[0:6], u[6:12]])
        f = m.element_forces(BEAM, u)[0]
        G = E / (2 * (1 + nu))
        L = np.linalg.norm(p2 - p1)
        c = [E * A, G * A * k, G * A * k, G * it, E * i22, E * i11]  # I12 = 0: I11' = I11, I22' = I22
        comps = [f[0], f[1], f[2], f[3], f[5], f[8]]
        energy = L * sum(x * x / ci for x, ci in zip(comps, c))
        assert ue @ K @ ue == pytest.approx(energy, rel=1e-9)
        assert f[4] - f[6] == pytest.approx(L * f[2], rel=1e-10, abs=1e-9 * abs(L * f[2]) + 1e-12)
        assert f[7] - f[9] == pytest.approx(L * f[1], rel=1e-10, abs=1e-9 * abs(L * f[1]) + 1e-12)


def plate(xy, E=7e10, nu=0.3, t=0.01, k=5 / 6):
    """One plate in z = 0 with nodes 1(+,+), 2(-,+), 3(-,-), 4(+,-) at xy."""
    m = OracleModel(1e-4, 1e-12, 4)
    for i, (x, y) in enumerate(xy):
        m.add_node(i + 1, x, y, 0.0)
    m.add_plate(1, 1, 2, 3, 4, E, nu, t, k)
    return m


RECT = [(2.0, 1.0), (0.0, 1.0), (0.0, 0.0), (2.0, 0.0)]
DIST = [(2.1, 1.2), (-0.1, 0.9), (0.0, 0.0), (1.9, -0.2)]


def field(xy, f):
    u = np.zeros(24)
    for a, (x, y) in enumerate(xy):
        u[6 * a:6 * a + 6] = f(x, y)
    return u


@pytest.mark.parametrize("xy", [RECT, DIST])
def test_plate_membrane_and_bending_patch(xy):
    E, nu, t = 7e10, 0.3, 0.01
    m = plate(xy, E, nu, t)
    C = np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])
    a, b, c, d = 1e-4, 2e-5, -3e-5, 5e-5  # u = a x + b y, v = c x + d y
    f = m.element_forces(PLATE, field(xy, lambda x, y: [a * x + b * y, c * x + d * y, 0, 0, 0, 0]))[0]
    want = E * t / (1 - nu ** 2) * (C @ [a, d, b + c])
    assert np.allclose(f[0:3], want, rtol=1e-10, atol=1e-10 * np.abs(want).max())
    assert np.allclose(f[3:8], 0.0, atol=1e-10 * np.abs(want).max())
    p_, q_, r_, s_ = 1e-3, -4e-4, 2e-4, 6e-4  # theta_x = p x + q y, theta_y = r x + s y
    f = m.element_forces(PLATE, field(xy, lambda x, y: [0, 0, 0, p_ * x + q_ * y, r_ * x + s_ * y, 0]))[0]
    kap = np.array([-r_, q_, p_ - s_])  # B_bend rows (plate.rs:276-390)
    mom = E * t ** 3 / (12 * (1 - nu ** 2)) * (C @ kap)
    assert np.allclose([f[4], f[3], f[5]], mom, rtol=1e-10, atol=1e-10 * np.abs(mom).max())
    assert np.allclose(f[0:3], 0.0, atol=1e-10 * np.abs(mom).max())


def test_plate_shear_rectangle():
    """Constant transverse shear on a rectangle: MITC4 reproduces it, and the
    natural r, s directions are x, y."""
    E, nu, t, k = 7e10, 0.3, 0.01, 5 / 6
    m = plate(RECT, E, nu, t, k)
    g, h = 1e-4, -3e-4
    f = m.element_forces(PLATE, field(RECT, lambda x, y: [0, 0, g * x + h * y, 0, 0, 0]))[0]
    G = E / (2 * (1 + nu))
    assert f[6] == pytest.approx(G * t * k * g, rel=1e-10)
    assert f[7] == pytest.approx(G * t * k * h, rel=1e-10)


@pytest.mark.parametrize("xy", [RECT, DIST])
def test_plate_rigid_modes(xy):
    m = plate(xy)
    rng = np.random.default_rng(3)
    for _ in range(5):
        tr = rng.normal(size=3) * 1e-3
        om = rng.normal(size=3) * 1e-3
        u = field(xy, lambda x, y: list(tr + np.cross(om, [x, y, 0.0])) + list(om))
        f = m.element_forces(PLATE, u)[0]
        assert np.abs(f).max() <= 1e-6 * 7e10 * 0.01 * 1e-3
