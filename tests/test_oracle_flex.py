"""Pins the FP64 flexible-water restatement (oracle/flexwater.py): closed
forms of a stretched bond and a bent angle, finite differences of the
intramolecular energy, the M site on the bisector, and the M-force spread
conserving force and torque."""
import numpy as np
import pytest

from oracle import flexwater as FW


def one_water(r1=0.9419, r2=0.9419, theta=107.4):
    th = np.deg2rad(theta)
    x = np.array([[5.0, 5.0, 5.0], [5.0 + r1, 5.0, 5.0], [5.0 + r2 * np.cos(th), 5.0 + r2 * np.sin(th), 5.0],
                  [5.0, 5.0, 5.0]])
    return np.array([[0, 1, 2, 3]]), x, np.array([20.0, 20.0, 20.0])


def test_equilibrium_has_no_force():
    mol, x, L = one_water()
    f, e, _ = FW.bonded(mol, x, L)
    assert e == pytest.approx(0.0, abs=1e-20)
    assert np.abs(f).max() < 1e-10


def test_stretched_bond_closed_form():
    p = FW.QTIP4PF
    dr = 0.05
    mol, x, L = one_water(r1=p["r0"] + dr)
    f, e, _ = FW.bonded(mol, x, L)
    assert e == pytest.approx(p["k2"] * dr ** 2 + p["k3"] * dr ** 3 + p["k4"] * dr ** 4, rel=1e-12)
    # the bond lies along x: F_H1 = -dU/dr x_hat, and the angle term is at rest
    assert f[1, 0] == pytest.approx(-(2 * p["k2"] * dr + 3 * p["k3"] * dr ** 2 + 4 * p["k4"] * dr ** 3), rel=1e-12)
    assert abs(f[1, 1]) < 1e-9 and abs(f[1, 2]) < 1e-12


def test_bent_angle_closed_form():
    p = FW.QTIP4PF
    mol, x, L = one_water(theta=107.4 + 5.0)
    f, e, _ = FW.bonded(mol, x, L)
    dth = np.deg2rad(5.0)
    assert e == pytest.approx(0.5 * p["ktheta"] * dth ** 2, rel=1e-9)
    # H1 on the x axis: the angle force is along y with magnitude ktheta dth / r
    assert f[1, 1] == pytest.approx(p["ktheta"] * dth / p["r0"], rel=1e-9)
    assert abs(f[1, 0]) < 1e-9


def test_forces_are_minus_gradient():
    rng = np.random.default_rng(3)
    mol, x, L = one_water(r1=0.97, r2=0.91, theta=101.0)
    x = x + rng.normal(0, 0.02, x.shape)
    x[3] = x[0]
    f, _, W = FW.bonded(mol, x, L)
    h = 1e-6
    for a in range(3):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a, d] += h
            xm[a, d] -= h
            fd = -(FW.bonded(mol, xp, L)[1] - FW.bonded(mol, xm, L)[1]) / (2 * h)
            assert f[a, d] == pytest.approx(fd, rel=1e-6, abs=1e-7)
    assert np.allclose(f.sum(0), 0.0, atol=1e-10)
    # virial trace = -dU/dlambda under scaling of the molecule about O
    up = FW.bonded(mol, x[0] + (x - x[0]) * (1 + h), L)[1]
    um = FW.bonded(mol, x[0] + (x - x[0]) * (1 - h), L)[1]
    assert W[:3].sum() == pytest.approx(-(up - um) / (2 * h), rel=1e-6)


def test_m_site_on_bisector_and_spread_conserves():
    g = FW.QTIP4PF["gamma"]
    mol, x, L = one_water(r1=0.95, r2=0.95, theta=104.0)
    y = FW.place_m(mol, x, L, g)
    b = (x[1] - x[0]) + (x[2] - x[0])
    assert np.allclose(y[3] - y[0], 0.5 * (1 - g) * b, atol=1e-14)
    rng = np.random.default_rng(1)
    f = rng.normal(size=(4, 3))
    s = FW.spread_m(mol, f, g)
    assert np.allclose(s.sum(0), f.sum(0), atol=1e-12)
    assert np.all(s[3] == 0)
    tq = lambda F: np.cross(y - y[0], F).sum(0)  # noqa: E731
    assert np.allclose(tq(s), tq(f), atol=1e-12)
