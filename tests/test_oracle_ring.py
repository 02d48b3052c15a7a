"""Pins the ring-polymer restatement (oracle/ring.py): the normal-mode matrix
is orthonormal and diagonalises the ring Laplacian with eigenvalues
(w_k / omegaP)^2, and one exact step equals many small velocity-Verlet steps
of the spring chain."""
import numpy as np
import pytest

from oracle import ring as R


@pytest.mark.parametrize("P", [2, 4, 8, 32])
def test_modes_orthonormal_and_diagonalise_the_ring(P):
    C = R.modes(P)
    assert np.allclose(C.T @ C, np.eye(P), atol=1e-13)
    A = 2 * np.eye(P) - np.roll(np.eye(P), 1, axis=0) - np.roll(np.eye(P), -1, axis=0)
    D = C.T @ A @ C
    w = R.frequencies(P, 1.0)
    assert np.allclose(D, np.diag(w ** 2), atol=1e-12)


def test_exact_step_matches_fine_verlet():
    rng = np.random.default_rng(2)
    P, N, wP, dt = 8, 5, 0.07, 1.0
    x = rng.normal(0, 0.3, (P, N, 3))
    v = rng.normal(0, 0.01, (P, N, 3))
    xe, ve = R.free_step(x, v, wP, dt)
    m = 4000
    h = dt / m
    xv, vv = x.copy(), v.copy()
    a = R.spring_accel(xv, wP)
    for _ in range(m):
        vv += 0.5 * h * a
        xv += h * vv
        a = R.spring_accel(xv, wP)
        vv += 0.5 * h * a
    assert np.allclose(xe, xv, atol=1e-7)
    assert np.allclose(ve, vv, atol=1e-8)
    # the centroid drifts freely
    assert np.allclose(xe.mean(0), x.mean(0) + dt * v.mean(0), atol=1e-13)
