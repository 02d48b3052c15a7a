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


def test_centroids_across_the_box_edge():
    """Beads of a ring that straddles a periodic face: the centroid is the
    plain mean of the unwrapped beads, put back in the box."""
    rng = np.random.default_rng(3)
    L = np.array([10.0, 12.0, 14.0])
    lo = np.array([-1.0, 0.0, 2.0])
    P, N = 8, 40
    c0 = lo + rng.uniform(0, 1, (N, 3)) * L
    xs_unwrapped = c0[None] + rng.normal(0, 0.4, (P, N, 3))
    xs = xs_unwrapped - L * np.floor((xs_unwrapped - lo) / L)  # every bead wrapped on its own
    want = xs_unwrapped.mean(0)
    want -= L * np.floor((want - lo) / L)
    got = R.centroids(xs, lo, L)
    d = got - want
    d -= L * np.rint(d / L)  # a centroid on the face itself may land on either side
    assert np.abs(d).max() < 1e-12
    assert np.all(got >= lo) and np.all(got < lo + L)


@pytest.fixture(scope="module")
def charged_ring(orc):
    rng = np.random.default_rng(5)
    L = np.full(3, 18.0)
    lo = np.zeros(3)
    N, P = 60, 4
    x = rng.uniform(0, 18.0, (N, 3))
    q = np.where(np.arange(N) % 2 == 0, 0.8, -0.8)
    xs = x[None] + rng.normal(0, 0.15, (P, N, 3))
    return dict(L=L, lo=lo, x=x, q=q, xs=xs, sz=np.array([32, 32, 32]), order=5, qqr=332.06371, rc=6.0)


def test_contraction_of_coincident_beads_is_the_plain_mesh_part(orc, charged_ring):
    """P beads on top of each other: every bead takes the single-system P3M
    force (alpha solved for P n atoms), not a 1/P share of it."""
    r = charged_ring
    P = r["xs"].shape[0]
    xs = np.repeat(r["x"][None], P, axis=0)
    alpha, f = R.contracted_pppm(xs, r["q"], r["lo"], r["L"], r["sz"], r["order"], r["qqr"], r["rc"])
    q2 = orc.total_q2(r["q"], r["qqr"])
    a_ref, _ = orc.pppm_alpha(P * len(r["x"]), q2, r["L"], r["sz"], r["order"], r["rc"])
    a_one, _ = orc.pppm_alpha(len(r["x"]), q2, r["L"], r["sz"], r["order"], r["rc"])
    # both error estimates scale as N^-1/2, so their balance point does not
    # move with the atom count (and a power-of-two P scales exactly)
    assert alpha == a_ref == a_one
    G = orc.pppm_green(r["L"], r["sz"], alpha, r["order"])
    ref = orc.pppm(r["x"], r["q"], r["lo"], r["L"], r["sz"], r["order"], r["qqr"], alpha, q2, G)["f"]
    assert np.allclose(f, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_contracted_force_is_minus_the_ring_energy_gradient(orc, charged_ring):
    """The ring Hamiltonian's mesh term is P E_mesh(centroids); its derivative
    in one bead's coordinate is the whole centroid force (the centroid moves
    by 1/P of the bead).  P3M's ik-differentiated force is the gradient of its
    energy only to the method's accuracy, so the bar is 2 % of the force,
    far inside the factor P a wrong share would give."""
    r = charged_ring
    xs = r["xs"].copy()
    P = xs.shape[0]
    _, f = R.contracted_pppm(xs, r["q"], r["lo"], r["L"], r["sz"], r["order"], r["qqr"], r["rc"])
    h = 1e-4
    k, atoms = 2, (0, 7, 31)
    for i in atoms:
        for d in range(3):
            es = []
            for s in (1, -1):
                xx = xs.copy()
                xx[k, i, d] += s * h
                es.append(P * R.contracted_pppm(xx, r["q"], r["lo"], r["L"], r["sz"], r["order"], r["qqr"],
                                                r["rc"], energy=True)[2])
            fd = -(es[0] - es[1]) / (2 * h)
            assert fd == pytest.approx(f[i, d], rel=0.02, abs=0.02 * np.abs(f[i]).max())
