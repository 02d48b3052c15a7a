"""FP64 CPU restatement of the ring-polymer drift (test infrastructure).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
kernel in md_engine_amd/csrc/ring.hip; the product path never imports it.

The free ring polymer of one atom, P beads coupled by springs
1/2 m omegaP^2 |x_n - x_{n+1}|^2, is propagated exactly in the normal modes
of nve_xPIMD_cu (src/Integrators/IntegratorVerlet.cu:63-304): the
orthonormal matrix C (k = 0: 1/sqrt P; 1 <= k < P/2: sqrt(2/P) cos(2 pi k n/P);
k = P/2: (-1)^n/sqrt P; P/2 < k < P: sqrt(2/P) sin(2 pi k n/P), n = 1..P),
mode 0 drifting and mode k rotating at w_k = 2 omegaP sin(k pi/P).

Ring-polymer contraction of the mesh part (FixChargeEwald.cu:1255-1266,
1395-1405): ``centroids`` follows computeCentroid (FixChargeEwald.cu:22-50),
bead 0 plus the mean of the other beads' minimum-image offsets from it,
wrapped into the box; ``contracted_pppm`` spreads the charges from the
centroids, interpolates the field there and gives every bead the whole
force q E, with alpha from the secant solve over all P n ring atoms and one
bead's Q^2 (setTotalQ2 divides by nPerRingPoly, FixChargeEwald.cu:885-900).

Pinned by: C orthonormal; w_k^2 the eigenvalues of the ring Laplacian times
omegaP^2 with C's columns its eigenvectors; one exact step against many small
velocity-Verlet steps of the spring chain itself; centroids of rings that
straddle the box edge against the plain mean of the unwrapped beads; the
contracted force of P coincident beads against the single-system P3M; and
the contracted force on each bead against minus the derivative of P times
the P3M energy at the centroids (the ring Hamiltonian's mesh term).
"""
from __future__ import annotations

import numpy as np


def modes(P):
    C = np.zeros((P, P))
    h = P // 2
    for nb in range(P):
        n = nb + 1
        for k in range(P):
            if k == 0:
                C[nb, k] = np.sqrt(1.0 / P)
            elif k < h:
                C[nb, k] = np.sqrt(2.0 / P) * np.cos(2 * np.pi * k * n / P)
            elif k == h:
                C[nb, k] = np.sqrt(1.0 / P) * (-1.0 if n % 2 else 1.0)
            else:
                C[nb, k] = np.sqrt(2.0 / P) * np.sin(2 * np.pi * k * n / P)
    return C


def frequencies(P, omega_p):
    return 2.0 * omega_p * np.sin(np.arange(P) * np.pi / P)


def free_step(x, v, omega_p, dt):
    """x, v: (P, N, 3) bead positions (already in one image) and velocities"""
    P = x.shape[0]
    C = modes(P)
    w = frequencies(P, omega_p)
    xt = np.einsum("nk,nad->kad", C, x)
    vt = np.einsum("nk,nad->kad", C, v)
    x1, v1 = np.empty_like(xt), np.empty_like(vt)
    x1[0] = xt[0] + vt[0] * dt
    v1[0] = vt[0]
    c, s = np.cos(w[1:] * dt)[:, None, None], np.sin(w[1:] * dt)[:, None, None]
    wk = w[1:, None, None]
    x1[1:] = xt[1:] * c + vt[1:] * s / wk
    v1[1:] = vt[1:] * c - xt[1:] * s * wk
    return np.einsum("nk,kad->nad", C, x1), np.einsum("nk,kad->nad", C, v1)


def spring_accel(x, omega_p):
    """-omegaP^2 (2 x_n - x_{n-1} - x_{n+1}) per bead (mass cancels)"""
    return -omega_p ** 2 * (2 * x - np.roll(x, 1, axis=0) - np.roll(x, -1, axis=0))


def spring_energy(x, mass, omega_p, mvv_to_eng=1.0):
    """sum over beads and atoms of 1/2 m omegaP^2 |x_n - x_{n+1}|^2"""
    d = x - np.roll(x, -1, axis=0)
    return 0.5 * mvv_to_eng * omega_p ** 2 * float((mass[None, :, None] * d ** 2).sum())


def centroids(xs, lo, L):
    """xs: (P, N, 3) bead positions; centroid of every atom's ring in [lo, lo + L)"""
    P = xs.shape[0]
    x0 = xs[0]
    d = xs[1:] - x0
    d = d - L * np.rint(d / L)
    c = x0 + d.sum(0) / P
    return c - L * np.floor((c - lo) / L)


def contracted_pppm(xs, q, lo, L, sz, order, qqr, rcut, energy=False):
    """Mesh part of a ring of P beads (xs: (P, N, 3)): alpha, the force every
    bead takes (N, 3) and, with energy, the P3M energy at the centroids (of
    one configuration; the ring Hamiltonian carries P times it)."""
    from . import pppm, pppm_alpha, pppm_green, total_q2

    P, n = xs.shape[:2]
    sz = np.asarray(sz)
    q2 = total_q2(q, qqr)
    alpha, _ = pppm_alpha(P * n, q2, L, sz, order, rcut)
    G = pppm_green(L, sz, alpha, order)
    r = pppm(centroids(xs, lo, L), q, lo, L, sz, order, qqr, alpha, q2, G, energy=energy)
    return (alpha, r["f"], r["energy"]) if energy else (alpha, r["f"])
