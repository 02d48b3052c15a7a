"""FP64 CPU restatement of the ring-polymer drift (test infrastructure).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
kernel in md_engine_amd/csrc/ring.hip; the product path never imports it.

The free ring polymer of one atom, P beads coupled by springs
1/2 m omegaP^2 |x_n - x_{n+1}|^2, is propagated exactly in the normal modes
of nve_xPIMD_cu (src/Integrators/IntegratorVerlet.cu:63-304): the
orthonormal matrix C (k = 0: 1/sqrt P; 1 <= k < P/2: sqrt(2/P) cos(2 pi k n/P);
k = P/2: (-1)^n/sqrt P; P/2 < k < P: sqrt(2/P) sin(2 pi k n/P), n = 1..P),
mode 0 drifting and mode k rotating at w_k = 2 omegaP sin(k pi/P).

Pinned by: C orthonormal; w_k^2 the eigenvalues of the ring Laplacian times
omegaP^2 with C's columns its eigenvectors; one exact step against many small
velocity-Verlet steps of the spring chain itself.
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
