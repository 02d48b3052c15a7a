"""FP64 CPU restatement of the flexible q-TIP4P/F water terms (test infrastructure).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
kernels in md_engine_amd/csrc/flex.hip; the product path never imports it.

* quartic O-H bonds, U = k2 dr^2 + k3 dr^3 + k4 dr^4, dr = r - r0
  (BondEvaluatorQuartic::force/energy, src/Evaluators/BondEvaluatorQuartic.h:9-36);
* harmonic H-O-H angle, U = ktheta/2 (theta - theta0)^2
  (AngleEvaluatorHarmonic, src/Evaluators/AngleEvaluatorHarmonic.h:12-78,
  whose per-atom energy share is ktheta/6 dtheta^2);
* massless M site at gamma r_O + (1 - gamma)/2 (r_H1 + r_H2), hydrogens in
  the oxygen's image, and its force spread gamma F_M onto O and
  (1 - gamma)/2 F_M onto each H (setMSiteFlexible / distributeMSiteFlexible,
  src/Fixes/FixTIP4PFlexible.cu:92-300).

Parameters of q-TIP4P/F as the reference's example sets them
(examples/PIMD_TIP4P.py: k2 607.19, k3 -1388.65, k4 1852.58 kcal/mol/A^n,
r0 0.9419 A, ktheta 87.85 kcal/mol/rad^2, theta0 107.4 deg) and
FixTIP4PFlexible's q-TIP4P/F gamma 0.73612 (FixTIP4PFlexible.cu:325-330).

Pinned by: finite differences of ``bonded`` energies; closed forms for a
stretched bond and a bent angle; the M site on the H-O-H bisector at
(1 - gamma) |r_OH + r_OH'| / 2 from O; the spread conserving total force and
torque about O.
"""
from __future__ import annotations

import numpy as np

QTIP4PF = dict(k2=607.19, k3=-1388.65, k4=1852.58, r0=0.9419, ktheta=87.85, theta0=107.4 / 180.0 * np.pi,
               gamma=0.73612)


def _mi(d, L):
    return d - L * np.rint(d / L)


def place_m(mol, x, L, gamma):
    """positions with every M site (mol[:, 3] >= 0) placed"""
    x = np.array(x, np.float64, copy=True)
    m = mol[mol[:, 3] >= 0]
    o = x[m[:, 0]]
    r1 = _mi(x[m[:, 1]] - o, L)
    r2 = _mi(x[m[:, 2]] - o, L)
    x[m[:, 3]] = o + 0.5 * (1.0 - gamma) * (r1 + r2)
    return x


def spread_m(mol, f, gamma):
    """forces with every M-site force moved onto O and the hydrogens"""
    f = np.array(f, np.float64, copy=True)
    m = mol[mol[:, 3] >= 0]
    fm = f[m[:, 3]].copy()
    f[m[:, 0]] += gamma * fm
    f[m[:, 1]] += 0.5 * (1.0 - gamma) * fm
    f[m[:, 2]] += 0.5 * (1.0 - gamma) * fm
    f[m[:, 3]] = 0.0
    return f


def bonded(mol, x, L, p=None):
    """forces (n, 3), energy and virial sum of (r_H - r_O) (x) F_H
    (xx, yy, zz, xy, xz, yz) of the quartic bonds and the harmonic angle"""
    p = p or QTIP4PF
    f = np.zeros_like(np.asarray(x, np.float64))
    o = x[mol[:, 0]]
    d = [_mi(x[mol[:, 1]] - o, L), _mi(x[mol[:, 2]] - o, L)]
    r = [np.linalg.norm(v, axis=1) for v in d]
    fh = []
    e = 0.0
    for b in range(2):
        dr = r[b] - p["r0"]
        e += float((p["k2"] * dr ** 2 + p["k3"] * dr ** 3 + p["k4"] * dr ** 4).sum())
        dU = 2 * p["k2"] * dr + 3 * p["k3"] * dr ** 2 + 4 * p["k4"] * dr ** 3
        fh.append(-(dU / r[b])[:, None] * d[b])
    u1, u2 = d[0] / r[0][:, None], d[1] / r[1][:, None]
    c = np.clip((u1 * u2).sum(1), -1.0, 1.0)
    th = np.arccos(c)
    dth = th - p["theta0"]
    e += float((0.5 * p["ktheta"] * dth ** 2).sum())
    pre = p["ktheta"] * dth / np.sqrt(1.0 - c * c)
    fh[0] = fh[0] + (pre / r[0])[:, None] * (u2 - c[:, None] * u1)
    fh[1] = fh[1] + (pre / r[1])[:, None] * (u1 - c[:, None] * u2)
    np.add.at(f, mol[:, 1], fh[0])
    np.add.at(f, mol[:, 2], fh[1])
    np.add.at(f, mol[:, 0], -fh[0] - fh[1])
    W = np.zeros(6)
    for b in range(2):
        v, F = d[b], fh[b]
        W += [(v[:, 0] * F[:, 0]).sum(), (v[:, 1] * F[:, 1]).sum(), (v[:, 2] * F[:, 2]).sum(),
              (v[:, 0] * F[:, 1]).sum(), (v[:, 0] * F[:, 2]).sum(), (v[:, 1] * F[:, 2]).sum()]
    return f, e, W
