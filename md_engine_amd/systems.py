"""Synthetic input systems for the force path (host-side numpy, no compute).

* ``fcc_lj``: single-type Lennard-Jones FCC lattice (the 32,000-atom LJ fluid
  at reduced density 0.8442 is ``fcc_lj(20, 0.8442)``).
* ``tip4p2005_box``: rigid TIP4P/2005 water on a cubic lattice with seeded
  random orientations, four sites per molecule in the order O, H1, H2, M
  (the order FixE3B::addMolecule requires, src/Fixes/FixE3B.cu:753-778).
* ``maxwell_boltzmann``: the InitializeAtoms::initTemp recipe
  (src/InitializeAtoms.cpp:99-150): N(0, sqrt(1/m)) per component, remove the
  centre-of-mass velocity, rescale to T with 3N degrees of freedom.

TIP4P/2005 geometry and charges are those of util_py/water.py:196-230
(q_H = 0.5564, q_M = -2 q_H, r_OM = 0.1546, r_OH = 0.9572,
angle HOH = 1.824218134 rad); LJ O-O parameters eps = 0.185207422 kcal/mol,
sigma = 3.1589 A (examples/e3b3_speed/e3b3_8000.py:133-147).
"""
from __future__ import annotations

import numpy as np

# units (src/Units.cpp:5-32)
REAL_BOLTZ = 0.0019872067
REAL_HBAR = 15.178728314015  # kcal/mol fs (Units.cpp:22)
REAL_MVV_TO_ENG = 48.88821291 * 48.88821291
REAL_QQR_TO_ENG = 332.06371
REAL_FTM_TO_V = 1.0 / (48.88821291 * 48.88821291)

TIP4P2005 = dict(
    q_H=0.5564, q_M=-2.0 * 0.5564, r_OM=0.1546, r_OH=0.9572, theta=1.824218134,
    eps_OO=0.185207422, sig_OO=3.1589, m_O=15.9994, m_H=1.008, m_M=0.0,
    # rigid geometry kept by SETTLE and the M-site weight, as FixRigid sets
    # them for TIP4P/2005 (src/Fixes/FixRigid.cu:1372-1379)
    r_HH=1.5139, gamma=0.73612446364836,
)

WATER_DENSITY_G_CM3 = 0.997
AVOGADRO = 6.02214076e23
WATER_MOLAR_MASS = 15.9994 + 2 * 1.008


def water_box_length(nmol, density=WATER_DENSITY_G_CM3):
    """Cubic box edge (A) holding nmol molecules at the given mass density."""
    vol_cm3 = nmol * WATER_MOLAR_MASS / AVOGADRO / density
    return (vol_cm3 * 1e24) ** (1.0 / 3.0)


def fcc_lj(ncell, density):
    """FCC lattice of 4 * ncell^3 atoms at reduced number density `density`.

    Returns (x (N,3) float64, L (3,), lo (3,))."""
    a = (4.0 / density) ** (1.0 / 3.0)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.arange(ncell)
    cells = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    x = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3) * a
    L = np.full(3, ncell * a)
    lo = np.zeros(3)
    return x + 0.25 * a, L, lo


def maxwell_boltzmann(masses, temp, seed, boltz=1.0, mvv_to_eng=1.0):
    """Velocities by the initTemp recipe (InitializeAtoms.cpp:99-150).

    Massless sites (m = 0) get zero velocity here; the reference draws them
    from N(0, 1) and they carry no kinetic energy once constrained."""
    masses = np.asarray(masses, dtype=np.float64)
    rng = np.random.default_rng(seed)
    v = np.zeros((len(masses), 3))
    heavy = masses > 0
    v[heavy] = rng.normal(size=(heavy.sum(), 3)) * np.sqrt(1.0 / masses[heavy])[:, None]
    mtot = masses[heavy].sum()
    v[heavy] -= (masses[heavy, None] * v[heavy]).sum(0) / mtot
    ke2 = (masses[heavy, None] * v[heavy] ** 2).sum()
    cur = (mvv_to_eng / boltz) * ke2 / (3 * heavy.sum())
    v[heavy] *= np.sqrt(temp / cur)
    return v


def _random_rotations(rng, n):
    """Uniform random rotation matrices from unit quaternions."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def tip4p2005_frame():
    """Body-frame site offsets (O, H1, H2, M) relative to O: |OH| = r_OH,
    |HH| = r_HH exactly (the constraint SETTLE keeps), M = O + (1 - gamma)/2
    (r_OH1 + r_OH2), which puts M 0.1546 A along the bisector."""
    p = TIP4P2005
    half = 0.5 * p["r_HH"]
    mid = np.sqrt(p["r_OH"] ** 2 - half ** 2)
    h1 = np.array([half, mid, 0.0])
    h2 = np.array([-half, mid, 0.0])
    m = 0.5 * (1.0 - p["gamma"]) * (h1 + h2)
    return np.stack([np.zeros(3), h1, h2, m])


def tip4p2005_box(nmol, L=None, seed=20261202, min_sep=1.5):
    """nmol rigid TIP4P/2005 molecules in a cubic box of edge L (default: the
    edge giving 0.997 g/cm^3), oxygens on the smallest cubic lattice that holds
    them (sites chosen with the seed when the lattice has spare sites),
    orientations uniformly random from the seed.  Orientations are redrawn
    (up to 200 sweeps) until no two sites of different molecules are closer
    than `min_sep` A.

    Returns dict with x (4*nmol, 3), q, type (0 = O, 1 = H, 2 = M), mass,
    mol (nmol, 4) site indices, L (3,), lo (3,)."""
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    if L is None:
        L = water_box_length(nmol)
    m = int(np.ceil(nmol ** (1.0 / 3.0) - 1e-9))
    g = np.arange(m)
    sites = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    if len(sites) > nmol:
        sites = sites[np.sort(rng.choice(len(sites), nmol, replace=False))]
    O = (sites + 0.5) * (L / m)
    frame = tip4p2005_frame()
    R = _random_rotations(rng, nmol)
    box = np.full(3, L)

    def build(R):
        return (O[:, None, :] + np.einsum("mij,sj->msi", R, frame)) % L

    x = build(R)
    for _ in range(200):
        flat = x.reshape(-1, 3)
        tree = cKDTree(flat, boxsize=box)
        pairs = tree.query_pairs(min_sep, output_type="ndarray")
        if len(pairs):
            pairs = pairs[(pairs[:, 0] // 4) != (pairs[:, 1] // 4)]
        if len(pairs) == 0:
            break
        bad = np.unique(pairs[:, 0] // 4)
        R[bad] = _random_rotations(rng, len(bad))
        x = build(R)
    x = x.reshape(-1, 3)
    p = TIP4P2005
    q = np.tile([0.0, p["q_H"], p["q_H"], p["q_M"]], nmol)
    typ = np.tile([0, 1, 1, 2], nmol).astype(np.int32)
    mass = np.tile([p["m_O"], p["m_H"], p["m_H"], p["m_M"]], nmol)
    mol = np.arange(4 * nmol, dtype=np.int32).reshape(nmol, 4)
    return dict(x=x, q=q, type=typ, mass=mass, mol=mol, L=box, lo=np.zeros(3))


def water_lj_tables():
    """Type-pair (O, H, M) LJ tables (rc2 filled by the caller): eps24, sig6.

    Only O-O interacts; the reference sets every other eps and sigma to 0
    (examples/e3b3_speed/e3b3_8000.py:139-147) and geometric mixing
    (FixLJCut.cu:85-116) keeps them 0."""
    p = TIP4P2005
    eps = np.zeros((3, 3))
    sig = np.zeros((3, 3))
    eps[0, 0] = p["eps_OO"]
    sig[0, 0] = p["sig_OO"]
    return 24.0 * eps, sig ** 6


def water_bonds(mol):
    """Constraint bonds FixRigid registers for exclusions (FixRigid.cu:1631-1681):
    O-H1, O-H2, H1-H2, O-M."""
    b = []
    for o, h1, h2, m in np.asarray(mol):
        b += [(o, h1), (o, h2), (h1, h2), (o, m)]
    return b


# q-TIP4P/F flexible water (mdx_flexible_water_set): the quartic O-H bond and
# harmonic angle constants of the reference's example (examples/PIMD_TIP4P.py:
# FixBondQuartic k2, k3, k4, r0; FixAngleHarmonic k, theta0) and the M-site
# weight FixTIP4PFlexible uses for q-TIP4P/F (FixTIP4PFlexible.cu:325-330).
# Charges, LJ and masses equal TIP4P/2005's.
QTIP4PF = dict(k2=607.19, k3=-1388.65, k4=1852.58, r0=0.9419, ktheta=87.85, theta0=107.4 / 180.0 * np.pi,
               gamma=0.73612)


def ring_modes(P):
    """orthonormal normal-mode matrix C[n, k] of a P-bead ring polymer
    (nve_xPIMD_cu, IntegratorVerlet.cu:63-304; beads n = 1..P)"""
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


def ring_beads(x, mass, P, temp_beads, seed, boltz=REAL_BOLTZ, hbar=REAL_HBAR, mvv_to_eng=REAL_MVV_TO_ENG):
    """P bead copies of positions x with the free ring polymer's normal modes
    sampled at the bead temperature and the centroid at x, as
    State::preparePIMD does (src/State.cpp:1066-1130): mode 0 = sqrt(P) x,
    mode k ~ N(0, sqrt(kB T / m) / w_k), w_k = 2 (kB T / hbar) sin(k pi / P).
    Massless sites stay on the centroid."""
    rng = np.random.default_rng(seed)
    n = len(x)
    wp = boltz * temp_beads / hbar
    xt = np.zeros((P, n, 3))
    xt[0] = np.sqrt(P) * x
    live = mass > 0
    for k in range(1, P):
        wk = 2.0 * wp * np.sin(k * np.pi / P)
        sig = np.zeros(n)
        sig[live] = np.sqrt(boltz * temp_beads / (mass[live] * mvv_to_eng)) / wk
        xt[k] = rng.normal(size=(n, 3)) * sig[:, None]
    return list(np.einsum("nk,kad->nad", ring_modes(P), xt))
