"""FP64 CPU restatement of the learned E3B surrogate (test infrastructure).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
surrogate kernels; the product path never imports it.

The surrogate has no counterpart in the reference: it is a learned
replacement for the E3B three-body energy of FixE3B (src/Fixes/FixE3B.cu,
src/Evaluators/ThreeBodyE3B.h).  Its definition, shared with the HIP kernels:

* descriptors per molecule i (48): for 16 radial functions
  g_k(r) = exp(-KAPPA_k r) fc(r), KAPPA_k = k3 (0.5 + 0.1 k) (basis 5 is the
  E3B hydrogen-bond function itself), fc the E3B switching polynomial between
  RS_C = 5.0 and RF_C = 5.2,
    acc_k = sum_{j != i} sum_{H of j} g_k(|H - O_i|)   (hydrogens donated to O_i)
    d1_k  = sum_{j != i} g_k(|H1_i - O_j|),  d2_k likewise for H2_i,
  features F = [acc, d1 + d2, d1 * d2] (symmetric in H1 <-> H2);
* x = (F - mu) / sd; h1 = silu(x W1 + b1); h2 = silu(h1 W2 + b2);
  e = h2 . w3 + b3; E_i = escale e + eshift; E = sum_i E_i;
* forces = -dE/dpositions through the descriptors (chain rule below).

Pinned by: finite differences of ``energy`` against ``forces`` (FP64);
``descriptors`` against a per-pair loop (``descriptors_loop``); the per-molecule
three-body energies ``e3b_molecule_energies`` sum to the C oracle's E3B
three-body energy, which is itself pinned by brute-force enumeration.
"""
from __future__ import annotations

import numpy as np

NB = 16
NF = 3 * NB
NH = 64
KAPPA = 1.907 * (0.5 + 0.1 * np.arange(NB))
RS_C, RF_C = 5.0, 5.2
NPARAM = NF * NH + NH + NH * NH + NH + NH + 1 + NF + NF + 2


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0

    def take(n, shape=None):
        nonlocal o
        a = p[o:o + n]
        o += n
        return a.reshape(shape) if shape else a

    W1 = take(NF * NH, (NF, NH))
    b1 = take(NH)
    W2 = take(NH * NH, (NH, NH))
    b2 = take(NH)
    w3 = take(NH)
    b3 = take(1)[0]
    mu = take(NF)
    sd = take(NF)
    escale, eshift = take(2)
    return dict(W1=W1, b1=b1, W2=W2, b2=b2, w3=w3, b3=b3, mu=mu, sd=sd, escale=escale, eshift=eshift)


def pack(m):
    return np.concatenate([m["W1"].ravel(), m["b1"], m["W2"].ravel(), m["b2"], m["w3"], [m["b3"]],
                           m["mu"], m["sd"], [m["escale"], m["eshift"]]])


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    m = dict(W1=rng.normal(0, 1 / np.sqrt(NF), (NF, NH)), b1=np.zeros(NH),
             W2=rng.normal(0, 1 / np.sqrt(NH), (NH, NH)), b2=np.zeros(NH),
             w3=rng.normal(0, 1 / np.sqrt(NH), NH), b3=0.0, mu=np.zeros(NF), sd=np.ones(NF),
             escale=1.0, eshift=0.0)
    return pack(m)


def fc(r):
    """switching polynomial (EvaluatorE3B.h:91-117 form) and its derivative"""
    s = np.where(r <= RS_C, 1.0, np.where(r >= RF_C, 0.0,
                 (RF_C - r) ** 2 * (RF_C + 2 * r - 3 * RS_C) / (RF_C - RS_C) ** 3))
    ds = np.where((r > RS_C) & (r < RF_C), 6 * (RF_C - r) * (RS_C - r) / (RF_C - RS_C) ** 3, 0.0)
    return s, ds


def basis(r):
    """g (n, NB) and dg/dr (n, NB) at distances r (n,)"""
    s, ds = fc(r)
    ex = np.exp(-KAPPA[None, :] * r[:, None])
    g = ex * s[:, None]
    dg = ex * (ds[:, None] - KAPPA[None, :] * s[:, None])
    return g, dg


def _pairs(mol, x, L):
    """(i, j) molecule pairs i != j with O-O within 6.2 A (full list)."""
    from scipy.spatial import cKDTree

    O = x[mol[:, 0]] % L
    t = cKDTree(O, boxsize=L)
    p = t.query_pairs(6.2, output_type="ndarray")
    return np.concatenate([p, p[:, ::-1]]) if len(p) else np.zeros((0, 2), int)


def _mi(d, L):
    return d - L * np.rint(d / L)


def descriptors(mol, x, L):
    """raw per-molecule channels acc, d1, d2 (M, NB each)"""
    M = len(mol)
    pr = _pairs(mol, x, L)
    i, j = pr[:, 0], pr[:, 1]
    acc = np.zeros((M, NB))
    d1 = np.zeros((M, NB))
    d2 = np.zeros((M, NB))
    for h, dh in ((1, d1), (2, d2)):
        dr = _mi(x[mol[i, h]] - x[mol[j, 0]], L)      # H_h(i) -> O(j)
        g, _ = basis(np.linalg.norm(dr, axis=1))
        np.add.at(dh, i, g)
        np.add.at(acc, j, g)
    return acc, d1, d2


def descriptors_loop(mol, x, L):
    """the same channels by a plain double loop (small systems)"""
    M = len(mol)
    acc = np.zeros((M, NB))
    d1 = np.zeros((M, NB))
    d2 = np.zeros((M, NB))
    for i in range(M):
        for j in range(M):
            if i == j:
                continue
            for h, dh in ((1, d1), (2, d2)):
                r = np.linalg.norm(_mi(x[mol[i, h]] - x[mol[j, 0]], L))
                if r < RF_C:
                    g, _ = basis(np.array([r]))
                    dh[i] += g[0]
                    acc[j] += g[0]
    return acc, d1, d2


def features(acc, d1, d2):
    return np.concatenate([acc, d1 + d2, d1 * d2], axis=1)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def mlp(p, F):
    """per-molecule energies (M,) and dE_i/dF_i (M, NF)"""
    m = unpack(p)
    xh = (F - m["mu"]) / m["sd"]
    z1 = xh @ m["W1"] + m["b1"]
    h1 = silu(z1)
    z2 = h1 @ m["W2"] + m["b2"]
    h2 = silu(z2)
    e = h2 @ m["w3"] + m["b3"]
    E = m["escale"] * e + m["eshift"]
    dz2 = m["w3"][None, :] * dsilu(z2)
    dz1 = (dz2 @ m["W2"].T) * dsilu(z1)
    dF = m["escale"] * (dz1 @ m["W1"].T) / m["sd"]
    return E, dF


def energy(p, mol, x, L):
    acc, d1, d2 = descriptors(mol, x, L)
    E, _ = mlp(p, features(acc, d1, d2))
    return E.sum()


def forces(p, mol, x, L):
    """-dE/dx (n, 3) for the site positions x (M sites unaffected)."""
    acc, d1, d2 = descriptors(mol, x, L)
    E, dF = mlp(p, features(acc, d1, d2))
    wa = dF[:, :NB]
    w1 = dF[:, NB:2 * NB] + dF[:, 2 * NB:] * d2
    w2 = dF[:, NB:2 * NB] + dF[:, 2 * NB:] * d1
    f = np.zeros_like(x)
    pr = _pairs(mol, x, L)
    i, j = pr[:, 0], pr[:, 1]
    for h, wh in ((1, w1), (2, w2)):
        dr = _mi(x[mol[i, h]] - x[mol[j, 0]], L)
        r = np.linalg.norm(dr, axis=1)
        _, dg = basis(r)
        c = ((wh[i] + wa[j]) * dg).sum(1)             # dE/dr of the pair H_h(i) - O(j)
        fv = -(c / np.where(r > 0, r, 1.0))[:, None] * dr
        np.add.at(f, mol[i, h], fv)
        np.add.at(f, mol[j, 0], -fv)
    return f, E


def e3b_molecule_energies(mol, x, L, params=None):
    """E3B three-body energy owned by each molecule (donor for type A,
    centre for B, acceptor for C), the decomposition the GPU E3B kernel
    uses; sums to the total three-body energy."""
    from . import e3b_params

    P = params or e3b_params()
    Ea, Eb, Ec, k3, rs, rf, rmax = P.Ea, P.Eb, P.Ec, P.k3, P.rs, P.rf, P.rmax
    M = len(mol)
    O = x[mol[:, 0]] % L
    from scipy.spatial import cKDTree

    pp = cKDTree(O, boxsize=L).query_pairs(rmax, output_type="ndarray")
    pr = np.concatenate([pp, pp[:, ::-1]]) if len(pp) else np.zeros((0, 2), int)
    i, j = pr[:, 0], pr[:, 1]

    def f(dr):
        r = np.linalg.norm(dr, axis=1)
        s = np.where(r <= rs, 1.0, np.where(r > rf, 0.0, (rf - r) ** 2 * (rf + 2 * r - 3 * rs) / (rf - rs) ** 3))
        return np.exp(-k3 * r) * s

    fA1 = f(_mi(x[mol[i, 1]] - x[mol[j, 0]], L))    # H1_i -> O_j
    fA2 = f(_mi(x[mol[i, 2]] - x[mol[j, 0]], L))
    Dij = fA1 + fA2                                  # i donates to j
    xt = np.zeros(M)
    y = np.zeros(M)
    z = np.zeros(M)
    np.add.at(xt, j, Dij)
    np.add.at(y, i, fA1)
    np.add.at(z, i, fA2)
    # D(j -> i) for each directed pair (i, j): look up the reverse pair
    key = {(a, b): k for k, (a, b) in enumerate(pr)}
    rev = np.array([key[(b, a)] for a, b in pr], int) if len(pr) else np.zeros(0, int)
    Dji = Dij[rev] if len(pr) else Dij
    ea = np.zeros(M)
    eb = np.zeros(M)
    ec = np.zeros(M)
    np.add.at(ea, i, fA1 * fA2)
    np.add.at(eb, i, Dij * Dji)
    np.add.at(ec, i, Dji * Dji)
    return Ea * (y * z - ea) + Eb * (xt * (y + z) - eb) + 0.5 * Ec * (xt * xt - ec)
