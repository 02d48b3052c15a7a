"""FP64 CPU restatement of the learned E3B surrogate (test infrastructure).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
surrogate kernels; the product path never imports it.

The surrogate has no counterpart in the reference: it is a learned
replacement for the E3B three-body energy of FixE3B (src/Fixes/FixE3B.cu,
src/Evaluators/ThreeBodyE3B.h).  Its definition, shared with the HIP kernels:

* raw channels per molecule i (3 NB + 3): for NB radial functions
  g_k(r) = exp(-KAPPA_k r) fc(r), KAPPA_k = k3 (1 - 0.1 (KE - k)) for
  k <= KE = NB - 2 (basis KE is the E3B hydrogen-bond function f itself) and
  KAPPA = 2 k3 for basis K2 = NB - 1 (f^2 up to the switch), fc the E3B
  switching polynomial between RS_C = 5.0 and RF_C = 5.2,
    acc_k = sum_{j != i} sum_{H of j} g_k(|H - O_i|)   (hydrogens donated to O_i)
    d1_k  = sum_{j != i} g_k(|H1_i - O_j|),  d2_k likewise for H2_i,
  and three pair-product channels with f = g_KE over the same neighbours j,
    pAA = sum_j f(H1_i - O_j) f(H2_i - O_j)      (both my H towards one O_j)
    pBB = sum_j f(H1_j - O_i) f(H2_j - O_i)      (both H of j towards my O)
    pAB = sum_j D_ij D_ji,  D_ij = f(H1_i - O_j) + f(H2_i - O_j);
* network input F (3 NB + 6) = [acc, d1 + d2, d1 * d2, pAA, pBB, pAB,
  acc_KE (d1 + d2)_KE, acc_KE^2, (d1 + d2)_KE^2] (symmetric in H1 <-> H2);
* x = (F - mu) / sd; h1 = silu(x W1 + b1); h2 = silu(h1 W2 + b2);
  e = h2 . w3 + b3; E_i = escale e + eshift; E = sum_i E_i;
* forces = -dE/dpositions through the raw channels (chain rule below).

The pair-product channels and the products of the KE channels are the
three-body correlations the E3B energy is built from (its per-molecule
energy is linear in F), so the network can reach the E3B forces closely.

Training loss (the GPU's ``k_sur_grad``): mean over a mini-batch of the
squared difference between the network output e and the standardised
target (t - eshift) / escale; ``loss_grad`` gives it with its gradient in the
trained parameters (W1 b1 W2 b2 w3 b3, packed order).

Pinned by: finite differences of ``energy`` against ``forces`` (FP64);
finite differences of ``loss_grad``'s loss against its gradient along random
directions;
``descriptors`` against a per-pair loop (``descriptors_loop``); the per-molecule
three-body energies ``e3b_molecule_energies`` sum to the C oracle's E3B
three-body energy, which is itself pinned by brute-force enumeration.
"""
from __future__ import annotations

import numpy as np

NB = 2
KE = NB - 2    # basis with kappa = k3 (the E3B function)
K2 = NB - 1    # basis with kappa = 2 k3
NR = 3 * NB + 3
NF = 3 * NB + 6
NH = 64
KAPPA = 1.907 * np.array([1.0 - 0.1 * (KE - k) for k in range(K2)] + [2.0])
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


def _pair_channels(mol, x, L, pr):
    """g (n, NB, 4) of the four H-O distances of each directed pair (i, j):
    H1_i - O_j, H2_i - O_j, H1_j - O_i, H2_j - O_i; and the distance vectors"""
    i, j = pr[:, 0], pr[:, 1]
    dr = [_mi(x[mol[i, 1]] - x[mol[j, 0]], L), _mi(x[mol[i, 2]] - x[mol[j, 0]], L),
          _mi(x[mol[j, 1]] - x[mol[i, 0]], L), _mi(x[mol[j, 2]] - x[mol[i, 0]], L)]
    return dr


def descriptors(mol, x, L):
    """raw per-molecule channels (M, NR): acc, d1, d2 (NB each), pAA, pBB, pAB"""
    M = len(mol)
    pr = _pairs(mol, x, L)
    i = pr[:, 0]
    dr = _pair_channels(mol, x, L, pr)
    g = [basis(np.linalg.norm(d, axis=1))[0] for d in dr]
    raw = np.zeros((M, NR))
    np.add.at(raw[:, NB:2 * NB], i, g[0])
    np.add.at(raw[:, 2 * NB:3 * NB], i, g[1])
    np.add.at(raw[:, :NB], i, g[2] + g[3])
    a1, a2, b1, b2 = (gq[:, KE] for gq in g)
    np.add.at(raw[:, 3 * NB], i, a1 * a2)
    np.add.at(raw[:, 3 * NB + 1], i, b1 * b2)
    np.add.at(raw[:, 3 * NB + 2], i, (a1 + a2) * (b1 + b2))
    return raw


def descriptors_loop(mol, x, L):
    """the same channels by a plain double loop (small systems)"""
    M = len(mol)
    raw = np.zeros((M, NR))

    def gk(a, b):
        r = np.linalg.norm(_mi(a - b, L))
        return basis(np.array([r]))[0][0] if r < RF_C else np.zeros(NB)

    for i in range(M):
        for j in range(M):
            if i == j:
                continue
            if np.linalg.norm(_mi(x[mol[i, 0]] - x[mol[j, 0]], L)) >= 6.2:
                continue
            a1, a2 = gk(x[mol[i, 1]], x[mol[j, 0]]), gk(x[mol[i, 2]], x[mol[j, 0]])
            b1, b2 = gk(x[mol[j, 1]], x[mol[i, 0]]), gk(x[mol[j, 2]], x[mol[i, 0]])
            raw[i, NB:2 * NB] += a1
            raw[i, 2 * NB:3 * NB] += a2
            raw[i, :NB] += b1 + b2
            raw[i, 3 * NB] += a1[KE] * a2[KE]
            raw[i, 3 * NB + 1] += b1[KE] * b2[KE]
            raw[i, 3 * NB + 2] += (a1[KE] + a2[KE]) * (b1[KE] + b2[KE])
    return raw


def features(raw):
    acc, d1, d2 = raw[:, :NB], raw[:, NB:2 * NB], raw[:, 2 * NB:3 * NB]
    s = d1 + d2
    x5, s5 = acc[:, KE], s[:, KE]
    return np.concatenate([acc, s, d1 * d2, raw[:, 3 * NB:], (x5 * s5)[:, None], (x5 * x5)[:, None],
                           (s5 * s5)[:, None]], axis=1)


def raw_weights(dF, raw):
    """dE/d(raw channel) (M, NR) from dE/dF (M, NF)"""
    acc, d1, d2 = raw[:, :NB], raw[:, NB:2 * NB], raw[:, 2 * NB:3 * NB]
    x5, s5 = acc[:, KE], d1[:, KE] + d2[:, KE]
    w = np.zeros_like(raw)
    w[:, :NB] = dF[:, :NB]
    w[:, NB:2 * NB] = dF[:, NB:2 * NB] + dF[:, 2 * NB:3 * NB] * d2
    w[:, 2 * NB:3 * NB] = dF[:, NB:2 * NB] + dF[:, 2 * NB:3 * NB] * d1
    w[:, 3 * NB:] = dF[:, 3 * NB:3 * NB + 3]
    gx = dF[:, 3 * NB + 3] * s5 + 2.0 * dF[:, 3 * NB + 4] * x5
    gs = dF[:, 3 * NB + 3] * x5 + 2.0 * dF[:, 3 * NB + 5] * s5
    w[:, KE] += gx
    w[:, NB + KE] += gs
    w[:, 2 * NB + KE] += gs
    return w


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


NTRAIN = NF * NH + NH + NH * NH + NH + NH + 1


def loss_grad(p, raw, target):
    """Energy-regression loss of one mini-batch (raw channels (B, NR), E3B
    per-molecule three-body energies (B,)) and its gradient (NTRAIN,) in the
    packed order W1, b1, W2, b2, w3, b3."""
    m = unpack(p)
    B = len(raw)
    xh = (features(raw) - m["mu"]) / m["sd"]
    z1 = xh @ m["W1"] + m["b1"]
    h1 = silu(z1)
    z2 = h1 @ m["W2"] + m["b2"]
    h2 = silu(z2)
    e = h2 @ m["w3"] + m["b3"]
    t = (np.asarray(target, np.float64) - m["eshift"]) / m["escale"]
    diff = e - t
    loss = float((diff * diff).sum() / B)
    g = 2.0 * diff / B                      # dL/de
    gw3 = h2.T @ g
    gb3 = g.sum()
    dz2 = (g[:, None] * m["w3"][None, :]) * dsilu(z2)
    gW2 = h1.T @ dz2
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ m["W2"].T) * dsilu(z1)
    gW1 = xh.T @ dz1
    gb1 = dz1.sum(0)
    return loss, np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])


def energy(p, mol, x, L):
    E, _ = mlp(p, features(descriptors(mol, x, L)))
    return E.sum()


def forces(p, mol, x, L, virial=False):
    """-dE/dx (n, 3) for the site positions x (M sites unaffected); with
    virial=True also the virial sum over H-O bonds of d (x) F_H (d = r_H - r_O),
    as (xx, yy, zz, xy, xz, yz).

    Every H-O bond H_h(i) - O_j belongs to one directed pair (i, j); its
    value a_h = f feeds i's donor channel h, j's acceptor channel, the
    products a1 a2 (pAA of i, pBB of j) and D_ij D_ji (pAB of i and of j)."""
    raw = descriptors(mol, x, L)
    E, dF = mlp(p, features(raw))
    w = raw_weights(dF, raw)
    wacc, wd = w[:, :NB], (w[:, NB:2 * NB], w[:, 2 * NB:3 * NB])
    wAA, wBB, wAB = w[:, 3 * NB], w[:, 3 * NB + 1], w[:, 3 * NB + 2]
    f = np.zeros_like(x)
    pr = _pairs(mol, x, L)
    i, j = pr[:, 0], pr[:, 1]
    dr = _pair_channels(mol, x, L, pr)
    r = [np.linalg.norm(d, axis=1) for d in dr]
    gg = [basis(rq) for rq in r]
    a = [gg[q][0][:, KE] for q in range(4)]
    W = np.zeros(6)
    for h in (0, 1):
        g, dg = gg[h]
        c = ((wd[h][i] + wacc[j]) * dg).sum(1)
        c += (wAA[i] + wBB[j]) * a[1 - h] * dg[:, KE]
        c += (wAB[i] + wAB[j]) * (a[2] + a[3]) * dg[:, KE]
        fv = -(c / np.where(r[h] > 0, r[h], 1.0))[:, None] * dr[h]
        np.add.at(f, mol[i, h + 1], fv)
        np.add.at(f, mol[j, 0], -fv)
        d = dr[h]
        W += [(d[:, 0] * fv[:, 0]).sum(), (d[:, 1] * fv[:, 1]).sum(), (d[:, 2] * fv[:, 2]).sum(),
              (d[:, 0] * fv[:, 1]).sum(), (d[:, 0] * fv[:, 2]).sum(), (d[:, 1] * fv[:, 2]).sum()]
    if virial:
        return f, E, W
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
