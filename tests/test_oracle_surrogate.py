"""Pins the FP64 surrogate restatement (oracle/surrogate.py): descriptors
against a plain loop, analytic forces against finite differences of the
energy, translation invariance, and the per-molecule E3B three-body energies
against the C oracle's total (itself pinned by brute-force enumeration)."""
import numpy as np
import pytest

from md_engine_amd import systems
from oracle import surrogate as S


@pytest.fixture(scope="module")
def small_box():
    return systems.tip4p2005_box(27, seed=3, min_sep=1.6)


def test_descriptors_match_loop(small_box):
    b = small_box
    a1 = S.descriptors(b["mol"], b["x"], b["L"])
    a2 = S.descriptors_loop(b["mol"], b["x"], b["L"])
    for u, v in zip(a1, a2):
        assert np.allclose(u, v, rtol=1e-12, atol=1e-14)
    assert a1[0].sum() > 0 and a1[1].sum() > 0


def test_forces_are_minus_gradient(small_box):
    b = small_box
    rng = np.random.default_rng(4)
    p = S.init_params(7)
    m = S.unpack(p)
    m["escale"] = 0.3
    m["mu"] = rng.normal(0, 0.1, S.NF)
    m["sd"] = rng.uniform(0.5, 2.0, S.NF)
    p = S.pack(m)
    x = b["x"].copy()
    f, _ = S.forces(p, b["mol"], x, b["L"])
    h = 1e-6
    movable = np.concatenate([b["mol"][:, 0], b["mol"][:, 1], b["mol"][:, 2]])
    for a in rng.choice(movable, 8, replace=False):
        for d in range(3):
            xp = x.copy()
            xm = x.copy()
            xp[a, d] += h
            xm[a, d] -= h
            fd = -(S.energy(p, b["mol"], xp, b["L"]) - S.energy(p, b["mol"], xm, b["L"])) / (2 * h)
            assert f[a, d] == pytest.approx(fd, rel=1e-6, abs=1e-8)
    assert np.allclose(f.sum(0), 0.0, atol=1e-10)
    assert np.all(f[b["mol"][:, 3]] == 0)


def test_energy_translation_invariant(small_box):
    b = small_box
    p = S.init_params(2)
    e0 = S.energy(p, b["mol"], b["x"], b["L"])
    e1 = S.energy(p, b["mol"], b["x"] + np.array([1.3, -7.1, 40.2]), b["L"])
    assert e1 == pytest.approx(e0, rel=1e-12)


def test_molecule_energies_sum_to_total(orc):
    box = systems.tip4p2005_box(64, seed=6, min_sep=1.6)
    em = S.e3b_molecule_energies(box["mol"], box["x"], box["L"])
    _, _, e3 = orc.e3b(box["mol"], box["x"], box["L"])
    assert abs(e3) > 0.1
    assert em.sum() == pytest.approx(e3, rel=1e-10)


def test_basis_five_is_the_e3b_bond_function(orc):
    P = orc.e3b_params()
    r = np.linspace(0.8, 5.3, 200)
    g, dg = S.basis(r)
    s = np.where(r <= P.rs, 1.0, np.where(r > P.rf, 0.0, (P.rf - r) ** 2 * (P.rf + 2 * r - 3 * P.rs) / (P.rf - P.rs) ** 3))
    assert np.allclose(g[:, 5], np.exp(-P.k3 * r) * s, rtol=1e-12, atol=0)
    h = 1e-6
    gp, _ = S.basis(r + h)
    gm, _ = S.basis(r - h)
    inner = (r > 0.9) & (np.abs(r - P.rs) > 1e-3) & (np.abs(r - P.rf) > 1e-3)
    assert np.allclose(dg[inner], ((gp - gm) / (2 * h))[inner], rtol=1e-6, atol=1e-12)
