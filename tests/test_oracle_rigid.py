"""Pins the oracle's rigid-water pieces: SETTLE against iterative SHAKE,
constraint and momentum invariants, RATTLE orthogonality, NVE energy
conservation of the FP64 water integrator (CPU only)."""
import numpy as np
import pytest

from md_engine_amd import systems

G = dict(mO=15.9994, mH=1.008, dOH=0.9572, dHH=1.5139)


def _rigid_molecule(rng):
    fr = systems.tip4p2005_frame()[:3]
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    R = systems._random_rotations(np.random.default_rng(int(rng.integers(1 << 30))), 1)[0]
    return fr @ R.T


@pytest.mark.parametrize("seed", range(5))
def test_settle_matches_shake(orc, seed):
    rng = np.random.default_rng(seed)
    old = _rigid_molecule(rng)
    B0, C0 = old[1] - old[0], old[2] - old[0]
    disp = rng.normal(scale=0.03, size=(3, 3))       # unconstrained drift dt*v
    A1, B1, C1 = disp[0], B0 + disp[1], C0 + disp[2]
    s = orc.settle_one(B0, C0, A1, B1, C1, **G)
    k = orc.shake_one(B0, C0, A1, B1, C1, **G)
    for a, b in zip(s, k):
        assert np.allclose(a, b, atol=1e-10, rtol=0)
    a3, b3, c3 = s
    assert np.linalg.norm(b3 - a3) == pytest.approx(G["dOH"], abs=1e-12)
    assert np.linalg.norm(c3 - a3) == pytest.approx(G["dOH"], abs=1e-12)
    assert np.linalg.norm(b3 - c3) == pytest.approx(G["dHH"], abs=1e-12)
    m = np.array([G["mO"], G["mH"], G["mH"]])
    com_before = (m[:, None] * np.array([A1, B1, C1])).sum(0)
    com_after = (m[:, None] * np.array(s)).sum(0)
    assert np.allclose(com_before, com_after, atol=1e-12)
    # corrections lie in the plane of the old molecule (constraint forces along old bonds)
    n0 = np.cross(B0, C0)
    for new, unc in zip(s, (A1, B1, C1)):
        assert abs(np.dot(new - unc, n0)) < 1e-12


def test_rattle_zeroes_bond_velocities(orc):
    rng = np.random.default_rng(3)
    r = _rigid_molecule(rng)
    v = rng.normal(scale=0.01, size=(3, 3))
    vo, v1, v2 = orc.rattle_one(r[0], r[1], r[2], v[0], v[1], v[2], G["mO"], G["mH"])
    w = np.array([vo, v1, v2])
    for i, j in ((0, 1), (0, 2), (1, 2)):
        assert abs(np.dot(w[i] - w[j], r[i] - r[j])) < 1e-14
    m = np.array([G["mO"], G["mH"], G["mH"]])
    assert np.allclose((m[:, None] * w).sum(0), (m[:, None] * v).sum(0), atol=1e-15)
    # idempotent
    again = orc.rattle_one(r[0], r[1], r[2], vo, v1, v2, G["mO"], G["mH"])
    assert np.allclose(np.array(again), w, atol=1e-15)


def test_water_md_conserves_energy(orc):
    box = systems.tip4p2005_box(64, seed=4, min_sep=1.8)
    v = systems.maxwell_boltzmann(box["mass"], 298.0, seed=5, boltz=systems.REAL_BOLTZ,
                                  mvv_to_eng=systems.REAL_MVV_TO_ENG)
    # start from constrained velocities
    for o, h1, h2, m in box["mol"]:
        v[o], v[h1], v[h2] = orc.rattle_one(box["x"][o], box["x"][h1], box["x"][h2], v[o], v[h1], v[h2],
                                            G["mO"], G["mH"])
    x, vf, et = orc.water_md(box, v, 60, dt=1.0, rc=5.5, skin=0.4, sz=(32, 32, 32), order=5,
                             e3b_on=True)
    # geometry kept
    for o, h1, h2, m in box["mol"]:
        d = x[[h1, h2]] - x[o]
        d -= box["L"] * np.rint(d / box["L"])
        assert np.allclose(np.linalg.norm(d, axis=1), G["dOH"], atol=1e-9)
        dm = x[m] - x[o]
        dm -= box["L"] * np.rint(dm / box["L"])
        assert np.linalg.norm(dm) == pytest.approx(0.1546, abs=1e-6)
    # total energy fluctuates far below the kinetic energy scale
    ke = orc.water_kinetic(box, v)
    assert (et.max() - et.min()) < 0.02 * ke
