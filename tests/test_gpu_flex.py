"""Flexible q-TIP4P/F water on the GPU (md_engine_amd/csrc/flex.hip) against
the FP64 restatement (oracle/flexwater.py) and the C oracle's pair and P3M
terms: M-site placement, the intramolecular forces, energy and virial, the
M-force spread, and energy conservation of NVE flexible-water dynamics."""
import sys

import numpy as np
import pytest

from md_engine_amd import systems
from oracle import flexwater as FW

from test_gpu_parity import oracle_water
from test_gpu_water_md import start_velocities

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def flex_engine(make, box, v, rc=7.0, skin=1.0, sz=(32, 32, 32), order=5, dt=0.5):
    from md_engine_amd import BONDED, PAIR, PPPM

    e = make()
    inv = np.where(box["mass"] > 0, 1.0 / np.maximum(box["mass"], 1e-30), 0.0)
    e.set_system(box["x"], box["lo"], box["L"], v=v, inv_mass=inv, q=box["q"], types=box["type"])
    e.exclusions_from_bonds(systems.water_bonds(box["mol"]))
    e.nlist_config(rc + skin, skin)
    eps24, sig6 = systems.water_lj_tables()
    e.lj_set(np.full((3, 3), rc * rc), eps24, sig6)
    e.pppm_set(sz, order, rc, systems.REAL_QQR_TO_ENG)
    e.ewald_real_set(True, -1.0, systems.REAL_QQR_TO_ENG, rc)
    p = FW.QTIP4PF
    e.flexible_water_set(box["mol"], p["k2"], p["k3"], p["k4"], p["r0"], p["ktheta"], p["theta0"], p["gamma"])
    e.integrator_set(dt, systems.REAL_FTM_TO_V)
    return e, PAIR | PPPM | BONDED


def distorted_box(n, seed):
    box = systems.tip4p2005_box(n, seed=seed, min_sep=1.8)
    rng = np.random.default_rng(seed)
    x = box["x"].copy()
    mol = box["mol"]
    for h in (1, 2):  # stretch and bend the O-H bonds a little
        x[mol[:, h]] += rng.normal(0, 0.03, (len(mol), 3))
    x = FW.place_m(mol, x, box["L"], FW.QTIP4PF["gamma"])
    return dict(box, x=x)


def test_flexible_forces_vs_oracle(orc, gpu_engine_factory):
    say("flexible water forces vs FP64")
    from md_engine_amd import BONDED, ENERGY, VIRIAL

    box = distorted_box(216, 21)
    e, flags = flex_engine(gpu_engine_factory, box, np.zeros_like(box["x"]))
    # intramolecular terms alone
    e.compute(BONDED | ENERGY | VIRIAL)
    fb = e.forces()
    _, vir = e.energies()
    fo, eo, wo = FW.bonded(box["mol"], box["x"], box["L"])
    assert np.abs(fb - fo).max() < 1e-6 * np.abs(fo).max()
    assert e.bonded_energy() == pytest.approx(eo, rel=1e-9)
    assert np.allclose(vir, wo, rtol=0, atol=1e-8 * np.abs(wo).max())
    # everything, with the M-site forces spread onto O and H
    e.compute(flags)
    f = e.forces()
    o = oracle_water(orc, box, rc=7.0, skin=1.0, sz=(32, 32, 32), order=5)
    ref = FW.spread_m(box["mol"], o["fp"] + o["fk"], FW.QTIP4PF["gamma"]) + fo
    rel = np.sqrt(((f - ref) ** 2).sum() / (ref ** 2).sum())
    say(f"relative force RMSE {rel:.2e}")
    assert rel < 1e-4
    assert np.all(f[box["mol"][:, 3]] == 0)


def test_m_sites_follow_the_molecule(orc, gpu_engine_factory):
    say("flexible water M placement during MD")
    box = distorted_box(216, 22)
    v = start_velocities(orc, box, seed=4)
    e, flags = flex_engine(gpu_engine_factory, box, v)
    e.run(50, flags)
    x, _ = e.get_system()
    placed = FW.place_m(box["mol"], x, box["L"], FW.QTIP4PF["gamma"])
    d = x[box["mol"][:, 3]] - placed[box["mol"][:, 3]]
    d -= box["L"] * np.rint(d / box["L"])
    assert np.abs(d).max() < 1e-9


@pytest.mark.parametrize("graph", [0, 1])
def test_flexible_nve_energy_error_is_second_order(orc, gpu_engine_factory, graph):
    """Velocity Verlet conserves a shadow energy: the deviation of the total
    energy over the same 60 fs falls about fourfold when dt halves."""
    say(f"flexible water NVE energy error vs dt (graph {graph})")
    from md_engine_amd import ENERGY

    box = distorted_box(216, 23)
    v = start_velocities(orc, box, seed=6)
    mvv = systems.REAL_MVV_TO_ENG
    errs = []
    for dt, nsteps in ((0.2, 300), (0.1, 600)):
        e, flags = flex_engine(gpu_engine_factory, box, v, dt=dt)
        e.option("graph", graph)

        def total():
            e.compute(flags | ENERGY)
            en, _ = e.energies()
            _, vv = e.get_system()
            return en.sum() + e.bonded_energy() + 0.5 * mvv * (box["mass"][:, None] * vv ** 2).sum()

        e0 = total()
        dev = 0.0
        for _ in range(12):
            e.run(nsteps // 12, flags)
            dev = max(dev, abs(total() - e0))
        errs.append(dev)
        e.close()
    say(f"max |E - E0|: dt 0.2 fs {errs[0]:.4f}, dt 0.1 fs {errs[1]:.4f} kcal/mol")
    assert errs[0] > 0
    assert errs[1] < 0.4 * errs[0]
