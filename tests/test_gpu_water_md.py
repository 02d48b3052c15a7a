"""Rigid-water MD on the GPU (FP64 SETTLE / RATTLE, M site) against the FP64
oracle integrator: one step, a short trajectory, constraints after many
steps, and NVE energy conservation."""
import sys

import numpy as np
import pytest

from md_engine_amd import systems

pytestmark = pytest.mark.gpu

G = dict(mO=15.9994, mH=1.008, dOH=0.9572, dHH=1.5139, gamma=0.73612446364836)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def start_velocities(orc, box, seed=5):
    v = systems.maxwell_boltzmann(box["mass"], 298.0, seed=seed, boltz=systems.REAL_BOLTZ,
                                  mvv_to_eng=systems.REAL_MVV_TO_ENG)
    for o, h1, h2, m in box["mol"]:
        v[o], v[h1], v[h2] = orc.rattle_one(box["x"][o], box["x"][h1], box["x"][h2], v[o], v[h1], v[h2],
                                            G["mO"], G["mH"])
    return v


def water_md_engine(make, box, v, rc, skin, sz, order, e3b, dt):
    from md_engine_amd import E3B, PAIR, PPPM

    e = make()
    inv = np.where(box["mass"] > 0, 1.0 / np.maximum(box["mass"], 1e-30), 0.0)
    e.set_system(box["x"], box["lo"], box["L"], v=v, inv_mass=inv, q=box["q"], types=box["type"])
    e.exclusions_from_bonds(systems.water_bonds(box["mol"]))
    e.nlist_config(rc + skin, skin)
    eps24, sig6 = systems.water_lj_tables()
    e.lj_set(np.full((3, 3), rc * rc), eps24, sig6)
    flags = PAIR
    if sz is not None:
        e.pppm_set(sz, order, rc, systems.REAL_QQR_TO_ENG)
        e.ewald_real_set(True, -1.0, systems.REAL_QQR_TO_ENG, rc)
        flags |= PPPM
    else:
        # no mesh: Ewald real space alone at a fixed splitting
        e.ewald_real_set(True, 0.3, systems.REAL_QQR_TO_ENG, rc)
    if e3b:
        k = 0.23900573614
        e.e3b_set(box["mol"], [453000 * k, 150 * k, -1005 * k, 1880 * k, 4.872, 1.907, 5.0, 5.2])
        flags |= E3B
    e.rigid_set(box["mol"], G["mO"], G["mH"], G["dOH"], G["dHH"], G["gamma"])
    e.integrator_set(dt, systems.REAL_FTM_TO_V)
    return e, flags


def min_image_diff(a, b, L):
    d = a - b
    return d - L * np.rint(d / L)


@pytest.mark.parametrize("nsteps", [1, 15])
def test_water_trajectory_matches_oracle(orc, gpu_engine_factory, nsteps):
    say(f"water trajectory {nsteps} steps vs oracle")
    box = systems.tip4p2005_box(216, seed=21, min_sep=1.8)
    v = start_velocities(orc, box)
    kw = dict(rc=7.0, skin=1.0, sz=(32, 32, 32), order=5)
    e, flags = water_md_engine(gpu_engine_factory, box, v, e3b=True, dt=1.0, **kw)
    e.run(nsteps, flags)
    xg, vg = e.get_system()
    xo, vo, _ = orc.water_md(box, v, nsteps, dt=1.0, e3b_on=True, **kw)
    dx = np.abs(min_image_diff(xg, xo, box["L"])).max()
    assert dx < (1e-8 if nsteps == 1 else 1e-6), dx
    heavy = box["mass"] > 0
    rel_v = np.abs(vg[heavy] - vo[heavy]).max() / np.abs(vo[heavy]).max()
    assert rel_v < (1e-6 if nsteps == 1 else 1e-4), rel_v


def test_water_constraints_and_energy(orc, gpu_engine_factory):
    say("water NVE 512 molecules, 1500 steps")
    from md_engine_amd import ENERGY

    box = systems.tip4p2005_box(512, seed=22, min_sep=1.8)
    v = start_velocities(orc, box, seed=9)
    e, flags = water_md_engine(gpu_engine_factory, box, v, rc=9.0, skin=1.0, sz=(32, 32, 32), order=5,
                               e3b=True, dt=1.0)

    def total():
        e.compute(flags | ENERGY)
        en, _ = e.energies()
        _, vv = e.get_system()
        return en.sum() + orc.water_kinetic(box, vv)

    e.run(100, flags)          # let the lattice start relax a little
    e0 = total()
    energies = []
    for _ in range(7):
        e.run(200, flags)
        energies.append(total())
    st = e.nlist_stats()
    assert st["violations"] == 0
    assert st["rebuilds"] > 5
    x, vv = e.get_system()
    for o, h1, h2, m in box["mol"]:
        d = min_image_diff(x[[h1, h2]], x[o], box["L"])
        assert np.allclose(np.linalg.norm(d, axis=1), G["dOH"], atol=1e-9)
        assert np.linalg.norm(min_image_diff(x[h1], x[h2], box["L"])) == pytest.approx(G["dHH"], abs=1e-9)
        dm = min_image_diff(x[m], x[o], box["L"])
        assert np.linalg.norm(dm) == pytest.approx(0.1546, abs=1e-6)
        for i, j in ((o, h1), (o, h2), (h1, h2)):
            r = min_image_diff(x[i], x[j], box["L"])
            assert abs(np.dot(vv[i] - vv[j], r)) < 1e-9
    ke = orc.water_kinetic(box, vv)
    drift = abs(energies[-1] - e0)
    assert drift < 0.01 * ke, (e0, energies, ke)


def test_thermostat_rescale(orc, gpu_engine_factory):
    say("thermostat")
    box = systems.tip4p2005_box(216, seed=23, min_sep=1.8)
    v = start_velocities(orc, box, seed=3)
    e, flags = water_md_engine(gpu_engine_factory, box, v, rc=7.0, skin=1.0, sz=(32, 32, 32), order=5,
                               e3b=False, dt=1.0)
    kb, mvv = systems.REAL_BOLTZ, systems.REAL_MVV_TO_ENG
    e.thermostat(True, 350.0, 5, 0.0, kb, mvv)
    e.run(100, flags)          # ends right after a rescale (100 % 5 == 0)
    T, ke = e.temperature(kb, mvv)
    _, vv = e.get_system()
    ke_host = orc.water_kinetic(box, vv)
    ndf = 3 * 3 * len(box["mol"]) - 3 * len(box["mol"]) - 3
    assert ke == pytest.approx(ke_host, rel=1e-9)
    assert T == pytest.approx(2 * ke_host / (ndf * kb), rel=1e-9)
    assert T == pytest.approx(350.0, rel=1e-6)


@pytest.mark.parametrize("graph", [0, 1])
def test_nose_hoover_chain(orc, gpu_engine_factory, graph):
    """Nose-Hoover chain at 298 K from a 350 K start: the temperature settles
    at the set point and the extended energy (kinetic + potential + chain)
    is conserved, eager and captured alike."""
    say(f"Nose-Hoover chain, graph={graph}")
    from md_engine_amd import ENERGY

    box = systems.tip4p2005_box(343, seed=24, min_sep=1.8)
    v = start_velocities(orc, box, seed=6) * np.sqrt(350.0 / 298.0)
    e, flags = water_md_engine(gpu_engine_factory, box, v, rc=8.0, skin=1.0, sz=(32, 32, 32), order=5,
                               e3b=True, dt=1.0)
    kb, mvv = systems.REAL_BOLTZ, systems.REAL_MVV_TO_ENG
    e.nose_hoover(True, 298.0, 50.0, 10, kb, mvv)
    if graph:
        e.option("graph", 1)

    def extended():
        e.compute(flags | ENERGY)
        en, _ = e.energies()
        T, ke = e.temperature(kb, mvv)
        return en.sum() + ke + e.nose_hoover_energy(), T, ke

    e.run(1000, flags)
    h0, _, ke0 = extended()
    temps, hs = [], []
    for _ in range(10):
        e.run(100, flags)
        h, T, _ = extended()
        temps.append(T)
        hs.append(h)
    say(f"T mean {np.mean(temps):.1f} K, extended-energy span {np.ptp(hs + [h0]):.3f} (KE {ke0:.1f})")
    assert np.mean(temps) == pytest.approx(298.0, rel=0.05)
    assert np.ptp(hs + [h0]) < 0.01 * ke0


@pytest.mark.parametrize("graph", [0, 1])
def test_nose_hoover_chain_matches_oracle(orc, gpu_engine_factory, graph):
    """50 steps under a 10-link chain at 298 K from a 350 K start: particle
    velocities, positions and every chain variable against the FP64
    restatement of FixNoseHoover::thermostatIntegrate (oracle/thermostat.py)."""
    say(f"Nose-Hoover chain vs oracle, graph={graph}")
    from oracle.thermostat import water_md_nh

    box = systems.tip4p2005_box(216, seed=25, min_sep=1.8)
    v = start_velocities(orc, box, seed=8) * np.sqrt(350.0 / 298.0)
    kw = dict(rc=7.0, skin=1.0, sz=(32, 32, 32), order=5)
    e, flags = water_md_engine(gpu_engine_factory, box, v, e3b=False, dt=1.0, **kw)
    kb, mvv = systems.REAL_BOLTZ, systems.REAL_MVV_TO_ENG
    e.nose_hoover(True, 298.0, 20.0, 10, kb, mvv)
    if graph:
        e.option("graph", 1)
    e.run(50, flags)
    xg, vg = e.get_system()
    st = e.nose_hoover_state(10)
    xo, vo, ch = water_md_nh(box, v, 50, 298.0, 20.0, chain=10, dt=1.0, e3b_on=False, **kw)
    so = ch.state()
    e.close()
    dx = np.abs(min_image_diff(xg, xo, box["L"])).max()
    heavy = box["mass"] > 0
    rel_v = np.abs(vg[heavy] - vo[heavy]).max() / np.abs(vo[heavy]).max()
    say(f"dx {dx:.2e} A, rel v {rel_v:.2e}, chain v {st[:10]} vs {so[:10]}")
    assert dx < 1e-5, dx
    assert rel_v < 1e-4, rel_v
    # chain velocities, forces, masses and positions, sum m v^2, last scale
    for name, a, b in (("v", st[0:10], so[0:10]), ("G", st[10:20], so[10:20]), ("Q", st[20:30], so[20:30]),
                       ("xi", st[30:40], so[30:40])):
        err = np.abs(a - b).max() / np.abs(b).max()
        assert err < 1e-4, (name, err)
    assert st[40] == pytest.approx(so[40], rel=1e-5)
    assert st[41] == pytest.approx(so[41], rel=1e-9)
