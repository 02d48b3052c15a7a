"""Every benchmark configuration at its full size, HIP path against the FP64
oracle through the C ABI (bars: neighbor lists bit-exact per id with tags,
forces relative RMSE <= 1e-4, energies relative 1e-5, alpha bit-identical).

C1: LJ fluid, 32,000 atoms FCC, rho* 0.8442, T* 1.2, rc 2.5, skin 0.3,
    1,000 NVE steps at dt 0.005.
C2: rigid TIP4P/2005, 4,096 molecules, L 49.72 A, rc 9, skin 1, P3M 64^3
    order 5, Nose-Hoover 298 K, dt 2 fs.
C3: TIP4P/2005 + E3B3, 8,192 molecules, L 62.64 A, 298 K, 1 bar Berendsen.
C4: C3 with the learned three-body term, trained on the GPU from C3 frames.
C5: 48 uneven TIP4P/2005 + E3B3 boxes (512-4,096 molecules) in one batch
    against single-context runs, and the 32-bead ring of 1,024-molecule
    q-TIP4P/F beads (131,072 bead sites).
"""
import sys

import numpy as np
import pytest

from md_engine_amd import systems

pytestmark = pytest.mark.gpu

FORCE_RMSE = 1e-4
ENERGY_REL = 1e-5
KJ_TO_KCAL = 0.23900573614
TIP = systems.TIP4P2005


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_rmse(f, ref):
    return float(np.sqrt(((f - ref) ** 2).sum() / (ref ** 2).sum()))


def e3b3():
    k = KJ_TO_KCAL
    return [453000 * k, 150 * k, -1005 * k, 1880 * k, 4.872, 1.907, 5.0, 5.2]


def water_engine(make, box, e3b, rc=9.0, skin=1.0, sz=(64, 64, 64), dt=2.0, v=None):
    from md_engine_amd import E3B, PAIR, PPPM

    e = make()
    inv = np.where(box["mass"] > 0, 1.0 / np.maximum(box["mass"], 1e-30), 0.0)
    e.set_system(box["x"], box["lo"], box["L"], v=v, inv_mass=inv, q=box["q"], types=box["type"])
    e.exclusions_from_bonds(systems.water_bonds(box["mol"]))
    e.nlist_config(rc + skin, skin)
    eps24, sig6 = systems.water_lj_tables()
    e.lj_set(np.full((3, 3), rc * rc), eps24, sig6)
    e.pppm_set(sz, 5, rc, systems.REAL_QQR_TO_ENG)
    e.ewald_real_set(True, -1.0, systems.REAL_QQR_TO_ENG, rc)
    flags = PAIR | PPPM
    if e3b:
        e.e3b_set(box["mol"], e3b3())
        flags |= E3B
    e.rigid_set(box["mol"], TIP["m_O"], TIP["m_H"], TIP["r_OH"], TIP["r_HH"], TIP["gamma"])
    e.integrator_set(dt, systems.REAL_FTM_TO_V)
    return e, flags


def velocities(box, seed):
    import oracle

    v = systems.maxwell_boltzmann(box["mass"], 298.0, seed=seed, boltz=systems.REAL_BOLTZ,
                                  mvv_to_eng=systems.REAL_MVV_TO_ENG)
    for o, h1, h2, _ in box["mol"]:
        v[o], v[h1], v[h2] = oracle.rattle_one(box["x"][o], box["x"][h1], box["x"][h2], v[o], v[h1], v[h2],
                                               TIP["m_O"], TIP["m_H"])
    return v


def oracle_lists(orc, x, lo, L, rlist, bonds=None):
    n = len(x)
    xw = orc.wrap(x, lo, L)
    counts, tab = orc.nlist_cell(xw, lo, L, rlist)
    rows = [tab[i, : counts[i]].copy() for i in range(n)]
    if bonds is not None:
        rows = orc.tag_rows(rows, np.arange(n), orc.exclusions_bfs(n, bonds))
    else:
        rows = [r.astype(np.uint32) for r in rows]
    return counts, [np.sort(r) for r in rows], xw


def assert_same_lists(e, oc, orows):
    gc, grows = e.nlist_get()
    assert np.array_equal(gc, oc), f"count mismatch at {np.nonzero(gc != oc)[0][:10]}"
    bad = [i for i in range(len(orows)) if not np.array_equal(np.sort(grows[i]), orows[i])]
    assert not bad, f"{len(bad)} rows differ, first {bad[:5]}"
    return int(gc.sum())


def water_state_parity(orc, e, box, flags, e3b, rc=9.0, skin=1.0, sz=(64, 64, 64)):
    """At the engine's current state: lists, forces, energies, alpha."""
    from md_engine_amd import ENERGY

    x, _ = e.get_system()
    lo, L = e.get_box()
    st = dict(box, x=x, lo=np.asarray(lo), L=np.asarray(L))
    e.nlist_update(force=True)
    oc, orows, xw = oracle_lists(orc, x, st["lo"], st["L"], rc + skin, systems.water_bonds(box["mol"]))
    entries = assert_same_lists(e, oc, orows)
    n = len(x)
    qqr = systems.REAL_QQR_TO_ENG
    q2 = orc.total_q2(box["q"], qqr)
    alpha, _ = orc.pppm_alpha(n, q2, st["L"], np.array(sz), 5, rc)
    a_gpu, _ = e.pppm_info()
    assert a_gpu == alpha
    eps24, sig6 = systems.water_lj_tables()
    fp, ep, _ = orc.pair(xw, box["q"], box["type"], st["L"], oc, orows, np.full((3, 3), rc * rc), eps24, sig6,
                         with_charge=True, alpha=alpha, qqr=qqr, rq=rc)
    G = orc.pppm_green(st["L"], np.array(sz), alpha, 5)
    k = orc.pppm(x, box["q"], st["lo"], st["L"], np.array(sz), 5, qqr, alpha, q2, G, energy=True)
    ref = fp + k["f"]
    e_ref = [ep.sum(), k["energy"], 0.0, 0.0]
    if e3b:
        fe, e2, e3 = orc.e3b(box["mol"], x, st["L"])
        ref = ref + fe
        e_ref[2:] = [e2, e3]
    e.compute(flags | ENERGY)
    f = e.forces()
    en, _ = e.energies()
    rel = rel_rmse(f, ref)
    say(f"  {entries} list entries bit-exact; force RMSE {rel:.2e}; energies {en} vs {np.array(e_ref)}")
    assert rel <= FORCE_RMSE
    for a, b in zip(en, e_ref):
        if b != 0.0:
            assert a == pytest.approx(b, rel=ENERGY_REL)
    return x, st


def test_c1_lj_32000_nve_1000_steps(orc, gpu_engine_factory):
    say("C1: LJ 32,000 atoms, 1,000 NVE steps")
    from md_engine_amd import ENERGY, PAIR

    x0, L, lo = systems.fcc_lj(20, 0.8442)
    n = len(x0)
    v0 = systems.maxwell_boltzmann(np.ones(n), 1.2, seed=20261202)
    e = gpu_engine_factory()
    e.set_system(x0, lo, L, v=v0, inv_mass=np.ones(n))
    e.nlist_config(2.8, 0.3)
    e.lj_set([[6.25]], [[24.0]], [[1.0]])
    e.integrator_set(0.005, 1.0)
    e.option("graph", 1)

    def total():
        e.compute(PAIR | ENERGY)
        en, _ = e.energies()
        _, v = e.get_system()
        return en[0] + 0.5 * (v * v).sum()

    h0 = total()
    hs = []
    for _ in range(10):
        e.run(100, PAIR)
        hs.append(total())
    drift = (hs[-1] - h0) / n
    say(f"  energy per atom: start {h0 / n:.6f}, drift over 1000 steps {drift:.2e}, span {np.ptp(hs) / n:.2e}")
    assert abs(drift) < 2e-4 and np.ptp(hs) / n < 5e-4
    assert e.nlist_stats()["violations"] == 0
    # the list and forces at the final state against the oracle
    x, _ = e.get_system()
    e.nlist_update(force=True)
    oc, orows, xw = oracle_lists(orc, x, lo, L, 2.8)
    entries = assert_same_lists(e, oc, orows)
    fo, eo, _ = orc.pair(xw, np.zeros(n), np.zeros(n, np.int32), L, oc, orows, np.array([[6.25]]),
                         np.array([[24.0]]), np.array([[1.0]]))
    e.compute(PAIR | ENERGY)
    rel = rel_rmse(e.forces(), fo)
    en, _ = e.energies()
    say(f"  {entries} entries bit-exact, force RMSE {rel:.2e}")
    assert rel <= FORCE_RMSE
    assert en[0] == pytest.approx(eo.sum(), rel=ENERGY_REL)
    e.close()


def test_c2_tip4p_4096(orc, gpu_engine_factory):
    say("C2: TIP4P/2005, 4,096 molecules, 64^3 P3M, Nose-Hoover 298 K")
    box = systems.tip4p2005_box(4096, seed=20261202)
    e, flags = water_engine(gpu_engine_factory, box, e3b=False, v=velocities(box, 11))
    e.nose_hoover(True, 298.0, 100.0, 10, systems.REAL_BOLTZ, systems.REAL_MVV_TO_ENG)
    e.option("graph", 1)
    e.run(300, flags)
    assert e.nlist_stats()["violations"] == 0
    x, st = water_state_parity(orc, e, box, flags, e3b=False)
    # ten more steps against the FP64 integrator from the same state
    _, v = e.get_system()
    e2, _ = water_engine(gpu_engine_factory, dict(st, x=x), e3b=False, v=v)
    e2.run(10, flags)
    xg, vg = e2.get_system()
    xo, vo, _ = orc.water_md(dict(st, x=x), v, 10, dt=2.0, rc=9.0, skin=1.0, sz=(64, 64, 64), order=5)
    d = xg - xo
    d -= st["L"] * np.rint(d / st["L"])
    heavy = box["mass"] > 0
    rel_v = np.abs(vg[heavy] - vo[heavy]).max() / np.abs(vo[heavy]).max()
    say(f"  10-step trajectory: max |dx| {np.abs(d).max():.2e} A, rel v {rel_v:.2e}")
    assert np.abs(d).max() < 1e-5 and rel_v < 1e-4
    e.close()
    e2.close()


def test_c3_tip4p_e3b_8192_barostat(orc, gpu_engine_factory):
    say("C3: TIP4P/2005 + E3B3, 8,192 molecules, 1 bar Berendsen")
    box = systems.tip4p2005_box(8192, seed=20261202)
    e, flags = water_engine(gpu_engine_factory, box, e3b=True, v=velocities(box, 12))
    e.thermostat(True, 298.0, 10, 0.0, systems.REAL_BOLTZ, systems.REAL_MVV_TO_ENG)
    e.barostat(True, 1.0 / 1.01325, 1000.0, 1, mvv_to_eng=systems.REAL_MVV_TO_ENG)
    e.option("graph", 1)
    L0 = e.get_box()[1][0]
    e.run(300, flags)
    L1 = e.get_box()[1][0]
    say(f"  box {L0:.4f} -> {L1:.4f} A")
    assert L1 != L0
    assert e.nlist_stats()["violations"] == 0
    water_state_parity(orc, e, box, flags, e3b=True)
    e.close()


def test_c4_surrogate_8192(orc, gpu_engine_factory):
    """Trained on the GPU from frames of the C3 system, the learned three-body
    term meets its bar (force RMSE <= 0.5 kJ/mol/nm against E3B) on a frame
    recorded later, its forces match the FP64 restatement of the same network,
    and it beats dropping the three-body term by far."""
    say("C4: learned three-body term at 8,192 molecules")
    from md_engine_amd import E3B, SURROGATE
    from oracle import surrogate as S

    box = systems.tip4p2005_box(8192, seed=20261202)
    e, flags = water_engine(gpu_engine_factory, box, e3b=True, v=velocities(box, 13))
    e.thermostat(True, 298.0, 10, 0.0, systems.REAL_BOLTZ, systems.REAL_MVV_TO_ENG)
    e.option("graph", 1)
    e.run(1000, flags)
    nm = len(box["mol"])
    frames = 12
    for k in range(frames):
        e.run(100, flags)
        e.surrogate_record(reset=(k == 0), capacity=frames * nm)
    e.surrogate_set(S.init_params(3).astype(np.float32))
    e.surrogate_fit(1, batch=8192, lr=1e-3, seed=1, normalize=True)
    for k, (ep, lr) in enumerate(((60, 1e-3), (30, 3e-4), (30, 1e-4), (15, 3e-5))):
        loss = e.surrogate_fit(ep, batch=1024, lr=lr, seed=2 + k)
    e.run(500, flags)
    x, _ = e.get_system()
    L = np.asarray(e.get_box()[1])
    e.compute(E3B)
    f_e3b = e.forces()
    e.compute(SURROGATE)
    f_sur = e.forces()
    # both share the analytic two-body term: the difference is the three-body error
    rms = float(np.sqrt(((f_sur - f_e3b) ** 2).mean()))
    say(f"  fit loss {loss:.5f}; three-body force RMSE vs E3B {rms:.5f} kcal/mol/A = {rms * 41.84:.3f} kJ/mol/nm")
    assert rms * 41.84 <= 0.5
    # the GPU network (BF16 operands) against its FP64 restatement
    p = orc.e3b_params()
    two_only = orc.E3BParams(p.E2, 0.0, 0.0, 0.0, p.k2, p.k3, p.rs, p.rf, p.r2b, p.rmax)
    f_two, _, _ = orc.e3b(box["mol"], x, L, params=two_only)
    fo_sur, _ = S.forces(e.surrogate_get().astype(np.float64), box["mol"], x, L)
    rel = rel_rmse(f_sur - f_two, fo_sur)
    say(f"  surrogate three-body forces vs FP64 restatement: relative RMSE {rel:.2e}")
    assert rel < 3e-2
    drop = float(np.sqrt(((f_e3b - f_two) ** 2).mean()))
    say(f"  dropping the three-body term instead: RMSE {drop:.4f} kcal/mol/A")
    assert rms < 0.05 * drop
    e.close()


def test_c5_batch_48_boxes_equal_single_runs(orc, gpu_engine_factory):
    say("C5: 48 uneven boxes in one batch vs single runs")
    from md_engine_amd._lib import run_batch

    rng = np.random.default_rng(20261202)
    sizes = rng.integers(512, 4097, 48)
    steps = 16
    batch, singles = [], []
    for k, m in enumerate(sizes):
        box = systems.tip4p2005_box(int(m), seed=20261202 + k)
        sz = (32, 32, 32) if box["L"][0] < 40 else (64, 64, 64)
        v = velocities(box, 100 + k)
        e, flags = water_engine(gpu_engine_factory, box, e3b=True, sz=sz, v=v)
        batch.append(e)
        s, _ = water_engine(gpu_engine_factory, box, e3b=True, sz=sz, v=v)
        s.option("graph", 1)
        s.run(steps, flags)
        singles.append(s.get_system())
        s.close()
    say(f"  sizes {int(sizes.min())}-{int(sizes.max())} molecules, {4 * int(sizes.sum())} sites")
    run_batch(batch, steps, flags)
    assert batch[0].batch_merged(), "the 48 replicas did not run as one launch per kernel"
    for k, (e, (xs, vs)) in enumerate(zip(batch, singles)):
        xb, vb = e.get_system()
        assert np.array_equal(xb, xs), k
        assert np.array_equal(vb, vs), k
        e.close()


def test_c5_ring_32_beads_1024_molecules(orc, gpu_engine_factory):
    """The 32-bead ring at full size: the free ring polymer (no force terms)
    against the exact normal-mode propagation for every bead site, then the
    forces of two beads (pair, P3M, intramolecular) against the FP64 oracle."""
    say("C5: 32-bead q-TIP4P/F ring, 1,024 molecules per bead")
    from md_engine_amd._lib import ring_omega, run_ring
    from oracle import flexwater as FW
    from oracle import ring as R

    P = 32
    box = systems.tip4p2005_box(1024, seed=20261203)
    tb = P * 298.0
    xs = systems.ring_beads(box["x"], box["mass"], P, tb, 20261203)
    rng = np.random.default_rng(5)
    inv = np.where(box["mass"] > 0, 1.0, 0.0)[:, None]
    vs = [rng.normal(0, 0.005, box["x"].shape) * inv for _ in range(P)]
    es = []
    for x, v in zip(xs, vs):
        e = gpu_engine_factory()
        im = np.where(box["mass"] > 0, 1.0 / np.maximum(box["mass"], 1e-30), 0.0)
        e.set_system(x, box["lo"], box["L"], v=v, inv_mass=im, q=box["q"], types=box["type"])
        e.exclusions_from_bonds(systems.water_bonds(box["mol"]))
        e.nlist_config(10.0, 1.0)
        eps24, sig6 = systems.water_lj_tables()
        e.lj_set(np.full((3, 3), 81.0), eps24, sig6)
        e.pppm_set((32, 32, 32), 5, 9.0, systems.REAL_QQR_TO_ENG)
        e.ewald_real_set(True, -1.0, systems.REAL_QQR_TO_ENG, 9.0)
        p = FW.QTIP4PF
        e.flexible_water_set(box["mol"], p["k2"], p["k3"], p["k4"], p["r0"], p["ktheta"], p["theta0"], p["gamma"])
        e.integrator_set(0.5, systems.REAL_FTM_TO_V)
        es.append(e)
    for e in es:
        e.compute(0)
    wP = ring_omega(tb, systems.REAL_BOLTZ, systems.REAL_HBAR)
    nsteps = 4
    run_ring(es, nsteps, 0, wP)
    L = box["L"]
    x = np.array(xs)
    x = x[0] + (x - x[0]) - L * np.rint((x - x[0]) / L)
    v = np.array(vs)
    for _ in range(nsteps):
        x, v = R.free_step(x, v, wP, 0.5)
        x = x[0] + (x - x[0]) - L * np.rint((x - x[0]) / L)
    mol = box["mol"]
    massive = np.concatenate([mol[:, 0], mol[:, 1], mol[:, 2]])
    worst = 0.0
    for k, e in enumerate(es):
        xg, vg = e.get_system()
        d = xg[massive] - x[k][massive]
        d -= L * np.rint(d / L)
        worst = max(worst, float(np.abs(d).max()))
        assert np.abs(vg[massive] - v[k][massive]).max() < 1e-10, f"bead {k}"
    say(f"  free ring: max bead-site deviation {worst:.2e} A over {P * len(box['x'])} bead sites")
    assert worst < 1e-9
    # forces of two beads at their current positions
    from test_gpu_parity import oracle_water

    for k in (0, 17):
        xg, _ = es[k].get_system()
        xm = FW.place_m(mol, xg, L, FW.QTIP4PF["gamma"])
        st = dict(box, x=xm)
        from md_engine_amd import BONDED, PAIR, PPPM

        es[k].compute(PAIR | PPPM | BONDED)
        f = es[k].forces()
        o = oracle_water(orc, st, rc=9.0, skin=1.0, sz=(32, 32, 32), order=5)
        fb, _, _ = FW.bonded(mol, xm, L)
        ref = FW.spread_m(mol, o["fp"] + o["fk"], FW.QTIP4PF["gamma"]) + fb
        rel = rel_rmse(f, ref)
        say(f"  bead {k}: force RMSE {rel:.2e}")
        assert rel <= FORCE_RMSE
    for e in es:
        e.close()
