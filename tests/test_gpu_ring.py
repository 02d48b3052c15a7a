"""Ring-polymer (PIMD) steps on the GPU (mdx_ring_run, md_engine_amd/csrc/
ring.hip): the free ring polymer against the exact normal-mode propagation of
oracle/ring.py, every bead's forces against the FP64 oracle with the mesh
part contracted to the centroids (the reference's scheme) and per bead, the
neighbor list built on the centroids (the reference's) or per bead, and
the ring Hamiltonian of flexible q-TIP4P/F water beads conserved to second
order in dt in both schemes."""
import sys

import numpy as np
import pytest

from md_engine_amd import systems
from md_engine_amd._lib import PPPM_PER_BEAD, RING_BEAD_LISTS, ring_forces, ring_omega, run_ring
from oracle import flexwater as FW
from oracle import ring as R

from test_gpu_flex import flex_engine
from test_gpu_water_md import start_velocities

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bead_systems(box, P, seed, spread=0.05):
    """P copies of the box, the massive sites displaced a little per bead"""
    rng = np.random.default_rng(seed)
    mol = box["mol"]
    out = []
    for _ in range(P):
        x = box["x"].copy()
        for h in (0, 1, 2):
            x[mol[:, h]] += rng.normal(0, spread, (len(mol), 3))
        out.append(FW.place_m(mol, x, box["L"], FW.QTIP4PF["gamma"]))
    return out


def ring_engines(make, box, xs, vs, dt, rc=5.5, skin=0.5, sz=(32, 32, 32)):
    es = []
    for x, v in zip(xs, vs):
        e, flags = flex_engine(make, dict(box, x=x), v, rc=rc, skin=skin, sz=sz, order=5, dt=dt)
        es.append(e)
    return es, flags


def test_free_ring_matches_normal_modes(orc, gpu_engine_factory):
    say("free ring polymer, 4 beads")
    P = 4
    box = systems.tip4p2005_box(64, seed=31, min_sep=1.8)
    xs = bead_systems(box, P, 1)
    rng = np.random.default_rng(2)
    inv = np.where(box["mass"] > 0, 1.0, 0.0)[:, None]
    vs = [rng.normal(0, 0.005, box["x"].shape) * inv for _ in range(P)]
    es, _ = ring_engines(gpu_engine_factory, box, xs, vs, dt=0.5)
    for e in es[1:]:
        e.compute(0)  # lists and P3M plans made on the beads' own streams first
    wP = ring_omega(P * 298.0, systems.REAL_BOLTZ, systems.REAL_HBAR)
    nsteps = 12
    run_ring(es, nsteps, 0, wP)  # no force terms: only the ring springs act
    # oracle: the same beads put in bead 0's image, then exact steps
    L = box["L"]
    x = np.array(xs)
    x = x[0] + (x - x[0]) - L * np.rint((x - x[0]) / L)
    v = np.array(vs)
    for _ in range(nsteps):
        x, v = R.free_step(x, v, wP, 0.5)
        x = x[0] + (x - x[0]) - L * np.rint((x - x[0]) / L)
    mol = box["mol"]
    massive = np.concatenate([mol[:, 0], mol[:, 1], mol[:, 2]])
    for k, e in enumerate(es):
        xg, vg = e.get_system()
        d = xg[massive] - x[k][massive]
        d -= L * np.rint(d / L)
        assert np.abs(d).max() < 1e-9, f"bead {k}"
        assert np.abs(vg[massive] - v[k][massive]).max() < 1e-11, f"bead {k}"
        e.close()


@pytest.mark.parametrize("mode", ["contracted", "per-bead"])
def test_ring_forces_vs_oracle(orc, gpu_engine_factory, mode):
    """Every bead's force: LJ + Ewald real space of the bead's own
    configuration, the q-TIP4P/F intramolecular terms, and the mesh part
    either from the ring's centroids (whole force on every bead) or per bead,
    M-site forces spread onto O and H; relative RMSE <= 1e-4 per bead."""
    say(f"ring forces, 4 beads of 64 molecules, mesh part {mode}")
    from test_gpu_parity import oracle_rows_by_id

    P, rc, skin, sz = 4, 5.5, 0.5, (32, 32, 32)
    box = systems.tip4p2005_box(64, seed=37, min_sep=1.8)
    xs = bead_systems(box, P, 4, spread=0.08)
    vs = [np.zeros_like(x) for x in xs]
    es, flags = ring_engines(gpu_engine_factory, box, xs, vs, dt=0.5, rc=rc, skin=skin, sz=sz)
    ring_forces(es, flags | (PPPM_PER_BEAD if mode == "per-bead" else 0))
    L, lo, mol = box["L"], box["lo"], box["mol"]
    n = len(box["x"])
    qqr = systems.REAL_QQR_TO_ENG
    q2 = orc.total_q2(box["q"], qqr)
    xa = np.array(xs)
    if mode == "contracted":
        alpha, fmesh = R.contracted_pppm(xa, box["q"], lo, L, sz, 5, qqr, rc)
    excl = orc.exclusions_bfs(n, systems.water_bonds(mol))
    eps24, sig6 = systems.water_lj_tables()
    gamma = FW.QTIP4PF["gamma"]
    # one list on the ring's centroids, every bead evaluating its id pairs
    oc, orows, _ = oracle_rows_by_id(orc, R.centroids(xa, lo, L), lo, L, rc + skin, excl=excl, brute=False)
    for k, e in enumerate(es):
        x = xa[k]
        if mode == "per-bead":
            alpha, _ = orc.pppm_alpha(n, q2, L, np.array(sz), 5, rc)
            G = orc.pppm_green(L, np.array(sz), alpha, 5)
            fmesh = orc.pppm(x, box["q"], lo, L, np.array(sz), 5, qqr, alpha, q2, G)["f"]
        assert e.pppm_info()[0] == alpha
        xw = orc.wrap(x, lo, L)
        fp, _, _ = orc.pair(xw, box["q"], box["type"], L, oc, orows, np.full((3, 3), rc * rc), eps24, sig6,
                            with_charge=True, alpha=alpha, qqr=qqr, rq=rc)
        fb, _, _ = FW.bonded(mol, x, L)
        ref = FW.spread_m(mol, fp + fmesh, gamma) + fb
        f = e.forces()
        rel = float(np.sqrt(((f - ref) ** 2).sum() / (ref ** 2).sum()))
        say(f"bead {k}: relative force RMSE {rel:.2e}")
        assert rel < 1e-4, f"bead {k}"
    for e in es:
        e.close()


def test_ring_centroid_lists(orc, gpu_engine_factory):
    """The ring's neighbor list is built on the centroids (GridGPU.cu:163-189,
    888-895): every bead's rows, read back by id, are the oracle's cell list of
    the centroids bit for bit, and its forces pair bead k with bead k over
    those rows.  With beads spread far enough that lists of the bead positions
    differ from the centroid list, MDX_RING_BEAD_LISTS gives every bead the
    oracle's list of its own positions instead."""
    say("ring neighbor lists on the centroids, 4 beads of 64 molecules")
    from test_gpu_parity import assert_same_lists, oracle_rows_by_id

    P, rc, skin, sz = 4, 5.5, 0.5, (32, 32, 32)
    box = systems.tip4p2005_box(64, seed=41, min_sep=1.8)
    xs = bead_systems(box, P, 9, spread=0.25)
    vs = [np.zeros_like(x) for x in xs]
    es, flags = ring_engines(gpu_engine_factory, box, xs, vs, dt=0.5, rc=rc, skin=skin, sz=sz)
    L, lo, mol = box["L"], box["lo"], box["mol"]
    n = len(box["x"])
    xa = np.array(xs)
    excl = orc.exclusions_bfs(n, systems.water_bonds(mol))
    cen = R.centroids(xa, lo, L)
    oc, orows, _ = oracle_rows_by_id(orc, cen, lo, L, rc + skin, excl=excl, brute=False)
    bead_rows = [oracle_rows_by_id(orc, xa[k], lo, L, rc + skin, excl=excl, brute=False)[:2] for k in range(P)]
    differ = sum(not np.array_equal(bc, oc) for bc, _ in bead_rows)
    say(f"beads whose own list differs from the centroid list: {differ} of {P}")
    assert differ > 0, "the beads are too close to their centroids to tell the lists apart"

    ring_forces(es, flags)
    qqr = systems.REAL_QQR_TO_ENG
    alpha, fmesh = R.contracted_pppm(xa, box["q"], lo, L, sz, 5, qqr, rc)
    eps24, sig6 = systems.water_lj_tables()
    gamma = FW.QTIP4PF["gamma"]
    for k, e in enumerate(es):
        gc, grows = e.nlist_get()
        assert_same_lists(gc, grows, oc, orows)
        xw = orc.wrap(xa[k], lo, L)
        fp, _, _ = orc.pair(xw, box["q"], box["type"], L, oc, orows, np.full((3, 3), rc * rc), eps24, sig6,
                            with_charge=True, alpha=alpha, qqr=qqr, rq=rc)
        fb, _, _ = FW.bonded(mol, xa[k], L)
        ref = FW.spread_m(mol, fp + fmesh, gamma) + fb
        f = e.forces()
        rel = float(np.sqrt(((f - ref) ** 2).sum() / (ref ** 2).sum()))
        say(f"bead {k}: centroid list exact, relative force RMSE {rel:.2e}")
        assert rel < 1e-4, f"bead {k}"

    ring_forces(es, flags | RING_BEAD_LISTS)
    for k, e in enumerate(es):
        gc, grows = e.nlist_get()
        assert_same_lists(gc, grows, *bead_rows[k])
    say("per-bead lists: every bead's own list exact")
    for e in es:
        e.close()


@pytest.mark.parametrize("mode", ["contracted", "per-bead"])
def test_ring_hamiltonian_second_order(orc, gpu_engine_factory, mode):
    """H = sum over beads of (KE + V) + the ring springs; its deviation over the
    same 40 fs falls about fourfold when dt halves.  With the contraction the
    mesh part of H is P times the P3M energy at the centroids (evaluated by a
    separate single-system context at the ring's alpha)."""
    say(f"flexible water ring, 4 beads: energy error vs dt, mesh part {mode}")
    from md_engine_amd import ENERGY, PPPM

    P = 4
    box = systems.tip4p2005_box(64, seed=32, min_sep=1.8)
    xs = bead_systems(box, P, 3, spread=0.02)
    vs = [start_velocities(orc, dict(box, x=x), seed=10 + k) for k, x in enumerate(xs)]
    wP = ring_omega(P * 298.0, systems.REAL_BOLTZ, systems.REAL_HBAR)
    mvv = systems.REAL_MVV_TO_ENG
    L = box["L"]
    errs = []
    contracted = mode == "contracted"
    for dt, nsteps in ((0.2, 200), (0.1, 400)):
        es, flags = ring_engines(gpu_engine_factory, box, xs, vs, dt=dt)
        run_flags = flags | (0 if contracted else PPPM_PER_BEAD)
        ring_forces(es, run_flags)  # alpha of the scheme set on every bead
        cen = None
        if contracted:
            # the mesh term at the centroids, on a context of its own
            cen, _ = flex_engine(gpu_engine_factory, box, np.zeros_like(box["x"]), rc=5.5, skin=0.5,
                                 sz=(32, 32, 32), order=5, dt=dt)
            cen.pppm_set((32, 32, 32), 5, 5.5, systems.REAL_QQR_TO_ENG, alpha=es[0].pppm_info()[0])

        def total():
            h, xb = 0.0, []
            ke = 0.0
            for e in es:
                e.compute((flags & ~PPPM if contracted else flags) | ENERGY)
                en, _ = e.energies()
                x, v = e.get_system()
                k = 0.5 * mvv * (box["mass"][:, None] * v ** 2).sum()
                h += en.sum() + e.bonded_energy() + k
                ke += k
                xb.append(x)
            xb = np.array(xb)
            if contracted:
                cen.set_positions(R.centroids(xb, box["lo"], L))
                cen.compute(PPPM | ENERGY)
                h += P * cen.energies()[0][1]
            xb = xb[0] + (xb - xb[0]) - L * np.rint((xb - xb[0]) / L)
            total.ke = ke
            return h + R.spring_energy(xb, box["mass"], wP, mvv)

        h0 = total()
        ke0 = total.ke
        dev = 0.0
        for _ in range(10):
            run_ring(es, nsteps // 10, run_flags, wP)
            dev = max(dev, abs(total() - h0))
        errs.append(dev)
        for e in es + ([cen] if cen else []):
            e.close()
    say(f"max |H - H0|: dt 0.2 fs {errs[0]:.5f}, dt 0.1 fs {errs[1]:.5f} kcal/mol")
    assert 0 < errs[0] < 0.01 * ke0
    assert errs[1] < 0.4 * errs[0]


@pytest.mark.parametrize("mode", ["contracted", "per-bead", "bead-lists"])
def test_ring_merged_launches_equal_separate(orc, gpu_engine_factory, mode):
    """The ring's per-bead kernels as one launch per kernel over all beads
    (the default) give the trajectory of the beads launched one by one, bit
    for bit, with the mesh part contracted to the centroids or per bead, and
    with one list on the centroids or every bead its own."""
    say(f"ring: merged bead launches vs separate, mesh part {mode}")
    P = 4
    box = systems.tip4p2005_box(64, seed=33, min_sep=1.8)
    xs = bead_systems(box, P, 5, spread=0.02)
    vs = [start_velocities(orc, dict(box, x=x), seed=20 + k) for k, x in enumerate(xs)]
    wP = ring_omega(P * 298.0, systems.REAL_BOLTZ, systems.REAL_HBAR)
    out = []
    for merge in (1, 0):
        es, flags = ring_engines(gpu_engine_factory, box, xs, vs, dt=0.5)
        run_flags = flags | {"contracted": 0, "per-bead": PPPM_PER_BEAD, "bead-lists": RING_BEAD_LISTS}[mode]
        es[0].option("batch_merge", merge)
        run_ring(es, 40, run_flags, wP)
        assert es[0].batch_merged() == bool(merge)
        out.append([e.get_system() for e in es])
        for e in es:
            e.close()
    for (xa, va), (xb, vb) in zip(out[0], out[1]):
        assert np.array_equal(xa, xb)
        assert np.array_equal(va, vb)
