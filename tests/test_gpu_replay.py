"""Replayed chunks.  When a list rebuild inside a chunk overflows its rows,
the step loop restores the chunk's starting state and runs it again with
grown rows; every piece of state that the steps advance (particles, the
Nose-Hoover chain, the thermostat sums, the turn) must come back, so the
trajectory equals a run whose rows never overflowed, bit for bit."""
import sys

import numpy as np
import pytest

from md_engine_amd import systems

pytestmark = pytest.mark.gpu

G = dict(mO=15.9994, mH=1.008, dOH=0.9572, dHH=1.5139, gamma=0.73612446364836)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def water(make, nmol, seed, nh):
    import oracle
    from md_engine_amd import PAIR, PPPM

    box = systems.tip4p2005_box(nmol, seed=seed, min_sep=1.8)
    v = systems.maxwell_boltzmann(box["mass"], 298.0, seed=seed + 1, boltz=systems.REAL_BOLTZ,
                                  mvv_to_eng=systems.REAL_MVV_TO_ENG)
    for o, h1, h2, m in box["mol"]:
        v[o], v[h1], v[h2] = oracle.rattle_one(box["x"][o], box["x"][h1], box["x"][h2], v[o], v[h1], v[h2],
                                               G["mO"], G["mH"])
    e = make()
    inv = np.where(box["mass"] > 0, 1.0 / np.maximum(box["mass"], 1e-30), 0.0)
    e.set_system(box["x"], box["lo"], box["L"], v=v, inv_mass=inv, q=box["q"], types=box["type"])
    e.exclusions_from_bonds(systems.water_bonds(box["mol"]))
    rc = 7.0
    e.nlist_config(rc + 1.0, 1.0)
    eps24, sig6 = systems.water_lj_tables()
    e.lj_set(np.full((3, 3), rc * rc), eps24, sig6)
    e.pppm_set((32, 32, 32), 5, rc, systems.REAL_QQR_TO_ENG)
    e.ewald_real_set(True, -1.0, systems.REAL_QQR_TO_ENG, rc)
    e.rigid_set(box["mol"], G["mO"], G["mH"], G["dOH"], G["dHH"], G["gamma"])
    e.integrator_set(2.0, systems.REAL_FTM_TO_V)
    if nh:
        e.nose_hoover(True, 298.0, 20.0, 10, systems.REAL_BOLTZ, systems.REAL_MVV_TO_ENG)
    return e, PAIR | PPPM


def tight_rows(e, flags):
    """Rows exactly as long as the starting list's longest: a later rebuild
    in which any row grew overflows them and its chunk is replayed."""
    e.compute(flags)
    counts, _ = e.nlist_get()
    cap = max(16, int(counts.max()))
    e.option("row_cap", cap)
    return cap


@pytest.mark.parametrize("graph,trigger", [(0, "forced"), (1, "forced"), (1, "rows")])
def test_overflow_replay_with_nose_hoover_is_exact(gpu_engine_factory, graph, trigger):
    """A replayed chunk that starts with a list carried in from an earlier
    build: the replay builds that list again from the positions and box of
    its build, so the rows (and the FP32 summation order they set) are the
    same, and the Nose-Hoover chain and thermostat sums come back with the
    particles.  "forced" replays the third chunk (turns 128-191) through the
    replay_chunk option; "rows" makes the rows exactly as long as the
    starting list's longest, so a later rebuild overflows them."""
    say(f"overflow replay with Nose-Hoover, graph={graph}, trigger={trigger}")
    steps = 384
    ref, flags = water(gpu_engine_factory, 343, 31, nh=True)
    tst, _ = water(gpu_engine_factory, 343, 31, nh=True)
    for e in (ref, tst):
        e.option("graph", graph)
    if trigger == "rows":
        # both engines take the same set-up path and differ only in row length
        cap0 = tight_rows(tst, flags)
        ref.compute(flags)
        ref.option("row_cap", 2 * cap0)
    else:
        tst.option("replay_chunk", 3)
    ref.run(steps, flags)
    tst.run(steps, flags)
    if trigger == "rows":
        cap1 = tst.nlist_stats()["row_cap"]
        say(f"row capacity {cap0} -> {cap1}")
        assert cap1 > cap0, "no rebuild overflowed the tight rows: the replay was not exercised"
    xr, vr = ref.get_system()
    xt, vt = tst.get_system()
    sr, stt = ref.nose_hoover_state(10), tst.nose_hoover_state(10)
    er, et = ref.nose_hoover_energy(), tst.nose_hoover_energy()
    nr, nt = ref.nlist_stats()["rebuilds"], tst.nlist_stats()["rebuilds"]
    ref.close()
    tst.close()
    say(f"  max |dx| {np.abs(xr - xt).max():.2e}, rebuilds {nr} / {nt}")
    assert nr == nt
    assert np.array_equal(xr, xt)
    assert np.array_equal(vr, vt)
    assert np.array_equal(sr, stt)
    assert er == et


@pytest.mark.parametrize("trigger", ["rows", "forced"])
def test_batch_replay_keeps_replicas_in_step(gpu_engine_factory, trigger):
    """Three replicas in one batch, one with tight rows (or made to replay
    its second chunk): its replay replays the chunk for every replica, and
    each still equals its own single run."""
    from md_engine_amd._lib import run_batch

    say(f"batch replay ({trigger})")
    steps = 128
    seeds = (41, 42, 43)
    singles = []
    for s in seeds:
        e, flags = water(gpu_engine_factory, 216, s, nh=False)
        e.option("graph", 1)
        e.run(steps, flags)
        singles.append(e.get_system())
        e.close()
    batch = []
    for s in seeds:
        e, flags = water(gpu_engine_factory, 216, s, nh=False)
        batch.append(e)
    if trigger == "rows":
        cap0 = tight_rows(batch[1], flags)
    else:
        batch[1].option("replay_chunk", 2)
    run_batch(batch, steps, flags)
    if trigger == "rows":
        cap1 = batch[1].nlist_stats()["row_cap"]
        say(f"replica 1 row capacity {cap0} -> {cap1}")
        assert cap1 > cap0
    for e, (xs, vs) in zip(batch, singles):
        xb, vb = e.get_system()
        assert np.array_equal(xb, xs)
        assert np.array_equal(vb, vs)
    for e in batch:
        e.close()
