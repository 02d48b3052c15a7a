"""Captured MD steps (hipGraph) and batched replicas on the GPU.

A captured step decides list rebuilds and thermostat turns on the device, so
its trajectory is checked against the FP64 oracle integrator like the eager
path; replicas advanced together in one graph must equal the same replicas
advanced alone, bit for bit (fixed-point force sums make the step
deterministic)."""
import sys

import numpy as np
import pytest

from md_engine_amd import systems

from test_gpu_water_md import G, min_image_diff, start_velocities, water_md_engine

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("nsteps", [1, 15])
def test_graph_trajectory_matches_oracle(orc, gpu_engine_factory, nsteps):
    say(f"captured water trajectory {nsteps} steps vs oracle")
    box = systems.tip4p2005_box(216, seed=21, min_sep=1.8)
    v = start_velocities(orc, box)
    kw = dict(rc=7.0, skin=1.0, sz=(32, 32, 32), order=5)
    e, flags = water_md_engine(gpu_engine_factory, box, v, e3b=True, dt=1.0, **kw)
    e.option("graph", 1)
    e.run(nsteps, flags)
    xg, vg = e.get_system()
    xo, vo, _ = orc.water_md(box, v, nsteps, dt=1.0, e3b_on=True, **kw)
    dx = np.abs(min_image_diff(xg, xo, box["L"])).max()
    assert dx < (1e-8 if nsteps == 1 else 1e-6), dx
    heavy = box["mass"] > 0
    rel_v = np.abs(vg[heavy] - vo[heavy]).max() / np.abs(vo[heavy]).max()
    assert rel_v < (1e-6 if nsteps == 1 else 1e-4), rel_v


def test_graph_long_run_constraints_rebuilds(orc, gpu_engine_factory):
    """300 captured steps (37 full graphs + a 4-step tail): the device-side
    rebuild decision keeps rebuilding and the constraints hold."""
    say("captured water 300 steps")
    box = systems.tip4p2005_box(343, seed=31, min_sep=1.8)
    v = start_velocities(orc, box, seed=4)
    e, flags = water_md_engine(gpu_engine_factory, box, v, rc=8.0, skin=1.0, sz=(32, 32, 32), order=5,
                               e3b=True, dt=2.0)
    e.option("graph", 1)
    e.run(300, flags)
    st = e.nlist_stats()
    assert st["rebuilds"] > 3
    x, vv = e.get_system()
    for o, h1, h2, m in box["mol"]:
        d = min_image_diff(x[[h1, h2]], x[o], box["L"])
        assert np.allclose(np.linalg.norm(d, axis=1), G["dOH"], atol=1e-9)
        for i, j in ((o, h1), (o, h2), (h1, h2)):
            r = min_image_diff(x[i], x[j], box["L"])
            assert abs(np.dot(vv[i] - vv[j], r)) < 1e-9
    # the list built by the captured steps is still the oracle's list
    from test_gpu_parity import assert_same_lists, oracle_rows_by_id
    excl = orc.exclusions_bfs(len(box["x"]), systems.water_bonds(box["mol"]))
    e.nlist_update(force=True)
    gc, grows = e.nlist_get()
    oc, orows, _ = oracle_rows_by_id(orc, x, box["lo"], box["L"], 9.0, excl=excl, brute=False)
    assert_same_lists(gc, grows, oc, orows)


def test_graph_thermostat_turns(orc, gpu_engine_factory):
    """The device turn counter applies the rescale on the same turns as the
    eager loop: after 100 steps with every=5 the temperature is the target."""
    say("captured thermostat")
    box = systems.tip4p2005_box(216, seed=23, min_sep=1.8)
    v = start_velocities(orc, box, seed=3)
    e, flags = water_md_engine(gpu_engine_factory, box, v, rc=7.0, skin=1.0, sz=(32, 32, 32), order=5,
                               e3b=False, dt=1.0)
    kb, mvv = systems.REAL_BOLTZ, systems.REAL_MVV_TO_ENG
    e.thermostat(True, 350.0, 5, 0.0, kb, mvv)
    e.option("graph", 1)
    e.run(37, flags)
    e.run(63, flags)           # 100 % 5 == 0: ends right after a rescale
    T, _ = e.temperature(kb, mvv)
    assert T == pytest.approx(350.0, rel=1e-6)


@pytest.mark.parametrize("merge", [1, 0])
def test_batch_equals_single(orc, gpu_engine_factory, merge):
    """Three replicas of different sizes in one graph == each alone, with one
    launch per kernel across the replicas (merge 1) and with one stream per
    replica (merge 0)."""
    say(f"batched replicas vs single runs (merge {merge})")
    from md_engine_amd import run_batch

    # sizes whose replicas run the same kernels (replicas too small for the
    # tiled charge spread take the per-atom spread and are not merged)
    specs = [(512, 41), (729, 42), (1000, 43)]
    boxes = [systems.tip4p2005_box(m, seed=s, min_sep=1.8) for m, s in specs]
    vels = [start_velocities(orc, b, seed=s) for b, (_, s) in zip(boxes, specs)]

    def make(b, v):
        e, flags = water_md_engine(gpu_engine_factory, b, v, rc=6.5, skin=0.9, sz=(32, 32, 32), order=5,
                                   e3b=True, dt=2.0)
        e.option("graph", 1)
        return e, flags

    batch = [make(b, v) for b, v in zip(boxes, vels)]
    flags = batch[0][1]
    batch[0][0].option("batch_merge", merge)
    run_batch([e for e, _ in batch], 70, flags)
    assert batch[0][0].batch_merged() == bool(merge)
    for (eb, _), b, v in zip(batch, boxes, vels):
        es, fl = make(b, v)
        es.run(70, fl)
        xb, vb = eb.get_system()
        xs, vs = es.get_system()
        assert np.array_equal(xb, xs)
        assert np.array_equal(vb, vs)
        assert eb.nlist_stats()["rebuilds"] == es.nlist_stats()["rebuilds"]


def test_cached_graphs_across_calls(orc, gpu_engine_factory):
    """Step graphs are kept between mdx_run calls: 96 steps as one call, as
    twelve calls of 8 and as calls of odd lengths (the 1-step graph too) give
    the same trajectory bit for bit, also when the set-up changes between
    calls (the thermostat is switched on: the graphs are captured again),
    with the barostat moving the box."""
    say("cached step graphs across calls")
    box = systems.tip4p2005_box(343, seed=33, min_sep=1.8)
    v = start_velocities(orc, box, seed=7)
    out = []
    for pattern in ([48, 48], [8] * 6 + [48], [5, 13, 30, 3, 17, 28]):
        e, flags = water_md_engine(gpu_engine_factory, box, v, rc=8.0, skin=1.0, sz=(32, 32, 32), order=5,
                                   e3b=True, dt=2.0)
        e.barostat(True, 1.0, 500.0, 1, mvv_to_eng=systems.REAL_MVV_TO_ENG)
        e.option("graph", 1)
        done = 0
        for k in pattern:
            if done < 48 <= done + k:
                # split the call at step 48 so the set-up change lands there
                a = 48 - done
                if a:
                    e.run(a, flags)
                e.thermostat(True, 298.0, 5, 0.0, systems.REAL_BOLTZ, systems.REAL_MVV_TO_ENG)
                if k - a:
                    e.run(k - a, flags)
            else:
                e.run(k, flags)
            done += k
        assert done == 96
        x, vv = e.get_system()
        out.append((x, vv, e.get_box()[1]))
        e.close()
    for x, vv, L in out[1:]:
        assert np.array_equal(x, out[0][0])
        assert np.array_equal(vv, out[0][1])
        assert np.array_equal(L, out[0][2])


def test_batch_tuner_keeps_launches_merged(orc, gpu_engine_factory):
    """The batch autotuner times whole merged steps per candidate and leaves
    one choice on every replica: the batch still runs one launch per kernel
    afterwards, and the replicas keep their constraints."""
    say("batch autotuner")
    from md_engine_amd import run_batch, tune_batch

    specs = [(512, 51), (729, 52), (1000, 53)]
    boxes = [systems.tip4p2005_box(m, seed=s, min_sep=1.8) for m, s in specs]
    vels = [start_velocities(orc, b, seed=s) for b, (_, s) in zip(boxes, specs)]
    batch = []
    for b, v in zip(boxes, vels):
        e, flags = water_md_engine(gpu_engine_factory, b, v, rc=6.5, skin=0.9, sz=(32, 32, 32), order=5,
                                   e3b=True, dt=2.0)
        e.option("graph", 1)
        batch.append(e)
    choice, ms = tune_batch(batch, flags, steps=8)
    say(f"  chosen {choice}, {ms:.3f} ms per batch step")
    assert choice["build_team"] in (64, 8, 16) and choice["pair_team"] in (8, 16, 32)
    assert choice["e3b_team"] in (16, 32) and ms > 0
    run_batch(batch, 24, flags)
    assert batch[0].batch_merged()
    for e, b in zip(batch, boxes):
        x, _ = e.get_system()
        o, h1 = x[b["mol"][:, 0]], x[b["mol"][:, 1]]
        d = h1 - o
        d -= b["L"] * np.rint(d / b["L"])
        assert np.allclose(np.linalg.norm(d, axis=1), G["dOH"], atol=1e-8)
        e.close()


@pytest.mark.parametrize("barostat", [False, True])
def test_mesh_stream_matches_one_stream(orc, gpu_engine_factory, barostat):
    """The P3M chain on the side stream of a captured step (beside the pair
    and E3B kernels) gives the one-stream trajectory bit for bit: every term
    adds into fixed-point buffers."""
    say(f"mesh stream vs one stream, barostat {barostat}")
    box = systems.tip4p2005_box(343, seed=35, min_sep=1.8)
    v = start_velocities(orc, box, seed=9)
    out = []
    for side in (1, 0):
        e, flags = water_md_engine(gpu_engine_factory, box, v, rc=8.0, skin=1.0, sz=(32, 32, 32), order=5,
                                   e3b=True, dt=2.0)
        if barostat:
            e.barostat(True, 1.0, 500.0, 1, mvv_to_eng=systems.REAL_MVV_TO_ENG)
        e.option("graph", 1)
        e.option("mesh_stream", side)
        e.run(60, flags)
        x, vv = e.get_system()
        out.append((x, vv, e.get_box()[1], e.nlist_stats()["rebuilds"]))
        e.close()
    for x, vv, L, nb in out[1:]:
        assert np.array_equal(x, out[0][0])
        assert np.array_equal(vv, out[0][1])
        assert np.array_equal(L, out[0][2])
        assert nb == out[0][3]


def test_fused_step_launches_match_separate(orc, gpu_engine_factory):
    """The captured step's folded launches (step decisions and the position
    refresh in the first half kick, the cell scan in the binning, molecule
    map in the list build, row maxima in the pair launch, the mesh stream's
    fork after the rebuild's copy-back, the barostat step in the second half
    kick, the thermostat's rescale with the barostat's Green function) give
    the trajectory of separate launches bit for bit, rebuilds and box
    included."""
    say("fused step launches vs separate")
    box = systems.tip4p2005_box(343, seed=37, min_sep=1.8)
    v = start_velocities(orc, box, seed=11)
    out = []
    for fuse in (1, 0):
        e, flags = water_md_engine(gpu_engine_factory, box, v, rc=8.0, skin=1.0, sz=(32, 32, 32), order=5,
                                   e3b=True, dt=2.0)
        e.thermostat(True, 298.0, 5, 0.0, systems.REAL_BOLTZ, systems.REAL_MVV_TO_ENG)
        e.barostat(True, 1.0, 500.0, 1, mvv_to_eng=systems.REAL_MVV_TO_ENG)
        e.option("graph", 1)
        e.option("fuse_step", fuse)
        e.run(70, flags)
        x, vv = e.get_system()
        out.append((x, vv, e.get_box()[1], e.nlist_stats()["rebuilds"]))
        e.close()
    assert out[0][3] > 3
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])
    assert out[0][3] == out[1][3]
