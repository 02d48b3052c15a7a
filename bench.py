#!/usr/bin/env python3
"""Benchmark of the per-step force path on one MI355X.

Default workload: TIP4P/2005 + E3B3 water, 8,192 molecules (32,768 sites
counting M), cubic box at 0.997 g/cm^3 (L = 62.64 A), LJ O-O + Ewald real
space at rc = 9 A with a 1 A skin, P3M 64^3 order 5, E3B3 (5.0-5.2 A switch,
two-body O-O correction), rigid molecules (FP64 SETTLE / RATTLE), dt = 2 fs,
298 K velocity rescaling every 10 steps, 1 bar Berendsen barostat applied
every step (--barostat N: every N steps, 0: off).  One step = one MD step:
kick-drift and constraints, skin check and list rebuild when due, LJ + Ewald
real space, P3M spread/FFT/field/FFT/gather, E3B, kick and constraints, then
the barostat: molecular virial pressure, centre-of-mass dilation, alpha
re-solve and Green function for the new box.  The lattice start is
equilibrated (untimed) first; inputs are resident in HBM before timing.

Modes: md (default; --graph 1 replays captured 8-step hipGraphs with device-
side rebuild decisions), force (force path at fixed positions, list rebuilt
every --rebuild-every steps), batch (48 uneven TIP4P/2005 boxes of 512-4096
molecules advanced by one captured graph).  --three-body surrogate swaps the
E3B three-body term for the learned surrogate: it is trained on the GPU from
frames of the equilibration (untimed), the timed run uses it, and the line
reports its three-body force RMSE against E3B in kJ/mol/nm.

Prints one JSON line: atom-steps/s (sites x steps / s), ms per step, the
roofline of the dominant kernel (HIP events on the library stream during the
timed region) and an FP64 CPU baseline of the same formulas (the oracle,
OpenMP on the host cores) on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
E3B3_KJ = (453000.0, 150.0, -1005.0, 1880.0)
KJ_TO_KCAL = 0.23900573614
BAR_TO_ATM = 1.0 / 1.01325


def e3b3_params():
    E2, Ea, Eb, Ec = (v * KJ_TO_KCAL for v in E3B3_KJ)
    return np.array([E2, Ea, Eb, Ec, 4.872, 1.907, 5.0, 5.2])


def make_workload(name):
    from md_engine_amd import systems

    if name == "tip4p-e3b":
        box = systems.tip4p2005_box(8192, seed=20261202)
        return dict(box=box, rc=9.0, skin=1.0, sz=(64, 64, 64), order=5, e3b=True, dt_fs=2.0,
                    desc="TIP4P/2005+E3B3 water, 8192 molecules (32768 sites), L=%.3f A" % box["L"][0])
    if name == "tip4p":
        box = systems.tip4p2005_box(4096, seed=20261202)
        return dict(box=box, rc=9.0, skin=1.0, sz=(64, 64, 64), order=5, e3b=False, dt_fs=2.0,
                    desc="TIP4P/2005 water, 4096 molecules (16384 sites), L=%.3f A" % box["L"][0])
    if name == "lj":
        x, L, lo = systems.fcc_lj(20, 0.8442)
        box = dict(x=x, L=L, lo=lo, q=np.zeros(len(x)), type=np.zeros(len(x), np.int32),
                   mass=np.ones(len(x)), mol=None)
        return dict(box=box, rc=2.5, skin=0.3, sz=None, order=None, e3b=False, dt_fs=None,
                    desc="LJ fluid FCC 32000 atoms, rho*=0.8442")
    raise SystemExit(f"unknown config {name}")


def setup_engine(w):
    from md_engine_amd import E3B, PAIR, PPPM, Engine, systems

    box = w["box"]
    e = Engine(0)
    mass = box["mass"]
    inv = np.where(mass > 0, 1.0 / np.maximum(mass, 1e-30), 0.0)
    e.set_system(box["x"], box["lo"], box["L"], q=box["q"], types=box["type"], inv_mass=inv)
    flags = PAIR
    if box["mol"] is not None:
        e.exclusions_from_bonds(systems.water_bonds(box["mol"]))
        eps24, sig6 = systems.water_lj_tables()
        e.lj_set(np.full((3, 3), w["rc"] ** 2), eps24, sig6)
        e.pppm_set(w["sz"], w["order"], w["rc"], systems.REAL_QQR_TO_ENG)
        e.ewald_real_set(True, -1.0, systems.REAL_QQR_TO_ENG, w["rc"])
        flags |= PPPM
        if w["e3b"]:
            e.e3b_set(box["mol"], e3b3_params())
            flags |= E3B
    else:
        e.lj_set([[w["rc"] ** 2]], [[24.0]], [[1.0]])
    e.nlist_config(w["rc"] + w["skin"], w["skin"])
    return e, flags


def setup_md(e, w):
    """Rigid TIP4P/2005 MD: Maxwell-Boltzmann velocities at 298 K (seeded),
    SETTLE geometry, dt = 2 fs."""
    from md_engine_amd import systems

    box = w["box"]
    if box["mol"] is None:
        v = systems.maxwell_boltzmann(box["mass"], 1.2, seed=20261202)
        e.set_system(box["x"], box["lo"], box["L"], v=v, inv_mass=np.ones(len(v)), q=box["q"],
                     types=box["type"])
        e.integrator_set(0.005, 1.0)
        return
    p = systems.TIP4P2005
    v = systems.maxwell_boltzmann(box["mass"], 298.0, seed=20261202, boltz=systems.REAL_BOLTZ,
                                  mvv_to_eng=systems.REAL_MVV_TO_ENG)
    inv = np.where(box["mass"] > 0, 1.0 / np.maximum(box["mass"], 1e-30), 0.0)
    e.set_system(box["x"], box["lo"], box["L"], v=v, inv_mass=inv, q=box["q"], types=box["type"])
    e.rigid_set(box["mol"], p["m_O"], p["m_H"], p["r_OH"], p["r_HH"], p["gamma"])
    e.integrator_set(w["dt_fs"], systems.REAL_FTM_TO_V)
    # 298 K velocity rescaling every 10 steps (FixNVTRescale) during
    # equilibration and the timed run
    e.thermostat(True, 298.0, 10, 0.0, systems.REAL_BOLTZ, systems.REAL_MVV_TO_ENG)


def cpu_baseline(w, rebuild_every, budget_s=12.0):
    """FP64 oracle of the same step on the host cores (bounded sample)."""
    import oracle
    from md_engine_amd import systems

    box = w["box"]
    n = len(box["x"])
    L, lo, rc, skin = box["L"], box["lo"], w["rc"], w["skin"]
    x = oracle.wrap(box["x"], lo, L)
    t0 = time.perf_counter()
    counts, tab = oracle.nlist_cell(x, lo, L, rc + skin)
    t_build = time.perf_counter() - t0
    rows = [tab[i, : counts[i]] for i in range(n)]
    if box["mol"] is not None:
        excl = oracle.exclusions_bfs(n, systems.water_bonds(box["mol"]))
        rows = oracle.tag_rows(rows, np.arange(n), excl)
        eps24, sig6 = systems.water_lj_tables()
        rc2 = np.full((3, 3), rc * rc)
        qqr = systems.REAL_QQR_TO_ENG
        q2 = qqr * (box["q"] ** 2).sum()
        alpha, _ = oracle.pppm_alpha(n, q2, L, np.array(w["sz"]), w["order"], rc)
        G = oracle.pppm_green(L, np.array(w["sz"]), alpha, w["order"])
    else:
        rows = [r.astype(np.uint32) for r in rows]
        eps24, sig6, rc2 = np.array([[24.0]]), np.array([[1.0]]), np.array([[rc * rc]])
    rows = oracle.rows_to_table(counts, rows)

    def step():
        if box["mol"] is not None:
            oracle.pair(x, box["q"], box["type"], L, counts, rows, rc2, eps24, sig6, with_charge=True,
                        alpha=alpha, qqr=qqr, rq=rc)
            oracle.pppm(x, box["q"], lo, L, np.array(w["sz"]), w["order"], qqr, alpha, q2, G)
            if w["e3b"]:
                oracle.e3b(box["mol"], x, L)
        else:
            oracle.pair(x, box["q"], box["type"], L, counts, rows, rc2, eps24, sig6)

    step()  # warm
    steps = 0
    t0 = time.perf_counter()
    while True:
        step()
        steps += 1
        if time.perf_counter() - t0 > budget_s or steps >= 50:
            break
    t_step = (time.perf_counter() - t0) / steps
    per_step = t_step + (t_build / rebuild_every if rebuild_every else 0.0)
    return dict(value=n / per_step, unit="atom-steps/s", cores=oracle.num_threads(), kind="port",
                sample=f"{steps} force-path steps of the same system + 1 list build amortised over "
                       f"{rebuild_every} steps; FP64 oracle (C, OpenMP); {t_step * 1e3:.1f} ms/step, "
                       f"build {t_build * 1e3:.1f} ms")


def copy_probe(mib=256, reps=5):
    """Plain device-to-device copy of a known size, for scaling the FETCH_SIZE /
    WRITE_SIZE counters to bytes (profiles/ records the factor)."""
    import torch

    n = mib * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read_per_launch": n * 4, "bytes_written_per_launch": n * 4,
                      "launches": reps}))


def train_surrogate(e, w, flags, frames=20, every=100, batch=1024,
                    schedule=((100, 1e-3), (40, 3e-4), (40, 1e-4), (20, 3e-5))):
    """Record `frames` frames of the thermostatted run (every `every` steps)
    and fit the surrogate on the GPU (Adam, mini-batches of `batch`
    molecules, `schedule` = (epochs, learning rate) stages); returns fit
    statistics."""
    from md_engine_amd import SURROGATE  # noqa: F401
    from md_engine_amd._lib import Engine

    nm = len(w["box"]["mol"])
    t0 = time.perf_counter()
    for k in range(frames):
        e.run(every, flags)
        e.surrogate_record(reset=(k == 0), capacity=frames * nm)
    rng = np.random.default_rng(7)
    p = _init_surrogate(rng)
    e.surrogate_set(p)
    l0 = e.surrogate_fit(1, batch=8192, lr=1e-3, seed=1, normalize=True)
    loss = l0
    for k, (ep, lr) in enumerate(schedule):
        loss = e.surrogate_fit(ep, batch=batch, lr=lr, seed=2 + k)
    return dict(frames=frames, samples=frames * nm, epochs=1 + sum(ep for ep, _ in schedule), batch=batch,
                loss_first=l0, loss_final=loss, train_s=time.perf_counter() - t0, model_params=Engine.SUR_NPARAM)


def _init_surrogate(rng):
    nf, nh = 48, 64
    parts = [rng.normal(0, 1 / np.sqrt(nf), nf * nh), np.zeros(nh), rng.normal(0, 1 / np.sqrt(nh), nh * nh),
             np.zeros(nh), rng.normal(0, 1 / np.sqrt(nh), nh), [0.0], np.zeros(nf), np.ones(nf), [1.0, 0.0]]
    return np.concatenate([np.asarray(a, np.float64) for a in parts]).astype(np.float32)


def three_body_rmse(e):
    """RMSE over atoms and components of surrogate vs E3B forces (the two-body
    term is identical), in kcal/mol/A and kJ/mol/nm."""
    from md_engine_amd import E3B, SURROGATE

    e.compute(E3B)
    fe = e.forces()
    e.compute(SURROGATE)
    fs = e.forces()
    rms = float(np.sqrt(((fs - fe) ** 2).mean()))
    return rms, rms * 41.84


def batch_workload(seed=20261202, nbox=48):
    from md_engine_amd import systems

    rng = np.random.default_rng(seed)
    sizes = rng.integers(512, 4097, nbox)
    return [systems.tip4p2005_box(int(m), seed=int(seed + k)) for k, m in enumerate(sizes)]


def pimd_ring(nbeads=32, nmol=1024, seed=20261203):
    """The C5 ring: nbeads beads of nmol-molecule flexible q-TIP4P/F water at
    298 K (bead temperature nbeads x 298 K), beads sampled from the free ring
    polymer about a lattice start, velocity rescaling every 10 steps per bead."""
    from md_engine_amd import BONDED, PAIR, PPPM, systems
    from md_engine_amd._lib import ring_omega

    box = systems.tip4p2005_box(nmol, seed=seed)
    p = systems.QTIP4PF
    tb = nbeads * 298.0
    xs = systems.ring_beads(box["x"], box["mass"], nbeads, tb, seed)
    inv = np.where(box["mass"] > 0, 1.0 / np.maximum(box["mass"], 1e-30), 0.0)
    w = dict(box=box, rc=9.0, skin=1.0, sz=(32, 32, 32), order=5, e3b=False, dt_fs=0.5, desc="")
    es = []
    for k, x in enumerate(xs):
        e, _ = setup_engine(dict(w, box=dict(box, x=x)))
        v = systems.maxwell_boltzmann(box["mass"], tb, seed=seed + k, boltz=systems.REAL_BOLTZ,
                                      mvv_to_eng=systems.REAL_MVV_TO_ENG)
        e.set_system(x, box["lo"], box["L"], v=v, inv_mass=inv, q=box["q"], types=box["type"])
        e.flexible_water_set(box["mol"], p["k2"], p["k3"], p["k4"], p["r0"], p["ktheta"], p["theta0"], p["gamma"])
        e.integrator_set(0.5, systems.REAL_FTM_TO_V)
        e.thermostat(True, tb, 10, 0.0, systems.REAL_BOLTZ, systems.REAL_MVV_TO_ENG)
        es.append(e)
    return es, PAIR | PPPM | BONDED, ring_omega(tb, systems.REAL_BOLTZ, systems.REAL_HBAR), len(box["x"])


def run_batch_mode(args):
    from md_engine_amd import E3B, PAIR, PPPM, run_batch, systems

    boxes = batch_workload()
    engines = []
    flags = PAIR | PPPM | E3B
    for b in boxes:
        w = dict(box=b, rc=9.0, skin=1.0, sz=(32, 32, 32) if b["L"][0] < 40 else (64, 64, 64), order=5,
                 e3b=True, dt_fs=2.0, desc="")
        e, fl = setup_engine(w)
        setup_md(e, w)
        engines.append(e)
    n = sum(len(b["x"]) for b in boxes)
    run_batch(engines, args.equil, flags)
    run_batch(engines, args.warmup, flags)
    for e in engines:
        e.sync()
    t0 = time.perf_counter()
    run_batch(engines, args.steps, flags)
    for e in engines:
        e.sync()
    wall = time.perf_counter() - t0
    # the 32-bead PIMD ring, its own captured graphs
    from md_engine_amd._lib import run_ring

    beads, rflags, wp, nsite = pimd_ring()
    run_ring(beads, max(args.equil // 10, 20), rflags, wp)
    run_ring(beads, args.warmup, rflags, wp)
    for e in beads:
        e.sync()
    t1 = time.perf_counter()
    run_ring(beads, args.steps, rflags, wp)
    for e in beads:
        e.sync()
    wall_r = time.perf_counter() - t1
    nring = nsite * len(beads)
    out = {
        "metric": "atom-steps/s", "value": (n + nring) * args.steps / (wall + wall_r), "unit": "atom-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": (wall + wall_r) / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 pair/3-body arithmetic, f64 positions, i64 fixed-point accumulation",
        "data": "synthetic (seeded lattices, random orientations)",
        "config": {"workload": "48 TIP4P/2005+E3B3 boxes, %d-%d molecules (seeded uniform), %d sites in all"
                               % (min(len(b["mol"]) for b in boxes), max(len(b["mol"]) for b in boxes), n),
                   "step": "one captured hipGraph per 8 MD steps holding every replica; the ring in its own "
                           "captured graphs (value: both sets of sites over both wall times)",
                   "boxes_atom_steps_per_s": n * args.steps / wall,
                   "pimd_ring": {"beads": len(beads), "molecules": nsite // 4, "bead_sites": nring,
                                 "model": "flexible q-TIP4P/F, 298 K (bead temperature 32 x 298 K), dt 0.5 fs",
                                 "ms_per_step": wall_r / args.steps * 1e3,
                                 "bead_site_steps_per_s": nring * args.steps / wall_r}},
    }
    for e in engines + beads:
        e.close()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--config", default="tip4p-e3b", choices=["tip4p-e3b", "tip4p", "lj"])
    ap.add_argument("--rebuild-every", type=int, default=10,
                    help="force mode: forced list rebuild interval")
    ap.add_argument("--mode", default="md", choices=["md", "force", "batch"],
                    help="md: rigid-water MD steps; force: force path at fixed positions; batch: 48 replicas")
    ap.add_argument("--graph", type=int, default=1, help="md mode: 1 = captured hipGraph steps")
    ap.add_argument("--three-body", default="e3b", choices=["e3b", "surrogate"])
    ap.add_argument("--equil", type=int, default=3000,
                    help="md mode: untimed thermostatted steps before warm-up (lattice start)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--pair-team", type=int, default=0)
    ap.add_argument("--build-team", type=int, default=0)
    ap.add_argument("--e3b-team", type=int, default=0)
    ap.add_argument("--reach-xy", type=int, default=0, help="cell stencil reach in x/y (0: library default)")
    ap.add_argument("--reach-z", type=int, default=0, help="cell stencil reach in z (0: library default)")
    ap.add_argument("--gather-team", type=int, default=0, help="lanes per charged atom in the P3M gather")
    ap.add_argument("--barostat", type=int, default=1,
                    help="md mode, water: 1 bar Berendsen barostat applied every N steps (0: off)")
    ap.add_argument("--tune", action="store_true", help="run the autotuner before timing")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain device copy of known size (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for --gpus {args.gpus}); this build has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)
    if args.mode == "batch":
        run_batch_mode(args)
        return

    from md_engine_amd import E3B, SURROGATE

    w = make_workload(args.config)
    e, flags = setup_engine(w)
    e.teams_set(args.build_team, args.pair_team, args.e3b_team)
    if args.reach_xy:
        e.option("cell_reach_xy", args.reach_xy)
    if args.reach_z:
        e.option("cell_reach_z", args.reach_z)
    if args.gather_team:
        e.option("gather_team", args.gather_team)
    n = len(w["box"]["x"])
    R = args.rebuild_every
    surrogate = args.three_body == "surrogate"
    if surrogate and not (flags & E3B):
        raise SystemExit("--three-body surrogate needs the tip4p-e3b config")

    md = args.mode == "md"
    train = None
    if md:
        setup_md(e, w)
        if args.graph:
            e.option("graph", 1)
        if args.config != "tip4p-e3b":
            args.barostat = 0  # the pressure-coupled workload is the E3B water box
        if args.barostat:
            from md_engine_amd import systems as _sy

            # FixPressureBerendsen(state, handle, 1 bar, period 1 ps, applyEvery)
            e.barostat(True, BAR_TO_ATM, 1000.0, args.barostat, mvv_to_eng=_sy.REAL_MVV_TO_ENG)
        step = lambda k: e.run(k, flags)  # noqa: E731
        if args.equil:
            step(args.equil)
        if surrogate:
            train = train_surrogate(e, w, flags)
            flags = (flags & ~E3B) | SURROGATE
    else:
        step = lambda k: e.force_steps(k, flags, R)  # noqa: E731
    tuned = e.tune(flags, evals=10) if args.tune else None
    step(args.warmup)
    e.sync()
    rebuilds0 = e.nlist_stats()["rebuilds"]
    e.profile(True)
    t0 = time.perf_counter()
    step(args.steps)
    e.sync()
    t1 = time.perf_counter()
    wall = t1 - t0
    prof = {k: e.profile_get(k) for k in ("pair", "build", "spread", "gather", "fft", "field", "e3b1", "e3b2")}
    e.profile(False)
    timing_pass = "timed region"
    if md and args.graph:
        # kernels inside a captured graph carry no events: time the same
        # kernels over an eager pass of the same state right after
        st0 = e.nlist_stats()
        e.option("graph", 0)
        e.profile(True)
        nprof = min(args.steps, 200)
        step(nprof)
        e.sync()
        prof = {k: e.profile_get(k) for k in prof}
        e.profile(False)
        timing_pass = "eager pass of %d steps right after the timed graph run" % nprof
        rebuilds0 += e.nlist_stats()["rebuilds"] - st0["rebuilds"]
    pairs, mols = e.nlist_totals()
    read_entries, read_atoms = e.pair_entries_read()
    st = e.nlist_stats()
    rmse = three_body_rmse(e) if surrogate else None

    ms_step = wall / args.steps * 1e3
    value = n * args.steps / wall
    # dominant per-step kernel: pair (LJ + Ewald real space).  Algorithmic
    # bytes per launch: per atom xq 16 B + type 4 B + two row counts 8 B, per
    # atom with work a fixed-point force row 24 B, plus 4 B per list entry the
    # kernel reads (charged partners for charged atoms, charge-free partners for
    # atoms with LJ to them; the other part of a row is never touched).
    cnt, tot_ms = prof["pair"]
    avg_s = tot_ms / max(cnt, 1) * 1e-3
    bytes_launch = n * (16 + 4 + 8) + 24 * read_atoms + 4 * read_entries
    achieved = bytes_launch / avg_s / 1e9
    kernel_ms = {k: (v[1] / max(v[0], 1)) for k, v in prof.items() if v[0]}
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "pair_traffic_tip4p_e3b_8192.json")
    if args.config == "tip4p-e3b" and os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("bytes_per_launch")

    out = {
        "metric": "atom-steps/s",
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair/3-body arithmetic, f64 positions, i64 fixed-point accumulation",
        "data": "synthetic (seeded lattice, random orientations)",
        "config": {
            "workload": w["desc"],
            "step": ("MD step: velocity Verlet + FP64 SETTLE/RATTLE + M site, skin-driven list rebuilds, "
                     "LJ+Ewald real, P3M, %s%s%s" % ("E3B two-body + learned three-body" if surrogate else "E3B",
                                                    (", 1 bar Berendsen barostat every %d steps" % args.barostat)
                                                    if args.barostat else ", NVT (no barostat)",
                                                    ", captured hipGraph (8 steps)" if args.graph else "")
                     if md else "force path: list rebuild every %d steps, LJ+Ewald real, P3M, E3B" % R),
            "three_body": args.three_body,
            "surrogate_training": train,
            "three_body_force_rmse_vs_e3b": ({"kcal_mol_A": rmse[0], "kJ_mol_nm": rmse[1]} if rmse else None),
            "tuned": tuned,
            "list_rebuilds_timed": e.nlist_stats()["rebuilds"] - rebuilds0,
            "list_violations": e.nlist_stats()["violations"],
            "rc": w["rc"], "skin": w["skin"], "pppm_mesh": w["sz"], "pppm_order": w["order"],
            "e3b": w["e3b"], "sites": n, "pair_list_entries": pairs, "pair_entries_read": read_entries,
            "e3b_list_entries": mols,
            "list_rebuilds_in_run": st["rebuilds"],
            "ns_per_day_at_dt": (w["dt_fs"] * 1e-6 * args.steps / wall * 86400.0) if w["dt_fs"] else None,
            "kernel_avg_ms": kernel_ms,
            "kernel_timing": timing_pass,
        },
        "roofline": {
            "kernel": "k_pair (LJ + Ewald real space)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE, profiles/force_path_tip4p_e3b_8192.md"
            if traffic else None,
            "bytes_per_launch": bytes_launch,
            "avg_launch_ms": avg_s * 1e3,
            "launches": cnt,
        },
    }
    if not args.no_cpu:
        # in md mode the CPU sample amortises one build over the GPU run's
        # measured rebuild interval
        rb = out["config"]["list_rebuilds_timed"]
        out["cpu_baseline"] = cpu_baseline(w, max(1, round(args.steps / rb)) if md and rb else R)
    e.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
