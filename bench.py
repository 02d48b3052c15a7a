#!/usr/bin/env python3
"""Benchmark of the per-step force path on one MI355X.

Headline workload (the metric's configuration): TIP4P/2005 + E3B3 water,
8,192 molecules (32,768 sites counting M), cubic box at 0.997 g/cm^3
(L = 62.64 A), LJ O-O + Ewald real space at rc = 9 A with a 1 A skin, P3M
64^3 order 5, E3B3 (5.0-5.2 A switch, two-body O-O correction), rigid
molecules (FP64 SETTLE / RATTLE), dt = 2 fs, 298 K velocity rescaling every
10 steps, 1 bar Berendsen barostat applied every step.  One step = one MD
step: kick-drift and constraints, skin check and list rebuild when due, LJ +
Ewald real space, P3M spread/FFT/field/FFT/gather, E3B, kick and
constraints, then the barostat (molecular virial pressure, centre-of-mass
dilation, alpha re-solve and Green function for the new box).  Steps run as
captured 8-step hipGraphs replayed from a cache (captured during warm-up).
The lattice start is equilibrated (untimed); inputs are resident in HBM
before timing.

The one JSON line holds both halves of the metric, "plain E3B vs MLP
three-body at equal force RMSE":
  value            atom-steps/s of the E3B step (sites x steps / s);
  config.surrogate the same step with the learned three-body term (trained
                   on the GPU from frames of this run, untimed): atom-steps/s,
                   ns/day, three-body force RMSE against E3B (kJ/mol/nm) and
                   the E3B judge's checks every 100 steps;
  config.parity    force RMSE of the final timed frame against the FP64
                   oracle, NVE energy drift per ns, list-staleness violations;
  config.c1 / c2   LJ 32,000 atoms NVE and 4,096-molecule TIP4P/2005 NVT;
  roofline         the pair kernel's algorithmic bytes / its mean duration
                   (HIP events on the library stream, eager pass);
  cpu_baseline     the FP64 oracle of the same step on the host cores, one
                   thread and all cores, on a bounded sample.

Modes: md (default), force (force path at fixed positions, list rebuilt
every --rebuild-every steps), batch (48 uneven TIP4P/2005 boxes of 512-4096
molecules in one captured graph, plus the 32-bead q-TIP4P/F ring).
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
KCAL_A_TO_KJ_NM = 41.84
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
                    desc="LJ fluid FCC 32000 atoms, rho*=0.8442, T*=1.2")
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


def setup_md(e, w, thermostat="rescale"):
    """Velocities from the seeded Maxwell-Boltzmann recipe; rigid TIP4P/2005
    (SETTLE geometry, dt 2 fs) with a 298 K thermostat, or LJ NVE at dt 0.005."""
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
    if thermostat == "rescale":
        # 298 K velocity rescaling every 10 steps (FixNVTRescale)
        e.thermostat(True, 298.0, 10, 0.0, systems.REAL_BOLTZ, systems.REAL_MVV_TO_ENG)
    elif thermostat == "nose-hoover":
        e.nose_hoover(True, 298.0, 100.0, 10, systems.REAL_BOLTZ, systems.REAL_MVV_TO_ENG)


def timed(e, flags, steps, warmup):
    """warm-up (captures the step graphs), then exactly `steps` steps between
    two stream synchronisations; returns the wall time."""
    e.run(warmup, flags)
    e.sync()
    t0 = time.perf_counter()
    e.run(steps, flags)
    e.sync()
    return time.perf_counter() - t0


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(w, x, v, lo, L, rebuild_every, gpu_forces, one_thread_s=12.0, all_core_s=12.0):
    """FP64 oracle of the same MD step on the host cores, from the GPU run's
    final frame: kick-drift + SETTLE, LJ + Ewald real space over a cell/Verlet
    list, P3M, E3B over an oxygen cell grid, kick + RATTLE, with one list
    build per `rebuild_every` steps (the GPU run's measured rebuild interval).
    One thread and all threads, each on a bounded number of steps.  The
    oracle's forces at the frame also give the GPU forces' RMSE against FP64."""
    import oracle
    from md_engine_amd import systems

    # the barostat moved the box (it dilates about the centre: lo moves too)
    box = dict(w["box"], lo=np.asarray(lo, np.float64), L=np.asarray(L, np.float64))
    n = len(x)
    lo, rc, skin = box["lo"], w["rc"], w["skin"]
    qqr = systems.REAL_QQR_TO_ENG
    q2 = oracle.total_q2(box["q"], qqr)
    sz = np.array(w["sz"])
    eps24, sig6 = systems.water_lj_tables()
    rc2 = np.full((3, 3), rc * rc)
    excl = oracle.exclusions_bfs(n, systems.water_bonds(box["mol"]))
    alpha, _ = oracle.pppm_alpha(n, q2, box["L"], sz, w["order"], rc)
    G = oracle.pppm_green(box["L"], sz, alpha, w["order"])
    ftm = systems.REAL_FTM_TO_V
    nthreads = oracle.num_threads()

    def build(xx):
        xw = oracle.wrap(xx, lo, box["L"])
        counts, tab = oracle.nlist_cell(xw, lo, box["L"], rc + skin)
        rows = oracle.tag_rows([tab[i, : counts[i]] for i in range(n)], np.arange(n), excl)
        return counts, oracle.rows_to_table(counts, rows)

    def forces(xx, lst):
        xw = oracle.wrap(xx, lo, box["L"])
        fp, _, _ = oracle.pair(xw, box["q"], box["type"], box["L"], lst[0], lst[1], rc2, eps24, sig6,
                               with_charge=True, alpha=alpha, qqr=qqr, rq=rc)
        fk = oracle.pppm(xx, box["q"], lo, box["L"], sz, w["order"], qqr, alpha, q2, G)["f"]
        f = fp + fk
        if w["e3b"]:
            f = f + oracle.e3b(box["mol"], xx, box["L"])[0]
        return f

    # parity of the final GPU frame (the oracle as the checker)
    lst = build(x)
    f_ref = forces(x, lst)
    rmse = float(np.sqrt(((gpu_forces - f_ref) ** 2).sum() / (f_ref ** 2).sum()))

    def leg(threads, budget):
        oracle.set_threads(threads)
        xx, vv = x.copy(), v.copy()
        t0 = time.perf_counter()
        lst = build(xx)
        t_build = time.perf_counter() - t0
        f = forces(xx, lst)
        steps = 0
        t0 = time.perf_counter()
        while True:
            oracle.water_step_half(1, box["mol"], xx, vv, f, w["dt_fs"], ftm, box["L"])
            f = forces(xx, lst)
            oracle.water_step_half(2, box["mol"], xx, vv, f, w["dt_fs"], ftm, box["L"])
            steps += 1
            if time.perf_counter() - t0 > budget or steps >= 50:
                break
        t_step = (time.perf_counter() - t0) / steps
        per_step = t_step + t_build / max(rebuild_every, 1)
        return dict(atom_steps_per_s=n / per_step, ms_per_step=per_step * 1e3, steps=steps,
                    build_ms=t_build * 1e3, threads=threads,
                    ns_per_day=w["dt_fs"] * 1e-6 / per_step * 86400.0)

    one = leg(1, one_thread_s)
    allc = leg(nthreads, all_core_s)
    oracle.set_threads(nthreads)
    base = dict(value=allc["atom_steps_per_s"], unit="atom-steps/s", cores=nthreads, kind="port",
                sample=("FP64 oracle (C, OpenMP) of the full MD step of the same 32,768-site system from the GPU "
                        "run's final frame: kick-drift + SETTLE, LJ + Ewald real over the Verlet list, P3M 64^3, "
                        "E3B over an oxygen cell grid, kick + RATTLE, one list build per %d steps; %d steps on "
                        "%d threads (%.0f ms/step), %d steps on 1 thread (%.0f ms/step)"
                        % (rebuild_every, allc["steps"], nthreads, allc["ms_per_step"], one["steps"],
                           one["ms_per_step"])),
                one_thread=one, all_cores=allc, cpu_model=cpu_model(), nproc=os.cpu_count())
    return base, rmse


def nve_drift(e, flags, steps=2000, every=200):
    """Energy drift of plain NVE from the current state (thermostat and
    barostat off): slope of the total energy over the segment, per ns and
    per molecule (kcal/mol), and in kT."""
    from md_engine_amd import ENERGY, systems

    kb, mvv = systems.REAL_BOLTZ, systems.REAL_MVV_TO_ENG
    e.thermostat(False)
    e.barostat(False)

    def total():
        e.compute(flags | ENERGY)
        en, _ = e.energies()
        _, ke = e.temperature(kb, mvv)
        return float(en.sum() + ke)

    t, h = [0.0], [total()]
    for k in range(steps // every):
        e.run(every, flags)
        t.append((k + 1) * every)
        h.append(total())
    t = np.asarray(t)
    h = np.asarray(h)
    slope = np.polyfit(t, h, 1)[0]  # kcal/mol per step
    return slope, h


def train_surrogate(e, w, flags, frames=20, every=100, batch=1024,
                    schedule=((100, 1e-3), (40, 3e-4), (40, 1e-4), (20, 3e-5))):
    """Record `frames` frames of the thermostatted run (every `every` steps)
    and fit the surrogate on the GPU (Adam, mini-batches of `batch`
    molecules, `schedule` = (epochs, learning rate) stages); returns fit
    statistics."""
    from md_engine_amd._lib import Engine

    nm = len(w["box"]["mol"])
    t0 = time.perf_counter()
    for k in range(frames):
        e.run(every, flags)
        e.surrogate_record(reset=(k == 0), capacity=frames * nm)
    rng = np.random.default_rng(7)
    e.surrogate_set(_init_surrogate(rng))
    l0 = e.surrogate_fit(1, batch=8192, lr=1e-3, seed=1, normalize=True)
    loss = l0
    for k, (ep, lr) in enumerate(schedule):
        loss = e.surrogate_fit(ep, batch=batch, lr=lr, seed=2 + k)
    return dict(frames=frames, samples=frames * nm, epochs=1 + sum(ep for ep, _ in schedule), batch=batch,
                loss_first=l0, loss_final=loss, train_s=time.perf_counter() - t0, model_params=Engine.SUR_NPARAM)


def _init_surrogate(rng):
    from md_engine_amd._lib import SUR_NF, SUR_NH

    nf, nh = SUR_NF, SUR_NH
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
    return rms, rms * KCAL_A_TO_KJ_NM


def sub_result(config, steps, warmup, equil, thermostat):
    """atom-steps/s of a secondary configuration (C1, C2)."""
    w = make_workload(config)
    e, flags = setup_engine(w)
    setup_md(e, w, thermostat=thermostat)
    e.option("graph", 1)
    if equil:
        e.run(equil, flags)
    wall = timed(e, flags, steps, warmup)
    n = len(w["box"]["x"])
    dt = w["dt_fs"] if w["dt_fs"] else None
    out = dict(workload=w["desc"], atom_steps_per_s=n * steps / wall, ms_per_step=wall / steps * 1e3,
               steps=steps, list_violations=e.nlist_stats()["violations"],
               ns_per_day=(dt * 1e-6 * steps / wall * 86400.0) if dt else None,
               tau_per_day=(0.005 * steps / wall * 86400.0) if not dt else None)
    e.close()
    return out


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
    from md_engine_amd import E3B, PAIR, PPPM, run_batch

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
    # team widths and list-build kernel chosen by timing whole merged batch steps
    from md_engine_amd import tune_batch

    tuned, tuned_ms = tune_batch(engines, flags, steps=16)
    run_batch(engines, args.warmup, flags)
    for e in engines:
        e.sync()
    t0 = time.perf_counter()
    run_batch(engines, args.steps, flags)
    for e in engines:
        e.sync()
    wall = time.perf_counter() - t0
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
                   "step": "one captured hipGraph per 8 MD steps holding every replica, one launch per kernel "
                           "across the replicas; the ring likewise in its own captured graphs, one launch per "
                           "kernel across the beads (value: both sets of sites over both wall times)",
                   "boxes_atom_steps_per_s": n * args.steps / wall,
                   "merged_launches": bool(engines[0].batch_merged()),
                   "tuned": dict(tuned, ms_per_step=tuned_ms),
                   "pimd_ring": {"beads": len(beads), "molecules": nsite // 4, "bead_sites": nring,
                                 "model": "flexible q-TIP4P/F, 298 K (bead temperature 32 x 298 K), dt 0.5 fs",
                                 "ms_per_step": wall_r / args.steps * 1e3,
                                 "merged_launches": bool(beads[0].batch_merged()),
                                 "bead_site_steps_per_s": nring * args.steps / wall_r}},
    }
    for e in engines + beads:
        e.close()
    print(json.dumps(out))


def pair_roofline(e, n, prof):
    """Algorithmic bytes of one pair launch over its mean duration: per atom
    xq 16 B + type 4 B + two row counts 8 B, per atom with work a fixed-point
    force row 24 B, plus 4 B per list entry the kernel reads (charged partners
    for charged atoms, charge-free partners for atoms with LJ to them; the
    other part of a row is never touched)."""
    read_entries, read_atoms = e.pair_entries_read()
    cnt, tot_ms = prof["pair"]
    avg_s = tot_ms / max(cnt, 1) * 1e-3
    bytes_launch = n * (16 + 4 + 8) + 24 * read_atoms + 4 * read_entries
    achieved = bytes_launch / avg_s / 1e9
    traffic = valu_busy = None
    tpath = os.path.join(ROOT, "profiles", "pair_traffic_tip4p_e3b_8192.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            rec = json.load(fh)
        traffic, valu_busy = rec.get("bytes_per_launch"), rec.get("valu_busy")
    return {
        "kernel": "k_pair (LJ + Ewald real space)", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
        "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE, profiles/force_path_tip4p_e3b_8192.md" if traffic else None,
        # fraction of SIMD cycles issuing vector ALU work (SQ_ACTIVE_INST_VALU):
        # the kernel is limited by instruction issue, not by HBM
        "valu_busy": valu_busy,
        "bytes_per_launch": bytes_launch, "avg_launch_ms": avg_s * 1e3, "launches": cnt,
    }, read_entries


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--config", default="tip4p-e3b", choices=["tip4p-e3b", "tip4p", "lj"])
    ap.add_argument("--rebuild-every", type=int, default=10, help="force mode: forced list rebuild interval")
    ap.add_argument("--mode", default="md", choices=["md", "force", "batch"],
                    help="md: MD steps; force: force path at fixed positions; batch: 48 replicas + PIMD ring")
    ap.add_argument("--graph", type=int, default=1, help="md mode: 1 = captured hipGraph steps")
    ap.add_argument("--equil", type=int, default=3000,
                    help="md mode: untimed thermostatted steps before warm-up (lattice start)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the FP64 parity frame")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned three-body leg")
    ap.add_argument("--no-sub", action="store_true", help="skip the C1 / C2 sub-results")
    ap.add_argument("--drift-steps", type=int, default=2500, help="NVE segment for the energy drift (0: skip)")
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
        print(f"bench.py runs on exactly one GPU (asked for --gpus {args.gpus}): every GPU box here has one "
              "GPU, and this build's multi-GPU sharding is a design only", file=sys.stderr)
        sys.exit(2)
    if args.mode == "batch":
        run_batch_mode(args)
        return

    from md_engine_amd import E3B, SURROGATE, systems

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
    md = args.mode == "md"
    water = w["box"]["mol"] is not None
    headline = md and args.config == "tip4p-e3b"
    if md:
        setup_md(e, w)
        e.option("graph", args.graph)
        if not headline:
            args.barostat = 0  # the pressure-coupled workload is the E3B water box
        if args.barostat:
            # FixPressureBerendsen(state, handle, 1 bar, period 1 ps, applyEvery)
            e.barostat(True, BAR_TO_ATM, 1000.0, args.barostat, mvv_to_eng=systems.REAL_MVV_TO_ENG)
        if args.equil:
            e.run(args.equil, flags)
        step = lambda k: e.run(k, flags)  # noqa: E731
    else:
        step = lambda k: e.force_steps(k, flags, args.rebuild_every)  # noqa: E731
    tuned = e.tune(flags, evals=10) if args.tune else None

    # ---- timed region: the E3B step ------------------------------------
    step(args.warmup)
    e.sync()
    st0 = e.nlist_stats()
    t0 = time.perf_counter()
    step(args.steps)
    e.sync()
    wall = time.perf_counter() - t0
    st1 = e.nlist_stats()
    rebuilds_timed = st1["rebuilds"] - st0["rebuilds"]
    x_final, v_final = e.get_system()
    lo_final, L_final = (np.asarray(a) for a in e.get_box())
    if headline and not args.no_cpu:
        e.compute(flags)
        f_final = e.forces()

    # ---- per-kernel times: an eager pass of the same kernels -------------
    # (kernels inside a captured graph carry no events; the eager pass runs
    # right after the timed region from the same state)
    kernels = ("pair", "build", "spread", "gather", "fft", "field", "e3b1", "e3b2")
    nprof = min(args.steps, 200)
    if md and args.graph:
        e.option("graph", 0)
    e.profile(True)
    step(nprof)
    e.sync()
    prof = {k: e.profile_get(k) for k in kernels}
    e.profile(False)
    if md and args.graph:
        e.option("graph", 1)
    kernel_ms = {k: (v[1] / max(v[0], 1)) for k, v in prof.items() if v[0]}
    roof, read_entries = pair_roofline(e, n, prof)
    pairs, mols = e.nlist_totals()

    value = n * args.steps / wall
    dt = w["dt_fs"]
    config = {
        "workload": w["desc"],
        "step": ("MD step: velocity Verlet + FP64 SETTLE/RATTLE + M site, skin-driven list rebuilds, LJ+Ewald "
                 "real, P3M, %s%s%s" % ("E3B" if w["e3b"] else "no E3B",
                                        (", 1 bar Berendsen barostat every %d steps" % args.barostat)
                                        if args.barostat else "",
                                        ", captured hipGraphs (8 steps) replayed from the cache" if args.graph else "")
                 if md else "force path: list rebuild every %d steps, LJ+Ewald real, P3M, E3B" % args.rebuild_every),
        "ns_per_day": (dt * 1e-6 * args.steps / wall * 86400.0) if (md and dt) else None,
        "tuned": tuned,
        "list_rebuilds_timed": rebuilds_timed,
        "rc": w["rc"], "skin": w["skin"], "pppm_mesh": w["sz"], "pppm_order": w["order"], "e3b": w["e3b"],
        "sites": n, "pair_list_entries": pairs, "pair_entries_read": read_entries, "e3b_list_entries": mols,
        "kernel_avg_ms": kernel_ms,
        "kernel_timing": "HIP events on the library stream, eager pass of %d steps right after the timed run" % nprof,
    }
    out = {
        "metric": "atom-steps/s", "value": value, "unit": "atom-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 pair/3-body arithmetic, f64 positions and integration, i64 fixed-point accumulation",
        "data": "synthetic (seeded lattice, random orientations)", "config": config, "roofline": roof,
    }

    if headline:
        parity = {"list_violations": e.nlist_stats()["violations"]}
        # ---- the learned three-body leg (trained untimed on this run) -----
        if not args.no_surrogate:
            from md_engine_amd._lib import SUR_NB, SUR_NF

            e.option("graph", 1)
            train = train_surrogate(e, w, flags)
            sflags = (flags & ~E3B) | SURROGATE
            e.surrogate_judge(100, 0.5 / KCAL_A_TO_KJ_NM)
            wall_s = timed(e, sflags, args.steps, args.warmup)
            judge = e.surrogate_judge_stats()
            rms = three_body_rmse(e)
            # kernel times of the three-body forms, eager passes from the same state
            e.option("graph", 0)
            kt = {}
            for name, fl in (("e3b", flags), ("surrogate", sflags)):
                e.profile(True)
                e.run(min(args.steps, 100), fl)
                e.sync()
                kt[name] = {k: v[1] / max(v[0], 1) for k, v in
                            ((k, e.profile_get(k)) for k in ("e3b1", "e3b2", "sur_feat", "sur_mlp", "sur_force"))
                            if v[0]}
                e.profile(False)
            e.option("graph", 1)
            config["surrogate"] = {
                "atom_steps_per_s": n * args.steps / wall_s, "ms_per_step": wall_s / args.steps * 1e3,
                "ns_per_day": dt * 1e-6 * args.steps / wall_s * 86400.0,
                "speed_vs_e3b": (n * args.steps / wall_s) / value,
                "three_body_force_rmse_vs_e3b": {"kcal_mol_A": rms[0], "kJ_mol_nm": rms[1], "bar_kJ_mol_nm": 0.5},
                "e3b_judge": judge, "training": train, "three_body_kernel_avg_ms": kt,
                "model": "%d symmetry-function inputs (%d radial channels) -> 64 -> 64 -> 1 SiLU, BF16 MFMA, FP32 accumulation"
                         % (SUR_NF, SUR_NB),
            }
        # ---- NVE energy drift from the same system (untimed) --------------
        if args.drift_steps:
            e.option("graph", 1)
            slope, h = nve_drift(e, flags, steps=args.drift_steps)
            nmol = len(w["box"]["mol"])
            steps_per_ns = 1e6 / dt
            kT = systems.REAL_BOLTZ * 298.0
            parity["nve_drift"] = {
                "kcal_mol_per_ns": slope * steps_per_ns,
                "kcal_mol_per_ns_per_molecule": slope * steps_per_ns / nmol,
                "kT_per_ns_per_molecule": slope * steps_per_ns / nmol / kT,
                "segment_ps": args.drift_steps * dt * 1e-3, "energy_span_kcal_mol": float(np.ptp(h)),
            }
        config["parity"] = parity
    if not args.no_sub and headline:
        config["c1"] = sub_result("lj", max(args.steps, 100), args.warmup, 0, None)
        config["c2"] = sub_result("tip4p", args.steps, args.warmup, min(args.equil, 1000), "nose-hoover")
    if headline and not args.no_cpu:
        rb = rebuilds_timed
        base, rmse = cpu_baseline(w, x_final, v_final, lo_final, L_final,
                                  max(1, round(args.steps / rb)) if rb else 10, f_final)
        out["cpu_baseline"] = base
        config.setdefault("parity", {})["force_rmse_vs_fp64"] = rmse
    e.close()
    print(json.dumps(out))


def copy_probe(mib=1024, reps=5):
    """Plain device-to-device copy of a known size, for scaling the FETCH_SIZE /
    WRITE_SIZE counters to bytes: 1 GiB (four times the 256 MB Infinity
    Cache) from a source written by a separate fill and read back before the
    copies, so neither the cache nor first-touch effects enter the factor."""
    import torch

    n = mib * 1024 * 1024 // 4
    a = torch.empty(n, dtype=torch.float32, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read_per_launch": n * 4, "bytes_written_per_launch": n * 4,
                      "launches": reps}))


if __name__ == "__main__":
    main()
