"""The FixChargeEwald surface beyond setParameters, through the C ABI:
setError (the mesh the doubling search of setGridToErrorTolerance picks,
FixChargeEwald.cu:1024-1096) and setLongRangeInterval (the mesh part every
n-th step, its stored forces added on the steps between,
FixChargeEwald.cu:804-814, 1251, 1402-1451)."""
import sys

import numpy as np
import pytest

from md_engine_amd import systems

from test_gpu_parity import water_engine

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_error_mode_picks_the_oracle_mesh(orc, gpu_engine_factory):
    from md_engine_amd._lib import PPPM

    box = systems.tip4p2005_box(1000, seed=15)
    qqr = systems.REAL_QQR_TO_ENG
    q2 = orc.total_q2(box["q"], qqr)
    for target in (5e-2, 1e-2, 3e-3):
        say(f"error mode, target {target}")
        e = water_engine(gpu_engine_factory, box, rc=8.0, sz=(32, 32, 32))
        e.pppm_set_error(target, 8.0, 5, qqr)  # the search starts from the 32^3 set before
        sz, interval = e.pppm_mesh()
        so, err = orc.pppm_mesh_for_error(len(box["x"]), q2, box["L"], (32, 32, 32), 5, 8.0, target)
        say(f"  mesh {sz} (oracle {so}, estimate {err:.2e})")
        assert sz == so and interval == 1
        a, rms = e.pppm_info()
        assert rms == pytest.approx(err, rel=1e-12)
        e.compute(PPPM)
        f_err = e.forces()
        # the same mesh set directly gives the same forces, bit for bit
        e2 = water_engine(gpu_engine_factory, box, rc=8.0, sz=sz)
        e2.compute(PPPM)
        assert np.array_equal(e2.forces(), f_err)
        assert e2.pppm_info()[0] == a
        e.close()
        e2.close()


def ionic_fluid(make, n=1000, L=30.0, seed=3, interval=1):
    """point charges +-1 with LJ (eps 0.1 kcal/mol, sigma 3 A), 20 amu, real
    units, rc 9 A: plain velocity Verlet (no constraints) so the host can
    replay the steps with forces the engine reports"""
    from md_engine_amd import PAIR, PPPM

    rng = np.random.default_rng(seed)
    m = int(np.ceil(n ** (1 / 3)))
    g = (np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n] + 0.5) * (L / m)
    x = g + rng.uniform(-0.3, 0.3, g.shape)
    q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    mass = np.full(n, 20.0)
    v = systems.maxwell_boltzmann(mass, 300.0, seed=seed, boltz=systems.REAL_BOLTZ,
                                  mvv_to_eng=systems.REAL_MVV_TO_ENG)
    e = make()
    e.set_system(x, np.zeros(3), np.full(3, L), v=v, inv_mass=1.0 / mass, q=q, types=np.zeros(n, np.int32))
    e.nlist_config(10.0, 1.0)
    e.lj_set([[81.0]], [[24.0 * 0.1]], [[3.0 ** 6]])
    e.pppm_set((32, 32, 32), 5, 9.0, systems.REAL_QQR_TO_ENG)
    e.ewald_real_set(True, -1.0, systems.REAL_QQR_TO_ENG, 9.0)
    e.integrator_set(1.0, systems.REAL_FTM_TO_V)
    if interval > 1:
        e.pppm_interval(interval)
    return e, x, v, mass, PAIR | PPPM


def test_long_range_interval_replays_on_the_host(gpu_engine_factory):
    """Two steps with interval 2: the first evaluates the mesh at its
    positions and stores the forces, the second adds those stored forces to
    its own short-range ones.  The host replays both Verlet steps with
    forces the engine reports for each part and lands on the same state, to
    the FP32 rounding of the forces."""
    say("long-range interval 2 vs host replay")
    from md_engine_amd import PAIR, PPPM

    a, x0, v0, mass, flags = ionic_fluid(gpu_engine_factory, interval=2)
    assert a.pppm_mesh()[1] == 2
    a.run(2, flags)
    xa, va = a.get_system()
    b, _, _, _, _ = ionic_fluid(gpu_engine_factory)
    dt, ftm = 1.0, systems.REAL_FTM_TO_V
    k = 0.5 * dt * ftm / mass[:, None]

    def forces(x, fl):
        b.set_positions(x)
        b.compute(fl)
        return b.forces()

    f0 = forces(x0, PAIR | PPPM)
    vh = v0 + k * f0
    x1 = x0 + dt * vh
    f1 = forces(x1, PAIR | PPPM)
    fm1 = forces(x1, PPPM)           # the mesh forces the first step stores
    v1 = vh + k * f1
    vh2 = v1 + k * f1
    x2 = x1 + dt * vh2
    f2 = forces(x2, PAIR) + fm1      # short range now, mesh from the first step
    v2 = vh2 + k * f2
    f2_fresh = forces(x2, PAIR | PPPM)
    L = 30.0
    d = xa - x2
    d -= L * np.rint(d / L)
    say(f"  max |dx| {np.abs(d).max():.2e}, max |dv| {np.abs(va - v2).max():.2e}, "
        f"stored vs fresh mesh force difference {np.abs(f2 - f2_fresh).max():.2e}")
    # The host's FP64 positions differ from the device's in the last bits
    # (fused multiply-adds), which can flip the FP32 rounding of a pair
    # input: the replay agrees to FP32 force rounding, ~1e-10 here.  Using a
    # fresh mesh in the second step instead of the stored one would move the
    # velocities by k |f2 - f2_fresh| (~1e-6 here): the bound sits a thousand
    # times below that.
    dv_fresh = (k * np.abs(f2 - f2_fresh)).max()
    assert np.abs(va - v2).max() < 1e-3 * dv_fresh
    assert np.abs(d).max() < 1e-3 * dt * dv_fresh + 1e-12
    # the stored forces are really older than the step: they differ from a fresh mesh
    assert np.abs(f2 - f2_fresh).max() > 1e-6 * np.abs(f2).max()
    a.close()
    b.close()


def test_long_range_interval_one_is_the_default(gpu_engine_factory):
    say("long-range interval 1")
    a, _, _, _, flags = ionic_fluid(gpu_engine_factory)
    b, _, _, _, _ = ionic_fluid(gpu_engine_factory)
    b.pppm_interval(1)
    b.pppm_interval(0)  # 0 keeps the interval
    a.run(20, flags)
    b.run(20, flags)
    xa, va = a.get_system()
    xb, vb = b.get_system()
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    a.close()
    b.close()
