"""Molecular pressure and the Berendsen barostat on the GPU.

The virial pressure of rigid molecules (atomic virial minus sum F.(r - R_com))
must equal -dU/dV under centre-of-mass dilation, which is checked by finite
differences of the engine's own energy at fixed Ewald splitting; a barostat
run must keep the constraints and an oracle-exact neighbor list."""
import sys

import numpy as np
import pytest

from md_engine_amd import systems

from test_gpu_water_md import G, min_image_diff, start_velocities, water_md_engine

pytestmark = pytest.mark.gpu
NKTV2P = 68568.415


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def com_dilate(box, x, mu):
    """molecules translated with their centres of mass scaled about the box centre"""
    L, lo = box["L"], box["lo"]
    c = lo + 0.5 * L
    out = x.copy()
    m = np.array([G["mO"], G["mH"], G["mH"]])
    for o, h1, h2, ms in box["mol"]:
        rel = min_image_diff(x[[h1, h2, ms]], x[o], L)
        R = (m[0] * x[o] + m[1] * (x[o] + rel[0]) + m[2] * (x[o] + rel[1])) / m.sum()
        shift = c + mu * (R - c) - R
        out[[o, h1, h2, ms]] += shift
    return out, lo + 0.5 * L * (1 - mu), L * mu


@pytest.mark.parametrize("terms", ["pair_e3b", "all"])
def test_pressure_is_volume_derivative(orc, gpu_engine_factory, terms):
    say(f"virial pressure vs -dU/dV ({terms})")
    from md_engine_amd import E3B, ENERGY, PAIR, PPPM

    box = systems.tip4p2005_box(216, seed=33, min_sep=1.8)
    e, _ = water_md_engine(gpu_engine_factory, box, np.zeros_like(box["x"]), rc=7.0, skin=1.0, sz=(32, 32, 32),
                           order=5, e3b=True, dt=1.0)
    alpha, _ = e.pppm_info()
    e.pppm_set((32, 32, 32), 5, 7.0, systems.REAL_QQR_TO_ENG, alpha=alpha)  # fixed splitting
    flags = PAIR | E3B | (PPPM if terms == "all" else 0)
    p_vir = e.pressure(flags, NKTV2P, systems.REAL_MVV_TO_ENG)
    x0 = box["x"]
    us = []
    h = 2e-5
    for mu in (1 + h, 1 - h):
        x, lo, L = com_dilate(box, x0, mu)
        e.set_box(lo, L)
        e.set_positions(x)
        e.compute(flags | ENERGY)
        en, _ = e.energies()
        us.append(en.sum())
    V0 = np.prod(box["L"])
    dV = V0 * ((1 + h) ** 3 - (1 - h) ** 3)
    p_fd = -(us[0] - us[1]) / dV * NKTV2P
    say(f"P virial {p_vir:.2f} atm, -dU/dV {p_fd:.2f} atm")
    # FP32 energies and pairs crossing the cutoffs between the two volumes
    # limit the finite difference to ~1e-3; a wrong or missing virial term is
    # off by tens of percent
    tol = 0.005 * abs(p_fd) + 2.0 if terms == "pair_e3b" else 0.03 * abs(p_fd) + 20.0
    assert p_vir == pytest.approx(p_fd, abs=tol)


def test_barostat_run(orc, gpu_engine_factory):
    say("Berendsen barostat 400 steps")
    from md_engine_amd import ENERGY  # noqa: F401

    box = systems.tip4p2005_box(216, seed=34, min_sep=1.8)
    v = start_velocities(orc, box, seed=8)
    e, flags = water_md_engine(gpu_engine_factory, box, v, rc=7.0, skin=1.0, sz=(32, 32, 32), order=5,
                               e3b=True, dt=1.0)
    kb, mvv = systems.REAL_BOLTZ, systems.REAL_MVV_TO_ENG
    e.nose_hoover(True, 298.0, 100.0, 10, kb, mvv)
    e.barostat(True, 1.0, 200.0, 10, 10.0, 1e-4, NKTV2P, mvv)
    e.option("graph", 1)  # the barostat runs on the device inside captured steps
    L0 = box["L"].copy()
    e.run(400, flags)
    p = e.barostat_last()
    x, vv = e.get_system()
    lo, L = e.get_box()
    say(f"last pressure {p:.1f} atm, box {L0[0]:.4f} -> {L[0]:.4f}")
    assert np.isfinite(p)
    assert L[0] != L0[0] and abs(L[0] / L0[0] - 1) <= 40 * 1e-4 + 1e-12
    assert np.allclose(L, L[0]) and np.allclose(lo + 0.5 * L, box["lo"] + 0.5 * L0)
    for o, h1, h2, m in box["mol"]:
        d = min_image_diff(x[[h1, h2]], x[o], L)
        assert np.allclose(np.linalg.norm(d, axis=1), G["dOH"], atol=1e-8)
        for i, j in ((o, h1), (o, h2), (h1, h2)):
            r = min_image_diff(x[i], x[j], L)
            assert abs(np.dot(vv[i] - vv[j], r)) < 1e-9
    from test_gpu_parity import assert_same_lists, oracle_rows_by_id
    e.nlist_update(force=True)
    gc, grows = e.nlist_get()
    excl = orc.exclusions_bfs(len(x), systems.water_bonds(box["mol"]))
    oc, orows, _ = oracle_rows_by_id(orc, x, lo, L, 8.0, excl=excl, brute=False)
    assert_same_lists(gc, grows, oc, orows)


@pytest.mark.parametrize("thermo,every,mesh", [(None, 1, True), ((350.0, 3, 0.0), 1, True),
                                               ((330.0, 1, 40.0), 1, True), (None, 3, True), (None, 1, False)],
                         ids=["nve", "rescale-every-3", "berendsen-every-1", "barostat-every-3", "no-mesh"])
def test_barostat_graph_matches_eager(orc, gpu_engine_factory, thermo, every, mesh):
    """The barostat runs on the device inside captured steps: box, pressure,
    positions and the re-solved Ewald splitting agree with eager steps, and
    alpha equals the oracle's secant solve for the final box.  Captured steps
    sum the centre-of-mass kinetic term before a velocity rescale of the same
    step and scale it by the rescale's lambda^2; the thermostat cases pin that.
    Every 3 steps: the virial-trace reductions and the k-space virial are
    skipped on the steps in between (device gate).  Without a mesh the
    dilation runs as its own kernel instead of riding on the Green function's."""
    say(f"Berendsen barostat every {every} step(s), thermostat {thermo}, mesh {mesh}: captured graph vs eager")
    box = systems.tip4p2005_box(216, seed=35, min_sep=1.8)
    v = start_velocities(orc, box, seed=9)
    mvv = systems.REAL_MVV_TO_ENG
    sz = (32, 32, 32) if mesh else None
    out = []
    for graph in (0, 1):
        e, flags = water_md_engine(gpu_engine_factory, box, v, rc=7.0, skin=1.0, sz=sz, order=5,
                                   e3b=True, dt=1.0)
        e.barostat(True, 1.0, 200.0, every, 10.0, 1e-4, NKTV2P, mvv)
        if thermo:
            e.thermostat(True, thermo[0], thermo[1], thermo[2], systems.REAL_BOLTZ, mvv)
        e.option("graph", graph)
        e.run(40, flags)
        x, vv = e.get_system()
        lo, L = e.get_box()
        out.append((x, L, e.barostat_last(), e.pppm_info()[0] if mesh else None))
        e.close()
    (xa, La, pa, aa), (xb, Lb, pb, ab) = out
    say(f"box {La[0]:.9f} / {Lb[0]:.9f}, P {pa:.3f} / {pb:.3f} atm, alpha {aa} / {ab}")
    assert La[0] != box["L"][0]
    assert np.allclose(La, Lb, rtol=1e-10, atol=0)
    assert pb == pytest.approx(pa, rel=1e-4, abs=0.05)
    assert np.abs(min_image_diff(xa, xb, La)).max() < 1e-6
    if mesh:
        q2 = orc.total_q2(box["q"], systems.REAL_QQR_TO_ENG)
        a_orc, _ = orc.pppm_alpha(len(box["x"]), q2, Lb, np.array([32, 32, 32]), 5, 7.0)
        assert ab == pytest.approx(a_orc, rel=1e-12)
        assert aa == pytest.approx(a_orc, rel=1e-12)


@pytest.mark.parametrize("terms", ["pair", "pppm", "e3b", "all"])
def test_virial_trace_matches_full_tensor(gpu_engine_factory, terms):
    """The barostat reads the virial through the trace-only kernels (one r.F
    sum per pair or bond into its own slot); their trace must equal the trace
    of the full tensor of the same configuration, term by term."""
    say(f"trace-only virial vs trace of the full tensor ({terms})")
    from md_engine_amd import E3B, PAIR, PPPM, VIRIAL, VIRIAL_TRACE

    box = systems.tip4p2005_box(343, seed=36, min_sep=1.8)
    e, _ = water_md_engine(gpu_engine_factory, box, np.zeros_like(box["x"]), rc=7.0, skin=1.0, sz=(32, 32, 32),
                           order=5, e3b=True, dt=1.0)
    flags = {"pair": PAIR, "pppm": PPPM, "e3b": E3B, "all": PAIR | PPPM | E3B}[terms]
    e.compute(flags | VIRIAL)
    _, vir = e.energies()
    w_full = e.virial_trace()
    assert w_full == pytest.approx(vir[:3].sum(), rel=1e-12, abs=1e-9)
    e.compute(flags | VIRIAL_TRACE)
    w_tr = e.virial_trace()
    scale = np.abs(vir[:3]).sum()
    say(f"trace {w_full:.6f} full tensor, {w_tr:.6f} trace-only (scale {scale:.3f})")
    assert scale > 0
    # FP32 pair arithmetic summed in another order: ~1e-6 of the diagonal's size
    assert abs(w_tr - w_full) <= 1e-5 * scale
    e.close()
