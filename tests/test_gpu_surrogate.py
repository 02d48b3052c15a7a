"""Learned E3B three-body surrogate on the GPU against the FP64 restatement
(oracle/surrogate.py): descriptors, per-molecule energies and forces of the
BF16-MFMA network, the per-molecule E3B targets, and on-device training
(loss falls; the trained surrogate's three-body forces beat dropping the
three-body term)."""
import sys

import numpy as np
import pytest

from md_engine_amd import systems
from oracle import surrogate as S

from test_gpu_water_md import start_velocities, water_md_engine

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_rmse(f, ref):
    return float(np.sqrt(((f - ref) ** 2).sum() / (ref ** 2).sum()))


def e2_forces(box, params):
    """two-body E3B forces (EvaluatorE3B.h:67-87) in FP64 numpy"""
    from scipy.spatial import cKDTree

    mol, x, L = box["mol"], box["x"], box["L"]
    O = x[mol[:, 0]]
    pairs = cKDTree(O % L, boxsize=L).query_pairs(params.r2b, output_type="ndarray")
    f = np.zeros_like(x)
    d = O[pairs[:, 0]] - O[pairs[:, 1]]
    d -= L * np.rint(d / L)
    r = np.linalg.norm(d, axis=1)
    s = params.k2 * params.E2 * np.exp(-params.k2 * r) / r
    np.add.at(f, mol[pairs[:, 0], 0], d * s[:, None])
    np.add.at(f, mol[pairs[:, 1], 0], -d * s[:, None])
    return f


def e2_virial(box, params):
    """two-body E3B virial sum d (x) F over O-O pairs (xx, yy, zz, xy, xz, yz)"""
    from scipy.spatial import cKDTree

    mol, x, L = box["mol"], box["x"], box["L"]
    O = x[mol[:, 0]]
    pairs = cKDTree(O % L, boxsize=L).query_pairs(params.r2b, output_type="ndarray")
    d = O[pairs[:, 0]] - O[pairs[:, 1]]
    d -= L * np.rint(d / L)
    r = np.linalg.norm(d, axis=1)
    F = d * (params.k2 * params.E2 * np.exp(-params.k2 * r) / r)[:, None]
    return np.array([(d[:, 0] * F[:, 0]).sum(), (d[:, 1] * F[:, 1]).sum(), (d[:, 2] * F[:, 2]).sum(),
                     (d[:, 0] * F[:, 1]).sum(), (d[:, 0] * F[:, 2]).sum(), (d[:, 1] * F[:, 2]).sum()])


def random_model(seed, feats):
    rng = np.random.default_rng(seed)
    m = S.unpack(S.init_params(seed))
    m["mu"] = feats.mean(0)
    m["sd"] = np.maximum(feats.std(0), 1e-3)
    m["escale"] = 0.05
    m["eshift"] = -0.01
    m["b1"] = rng.normal(0, 0.1, S.NH)
    m["b2"] = rng.normal(0, 0.1, S.NH)
    return S.pack(m).astype(np.float32)


def test_e3b_molecule_energies(orc, gpu_engine_factory):
    say("per-molecule E3B three-body energies")
    box = systems.tip4p2005_box(216, seed=8, min_sep=1.8)
    e, _ = water_md_engine(gpu_engine_factory, box, np.zeros_like(box["x"]), rc=7.0, skin=1.0,
                           sz=(32, 32, 32), order=5, e3b=True, dt=1.0)
    g = e.e3b_molecule_energies(len(box["mol"]))
    o = S.e3b_molecule_energies(box["mol"], box["x"], box["L"])
    assert np.abs(o).max() > 1e-3
    assert np.allclose(g, o, rtol=1e-4, atol=1e-5 * np.abs(o).max())


def test_surrogate_forward_backward_vs_oracle(orc, gpu_engine_factory):
    say("surrogate descriptors, energies, forces vs FP64")
    from md_engine_amd import ENERGY, SURROGATE

    box = systems.tip4p2005_box(216, seed=9, min_sep=1.8)
    raw = S.descriptors(box["mol"], box["x"], box["L"])
    p = random_model(1, S.features(raw))
    e, _ = water_md_engine(gpu_engine_factory, box, np.zeros_like(box["x"]), rc=7.0, skin=1.0,
                           sz=(32, 32, 32), order=5, e3b=True, dt=1.0)
    e.surrogate_set(p)
    eg, fg = e.surrogate_eval(len(box["mol"]))
    for c0, c1 in ((0, 3 * S.NB), (3 * S.NB, S.NR)):  # radial channels, pair products
        assert np.allclose(fg[:, c0:c1], raw[:, c0:c1], rtol=1e-4, atol=1e-5 * np.abs(raw[:, c0:c1]).max())
    fo, eo = S.forces(p.astype(np.float64), box["mol"], box["x"], box["L"])
    # BF16 operands, FP32 accumulation: a few 1e-3 of the energy spread
    assert np.abs(eg - eo).max() < 2e-2 * (np.abs(eo - eo.mean()).max() + 1e-3)
    e.compute(SURROGATE | ENERGY)
    f = e.forces() - e2_forces(box, orc.e3b_params())
    assert rel_rmse(f, fo) < 3e-2
    en, _ = e.energies()
    assert en[3] == pytest.approx(eo.sum(), rel=2e-2)
    # virial of the surrogate term (FP32 bond sums against the FP64 oracle)
    from md_engine_amd import VIRIAL

    e.compute(SURROGATE | VIRIAL)
    _, vir = e.energies()
    _, _, wo = S.forces(p.astype(np.float64), box["mol"], box["x"], box["L"], virial=True)
    wg = vir - e2_virial(box, orc.e3b_params())
    assert np.abs(wo[:3]).max() > 1e-3
    assert np.allclose(wg, wo, rtol=0, atol=3e-2 * np.abs(wo).max())


def test_surrogate_training_gain(orc, gpu_engine_factory):
    """Record MD frames, fit on the GPU, and compare three-body forces on a
    later frame: surrogate vs exact E3B, against dropping the term."""
    say("surrogate training on 512-molecule water")
    from md_engine_amd import E3B, SURROGATE

    box = systems.tip4p2005_box(512, seed=12, min_sep=1.8)
    v = start_velocities(orc, box, seed=2)
    e, flags = water_md_engine(gpu_engine_factory, box, v, rc=8.0, skin=1.0, sz=(32, 32, 32), order=5,
                               e3b=True, dt=2.0)
    e.run(200, flags)
    nm = len(box["mol"])
    for k in range(8):
        e.surrogate_record(reset=(k == 0), capacity=8 * nm)
        e.run(50, flags)
    e.surrogate_set(S.init_params(3).astype(np.float32))
    l0 = e.surrogate_fit(1, batch=512, lr=1e-3, seed=1, normalize=True)
    l1 = e.surrogate_fit(300, batch=512, lr=1e-3, seed=2)
    l1 = e.surrogate_fit(100, batch=512, lr=1e-4, seed=3)
    say(f"loss {l0:.4f} -> {l1:.6f}")
    assert l1 < 0.01 * l0
    e.run(100, flags)
    e.compute(E3B)
    f_exact = e.forces()
    e.compute(SURROGATE)
    f_sur = e.forces()
    x, _ = e.get_system()
    b2 = dict(box, x=x)
    f2 = e2_forces(b2, orc.e3b_params())
    three = f_exact - f2
    err_sur = np.sqrt(((f_sur - f_exact) ** 2).sum(1).mean())
    err_drop = np.sqrt((three ** 2).sum(1).mean())
    say(f"three-body force rms: dropped {err_drop:.4f}, surrogate error {err_sur:.4f} kcal/mol/A")
    # the network inputs hold the correlations the E3B energy is built from:
    # a fit on eight frames reproduces the three-body forces of a later frame
    # to a few percent
    assert err_sur < 0.1 * err_drop


def test_training_gradient_vs_oracle(orc, gpu_engine_factory):
    """The hand-written backward pass of training (k_sur_grad: BF16 MFMA
    tiles, FP32 accumulation) against the FP64 loss gradient
    (oracle/surrogate.py loss_grad) on the same recorded samples."""
    say("surrogate training gradient vs FP64")
    box = systems.tip4p2005_box(216, seed=13, min_sep=1.8)
    e, _ = water_md_engine(gpu_engine_factory, box, np.zeros_like(box["x"]), rc=7.0, skin=1.0,
                           sz=(32, 32, 32), order=5, e3b=True, dt=1.0)
    nm = len(box["mol"])
    raw = S.descriptors(box["mol"], box["x"], box["L"])
    p = random_model(2, S.features(raw))
    e.surrogate_set(p)
    e.surrogate_record(reset=True, capacity=nm)
    target = e.e3b_molecule_energies(nm).astype(np.float64)
    _, feat = e.surrogate_eval(nm)
    for bstart, bsize, mul, add in ((0, nm, 1, 0), (16, 96, 7, 3)):
        g = e.surrogate_grad(bstart, bsize, mul, add)
        rows = ((bstart + np.arange(bsize)) * mul + add) % nm
        lo, go = S.loss_grad(p.astype(np.float64), feat[rows].astype(np.float64), target[rows])
        rel = np.linalg.norm(g[:-1] - go) / np.linalg.norm(go)
        say(f"batch {bstart}+{bsize}: loss {g[-1]:.5f} vs {lo:.5f}, gradient rel err {rel:.2e}")
        assert g[-1] == pytest.approx(lo, rel=2e-2)
        assert rel < 2e-2
        # every block of the gradient, not just the largest
        blocks = np.cumsum([0, S.NF * S.NH, S.NH, S.NH * S.NH, S.NH, S.NH, 1])
        for a, b in zip(blocks[:-1], blocks[1:]):
            assert np.linalg.norm(g[a:b] - go[a:b]) <= 3e-2 * np.linalg.norm(go[a:b]) + 1e-6, (a, b)


@pytest.mark.parametrize("graph", [0, 1])
def test_e3b_judge_and_fallback(orc, gpu_engine_factory, graph):
    """E3B judges the learned term every 100 steps: the checks are counted,
    the last RMSE is the three-body force RMSE of the frame it saw, and a
    threshold it cannot meet makes the run continue with E3B exactly as a run
    that switched to E3B at that step."""
    say(f"E3B judge and fallback, graph={graph}")
    from md_engine_amd import E3B, SURROGATE

    box = systems.tip4p2005_box(343, seed=14, min_sep=1.8)
    v = start_velocities(orc, box, seed=3)
    raw = S.descriptors(box["mol"], box["x"], box["L"])
    p = random_model(4, S.features(raw))

    def engine():
        e, flags = water_md_engine(gpu_engine_factory, box, v, rc=8.0, skin=1.0, sz=(32, 32, 32), order=5,
                                   e3b=True, dt=1.0)
        e.surrogate_set(p)
        e.option("graph", graph)
        return e, (flags & ~E3B) | SURROGATE

    # generous bar: three checks, no fallback, the last one at step 300
    e, sflags = engine()
    e.surrogate_judge(100, 1e9)
    e.run(300, sflags)
    st = e.surrogate_judge_stats()
    e.compute(E3B)
    fe = e.forces()
    e.compute(SURROGATE)
    fs = e.forces()
    rms = float(np.sqrt(((fs - fe) ** 2).mean()))
    say(f"judge {st}, host RMSE {rms:.6f}")
    assert st["checks"] == 3 and not st["fell_back"]
    assert st["last_rmse"] == pytest.approx(rms, rel=1e-9)
    e.close()
    # a bar no model meets: the run falls back at step 100
    a, sflags = engine()
    a.surrogate_judge(100, 1e-12)
    a.run(150, sflags)
    st = a.surrogate_judge_stats()
    assert st["checks"] == 1 and st["fell_back"]
    b, _ = engine()
    b.run(100, sflags)
    b.run(50, (sflags & ~SURROGATE) | E3B)
    xa, va = a.get_system()
    xb, vb = b.get_system()
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    a.close()
    b.close()
