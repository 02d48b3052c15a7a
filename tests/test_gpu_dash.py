"""DASH-style front end on the GPU: the reference's EvaluatorLJ analytic
vectors (tests/Evaluators/test_PairEvaluatorLJ.py:72-151) through the pair
kernel, and a water script built with State/Fix objects against the same
system driven through the Engine directly."""
import sys

import numpy as np
import pytest

from md_engine_amd import dash, systems

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_evaluator_lj_vectors():
    say("EvaluatorLJ analytic vectors")
    sigma, epsilon, cutoff = 3.0, 6.0, 10.0
    sig6 = sigma ** 6
    sig12 = sig6 * sig6
    A = [dash.Vector(10.0, 10.0, 10.0)] * 6
    B = [dash.Vector(9.0, 10.0, 10.0), dash.Vector(8.0, 10.0, 10.0), dash.Vector(7.0, 10.0, 10.0),
         dash.Vector(10.0 - sigma * 2.0 ** (1.0 / 6.0), 10.0, 10.0),
         dash.Vector(9.23472, 9.84752, 8.293478), dash.Vector(11.0, 9.0, 12.0)]
    bounds = dash.BoundsGPU(dash.Vector(0, 0, 0), dash.Vector(30, 30, 30), dash.Vector(1, 1, 1))
    ev = dash.EvaluatorLJ()
    for a, b in zip(A, B):
        dr = bounds.minImage(a - b)
        r = dr.len()
        r2i = 1.0 / (r * r)
        r6i = r2i ** 3
        expect = dr.v * (r6i * r2i * 24.0 * epsilon * (2.0 * sig12 * r6i - sig6))
        got = ev.force(sigma, epsilon, cutoff, dr).v
        diff = np.linalg.norm(got - expect)
        # FP32 kernel: relative 2e-6, except at the zero-force point r = 2^(1/6) sigma
        # where the two terms (each ~24 eps sig^6 / r^7 ~ 10) cancel: absolute 1e-4
        if np.linalg.norm(expect) < 1e-3:
            assert diff < 1e-4, (got, expect)
        else:
            assert diff / np.linalg.norm(expect) <= 2e-6, (got, expect)
    # energy with the shift at the cutoff (PairEvaluatorLJ.h:17-60)
    r = 3.7
    e = ev.energy(sigma, epsilon, cutoff, r)
    s6 = (sigma / r) ** 6
    sc6 = (sigma / cutoff) ** 6
    expect = 4.0 * epsilon * ((s6 * s6 - s6) - (sc6 * sc6 - sc6))
    assert e == pytest.approx(expect, rel=1e-5)


def build_state(box, v):
    st = dash.State()
    st.units.setReal()
    st.atomParams.addSpecies(handle="OW", mass=15.9994, atomicNum=8)
    st.atomParams.addSpecies(handle="HY", mass=1.008, atomicNum=1)
    st.atomParams.addSpecies(handle="OM", mass=0.0, atomicNum=0)
    st.bounds = dash.Bounds(st, lo=dash.Vector(0, 0, 0), hi=dash.Vector(*box["L"]))
    st.rCut = 7.0
    st.padding = 1.0
    st.dt = 1.0
    lj = dash.FixLJCut(st, "cut")
    lj.setParameter("sig", "OW", "OW", 3.1589)
    lj.setParameter("eps", "OW", "OW", 0.185207422)
    for h in ("HY", "OM"):
        lj.setParameter("sig", h, h, 0.0)
        lj.setParameter("eps", h, h, 0.0)
    rigid = dash.FixRigid(st, handle="rigid", style="TIP4P/2005")
    e3b = dash.FixE3B(st, "e3b3", "E3B3")
    x = box["x"]
    for o, h1, h2, m in box["mol"]:
        mol = dash.create_TIP4P_2005(st, "OW", "HY", "OM", OPos=x[o], H1Pos=x[h1], H2Pos=x[h2], MPos=x[m])
        rigid.createRigid(*mol.ids)
        e3b.addMolecule(*mol.ids)
    for a in st.atoms:
        a.vel = dash.Vector(*v[a.id])
    charge = dash.FixChargeEwald(st, "charge", "all")
    charge.setParameters(32, -1, 5)
    for f in (charge, lj, rigid, e3b):
        st.activateFix(f)
    return st


def test_water_script_matches_engine(orc, gpu_engine_factory):
    say("DASH-style water script vs direct engine, 20 steps")
    from test_gpu_water_md import start_velocities, water_md_engine

    box = systems.tip4p2005_box(216, seed=21, min_sep=1.8)
    v = start_velocities(orc, box)
    st = build_state(box, v)
    integ = dash.IntegratorVerlet(st)
    rate = integ.run(20)
    assert rate > 0
    xs = np.array([a.pos.v for a in st.atoms])
    vs = np.array([a.vel.v for a in st.atoms])
    e, flags = water_md_engine(gpu_engine_factory, box, v, rc=7.0, skin=1.0, sz=(32, 32, 32), order=5,
                               e3b=True, dt=1.0)
    e.run(20, flags)
    xe, ve = e.get_system()
    assert np.array_equal(xs, xe)
    assert np.array_equal(vs, ve)
    nl = st.grid.getNeighborList()
    assert len(nl) == len(st.atoms) and all(i not in row for i, row in enumerate(nl))
    # GridGPU::verifyNeighborlists: the list against an all-pairs host search
    # at its own cutoff (r_cut + padding); a shorter cutoff must disagree
    assert st.grid.verifyNeighborlists(st.rCut + st.padding)
    assert not st.grid.verifyNeighborlists(st.rCut)
