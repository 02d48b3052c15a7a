"""Host-side logic of the DASH-style front end (no GPU): parameter tables,
mixing rules, mesh parsing, molecule construction, minImage."""
import numpy as np
import pytest

from md_engine_amd import dash, systems


def water_state(nmol=2):
    st = dash.State()
    st.units.setReal()
    st.atomParams.addSpecies(handle="OW", mass=15.9994, atomicNum=8)
    st.atomParams.addSpecies(handle="HY", mass=1.008, atomicNum=1)
    st.atomParams.addSpecies(handle="OM", mass=0.0, atomicNum=0)
    st.bounds = dash.Bounds(st, lo=dash.Vector(0, 0, 0), hi=dash.Vector(20, 20, 20))
    for k in range(nmol):
        dash.create_TIP4P_2005(st, "OW", "HY", "OM", center=dash.Vector(5 + 8 * k, 5, 5))
    return st


def test_lj_tables_geometric_and_explicit():
    st = water_state()
    lj = dash.FixLJCut(st, "cut")
    lj.setParameter("sig", "OW", "OW", 3.1589)
    lj.setParameter("eps", "OW", "OW", 0.185207422)
    lj.setParameter("sig", "HY", "HY", 0.0)
    lj.setParameter("eps", "HY", "HY", 0.0)
    rc2, eps24, sig6 = lj.tables(st, 9.0)
    ref_eps24, ref_sig6 = systems.water_lj_tables()
    assert np.allclose(eps24, ref_eps24) and np.allclose(sig6, ref_sig6)
    assert np.allclose(rc2, 81.0)


def test_lj_mixing_rules():
    st = dash.State()
    st.atomParams.addSpecies("A", 1.0)
    st.atomParams.addSpecies("B", 2.0)
    for rule, sab in (("geometric", np.sqrt(1.0 * 2.0)), ("arithmetic", 1.5)):
        lj = dash.FixLJCut(st, "cut", mixingRules=rule)
        lj.setParameter("sig", "A", "A", 1.0)
        lj.setParameter("sig", "B", "B", 2.0)
        lj.setParameter("eps", "A", "A", 1.0)
        lj.setParameter("eps", "B", "B", 4.0)
        lj.setParameter("rCut", "A", "A", 2.5)
        lj.setParameter("rCut", "B", "B", 3.0)
        rc2, eps24, sig6 = lj.tables(st, -1)
        assert eps24[0, 1] == pytest.approx(24.0 * 2.0)
        assert sig6[0, 1] == pytest.approx(sab ** 6)
        assert rc2[0, 1] == pytest.approx(9.0)
        assert lj.maxRCut(st) == 3.0


def test_ewald_parameter_forms():
    st = water_state()
    ew = dash.FixChargeEwald(st, "charge", "all")
    ew.setParameters(64, -1, 5)
    assert ew.sz == (64, 64, 64) and ew.r_cut == -1 and ew.order == 5
    ew.setParameters(32, 64, 128, 9.0, 7)
    assert ew.sz == (32, 64, 128) and ew.r_cut == 9.0 and ew.order == 7
    with pytest.raises(ValueError):
        ew.setParameters(48, -1, 5)
    with pytest.raises(NotImplementedError):
        ew.setLongRangeInterval(4)


def test_create_tip4p2005_geometry_and_charges():
    st = water_state(1)
    x = np.array([a.pos.v for a in st.atoms])
    q = np.array([a.q for a in st.atoms])
    assert np.linalg.norm(x[1] - x[0]) == pytest.approx(0.9572, abs=1e-12)
    assert np.linalg.norm(x[2] - x[0]) == pytest.approx(0.9572, abs=1e-12)
    assert np.linalg.norm(x[3] - x[0]) == pytest.approx(0.1546, abs=2e-5)
    assert q.tolist() == [0.0, 0.5564, 0.5564, -1.1128]
    assert st.molecules[0].ids == [0, 1, 2, 3]


def test_e3b_params_both_styles():
    st = water_state()
    k = 0.23900573614
    p3 = dash.FixE3B(st, "e3b", "E3B3").params(st)
    p2 = dash.FixE3B(st, "e3b", "E3B2").params(st)
    assert p3[:4] == pytest.approx([453000 * k, 150 * k, -1005 * k, 1880 * k])
    assert p2[:4] == pytest.approx([2349000 * k, 1745.7 * k, -4565 * k, 7606.8 * k])
    assert p3[4:] == p2[4:] == [4.872, 1.907, 5.0, 5.2]
    st.units.setLJ()
    with pytest.raises(ValueError):
        dash.FixE3B(st, "e3b").params(st)


def test_boundsgpu_minimage():
    b = dash.BoundsGPU(dash.Vector(0, 0, 0), dash.Vector(10, 10, 10), dash.Vector(1, 1, 0))
    v = b.minImage(dash.Vector(7.0, -6.0, 7.0))
    assert (v[0], v[1], v[2]) == (-3.0, 4.0, 7.0)


def test_init_temp_recipe():
    st = water_state(3)
    st.seedRNG(5)
    dash.InitializeAtoms.initTemp(st, "all", 300.0)
    m = np.array([a.mass for a in st.atoms])
    v = np.array([a.vel.v for a in st.atoms])
    assert np.allclose((m[:, None] * v).sum(0), 0.0, atol=1e-12)
    ke = 0.5 * (m[:, None] * v * v).sum() * st.units.mvv_to_eng
    nmass = (m > 0).sum()
    assert 2 * ke / (3 * nmass * st.units.boltz) == pytest.approx(300.0, rel=1e-9)
    assert np.all(v[m == 0] == 0)
