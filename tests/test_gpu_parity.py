"""GPU parity of the HIP force path against the FP64 oracle, through the C ABI.

Bars: neighbor lists bit-exact (per-id sorted id sets with exclusion tags);
forces relative RMSE <= 1e-4 against FP64 (sqrt(sum|F - F_ref|^2 / sum|F_ref|^2));
energies and virials relative 1e-5.
"""
import sys

import numpy as np
import pytest

from md_engine_amd import systems

pytestmark = pytest.mark.gpu

FORCE_RMSE = 1e-4
ENERGY_REL = 1e-5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_rmse(f, ref):
    return float(np.sqrt(((f - ref) ** 2).sum() / (ref ** 2).sum()))


def perturbed_fcc(ncell, density=0.8442, amp=0.3, seed=7):
    x, L, lo = systems.fcc_lj(ncell, density)
    rng = np.random.default_rng(seed)
    return x + rng.uniform(-amp, amp, x.shape), L, lo


def oracle_rows_by_id(orc, x, lo, L, rlist, excl=None, brute=True):
    xw = orc.wrap(x, lo, L)
    if brute:
        counts, rows = orc.nlist_brute(xw, L, rlist)
    else:
        counts, tab = orc.nlist_cell(xw, lo, L, rlist)
        rows = [tab[i, : counts[i]].copy() for i in range(len(x))]
    ids = np.arange(len(x))
    if excl is not None:
        rows = orc.tag_rows(rows, ids, excl)
    else:
        rows = [r.astype(np.uint32) for r in rows]
    return counts, rows, xw


def lj_engine(make, x, L, lo, rc=2.5, skin=0.3, v=None):
    e = make()
    e.set_system(x, lo, L, v=v, inv_mass=np.ones(len(x)))
    e.nlist_config(rc + skin, skin)
    e.lj_set([[rc * rc]], [[24.0]], [[1.0]])
    return e


def assert_same_lists(gc, grows, oc, orows):
    assert np.array_equal(gc, oc), f"count mismatch at {np.nonzero(gc != oc)[0][:10]}"
    bad = [i for i in range(len(orows)) if not np.array_equal(grows[i], orows[i])]
    assert not bad, f"{len(bad)} rows differ, first {bad[:5]}"


# ---------------------------------------------------------------------------
# neighbor lists
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("ncell,team", [(6, 8), (10, 8), (10, 64)])
def test_nlist_lj_bitexact(orc, gpu_engine_factory, ncell, team):
    say(f"nlist LJ fcc ncell={ncell} build team {team}")
    x, L, lo = perturbed_fcc(ncell)
    e = lj_engine(gpu_engine_factory, x, L, lo)
    e.teams_set(build=team)
    assert e.nlist_update(force=True)
    gc, grows = e.nlist_get()
    oc, orows, _ = oracle_rows_by_id(orc, x, lo, L, 2.8)
    assert_same_lists(gc, grows, oc, orows)


def test_nlist_lj_full_size_bitexact(orc, gpu_engine_factory):
    say("nlist LJ 32000 atoms")
    x, L, lo = perturbed_fcc(20, amp=0.25, seed=11)
    e = lj_engine(gpu_engine_factory, x, L, lo)
    e.nlist_update(force=True)
    gc, grows = e.nlist_get()
    oc, orows, _ = oracle_rows_by_id(orc, x, lo, L, 2.8, brute=False)
    assert_same_lists(gc, grows, oc, orows)


@pytest.mark.parametrize("team", [8, 32, 64])
def test_nlist_small_noncubic_box(orc, gpu_engine_factory, team):
    """Axes with fewer than 5 cells scan every cell once (no duplicates)."""
    say(f"nlist small non-cubic box, build team {team}")
    rng = np.random.default_rng(3)
    L = np.array([6.5, 12.0, 9.0])
    n = int(0.8 * L.prod())
    x = rng.uniform(-3.0, 15.0, size=(n, 3))  # outside the box on purpose: wrapped by the build
    e = lj_engine(gpu_engine_factory, x, L, np.zeros(3))
    e.teams_set(build=team)
    e.nlist_update(force=True)
    st = e.nlist_stats()
    assert min(st["cells"]) < 5
    gc, grows = e.nlist_get()
    oc, orows, _ = oracle_rows_by_id(orc, x, np.zeros(3), L, 2.8)
    assert_same_lists(gc, grows, oc, orows)


@pytest.mark.parametrize("team", [8, 64])
def test_nlist_wide_neighbourhood(orc, gpu_engine_factory, team):
    """r_list = 7 sigma at liquid density: a 5x5x5-cell neighbourhood holds
    ~4500 candidates, more than one LDS stage of the per-cell build."""
    say(f"nlist wide neighbourhood (r_list 7), build team {team}")
    rng = np.random.default_rng(21)
    L = np.array([17.6, 17.9, 18.3])
    n = int(0.85 * L.prod())
    x = rng.uniform(0.0, 1.0, size=(n, 3)) * L
    e = lj_engine(gpu_engine_factory, x, L, np.zeros(3), rc=6.7, skin=0.3)
    e.teams_set(build=team)
    e.nlist_update(force=True)
    gc, grows = e.nlist_get()
    oc, orows, _ = oracle_rows_by_id(orc, x, np.zeros(3), L, 7.0, brute=False)
    assert gc.mean() > 1000
    assert_same_lists(gc, grows, oc, orows)


@pytest.mark.parametrize("team", [8, 16, 64])
def test_nlist_water_tags_bitexact(orc, gpu_engine_factory, team):
    say(f"nlist water 512 molecules with exclusion tags, build team {team}")
    box = systems.tip4p2005_box(512, seed=5)
    e = gpu_engine_factory()
    e.set_system(box["x"], box["lo"], box["L"], q=box["q"], types=box["type"])
    e.exclusions_from_bonds(systems.water_bonds(box["mol"]))
    e.teams_set(build=team)
    e.nlist_config(10.0, 1.0)
    e.nlist_update(force=True)
    gc, grows = e.nlist_get()
    excl = orc.exclusions_bfs(len(box["x"]), systems.water_bonds(box["mol"]))
    oc, orows, _ = oracle_rows_by_id(orc, box["x"], box["lo"], box["L"], 10.0, excl=excl)
    assert_same_lists(gc, grows, oc, orows)
    tags = np.concatenate(grows) >> 30
    assert set(np.unique(tags)) == {0, 1, 2}


@pytest.mark.parametrize("team", [8, 64])
def test_nlist_capacity_growth(orc, gpu_engine_factory, team):
    """A dense cluster overflows the initial row capacity and crowds one cell
    past 256 atoms: rows grow with no loss (narrow and wide teams)."""
    say(f"nlist capacity growth, build team {team}")
    rng = np.random.default_rng(9)
    L = np.full(3, 20.0)
    x = rng.uniform(0, 20, size=(3000, 3))
    x[:400] = rng.uniform(8, 10, size=(400, 3))
    x[400:700] = rng.uniform(12.1, 12.9, size=(300, 3))
    e = lj_engine(gpu_engine_factory, x, L, np.zeros(3))
    e.teams_set(build=team)
    e.nlist_update(force=True)
    gc, grows = e.nlist_get()
    oc, orows, _ = oracle_rows_by_id(orc, x, np.zeros(3), L, 2.8)
    assert gc.max() > 100
    assert_same_lists(gc, grows, oc, orows)


# ---------------------------------------------------------------------------
# LJ
# ---------------------------------------------------------------------------
def test_lj_forces_full_size(orc, gpu_engine_factory):
    say("LJ forces 32000 atoms")
    from md_engine_amd._lib import ENERGY, PAIR, VIRIAL

    x, L, lo = perturbed_fcc(20, amp=0.15, seed=2)
    e = lj_engine(gpu_engine_factory, x, L, lo)
    e.compute(PAIR | ENERGY | VIRIAL)
    f = e.forces()
    en, vir = e.energies()
    oc, orows, xw = oracle_rows_by_id(orc, x, lo, L, 2.8, brute=False)
    fo, eo, vo = orc.pair(xw, np.zeros(len(x)), np.zeros(len(x), np.int32), L, oc, orows,
                          [[6.25]], [[24.0]], [[1.0]])
    assert rel_rmse(f, fo) <= FORCE_RMSE
    assert en[0] == pytest.approx(eo.sum(), rel=ENERGY_REL)
    assert np.allclose(vir, vo, rtol=ENERGY_REL, atol=1e-6 * np.abs(vo).max())
    assert np.abs(f.sum(0)).max() < 1e-3


def test_forces_deterministic(gpu_engine_factory):
    say("determinism")
    from md_engine_amd._lib import PAIR

    x, L, lo = perturbed_fcc(10)
    a = lj_engine(gpu_engine_factory, x, L, lo)
    a.compute(PAIR)
    f1 = a.forces()
    a.set_positions(x)
    a.compute(PAIR)
    f2 = a.forces()
    b = lj_engine(gpu_engine_factory, x[::-1].copy(), L, lo)  # other id order, same physics
    b.compute(PAIR)
    f3 = b.forces()[::-1]
    assert np.array_equal(f1, f2)
    # a different id order changes the FP32 summation order only
    assert rel_rmse(f3, f1) < 1e-6


# ---------------------------------------------------------------------------
# water: LJ + Ewald real space + P3M, E3B
# ---------------------------------------------------------------------------
def water_engine(make, box, rc=9.0, skin=1.0, sz=(32, 32, 32), order=5, e3b=False):
    e = make()
    e.set_system(box["x"], box["lo"], box["L"], q=box["q"], types=box["type"],
                 inv_mass=np.where(box["mass"] > 0, 1.0 / np.maximum(box["mass"], 1e-30), 0.0))
    e.exclusions_from_bonds(systems.water_bonds(box["mol"]))
    e.nlist_config(rc + skin, skin)
    eps24, sig6 = systems.water_lj_tables()
    e.lj_set(np.full((3, 3), rc * rc), eps24, sig6)
    e.pppm_set(sz, order, rc, systems.REAL_QQR_TO_ENG)
    e.ewald_real_set(True, -1.0, systems.REAL_QQR_TO_ENG, rc)
    if e3b:
        e.e3b_set(box["mol"], e3b_param_vector())
    return e


def e3b_param_vector():
    k = 0.23900573614
    return np.array([453000.0 * k, 150.0 * k, -1005.0 * k, 1880.0 * k, 4.872, 1.907, 5.0, 5.2])


def oracle_water(orc, box, rc=9.0, skin=1.0, sz=(32, 32, 32), order=5):
    n = len(box["x"])
    q2 = orc.total_q2(box["q"], systems.REAL_QQR_TO_ENG)
    alpha, err = orc.pppm_alpha(n, q2, box["L"], np.array(sz), order, rc)
    excl = orc.exclusions_bfs(n, systems.water_bonds(box["mol"]))
    oc, orows, xw = oracle_rows_by_id(orc, box["x"], box["lo"], box["L"], rc + skin, excl=excl,
                                      brute=False)
    eps24, sig6 = systems.water_lj_tables()
    fp, ep, vp = orc.pair(xw, box["q"], box["type"], box["L"], oc, orows, np.full((3, 3), rc * rc), eps24,
                          sig6, with_charge=True, alpha=alpha, qqr=systems.REAL_QQR_TO_ENG, rq=rc)
    G = orc.pppm_green(box["L"], np.array(sz), alpha, order)
    k = orc.pppm(box["x"], box["q"], box["lo"], box["L"], np.array(sz), order, systems.REAL_QQR_TO_ENG,
                 alpha, q2, G, energy=True, virial=True)
    return dict(alpha=alpha, fp=fp, ep=ep.sum(), vp=vp, fk=k["f"], ek=k["energy"], vk=k["virial"])


def test_water_alpha_bit_identical(orc, gpu_engine_factory):
    say("water alpha")
    box = systems.tip4p2005_box(512, seed=5)
    e = water_engine(gpu_engine_factory, box)
    a, _ = e.pppm_info()
    q2 = orc.total_q2(box["q"], systems.REAL_QQR_TO_ENG)
    ao, _ = orc.pppm_alpha(len(box["x"]), q2, box["L"], np.array([32, 32, 32]), 5, 9.0)
    assert a == ao


def test_water_pair_and_pppm(orc, gpu_engine_factory):
    say("water LJ + Ewald real + P3M, 512 molecules")
    from md_engine_amd._lib import ENERGY, PAIR, PPPM, VIRIAL

    box = systems.tip4p2005_box(512, seed=5)
    ref = oracle_water(orc, box)
    e = water_engine(gpu_engine_factory, box)
    e.compute(PAIR | ENERGY | VIRIAL)
    fp = e.forces()
    en, vp = e.energies()
    assert rel_rmse(fp, ref["fp"]) <= FORCE_RMSE
    assert en[0] == pytest.approx(ref["ep"], rel=ENERGY_REL)
    e.compute(PPPM | ENERGY | VIRIAL)
    fk = e.forces()
    en, vk = e.energies()
    assert rel_rmse(fk, ref["fk"]) <= FORCE_RMSE
    assert en[1] == pytest.approx(ref["ek"], rel=ENERGY_REL)
    assert np.allclose(vk, ref["vk"], rtol=1e-4, atol=1e-4 * np.abs(ref["vk"]).max())
    e.compute(PAIR | PPPM)
    f = e.forces()
    assert rel_rmse(f, ref["fp"] + ref["fk"]) <= FORCE_RMSE


@pytest.mark.parametrize("sz", [(32, 32, 32), (32, 32, 64)])
def test_spread_paths_bit_identical(gpu_engine_factory, sz):
    """Owner-computes cell blocks, LDS tiles with global atomics and direct
    global atomics sum the same 64-bit integers: the same force bits."""
    say(f"spread paths, mesh {sz}")
    from md_engine_amd._lib import PPPM

    box = systems.tip4p2005_box(512, seed=12)
    e = water_engine(gpu_engine_factory, box, sz=sz)
    out = {}
    for mode in (2, 1, 0):
        e.option("p3m_spread", mode)
        e.compute(PPPM)
        assert e.pppm_paths()[0] == mode
        out[mode] = e.forces()
    assert np.array_equal(out[2], out[1])
    assert np.array_equal(out[2], out[0])
    # tiles of 2 x 2 cell columns (spread and gather) and a smaller spread
    # tile (where it does not fit the mesh's z extent, the plain spread runs)
    e.option("p3m_spread", 1)
    for cols, kb in ((2, 48), (1, 24)):
        e.option("p3m_tile_cols", cols)
        e.option("spread_tile_kb", kb)
        e.compute(PPPM)
        assert e.pppm_paths()[0] in ((1,) if kb == 48 else (1, 0))
        assert np.array_equal(e.forces(), out[1]), f"tile columns {cols}, spread tile {kb} KB"
    e.option("p3m_tile_cols", 1)
    e.option("spread_tile_kb", 48)


@pytest.mark.parametrize("sz", [(32, 32, 32), (64, 32, 32), (128, 32, 32), (64, 64, 64), (32, 64, 128)])
def test_mesh_solve_paths(orc, gpu_engine_factory, sz):
    """The mesh solve in three LDS passes (real FFT along z two rows at a time,
    complex FFTs along y and x, the field and the k-space energy and virial
    beside rho_hat, the inverse likewise) against the FP64 oracle's P3M and
    against the hipFFT path; a mesh whose planes do not fit a workgroup's LDS
    (32 x 64 x 128) takes the hipFFT path."""
    say(f"mesh solve paths, mesh {sz}")
    from md_engine_amd._lib import ENERGY, PPPM, VIRIAL

    box = systems.tip4p2005_box(512, seed=15)
    ref = oracle_water(orc, box, sz=sz)
    e = water_engine(gpu_engine_factory, box, sz=sz)
    lds = sz != (32, 64, 128)
    out = {}
    for mode in (1, 0):
        e.option("p3m_fft", mode)
        e.compute(PPPM | ENERGY | VIRIAL)
        assert e.pppm_paths()[2] == (mode if lds else 0)
        f = e.forces()
        en, vk = e.energies()
        say(f"fft {mode}: force RMSE {rel_rmse(f, ref['fk']):.2e}, energy {en[1]:.6f} vs {ref['ek']:.6f}")
        assert rel_rmse(f, ref["fk"]) <= FORCE_RMSE
        assert en[1] == pytest.approx(ref["ek"], rel=ENERGY_REL)
        assert np.allclose(vk, ref["vk"], rtol=1e-4, atol=1e-4 * np.abs(ref["vk"]).max())
        out[mode] = (f, en[1], vk)
    # both FP32 transforms of the same mesh: agreement to rounding
    assert rel_rmse(out[1][0], out[0][0]) <= 2e-6
    assert out[1][1] == pytest.approx(out[0][1], rel=1e-6)
    # a second evaluation starts from the cleared charge mesh
    e.option("p3m_fft", 1)
    e.compute(PPPM)
    assert np.array_equal(e.forces(), out[1][0])


@pytest.mark.parametrize("order", [1, 3, 5, 7])
def test_gather_paths(orc, gpu_engine_factory, order):
    """The gathers (LDS tiles of the field, row teams from the per-site mesh
    records, row teams from the positions) take the same weights, field
    points and summation order: the same bits, and the oracle's forces."""
    say(f"gather paths, order {order}")
    from md_engine_amd._lib import PPPM

    box = systems.tip4p2005_box(512, seed=14)
    ref = oracle_water(orc, box, order=order)
    e = water_engine(gpu_engine_factory, box, order=order)
    out = {}
    for mode in (2, 1, 0):
        e.option("p3m_gather", mode)
        e.compute(PPPM)
        # order 7's field box does not fit the LDS here: the records' row teams
        assert e.pppm_paths()[:2] == (1, 1 if (mode == 2 and order == 7) else mode)
        out[mode] = e.forces()
    assert np.array_equal(out[2], out[1])
    assert np.array_equal(out[1], out[0])
    assert rel_rmse(out[1], ref["fk"]) <= FORCE_RMSE
    # the tile gather from the interleaved field, on smaller tiles
    e.option("p3m_gather", 2)
    e.option("gather_tile_planar", 0)
    e.option("gather_tile_kb", 24)
    e.compute(PPPM)
    assert np.array_equal(e.forces(), out[0])


@pytest.mark.parametrize("order", [1, 3, 7])
def test_water_pppm_orders(orc, gpu_engine_factory, order):
    say(f"water P3M order {order}")
    from md_engine_amd._lib import PPPM

    box = systems.tip4p2005_box(216, seed=8)
    ref = oracle_water(orc, box, rc=7.0, sz=(32, 32, 64), order=order)
    e = water_engine(gpu_engine_factory, box, rc=7.0, sz=(32, 32, 64), order=order)
    e.compute(PPPM)
    assert rel_rmse(e.forces(), ref["fk"]) <= FORCE_RMSE


def test_water_e3b2(orc, gpu_engine_factory):
    """The E3B2 parameter set (FixE3B.cu:557-579) through the same kernels."""
    say("water E3B2 512 molecules")
    from md_engine_amd._lib import E3B, ENERGY

    box = systems.tip4p2005_box(512, seed=7)
    e = water_engine(gpu_engine_factory, box)
    p = orc.e3b_params("E3B2")
    e.e3b_set(box["mol"], [p.E2, p.Ea, p.Eb, p.Ec, p.k2, p.k3, p.rs, p.rf])
    e.compute(E3B | ENERGY)
    f = e.forces()
    en, _ = e.energies()
    fo, e2, e3 = orc.e3b(box["mol"], box["x"], box["L"], params=p)
    assert abs(e3) > 1.0
    assert rel_rmse(f, fo) <= FORCE_RMSE
    assert en[2] == pytest.approx(e2, rel=ENERGY_REL)
    assert en[3] == pytest.approx(e3, rel=ENERGY_REL)


def test_water_e3b(orc, gpu_engine_factory):
    say("water E3B 512 molecules")
    from md_engine_amd._lib import E3B, ENERGY

    box = systems.tip4p2005_box(512, seed=6)
    e = water_engine(gpu_engine_factory, box, e3b=True)
    e.compute(E3B | ENERGY)
    f = e.forces()
    en, _ = e.energies()
    fo, e2, e3 = orc.e3b(box["mol"], box["x"], box["L"])
    assert abs(e3) > 1.0
    assert rel_rmse(f, fo) <= FORCE_RMSE
    assert en[2] == pytest.approx(e2, rel=ENERGY_REL)
    assert en[3] == pytest.approx(e3, rel=ENERGY_REL)


# ---------------------------------------------------------------------------
# dynamics
# ---------------------------------------------------------------------------
def test_lj_md_short_trajectory_and_rebuilds(orc, gpu_engine_factory):
    say("LJ NVE 4000 atoms")
    from md_engine_amd._lib import PAIR

    x, L, lo = perturbed_fcc(10, amp=0.05, seed=1)
    v = systems.maxwell_boltzmann(np.ones(len(x)), 1.2, seed=20261202)
    e = lj_engine(gpu_engine_factory, x, L, lo, v=v)
    e.integrator_set(0.005, 1.0)
    e.run(20, PAIR)
    xg, vg = e.get_system()
    xo, vo, _, _ = orc.lj_verlet(x, v, 1.0, lo, L, 1.0, 1.0, 2.5, 0.3, 0.005, 20)
    d = xg - xo
    d -= L * np.rint(d / L)
    assert np.abs(d).max() < 1e-4
    e.run(400, PAIR)
    st = e.nlist_stats()
    assert st["rebuilds"] >= 3
    assert st["violations"] == 0
    # the list after many rebuilds is still bit-exact for the current positions
    e.nlist_update(force=True)
    xg, _ = e.get_system()
    gc, grows = e.nlist_get()
    oc, orows, _ = oracle_rows_by_id(orc, xg, lo, L, 2.8, brute=False)
    assert_same_lists(gc, grows, oc, orows)


# ---------------------------------------------------------------------------
# autotuner
# ---------------------------------------------------------------------------
def test_tune_keeps_lists_and_forces(orc, gpu_engine_factory):
    """The autotuner only changes launch geometry: after it, the list is still
    the oracle's and the full force matches FP64."""
    say("autotuner on 512-molecule water")
    from md_engine_amd._lib import E3B, PAIR, PPPM

    box = systems.tip4p2005_box(512, seed=5)
    e = water_engine(gpu_engine_factory, box, e3b=True)
    e.compute(PAIR | PPPM | E3B)
    choice = e.tune(PAIR | PPPM | E3B, evals=3)
    assert choice["build_mode"] == 2 and choice["build_team"] in (8, 16, 32, 64)
    assert choice["pair_team"] in (4, 8, 16, 32) and choice["e3b_team"] in (4, 8, 16, 32)
    assert choice["build_ms"] > 0 and choice["pair_ms"] > 0 and choice["e3b_ms"] > 0
    gc, grows = e.nlist_get()
    excl = orc.exclusions_bfs(len(box["x"]), systems.water_bonds(box["mol"]))
    oc, orows, _ = oracle_rows_by_id(orc, box["x"], box["lo"], box["L"], 10.0, excl=excl, brute=False)
    assert_same_lists(gc, grows, oc, orows)
    ref = oracle_water(orc, box)
    fo, _, _ = orc.e3b(box["mol"], box["x"], box["L"])
    e.compute(PAIR | PPPM | E3B)
    assert rel_rmse(e.forces(), ref["fp"] + ref["fk"] + fo) <= FORCE_RMSE
