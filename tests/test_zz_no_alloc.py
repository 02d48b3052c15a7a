"""mdx_forces_compute_noalloc against mdx_forces_compute through the C ABI:
the same forces, energies and virial bit for bit (the same kernels, integer
accumulation; only the P3M work arrays live in the library's device scratch),
on two calls in a row, the second on a smaller system with a smaller mesh;
its refusals (no set-up yet, a mesh above 64^3) are clear errors."""
import sys

import numpy as np
import pytest

from md_engine_amd import systems
from md_engine_amd._lib import E3B, ENERGY, PAIR, PPPM, VIRIAL, MdxError

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def water(make, nmol, seed, mesh):
    box = systems.tip4p2005_box(nmol, seed=seed, min_sep=1.8)
    e = make()
    e.set_system(box["x"], box["lo"], box["L"], q=box["q"], types=box["type"])
    e.exclusions_from_bonds(systems.water_bonds(box["mol"]))
    rc = 8.0
    e.nlist_config(rc + 1.0, 1.0)
    eps24, sig6 = systems.water_lj_tables()
    e.lj_set(np.full((3, 3), rc * rc), eps24, sig6)
    e.pppm_set((mesh, mesh, mesh), 5, rc, systems.REAL_QQR_TO_ENG)
    e.ewald_real_set(True, -1.0, systems.REAL_QQR_TO_ENG, rc)
    k = 0.23900573614
    e.e3b_set(box["mol"], [453000 * k, 150 * k, -1005 * k, 1880 * k, 4.872, 1.907, 5.0, 5.2])
    return e


def test_noalloc_refuses_before_setup(gpu_engine_factory):
    say("no-allocation entry: refused before set-up")
    e = water(gpu_engine_factory, 64, 3, 32)
    with pytest.raises(MdxError, match="set the context up"):
        e.compute_noalloc(PAIR | PPPM)
    e.close()


def test_noalloc_rejects_large_mesh(gpu_engine_factory):
    say("no-allocation entry: 128^3 mesh rejected")
    e = water(gpu_engine_factory, 216, 4, 128)
    e.compute(PAIR)
    e.compute(PAIR | PPPM)
    with pytest.raises(MdxError, match="64\\^3"):
        e.compute_noalloc(PAIR | PPPM)
    e.close()


def test_noalloc_matches_compute(gpu_engine_factory):
    flags = PAIR | PPPM | E3B | ENERGY | VIRIAL
    for nmol, mesh, seed in ((1728, 64, 5), (343, 32, 6)):
        say(f"no-allocation entry vs mdx_forces_compute: {nmol} molecules, {mesh}^3 mesh")
        e = water(gpu_engine_factory, nmol, seed, mesh)
        e.compute(flags)
        f_ref = e.forces()
        en_ref, vir_ref = e.energies()
        for _ in range(2):
            e.compute_noalloc(flags)
            f = e.forces()
            en, vir = e.energies()
            assert np.array_equal(f, f_ref)
            assert np.array_equal(en, en_ref)
            assert np.array_equal(vir, vir_ref)
        # and the regular entry point is unchanged after it
        e.compute(flags)
        assert np.array_equal(e.forces(), f_ref)
        e.close()
