"""ctypes binding of libmdx.so (the C ABI in include/mdx.h).

The library is loaded from this package directory only.  There is no CPU
fallback: if the HIP library is missing or cannot be loaded, every call raises
MdxError.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmdx.so")

PAIR, PPPM, E3B, ENERGY, VIRIAL, SURROGATE, BONDED = 1, 2, 4, 8, 16, 32, 64
VIRIAL_TRACE = 4096  # only the virial's trace (the barostat's form); read with Engine.virial_trace()
PPPM_PER_BEAD = 8192  # ring polymers: per-bead P3M instead of the centroid contraction
RING_BEAD_LISTS = 16384  # ring polymers: per-bead neighbor lists instead of one on the centroids

# learned three-body model: SUR_NB radial channels, SUR_NF network inputs,
# SUR_NF-64-64-1 network; trained parameters W1 b1 W2 b2 w3 b3, then the fixed
# input mean and scale and the energy scale and shift
SUR_NB = 2
SUR_NF = 3 * SUR_NB + 6
SUR_NH = 64
NTRAIN = SUR_NF * SUR_NH + SUR_NH + SUR_NH * SUR_NH + SUR_NH + SUR_NH + 1
SUR_NPARAM = NTRAIN + 2 * SUR_NF + 2

# every symbol declared in include/mdx.h
SYMBOLS = [
    "mdx_version", "mdx_create", "mdx_destroy", "mdx_last_error", "mdx_system_set",
    "mdx_positions_set", "mdx_system_get", "mdx_box_set", "mdx_nlist_config",
    "mdx_exclusions_set", "mdx_exclusions_from_bonds", "mdx_nlist_update", "mdx_nlist_get",
    "mdx_nlist_stats", "mdx_lj_set", "mdx_special_set", "mdx_ewald_real_set", "mdx_pppm_set",
    "mdx_pppm_info", "mdx_pppm_paths", "mdx_pppm_set_error", "mdx_pppm_interval_set", "mdx_pppm_mesh_get", "mdx_e3b_set", "mdx_forces_compute", "mdx_forces_get", "mdx_energy_get", "mdx_virial_trace_get",
    "mdx_integrator_set", "mdx_run", "mdx_forces_compute_noalloc", "mdx_teams_set", "mdx_sync", "mdx_force_steps",
    "mdx_profile_enable", "mdx_profile_get", "mdx_nlist_totals", "mdx_pair_entries_read", "mdx_option_set", "mdx_rigid_set",
    "mdx_thermostat_set", "mdx_temperature", "mdx_tune", "mdx_run_batch", "mdx_batch_merged", "mdx_tune_batch", "mdx_surrogate_set",
    "mdx_surrogate_get", "mdx_surrogate_record", "mdx_surrogate_fit", "mdx_surrogate_eval", "mdx_surrogate_grad",
    "mdx_surrogate_judge_set", "mdx_surrogate_judge_stats",
    "mdx_e3b_molecule_energies", "mdx_nose_hoover_set", "mdx_nose_hoover_energy", "mdx_nose_hoover_state", "mdx_barostat_set",
    "mdx_pressure", "mdx_barostat_last", "mdx_box_get", "mdx_flexible_water_set", "mdx_bonded_energy_get", "mdx_ring_run",
    "mdx_ring_forces_compute",
]

# kernel ids of mdx_profile_get
PROF_KERNELS = {"build": 0, "pair": 1, "spread": 2, "gather": 3, "e3b1": 4, "e3b2": 5, "fft": 6,
                "field": 7, "sur_feat": 8, "sur_mlp": 9, "sur_force": 10}


class MdxError(RuntimeError):
    pass


_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_vp = C.c_void_p


def load():
    """Load libmdx.so (raises MdxError when it is absent or broken)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MdxError(f"{LIB_PATH} not built: run md_engine_amd/build.py (hipcc, gfx950)")
    try:
        L = C.CDLL(LIB_PATH)
    except OSError as e:
        raise MdxError(f"cannot load {LIB_PATH}: {e}") from e
    c = C.c_int
    d = C.c_double
    L.mdx_version.restype = c
    L.mdx_create.argtypes = [c, C.POINTER(_vp)]
    L.mdx_destroy.argtypes = [_vp]
    L.mdx_last_error.argtypes = [_vp]
    L.mdx_last_error.restype = C.c_char_p
    L.mdx_system_set.argtypes = [_vp, c, _dp, _vp, _vp, _vp, _vp, _dp, _dp, _ip]
    L.mdx_positions_set.argtypes = [_vp, _dp]
    L.mdx_system_get.argtypes = [_vp, _vp, _vp]
    L.mdx_box_set.argtypes = [_vp, _dp, _dp]
    L.mdx_nlist_config.argtypes = [_vp, d, d]
    L.mdx_exclusions_set.argtypes = [_vp, c, _ip, _up]
    L.mdx_exclusions_from_bonds.argtypes = [_vp, c, _ip, c]
    L.mdx_nlist_update.argtypes = [_vp, c, C.POINTER(c)]
    L.mdx_nlist_get.argtypes = [_vp, _vp, _vp, c, C.POINTER(c)]
    L.mdx_nlist_stats.argtypes = [_vp, C.POINTER(C.c_int64), C.POINTER(c), _ip, C.POINTER(C.c_int64)]
    L.mdx_option_set.argtypes = [_vp, C.c_char_p, c]
    L.mdx_rigid_set.argtypes = [_vp, c, _ip, d, d, d, d, d]
    L.mdx_flexible_water_set.argtypes = [_vp, c, _ip, d, d, d, d, d, d, d]
    L.mdx_bonded_energy_get.argtypes = [_vp, C.POINTER(d)]
    L.mdx_ring_run.argtypes = [C.POINTER(_vp), c, c, c, d]
    L.mdx_ring_forces_compute.argtypes = [C.POINTER(_vp), c, c]
    L.mdx_thermostat_set.argtypes = [_vp, c, d, c, d, d, d]
    L.mdx_temperature.argtypes = [_vp, d, d, C.POINTER(d), C.POINTER(d)]
    L.mdx_lj_set.argtypes = [_vp, c, _dp, _dp, _dp]
    L.mdx_special_set.argtypes = [_vp, d, d, d]
    L.mdx_ewald_real_set.argtypes = [_vp, c, d, d, d]
    L.mdx_pppm_set.argtypes = [_vp, c, c, c, c, d, d, d]
    L.mdx_pppm_info.argtypes = [_vp, C.POINTER(d), C.POINTER(d)]
    L.mdx_pppm_paths.argtypes = [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.mdx_pppm_set_error.argtypes = [_vp, d, d, c, d]
    L.mdx_pppm_interval_set.argtypes = [_vp, c]
    L.mdx_pppm_mesh_get.argtypes = [_vp, _ip, C.POINTER(c)]
    L.mdx_e3b_set.argtypes = [_vp, c, _ip, _dp, d, d]
    L.mdx_forces_compute.argtypes = [_vp, c]
    L.mdx_forces_compute_noalloc.argtypes = [_vp, c]
    L.mdx_forces_get.argtypes = [_vp, _dp]
    L.mdx_energy_get.argtypes = [_vp, _dp, _dp]
    L.mdx_virial_trace_get.argtypes = [_vp, C.POINTER(d)]
    L.mdx_integrator_set.argtypes = [_vp, d, d]
    L.mdx_run.argtypes = [_vp, c, c]
    L.mdx_teams_set.argtypes = [_vp, c, c, c]
    L.mdx_sync.argtypes = [_vp]
    L.mdx_force_steps.argtypes = [_vp, c, c, c]
    L.mdx_profile_enable.argtypes = [_vp, c]
    L.mdx_profile_get.argtypes = [_vp, c, C.POINTER(c), C.POINTER(d)]
    L.mdx_nlist_totals.argtypes = [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.mdx_pair_entries_read.argtypes = [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.mdx_tune.argtypes = [_vp, c, c, _ip, _dp]
    L.mdx_run_batch.argtypes = [C.POINTER(_vp), c, c, c]
    L.mdx_batch_merged.argtypes = [_vp, C.POINTER(c)]
    L.mdx_tune_batch.argtypes = [C.POINTER(_vp), c, c, c, _ip, C.POINTER(C.c_double)]
    _fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    L.mdx_surrogate_set.argtypes = [_vp, _fp, c]
    L.mdx_surrogate_get.argtypes = [_vp, _fp, c]
    L.mdx_surrogate_record.argtypes = [_vp, c, c]
    L.mdx_surrogate_fit.argtypes = [_vp, c, c, d, c, c, C.POINTER(d)]
    L.mdx_surrogate_eval.argtypes = [_vp, _vp, _vp]
    L.mdx_surrogate_grad.argtypes = [_vp, c, c, C.c_uint32, C.c_uint32, _vp, c]
    L.mdx_surrogate_judge_set.argtypes = [_vp, c, d]
    L.mdx_surrogate_judge_stats.argtypes = [_vp, C.POINTER(C.c_int64), C.POINTER(d), C.POINTER(d), C.POINTER(c)]
    L.mdx_e3b_molecule_energies.argtypes = [_vp, _fp]
    L.mdx_nose_hoover_set.argtypes = [_vp, c, d, d, c, d, d]
    L.mdx_nose_hoover_energy.argtypes = [_vp, C.POINTER(d)]
    L.mdx_nose_hoover_state.argtypes = [_vp, _dp, c]
    L.mdx_barostat_set.argtypes = [_vp, c, d, d, c, d, d, d, d]
    L.mdx_pressure.argtypes = [_vp, c, d, d, C.POINTER(d)]
    L.mdx_barostat_last.argtypes = [_vp, C.POINTER(d)]
    L.mdx_box_get.argtypes = [_vp, _dp, _dp]
    for name in SYMBOLS:
        fn = getattr(L, name)
        if name not in ("mdx_last_error", "mdx_version"):
            fn.restype = c
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Engine:
    """One libmdx context (one system on one GPU)."""

    def __init__(self, device: int = 0):
        L = load()
        h = C.c_void_p()
        rc = L.mdx_create(int(device), C.byref(h))
        if rc != 0:
            raise MdxError(f"mdx_create(device={device}) failed with code {rc}")
        self._h = h
        self._L = L
        self.n = 0

    def close(self):
        if getattr(self, "_h", None):
            self._L.mdx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            msg = self._L.mdx_last_error(self._h).decode()
            raise MdxError(f"{what}: {msg} (code {rc})")

    # system -----------------------------------------------------------------
    def set_system(self, x, lo, L, v=None, inv_mass=None, q=None, types=None, periodic=(1, 1, 1)):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        n = len(x)
        v = None if v is None else np.ascontiguousarray(v, np.float64).reshape(-1, 3)
        im = None if inv_mass is None else np.ascontiguousarray(inv_mass, np.float64)
        q = None if q is None else np.ascontiguousarray(q, np.float64)
        t = None if types is None else np.ascontiguousarray(types, np.int32)
        self._keep = (v, im, q, t)
        self._chk(self._L.mdx_system_set(self._h, n, x, _ptr(v), _ptr(im), _ptr(q), _ptr(t),
                                         np.asarray(lo, np.float64), np.asarray(L, np.float64),
                                         np.asarray(periodic, np.int32)), "mdx_system_set")
        self.n = n

    def set_positions(self, x):
        self._chk(self._L.mdx_positions_set(self._h, np.ascontiguousarray(x, np.float64).reshape(-1)),
                  "mdx_positions_set")

    def get_system(self):
        x = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        self._chk(self._L.mdx_system_get(self._h, _ptr(x), _ptr(v)), "mdx_system_get")
        return x, v

    def set_box(self, lo, L):
        self._chk(self._L.mdx_box_set(self._h, np.asarray(lo, np.float64), np.asarray(L, np.float64)),
                  "mdx_box_set")

    def get_box(self):
        lo = np.zeros(3)
        L = np.zeros(3)
        self._chk(self._L.mdx_box_get(self._h, lo, L), "mdx_box_get")
        return lo, L

    # neighbor list ------------------------------------------------------------
    def nlist_config(self, r_list, skin):
        self._chk(self._L.mdx_nlist_config(self._h, float(r_list), float(skin)), "mdx_nlist_config")

    def exclusions_from_bonds(self, bonds, max_depth=3):
        b = np.ascontiguousarray(np.asarray(bonds, np.int32).reshape(-1))
        self._chk(self._L.mdx_exclusions_from_bonds(self._h, len(b) // 2, b, int(max_depth)),
                  "mdx_exclusions_from_bonds")

    def exclusions_set(self, offsets, tagged):
        off = np.ascontiguousarray(offsets, np.int32)
        tg = np.ascontiguousarray(tagged, np.uint32)
        if tg.size == 0:
            tg = np.zeros(1, np.uint32)
        self._chk(self._L.mdx_exclusions_set(self._h, len(off) - 1, off, tg), "mdx_exclusions_set")

    def nlist_update(self, force=False):
        r = C.c_int(0)
        self._chk(self._L.mdx_nlist_update(self._h, int(bool(force)), C.byref(r)), "mdx_nlist_update")
        return bool(r.value)

    def nlist_get(self):
        """Per-id neighbor rows: (counts, list of uint32 arrays id | tag << 30, ids ascending)."""
        counts = np.zeros(self.n, np.int32)
        mx = C.c_int(0)
        self._chk(self._L.mdx_nlist_get(self._h, _ptr(counts), None, 0, C.byref(mx)), "mdx_nlist_get")
        cap = max(1, mx.value)
        tab = np.zeros(self.n * cap, np.uint32)
        self._chk(self._L.mdx_nlist_get(self._h, _ptr(counts), _ptr(tab), cap, C.byref(mx)),
                  "mdx_nlist_get")
        tab = tab.reshape(self.n, cap)
        return counts, [tab[i, : counts[i]].copy() for i in range(self.n)]

    def nlist_stats(self):
        nb = C.c_int64(0)
        cap = C.c_int(0)
        viol = C.c_int64(0)
        cells = np.zeros(3, np.int32)
        self._chk(self._L.mdx_nlist_stats(self._h, C.byref(nb), C.byref(cap), cells, C.byref(viol)),
                  "mdx_nlist_stats")
        return dict(rebuilds=nb.value, row_cap=cap.value, cells=cells.tolist(), violations=viol.value)

    def batch_merged(self):
        """Whether the last batch led by this engine ran as one launch per kernel."""
        m = C.c_int(0)
        self._chk(self._L.mdx_batch_merged(self._h, C.byref(m)), "mdx_batch_merged")
        return bool(m.value)

    def option(self, name, value):
        self._chk(self._L.mdx_option_set(self._h, name.encode(), int(value)), "mdx_option_set")

    # force terms ----------------------------------------------------------------
    def lj_set(self, rc2, eps24, sig6):
        rc2 = np.ascontiguousarray(rc2, np.float64).ravel()
        T = int(round(np.sqrt(rc2.size)))
        self._chk(self._L.mdx_lj_set(self._h, T, rc2, np.ascontiguousarray(eps24, np.float64).ravel(),
                                     np.ascontiguousarray(sig6, np.float64).ravel()), "mdx_lj_set")

    def special_set(self, c12, c13, c14):
        self._chk(self._L.mdx_special_set(self._h, c12, c13, c14), "mdx_special_set")

    def ewald_real_set(self, enable, alpha, qqr, r_cut):
        self._chk(self._L.mdx_ewald_real_set(self._h, int(enable), float(alpha), float(qqr), float(r_cut)),
                  "mdx_ewald_real_set")

    def pppm_set(self, sz, order, r_cut, qqr, alpha=-1.0):
        self._chk(self._L.mdx_pppm_set(self._h, int(sz[0]), int(sz[1]), int(sz[2]), int(order),
                                       float(r_cut), float(qqr), float(alpha)), "mdx_pppm_set")

    def pppm_info(self):
        a = C.c_double(0)
        e = C.c_double(0)
        self._chk(self._L.mdx_pppm_info(self._h, C.byref(a), C.byref(e)), "mdx_pppm_info")
        return a.value, e.value

    def pppm_paths(self):
        """Kernels of the last mesh evaluation: (spread, gather, fft), see mdx_pppm_paths."""
        a = C.c_int(-1)
        b = C.c_int(-1)
        f = C.c_int(-1)
        self._chk(self._L.mdx_pppm_paths(self._h, C.byref(a), C.byref(b), C.byref(f)), "mdx_pppm_paths")
        return a.value, b.value, f.value

    def pppm_set_error(self, error, r_cut, order, qqr):
        """FixChargeEwald.setError: the mesh is picked for this RMS force error."""
        self._chk(self._L.mdx_pppm_set_error(self._h, float(error), float(r_cut), int(order), float(qqr)),
                  "mdx_pppm_set_error")

    def pppm_interval(self, interval):
        """FixChargeEwald.setLongRangeInterval."""
        self._chk(self._L.mdx_pppm_interval_set(self._h, int(interval)), "mdx_pppm_interval_set")

    def pppm_mesh(self):
        sz = np.zeros(3, np.int32)
        iv = C.c_int(0)
        self._chk(self._L.mdx_pppm_mesh_get(self._h, sz, C.byref(iv)), "mdx_pppm_mesh_get")
        return tuple(int(v) for v in sz), iv.value

    def e3b_set(self, mol_ids, params, r_two_body=6.2, r_max=6.1572):
        m = np.ascontiguousarray(mol_ids, np.int32).reshape(-1)
        p = np.ascontiguousarray(params, np.float64)
        self._chk(self._L.mdx_e3b_set(self._h, len(m) // 4, m, p, float(r_two_body), float(r_max)),
                  "mdx_e3b_set")

    def rigid_set(self, mol_ids, m_O, m_H, d_OH, d_HH, gamma):
        m = np.ascontiguousarray(mol_ids, np.int32).reshape(-1)
        self._chk(self._L.mdx_rigid_set(self._h, len(m) // 4, m, float(m_O), float(m_H), float(d_OH),
                                        float(d_HH), float(gamma)), "mdx_rigid_set")

    def flexible_water_set(self, mol_ids, k2, k3, k4, r0, ktheta, theta0, gamma):
        m = np.ascontiguousarray(mol_ids, np.int32).reshape(-1)
        self._chk(self._L.mdx_flexible_water_set(self._h, len(m) // 4, m, float(k2), float(k3), float(k4),
                                                 float(r0), float(ktheta), float(theta0), float(gamma)),
                  "mdx_flexible_water_set")

    def bonded_energy(self):
        e = C.c_double(0.0)
        self._chk(self._L.mdx_bonded_energy_get(self._h, C.byref(e)), "mdx_bonded_energy_get")
        return e.value

    def thermostat(self, enable, temperature=0.0, every=1, tau=0.0, boltz=1.0, mvv_to_eng=1.0):
        self._chk(self._L.mdx_thermostat_set(self._h, int(enable), float(temperature), int(every),
                                             float(tau), float(boltz), float(mvv_to_eng)),
                  "mdx_thermostat_set")

    def temperature(self, boltz=1.0, mvv_to_eng=1.0):
        t = C.c_double(0)
        k = C.c_double(0)
        self._chk(self._L.mdx_temperature(self._h, float(boltz), float(mvv_to_eng), C.byref(t), C.byref(k)),
                  "mdx_temperature")
        return t.value, k.value

    def compute(self, flags):
        self._chk(self._L.mdx_forces_compute(self._h, int(flags)), "mdx_forces_compute")

    def compute_noalloc(self, flags):
        """mdx_forces_compute without device allocation (P3M work arrays in
        the library's scratch; needs one compute() first)."""
        self._chk(self._L.mdx_forces_compute_noalloc(self._h, int(flags)), "mdx_forces_compute_noalloc")

    def forces(self):
        f = np.zeros(self.n * 3)
        self._chk(self._L.mdx_forces_get(self._h, f), "mdx_forces_get")
        return f.reshape(-1, 3)

    def energies(self):
        e = np.zeros(4)
        v = np.zeros(6)
        self._chk(self._L.mdx_energy_get(self._h, e, v), "mdx_energy_get")
        return e, v

    def virial_trace(self):
        """xx + yy + zz of the last compute's virial (MDX_VIRIAL or VIRIAL_TRACE)."""
        w = C.c_double()
        self._chk(self._L.mdx_virial_trace_get(self._h, C.byref(w)), "mdx_virial_trace_get")
        return w.value

    # integration ------------------------------------------------------------------
    def integrator_set(self, dt, ftm_to_v):
        self._chk(self._L.mdx_integrator_set(self._h, float(dt), float(ftm_to_v)), "mdx_integrator_set")

    def run(self, nsteps, flags):
        self._chk(self._L.mdx_run(self._h, int(nsteps), int(flags)), "mdx_run")

    def teams_set(self, build=0, pair=0, e3b=0):
        self._chk(self._L.mdx_teams_set(self._h, int(build), int(pair), int(e3b)), "mdx_teams_set")

    def nose_hoover(self, enable, temperature, time_constant, chain_length=10, boltz=1.0, mvv_to_eng=1.0):
        self._chk(self._L.mdx_nose_hoover_set(self._h, int(bool(enable)), float(temperature),
                                              float(time_constant), int(chain_length), float(boltz),
                                              float(mvv_to_eng)), "mdx_nose_hoover_set")

    def nose_hoover_energy(self):
        e = C.c_double(0.0)
        self._chk(self._L.mdx_nose_hoover_energy(self._h, C.byref(e)), "mdx_nose_hoover_energy")
        return e.value

    def nose_hoover_state(self, chain_length):
        """Chain block: velocities, forces, masses, positions (chain_length
        each), sum m v^2, last scale factor."""
        st = np.zeros(4 * chain_length + 2)
        self._chk(self._L.mdx_nose_hoover_state(self._h, st, st.size), "mdx_nose_hoover_state")
        return st

    def barostat(self, enable, pressure=1.0, period=1000.0, every=10, bulk_modulus=10.0, max_dilation=1e-5,
                 nktv_to_press=68568.415, mvv_to_eng=1.0):
        self._chk(self._L.mdx_barostat_set(self._h, int(bool(enable)), float(pressure), float(period), int(every),
                                           float(bulk_modulus), float(max_dilation), float(nktv_to_press),
                                           float(mvv_to_eng)), "mdx_barostat_set")

    def pressure(self, flags, nktv_to_press=68568.415, mvv_to_eng=1.0):
        p = C.c_double(0.0)
        self._chk(self._L.mdx_pressure(self._h, int(flags), float(nktv_to_press), float(mvv_to_eng), C.byref(p)),
                  "mdx_pressure")
        return p.value

    def barostat_last(self):
        p = C.c_double(0.0)
        self._chk(self._L.mdx_barostat_last(self._h, C.byref(p)), "mdx_barostat_last")
        return p.value

    # -- learned E3B three-body surrogate -----------------------------------
    SUR_NPARAM = SUR_NPARAM
    SUR_NRAW = 3 * SUR_NB + 3  # raw descriptor channels per molecule

    def surrogate_set(self, params):
        p = np.ascontiguousarray(params, np.float32)
        self._chk(self._L.mdx_surrogate_set(self._h, p, len(p)), "mdx_surrogate_set")

    def surrogate_get(self):
        p = np.zeros(self.SUR_NPARAM, np.float32)
        self._chk(self._L.mdx_surrogate_get(self._h, p, len(p)), "mdx_surrogate_get")
        return p

    def surrogate_record(self, reset=False, capacity=0):
        self._chk(self._L.mdx_surrogate_record(self._h, int(bool(reset)), int(capacity)),
                  "mdx_surrogate_record")

    def surrogate_fit(self, epochs, batch=4096, lr=1e-3, seed=0, normalize=False):
        loss = C.c_double(0.0)
        self._chk(self._L.mdx_surrogate_fit(self._h, int(epochs), int(batch), float(lr), int(seed),
                                            int(bool(normalize)), C.byref(loss)), "mdx_surrogate_fit")
        return loss.value

    def surrogate_grad(self, bstart, bsize, mul=1, add=0):
        """Summed gradient of the trained parameters over one mini-batch of
        recorded samples and the batch loss (last entry)."""
        g = np.zeros(NTRAIN + 1, np.float32)
        self._chk(self._L.mdx_surrogate_grad(self._h, int(bstart), int(bsize), int(mul), int(add), g.ctypes.data,
                                             g.size), "mdx_surrogate_grad")
        return g

    def surrogate_judge(self, every, max_rmse):
        """E3B judges the learned term every `every` steps of a run; above
        max_rmse (kcal/mol/A) the run falls back to E3B."""
        self._chk(self._L.mdx_surrogate_judge_set(self._h, int(every), float(max_rmse)), "mdx_surrogate_judge_set")

    def surrogate_judge_stats(self):
        n = C.c_int64(0)
        last = C.c_double(0)
        worst = C.c_double(0)
        fb = C.c_int(0)
        self._chk(self._L.mdx_surrogate_judge_stats(self._h, C.byref(n), C.byref(last), C.byref(worst), C.byref(fb)),
                  "mdx_surrogate_judge_stats")
        return dict(checks=n.value, last_rmse=last.value, worst_rmse=worst.value, fell_back=bool(fb.value))

    def surrogate_eval(self, nmol):
        e = np.zeros(nmol, np.float32)
        f = np.zeros((nmol, self.SUR_NRAW), np.float32)
        self._chk(self._L.mdx_surrogate_eval(self._h, e.ctypes.data, f.ctypes.data), "mdx_surrogate_eval")
        return e, f

    def e3b_molecule_energies(self, nmol):
        e = np.zeros(nmol, np.float32)
        self._chk(self._L.mdx_e3b_molecule_energies(self._h, e), "mdx_e3b_molecule_energies")
        return e

    def tune(self, flags, evals=10):
        """Time every launch variant of the list build, pair and E3B kernels and
        keep the fastest (the reference's Integrator::tune).  Returns the choice
        and the winners' average kernel times in ms."""
        ch = np.zeros(4, np.int32)
        ms = np.zeros(3)
        self._chk(self._L.mdx_tune(self._h, int(flags), int(evals), ch, ms), "mdx_tune")
        return dict(build_mode=int(ch[0]), build_team=int(ch[1]), pair_team=int(ch[2]),
                    e3b_team=int(ch[3]), build_ms=float(ms[0]), pair_ms=float(ms[1]),
                    e3b_ms=float(ms[2]))

    def sync(self):
        self._chk(self._L.mdx_sync(self._h), "mdx_sync")

    def force_steps(self, nsteps, flags, rebuild_every=0):
        self._chk(self._L.mdx_force_steps(self._h, int(nsteps), int(flags), int(rebuild_every)),
                  "mdx_force_steps")

    def profile(self, enable=True):
        self._chk(self._L.mdx_profile_enable(self._h, int(bool(enable))), "mdx_profile_enable")

    def profile_get(self, kernel):
        n = C.c_int(0)
        t = C.c_double(0.0)
        k = PROF_KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
        self._chk(self._L.mdx_profile_get(self._h, k, C.byref(n), C.byref(t)), "mdx_profile_get")
        return n.value, t.value

    def nlist_totals(self):
        a = C.c_int64(0)
        b = C.c_int64(0)
        self._chk(self._L.mdx_nlist_totals(self._h, C.byref(a), C.byref(b)), "mdx_nlist_totals")
        return a.value, b.value

    def pair_entries_read(self):
        """(row entries the pair kernel reads, atoms that read any) on the current list."""
        a, b = C.c_int64(0), C.c_int64(0)
        self._chk(self._L.mdx_pair_entries_read(self._h, C.byref(a), C.byref(b)), "mdx_pair_entries_read")
        return a.value, b.value


def tune_batch(engines, flags, steps=16):
    """Time whole batch steps for every candidate team width and list-build
    kernel (mdx_tune_batch) and keep the fastest on every replica.  Returns
    the choice (build team, pair team, E3B team) and the ms per batch step."""
    engines = list(engines)
    L = load()
    arr = (_vp * len(engines))(*[e._h for e in engines])
    ch = np.zeros(3, np.int32)
    ms = C.c_double(0.0)
    rc = L.mdx_tune_batch(arr, len(engines), int(flags), int(steps), ch, C.byref(ms))
    if rc != 0:
        msg = L.mdx_last_error(engines[0]._h)
        raise MdxError(f"mdx_tune_batch failed ({rc}): {msg.decode() if msg else ''}")
    return dict(build_team=int(ch[0]), pair_team=int(ch[1]), e3b_team=int(ch[2])), ms.value


def run_batch(engines, nsteps, flags):
    """Advance several engines (one system each, same device) by nsteps with
    one captured graph holding every replica (mdx_run_batch)."""
    engines = list(engines)
    if not engines:
        return
    L = load()
    arr = (_vp * len(engines))(*[e._h for e in engines])
    rc = L.mdx_run_batch(arr, len(engines), int(nsteps), int(flags))
    if rc != 0:
        msg = L.mdx_last_error(engines[0]._h)
        raise MdxError(f"mdx_run_batch failed ({rc}): {msg.decode() if msg else ''}")


def ring_omega(temperature_beads, boltz, hbar):
    """Ring-polymer frequency omegaP = kB T / hbar at the bead temperature
    (the reference's convention: T = P x physical temperature,
    IntegratorVerlet.cu:794)."""
    return boltz * temperature_beads / hbar


def run_ring(beads, nsteps, flags, omega_p):
    """Advance P bead engines as ring polymers (mdx_ring_run)."""
    beads = list(beads)
    if not beads:
        return
    L = load()
    arr = (_vp * len(beads))(*[e._h for e in beads])
    rc = L.mdx_ring_run(arr, len(beads), int(nsteps), int(flags), float(omega_p))
    if rc != 0:
        msg = L.mdx_last_error(beads[0]._h)
        raise MdxError(f"mdx_ring_run failed ({rc}): {msg.decode() if msg else ''}")


def ring_forces(beads, flags):
    """Forces of every bead of a ring at the current positions
    (mdx_ring_forces_compute); read each bead's with Engine.forces()."""
    beads = list(beads)
    L = load()
    arr = (_vp * len(beads))(*[e._h for e in beads])
    rc = L.mdx_ring_forces_compute(arr, len(beads), int(flags))
    if rc != 0:
        msg = L.mdx_last_error(beads[0]._h)
        raise MdxError(f"mdx_ring_forces_compute failed ({rc}): {msg.decode() if msg else ''}")
