"""FP64 CPU oracle for DASH's per-step force path.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker; the product path (``md_engine_amd``) never imports it.

The arithmetic lives in ``mdx_oracle.c`` (built by ``oracle/Makefile`` into
``liboracle.so``); this module marshals numpy arrays through ctypes and holds
the two small pieces that are easier to state in Python (the exclusion BFS and
the neighbor-tagging step).  Each function cites the reference file:line it
restates.

How each function is pinned independently of this restatement:

* ``wrap`` / min-image / ``dist2``: closed forms (tests/test_oracle_geometry.py).
* ``nlist_cell``: brute force ``nlist_brute`` on the same snapshot, and the
  per-atom count from a numpy distance matrix.
* ``exclusions_bfs``: closed form for rigid water (every intramolecular pair is
  1-2 or 1-3) and hand-enumerated chains.
* ``pair`` (LJ): the reference's analytic vectors
  (tests/Evaluators/test_PairEvaluatorLJ.py:90-151: sigma=3, eps=6, r=2^(1/6)
  sigma gives F=0), Newton's third law, finite differences of the energy, the
  cut-and-shift closed form at r = rc.
* ``pair`` (Ewald real space): scipy's erfc closed form, finite differences,
  the exclusion limit -erf(alpha r)/r.
* ``deltaf_k`` / ``deltaf_real`` (the RMS error estimates behind
  ``pppm_alpha``): the measured RMS force error of P3M against the converged
  reciprocal Ewald sum (orders 1, 3, 5, 7, h alpha 0.25-0.7), and of the cut
  real-space sum against the converged one, on random charges
  (tests/test_oracle_pppm_error.py); doubling or halving any entry of the
  coefficient table moves the estimate out of the tolerance band.  The
  reference's last order-7 entry is 10x the published value (see that test);
  the oracle keeps the reference's entry so that alpha matches it.
* ``pppm_alpha``: the two RMS error estimates are equal at the returned alpha.
* ``pppm_green`` / ``pppm``: converged classical Ewald (``ewald_direct``) on
  small charge-neutral systems; at h alpha 0.3-0.7 the optimal influence
  function gives a smaller error than the plain deconvolved kernel
  4 pi exp(-k^2/4 alpha^2) / (k^2 U^2(k)) and the undeconvolved one, and its
  measured error matches Deserno-Holm's estimate; sum of the spread mesh =
  Qunit * sum q; partition of unity of the weights; the FFT against numpy.fft.
* ``e3b``: brute-force bond-pair enumeration (``e3b_brute_energy``) for the
  energy, finite differences for the forces, Newton's third law.
* ``lj_verlet``: energy conservation over many steps.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> str:
    """Compile liboracle.so from mdx_oracle.c (gcc, a few seconds)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


class E3BParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("E2", "Ea", "Eb", "Ec", "k2", "k3", "rs", "rf", "r2b", "rmax")]


def _declare(L):
    L.orc_wrap.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _ip]
    L.orc_nlist_brute.argtypes = [C.c_int, _dp, _dp, _dp, _ip, C.c_double, C.c_int, _ip, _ip]
    L.orc_nlist_brute.restype = C.c_int
    L.orc_nlist_cell.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _ip, C.c_double, C.c_int, _ip, _ip]
    L.orc_nlist_cell.restype = C.c_int
    L.orc_pair.argtypes = [C.c_int, _dp, _dp, _ip, C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int,
                           C.c_double, C.c_double, C.c_double, _dp, _dp, _ip, _ip, _up, C.c_int,
                           C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_wp.argtypes = [C.c_int, C.c_int, C.c_double]
    L.orc_wp.restype = C.c_double
    L.orc_deltaf_k.argtypes = [C.c_double, C.c_int, C.c_double, _dp, _ip, C.c_int]
    L.orc_deltaf_k.restype = C.c_double
    L.orc_deltaf_real.argtypes = [C.c_double, C.c_int, C.c_double, _dp, C.c_double]
    L.orc_deltaf_real.restype = C.c_double
    L.orc_pppm_amp.argtypes = [C.c_int, C.c_int]
    L.orc_pppm_amp.restype = C.c_double
    L.orc_pppm_alpha.argtypes = [C.c_int, C.c_double, _dp, _ip, C.c_int, C.c_double, C.c_void_p]
    L.orc_pppm_alpha.restype = C.c_double
    L.orc_pppm_green.argtypes = [_dp, _ip, C.c_double, C.c_int, _dp]
    L.orc_fft3.argtypes = [_dp, _ip, C.c_int]
    L.orc_pppm.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _ip, C.c_int, C.c_double, C.c_double,
                           C.c_double, _dp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_ewald_direct.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                   C.c_int, _dp]
    L.orc_ewald_direct.restype = C.c_double
    L.orc_e3b.argtypes = [C.c_int, _ip, _dp, _dp, _ip, C.POINTER(E3BParams), _dp, C.c_void_p,
                          C.c_void_p]
    L.orc_e3b_allpairs.argtypes = L.orc_e3b.argtypes
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_e3b_brute_energy.argtypes = [C.c_int, _ip, _dp, _dp, _ip, C.POINTER(E3BParams)]
    L.orc_e3b_brute_energy.restype = C.c_double
    L.orc_lj_verlet.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp, _dp, C.c_double, C.c_double,
                                C.c_double, C.c_double, C.c_double, C.c_int, C.c_void_p]
    L.orc_lj_verlet.restype = C.c_int
    L.orc_num_threads.restype = C.c_int


def _d(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def num_threads() -> int:
    return lib().orc_num_threads()


# --------------------------------------------------------------------------
# geometry and neighbor lists
# --------------------------------------------------------------------------

def wrap(x, lo, L, periodic=(1, 1, 1)):
    """periodicWrap (GridGPU.cu:138-161), FP64, returns a new array."""
    x = _d(x).copy().reshape(-1, 3)
    L = _d(L)
    lib().orc_wrap(len(x), x, _d(lo), L, 1.0 / L, _i(periodic))
    return x


def nlist_brute(x, L, rlist, periodic=(1, 1, 1), cap=None):
    """O(N^2) full neighbor list (GridGPU::verifyNeighborlists, GridGPU.cu:1109-1200).

    Returns (counts, rows) where rows[i] is the sorted int32 array of neighbor
    indices of atom i; membership: j != i and the canonical FP64 min-image
    r^2 < rlist^2."""
    x = _d(x).reshape(-1, 3)
    n = len(x)
    L = _d(L)
    cap = cap or n
    counts = np.zeros(n, np.int32)
    nbr = np.zeros(n * cap, np.int32)
    lib().orc_nlist_brute(n, x, L, 1.0 / L, _i(periodic), float(rlist) * float(rlist), cap, counts, nbr)
    nbr = nbr.reshape(n, cap)
    return counts, [nbr[i, : counts[i]].copy() for i in range(n)]


def nlist_cell(x, lo, L, rlist, periodic=(1, 1, 1), cap=1024):
    """Cell-list full neighbor list with the same membership as nlist_brute
    (countNumNeighbors / assignNeighbors, GridGPU.cu:326-734).  x must be wrapped."""
    x = _d(x).reshape(-1, 3)
    n = len(x)
    L = _d(L)
    counts = np.zeros(n, np.int32)
    nbr = np.zeros(n * cap, np.int32)
    m = lib().orc_nlist_cell(n, x, _d(lo), L, 1.0 / L, _i(periodic), float(rlist), cap, counts, nbr)
    if m > cap:
        return nlist_cell(x, lo, L, rlist, periodic, cap=int(m * 1.25) + 16)
    return counts, nbr.reshape(n, cap)


def exclusions_bfs(n_ids, bonds, max_depth=3):
    """GridGPU::generateExclusionList / handleExclusionsDistance
    (GridGPU.cu:1377-1470): breadth-first search over bonds to depth 3.

    Returns {id: {other_id: tag}} with tag 1 (1-2), 2 (1-3), 3 (1-4); an id is
    never excluded from itself and keeps only its closest depth."""
    adj = [set() for _ in range(n_ids)]
    for a, b in bonds:
        adj[a].add(b)
        adj[b].add(a)
    out = {}
    for i in range(n_ids):
        levels = [set(adj[i])]
        for depth in range(1, max_depth):
            nxt = set()
            for k in levels[depth - 1]:
                nxt |= adj[k]
            closer = set().union(*levels) | {i}
            levels.append(nxt - closer)
        tags = {}
        for d, lev in enumerate(levels):
            for j in lev:
                tags[j] = d + 1
        if tags:
            out[i] = tags
    return out


def tag_rows(rows, ids, excl):
    """Apply exclusion tags to neighbor index rows: entry = j | tag << 30
    (assignFromCell + addExclusion, GridGPU.cu:471-520).  ids maps index -> id."""
    ids = np.asarray(ids, np.int64)
    n = len(ids)
    keys, tags = [], []
    for a, d in excl.items():
        for b, t in d.items():
            keys.append(a * (n + 1) + b)
            tags.append(t)
    keys = np.asarray(keys, np.int64)
    tags = np.asarray(tags, np.uint32)
    order = np.argsort(keys)
    keys, tags = keys[order], tags[order]
    lens = np.array([len(r) for r in rows])
    if lens.sum() == 0:
        return [r.astype(np.uint32) for r in rows]
    flat = np.concatenate(rows).astype(np.int64)
    owner = np.repeat(np.arange(n), lens)
    k = ids[owner] * (n + 1) + ids[flat]
    pos = np.clip(np.searchsorted(keys, k), 0, max(len(keys) - 1, 0))
    hit = (keys[pos] == k) if len(keys) else np.zeros(len(k), bool)
    t = np.where(hit, tags[pos] if len(keys) else 0, 0).astype(np.uint32)
    out = (flat.astype(np.uint32) | (t << np.uint32(30))).astype(np.uint32)
    return np.split(out, np.cumsum(lens)[:-1])


class NbrTable(tuple):
    """Pre-marshalled (counts int32, table uint32 n*cap, cap) for orc_pair."""


def rows_to_table(counts, rows, cap=None):
    if isinstance(rows, NbrTable):
        return rows
    n = len(rows)
    lens = np.array([len(r) for r in rows], np.int64)
    cap = cap or max(1, int(lens.max()) if n else 1)
    tab = np.zeros(n * cap, np.uint32)
    if lens.sum():
        flat = np.concatenate(rows).astype(np.uint32)
        start = np.cumsum(lens) - lens
        within = np.arange(len(flat)) - np.repeat(start, lens)
        tab[np.repeat(np.arange(n, dtype=np.int64) * cap, lens) + within] = flat
    return NbrTable((_i(counts), tab, cap))


# --------------------------------------------------------------------------
# pair forces
# --------------------------------------------------------------------------

def pair(x, q, types, L, counts, rows_tagged, rc2, eps24, sig6, special=(0.0, 0.0, 0.5),
         with_lj=True, with_charge=False, alpha=0.0, qqr=1.0, rq=0.0, periodic=(1, 1, 1)):
    """Fused LJ + Ewald real-space forces, per-atom energies and virial
    (PairEvaluateIso.h:8-358 with PairEvaluatorLJ.h:17-60 and
    ChargeEvaluatorEwald.h:11-63).  rows_tagged: per-atom uint32 rows
    (index | tag << 30).  rc2/eps24/sig6: T x T type-pair matrices."""
    x = _d(x).reshape(-1, 3)
    n = len(x)
    L = _d(L)
    cnt, tab, cap = rows_to_table(counts, rows_tagged)
    f = np.zeros((n, 3))
    e = np.zeros(n)
    v = np.zeros(6)
    T = int(np.sqrt(np.asarray(rc2).size))
    lib().orc_pair(n, x, _d(q), _i(types), T, _d(rc2).ravel(), _d(eps24).ravel(), _d(sig6).ravel(),
                   _d(special), int(with_lj), int(with_charge), float(alpha), float(qqr),
                   float(rq) ** 2, L, 1.0 / L, _i(periodic), cnt, tab, cap, _ptr(f), _ptr(e), _ptr(v))
    return f, e, v


# --------------------------------------------------------------------------
# P3M
# --------------------------------------------------------------------------

def wp(order, i, d):
    """Deserno-Holm charge-assignment weight (FixChargeEwald.cu:111-115, 163-172, 220-234)."""
    return lib().orc_wp(int(order), int(i), float(d))


def total_q2(q, qqr):
    """qqr * sum q^2 (setTotalQ2, FixChargeEwald.cu:903-915) over the charges
    as the device stores them (FP32), summed in FP64 in atom-id order: the
    same value the library computes (the reference sums the FP32 squares on
    the GPU in FP32, in no fixed order)."""
    q32 = np.asarray(q, np.float64).astype(np.float32).astype(np.float64)
    return float(qqr * np.cumsum(q32 * q32)[-1]) if len(q32) else 0.0


def pppm_alpha(natoms, total_q2, L, sz, order, rcut):
    """Secant solve for the Ewald splitting parameter (FixChargeEwald.cu:853-967)."""
    err = C.c_double(0.0)
    a = lib().orc_pppm_alpha(int(natoms), float(total_q2), _d(L), _i(sz), int(order), float(rcut),
                             C.cast(C.pointer(err), C.c_void_p))
    return a, err.value


def deltaf_k(a, natoms, total_q2, L, sz, order):
    return lib().orc_deltaf_k(float(a), int(natoms), float(total_q2), _d(L), _i(sz), int(order))


def pppm_amp(order, m):
    """Entry m of the order's row of the RMS-error coefficient table."""
    return lib().orc_pppm_amp(int(order), int(m))


def deltaf_real(a, natoms, total_q2, L, rcut):
    return lib().orc_deltaf_real(float(a), int(natoms), float(total_q2), _d(L), float(rcut))


def pppm_mesh_for_error(natoms, total_q2, L, sz, order, rcut, target):
    """setGridToErrorTolerance (FixChargeEwald.cu:1024-1058): from mesh sz,
    double the axis with the fewest points per length (the first such axis
    on ties) until the estimate DeltaF_k + DeltaF_real at the solved alpha
    is at most `target` (at most 100 doublings).  Returns (mesh, error)."""
    sz = [int(v) for v in sz]
    L = np.asarray(L, np.float64)
    _, err = pppm_alpha(natoms, total_q2, L, np.array(sz), order, rcut)
    tries = 0
    while tries < 100 and (err > target or err != err or err < 0):
        k = 0
        for a in (1, 2):
            if sz[a] / L[a] < sz[k] / L[k]:
                k = a
        sz[k] *= 2
        _, err = pppm_alpha(natoms, total_q2, L, np.array(sz), order, rcut)
        tries += 1
    return tuple(sz), err


def pppm_green(L, sz, alpha, order):
    """Optimal influence function on the full mesh (FixChargeEwald.cu:314-366, 1097-1103)."""
    G = np.zeros(int(np.prod(sz)))
    lib().orc_pppm_green(_d(L), _i(sz), float(alpha), int(order), G)
    return G.reshape(tuple(sz))


def fft3(c, sign):
    """Unnormalised 3-D complex FFT, sign -1 forward / +1 inverse."""
    c = np.ascontiguousarray(c, dtype=np.complex128)
    buf = c.view(np.float64).ravel().copy()
    lib().orc_fft3(buf, _i(c.shape), int(sign))
    return buf.view(np.complex128).reshape(c.shape)


def pppm(x, q, lo, L, sz, order, qqr, alpha, total_q2, G, forces=True, energy=False, virial=False,
         rho=False):
    """P3M long-range forces/energy/virial (FixChargeEwald.cu:1231-1611)."""
    x = _d(x).reshape(-1, 3)
    n = len(x)
    f = np.zeros((n, 3)) if forces else None
    e = np.zeros(1) if energy else None
    v = np.zeros(6) if virial else None
    r = np.zeros(int(np.prod(sz))) if rho else None
    lib().orc_pppm(n, x, _d(q), _d(lo), _d(L), _i(sz), int(order), float(qqr), float(alpha),
                   float(total_q2), _d(G).ravel(), _ptr(f), _ptr(e), _ptr(v), _ptr(r))
    out = {"f": f}
    if energy:
        out["energy"] = float(e[0])
    if virial:
        out["virial"] = v
    if rho:
        out["rho"] = r.reshape(tuple(sz))
    return out


def ewald_direct(x, q, L, alpha, qqr, rcut_real, kmax):
    """Converged classical Ewald sum (ground truth for P3M tests)."""
    x = _d(x).reshape(-1, 3)
    f = np.zeros((len(x), 3))
    e = lib().orc_ewald_direct(len(x), x, _d(q), _d(L), float(alpha), float(qqr), float(rcut_real),
                               int(kmax), f)
    return e, f


# --------------------------------------------------------------------------
# E3B
# --------------------------------------------------------------------------

KJ_TO_KCAL = 0.23900573614


def e3b_params(style="E3B3", r2b=6.2, rmax=6.1572):
    """E3B3/E3B2 constants (FixE3B.cu:515-593; EvaluatorE3B.h:46-64) in kcal/mol, Angstrom."""
    if style == "E3B3":
        E2, Ea, Eb, Ec = 453000.0, 150.0, -1005.0, 1880.0
    elif style == "E3B2":
        E2, Ea, Eb, Ec = 2349000.0, 1745.7, -4565.0, 7606.8
    else:
        raise ValueError("Unknown style in FixE3B")
    k = KJ_TO_KCAL
    return E3BParams(E2 * k, Ea * k, Eb * k, Ec * k, 4.872, 1.907, 5.0, 5.2, r2b, rmax)


def e3b(mol, x, L, params=None, periodic=(1, 1, 1)):
    """E3B two- and three-body forces and energies (FixE3B.cu:85-347,
    ThreeBodyE3B.h:18-549, EvaluatorE3B.h:67-345).  mol: (M, 4) atom indices
    (O, H1, H2, M).  Returns (forces (N,3), E_two_body, E_three_body)."""
    p = params or e3b_params()
    x = _d(x).reshape(-1, 3)
    mol = _i(mol).reshape(-1, 4)
    f = np.zeros((len(x), 3))
    e2 = np.zeros(1)
    e3 = np.zeros(1)
    lib().orc_e3b(len(mol), mol.ravel(), x, _d(L), _i(periodic), C.byref(p), f, _ptr(e2), _ptr(e3))
    return f, float(e2[0]), float(e3[0])


def e3b_allpairs(mol, x, L, params=None, periodic=(1, 1, 1)):
    """``e3b`` with the molecule pairs found by an all-pairs scan instead of
    the oxygen cell grid (same sums in the same order)."""
    p = params or e3b_params()
    x = _d(x).reshape(-1, 3)
    mol = _i(mol).reshape(-1, 4)
    f = np.zeros((len(x), 3))
    e2 = np.zeros(1)
    e3 = np.zeros(1)
    lib().orc_e3b_allpairs(len(mol), mol.ravel(), x, _d(L), _i(periodic), C.byref(p), f, _ptr(e2), _ptr(e3))
    return f, float(e2[0]), float(e3[0])


def set_threads(n):
    """OpenMP threads for the following oracle calls."""
    lib().orc_set_threads(int(n))


def e3b_brute_energy(mol, x, L, params=None, periodic=(1, 1, 1)):
    """E3B three-body energy by direct bond-pair enumeration (pins e3b)."""
    p = params or e3b_params()
    x = _d(x).reshape(-1, 3)
    mol = _i(mol).reshape(-1, 4)
    return lib().orc_e3b_brute_energy(len(mol), mol.ravel(), x, _d(L), _i(periodic), C.byref(p))


# --------------------------------------------------------------------------
# integration
# --------------------------------------------------------------------------

def lj_verlet(x, v, mass, lo, L, eps, sigma, rc, skin, dt, nsteps, record_energy=False):
    """FP64 NVE velocity Verlet for an LJ fluid (IntegratorVerlet.cu:307-338, 610-627,
    649-770).  Returns (x, v, n_rebuilds, etot or None); x, v are new arrays."""
    x = _d(x).copy().reshape(-1, 3)
    v = _d(v).copy().reshape(-1, 3)
    et = np.zeros(nsteps + 1) if record_energy else None
    nb = lib().orc_lj_verlet(len(x), x, v, float(mass), _d(lo), _d(L), float(eps), float(sigma),
                             float(rc), float(skin), float(dt), int(nsteps), _ptr(et))
    return x, v, nb, et


# --------------------------------------------------------------------------
# rigid water (SETTLE / RATTLE / M site) and an FP64 water MD driver
# --------------------------------------------------------------------------

_rigid_declared = False


def _declare_rigid():
    global _rigid_declared
    if _rigid_declared:
        return
    L = lib()
    L.orc_settle_one.argtypes = [_dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double]
    L.orc_shake_one.argtypes = [_dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_int]
    L.orc_rattle_one.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double]
    L.orc_water_step_half.argtypes = [C.c_int, C.c_int, _ip, _dp, _dp, _dp, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, _dp, _ip]
    _rigid_declared = True


def settle_one(B0, C0, A1, B1, C1, mO, mH, dOH, dHH):
    """SETTLE for one molecule (FixRigid.cu:850-1136), positions relative to the old O."""
    _declare_rigid()
    a, b, c = (_d(v).copy() for v in (A1, B1, C1))
    lib().orc_settle_one(_d(B0), _d(C0), a, b, c, mO, mH, dOH, dHH)
    return a, b, c


def shake_one(B0, C0, A1, B1, C1, mO, mH, dOH, dHH, tol=1e-15, maxit=1000):
    """Iterative SHAKE on the same constraints (pins SETTLE)."""
    _declare_rigid()
    a, b, c = (_d(v).copy() for v in (A1, B1, C1))
    lib().orc_shake_one(_d(B0), _d(C0), a, b, c, mO, mH, dOH, dHH, tol, maxit)
    return a, b, c


def rattle_one(rO, rH1, rH2, vO, vH1, vH2, mO, mH):
    """Velocity constraint of a rigid water (FixRigid.cu:658-849)."""
    _declare_rigid()
    a, b, c = (_d(v).copy() for v in (vO, vH1, vH2))
    lib().orc_rattle_one(_d(rO), _d(rH1), _d(rH2), a, b, c, mO, mH)
    return a, b, c


WATER_GEOM = dict(mO=15.9994, mH=1.008, dOH=0.9572, dHH=1.5139, gamma=0.73612446364836)


def water_step_half(phase, mol, x, v, f, dt, ftm, L, geom=WATER_GEOM, periodic=(1, 1, 1)):
    """In-place half step of rigid-water velocity Verlet (phase 1: kick, drift,
    SETTLE, M site; phase 2: kick, RATTLE).  f: raw per-site forces."""
    _declare_rigid()
    lib().orc_water_step_half(int(phase), len(mol), _i(mol).ravel(), x, v, _d(f).reshape(-1),
                              geom["mO"], geom["mH"], geom["dOH"], geom["dHH"], geom["gamma"],
                              float(dt), float(ftm), _d(L), _i(periodic))


def water_forces(box, x, rc, skin, sz, order, e3b_on, alpha=None):
    """Raw per-site forces and energies of rigid water (LJ O-O + Ewald real
    space + P3M + optional E3B3) at positions x, FP64."""
    from md_engine_amd import systems  # input tables only

    n = len(x)
    L, lo = box["L"], box["lo"]
    qqr = systems.REAL_QQR_TO_ENG
    q2 = total_q2(box["q"], qqr)
    if alpha is None:
        alpha, _ = pppm_alpha(n, q2, L, np.array(sz), order, rc)
    xw = wrap(x, lo, L)
    counts, tab = nlist_cell(xw, lo, L, rc + skin)
    rows = [tab[i, : counts[i]] for i in range(n)]
    excl = exclusions_bfs(n, systems.water_bonds(box["mol"]))
    rows = tag_rows(rows, np.arange(n), excl)
    eps24, sig6 = systems.water_lj_tables()
    fp, ep, _ = pair(xw, box["q"], box["type"], L, counts, rows, np.full((3, 3), rc * rc), eps24, sig6,
                     with_charge=True, alpha=alpha, qqr=qqr, rq=rc)
    G = pppm_green(L, np.array(sz), alpha, order)
    k = pppm(x, box["q"], lo, L, np.array(sz), order, qqr, alpha, q2, G, energy=True)
    f = fp + k["f"]
    e = ep.sum() + k["energy"]
    if e3b_on:
        fe, e2, e3 = e3b(box["mol"], x, L)
        f = f + fe
        e += e2 + e3
    return f, e, alpha


def water_kinetic(box, v):
    from md_engine_amd import systems

    return 0.5 * (box["mass"][:, None] * v * v).sum() * systems.REAL_MVV_TO_ENG


def water_md(box, v, nsteps, dt=2.0, rc=6.5, skin=0.5, sz=(32, 32, 32), order=5, e3b_on=False):
    """FP64 rigid-water NVE (velocity Verlet + SETTLE + RATTLE), forces
    rebuilt from a fresh list every step.  Returns (x, v, total energies)."""
    from md_engine_amd import systems

    x = _d(box["x"]).copy()
    v = _d(v).copy()
    ftm = systems.REAL_FTM_TO_V
    f, e, alpha = water_forces(box, x, rc, skin, sz, order, e3b_on)
    et = [e + water_kinetic(box, v)]
    for _ in range(nsteps):
        water_step_half(1, box["mol"], x, v, f, dt, ftm, box["L"])
        f, e, _ = water_forces(box, x, rc, skin, sz, order, e3b_on, alpha=alpha)
        water_step_half(2, box["mol"], x, v, f, dt, ftm, box["L"])
        et.append(e + water_kinetic(box, v))
    return x, v, np.array(et)
