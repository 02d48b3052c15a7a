"""DASH-style Python front end over the MI355X engine.

Mirrors the objects the reference exports for this path (python/DASH.cpp:34-131)
so water and LJ scripts keep their shape: ``State`` with ``units``,
``atomParams``, ``bounds``, ``rCut``/``padding``/``dt``, ``addAtom`` and
``createMolecule``; ``FixLJCut``, ``FixChargeEwald``, ``FixE3B``,
``FixRigid``, ``FixNVTRescale``, ``FixNoseHoover``, ``FixPressureBerendsen``;
``IntegratorVerlet.run`` (``state.graph = True`` replays captured steps);
``InitializeAtoms.initTemp``; ``EvaluatorLJ``; ``BoundsGPU.minImage``;
``state.grid.getNeighborList()``.  Everything runs through
``md_engine_amd._lib.Engine`` (the C ABI of libmdx.so); the objects here only
hold parameters and translate them (handles -> type indices, mixing rules,
molecules -> constraint bonds and exclusions) the way the reference's
prepareForRun code does.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import _lib, systems


class Vector:
    """3-vector with the arithmetic the reference's Vector exposes to Python."""

    def __init__(self, x=0.0, y=0.0, z=0.0):
        self.v = np.array([x, y, z], dtype=np.float64)

    @classmethod
    def of(cls, a):
        if isinstance(a, Vector):
            return cls(*a.v)
        return cls(*np.asarray(a, dtype=np.float64))

    def __getitem__(self, i):
        return float(self.v[i])

    def __setitem__(self, i, val):
        self.v[i] = val

    def __add__(self, o):
        return Vector(*(self.v + Vector.of(o).v))

    def __sub__(self, o):
        return Vector(*(self.v - Vector.of(o).v))

    def __mul__(self, s):
        return Vector(*(self.v * s))

    __rmul__ = __mul__

    def __truediv__(self, s):
        return Vector(*(self.v / s))

    def len(self):
        return float(np.linalg.norm(self.v))

    def lenSqr(self):
        return float(self.v @ self.v)

    def normalized(self):
        return self / self.len()

    def dot(self, o):
        return float(self.v @ Vector.of(o).v)

    def __repr__(self):
        return "Vector(%r, %r, %r)" % tuple(self.v)


class Bounds:
    """Orthorhombic box (Bounds(state, lo=, hi=)); all axes periodic by default."""

    def __init__(self, state=None, lo=None, hi=None):
        self.lo = Vector.of(lo if lo is not None else (0, 0, 0))
        self.hi = Vector.of(hi if hi is not None else (0, 0, 0))
        self.periodic = [True, True, True]

    @property
    def rect(self):
        return self.hi - self.lo


class BoundsGPU:
    """BoundsGPU(lo, rect, periodic).minImage(v): v -= L rint(v / L) on
    periodic axes (src/BoundsGPU.h:91-154)."""

    def __init__(self, lo, rect, periodic=Vector(1, 1, 1)):
        self.lo = Vector.of(lo)
        self.rect = Vector.of(rect)
        self.periodic = Vector.of(periodic)

    def minImage(self, v):
        v = Vector.of(v)
        L = self.rect.v
        return Vector(*(v.v - L * np.rint(v.v / L) * self.periodic.v))


class Units:
    """setReal() / setLJ() (src/Units.cpp:5-32)."""

    def __init__(self):
        self.setLJ()

    def setLJ(self):
        self.unitType = "LJ"
        self.boltz = 1.0
        self.mvv_to_eng = 1.0
        self.qqr_to_eng = 1.0
        self.ftm_to_v = 1.0

    def setReal(self):
        self.unitType = "REAL"
        self.boltz = systems.REAL_BOLTZ
        self.mvv_to_eng = systems.REAL_MVV_TO_ENG
        self.qqr_to_eng = systems.REAL_QQR_TO_ENG
        self.ftm_to_v = systems.REAL_FTM_TO_V


class AtomParams:
    """addSpecies(handle, mass, atomicNum): species become type indices in
    the order they are added."""

    def __init__(self):
        self.handles = []
        self.masses = []
        self.atomicNums = []

    def addSpecies(self, handle, mass, atomicNum=0):
        if handle in self.handles:
            raise ValueError(f"species {handle!r} exists")
        self.handles.append(handle)
        self.masses.append(float(mass))
        self.atomicNums.append(int(atomicNum))
        return len(self.handles) - 1

    def typeFromHandle(self, handle):
        return self.handles.index(handle)


class Atom:
    def __init__(self, id_, type_, pos, q, mass):
        self.id = id_
        self.type = type_
        self.pos = Vector.of(pos)
        self.vel = Vector()
        self.q = float(q)
        self.mass = float(mass)


class Molecule:
    def __init__(self, ids):
        self.ids = list(ids)


class DeviceManager:
    def __init__(self):
        self.device = 0

    def setDevice(self, i, output=True):
        self.device = int(i)
        return True


class Grid:
    """state.grid: neighbor lists of the engine the state last ran on (or a
    fresh build at the current positions)."""

    def __init__(self, state):
        self.state = state

    def buildNeighborlists(self):
        self.state._engine(rebuild=True)

    def getNeighborList(self):
        """Per-id lists of neighbor ids (tags stripped), sorted by id, as
        GridGPU::getNeighborList returns them (GridGPU.cu:1520-1648)."""
        eng = self.state._engine(rebuild=True)
        counts, rows = eng.nlist_get()
        return [sorted(int(x) for x in (r & 0x3FFFFFFF)) for r in rows]

    def verifyNeighborlists(self, neighCut):
        """GridGPU::verifyNeighborlists (GridGPU.cu:1109-1200): every atom's
        neighbor ids against an all-pairs host search (minimum image,
        r^2 < neighCut^2) at the positions the list was built from; the first
        atom that differs is printed.  The list holds every pair within its
        own cutoff (r_cut + padding), so neighCut is that cutoff.  Returns
        whether all rows agree (the reference prints and returns true)."""
        eng = self.state._engine(rebuild=True)
        eng.nlist_update(force=True)
        counts, rows = eng.nlist_get()
        x, _ = eng.get_system()
        lo, L = (np.asarray(a, np.float64) for a in eng.get_box())
        per = np.array([bool(p) for p in self.state.bounds.periodic[:3]])
        cut2 = float(neighCut) ** 2
        n = len(x)
        invL = 1.0 / L
        for i in range(n):
            d = x - x[i]
            d = np.where(per, d - L * np.rint(d * invL), d)
            r2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
            r2[i] = np.inf
            cpu = np.nonzero(r2 < cut2)[0]
            gpu = np.sort((rows[i] & 0x3FFFFFFF).astype(np.int64))
            if not np.array_equal(cpu, gpu):
                print(f"problem at id {i}")
                print(f"cpu {len(cpu)} gpu {len(gpu)}")
                print("cpu neighbor ids\n" + " ".join(map(str, cpu)))
                print("gpu neighbor ids\n" + " ".join(map(str, gpu)))
                return False
        return True


class State:
    """Holds the system and the active fixes; the engine is built from them at
    the first run (State::prepareForRun packing, src/State.cpp:590-608)."""

    def __init__(self):
        self.units = Units()
        self.atomParams = AtomParams()
        self.atoms = []
        self.molecules = []
        self.bounds = Bounds()
        self.rCut = -1.0
        self.padding = 0.5
        # The reference wraps atoms and tests the skin criterion every
        # periodicInterval turns (IntegratorVerlet.cu:696); the engine tests it
        # on the device every step, so a list is never older than the
        # reference's.  Kept so scripts that set it run unchanged.
        self.periodicInterval = 50
        self.shoutEvery = 0
        self.dt = 0.005
        self.turn = 0
        self.specialNeighborCoefs = [0.0, 0.0, 0.5]
        self.nThreadPerBlock = 256
        self.nThreadPerAtom = 0
        self.tuning = False
        self.deviceManager = DeviceManager()
        self.fixes = []
        self.grid = Grid(self)
        self._eng = None
        self._eng_key = None
        self._rng = np.random.default_rng(0)

    def seedRNG(self, seed=0):
        self._rng = np.random.default_rng(seed)

    def Vector(self, x=0.0, y=0.0, z=0.0):
        return Vector(x, y, z)

    def addAtom(self, handle="", pos=None, q=0.0):
        t = self.atomParams.typeFromHandle(handle)
        a = Atom(len(self.atoms), t, pos if pos is not None else (0.0, 0.0, 0.0), q, self.atomParams.masses[t])
        self.atoms.append(a)
        self._eng = None
        return a.id

    def createMolecule(self, ids):
        m = Molecule(ids)
        self.molecules.append(m)
        return m

    def activateFix(self, fix):
        if fix not in self.fixes:
            self.fixes.append(fix)
        self._eng = None
        return True

    def deactivateFix(self, fix):
        if fix in self.fixes:
            self.fixes.remove(fix)
        self._eng = None
        return True

    # -- packing -------------------------------------------------------------
    def _arrays(self):
        x = np.array([a.pos.v for a in self.atoms])
        v = np.array([a.vel.v for a in self.atoms])
        q = np.array([a.q for a in self.atoms])
        t = np.array([a.type for a in self.atoms], np.int32)
        m = np.array([a.mass for a in self.atoms])
        return x, v, q, t, m

    def _fix(self, cls):
        for f in self.fixes:
            if isinstance(f, cls):
                return f
        return None

    def _rcut(self):
        rc = self.rCut
        lj = self._fix(FixLJCut)
        if rc <= 0 and lj is not None:
            rc = lj.maxRCut(self)
        if rc <= 0:
            raise ValueError("state.rCut is not set")
        return rc

    def _engine(self, rebuild=False):
        """The engine for the current atoms and fixes (built on first use)."""
        if self._eng is None:
            self._eng, self._flags = self._build_engine()
        if rebuild:
            self._push()
            self._eng.nlist_update(force=True)
        return self._eng

    def _push(self):
        x, v, q, t, m = self._arrays()
        inv = np.where(m > 0, 1.0 / np.maximum(m, 1e-300), 0.0)
        lo = self.bounds.lo.v
        L = self.bounds.rect.v
        self._eng.set_system(x, lo, L, v=v, inv_mass=inv, q=q, types=t,
                             periodic=tuple(int(p) for p in self.bounds.periodic))

    def _build_engine(self):
        if not self.atoms:
            raise ValueError("no atoms")
        eng = _lib.Engine(self.deviceManager.device)
        self._eng = eng
        self._push()
        rc = self._rcut()
        flags = 0
        rigid = self._fix(FixRigid)
        bonds = []
        if rigid is not None:
            for o, h1, h2, m in rigid.molecules:
                bonds += [(o, h1), (o, h2), (h1, h2)] + ([(o, m)] if m >= 0 else [])
        if bonds:
            eng.exclusions_from_bonds(np.array(bonds, np.int32))
        eng.special_set(*self.specialNeighborCoefs)
        eng.nlist_config(rc + self.padding, self.padding)
        lj = self._fix(FixLJCut)
        if lj is not None:
            rc2, eps24, sig6 = lj.tables(self, rc)
            eng.lj_set(rc2, eps24, sig6)
            flags |= _lib.PAIR
        ew = self._fix(FixChargeEwald)
        if ew is not None:
            rq = ew.r_cut if ew.r_cut > 0 else rc
            eng.pppm_set(ew.sz, ew.order, rq, self.units.qqr_to_eng)
            eng.ewald_real_set(True, -1.0, self.units.qqr_to_eng, rq)
            flags |= _lib.PPPM | _lib.PAIR
        e3b = self._fix(FixE3B)
        if e3b is not None:
            eng.e3b_set(np.array(e3b.molecules, np.int32), e3b.params(self))
            flags |= _lib.E3B
        if rigid is not None:
            p = systems.TIP4P2005
            mO = self.atoms[rigid.molecules[0][0]].mass
            mH = self.atoms[rigid.molecules[0][1]].mass
            eng.rigid_set(np.array(rigid.molecules, np.int32), mO, mH, p["r_OH"], p["r_HH"], p["gamma"])
        nvt = self._fix(FixNVTRescale)
        if nvt is not None:
            eng.thermostat(True, nvt.temp, nvt.applyEvery, 0.0, self.units.boltz, self.units.mvv_to_eng)
        nh = self._fix(FixNoseHoover)
        if nh is not None:
            if nh.temp is None:
                raise ValueError("FixNoseHoover: call setTemperature first")
            eng.nose_hoover(True, nh.temp, nh.timeConstant, nh.chainLength, self.units.boltz,
                            self.units.mvv_to_eng)
        ber = self._fix(FixPressureBerendsen)
        if ber is not None:
            nktv2p = 68568.415 if self.units.unitType == "REAL" else 1.0
            eng.barostat(True, ber.pressure, ber.period, ber.applyEvery, ber.bulkModulus, ber.maxDilation,
                         nktv2p, self.units.mvv_to_eng)
        if getattr(self, "graph", False):
            eng.option("graph", 1)
        eng.integrator_set(self.dt, self.units.ftm_to_v)
        if self.nThreadPerAtom:
            eng.teams_set(build=0, pair=int(self.nThreadPerAtom))
        if self.tuning:
            eng.tune(flags, evals=10)
        return eng, flags

    def _pull(self):
        x, v = self._eng.get_system()
        for a in self.atoms:
            a.pos = Vector(*x[a.id])
            a.vel = Vector(*v[a.id])
        lo, L = self._eng.get_box()
        self.bounds.lo = Vector(*lo)
        self.bounds.hi = Vector(*(lo + L))


class Fix:
    def __init__(self, state, handle, groupHandle="all"):
        self.state = state
        self.handle = handle
        self.groupHandle = groupHandle


class FixLJCut(Fix):
    """setParameter('eps'|'sig'|'rCut', handleA, handleB, value); missing
    cross terms come from the mixing rule (geometric: sqrt(e_i e_j),
    sqrt(s_i s_j); arithmetic sigma: (s_i + s_j)/2), rCut pair = max of the
    two (FixLJCut.cu:83-122, FixPair.cpp:13-35)."""

    def __init__(self, state, handle, mixingRules="geometric"):
        super().__init__(state, handle)
        if mixingRules not in ("geometric", "arithmetic"):
            raise ValueError("mixingRules: geometric or arithmetic")
        self.mixingRules = mixingRules
        self.p = {"eps": {}, "sig": {}, "rCut": {}}

    def setParameter(self, param, handleA, handleB, val):
        if param not in self.p:
            raise ValueError(f"unknown LJ parameter {param!r}")
        self.p[param][(handleA, handleB)] = float(val)
        self.p[param][(handleB, handleA)] = float(val)
        return True

    def getParameter(self, param, handleA, handleB):
        return self.p[param].get((handleA, handleB), None)

    def maxRCut(self, state):
        return max(self.p["rCut"].values(), default=-1.0)

    def tables(self, state, rc_default):
        hs = state.atomParams.handles
        T = len(hs)
        eps = np.zeros((T, T))
        sig = np.zeros((T, T))
        rc = np.zeros((T, T))
        for i, a in enumerate(hs):
            for j, b in enumerate(hs):
                e = self.p["eps"].get((a, b))
                s = self.p["sig"].get((a, b))
                ea, eb = self.p["eps"].get((a, a), 0.0), self.p["eps"].get((b, b), 0.0)
                sa, sb = self.p["sig"].get((a, a), 0.0), self.p["sig"].get((b, b), 0.0)
                if e is None:
                    e = math.sqrt(ea * eb)
                if s is None:
                    s = math.sqrt(sa * sb) if self.mixingRules == "geometric" else 0.5 * (sa + sb)
                r = self.p["rCut"].get((a, b))
                if r is None:
                    r = max(self.p["rCut"].get((a, a), rc_default), self.p["rCut"].get((b, b), rc_default))
                eps[i, j], sig[i, j], rc[i, j] = e, s, r
        return rc * rc, 24.0 * eps, sig ** 6


class FixChargeEwald(Fix):
    """P3M long range plus Ewald real space fused into the pair kernel.
    setParameters(sz, r_cut=-1, interpolation_order=5) or
    setParameters(szx, szy, szz, r_cut, interpolation_order); the splitting
    parameter is solved from the Deserno-Holm error estimates as in
    FixChargeEwald::find_optimal_parameters (FixChargeEwald.cu:853-967)."""

    def __init__(self, state, handle, groupHandle="all"):
        super().__init__(state, handle, groupHandle)
        self.sz = (64, 64, 64)
        self.r_cut = -1.0
        self.order = 5
        self.longRangeInterval = 1

    def setParameters(self, *args, **kw):
        # (sz, r_cut, order) or (szx, szy, szz, r_cut, order), told apart by
        # the argument count as the two bound overloads are
        if len(args) >= 4:
            self.sz = tuple(int(a) for a in args[:3])
            rest = list(args[3:])
        else:
            self.sz = (int(args[0]),) * 3
            rest = list(args[1:])
        self.r_cut = float(rest[0]) if len(rest) > 0 else float(kw.get("r_cut", -1.0))
        self.order = int(rest[1]) if len(rest) > 1 else int(kw.get("interpolation_order", 5))
        for s in self.sz:
            if s < 32 or s > 1024 or (s & (s - 1)):
                raise ValueError("mesh sizes must be powers of two in 32..1024")
        if self.order not in (1, 3, 5, 7):
            raise ValueError("interpolation order must be 1, 3, 5 or 7")
        self.state._eng = None

    def setLongRangeInterval(self, interval=0):
        if interval not in (0, 1):
            raise NotImplementedError("long-range forces are evaluated every step")
        self.longRangeInterval = 1
        return 1


class FixE3B(Fix):
    """E3B3 (default) or E3B2 three-body water (FixE3B.cu:515-593): molecules
    are added as (O, H1, H2, M) ids."""

    def __init__(self, state, handle, style="E3B3"):
        super().__init__(state, handle)
        if style not in ("E3B3", "E3B2"):
            raise ValueError("style: E3B3 or E3B2")
        self.style = style
        self.molecules = []

    def addMolecule(self, O, H1, H2, M):
        self.molecules.append((int(O), int(H1), int(H2), int(M)))
        self.state._eng = None

    @property
    def nMolecules(self):
        return len(self.molecules)

    def params(self, state):
        if state.units.unitType != "REAL":
            raise ValueError("FixE3B needs real units")
        k = 0.23900573614  # kJ/mol -> kcal/mol
        if self.style == "E3B3":
            e2, ea, eb, ec = 453000.0, 150.0, -1005.0, 1880.0
        else:
            e2, ea, eb, ec = 2349000.0, 1745.7, -4565.0, 7606.8
        # k2, k3, rs, rf are shared by both styles (FixE3B.cu:515-593)
        return [e2 * k, ea * k, eb * k, ec * k, 4.872, 1.907, 5.0, 5.2]


class FixRigid(Fix):
    """Rigid TIP4P/2005 molecules (SETTLE / RATTLE, M site from the hydrogens)."""

    def __init__(self, state, handle="rigid", style="TIP4P/2005"):
        super().__init__(state, handle)
        if style != "TIP4P/2005":
            raise NotImplementedError("only TIP4P/2005 rigid water")
        self.style = style
        self.molecules = []

    def createRigid(self, O, H1, H2, M=-1):
        self.molecules.append((int(O), int(H1), int(H2), int(M)))
        self.state._eng = None


class FixNVTRescale(Fix):
    """FixNVTRescale(state, handle, groupHandle, temp, applyEvery=10): sets the
    temperature to `temp` every applyEvery turns."""

    def __init__(self, state, handle, groupHandle, temp, applyEvery=10):
        super().__init__(state, handle, groupHandle)
        self.temp = float(temp)
        self.applyEvery = int(applyEvery)


class FixNoseHoover(Fix):
    """FixNoseHoover(state, handle, groupHandle) with setTemperature(T,
    timeConstant): Nose-Hoover chain of length 10 (FixNoseHoover.cu:120-1098)."""

    def __init__(self, state, handle, groupHandle="all"):
        super().__init__(state, handle, groupHandle)
        self.temp = None
        self.timeConstant = None
        self.chainLength = 10

    def setTemperature(self, temperature, timeConstant):
        self.temp = float(temperature)
        self.timeConstant = float(timeConstant)
        self.state._eng = None

    def setPressure(self, *args, **kw):
        raise NotImplementedError("use FixPressureBerendsen for pressure control")


class FixPressureBerendsen(Fix):
    """FixPressureBerendsen(state, handle, pressure, period, applyEvery):
    isotropic Berendsen dilation of the box and the molecules' centres of mass
    (bulkModulus 10 and maxDilation 1e-5 by default, as the reference)."""

    def __init__(self, state, handle, pressure, period, applyEvery=10):
        super().__init__(state, handle)
        self.pressure = float(pressure)
        self.period = float(period)
        self.applyEvery = int(applyEvery)
        self.bulkModulus = 10.0
        self.maxDilation = 1e-5

    def setParameters(self, maxDilation=-1):
        if maxDilation > 0:
            self.maxDilation = float(maxDilation)


class IntegratorVerlet:
    """run(numTurns) advances the state and returns particle-timesteps per
    second, as IntegratorVerlet::run does (IntegratorVerlet.cu:649-770)."""

    def __init__(self, state):
        self.state = state

    def run(self, numTurns):
        st = self.state
        eng = st._engine()
        if st.turn == 0 or getattr(st, "_dirty", True):
            st._push()
            st._dirty = False
        t0 = time.perf_counter()
        eng.run(int(numTurns), st._flags)
        eng.sync()
        dt = time.perf_counter() - t0
        st.turn += int(numTurns)
        st._pull()
        return len(st.atoms) * numTurns / dt if dt > 0 else 0.0


class InitializeAtoms:
    @staticmethod
    def initTemp(state, groupHandle, temp):
        """InitializeAtoms::initTemp (src/InitializeAtoms.cpp:99-150):
        N(0, sqrt(1/m)) per component for massive atoms, centre-of-mass
        velocity removed, rescaled to `temp` with 3N degrees of freedom."""
        m = np.array([a.mass for a in state.atoms])
        seed = int(state._rng.integers(1 << 31))
        v = systems.maxwell_boltzmann(m, temp, seed, boltz=state.units.boltz,
                                      mvv_to_eng=state.units.mvv_to_eng)
        for a in state.atoms:
            a.vel = Vector(*v[a.id])
        state._dirty = True


class EvaluatorLJ:
    """force(sigma, epsilon, rcut, dr) / energy(sigma, epsilon, rcut, r):
    evaluated by the engine's pair kernel on a two-atom system (the reference
    calls its device evaluator on the host, PairEvaluatorLJ.cu)."""

    def _pair(self, sigma, epsilon, rcut, dr):
        dr = Vector.of(dr)
        L = 4.0 * (rcut + abs(dr.v).max() + 1.0)
        eng = _lib.Engine(0)
        x = np.array([[0.5 * L] * 3, (np.full(3, 0.5 * L) - dr.v)])
        eng.set_system(x, np.zeros(3), np.full(3, L), inv_mass=np.ones(2))
        eng.nlist_config(rcut + 0.5, 0.5)
        eng.lj_set([[rcut * rcut]], [[24.0 * epsilon]], [[sigma ** 6]])
        eng.compute(_lib.PAIR | _lib.ENERGY)
        f = eng.forces()
        en, _ = eng.energies()
        return f[0], en[0]

    def force(self, sigma, epsilon, rcut, dr):
        f, _ = self._pair(sigma, epsilon, rcut, dr)
        return Vector(*f)

    def energy(self, sigma, epsilon, rcut, r):
        _, e = self._pair(sigma, epsilon, rcut, Vector(r, 0.0, 0.0))
        return e  # total of the two atoms' halves = the pair energy


def create_TIP4P_2005(state, oxygenHandle, hydrogenHandle, mSiteHandle, center=None, orientation=None,
                      **kwargs):
    """util_py/water.py:196-230: four sites O, H1, H2, M with q_H = 0.5564,
    q_M = -2 q_H, either at given positions (OPos, H1Pos, H2Pos, MPos) or
    built about `center` on the rigid geometry (random orientation from the
    state's RNG when orientation == 'random')."""
    p = systems.TIP4P2005
    qh, qm = p["q_H"], p["q_M"]
    if {"OPos", "H1Pos", "H2Pos", "MPos"} <= set(kwargs):
        pos = [kwargs["OPos"], kwargs["H1Pos"], kwargs["H2Pos"], kwargs["MPos"]]
    else:
        c = Vector.of(center if center is not None else (0, 0, 0)).v
        fr = systems.tip4p2005_frame()
        if orientation == "random":
            R = systems._random_rotations(state._rng, 1)[0]
            fr = fr @ R.T
        pos = [c + fr[k] for k in range(4)]
    ids = [state.addAtom(oxygenHandle, pos[0], 0.0), state.addAtom(hydrogenHandle, pos[1], qh),
           state.addAtom(hydrogenHandle, pos[2], qh), state.addAtom(mSiteHandle, pos[3], qm)]
    return state.createMolecule(ids)
