"""FP64 restatement of the Nose-Hoover chain thermostat (TEST INFRASTRUCTURE
ONLY: imported by tests/ as the checker, never by the product path).

Follows FixNoseHoover (src/Fixes/FixNoseHoover.cu) without a barostat:

* chain masses, ``updateThermalMasses`` (FixNoseHoover.cu:1034-1042):
  Q_0 = ndf kT / w^2, Q_k = kT / w^2 with w = 1 / timeConstant
  (``setTemperature``, FixNoseHoover.cu:221-224);
* start, ``prepareFinal`` (FixNoseHoover.cu:241-283): chain velocities 0,
  G_0 = (K2 - ndf kT) / Q_0, G_k = (Q_{k-1} v_{k-1}^2 - kT) / Q_k, with K2
  the instantaneous sum m v^2 (``calculateKineticEnergy``,
  FixNoseHoover.cu:1044-1053: ``totalKEScalar`` is mvv_to_eng times the sum
  of m v^2, DataComputerTemperature.cu:116-121);
* half step, ``thermostatIntegrate`` (FixNoseHoover.cu:795-866) with
  nTimesteps = n_ys = 1 (the constructor's values, FixNoseHoover.cu:121-124):
  the chain is swept from its end to v_0 with quarter-step force kicks
  between eighth-step damping factors exp(-dt/8 v_{k+1}), the particle
  velocities scale by exp(-dt/2 v_0) (K2 by its square), and the sweep runs
  back up the chain with refreshed forces.  ``stepInit`` (FixNoseHoover.cu:
  481-496) calls it with the stored K2 and ``stepFinal`` (FixNoseHoover.cu:
  592-597) after measuring K2.  The reference defers the velocity scale of
  the final half to the next step's ``rescale`` (FixNoseHoover.cu:1076-1096);
  no force is evaluated in between, so scaling at once gives the same
  trajectory.
* The chain positions are commented out in the reference
  (FixNoseHoover.cu:829-832); they are integrated here (dt/2 v_k per half
  step) because they enter the conserved extended energy
  K + U + sum_k Q_k v_k^2 / 2 + ndf kT xi_0 + sum_{k>0} kT xi_k.

Pinned in tests/test_oracle_thermostat.py, independently of this text: the
extended energy of a thermostatted set of harmonic oscillators is conserved
to O(dt^2) (a 4x smaller drift at dt / 2); one half step followed by the same
half step with -dt restores the chain exactly (the splitting is
palindromic); the time-averaged kinetic temperature of a long run equals the
set point.
"""
from __future__ import annotations

import numpy as np


class Chain:
    """State of one Nose-Hoover chain: velocities v, forces G, masses Q,
    positions xi, the current K2 = sum m v^2 (energy units) and the last
    velocity scale factor."""

    def __init__(self, M, ndf, kT, freq):
        self.M = int(M)
        self.ndf = float(ndf)
        self.kT = float(kT)
        self.v = np.zeros(M)
        self.G = np.zeros(M)
        self.xi = np.zeros(M)
        self.Q = np.full(M, kT / (freq * freq))
        self.Q[0] = ndf * kT / (freq * freq)
        self.k2 = 0.0
        self.scale = 1.0

    def start(self, k2):
        """prepareFinal: forces from the measured K2 and zero chain velocities."""
        self.k2 = float(k2)
        self.G[0] = (self.k2 - self.ndf * self.kT) / self.Q[0]
        for k in range(1, self.M):
            self.G[k] = (self.Q[k - 1] * self.v[k - 1] ** 2 - self.kT) / self.Q[k]

    def half(self, dt, k2=None):
        """thermostatIntegrate over dt/2; k2 given: refresh K2 and G_0 first
        (the stepFinal half).  Returns the particle velocity scale factor."""
        M, v, G, Q, kT = self.M, self.v, self.G, self.Q, self.kT
        nkt = self.ndf * kT
        if k2 is not None:
            self.k2 = float(k2)
            G[0] = (self.k2 - nkt) / Q[0]
        t2, t4, t8 = 0.5 * dt, 0.25 * dt, 0.125 * dt
        v[M - 1] += t4 * G[M - 1]
        for k in range(M - 2, 0, -1):
            pf = np.exp(-t8 * v[k + 1])
            v[k] = (v[k] * pf + t4 * G[k]) * pf
        pf = np.exp(-t8 * v[1])
        v[0] = (v[0] * pf + t4 * G[0]) * pf
        s = np.exp(-t2 * v[0])
        self.k2 *= s * s
        self.xi += t2 * v
        v[0] *= pf
        G[0] = (self.k2 - nkt) / Q[0]
        v[0] = (v[0] + t4 * G[0]) * pf
        for k in range(1, M - 1):
            pf = np.exp(-t8 * v[k + 1])
            v[k] *= pf
            G[k] = (Q[k - 1] * v[k - 1] ** 2 - kT) / Q[k]
            v[k] = (v[k] + t4 * G[k]) * pf
        G[M - 1] = (Q[M - 2] * v[M - 2] ** 2 - kT) / Q[M - 1]
        v[M - 1] += t4 * G[M - 1]
        self.scale = float(s)
        return self.scale

    def energy(self):
        """Thermostat part of the conserved extended energy."""
        e = 0.5 * float((self.Q * self.v * self.v).sum())
        e += self.ndf * self.kT * self.xi[0] + self.kT * float(self.xi[1:].sum())
        return e

    def state(self):
        """Packed like the device chain block: v | G | Q | xi | K2 | scale."""
        return np.concatenate([self.v, self.G, self.Q, self.xi, [self.k2, self.scale]])


def water_md_nh(box, v, nsteps, T, time_constant, chain=10, dt=2.0, rc=6.5, skin=0.5, sz=(32, 32, 32), order=5,
                e3b_on=False):
    """FP64 rigid-water MD under a Nose-Hoover chain, step order of
    IntegratorVerlet::run (src/Integrators/IntegratorVerlet.cu:694-754):
    stepInit (chain half, scale), kick-drift + SETTLE, forces, kick + RATTLE,
    stepFinal (measure K2, chain half, scale).  Degrees of freedom of rigid
    water: 6 per molecule - 3.  Returns (x, v, chain)."""
    import oracle
    from md_engine_amd import systems

    x = np.ascontiguousarray(box["x"], dtype=np.float64).copy()
    v = np.ascontiguousarray(v, dtype=np.float64).copy()
    ftm, mvv, kb = systems.REAL_FTM_TO_V, systems.REAL_MVV_TO_ENG, systems.REAL_BOLTZ
    m = box["mass"]
    nmol = len(box["mol"])
    ch = Chain(chain, 6 * nmol - 3, kb * T, 1.0 / time_constant)

    def k2():
        return float((m[:, None] * v * v).sum()) * mvv

    ch.start(k2())
    f, _, alpha = oracle.water_forces(box, x, rc, skin, sz, order, e3b_on)
    for _ in range(nsteps):
        v *= ch.half(dt)
        oracle.water_step_half(1, box["mol"], x, v, f, dt, ftm, box["L"])
        f, _, _ = oracle.water_forces(box, x, rc, skin, sz, order, e3b_on, alpha=alpha)
        oracle.water_step_half(2, box["mol"], x, v, f, dt, ftm, box["L"])
        v *= ch.half(dt, k2())
    return x, v, ch
