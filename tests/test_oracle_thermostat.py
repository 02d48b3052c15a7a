"""Pins of the oracle's Nose-Hoover chain (oracle/thermostat.py) by what the
mathematics fixes, not by its own formulas: time reversibility of the
palindromic splitting, O(dt^2) conservation of the extended energy of a
thermostatted harmonic system, and the canonical mean temperature."""
import numpy as np
import pytest

from oracle.thermostat import Chain


def _random_chain(seed, M=6):
    rng = np.random.default_rng(seed)
    ch = Chain(M, ndf=30.0, kT=0.7, freq=0.3)
    ch.v[:] = rng.normal(0, 0.2, M)
    ch.xi[:] = rng.normal(0, 1.0, M)
    ch.start(30.0 * 0.7 * rng.uniform(0.5, 1.5))
    return ch


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_half_step_is_time_reversible(seed):
    ch = _random_chain(seed)
    before = ch.state().copy()
    ch.half(0.37)
    assert not np.allclose(ch.state()[:-1], before[:-1])
    ch.half(-0.37)
    after = ch.state()
    np.testing.assert_allclose(after[:-1], before[:-1], rtol=1e-12, atol=1e-12)


def test_reversibility_detects_a_wrong_damping_index():
    """A plausible slip (damping v_k by exp(-dt/8 v_k) instead of v_{k+1})
    breaks the reversibility the test above checks."""
    class Slipped(Chain):
        def half(self, dt, k2=None):
            M, v, G, Q, kT = self.M, self.v, self.G, self.Q, self.kT
            t4, t8 = 0.25 * dt, 0.125 * dt
            v[M - 1] += t4 * G[M - 1]
            for k in range(M - 2, 0, -1):
                pf = np.exp(-t8 * v[k])
                v[k] = (v[k] * pf + t4 * G[k]) * pf
            for k in range(1, M - 1):
                pf = np.exp(-t8 * v[k])
                v[k] *= pf
                G[k] = (Q[k - 1] * v[k - 1] ** 2 - kT) / Q[k]
                v[k] = (v[k] + t4 * G[k]) * pf
            G[M - 1] = (Q[M - 2] * v[M - 2] ** 2 - kT) / Q[M - 1]
            v[M - 1] += t4 * G[M - 1]
            return 1.0

    ch = Slipped(6, ndf=30.0, kT=0.7, freq=0.3)
    ch.v[:] = np.linspace(-0.3, 0.4, 6)
    ch.start(25.0)
    before = ch.state().copy()
    ch.half(0.37)
    ch.half(-0.37)
    assert np.abs(ch.state()[:6] - before[:6]).max() > 1e-6


def _oscillators(dt, t_end, seed=3, n=64, M=4, T=1.0, tau=2.0):
    """Velocity Verlet of n 3-D harmonic oscillators (unit spring, masses in
    [1, 3]) between the two chain halves, as IntegratorVerlet orders them.
    Returns the extended energy along the run and the mean temperature."""
    rng = np.random.default_rng(seed)
    m = rng.uniform(1.0, 3.0, n)[:, None]
    x = rng.normal(0, 1.0, (n, 3))
    v = rng.normal(0, 1.0, (n, 3)) / np.sqrt(m) * 1.5
    ndf = 3 * n
    ch = Chain(M, ndf=ndf, kT=T, freq=1.0 / tau)
    k2 = lambda: float((m * v * v).sum())
    ch.start(k2())
    f = -x
    H, temps = [], []
    for _ in range(int(round(t_end / dt))):
        v *= ch.half(dt)
        v += 0.5 * dt * f / m
        x += dt * v
        f = -x
        v += 0.5 * dt * f / m
        v *= ch.half(dt, k2())
        H.append(0.5 * k2() + 0.5 * float((x * x).sum()) + ch.energy())
        temps.append(k2() / ndf)
    return np.array(H), float(np.mean(temps))


def test_extended_energy_is_conserved_to_second_order():
    h1, _ = _oscillators(0.05, 40.0)
    h2, _ = _oscillators(0.025, 40.0)
    d1, d2 = np.ptp(h1), np.ptp(h2)
    assert d1 < 2e-3 * abs(h1[0])
    ratio = d1 / d2
    assert 3.0 < ratio < 5.0, ratio


def test_mean_temperature_is_the_set_point():
    _, tmean = _oscillators(0.05, 2000.0, n=128, T=1.3, tau=1.0)
    assert tmean == pytest.approx(1.3, rel=0.03)
