"""Pinning of the oracle's velocity Verlet (CPU only).

  * a harmonic bond released at rest follows velocity Verlet's exact discrete
    solution r_n - r0 = A cos(n theta), cos theta = 1 - (omega dt)^2 / 2, with
    omega^2 = 2 k 418.4 / mu (a wrong unit factor, mass or force constant
    changes theta; a first-order update breaks the form), and the centre of
    mass stays put;
  * reversing the velocities retraces the trajectory (time reversibility);
  * the energy fluctuation of the oscillator shrinks fourfold when dt halves
    (second order).
"""
import numpy as np

import oracle
from md_amd import native as N

from md_systems import discrete_stretch, harmonic_pair, oscillator_params, thermal_velocities


def _forces(desc, prm, x, box):
    _, f, _ = oracle.evaluate(desc, prm, x, box, skip_all_pairs=True)
    return f


def test_harmonic_bond_follows_the_discrete_solution():
    desc, keep, x0, box, m, w = harmonic_pair()
    prm = oscillator_params()
    dt, n, amp = 0.001, 400, 0.1
    f0 = _forces(desc, prm, x0, box)
    x, v, f, tr = oracle.md_run(desc, prm, m, x0, np.zeros_like(x0), f0, box, dt, n)
    want = discrete_stretch(w, dt, n, amp)
    # replay step by step to see every r_n
    xs, vs, fs = x0.copy(), np.zeros_like(x0), f0
    got = []
    for _ in range(n):
        xs, vs, fs, _ = oracle.md_run(desc, prm, m, xs, vs, fs, box, dt, 1)
        got.append(xs[1, 0] - xs[0, 0] - 1.5)
    got = np.array(got)
    assert np.abs(got - want).max() <= 1e-9, np.abs(got - want).max()
    assert np.allclose(x, xs, rtol=0, atol=1e-12)
    com = (m[:, None] * x).sum(0) / m.sum()
    com0 = (m[:, None] * x0).sum(0) / m.sum()
    assert np.abs(com - com0).max() <= 1e-12
    # energy: potential + kinetic stays near k A^2
    e = tr.sum(1)
    assert abs(e.mean() - 300.0 * amp ** 2) <= 0.02 * 300.0 * amp ** 2


def test_velocity_reversal_retraces():
    desc, keep, x0, box, m, w = harmonic_pair(stretch=0.05)
    prm = oscillator_params()
    v0 = thermal_velocities(m, 3)
    f0 = _forces(desc, prm, x0, box)
    x1, v1, f1, _ = oracle.md_run(desc, prm, m, x0, v0, f0, box, 0.0005, 50)
    x2, v2, _, _ = oracle.md_run(desc, prm, m, x1, -v1, f1, box, 0.0005, 50)
    assert np.abs(x2 - x0).max() <= 1e-10
    assert np.abs(v2 + v0).max() <= 1e-8 * np.abs(v0).max()


def test_energy_error_is_second_order():
    desc, keep, x0, box, m, w = harmonic_pair()
    prm = oscillator_params()
    f0 = _forces(desc, prm, x0, box)
    spread = []
    for dt, n in ((0.001, 100), (0.0005, 200)):
        _, _, _, tr = oracle.md_run(desc, prm, m, x0, np.zeros_like(x0), f0, box, dt, n)
        e = tr.sum(1)
        spread.append(e.max() - e.min())
    ratio = spread[0] / spread[1]
    assert 3.6 <= ratio <= 4.4, ratio
