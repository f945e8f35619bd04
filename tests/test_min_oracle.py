"""Pinning of the oracle's L-BFGS minimiser (CPU only).

  * a stretched harmonic bond relaxes to r = r0; after the first (steepest
    descent) step one stored pair makes H0 = s.y / y.y exact for the single
    coordinate, so the next step lands on the minimum;
  * an uncharged bonded chain reaches its known minimum energy 0 with the RMS
    force at the tolerance;
  * the energy never rises from one accepted iteration to the next (Armijo);
  * no atom moves more than max_step in one iteration.
"""
import numpy as np

import oracle
from md_amd import native as N

from md_systems import bonded_chain, harmonic_pair, oscillator_params


def test_harmonic_bond_relaxes_in_two_steps():
    desc, keep, x0, box, m, w = harmonic_pair(stretch=0.1)
    x, en, res = oracle.minimize(desc, oscillator_params(), x0, box, max_iter=10, rms_tol=1e-9, max_step=1.0)
    r = np.linalg.norm(x[1] - x[0])
    assert res["converged"]
    assert abs(r - 1.5) <= 1e-10
    assert res["iterations"] <= 3, res


def test_bonded_chain_reaches_zero_energy():
    desc, keep, x0, box = bonded_chain()
    prm = oscillator_params()
    x, en, res = oracle.minimize(desc, prm, x0, box, max_iter=2000, rms_tol=1e-6, max_step=0.2)
    assert res["converged"], res
    assert en[0] > 10.0
    assert res["energy"] <= 1e-9, res
    assert np.all(np.diff(en) <= 0.0)


def test_energy_never_rises_and_steps_are_capped(polymer_2k):
    prm_path, rst_path = polymer_2k
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-5,
                        pme_grid=(32, 32, 32))
    x0 = np.asarray(top.coords, dtype=np.float64).reshape(-1, 3)
    xs, steps = x0, []
    for k in range(6):
        x1, en, res = oracle.minimize(top.desc, prm, xs, top.box, max_iter=1, rms_tol=0.0, max_step=0.05)
        steps.append(np.linalg.norm(x1 - xs, axis=1).max())
        xs = x1
    assert max(steps) <= 0.05 + 1e-12
    _, en, res = oracle.minimize(top.desc, prm, x0, top.box, max_iter=25, rms_tol=0.0, max_step=0.2)
    assert len(en) == 26
    assert np.all(np.diff(en) <= 0.0)
    assert en[-1] < en[0] - 100.0
