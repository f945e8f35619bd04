"""Pinning of the oracle's full virial (CPU only).

W_ab = -dE/d(strain_ab): bonded terms sum_i x_i f_i over their own forces,
pairs sum g d_a d_b, plus the reciprocal-space part of the PME pass.  Checked
against central finite differences of the total energy under strain, with an
Ewald tolerance of 1e-12 so that no pair crossing the cutoff makes the energy
jump:
  * isotropic and axial strains of the 2.1k-atom periodic polymer with PME
    (the trace and the diagonal);
  * shears x_a += h x_b of an isolated polymer in a large box without
    reciprocal space (the off-diagonal entries; no pair crosses a face).
"""
import os

import numpy as np
import pytest

import oracle
from md_amd import native as N
from md_amd import systems


def _fd(energy, h=1e-6):
    return -(energy(h) - energy(-h)) / (2 * h)


def test_diagonal_virial_is_strain_derivative(polymer_2k):
    prm_path, rst_path = polymer_2k
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-12,
                        pme_grid=(32, 32, 32))
    x, L = top.coords, np.asarray(top.box)
    _, _, w = oracle.evaluate(top.desc, prm, x, L, skip_all_pairs=True, full_virial=True)

    def e_at(scale):
        return oracle.evaluate(top.desc, prm, x * scale, L * scale, skip_all_pairs=True,
                               want_forces=False)[0]["total"]

    iso = _fd(lambda h: e_at(np.full(3, 1.0 + h)))
    assert iso == pytest.approx(np.trace(w), rel=1e-7)
    for a in range(3):
        def axial(h, a=a):
            s = np.ones(3)
            s[a] += h
            return e_at(s)
        assert _fd(axial) == pytest.approx(w[a, a], rel=1e-7)
    assert np.allclose(w, w.T, rtol=0, atol=1e-9 * np.abs(w).max())


def test_shear_virial_of_an_isolated_polymer(tmp_path):
    top_g = systems.solvated_polymer(seed=7, box=80.0, n_poly=120, n_water=0, n_cmap=8, n_types=6)
    c = top_g.coords - top_g.coords.mean(axis=0) + 40.0
    assert c.min() > 10.0 and c.max() < 70.0  # 9 A cutoff: no pair reaches a periodic image
    top_g.coords = c
    p, r = systems.write_system(top_g, str(tmp_path / "iso"))
    top = N.Topology(p, r, N.SEM_AMBER)
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, lj_cutoff=9.0, lj_switch=8.0, ewald_tol=1e-12,
                        pme_grid=(0, 0, 0))
    x, L = top.coords.copy(), np.asarray(top.box)
    _, _, w = oracle.evaluate(top.desc, prm, x, L, skip_all_pairs=True, full_virial=True)
    assert np.abs(w).max() > 1.0
    for a, b in ((0, 1), (0, 2), (1, 2)):
        def shear(h, a=a, b=b):
            xs = x.copy()
            xs[:, a] += h * x[:, b]
            return oracle.evaluate(top.desc, prm, xs, L, skip_all_pairs=True, want_forces=False)[0]["total"]
        # -dE/dh = sum_i x_ib F_ia = W_ba
        assert _fd(shear) == pytest.approx(w[b, a], rel=1e-6, abs=1e-6 * np.abs(w).max())
