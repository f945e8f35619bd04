"""Molecular dynamics on the GPU (mdgpu_md_run): velocity Verlet, NVE.

Bars:
  * a harmonic bond released at rest follows velocity Verlet's exact discrete
    solution (FP64 form 1e-9 A; mixed form, FP32 bonded arithmetic and
    fixed-point forces, 2e-6 A);
  * on the 2.1k-atom polymer (AMBER, PME) the FP64 trajectory matches the
    oracle's velocity Verlet: positions 1e-9 A, velocities 1e-8 relative,
    potential 1e-9 and kinetic energy 1e-10 relative per step;
  * with a Verlet list the trajectory is the one without (FP64 1e-9 A) and the
    list is built fewer times than there are steps;
  * the mixed form is bit-reproducible and retraces its path when the
    velocities are reversed (1e-6 A);
  * a run whose list overflows restarts from the entry state and gives the
    same result as one that never overflowed.
"""
import numpy as np
import pytest
import torch

import oracle
from md_amd import native as N

from md_systems import discrete_stretch, harmonic_pair, oscillator_params, thermal_velocities

pytestmark = pytest.mark.gpu


class Md:
    def __init__(self, desc, prm, masses, skin=0.0):
        self.ev = N.Evaluator(0)
        self.ev.upload(desc)
        self.ev.set_params(prm)
        self.ev.set_verlet(skin)
        self.ev.md_set_masses(masses)
        n3 = 3 * int(desc.n_atoms)
        dev = torch.device("cuda", 0)
        self.x = torch.zeros(n3, dtype=torch.float64, device=dev)
        self.v = torch.zeros_like(self.x)
        self.f = torch.zeros_like(self.x)

    def start(self, x, v, box):
        self.box = np.asarray(box, dtype=np.float64)
        self.x.copy_(torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64).reshape(-1)))
        self.v.copy_(torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64).reshape(-1)))
        torch.cuda.synchronize()
        e, _ = self.ev.evaluate_device(self.x.data_ptr(), self.box, self.f.data_ptr())
        torch.cuda.synchronize()
        return e

    def run(self, dt, n):
        tr, last = self.ev.md_run(self.x.data_ptr(), self.v.data_ptr(), self.f.data_ptr(), self.box, dt, n)
        torch.cuda.synchronize()
        return tr, last

    def state(self):
        return (self.x.cpu().numpy().reshape(-1, 3).copy(), self.v.cpu().numpy().reshape(-1, 3).copy(),
                self.f.cpu().numpy().reshape(-1, 3).copy())

    def close(self):
        self.ev.close()


@pytest.mark.parametrize("prec,tol", [(N.PREC_FP64, 1e-9), (N.PREC_MIXED, 2e-6)], ids=["fp64", "mixed"])
def test_harmonic_bond(prec, tol):
    print(f"case harmonic bond prec={prec}", flush=True)
    desc, keep, x0, box, m, w = harmonic_pair()
    md = Md(desc, oscillator_params(prec), m)
    try:
        md.start(x0, np.zeros_like(x0), box)
        dt, n = 0.001, 120
        got = []
        for _ in range(n):
            md.run(dt, 1)
            x, _, _ = md.state()
            got.append(x[1, 0] - x[0, 0] - 1.5)
        err = np.abs(np.array(got) - discrete_stretch(w, dt, n, 0.1)).max()
        print(f"  max deviation from the discrete solution {err:.3e} A")
        assert err <= tol
    finally:
        md.close()


def _polymer(polymer_2k, prec):
    prm_path, rst_path = polymer_2k
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    prm = N.make_params(N.SEM_AMBER, prec, lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-5, pme_grid=(32, 32, 32))
    x0 = np.asarray(top.coords, dtype=np.float64).reshape(-1, 3)
    v0 = thermal_velocities(top.masses, 9)
    return top, prm, x0, v0


def test_fp64_trajectory_matches_oracle(polymer_2k):
    print("case polymer_2k FP64 trajectory vs oracle", flush=True)
    top, prm, x0, v0 = _polymer(polymer_2k, N.PREC_FP64)
    dt, n = 0.0005, 12
    md = Md(top.desc, prm, top.masses)
    try:
        md.start(x0, v0, top.box)
        tr, _ = md.run(dt, n)
        x, v, f = md.state()
    finally:
        md.close()
    _, f0, _ = oracle.evaluate(top.desc, prm, x0, top.box, skip_all_pairs=True)
    xo, vo, fo, tro = oracle.md_run(top.desc, prm, top.masses, x0, v0, f0, top.box, dt, n)
    dx = np.abs(x - xo).max()
    dv = np.abs(v - vo).max() / np.abs(vo).max()
    print(f"  positions {dx:.2e} A, velocities {dv:.2e}")
    assert dx <= 1e-9
    assert dv <= 1e-8
    assert np.all(np.abs(tr[:, 0] - tro[:, 0]) <= 1e-9 * np.abs(tro[:, 0]))
    assert np.all(np.abs(tr[:, 1] - tro[:, 1]) <= 1e-10 * np.abs(tro[:, 1]))


def test_verlet_list_trajectory(polymer_2k):
    print("case polymer_2k FP64 trajectory with a Verlet list", flush=True)
    top, prm, x0, v0 = _polymer(polymer_2k, N.PREC_FP64)
    dt, n = 0.001, 40
    out = []
    for skin in (0.0, 0.8):
        md = Md(top.desc, prm, top.masses, skin)
        try:
            md.start(x0, v0, top.box)
            tr, _ = md.run(dt, n)
            out.append((md.state(), tr, md.ev.list_builds()))
        finally:
            md.close()
    (s0, tr0, b0), (s1, tr1, b1) = out
    print(f"  builds without / with skin: {b0} / {b1}")
    assert b0 == (n + 1, n + 1)
    assert 2 <= b1[0] < n // 2 and b1[1] == n + 1
    assert np.abs(s0[0] - s1[0]).max() <= 1e-9
    assert np.all(np.abs(tr0[:, 0] - tr1[:, 0]) <= 1e-9 * np.abs(tr0[:, 0]))


def test_mixed_reproducible_and_reversible(polymer_2k):
    print("case polymer_2k mixed reproducibility and reversal", flush=True)
    top, prm, x0, v0 = _polymer(polymer_2k, N.PREC_MIXED)
    dt, n = 0.0005, 20
    runs = []
    for rep in range(2):
        md = Md(top.desc, prm, top.masses, 1.0)
        try:
            md.start(x0, v0, top.box)
            tr, _ = md.run(dt, n)
            x1, v1, f1 = md.state()
            runs.append((x1, v1, tr))
            if rep == 0:
                # reverse: start from (x1, -v1) with the forces at x1
                md.v.mul_(-1.0)
                torch.cuda.synchronize()
                md.run(dt, n)
                x2, v2, _ = md.state()
                dx = np.abs(x2 - x0).max()
                print(f"  reversal returns within {dx:.2e} A")
                assert dx <= 1e-6
        finally:
            md.close()
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    assert np.array_equal(runs[0][2], runs[1][2], equal_nan=True)


def test_overflow_restart(water_6k):
    print("case water_6k restart after list overflow", flush=True)
    prm_path, rst_path = water_6k
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    prm = N.make_params(N.SEM_AMBER, N.PREC_MIXED, lj_cutoff=9.0, lj_switch=8.0, ewald_tol=1e-5,
                        pme_grid=(48, 48, 48))
    x0 = np.asarray(top.coords, dtype=np.float64).reshape(-1, 3)
    v0 = thermal_velocities(top.masses, 4)
    res = []
    for skin in (0.0, 5.0):
        # the list capacity is sized at the first evaluation (no skin); a
        # skin of 5 A needs about twice as many tiles, so the first step of
        # the run overflows and the run restarts with a larger list
        md = Md(top.desc, prm, top.masses, 0.0)
        try:
            md.start(x0, v0, top.box)
            md.ev.set_verlet(skin)
            tr, _ = md.run(0.00005, 4)
            res.append((md.state(), tr))
        finally:
            md.close()
    (sa, tra), (sb, trb) = res
    assert np.abs(sa[0] - sb[0]).max() <= 1e-6
    assert np.all(np.abs(tra[:, 0] - trb[:, 0]) <= 1e-6 * np.abs(tra[:, 0]))
