"""L-BFGS minimisation on the GPU (mdgpu_minimize), and NVE dynamics from a
minimised state.

Bars:
  * a stretched harmonic bond relaxes to r0 (1e-10 A) in at most three
    iterations, as the oracle's does; an uncharged bonded chain reaches
    energy 0 (FP64 1e-9, mixed 1e-5);
  * on the 2.1k-atom polymer the FP64 minimiser follows the oracle's: the
    same iteration and evaluation counts and the final energy to 1e-9
    relative after 12 iterations;
  * with a Verlet list the result is the one without (FP64 1e-9 relative);
  * minimising the clashing 6.45k-atom water box lowers its energy and its
    largest force by orders of magnitude, and NVE dynamics from there conserves
    energy with an error that shrinks about fourfold when dt halves.
"""
import numpy as np
import pytest
import torch

import oracle
from md_amd import native as N

from md_systems import bonded_chain, harmonic_pair, oscillator_params, thermal_velocities

pytestmark = pytest.mark.gpu


def _dev(x):
    t = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64).reshape(-1)).to("cuda:0")
    torch.cuda.synchronize()
    return t


def _ev(desc, prm, skin=0.0):
    ev = N.Evaluator(0)
    ev.upload(desc)
    ev.set_params(prm)
    ev.set_verlet(skin)
    return ev


def test_harmonic_bond_relaxes():
    print("case harmonic bond minimisation", flush=True)
    desc, keep, x0, box, m, w = harmonic_pair(stretch=0.1)
    ev = _ev(desc, oscillator_params())
    try:
        xd = _dev(x0)
        res = ev.minimize(xd.data_ptr(), box, max_iter=10, rms_force_tol=1e-9, max_step=1.0)
        x = xd.cpu().numpy().reshape(-1, 3)
    finally:
        ev.close()
    _, _, ro = oracle.minimize(desc, oscillator_params(), x0, box, max_iter=10, rms_tol=1e-9, max_step=1.0)
    print(f"  {res}")
    assert res["converged"] and res["iterations"] == ro["iterations"] <= 3
    assert abs(np.linalg.norm(x[1] - x[0]) - 1.5) <= 1e-10


@pytest.mark.parametrize("prec,tol", [(N.PREC_FP64, 1e-9), (N.PREC_MIXED, 1e-5)], ids=["fp64", "mixed"])
def test_bonded_chain_reaches_zero(prec, tol):
    print(f"case bonded chain prec={prec}", flush=True)
    desc, keep, x0, box = bonded_chain()
    ev = _ev(desc, oscillator_params(prec))
    try:
        xd = _dev(x0)
        res = ev.minimize(xd.data_ptr(), box, max_iter=2000, rms_force_tol=1e-6 if prec == N.PREC_FP64 else 1e-3,
                          max_step=0.2)
    finally:
        ev.close()
    print(f"  {res}")
    assert res["energy"] <= tol


def _polymer(polymer_2k, prec=N.PREC_FP64):
    prm_path, rst_path = polymer_2k
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    prm = N.make_params(N.SEM_AMBER, prec, lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-5, pme_grid=(32, 32, 32))
    return top, prm, np.asarray(top.coords, dtype=np.float64).reshape(-1, 3)


def test_fp64_minimiser_follows_oracle(polymer_2k):
    print("case polymer_2k FP64 minimiser vs oracle", flush=True)
    top, prm, x0 = _polymer(polymer_2k)
    ev = _ev(top.desc, prm)
    try:
        xd = _dev(x0)
        res = ev.minimize(xd.data_ptr(), top.box, max_iter=12, rms_force_tol=0.0, max_step=0.2)
        x = xd.cpu().numpy().reshape(-1, 3)
    finally:
        ev.close()
    xo, en, ro = oracle.minimize(top.desc, prm, x0, top.box, max_iter=12, rms_tol=0.0, max_step=0.2)
    print(f"  gpu {res}\n  oracle {ro}")
    assert res["iterations"] == ro["iterations"] and res["evaluations"] == ro["evaluations"]
    assert abs(res["energy"] - ro["energy"]) <= 1e-9 * abs(ro["energy"])
    assert np.abs(x - xo).max() <= 1e-7


def test_verlet_list_minimisation(polymer_2k):
    print("case polymer_2k minimiser with a Verlet list", flush=True)
    top, prm, x0 = _polymer(polymer_2k)
    out = []
    for skin in (0.0, 1.0):
        ev = _ev(top.desc, prm, skin)
        try:
            xd = _dev(x0)
            out.append((ev.minimize(xd.data_ptr(), top.box, max_iter=15, rms_force_tol=0.0), ev.list_builds()))
        finally:
            ev.close()
    (r0, b0), (r1, b1) = out
    print(f"  builds {b0} / {b1}")
    assert b1[0] < b0[0]
    assert r0["iterations"] == r1["iterations"]
    assert abs(r0["energy"] - r1["energy"]) <= 1e-9 * abs(r0["energy"])


def test_minimised_water_conserves_energy(water_6k):
    print("case water_6k minimise, then NVE", flush=True)
    prm_path, rst_path = water_6k
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    prm = N.make_params(N.SEM_AMBER, N.PREC_MIXED, lj_cutoff=9.0, lj_switch=8.0, ewald_tol=1e-5,
                        pme_grid=(48, 48, 48))
    ev = _ev(top.desc, prm, 1.0)
    try:
        x0 = np.asarray(top.coords, dtype=np.float64).reshape(-1, 3)
        xd = _dev(x0)
        fd = torch.zeros_like(xd)
        e0, _ = ev.evaluate_device(xd.data_ptr(), top.box, fd.data_ptr())
        res = ev.minimize(xd.data_ptr(), top.box, max_iter=300, rms_force_tol=0.5, max_step=0.1)
        print(f"  {e0['total']:.1f} -> {res}")
        assert res["energy"] < e0["total"] - 0.5 * abs(e0["total"])
        assert res["max_force"] < 1e3
        xm = xd.cpu().numpy().copy()
        ev.md_set_masses(top.masses)
        v0 = thermal_velocities(top.masses, 12, temperature=100.0).reshape(-1)
        spread = []
        for dt, n in ((0.0005, 80), (0.00025, 160)):
            xd.copy_(torch.from_numpy(xm))
            vd = _dev(v0)
            torch.cuda.synchronize()
            ev.evaluate_device(xd.data_ptr(), top.box, fd.data_ptr())
            tr, _ = ev.md_run(xd.data_ptr(), vd.data_ptr(), fd.data_ptr(), top.box, dt, n)
            et = tr[:, 0] + tr[:, 1]
            spread.append(et.max() - et.min())
            print(f"  dt {dt * 1000:.2f} fs: total energy spread {spread[-1]:.4f}, kinetic {tr[-1, 1]:.1f}")
    finally:
        ev.close()
    ratio = spread[0] / spread[1]
    print(f"  spread ratio {ratio:.2f}")
    assert 2.5 <= ratio <= 5.5


def test_cpp_driver_minimize(polymer_2k):
    """md_main --minimize: the reference's driver with its commented-out
    sys.minimize() line enabled, through the C++ System API (AMBER semantics,
    rc 10 A, switch 9 A, skin 2 A, Ewald tolerance 1e-6, the reference's PME
    grid formula rounded up to 2^a 3^b).  The same minimisation through the C
    ABI from Python gives the same iterations and energy."""
    import math
    import os
    import re
    import subprocess
    from md_amd import systems
    print("case md_main --minimize 20 polymer_2k", flush=True)
    prm_path, rst_path = polymer_2k
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "md_amd", "md_main")
    out = subprocess.run([exe, prm_path, rst_path, "--minimize", "20"], capture_output=True, text=True, timeout=120)
    print(out.stdout[-500:], out.stderr[-300:])
    assert out.returncode == 0
    e_cpp = float(re.search(r"Minimized:.*\nEnergy: (\S+)", out.stdout).group(1))
    it_cpp = int(re.search(r"Minimized: iterations (\d+)", out.stdout).group(1))
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    tol, rc = 1e-6, 10.0
    alpha = math.sqrt(-math.log(2.0 * tol)) / rc
    grid = tuple(systems.smallest_2a3b(int(2.0 * alpha * L / (3.0 * tol ** 0.2))) for L in top.box)
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, lj_cutoff=rc, lj_switch=9.0, ewald_alpha=alpha, pme_grid=grid)
    ev = _ev(top.desc, prm, 2.0)
    try:
        xd = _dev(top.coords)
        res = ev.minimize(xd.data_ptr(), top.box, max_iter=20, rms_force_tol=1e-2, max_step=0.2)
    finally:
        ev.close()
    print(f"  C++ driver {e_cpp:.6f} in {it_cpp} iterations; C ABI {res}")
    assert it_cpp == res["iterations"] == 20
    # the driver prints six significant digits
    assert abs(e_cpp - res["energy"]) <= 1e-5 * abs(res["energy"])
