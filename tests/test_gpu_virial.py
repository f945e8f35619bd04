"""Full virial on the GPU (mdgpu_set_virial / mdgpu_get_virial) and the MD
pressure.

Bars:
  * FP64 full virial (pair, bonded, correction and reciprocal parts) against
    the oracle's: 1e-9 of the largest component (AMBER, PME; polymer and water);
  * mixed against FP64: 1e-5 of the largest component;
  * forming the virial changes no energy or force (mixed bit for bit, FP64 to
    1e-12);
  * the MD trace's pressure is (2 KE + trace W) / 3V of the step's virial.
"""
import numpy as np
import pytest
import torch

import oracle
from md_amd import native as N

from md_systems import thermal_velocities

pytestmark = pytest.mark.gpu

CASES = [("polymer_2k", 10.0, 9.0, (32, 32, 32)), ("water_6k", 9.0, 8.0, (48, 48, 48))]


def _eval(top, prm, virial):
    ev = N.Evaluator(0)
    try:
        ev.upload(top.desc)
        ev.set_params(prm)
        ev.set_virial(virial)
        e, f, v = ev.evaluate(top.coords, top.box)
        w = ev.virial() if virial else None
        return e, f, w
    finally:
        ev.close()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_full_virial_matches_oracle(case, request):
    fix, rc, sw, grid = case
    print(f"case {fix} full virial", flush=True)
    prm_path, rst_path = request.getfixturevalue(fix)
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    kw = dict(lj_cutoff=rc, lj_switch=sw, ewald_tol=1e-5, pme_grid=grid)
    p64 = N.make_params(N.SEM_AMBER, N.PREC_FP64, **kw)
    p32 = N.make_params(N.SEM_AMBER, N.PREC_MIXED, **kw)
    e64, f64, w64 = _eval(top, p64, True)
    _, _, wo = oracle.evaluate(top.desc, p64, top.coords, top.box, skip_all_pairs=True, full_virial=True)
    scale = np.abs(wo).max()
    print(f"  FP64 vs oracle {np.abs(w64 - wo).max() / scale:.2e}")
    assert np.abs(w64 - wo).max() <= 1e-9 * scale
    e32, f32, w32 = _eval(top, p32, True)
    print(f"  mixed vs FP64 {np.abs(w32 - w64).max() / scale:.2e}")
    assert np.abs(w32 - w64).max() <= 1e-5 * scale
    # the statistics instance leaves energies and forces as they are
    e32n, f32n, _ = _eval(top, p32, False)
    assert np.array_equal(f32, f32n) and all(e32[k] == e32n[k] for k in e32 if k != "total" or True)
    e64n, f64n, _ = _eval(top, p64, False)
    assert np.abs(f64 - f64n).max() <= 1e-12 * np.abs(f64n).max()


def test_md_pressure(polymer_2k):
    print("case polymer_2k MD pressure", flush=True)
    prm_path, rst_path = polymer_2k
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    prm = N.make_params(N.SEM_AMBER, N.PREC_MIXED, lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-5,
                        pme_grid=(32, 32, 32))
    ev = N.Evaluator(0)
    try:
        ev.upload(top.desc)
        ev.set_params(prm)
        ev.set_virial(True)
        ev.md_set_masses(top.masses)
        x = torch.tensor(top.coords.reshape(-1), dtype=torch.float64, device="cuda:0")
        v = torch.tensor(thermal_velocities(top.masses, 5).reshape(-1), dtype=torch.float64, device="cuda:0")
        f = torch.zeros_like(x)
        ev.evaluate_device(x.data_ptr(), top.box, f.data_ptr())
        tr, _ = ev.md_run(x.data_ptr(), v.data_ptr(), f.data_ptr(), top.box, 0.0005, 5)
        w = ev.virial()
        vol = float(np.prod(top.box))
        want = (2.0 * tr[-1, 1] + np.trace(w)) / (3.0 * vol) * 69476.95
        print(f"  pressure trace {tr[:, 2]} bar")
        assert np.all(np.isfinite(tr[:, 2]))
        assert tr[-1, 2] == pytest.approx(want, rel=1e-12)
        ev.set_virial(False)
        tr2, _ = ev.md_run(x.data_ptr(), v.data_ptr(), f.data_ptr(), top.box, 0.0005, 2)
        assert np.all(np.isnan(tr2[:, 2]))
    finally:
        ev.close()
