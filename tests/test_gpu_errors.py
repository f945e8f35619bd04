"""Error paths forced from the host: the fixed-point force overflow.

The mixed form accumulates forces in 64-bit fixed point at 2^32 (AMBER), so a
single pair force beyond 2^31 kcal/mol/A cannot be represented.  Two water
oxygens of different molecules 0.05 A apart (neither excluded nor 1-4, so
their LJ repulsion is ~6e23 kcal/mol/A) make every kernel finish normally and
set the overflow flag; the call returns MDGPU_ERR_OVERFLOW.  Checked:
  * mdgpu_eval returns the error, and the same context then evaluates the
    original coordinates bit-identically to a fresh context;
  * mdgpu_md_run whose first step drives the pair to 0.05 A returns the error
    and leaves x, v and f at their entry values, bit for bit.
"""
import numpy as np
import pytest

from md_amd import native as N

pytestmark = pytest.mark.gpu

KW = dict(lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-5, pme_grid=(32, 32, 32))
OVERFLOW = 4   # MDGPU_ERR_OVERFLOW
# polymer_2k: 300 polymer atoms, then 600 waters (O, H, H); the first two water oxygens
O1, O2 = 300, 303


def _mixed(top):
    ev = N.Evaluator(0)
    ev.upload(top.desc)
    ev.set_params(N.make_params(N.SEM_AMBER, N.PREC_MIXED, **KW))
    return ev


def _top(polymer_2k):
    prm, rst = polymer_2k
    top = N.Topology(prm, rst, N.SEM_AMBER)
    a = top.arrays
    assert a["type"][O1] == a["type"][O2]
    ex = a["excl_list"][a["excl_start"][O1]:a["excl_start"][O1 + 1]]
    assert O2 not in set(ex.tolist())
    return top


def test_eval_overflow_then_recovers(polymer_2k):
    top = _top(polymer_2k)
    print("case overflow: two water oxygens 0.05 A apart, mixed", flush=True)
    x_bad = top.coords.copy()
    x_bad[O2] = x_bad[O1] + np.array([0.05, 0.0, 0.0])
    ev = _mixed(top)
    try:
        with pytest.raises(N.GpuError) as ei:
            ev.evaluate(x_bad, top.box)
        assert ei.value.status == OVERFLOW, str(ei.value)
        e1, f1, _ = ev.evaluate(top.coords, top.box)
    finally:
        ev.close()
    ev = _mixed(top)
    try:
        e0, f0, _ = ev.evaluate(top.coords, top.box)
    finally:
        ev.close()
    assert np.array_equal(f1, f0)
    assert all(e1[k] == e0[k] for k in e0)


def test_md_overflow_leaves_entry_state(polymer_2k):
    import torch
    top = _top(polymer_2k)
    print("case overflow in md_run: first step closes a pair from 1.5 A to 0.05 A", flush=True)
    dt = 1e-4
    x0 = top.coords.copy()
    x0[O2] = x0[O1] + np.array([1.5, 0.0, 0.0])
    dev = torch.device("cuda:0")
    x = torch.tensor(x0.reshape(-1), dtype=torch.float64, device=dev)
    f = torch.zeros_like(x)
    ev = _mixed(top)
    try:
        ev.md_set_masses(top.masses)
        ev.evaluate_device(x.data_ptr(), top.box, f.data_ptr())   # forces at the entry state
        torch.cuda.synchronize()
        # velocities that cancel the first half kick of the two oxygens, so
        # that the drift of the first step puts them 0.05 A apart
        acc = 418.4 * f.cpu().numpy().reshape(-1, 3) / top.masses[:, None]
        v0 = np.zeros_like(x0)
        v0[O1] = -0.5 * dt * acc[O1]
        v0[O2] = np.array([-1.45 / dt, 0.0, 0.0]) - 0.5 * dt * acc[O2]
        v = torch.tensor(v0.reshape(-1), dtype=torch.float64, device=dev)
        xs, vs, fs = x.cpu().numpy().copy(), v.cpu().numpy().copy(), f.cpu().numpy().copy()
        assert np.all(np.isfinite(fs))
        with pytest.raises(N.GpuError) as ei:
            ev.md_run(x.data_ptr(), v.data_ptr(), f.data_ptr(), top.box, dt, 3)
        assert ei.value.status == OVERFLOW, str(ei.value)
        torch.cuda.synchronize()
        assert np.array_equal(x.cpu().numpy(), xs)
        assert np.array_equal(v.cpu().numpy(), vs)
        assert np.array_equal(f.cpu().numpy(), fs)
        # the context still runs dynamics from a sound state
        v.zero_()
        tr, _ = ev.md_run(x.data_ptr(), v.data_ptr(), f.data_ptr(), top.box, dt, 2)
        assert np.all(np.isfinite(tr[:, :2]))
    finally:
        ev.close()
