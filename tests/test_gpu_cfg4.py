"""The batched workload at full size, as the benchmark runs it.

256 water boxes (the 6,450-atom box's water model), atom counts uniform in
[3,000, 30,000] (seeds 4000 + i), cubic boxes at 0.1 atoms/A^3, PME grid the
smallest 2^a 3^b >= L / 1 A (32^3 ... 72^3), AMBER semantics, rc 9 A (switch
8 A), Ewald tol 1e-5: 4.28 M atoms in one batch.

Checked:
  * mixed form: the one graph launch (two replays) equals the serial mode
    (one synchronous evaluation per system) bit for bit, energies and forces;
  * FP64 form, one graph launch: the smallest, the median and the largest
    system against the oracle, per term |dE| <= 1e-9 |E| (+1e-12) and force
    relative RMS <= 1e-8.
"""
import numpy as np
import pytest
import torch

import oracle
from md_amd import native as N
from md_amd import systems

pytestmark = pytest.mark.gpu

TERMS = ["bond", "angle", "vdw", "ewald_direct", "ewald_self", "ewald_recip", "ewald_excl", "total"]


def _params(prec, k):
    return N.make_params(N.SEM_AMBER, prec, lj_cutoff=9.0, lj_switch=8.0, ewald_tol=1e-5, pme_grid=(k, k, k))


@pytest.fixture(scope="module")
def workload():
    work = systems.batch_workload(256)
    keep = []
    descs = [N.desc_from_arrays(w[0], keep) for w in work]
    x = np.concatenate([w[1] for w in work]).reshape(-1)
    yield work, descs, x, keep


def _batch(work, descs, prec):
    b = N.Batch(0)
    for (arrays, coords, box, k), d in zip(work, descs):
        b.add(d, _params(prec, k), box)
    return b


def test_cfg4_graph_equals_serial_mixed(workload):
    work, descs, x, _ = workload
    print(f"case 256-system batch mixed, {x.size // 3} atoms", flush=True)
    dev = torch.device("cuda:0")
    xd = torch.tensor(x, dtype=torch.float64, device=dev)
    fd = torch.zeros_like(xd)
    b = _batch(work, descs, N.PREC_MIXED)
    try:
        b.configure(mode=1)
        e_ser = b.evaluate_device(xd.data_ptr(), fd.data_ptr())
        torch.cuda.synchronize()
        f_ser = fd.cpu().numpy()
        b.configure(mode=0)
        for rep in range(2):
            fd.zero_()
            torch.cuda.synchronize()
            e_g = b.evaluate_device(xd.data_ptr(), fd.data_ptr())
            torch.cuda.synchronize()
            assert b.graph_state() == 1
            assert np.array_equal(fd.cpu().numpy(), f_ser), f"replay {rep}"
            for i, (a, s) in enumerate(zip(e_g, e_ser)):
                assert all(a[k] == s[k] for k in s), (rep, i)
    finally:
        b.close()


def test_cfg4_fp64_smallest_median_largest_match_oracle(workload):
    work, descs, x, _ = workload
    print(f"case 256-system batch fp64 vs oracle, {x.size // 3} atoms", flush=True)
    dev = torch.device("cuda:0")
    xd = torch.tensor(x, dtype=torch.float64, device=dev)
    fd = torch.zeros_like(xd)
    b = _batch(work, descs, N.PREC_FP64)
    try:
        e_g = b.evaluate_device(xd.data_ptr(), fd.data_ptr())
        torch.cuda.synchronize()
        assert b.graph_state() == 1
        f_all = fd.cpu().numpy()
    finally:
        b.close()
    sizes = np.array([w[1].shape[0] for w in work])
    off = np.concatenate([[0], np.cumsum(sizes)])
    order = np.argsort(sizes, kind="stable")
    bad = []
    for i in (int(order[0]), int(order[len(order) // 2]), int(order[-1])):
        arrays, coords, box, k = work[i]
        print(f"  system {i}: {sizes[i]} atoms, box {box:.2f} A, PME {k}^3", flush=True)
        eo, fo, _ = oracle.evaluate(descs[i], _params(N.PREC_FP64, k), coords, [box] * 3, skip_all_pairs=True)
        f = f_all[3 * off[i]:3 * off[i + 1]].reshape(-1, 3)
        for t in TERMS:
            err = abs(e_g[i][t] - eo[t])
            if not err <= 1e-9 * abs(eo[t]) + 1e-12:
                bad.append((i, t, e_g[i][t], eo[t]))
        rms = float(np.sqrt(np.sum((f - fo) ** 2) / np.sum(fo ** 2)))
        print(f"    total rel {abs(e_g[i]['total'] - eo['total']) / abs(eo['total']):.2e}, force rel-RMS {rms:.2e}")
        if not rms <= 1e-8:
            bad.append((i, "forces", rms))
    assert not bad, bad
