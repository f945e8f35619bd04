"""The allocation-free rebuild query (mdgpu_pairs_need_rebuild_noalloc)
against the context's query (mdgpu_pairs_need_rebuild), through the C ABI.

Two calls in a row share the GPU's scratch array, the second on a smaller
system: the 6,450-atom water box, then the 2,100-atom polymer.  For each, a
context with a Verlet skin of 2 A builds its list at x0; the allocation-free
query gets x0 as its reference positions.  At three displacement sets (small
Gaussian moves; one atom moved 1.2 A; one atom moved by a whole box edge plus
0.1 A, which the minimum image folds back) the decision must equal the
context's bit for bit, and the largest squared displacement must equal the
same arithmetic in numpy exactly.  A size beyond MDGPU_NOALLOC_MAX_ATOMS is
refused before anything is enqueued.
"""
import numpy as np
import pytest
import torch

from md_amd import native as N

pytestmark = pytest.mark.gpu

SKIN = 2.0


def _max_d2(x, x0, box):
    d = x - x0
    q = d / box
    r = np.trunc(q + np.copysign(0.5, q))   # C round(): half away from zero
    d = d - box * r
    return float(np.max(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]))


def _check_system(prm, rst, seed):
    top = N.Topology(prm, rst, N.SEM_AMBER)
    n = top.n_atoms
    box = np.asarray(top.box, dtype=np.float64)
    print(f"case noalloc query, {n} atoms", flush=True)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    x0h = top.coords.copy()
    x0 = torch.tensor(x0h.reshape(-1), dtype=torch.float64, device=dev)
    f = torch.zeros_like(x0)
    ev = N.Evaluator(0)
    try:
        ev.upload(top.desc)
        ev.set_params(N.make_params(N.SEM_AMBER, N.PREC_MIXED, lj_cutoff=9.0, lj_switch=8.0, ewald_tol=1e-5,
                                    pme_grid=(32, 32, 32)))
        ev.set_verlet(SKIN)
        ev.evaluate_device(x0.data_ptr(), box, f.data_ptr())   # builds the list at x0
        torch.cuda.synchronize()
        rng = np.random.default_rng(seed)
        x1 = x0h + rng.normal(0.0, 0.05, x0h.shape)
        x2 = x1.copy()
        x2[n // 2, 0] += 1.2
        x3 = x1.copy()
        x3[n // 3, 1] += box[1] + 0.1
        want_decisions = []
        for xh in (x1, x2, x3):
            x = torch.tensor(xh.reshape(-1), dtype=torch.float64, device=dev)
            want = ev.needs_rebuild(x.data_ptr())
            needs = torch.full((1,), 7, dtype=torch.int32, device=dev)
            m = torch.full((1,), -1.0, dtype=torch.float64, device=dev)
            N.pairs_need_rebuild_noalloc(0, x.data_ptr(), x0.data_ptr(), n, box, SKIN, stream, needs.data_ptr(),
                                         m.data_ptr())
            torch.cuda.synchronize()
            got, got_m = int(needs.item()), float(m.item())
            exp_m = _max_d2(xh, x0h, box)
            print(f"  context {int(want)} noalloc {got} max d2 {got_m!r} numpy {exp_m!r}")
            assert got == int(want)
            assert got_m == exp_m
            want_decisions.append(got)
        assert want_decisions == [0, 1, 0]
        # refused before anything is enqueued: the outputs keep their values
        needs = torch.full((1,), 7, dtype=torch.int32, device=dev)
        with pytest.raises(N.GpuError) as ei:
            N.pairs_need_rebuild_noalloc(0, x0.data_ptr(), x0.data_ptr(), N.NOALLOC_MAX_ATOMS + 1, box, SKIN,
                                         stream, needs.data_ptr(), 0)
        assert ei.value.status == 1
        torch.cuda.synchronize()
        assert int(needs.item()) == 7
    finally:
        ev.close()
        top.close()


def test_noalloc_query_matches_context_query(water_6k, polymer_2k):
    _check_system(*water_6k, seed=91)
    _check_system(*polymer_2k, seed=92)   # the second call, on a smaller system
