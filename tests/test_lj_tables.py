"""Combining rule of the autograd binding (md_amd.autograd.lj_tables), CPU.

Checks: the per-type sigma, epsilon of the 6,450-atom water box reproduce the
topology's LJ tables (A = 4 eps sigma^12, B = 4 eps sigma^6, closed form);
gradients through a type without LJ (TIP3P hydrogen, eps = 0) are finite and
the oxygen's gradient equals its closed form; torch.autograd.gradcheck on
positive parameters.
"""
import numpy as np
import torch

from md_amd import autograd as AG
from md_amd import native as N


def _water(water_6k):
    p, r = water_6k
    return N.Topology(p, r, N.SEM_AMBER)


def _per_type(a):
    t = int(a["n_types"])
    idx = np.asarray(a["nb_index"]).reshape(t, t) - 1
    sig, eps = np.zeros(t), np.zeros(t)
    for i in range(t):
        k = idx[i, i]
        A, B = (a["lj_a"][k], a["lj_b"][k]) if k >= 0 else (0.0, 0.0)
        if A > 0 and B > 0:
            sig[i] = (A / B) ** (1.0 / 6.0)
            eps[i] = B * B / (4.0 * A)
    return sig, eps


def test_lj_tables_reproduce_water_tables_and_gradients(water_6k):
    top = _water(water_6k)
    a = top.arrays
    sig, eps = _per_type(a)
    o = int(a["type"][0])    # the first atom is a water oxygen
    h = int(a["type"][1])
    assert eps[h] == 0.0 and eps[o] > 0.0
    s = torch.tensor(sig, dtype=torch.float64, requires_grad=True)
    e = torch.tensor(eps, dtype=torch.float64, requires_grad=True)
    A, B = AG.lj_tables(s, e, a["nb_index"], a["lj_a"].size)
    np.testing.assert_allclose(A.detach().numpy(), a["lj_a"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(B.detach().numpy(), a["lj_b"], rtol=1e-12, atol=0)
    # d(sum A + sum B)/d(eps_O), d/d(sigma_O): only the O-O entry depends on them
    (A.sum() + B.sum()).backward()
    assert torch.all(torch.isfinite(s.grad)) and torch.all(torch.isfinite(e.grad))
    so, eo = sig[o], eps[o]
    np.testing.assert_allclose(e.grad[o].item(), 4 * so ** 12 + 4 * so ** 6, rtol=1e-12)
    # sigma_O enters s = (s_O + s_O)/2 = s_O
    np.testing.assert_allclose(s.grad[o].item(), 48 * eo * so ** 11 + 24 * eo * so ** 5, rtol=1e-12)
    assert e.grad[h].item() == 0.0


def test_lj_tables_gradcheck():
    nb = np.array([[1, 2, 3], [2, 4, 5], [3, 5, 6]], dtype=np.int32)
    s = torch.tensor([3.1, 2.4, 1.9], dtype=torch.float64, requires_grad=True)
    e = torch.tensor([0.15, 0.05, 0.11], dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda s_, e_: AG.lj_tables(s_, e_, nb, 6), (s, e))
