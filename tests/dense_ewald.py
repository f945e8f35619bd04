"""Dense FP64 Ewald electrostatics in PyTorch, differentiable in the charges.

Test infrastructure: an implementation independent of the oracle and of the
GPU kernels, used to check their charge gradients (dE/dq) on small systems.

  E = sum_{i<j, not excluded, r < rc} k q_i q_j erfc(a r)/r          (direct)
    + sum_{m != 0, |m_d| <= mmax} k exp(-pi^2 m^2/a^2)/(2 pi V m^2) |S(m)|^2
    - k a/sqrt(pi) sum_i q_i^2                                      (self)
    - sum_{excluded non-1-4 pairs} k q_i q_j erf(a r)/r
    + sum_{1-4 pairs} k q_i q_j (1/(scee r) - erf(a r)/r)
with S(m) = sum_i q_i exp(2 pi i m.x_i), m = (m_x/L_x, m_y/L_y, m_z/L_z).
"""
import math

import numpy as np
import torch


def _min_image(d, box):
    return d - box * torch.round(d / box)


def energy(q, x, box, alpha, rc, mmax, kc, excl_start=None, excl_list=None, excl_pairs=None, pairs14=None,
           scee14=None):
    x = torch.as_tensor(np.asarray(x), dtype=torch.float64)
    L = torch.as_tensor(np.asarray(box), dtype=torch.float64)
    n = x.shape[0]
    i, j = torch.triu_indices(n, n, 1)
    d = _min_image(x[i] - x[j], L)
    r = torch.sqrt((d * d).sum(1))
    keep = r < rc
    if excl_start is not None:
        ex = set()
        for a in range(n):
            for b in excl_list[excl_start[a]:excl_start[a + 1]]:
                ex.add((min(a, int(b)), max(a, int(b))))
        mask = torch.tensor([(int(a), int(b)) not in ex for a, b in zip(i.tolist(), j.tolist())])
        keep = keep & mask
    i, j, r = i[keep], j[keep], r[keep]
    e_dir = (kc * q[i] * q[j] * torch.erfc(alpha * r) / r).sum()
    # reciprocal: structure factors over the integer cube
    rng = torch.arange(-mmax, mmax + 1, dtype=torch.float64)
    mx, my, mz = torch.meshgrid(rng, rng, rng, indexing="ij")
    m = torch.stack([mx.reshape(-1), my.reshape(-1), mz.reshape(-1)], 1)
    m = m[(m != 0).any(1)] / L
    m2 = (m * m).sum(1)
    V = float(L.prod())
    g = kc * torch.exp(-math.pi ** 2 * m2 / alpha ** 2) / (2.0 * math.pi * V * m2)
    ph = 2.0 * math.pi * (x @ m.T)  # (n, M)
    sr = (q[:, None] * torch.cos(ph)).sum(0)
    si = (q[:, None] * torch.sin(ph)).sum(0)
    e_rec = (g * (sr * sr + si * si)).sum()
    e_self = -kc * alpha / math.sqrt(math.pi) * (q * q).sum()
    e_corr = torch.zeros((), dtype=torch.float64)
    if excl_pairs is not None and len(excl_pairs):
        p = torch.as_tensor(np.asarray(excl_pairs).reshape(-1, 2), dtype=torch.long)
        dd = _min_image(x[p[:, 0]] - x[p[:, 1]], L)
        rr = torch.sqrt((dd * dd).sum(1))
        e_corr = e_corr - (kc * q[p[:, 0]] * q[p[:, 1]] * torch.erf(alpha * rr) / rr).sum()
    if pairs14 is not None and len(pairs14):
        p = torch.as_tensor(np.asarray(pairs14).reshape(-1, 2), dtype=torch.long)
        s = torch.as_tensor(np.asarray(scee14), dtype=torch.float64)
        dd = _min_image(x[p[:, 0]] - x[p[:, 1]], L)
        rr = torch.sqrt((dd * dd).sum(1))
        e_corr = e_corr + (kc * q[p[:, 0]] * q[p[:, 1]] * (1.0 / (s * rr) - torch.erf(alpha * rr) / rr)).sum()
    return e_dir + e_rec + e_self + e_corr


def charge_gradient(arrays, coords, box, alpha, rc, mmax, kc):
    """(E, dE/dq) of the dense Ewald electrostatics of a system's arrays."""
    q = torch.tensor(arrays["charge"], dtype=torch.float64, requires_grad=True)
    e = energy(q, coords, box, alpha, rc, mmax, kc, arrays.get("excl_start"), arrays.get("excl_list"),
               arrays.get("excl_pairs"), arrays.get("pairs14"), arrays.get("scee14"))
    (g,) = torch.autograd.grad(e, q)
    return float(e), g.numpy()
