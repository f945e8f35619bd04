"""Force-field parameters as differentiable inputs of the GPU energy.

``FieldEnergy.apply`` maps parameter tensors and a stack of frames to per-frame
total energies (AMBER semantics) computed by the mdgpu kernels; its backward
calls the hand-written adjoint (mdgpu_adjoint) once per frame, with the
frame's incoming gradient as the weight of the total energy.  Nothing from the
forward pass is stored on the device: the adjoint recomputes the pair list
and every pair contribution.

The parameters are the arrays of the uploaded system (``PARAMS``): charges,
LJ and 1-4 LJ table entries (A, B per type pair), bonded force constants and
equilibria per type, and the CMAP grid values.  Tensors are float64 on the
CPU; gradients come back the same way.  Combining rules (sigma, epsilon per
type to A, B per pair) are ordinary torch code on top, see ``lj_tables``.
"""
from __future__ import annotations

import numpy as np
import torch

from . import native as N

# differentiable parameter arrays of the system description (CMAP: grid values)
PARAMS = ["charge", "lj_a", "lj_b", "lj14_a", "lj14_b", "bond_k", "bond_r0", "ub_k", "ub_r0", "angle_k",
          "angle_theta0", "dihedral_k", "dihedral_phase", "improper_k", "improper_psi0", "cmap_grid"]


class Model:
    """One system on one GPU context; parameters are replaced per call."""

    def __init__(self, topology: N.Topology, params: N.Params, device: int = 0):
        if params.semantics != N.SEM_AMBER:
            raise ValueError("parameter gradients are defined for AMBER semantics")
        self.top = topology
        self.params = params
        self.ev = N.Evaluator(device)
        self.base = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in topology.arrays.items()}
        self._keep = []
        self._loaded = None

    def initial(self) -> dict:
        """The topology's parameter values as float64 tensors."""
        return {k: torch.tensor(self.base[k], dtype=torch.float64) for k in PARAMS}

    def _upload(self, values: dict):
        arr = dict(self.base)
        for k in PARAMS:
            v = values.get(k)
            if v is not None:
                arr[k] = v.detach().cpu().numpy().astype(np.float64).reshape(-1)
        res, ns = int(arr["cmap_resolution"]), int(arr["cmap_n_sets"])
        if ns and arr["cmap_set"].size:
            arr["cmap_coeff"] = N.cmap_coefficients(arr["cmap_grid"], res, ns)
        keep = []
        desc = N.desc_from_arrays(arr, keep)
        self.ev.upload(desc)
        self.ev.set_params(self.params)
        self._keep, self._desc = keep, desc

    def energies(self, values: dict, frames, box) -> np.ndarray:
        self._upload(values)
        return np.array([self.ev.evaluate(np.asarray(f), box)[0]["total"] for f in frames])

    def adjoint(self, frames, box, weights) -> dict:
        """sum_f w_f dE_f/dtheta for every entry of PARAMS (numpy arrays)."""
        acc = None
        for f, w in zip(frames, weights):
            if w == 0.0:
                continue
            g = self.ev.adjoint(np.asarray(f), box, {"total": float(w)}, self._desc)
            acc = g if acc is None else {k: acc[k] + g[k] for k in acc}
        if acc is None:
            acc = {k: np.zeros(n) for k, n in N.grad_lengths(self._desc).items()}
        res, ns = int(self.base["cmap_resolution"]), int(self.base["cmap_n_sets"])
        out = {k: acc[k] for k in PARAMS if k != "cmap_grid"}
        out["cmap_grid"] = (N.cmap_grid_grad(acc["cmap_coeff"], res, ns) if acc["cmap_coeff"].size
                            else np.zeros(self.base["cmap_grid"].size))
        return out

    def close(self):
        self.ev.close()


class FieldEnergy(torch.autograd.Function):
    """E_f = total energy of frame f under the given parameters."""

    @staticmethod
    def forward(ctx, model: Model, frames: torch.Tensor, box, *values):
        named = dict(zip(PARAMS, values))
        fr = frames.detach().cpu().numpy().astype(np.float64)
        e = model.energies(named, fr, np.asarray(box, dtype=np.float64))
        ctx.model, ctx.frames, ctx.box = model, fr, np.asarray(box, dtype=np.float64)
        return torch.tensor(e, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_e: torch.Tensor):
        # the parameters of the last forward are still uploaded in the context
        g = ctx.model.adjoint(ctx.frames, ctx.box, grad_e.detach().cpu().numpy().astype(np.float64))
        return (None, None, None) + tuple(torch.tensor(g[k], dtype=torch.float64) for k in PARAMS)


def energy(model: Model, values: dict, frames: torch.Tensor, box) -> torch.Tensor:
    """Per-frame total energies, differentiable in every tensor of ``values``."""
    vals = [values[k] for k in PARAMS]
    return FieldEnergy.apply(model, frames, box, *vals)


def lj_tables(sigma: torch.Tensor, epsilon: torch.Tensor, nb_index: np.ndarray, n_lj: int):
    """Lorentz-Berthelot A, B per LJ table entry from per-type sigma, epsilon
    (A = 4 eps s^12, B = 4 eps s^6, s = (s_i + s_j)/2, eps = sqrt(eps_i eps_j));
    entries no type pair maps to stay 0.  A pair with eps_i eps_j = 0 (a type
    without LJ, such as a TIP3P hydrogen) gets eps = 0 and a zero gradient: the
    square root is taken only of positive products, so its derivative at 0
    cannot turn the other type's gradient into NaN."""
    t = sigma.numel()
    idx = np.asarray(nb_index).reshape(t, t) - 1
    a = torch.zeros(n_lj, dtype=torch.float64)
    b = torch.zeros(n_lj, dtype=torch.float64)
    ii, jj = np.nonzero(idx >= 0)
    first = {}
    for i, j in zip(ii.tolist(), jj.tolist()):
        first.setdefault(int(idx[i, j]), (i, j))
    ent = torch.tensor(sorted(first), dtype=torch.long)
    ti = torch.tensor([first[k][0] for k in sorted(first)], dtype=torch.long)
    tj = torch.tensor([first[k][1] for k in sorted(first)], dtype=torch.long)
    s = 0.5 * (sigma[ti] + sigma[tj])
    prod = epsilon[ti] * epsilon[tj]
    pos = prod > 0
    e = torch.where(pos, torch.sqrt(torch.where(pos, prod, torch.ones_like(prod))), torch.zeros_like(prod))
    s6 = s ** 6
    a = a.index_put((ent,), 4.0 * e * s6 * s6)
    b = b.index_put((ent,), 4.0 * e * s6)
    return a, b
