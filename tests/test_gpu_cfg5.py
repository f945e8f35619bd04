"""Differentiation workload on seeded frames of the 6,450-atom water box.

Frames f = 0..3: the water box (2,150 three-site waters, 40 A cube) with every
coordinate perturbed by N(0, 0.1 A), seeds 5000 + f (the first four of the
benchmark's 64).  AMBER semantics, rc 9 A (switch 8 A), Ewald tol 1e-5, PME
48^3.  Loss L = sum_f (E_f - E_target,f)^2, with E_target from every parameter
array scaled by exp(N(0, 0.05)) (seed 5999).

The gradient comes from the torch.autograd.Function, whose backward is the
hand-written adjoint (mdgpu_adjoint), in the FP64 and the mixed form.  The
oxygen's LJ sigma and epsilon enter through lj_tables (the Lorentz-Berthelot
rule); the hydrogens have no LJ.  The reference for each checked entry is the
central difference of the FP64 loss, Richardson-extrapolated from the steps
h and h/2 (which removes the h^2 term; the loss is a polynomial of degree at
most 4 in a charge, an equilibrium or a force constant).  Bar: 1e-6 relative,
for both forms.  The steps are relative, h = 1e-4 |theta|; the same
extrapolation at h = 1e-6 |theta| is printed beside it: there the rounding of
the FP64 energies and of the fixed-point PME charge grid, divided by h, is of
the order of the bar itself.
"""
import numpy as np
import pytest
import torch

from md_amd import autograd as AG
from md_amd import native as N

pytestmark = pytest.mark.gpu

N_FRAMES = 4
KW = dict(lj_cutoff=9.0, lj_switch=8.0, ewald_tol=1e-5, pme_grid=(48, 48, 48))


def _sigma_eps(a):
    t = int(a["n_types"])
    idx = np.asarray(a["nb_index"]).reshape(t, t) - 1
    sig, eps = np.zeros(t), np.zeros(t)
    for i in range(t):
        k = idx[i, i]
        if k >= 0 and a["lj_a"][k] > 0 and a["lj_b"][k] > 0:
            sig[i] = (a["lj_a"][k] / a["lj_b"][k]) ** (1.0 / 6.0)
            eps[i] = a["lj_b"][k] ** 2 / (4.0 * a["lj_a"][k])
    return sig, eps


@pytest.fixture(scope="module")
def cfg5(water_6k):
    p, r = water_6k
    top = N.Topology(p, r, N.SEM_AMBER)
    frames = torch.tensor(np.stack([top.coords + np.random.default_rng(5000 + i).normal(0, 0.1, top.coords.shape)
                                    for i in range(N_FRAMES)]))
    model64 = AG.Model(top, N.make_params(N.SEM_AMBER, N.PREC_FP64, **KW))
    base = model64.initial()
    rng = np.random.default_rng(5999)
    tvals = {k: v * torch.tensor(np.exp(rng.normal(0, 0.05, v.shape))) for k, v in base.items()}
    with torch.no_grad():
        target = AG.energy(model64, tvals, frames, top.box)
    sig, eps = _sigma_eps(top.arrays)
    yield top, frames, target, model64, base, sig, eps
    model64.close()


def _values(top, base, sig, eps):
    v = {k: t.clone() for k, t in base.items() if k not in ("lj_a", "lj_b")}
    v["sigma"] = torch.tensor(sig, dtype=torch.float64)
    v["epsilon"] = torch.tensor(eps, dtype=torch.float64)
    return v


def _loss(model, top, frames, target, v):
    a = top.arrays
    vals = {k: t for k, t in v.items() if k not in ("sigma", "epsilon")}
    vals["lj_a"], vals["lj_b"] = AG.lj_tables(v["sigma"], v["epsilon"], a["nb_index"], a["lj_a"].size)
    return ((AG.energy(model, vals, frames, top.box) - target) ** 2).sum()


def _fd_checks(top, frames, target, model64, base, sig, eps):
    """(name, index, Richardson FD at h = 1e-4 |theta|, at h = 1e-6 |theta|)."""
    o = int(top.arrays["type"][0])

    def loss_at(name, idx, val):
        v = _values(top, base, sig, eps)
        v[name][idx] = val
        with torch.no_grad():
            return float(_loss(model64, top, frames, target, v))

    out = []
    for name, idx in [("charge", 0), ("charge", 1), ("charge", 3001), ("sigma", o), ("epsilon", o), ("bond_k", 0),
                      ("bond_r0", 0), ("angle_k", 0), ("angle_theta0", 0)]:
        base_v = float(_values(top, base, sig, eps)[name][idx])

        def rich(h):
            def cd(s):
                return (loss_at(name, idx, base_v + s) - loss_at(name, idx, base_v - s)) / (2 * s)
            return (4.0 * cd(h / 2) - cd(h)) / 3.0

        out.append((name, idx, rich(1e-4 * abs(base_v)), rich(1e-6 * abs(base_v))))
    return out


def _param_values(top, v):
    """The model's parameter arrays, LJ tables from sigma and epsilon."""
    a = top.arrays
    vals = {k: t.detach() for k, t in v.items() if k not in ("sigma", "epsilon")}
    A, B = AG.lj_tables(v["sigma"].detach(), v["epsilon"].detach(), a["nb_index"], a["lj_a"].size)
    vals["lj_a"], vals["lj_b"] = A, B
    return vals


def _grad_with_weights(model, top, frames, v, weights):
    """dL/dtheta = sum_f w_f dE_f/dtheta from Model.adjoint (the hand-written
    adjoint), chained through lj_tables for sigma and epsilon."""
    a = top.arrays
    model.energies(_param_values(top, v), frames.numpy()[:1], top.box)   # uploads the parameters
    g = model.adjoint(frames.numpy(), top.box, weights)
    s = v["sigma"].clone().requires_grad_(True)
    e = v["epsilon"].clone().requires_grad_(True)
    A, B = AG.lj_tables(s, e, a["nb_index"], a["lj_a"].size)
    torch.autograd.backward([A, B], [torch.tensor(g["lj_a"]), torch.tensor(g["lj_b"])])
    out = {k: g[k] for k in g}
    out["sigma"], out["epsilon"] = s.grad.numpy(), e.grad.numpy()
    return out


def test_cfg5_loss_gradient_matches_fd(cfg5):
    """FP64: the autograd.Function's loss gradient.  Mixed: the FP32 adjoint
    (Model.adjoint) under the same loss weights 2 (E_f - E_target,f), taken
    from the FP64 energies, so that the check is on the adjoint; the fully
    mixed loss gradient, whose weights carry the FP32 forward energies'
    error, is printed beside it."""
    top, frames, target, model64, base, sig, eps = cfg5
    h_ = int(top.arrays["type"][1])
    print(f"case water_6k, {N_FRAMES} frames, loss gradient vs Richardson FD", flush=True)
    v = {k: t.clone().requires_grad_(True) for k, t in _values(top, base, sig, eps).items()}
    loss = _loss(model64, top, frames, target, v)
    loss.backward()
    g64 = {k: t.grad.detach().numpy().copy() for k, t in v.items()}
    assert np.all(np.isfinite(g64["epsilon"])) and g64["epsilon"][h_] == 0.0
    vals = _values(top, base, sig, eps)
    e64 = model64.energies(_param_values(top, vals), frames.numpy(), top.box)
    w = 2.0 * (e64 - target.numpy())
    model32 = AG.Model(top, N.make_params(N.SEM_AMBER, N.PREC_MIXED, **KW))
    try:
        g32 = _grad_with_weights(model32, top, frames, vals, w)
        v = {k: t.clone().requires_grad_(True) for k, t in _values(top, base, sig, eps).items()}
        loss32 = _loss(model32, top, frames, target, v)
        loss32.backward()
        g32l = {k: t.grad.detach().numpy().copy() for k, t in v.items()}
        # per-term energies of frame 0 in both forms
        ev32, ev64 = model32.ev, model64.ev
        e_32 = ev32.evaluate(frames[0].numpy(), top.box)[0]
        e_64 = ev64.evaluate(frames[0].numpy(), top.box)[0]
        print("  frame 0 energies, mixed - fp64 (kcal/mol): " +
              ", ".join(f"{k} {e_32[k] - e_64[k]:+.2e}" for k in e_64 if e_64[k] != 0.0))
    finally:
        model32.close()
    print(f"  loss fp64 {float(loss):.10e} mixed forward {float(loss32):.10e}")
    bad = []
    for name, idx, fd, fd6 in _fd_checks(top, frames, target, model64, base, sig, eps):
        for form, g in (("fp64", g64), ("mixed", g32)):
            got = float(g[name][idx])
            rel = abs(got - fd) / abs(fd)
            extra = ""
            if form == "mixed":
                extra = f"  (fully mixed loss gradient rel {abs(float(g32l[name][idx]) - fd) / abs(fd):.2e})"
            print(f"  dL/d{name}[{idx}] {form:5s} {got: .12e} fd {fd: .12e} rel {rel:.2e}  "
                  f"fd(h=1e-6) rel {abs(got - fd6) / abs(fd6):.2e}{extra}")
            if not rel <= 1e-6:
                bad.append((name, idx, form, rel))
    assert not bad, bad
