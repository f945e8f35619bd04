"""GPU adjoint (mdgpu_adjoint, through the C ABI) against the oracle's
parameter gradients, finite differences of the GPU energies, a dense torch
Ewald reference, and through the torch.autograd.Function.

Bars: FP64 adjoint vs oracle 1e-9 (relative, per gradient block); central
finite differences of the FP64 GPU energy 1e-6; mixed vs FP64 per block as
stated in each test; bit-identical repeated mixed adjoints.
"""
import numpy as np
import pytest
import torch

import oracle
from md_amd import autograd as AG
from md_amd import native as N
from md_amd import systems

pytestmark = pytest.mark.gpu

AMBER_TERMS = ["bond", "angle", "dihedral", "urey_bradley", "improper", "cmap", "vdw", "ewald_direct",
               "ewald_self", "ewald_recip", "ewald_excl", "vdw14", "ee14"]
ELEC_TERMS = ["ewald_direct", "ewald_self", "ewald_recip", "ewald_excl", "ee14"]


def _weights(seed):
    rng = np.random.default_rng(seed)
    return {k: float(rng.uniform(0.5, 2.0)) * (1 if rng.random() < 0.5 else -1) for k in AMBER_TERMS}


def _gpu_adjoint(top, prm, weights, coords=None):
    ev = N.Evaluator(0)
    try:
        ev.upload(top.desc)
        ev.set_params(prm)
        return ev.adjoint(top.coords if coords is None else coords, top.box, weights, top.desc)
    finally:
        ev.close()


def _block_err(a, b):
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / den if den > 0 else np.linalg.norm(a)


@pytest.mark.parametrize("switch", [9.0, 0.0], ids=["switched", "shifted"])
def test_fp64_adjoint_matches_oracle(polymer_2k, switch):
    p, r = polymer_2k
    top = N.Topology(p, r, N.SEM_AMBER)
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, lj_cutoff=10.0, lj_switch=switch, ewald_tol=1e-5,
                        pme_grid=(32, 32, 32))
    w = _weights(11)
    print(f"case polymer_2k fp64 switch={switch}", flush=True)
    g = _gpu_adjoint(top, prm, w)
    go = oracle.param_grads(top.desc, prm, top.coords, top.box, w)
    bad = []
    for k in go:
        err = _block_err(g[k], go[k])
        print(f"  {k:16s} |g| {np.linalg.norm(go[k]):.4e} rel {err:.2e}")
        if not err <= 1e-9:
            bad.append((k, err))
    assert not bad, bad


def test_mixed_adjoint_close_to_fp64_and_reproducible(polymer_2k):
    p, r = polymer_2k
    top = N.Topology(p, r, N.SEM_AMBER)
    kw = dict(lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-5, pme_grid=(32, 32, 32))
    w = _weights(12)
    print("case polymer_2k mixed vs fp64", flush=True)
    g64 = _gpu_adjoint(top, N.make_params(N.SEM_AMBER, N.PREC_FP64, **kw), w)
    p32 = N.make_params(N.SEM_AMBER, N.PREC_MIXED, **kw)
    ev = N.Evaluator(0)
    ev.upload(top.desc)
    ev.set_params(p32)
    g32 = ev.adjoint(top.coords, top.box, w, top.desc)
    g32b = ev.adjoint(top.coords, top.box, w, top.desc)
    ev.close()
    # charge block: FP32 direct-space pair terms, FP64 reciprocal potential
    # (bar 1e-6); LJ blocks: FP32 pair terms from cluster-relative FP32
    # coordinates (2^-21 |x| per coordinate, so 1/r^12 of a close pair carries
    # ~1e-6 and a type pair with few close contacts keeps it; bar 2e-6); the
    # 1-4 LJ entries come from the FP64 corrections and the bonded and CMAP
    # blocks from the FP64 bonded adjoint in both forms, so they agree exactly
    bad = []
    for k in g64:
        err = _block_err(g32[k], g64[k])
        lim = 1e-6 if k == "charge" else (2e-6 if k in ("lj_a", "lj_b") else 1e-12)
        print(f"  {k:16s} rel {err:.2e} (bar {lim:.0e})")
        if not err <= lim:
            bad.append((k, err))
        assert np.array_equal(g32[k], g32b[k]), k
    assert not bad, bad


def test_adjoint_matches_fd_of_gpu_energies(water_6k):
    """Central differences of the FP64 GPU total energy (the 6,450-atom water
    box) for charges, LJ entries and bonded constants."""
    p, r = water_6k
    top = N.Topology(p, r, N.SEM_AMBER)
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, lj_cutoff=9.0, lj_switch=8.0, ewald_tol=1e-5,
                        pme_grid=(48, 48, 48))
    g = _gpu_adjoint(top, prm, {"total": 1.0})
    ev = N.Evaluator(0)
    ev.set_params(prm)
    arrays = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in top.arrays.items()}

    def energy():
        keep = []
        ev.upload(N.desc_from_arrays(arrays, keep))
        ev.set_params(prm)
        return ev.evaluate(top.coords, top.box)[0]["total"]

    checks = [("charge", 0, 1e-2), ("charge", 1, 1e-2), ("charge", 3001, 1e-2), ("lj_a", 0, None),
              ("lj_b", 0, None), ("bond_k", 0, 1.0), ("bond_r0", 0, 1e-3), ("angle_k", 0, 1.0),
              ("angle_theta0", 0, 1e-3)]
    for name, idx, h in checks:
        if idx >= arrays[name].size:
            continue
        base = arrays[name][idx]
        h = h if h is not None else 0.05 * max(abs(base), 1.0)
        arrays[name][idx] = base + h
        ep = energy()
        arrays[name][idx] = base - h
        em = energy()
        arrays[name][idx] = base
        fd = (ep - em) / (2 * h)
        tol = 1e-6 * abs(g[name][idx]) + 1e-13 * abs(ep) / h
        print(f"  {name}[{idx}] adjoint {g[name][idx]: .12e} fd {fd: .12e}")
        assert abs(fd - g[name][idx]) <= tol, (name, idx)
    ev.close()


@pytest.mark.parametrize("prec", [N.PREC_FP64, N.PREC_MIXED], ids=["fp64", "mixed"])
def test_charge_gradient_matches_dense_torch_ewald(tmp_path, prec):
    import dense_ewald
    top = systems.water_box(seed=5600, box=18.2, n_water=200, n_lattice=6, jitter=0.3)
    p, r = systems.write_system(top, str(tmp_path / "w600"))
    t = N.Topology(p, r, N.SEM_AMBER)
    rc = 8.0
    prm = N.make_params(N.SEM_AMBER, prec, lj_cutoff=rc, lj_switch=7.0, ewald_tol=1e-5, pme_grid=(96, 96, 96))
    g = _gpu_adjoint(t, prm, {k: 1.0 for k in ELEC_TERMS})["charge"]
    _, g_ref = dense_ewald.charge_gradient(t.arrays, t.coords, t.box, prm.ewald_alpha, rc, 14,
                                           oracle.coulomb_constant())
    rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    print(f"  dE/dq vs dense torch Ewald: rel {rel:.2e}")
    # 96^3 PME discretisation ~1e-6 in both forms; the mixed form's FP32 erfc
    # is relative-accurate to ~1e-6 (erfc_f32.h)
    assert rel <= 2e-6


def test_autograd_function_loss_gradient(polymer_2k):
    """loss = sum_f (E_f - E_target,f)^2 over perturbed frames; gradients from
    the autograd.Function (adjoint) against central differences of the loss."""
    p, r = polymer_2k
    top = N.Topology(p, r, N.SEM_AMBER)
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-5,
                        pme_grid=(32, 32, 32))
    model = AG.Model(top, prm)
    rng = np.random.default_rng(5000)
    frames = torch.tensor(np.stack([top.coords + rng.normal(0, 0.05, top.coords.shape) for _ in range(3)]))
    values = model.initial()
    with torch.no_grad():
        target = AG.energy(model, values, frames, top.box) + torch.tensor([3.0, -2.0, 1.5], dtype=torch.float64)
    params = {k: v.clone().requires_grad_(True) for k, v in values.items()}
    loss = ((AG.energy(model, params, frames, top.box) - target) ** 2).sum()
    loss.backward()

    def loss_at(name, idx, val):
        v = {k: t.detach().clone() for k, t in params.items()}
        v[name][idx] = val
        with torch.no_grad():
            return float(((AG.energy(model, v, frames, top.box) - target) ** 2).sum())

    # the loss is up to quartic in a charge or an equilibrium: small steps for those
    for name, idx, h in [("charge", 10, 1e-4), ("lj_a", 3, None), ("bond_r0", 2, 2e-5), ("dihedral_k", 1, 1e-3),
                         ("cmap_grid", None, 1e-2)]:
        if idx is None:  # the grid value with the largest gradient
            idx = int(torch.argmax(params[name].grad.abs()))
        base = float(params[name][idx])
        h = h if h is not None else 1e-2 * max(abs(base), 1.0)

        def cd(s):
            return (loss_at(name, idx, base + s) - loss_at(name, idx, base - s)) / (2 * s)

        fd = (4.0 * cd(h / 2) - cd(h)) / 3.0  # Richardson: the h^2 term cancels
        got = float(params[name].grad[idx])
        print(f"  dL/d{name}[{idx}] autograd {got: .10e} fd {fd: .10e}")
        assert abs(got - fd) <= 1e-6 * abs(got) + 1e-9 * abs(float(loss)) / h, (name, idx)
    model.close()
