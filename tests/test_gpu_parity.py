"""GPU parity: the HIP kernels (through the mdgpu C ABI) against the CPU oracle.

Bars (from the task's precision contract):
  * FP64 form vs oracle: per-term |dE| <= 1e-9 |E| (absolute floor 1e-12 for
    terms that vanish), force relative RMS <= 1e-8;
  * mixed form vs FP64 form (AMBER semantics): force relative RMS <= 1e-5,
    per-term energy errors reported and bounded;
  * mixed form: bit-identical forces and energies on repeated evaluation.
"""
import numpy as np
import pytest

import oracle
from md_amd import native as N

pytestmark = pytest.mark.gpu

TERMS_REF = ["bond", "angle", "dihedral", "urey_bradley", "improper", "cmap", "vdw", "ee_all", "ee_pairlist",
             "ee_rf", "ewald_direct", "ewald_self", "total"]
TERMS_AMBER = ["bond", "angle", "dihedral", "urey_bradley", "improper", "cmap", "vdw", "ewald_direct",
               "ewald_self", "ewald_recip", "ewald_excl", "vdw14", "ee14", "total"]


def rel_rms(f, g):
    return float(np.sqrt(np.sum((f - g) ** 2) / np.sum(g ** 2)))


def run_gpu(top, params):
    ev = N.Evaluator(0)
    try:
        ev.upload(top.desc)
        ev.set_params(params)
        return ev.evaluate(top.coords, top.box)
    finally:
        ev.close()


def check_terms(e, eo, terms, rtol, floor=1e-12):
    bad = []
    for k in terms:
        err = abs(e[k] - eo[k])
        lim = rtol * abs(eo[k]) + floor
        print(f"  {k:14s} gpu {e[k]: .15e} oracle {eo[k]: .15e} rel {err / max(abs(eo[k]), 1e-300):.2e}")
        if not err <= lim:
            bad.append((k, e[k], eo[k]))
    assert not bad, bad


CASES = [
    # name, system fixture, semantics, lj_cutoff, lj_switch, ewald_tol, pme grid
    ("polymer2k_ref_switch", "polymer_2k", N.SEM_REF, 10.0, 9.0, 1e-6, (0, 0, 0)),
    ("polymer2k_ref_shift", "polymer_2k", N.SEM_REF, 10.0, 0.0, 1e-6, (0, 0, 0)),
    ("water6k_ref", "water_6k", N.SEM_REF, 9.0, 8.0, 1e-5, (0, 0, 0)),
    ("polymer2k_amber_pme", "polymer_2k", N.SEM_AMBER, 10.0, 9.0, 1e-5, (32, 32, 32)),
    ("water6k_amber_pme", "water_6k", N.SEM_AMBER, 9.0, 8.0, 1e-5, (48, 48, 48)),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fp64_matches_oracle(case, request):
    name, fix, sem, rc, sw, tol, grid = case
    prm, rst = request.getfixturevalue(fix)
    print(f"case {name}", flush=True)
    top = N.Topology(prm, rst, sem)
    p = N.make_params(sem, N.PREC_FP64, lj_cutoff=rc, lj_switch=sw, ewald_tol=tol, pme_grid=grid)
    e, f, v = run_gpu(top, p)
    eo, fo, vo = oracle.evaluate(top.desc, p, top.coords, top.box)
    check_terms(e, eo, TERMS_AMBER if sem == N.SEM_AMBER else TERMS_REF, 1e-9)
    r = rel_rms(f, fo)
    print(f"  force rel-RMS {r:.3e}")
    assert r <= 1e-8
    if sem == N.SEM_AMBER:
        dv = np.abs(v - vo).max() / np.abs(vo).max()
        print(f"  recip virial rel {dv:.3e}")
        assert dv <= 1e-9


@pytest.mark.parametrize("case", [c for c in CASES if c[2] == N.SEM_AMBER], ids=lambda c: c[0])
def test_mixed_matches_fp64(case, request):
    name, fix, sem, rc, sw, tol, grid = case
    prm, rst = request.getfixturevalue(fix)
    print(f"case {name}", flush=True)
    top = N.Topology(prm, rst, sem)
    p64 = N.make_params(sem, N.PREC_FP64, lj_cutoff=rc, lj_switch=sw, ewald_tol=tol, pme_grid=grid)
    p32 = N.make_params(sem, N.PREC_MIXED, lj_cutoff=rc, lj_switch=sw, ewald_tol=tol, pme_grid=grid)
    e64, f64, _ = run_gpu(top, p64)
    e32, f32, _ = run_gpu(top, p32)
    for k in TERMS_AMBER:
        print(f"  {k:14s} mixed {e32[k]: .10e} fp64 {e64[k]: .10e} rel {abs(e32[k] - e64[k]) / max(abs(e64[k]), 1e-300):.2e}")
    r = rel_rms(f32, f64)
    print(f"  force rel-RMS mixed vs fp64 {r:.3e}")
    assert r <= 1e-5
    assert abs(e32["total"] - e64["total"]) <= 1e-5 * sum(abs(e64[k]) for k in TERMS_AMBER if k != "total")


def test_mixed_is_bit_reproducible(polymer_2k):
    prm, rst = polymer_2k
    top = N.Topology(prm, rst, N.SEM_AMBER)
    p = N.make_params(N.SEM_AMBER, N.PREC_MIXED, lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-5,
                      pme_grid=(32, 32, 32))
    ev = N.Evaluator(0)
    ev.upload(top.desc)
    ev.set_params(p)
    e1, f1, _ = ev.evaluate(top.coords, top.box)
    e2, f2, _ = ev.evaluate(top.coords, top.box)
    ev.close()
    assert np.array_equal(f1, f2)
    assert all(e1[k] == e2[k] for k in e1)


def test_ref_mixed_close_to_fp64(polymer_2k):
    """REF semantics in the mixed form (huge bonded-pair LJ) stays within 1e-5."""
    prm, rst = polymer_2k
    top = N.Topology(prm, rst, N.SEM_REF)
    p64 = N.make_params(N.SEM_REF, N.PREC_FP64, lj_cutoff=10.0, lj_switch=9.0)
    p32 = N.make_params(N.SEM_REF, N.PREC_MIXED, lj_cutoff=10.0, lj_switch=9.0)
    _, f64, _ = run_gpu(top, p64)
    _, f32, _ = run_gpu(top, p32)
    r = rel_rms(f32, f64)
    print(f"  REF force rel-RMS mixed vs fp64 {r:.3e}")
    assert r <= 1e-5
