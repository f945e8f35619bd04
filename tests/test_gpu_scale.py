"""Scale-up parity: ~100k atoms (100 A box, 16,000-atom polymer in ~28,000
waters), AMBER semantics, rc 10 A (switch 9 A), Ewald tol 1e-5, PME 108^3,
order 4; the pair list is rebuilt at every evaluation from positions
perturbed by 0.05 A Gaussian noise (seeds 3300 + i).

Bars: FP64 form vs the oracle per term 1e-9 and force rel-RMS 1e-8; mixed vs
FP64 force rel-RMS 1e-5; mixed bit-identical on repeat.
"""
import numpy as np
import pytest

import oracle
from md_amd import native as N

pytestmark = pytest.mark.gpu

TERMS = ["bond", "angle", "dihedral", "urey_bradley", "improper", "cmap", "vdw", "ewald_direct", "ewald_self",
         "ewald_recip", "ewald_excl", "vdw14", "ee14", "total"]
KW = dict(lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-5, pme_grid=(108, 108, 108))


def _frame(top, i):
    return top.coords + np.random.default_rng(3300 + i).normal(0.0, 0.05, top.coords.shape)


def _rms(f, g):
    return float(np.sqrt(np.sum((f - g) ** 2) / np.sum(g ** 2)))


def test_scale_100k_fp64_matches_oracle(scale_100k):
    p, r = scale_100k
    top = N.Topology(p, r, N.SEM_AMBER)
    print(f"case scale_100k fp64 vs oracle, {top.n_atoms} atoms", flush=True)
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, **KW)
    x = _frame(top, 0)
    ev = N.Evaluator(0)
    ev.upload(top.desc)
    ev.set_params(prm)
    e, f, v = ev.evaluate(x, top.box)
    ev.close()
    eo, fo, vo = oracle.evaluate(top.desc, prm, x, top.box)
    bad = []
    for k in TERMS:
        err = abs(e[k] - eo[k])
        print(f"  {k:14s} gpu {e[k]: .12e} oracle {eo[k]: .12e} rel {err / max(abs(eo[k]), 1e-300):.2e}")
        if not err <= 1e-9 * abs(eo[k]) + 1e-12:
            bad.append(k)
    r = _rms(f, fo)
    print(f"  force rel-RMS {r:.2e}; virial rel {np.abs(v - vo).max() / np.abs(vo).max():.2e}")
    assert not bad, bad
    assert r <= 1e-8


def test_scale_100k_mixed_vs_fp64_rebuilt_lists(scale_100k):
    p, r = scale_100k
    top = N.Topology(p, r, N.SEM_AMBER)
    print(f"case scale_100k mixed vs fp64, {top.n_atoms} atoms", flush=True)
    ev64, ev32 = N.Evaluator(0), N.Evaluator(0)
    for ev, prec in ((ev64, N.PREC_FP64), (ev32, N.PREC_MIXED)):
        ev.upload(top.desc)
        ev.set_params(N.make_params(N.SEM_AMBER, prec, **KW))
    for i in (1, 2):
        x = _frame(top, i)
        e64, f64, _ = ev64.evaluate(x, top.box)
        e32, f32, _ = ev32.evaluate(x, top.box)
        e32b, f32b, _ = ev32.evaluate(x, top.box)
        r = _rms(f32, f64)
        dt = abs(e32["total"] - e64["total"]) / sum(abs(e64[k]) for k in TERMS if k != "total")
        print(f"  frame {i}: force rel-RMS {r:.2e}, total energy rel {dt:.2e}")
        assert r <= 1e-5
        assert dt <= 1e-5
        assert np.array_equal(f32, f32b) and all(e32[k] == e32b[k] for k in e32)
    ev64.close()
    ev32.close()
