"""The 29k-atom workload exactly as the benchmark runs it.

66 A cube, one 4,800-atom random-walk heteropolymer in ~8,000 three-site
waters (28,800 atoms, 30 LJ types, angles + Urey-Bradley, cosine dihedrals,
impropers, 300 CMAP terms), AMBER semantics, rc 12 A with the LJ switch from
11 A, Ewald tolerance 1e-5, smooth PME on a 72^3 grid, order 4.

Bars:
  * FP64 form vs the oracle: per term |dE| <= 1e-9 |E| (+1e-12), force
    relative RMS <= 1e-8, reciprocal virial <= 1e-9 of its largest component;
  * mixed form vs FP64 form: force relative RMS <= 1e-5 and every energy term
    within the bar of MIXED_BARS (relative to the FP64 value);
  * mixed form bit-identical over repeated evaluations, direct launches and
    graph replays alike.
"""
import numpy as np
import pytest

import oracle
from md_amd import native as N

pytestmark = pytest.mark.gpu

KW = dict(lj_cutoff=12.0, lj_switch=11.0, ewald_tol=1e-5, pme_grid=(72, 72, 72))
TERMS = ["bond", "angle", "dihedral", "urey_bradley", "improper", "cmap", "vdw", "ewald_direct", "ewald_self",
         "ewald_recip", "ewald_excl", "vdw14", "ee14", "total"]

# Relative bars of the mixed form against the FP64 form, per term.  FP32 pair,
# PME and bonded arithmetic; the self, exclusion and 1-4 terms are FP64 in
# both forms.  ewald_direct is a cancelled sum (about 155 kcal/mol from
# 10.5 M pair terms of either sign), so its bar also bounds the FP32 per-pair
# error of erfc and its accumulation.
MIXED_BARS = {
    "bond": 1e-6, "angle": 1e-6, "dihedral": 1e-6, "urey_bradley": 1e-6, "improper": 1e-6, "cmap": 1e-5,
    "vdw": 1e-6, "ewald_direct": 1e-5, "ewald_self": 1e-14, "ewald_recip": 1e-6, "ewald_excl": 1e-12,
    "vdw14": 1e-12, "ee14": 1e-12, "total": 1e-6,
}


def _rms(f, g):
    return float(np.sqrt(np.sum((f - g) ** 2) / np.sum(g ** 2)))


def _eval(top, prec, reps=1):
    ev = N.Evaluator(0)
    try:
        ev.upload(top.desc)
        ev.set_params(N.make_params(N.SEM_AMBER, prec, **KW))
        out = [ev.evaluate(top.coords, top.box) for _ in range(reps)]
    finally:
        ev.close()
    return out


@pytest.fixture(scope="module")
def cfg2(polymer_29k):
    prm, rst = polymer_29k
    top = N.Topology(prm, rst, N.SEM_AMBER)
    yield top
    top.close()


@pytest.fixture(scope="module")
def fp64_result(cfg2):
    return _eval(cfg2, N.PREC_FP64)[0]


def test_cfg2_fp64_matches_oracle(cfg2, fp64_result):
    print(f"case polymer_29k fp64 vs oracle, {cfg2.n_atoms} atoms", flush=True)
    e, f, v = fp64_result
    p = N.make_params(N.SEM_AMBER, N.PREC_FP64, **KW)
    eo, fo, vo = oracle.evaluate(cfg2.desc, p, cfg2.coords, cfg2.box, skip_all_pairs=True)
    bad = []
    for k in TERMS:
        err = abs(e[k] - eo[k])
        print(f"  {k:14s} gpu {e[k]: .15e} oracle {eo[k]: .15e} rel {err / max(abs(eo[k]), 1e-300):.2e}")
        if not err <= 1e-9 * abs(eo[k]) + 1e-12:
            bad.append(k)
    r = _rms(f, fo)
    dv = float(np.abs(v - vo).max() / np.abs(vo).max())
    print(f"  force rel-RMS {r:.2e}; recip virial rel {dv:.2e}")
    assert not bad, bad
    assert r <= 1e-8
    assert dv <= 1e-9


def test_cfg2_mixed_vs_fp64(cfg2, fp64_result):
    print(f"case polymer_29k mixed vs fp64, {cfg2.n_atoms} atoms", flush=True)
    e64, f64, _ = fp64_result
    e32, f32, _ = _eval(cfg2, N.PREC_MIXED)[0]
    bad = []
    for k in TERMS:
        rel = abs(e32[k] - e64[k]) / max(abs(e64[k]), 1e-300)
        print(f"  {k:14s} mixed {e32[k]: .12e} fp64 {e64[k]: .12e} rel {rel:.2e} (bar {MIXED_BARS[k]:.0e})")
        if not rel <= MIXED_BARS[k]:
            bad.append((k, rel))
    r = _rms(f32, f64)
    print(f"  force rel-RMS mixed vs fp64 {r:.3e}")
    assert r <= 1e-5
    assert not bad, bad


def test_cfg2_mixed_bit_identical_on_repeat(cfg2):
    """Host-buffer evaluations (direct launches, then graph replays from the
    third call on) and the device-buffer steps of a second context (graph
    replays as in the benchmark) all give the same bits."""
    import torch
    print("case polymer_29k mixed repeat", flush=True)
    runs = _eval(cfg2, N.PREC_MIXED, reps=4)
    e0, f0, _ = runs[0]
    for e, f, _ in runs[1:]:
        assert np.array_equal(f, f0)
        assert all(e[k] == e0[k] for k in e0)
    ev = N.Evaluator(0)
    try:
        ev.upload(cfg2.desc)
        ev.set_params(N.make_params(N.SEM_AMBER, N.PREC_MIXED, **KW))
        x = torch.tensor(cfg2.coords.reshape(-1), dtype=torch.float64, device="cuda:0")
        f = torch.zeros_like(x)
        for i in range(4):
            e, _ = ev.evaluate_device(x.data_ptr(), cfg2.box, f.data_ptr())
            torch.cuda.synchronize()
            assert np.array_equal(f.cpu().numpy().reshape(f0.shape), f0), f"device step {i}"
            assert all(e[k] == e0[k] for k in e0), f"device step {i}"
        print(f"  graph state after device steps: {ev.graph_state()}")
        assert ev.graph_state() == 1
    finally:
        ev.close()


def test_long_candidate_streams_match_oracle(cfg2):
    """rc 18 A (switch 17 A) on the same system: a quarter of the
    i-clusters then have more candidate j-clusters (up to 874) than the
    search stages in LDS (768), so for them the search maps stream
    positions through its column ranges instead; the
    fast minimum-image search must still run (search mode 0) and give the
    oracle's result at the FP64 bars, and the mixed form its usual bar."""
    kw = dict(lj_cutoff=18.0, lj_switch=17.0, ewald_tol=1e-5, pme_grid=(72, 72, 72))
    print(f"case polymer_29k rc 18 A, {cfg2.n_atoms} atoms", flush=True)
    out = {}
    for prec in (N.PREC_FP64, N.PREC_MIXED):
        ev = N.Evaluator(0)
        try:
            ev.upload(cfg2.desc)
            ev.set_params(N.make_params(N.SEM_AMBER, prec, **kw))
            out[prec] = ev.evaluate(cfg2.coords, cfg2.box)
            assert ev.search_mode() == 0
        finally:
            ev.close()
    e, f, _ = out[N.PREC_FP64]
    p = N.make_params(N.SEM_AMBER, N.PREC_FP64, **kw)
    eo, fo, _ = oracle.evaluate(cfg2.desc, p, cfg2.coords, cfg2.box, skip_all_pairs=True)
    bad = [k for k in TERMS if not abs(e[k] - eo[k]) <= 1e-9 * abs(eo[k]) + 1e-12]
    r = _rms(f, fo)
    r32 = _rms(out[N.PREC_MIXED][1], f)
    print(f"  fp64 vs oracle force rel-RMS {r:.2e}, terms off {bad}; mixed vs fp64 {r32:.2e}")
    assert not bad, bad
    assert r <= 1e-8
    assert r32 <= 1e-5
