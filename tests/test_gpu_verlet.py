"""Verlet list with the GPU-resident rebuild criterion (mdgpu_set_verlet).

With a skin the context keeps its cluster-pair list (radius cutoff + skin)
until an atom has moved more than skin / 2 since the build, decided on the
GPU (System::ljpairs_needs_rebuild, NonBondedUtils.cpp:87-100).  Bars:
  * the GPU decision (mdgpu_pairs_need_rebuild) equals the oracle's
    criterion on every displacement set, including exact-boundary moves,
    moves across periodic faces and whole-box jumps;
  * FP64 evaluations from a kept list match the oracle at the displaced
    coordinates as fresh ones do: per-term 1e-9 relative, force rel-RMS 1e-8;
  * mixed evaluations from a kept list stay within the mixed bar (force
    rel-RMS 1e-5) of the FP64 result and are bit-reproducible;
  * along a trajectory, captured-graph replay gives what direct launches
    give (mixed bit for bit, FP64 to 1e-12), the number of list builds equals
    the oracle criterion's count, and the replay continues across rebuilds;
  * the kept list is pruned to cutoff + 0.2 A (again once an atom has moved
    0.1 A since the last prune): about the tiles of a list built at the
    cutoff, the same pair count and FP64 energies.
"""
import numpy as np
import pytest
import torch

import oracle
from md_amd import native as N

pytestmark = pytest.mark.gpu

TERMS_REF = ["bond", "angle", "dihedral", "urey_bradley", "improper", "cmap", "vdw", "ee_all", "ee_pairlist",
             "ee_rf", "ewald_direct", "ewald_self", "total"]
TERMS_AMBER = ["bond", "angle", "dihedral", "urey_bradley", "improper", "cmap", "vdw", "ewald_direct",
               "ewald_self", "ewald_recip", "ewald_excl", "vdw14", "ee14", "total"]


def rel_rms(f, g):
    return float(np.sqrt(np.sum((f - g) ** 2) / np.sum(g ** 2)))


def check_terms(e, eo, terms, rtol, floor=1e-12):
    bad = []
    for k in terms:
        err = abs(e[k] - eo[k])
        if not err <= rtol * abs(eo[k]) + floor:
            bad.append((k, e[k], eo[k]))
    assert not bad, bad


class Ctx:
    """One evaluator with device coordinate and force buffers."""

    def __init__(self, top, prm, skin, graph=True):
        self.ev = N.Evaluator(0)
        self.ev.upload(top.desc)
        self.ev.set_params(prm)
        self.ev.set_verlet(skin)
        self.ev.set_graph(graph)
        self.box = list(np.asarray(top.box, dtype=np.float64).reshape(-1)[:3])
        dev = torch.device("cuda", 0)
        self.xd = torch.zeros(3 * top.n_atoms, dtype=torch.float64, device=dev)
        self.fd = torch.zeros_like(self.xd)

    def put(self, x):
        self.xd.copy_(torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64).reshape(-1)))
        torch.cuda.synchronize()

    def eval(self, x):
        self.put(x)
        e, v = self.ev.evaluate_device(self.xd.data_ptr(), self.box, self.fd.data_ptr())
        torch.cuda.synchronize()
        return e, self.fd.cpu().numpy().reshape(-1, 3).copy(), v.copy()

    def needs(self, x):
        self.put(x)
        return self.ev.needs_rebuild(self.xd.data_ptr())

    def close(self):
        self.ev.close()


def displaced(x0, rng, maxd):
    """Every atom moved in a random direction by a length in [0, maxd)."""
    x = np.asarray(x0, dtype=np.float64).reshape(-1, 3)
    d = rng.normal(size=x.shape)
    d /= np.linalg.norm(d, axis=1)[:, None]
    d *= rng.uniform(0.0, maxd, (x.shape[0], 1))
    return x + d


def test_gpu_criterion_matches_oracle(water_6k):
    prm_path, rst_path = water_6k
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    print("case water_6k rebuild query, skin 2", flush=True)
    prm = N.make_params(N.SEM_AMBER, N.PREC_MIXED, lj_cutoff=9.0, lj_switch=8.0, ewald_tol=1e-5,
                        pme_grid=(48, 48, 48))
    skin = 2.0
    c = Ctx(top, prm, skin)
    try:
        x0 = np.asarray(top.coords, dtype=np.float64).reshape(-1, 3)
        assert c.needs(x0)  # nothing built yet
        c.eval(x0)
        assert c.ev.list_builds() == (1, 1)
        box = np.asarray(c.box)
        rng = np.random.default_rng(3)
        sets = [x0.copy(), displaced(x0, rng, 0.95), displaced(x0, rng, 1.05)]
        for ax in range(3):
            for dd in (1.0, np.nextafter(1.0, 2.0), 1.0 + 1e-9, -1.0 - 1e-9):
                x = x0.copy()
                x[17, ax] += dd
                sets.append(x)
        # across the faces (raw coordinates outside the box, and re-wrapped)
        x = displaced(x0, rng, 0.9)
        sets.append(x - box * np.floor(x / box))
        x = x0.copy()
        x[::7] += box * rng.integers(-2, 3, (x[::7].shape[0], 3))
        sets.append(x)
        for maxd in (0.99, 1.0, 1.01, 1.2):
            sets.append(displaced(x0, rng, maxd))
        x = displaced(x0, rng, 0.5)
        x[5, 1] = np.nan
        sets.append(x)
        got, want = [], []
        for x in sets:
            got.append(c.needs(x))
            want.append(oracle.pairs_need_rebuild(x, x0, box, skin))
        print(f"  decisions {got}")
        assert got == want
        assert any(got) and not all(got)
        # the queries changed nothing
        assert c.ev.list_builds() == (1, 1)
        assert not c.needs(x0)
    finally:
        c.close()


CASES = [
    # name, fixture, semantics, precision, cutoff, switch, tol, grid, skin
    ("polymer2k_ref_fp64", "polymer_2k", N.SEM_REF, N.PREC_FP64, 10.0, 9.0, 1e-6, (0, 0, 0), 2.0),
    ("water6k_amber_fp64", "water_6k", N.SEM_AMBER, N.PREC_FP64, 9.0, 8.0, 1e-5, (48, 48, 48), 2.0),
    ("polymer2k_amber_fp64_shift", "polymer_2k", N.SEM_AMBER, N.PREC_FP64, 10.0, 0.0, 1e-5, (32, 32, 32), 1.0),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_kept_list_matches_oracle(case, request):
    name, fix, sem, prec, rc, sw, tol, grid, skin = case
    prm_path, rst_path = request.getfixturevalue(fix)
    print(f"case {name} kept list", flush=True)
    top = N.Topology(prm_path, rst_path, sem)
    prm = N.make_params(sem, prec, lj_cutoff=rc, lj_switch=sw, ewald_tol=tol, pme_grid=grid)
    terms = TERMS_AMBER if sem == N.SEM_AMBER else TERMS_REF
    c = Ctx(top, prm, skin)
    try:
        x0 = np.asarray(top.coords, dtype=np.float64).reshape(-1, 3)
        box = np.asarray(c.box)
        rng = np.random.default_rng(11)
        c.eval(x0)
        # kept: every atom within skin / 2, a fifth of them re-wrapped into the box
        x1 = displaced(x0, rng, 0.49 * skin)
        w = rng.random(x1.shape[0]) < 0.2
        x1[w] -= box * np.floor(x1[w] / box)
        assert not oracle.pairs_need_rebuild(x1, x0, box, skin)
        e1, f1, v1 = c.eval(x1)
        assert c.ev.list_builds() == (1, 2)
        eo, fo, vo = oracle.evaluate(top.desc, prm, x1, top.box)
        check_terms(e1, eo, terms, 1e-9)
        r = rel_rms(f1, fo)
        print(f"  kept list: force rel-RMS {r:.3e}")
        assert r <= 1e-8
        if sem == N.SEM_AMBER:
            assert np.abs(v1 - vo).max() / np.abs(vo).max() <= 1e-9
        # one atom beyond skin / 2: rebuilt
        x2 = x1.copy()
        x2[3] = x0[3] + np.array([0.0, 0.51 * skin, 0.0])
        assert oracle.pairs_need_rebuild(x2, x0, box, skin)
        e2, f2, _ = c.eval(x2)
        assert c.ev.list_builds() == (2, 3)
        eo2, fo2, _ = oracle.evaluate(top.desc, prm, x2, top.box)
        check_terms(e2, eo2, terms, 1e-9)
        assert rel_rms(f2, fo2) <= 1e-8
    finally:
        c.close()


def test_kept_list_mixed(water_6k):
    prm_path, rst_path = water_6k
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    print("case water_6k mixed kept list", flush=True)
    kw = dict(lj_cutoff=9.0, lj_switch=8.0, ewald_tol=1e-5, pme_grid=(48, 48, 48))
    p32 = N.make_params(N.SEM_AMBER, N.PREC_MIXED, **kw)
    p64 = N.make_params(N.SEM_AMBER, N.PREC_FP64, **kw)
    x0 = np.asarray(top.coords, dtype=np.float64).reshape(-1, 3)
    x1 = displaced(x0, np.random.default_rng(5), 0.99)
    runs = []
    for rep in range(2):
        c = Ctx(top, p32, 2.0)
        try:
            c.eval(x0)
            runs.append(c.eval(x1))
            assert c.ev.list_builds() == (1, 2)
        finally:
            c.close()
    f = Ctx(top, p64, 0.0)
    try:
        e64, f64, _ = f.eval(x1)
    finally:
        f.close()
    (ea, fa, _), (eb, fb, _) = runs
    assert np.array_equal(fa, fb) and all(ea[k] == eb[k] for k in TERMS_AMBER)
    r = rel_rms(fa, f64)
    print(f"  mixed kept list vs FP64 fresh: force rel-RMS {r:.3e}")
    assert r <= 1e-5
    assert abs(ea["total"] - e64["total"]) <= 1e-5 * sum(abs(e64[k]) for k in TERMS_AMBER if k != "total")


@pytest.mark.parametrize("prec", [N.PREC_MIXED, N.PREC_FP64], ids=["mixed", "fp64"])
def test_trajectory_replay_and_build_count(prec, polymer_2k):
    prm_path, rst_path = polymer_2k
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    print(f"case polymer_2k trajectory prec={prec}", flush=True)
    prm = N.make_params(N.SEM_AMBER, prec, lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-5, pme_grid=(32, 32, 32))
    skin = 1.0
    rng = np.random.default_rng(21)
    x0 = np.asarray(top.coords, dtype=np.float64).reshape(-1, 3)
    v = rng.normal(0.0, 0.04, x0.shape)
    frames = [x0 + t * v for t in range(14)]
    box = np.asarray(top.box, dtype=np.float64).reshape(-1)[:3]
    # the oracle criterion's build sequence
    want_builds, ref = 0, None
    for x in frames:
        if ref is None or oracle.pairs_need_rebuild(x, ref, box, skin):
            want_builds += 1
            ref = x
    assert 3 <= want_builds <= 10, want_builds
    out = {}
    for key, graph, s in (("replay", True, skin), ("direct", False, skin), ("fresh", True, 0.0)):
        c = Ctx(top, prm, s, graph)
        try:
            res, states = [], []
            for x in frames:
                res.append(c.eval(x))
                states.append(c.ev.graph_state())
            out[key] = (res, states, c.ev.list_builds())
        finally:
            c.close()
        print(f"  {key}: graph states {out[key][1]} builds {out[key][2]}")
    assert out["replay"][2] == (want_builds, len(frames))
    assert out["direct"][2] == (want_builds, len(frames))
    assert out["replay"][1] == [0] + [1] * (len(frames) - 1)
    for k, ((eg, fg, _), (ed, fdd, _), (ef, ff, _)) in enumerate(zip(out["replay"][0], out["direct"][0],
                                                                     out["fresh"][0])):
        if prec == N.PREC_MIXED:
            assert np.array_equal(fg, fdd), k
            assert all(eg[t] == ed[t] for t in TERMS_AMBER), k
            assert rel_rms(fg, ff) <= 1e-5, k
        else:
            assert np.abs(fg - fdd).max() <= 1e-11 * np.abs(fdd).max(), k
            assert rel_rms(fg, ff) <= 1e-10, k
            check_terms(eg, ef, TERMS_AMBER, 1e-10)


def test_kept_list_is_pruned_to_the_cutoff(polymer_2k):
    """A kept list (radius cutoff + skin) is pruned to the tiles holding a pair
    within cutoff + 0.2 A: the pair kernel then runs about as many tiles as a
    list built at the cutoff alone (within 10 %), the same pairs, and the
    same FP64 energies (1e-12)."""
    print("case polymer_2k pruned Verlet list", flush=True)
    prm_path, rst_path = polymer_2k
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-5,
                        pme_grid=(32, 32, 32))
    out = {}
    for skin in (0.0, 1.5):
        ev = N.Evaluator(0)
        try:
            ev.upload(top.desc)
            ev.set_params(prm)
            ev.set_verlet(skin)
            ev.set_counting(True)
            e, _, _ = ev.evaluate(top.coords, top.box)
            out[skin] = (e, ev.counts())
        finally:
            ev.close()
    (e0, c0), (e1, c1) = out[0.0], out[1.5]
    print(f"  tiles without skin {c0['tiles']}, pruned from cutoff + 1.5 A {c1['tiles']}; pairs {c0['pairs']}")
    assert c1["pairs"] == c0["pairs"]
    assert c1["tiles"] <= 1.1 * c0["tiles"]
    for k in TERMS_AMBER:
        assert abs(e1[k] - e0[k]) <= 1e-12 * max(abs(e0[k]), 1.0), k


def _close_pair(top, rc, lo, hi):
    """Two atoms of the system, neither excluded nor 1-4 with the other, at a
    minimum-image distance in (rc + lo, rc + hi); and their unit separation."""
    x = np.asarray(top.coords, dtype=np.float64).reshape(-1, 3)
    box = np.asarray(top.box, dtype=np.float64)
    a = top.arrays
    st, ex = a["excl_start"], a["excl_list"]
    for i in range(x.shape[0]):
        d = x - x[i]
        d -= box * np.round(d / box)
        r = np.sqrt(np.sum(d * d, axis=1))
        near = set(ex[st[i]:st[i + 1]].tolist())
        for j in np.nonzero((r > rc + lo) & (r < rc + hi))[0]:
            if j > i and int(j) not in near:
                return i, int(j), d[j] / r[j], r[j]
    raise AssertionError("no pair in the shell")


def test_prune_interval(polymer_2k):
    """Reuse of a pruned list without a fresh prune.  Build and prune at x0
    (skin 1 A: tiles with a pair within cutoff + 0.2 A).  A pair at cutoff
    + 0.12..0.18 A is moved 0.09 A closer at each end (now inside the
    cutoff); no atom has moved 0.1 A since the prune, so the same pruned list
    is used (the prune count stays at 1), and FP64 still matches the oracle.
    Then one atom moves 0.12 A: the list is pruned again (count 2, still one
    build) and FP64 matches the oracle."""
    print("case polymer_2k prune interval", flush=True)
    prm_path, rst_path = polymer_2k
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    rc = 10.0
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, lj_cutoff=rc, lj_switch=9.0, ewald_tol=1e-5,
                        pme_grid=(32, 32, 32))
    i, j, u, r0 = _close_pair(top, rc, 0.12, 0.18)
    x0 = np.asarray(top.coords, dtype=np.float64).reshape(-1, 3)
    x1 = x0.copy()
    x1[i] += 0.09 * u
    x1[j] -= 0.09 * u
    x2 = x1.copy()
    k = (i + x0.shape[0] // 2) % x0.shape[0]
    x2[k] += np.array([0.12, 0.0, 0.0])
    print(f"  pair ({i}, {j}) at {r0:.4f} A -> {r0 - 0.18:.4f} A; atom {k} moves 0.12 A")
    ev = N.Evaluator(0)
    try:
        ev.upload(top.desc)
        ev.set_params(prm)
        ev.set_verlet(1.0)
        ev.evaluate(x0, top.box)
        assert ev.list_builds()[0] == 1 and ev.list_prunes() == 1
        e1, f1, _ = ev.evaluate(x1, top.box)
        p1 = ev.list_prunes()
        e2, f2, _ = ev.evaluate(x2, top.box)
        p2, b2 = ev.list_prunes(), ev.list_builds()[0]
    finally:
        ev.close()
    print(f"  prunes after the 0.09 A moves {p1}, after the 0.12 A move {p2}; builds {b2}")
    assert p1 == 1 and p2 == 2 and b2 == 1
    for x, e, f in ((x1, e1, f1), (x2, e2, f2)):
        eo, fo, _ = oracle.evaluate(top.desc, prm, x, top.box, skip_all_pairs=True)
        check_terms(e, eo, TERMS_AMBER, 1e-9)
        assert rel_rms(f, fo) <= 1e-8
