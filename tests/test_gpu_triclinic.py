"""Triclinic cells on the GPU (mdgpu_set_box_angles) against the oracle.

Three cells hold the 2.1k-atom polymer: the genuinely triclinic cells of
edges 30 A and angles 80, 85, 75 degrees and of the truncated octahedron
(109.47 degrees; the cubic system carried into each by x' = H x / 30), and
its own cubic box written in the sheared basis a, a + b, a + b + c (every
cell-matrix entry non-zero).  The library evaluates a cell in its reduced
basis (the same lattice; the PME grid runs along the reduced edges), so the
sheared basis is evaluated as the cube and the oracle is evaluated in the
reduced basis (tri_systems.reduced_box; the other two cells are reduced
already).  Bars, as for orthorhombic boxes:
  * FP64 vs oracle: per-term |dE| <= 1e-9 |E|, force relative RMS <= 1e-8;
    the full virial within 1e-9 of its largest component;
  * mixed vs FP64: force relative RMS <= 1e-5;
  * a Verlet list kept across moves below skin / 2 gives the oracle's result
    at the moved positions (FP64 1e-9) and is built once;
  * the rebuild criterion takes triclinic minimum images: a move of 0.99 and
    1.01 times skin / 2 plus a lattice vector answers as the oracle does;
  * twelve FP64 velocity-Verlet steps with a Verlet list follow the oracle's
    trajectory (positions 1e-9 A);
  * REF semantics, the adjoint and faces closer than twice the cutoff are
    refused with an error;
  * the C++ driver (md_main --minimize) takes the angles of a triclinic rst7
    and minimises as the C ABI does (1e-5, its six printed digits);
  * the 29k-atom polymer in a 66 A cell of the same angles keeps the
    minimum-image cluster search and matches the oracle (FP64 bars above).
"""
import numpy as np
import pytest
import torch

import oracle
from md_amd import native as N

from md_systems import thermal_velocities
from test_gpu_parity import TERMS_AMBER, check_terms, rel_rms
from tri_systems import cell_vectors, mapped, reduced_box, sheared_box

pytestmark = pytest.mark.gpu

TRI_ANGLES = [80.0, 85.0, 75.0]
OCTAHEDRON = [109.4712206, 109.4712206, 109.4712206]  # the truncated octahedron's cell (faces 24.5 A at 30 A)
KW = dict(lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-5, pme_grid=(32, 32, 32))


def _cell(polymer_2k, kind):
    prm_path, rst_path = polymer_2k
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    edge = float(top.box[0])
    if kind in ("triclinic", "octahedron"):
        box6 = [edge, edge, edge] + (TRI_ANGLES if kind == "triclinic" else OCTAHEDRON)
        x = mapped(top.coords, edge, box6)
    else:
        box6 = sheared_box(top.box)
        x = np.asarray(top.coords, dtype=np.float64).reshape(-1, 3)
    return top, np.ascontiguousarray(x), [float(v) for v in box6]


def _ev(top, prm, box6, skin=0.0, virial=False):
    ev = N.Evaluator(0)
    ev.upload(top.desc)
    ev.set_params(prm)
    ev.set_box_angles(box6[3:])
    ev.set_verlet(skin)
    ev.set_virial(virial)
    return ev


def _dev(x):
    t = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64).reshape(-1)).to("cuda:0")
    torch.cuda.synchronize()
    return t


@pytest.mark.parametrize("kind", ["triclinic", "sheared", "octahedron"])
def test_fp64_matches_oracle(polymer_2k, kind):
    print(f"case polymer_2k {kind} cell FP64 vs oracle", flush=True)
    top, x, box6 = _cell(polymer_2k, kind)
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, **KW)
    ev = _ev(top, prm, box6, virial=True)
    try:
        assert np.array_equal(ev.box_angles(), box6[3:])
        e, f, _ = ev.evaluate(x, box6[:3])
        w = ev.virial()
        mode = ev.search_mode()
    finally:
        ev.close()
    # a 30 A cell with a 10 A cutoff: the minimum-image search's bound on the
    # cluster half-extents (half of half the smallest face distance minus the
    # cutoff, over sqrt 3) is at most 0.6 A here, so the general search runs
    print(f"  search mode {mode}")
    assert mode == 1
    eo, fo, wo = oracle.evaluate(top.desc, prm, x, reduced_box(box6), skip_all_pairs=True, full_virial=True)
    check_terms(e, eo, TERMS_AMBER, 1e-9)
    r = rel_rms(f, fo)
    dw = np.abs(w - wo).max() / np.abs(wo).max()
    print(f"  force rel-RMS {r:.2e}, virial {dw:.2e}")
    assert r <= 1e-8
    assert dw <= 1e-9


@pytest.mark.parametrize("kind", ["triclinic", "sheared", "octahedron"])
def test_mixed_matches_fp64(polymer_2k, kind):
    print(f"case polymer_2k {kind} cell mixed vs FP64", flush=True)
    top, x, box6 = _cell(polymer_2k, kind)
    out = []
    for prec in (N.PREC_FP64, N.PREC_MIXED):
        ev = _ev(top, N.make_params(N.SEM_AMBER, prec, **KW), box6)
        try:
            out.append(ev.evaluate(x, box6[:3]))
        finally:
            ev.close()
    (e64, f64, _), (e32, f32, _) = out
    r = rel_rms(f32, f64)
    print(f"  force rel-RMS {r:.2e}, total {e32['total']:.6f} vs {e64['total']:.6f}")
    assert r <= 1e-5
    assert abs(e32["total"] - e64["total"]) <= 1e-5 * sum(abs(e64[k]) for k in TERMS_AMBER if k != "total")


def test_verlet_list_follows_moves(polymer_2k):
    print("case polymer_2k triclinic cell, Verlet list kept across moves", flush=True)
    top, x0, box6 = _cell(polymer_2k, "triclinic")
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, **KW)
    rng = np.random.default_rng(17)
    moves = rng.normal(0.0, 1.0, x0.shape)
    moves *= 0.45 / np.linalg.norm(moves, axis=1).max()  # largest move 0.45 A < skin / 2
    x1 = x0 + moves
    ev = _ev(top, prm, box6, skin=1.0)
    try:
        ev.evaluate(x0, box6[:3])
        e, f, _ = ev.evaluate(x1, box6[:3])
        builds = ev.list_builds()
    finally:
        ev.close()
    print(f"  builds {builds}")
    assert builds == (1, 2)
    eo, fo, _ = oracle.evaluate(top.desc, prm, x1, box6, skip_all_pairs=True)
    check_terms(e, eo, TERMS_AMBER, 1e-9)
    assert rel_rms(f, fo) <= 1e-8


def test_rebuild_criterion_takes_triclinic_images(polymer_2k):
    print("case polymer_2k triclinic cell rebuild criterion", flush=True)
    top, x0, box6 = _cell(polymer_2k, "triclinic")
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, **KW)
    skin = 1.0
    h = cell_vectors(box6)
    u = np.array([0.3, -0.5, 0.81])
    u /= np.linalg.norm(u)
    ev = _ev(top, prm, box6, skin=skin)
    try:
        xd = _dev(x0)
        fd = torch.zeros_like(xd)
        ev.evaluate_device(xd.data_ptr(), box6[:3], fd.data_ptr())
        for frac in (0.99, 1.01):
            x = x0.copy()
            x[123] += frac * 0.5 * skin * u + h @ np.array([1.0, -1.0, 2.0])
            got = ev.needs_rebuild(_dev(x).data_ptr())
            want = oracle.pairs_need_rebuild(x, x0, box6, skin)
            print(f"  move {frac} skin/2 plus a lattice vector: gpu {got} oracle {want}")
            assert got == want == (frac > 1.0)
    finally:
        ev.close()


def test_trajectory_matches_oracle(polymer_2k):
    print("case polymer_2k triclinic cell FP64 trajectory vs oracle", flush=True)
    top, x0, box6 = _cell(polymer_2k, "triclinic")
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, **KW)
    v0 = thermal_velocities(top.masses, 7, temperature=50.0)
    dt, n = 0.0002, 12
    ev = _ev(top, prm, box6, skin=0.8)
    try:
        ev.md_set_masses(top.masses)
        xd, vd = _dev(x0), _dev(v0)
        fd = torch.zeros_like(xd)
        ev.evaluate_device(xd.data_ptr(), box6[:3], fd.data_ptr())
        tr, _ = ev.md_run(xd.data_ptr(), vd.data_ptr(), fd.data_ptr(), box6[:3], dt, n)
        x = xd.cpu().numpy().reshape(-1, 3)
    finally:
        ev.close()
    _, f0, _ = oracle.evaluate(top.desc, prm, x0, box6, skip_all_pairs=True)
    xo, _, _, tro = oracle.md_run(top.desc, prm, top.masses, x0, v0, f0, box6, dt, n)
    dx = np.abs(x - xo).max()
    print(f"  positions {dx:.2e} A")
    assert dx <= 1e-9
    assert np.all(np.abs(tr[:, 0] - tro[:, 0]) <= 1e-9 * np.abs(tro[:, 0]))


def test_refusals(polymer_2k):
    print("case triclinic refusals", flush=True)
    top, x, box6 = _cell(polymer_2k, "triclinic")
    prm_ref = N.make_params(N.SEM_REF, N.PREC_FP64, lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-5,
                            pme_grid=(0, 0, 0))
    ev = _ev(top, prm_ref, box6)
    try:
        with pytest.raises(N.GpuError, match="AMBER"):
            ev.evaluate(x, box6[:3])
    finally:
        ev.close()
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, **KW)
    ev = _ev(top, prm, box6)
    try:
        with pytest.raises(N.GpuError, match="orthorhombic"):
            ev.adjoint(x, box6[:3], {"total": 1.0}, top.desc)
        with pytest.raises(N.GpuError):
            ev.set_box_angles([90.0, 90.0, 180.0])
        big = N.make_params(N.SEM_AMBER, N.PREC_FP64, lj_cutoff=14.5, lj_switch=13.0, ewald_tol=1e-5,
                            pme_grid=(32, 32, 32))
        ev.set_params(big)
        with pytest.raises(N.GpuError, match="faces"):
            ev.evaluate(x, box6[:3])
        # back to a right-angled cell: the orthorhombic path again
        ev.set_params(prm)
        ev.set_box_angles([90.0, 90.0, 90.0])
        e, _, _ = ev.evaluate(top.coords, top.box)
    finally:
        ev.close()
    eo, _, _ = oracle.evaluate(top.desc, prm, top.coords, top.box, skip_all_pairs=True)
    assert abs(e["total"] - eo["total"]) <= 1e-9 * abs(eo["total"])


def test_cpp_driver_reads_cell_angles(polymer_2k, tmp_path):
    """md_main --minimize on a triclinic rst7: the C++ System passes the file's
    angles to the GPU (AMBER semantics), so its minimisation equals the C ABI
    one with mdgpu_set_box_angles (same settings as the orthorhombic driver
    test: rc 10 A, switch 9 A, skin 2 A, Ewald tolerance 1e-6, the reference's
    PME grid formula rounded up to 2^a 3^b)."""
    import math
    import os
    import re
    import subprocess
    from md_amd import systems
    print("case md_main --minimize 10 on a triclinic rst7", flush=True)
    prm_path, _ = polymer_2k
    top, x, box6 = _cell(polymer_2k, "triclinic")
    rst = str(tmp_path / "tri.rst7")
    N.write_rst7(rst, x, box=box6)
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "md_amd", "md_main")
    out = subprocess.run([exe, prm_path, rst, "--minimize", "10"], capture_output=True, text=True, timeout=120)
    print(out.stdout[-400:], out.stderr[-300:])
    assert out.returncode == 0
    e_cpp = float(re.search(r"Minimized:.*\nEnergy: (\S+)", out.stdout).group(1))
    xr = N.read_rst7(rst)["coords"]
    tol, rc = 1e-6, 10.0
    alpha = math.sqrt(-math.log(2.0 * tol)) / rc
    grid = tuple(systems.smallest_2a3b(int(2.0 * alpha * L / (3.0 * tol ** 0.2))) for L in box6[:3])
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, lj_cutoff=rc, lj_switch=9.0, ewald_alpha=alpha, pme_grid=grid)
    ev = _ev(top, prm, box6, skin=2.0)
    try:
        xd = _dev(xr)
        res = ev.minimize(xd.data_ptr(), box6[:3], max_iter=10, rms_force_tol=1e-2, max_step=0.2)
    finally:
        ev.close()
    print(f"  C++ driver {e_cpp:.6f}; C ABI {res}")
    assert abs(e_cpp - res["energy"]) <= 1e-5 * abs(res["energy"])


def test_minimum_image_search_in_a_triclinic_cell(polymer_29k):
    """The 29k-atom polymer carried into a 66 A cell of angles 80, 85, 75: its
    clusters are narrow enough for the minimum-image search (fractional
    rounding of cluster separations), which stays in use (search mode 0) in
    every evaluation below:
      * FP64 matches the oracle as above;
      * mixed vs FP64: force rel-RMS <= 1e-5 and the per-term bars of the
        cubic 29k workload (test_gpu_cfg2.MIXED_BARS);
      * kept lists (skins 1, 0.5 and 0.25 A) built at x and pruned, then used
        at x + moves of at most 0.45 skin (FP64), match the oracle there.  The
        minimum-image search's bound tightens with the skin (half of half the
        smallest face distance minus cutoff + skin, over sqrt 3: 5.62 A without
        a skin, 5.33 A with 1 A), and the clusters of a kept list grow as their
        atoms move; the search each kept list ends with is printed."""
    from test_gpu_cfg2 import MIXED_BARS
    print("case polymer_29k triclinic cell, minimum-image search", flush=True)
    prm_path, rst_path = polymer_29k
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    edge = float(top.box[0])
    box6 = [edge, edge, edge] + TRI_ANGLES
    x = np.ascontiguousarray(mapped(top.coords, edge, box6))
    kw = dict(lj_cutoff=12.0, lj_switch=11.0, ewald_tol=1e-5, pme_grid=(72, 72, 72))
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, **kw)
    res = {}
    for name, prec in (("fp64", N.PREC_FP64), ("mixed", N.PREC_MIXED)):
        ev = _ev(top, N.make_params(N.SEM_AMBER, prec, **kw), box6)
        try:
            res[name] = ev.evaluate(x, box6[:3])
            mode = ev.search_mode()
        finally:
            ev.close()
        print(f"  {name}: search mode {mode}")
        assert mode == 0
    e, f, _ = res["fp64"]
    eo, fo, _ = oracle.evaluate(top.desc, prm, x, box6, skip_all_pairs=True)
    check_terms(e, eo, TERMS_AMBER, 1e-9)
    r = rel_rms(f, fo)
    print(f"  force rel-RMS {r:.2e}")
    assert r <= 1e-8
    e32, f32, _ = res["mixed"]
    bad = []
    for k in TERMS_AMBER:
        rel = abs(e32[k] - e[k]) / max(abs(e[k]), 1e-300)
        print(f"  {k:14s} mixed vs fp64 rel {rel:.2e} (bar {MIXED_BARS[k]:.0e})")
        if not rel <= MIXED_BARS[k]:
            bad.append((k, rel))
    r32 = rel_rms(f32, f)
    print(f"  mixed force rel-RMS {r32:.2e}")
    assert r32 <= 1e-5
    assert not bad, bad
    # kept lists across moves (FP64)
    modes = {}
    for skin in (1.0, 0.5, 0.25):
        rng = np.random.default_rng(29)
        moves = rng.normal(0.0, 1.0, x.shape)
        moves *= 0.45 * skin / np.linalg.norm(moves, axis=1).max()
        x1 = x + moves
        ev = _ev(top, prm, box6, skin=skin)
        try:
            ev.evaluate(x, box6[:3])
            e1, f1, _ = ev.evaluate(x1, box6[:3])
            builds, modes[skin], prunes = ev.list_builds(), ev.search_mode(), ev.list_prunes()
        finally:
            ev.close()
        print(f"  kept list, skin {skin}: builds {builds}, prunes {prunes}, search mode {modes[skin]}")
        assert builds == (1, 2)
        eo1, fo1, _ = oracle.evaluate(top.desc, prm, x1, box6, skip_all_pairs=True)
        check_terms(e1, eo1, TERMS_AMBER, 1e-9)
        assert rel_rms(f1, fo1) <= 1e-8


def test_sheared_basis_is_binned_in_its_reduced_cell(polymer_29k):
    """The cubic 29k workload written in the sheared basis a, a + b, a + b + c
    (edges 66, 93.3, 114.3 A; angles 35.3, 54.7, 45 degrees).  Its columns in
    that basis would be too wide for the minimum-image search; the library
    bins, searches and grids it in the reduced basis of the same lattice (the
    cube), so the minimum-image search runs (search mode 0).  Bars as above:
    FP64 vs the oracle in the reduced basis 1e-9 per term and 1e-8 force
    rel-RMS, virial 1e-9; mixed vs FP64 force rel-RMS 1e-5; the direct-space
    and bonded terms equal those of the cubic box (same lattice) to 1e-12,
    FP64."""
    print("case polymer_29k in a sheared basis, reduced binning cell", flush=True)
    prm_path, rst_path = polymer_29k
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    box6 = [float(v) for v in sheared_box(top.box)]
    x = np.ascontiguousarray(np.asarray(top.coords, dtype=np.float64).reshape(-1, 3))
    kw = dict(lj_cutoff=12.0, lj_switch=11.0, ewald_tol=1e-5, pme_grid=(72, 72, 72))
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, **kw)
    res = {}
    for name, prec in (("fp64", N.PREC_FP64), ("mixed", N.PREC_MIXED)):
        ev = _ev(top, N.make_params(N.SEM_AMBER, prec, **kw), box6, virial=name == "fp64")
        try:
            res[name] = ev.evaluate(x, box6[:3])
            if name == "fp64":
                w = ev.virial()
            mode = ev.search_mode()
        finally:
            ev.close()
        print(f"  {name}: search mode {mode}")
        assert mode == 0
    e, f, _ = res["fp64"]
    eo, fo, wo = oracle.evaluate(top.desc, prm, x, reduced_box(box6), skip_all_pairs=True, full_virial=True)
    check_terms(e, eo, TERMS_AMBER, 1e-9)
    r = rel_rms(f, fo)
    dw = np.abs(w - wo).max() / np.abs(wo).max()
    print(f"  FP64 vs oracle: force rel-RMS {r:.2e}, virial {dw:.2e}")
    assert r <= 1e-8
    assert dw <= 1e-9
    r32 = rel_rms(res["mixed"][1], f)
    print(f"  mixed force rel-RMS {r32:.2e}")
    assert r32 <= 1e-5
    # the same lattice as an orthorhombic box: every term but the reciprocal one
    ev = _ev(top, prm, [float(v) for v in top.box] + [90.0, 90.0, 90.0])
    try:
        ec, _, _ = ev.evaluate(x, [float(v) for v in top.box])
    finally:
        ev.close()
    for k in ("bond", "angle", "dihedral", "urey_bradley", "improper", "cmap", "vdw", "ewald_direct",
              "ewald_self", "ewald_excl", "vdw14", "ee14"):
        rel = abs(e[k] - ec[k]) / max(abs(ec[k]), 1e-300)
        print(f"  {k:14s} sheared vs cubic rel {rel:.2e}")
        assert rel <= 1e-12, k
