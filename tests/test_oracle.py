"""Pinning of the CPU oracle (CPU only).

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong sign, a transposed index) fails it:
  * golden vectors from the compiled reference parser + formulas
    (tests/golden/*.npz, written by tests/golden/make_golden.py);
  * the reference's scalar formula samples on two-atom systems;
  * central finite differences of the energies against the forces;
  * brute-force enumeration of the REF pair list;
  * alpha-independence of the total Ewald electrostatics with the
    structure-factor reciprocal sum;
  * smooth-PME convergence to the structure-factor sum, and the virial
    against a finite-strain derivative;
  * closed forms of the order-4 B-spline.
"""
import hashlib
import math
import os
import tempfile

import numpy as np
import pytest

import oracle
from md_amd import native as N
from md_amd import systems

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
GOLDEN_CASES = {"ref_polymer_2k": lambda: systems.small_polymer(seed=5),
                "ref_water_6k": lambda: systems.water_box(seed=11)}


def _write(top, d, name):
    p, r = systems.write_system(top, os.path.join(d, name))
    digest = hashlib.sha256(open(p, "rb").read() + open(r, "rb").read()).hexdigest()
    return p, r, digest


@pytest.fixture(scope="module", params=sorted(GOLDEN_CASES))
def golden(request, tmp_path_factory):
    name = request.param
    g = dict(np.load(os.path.join(GOLDEN, name + ".npz")))
    d = str(tmp_path_factory.mktemp(name))
    p, r, digest = _write(GOLDEN_CASES[name](), d, name)
    assert digest == str(g["files_sha256"]), "the generator changed: regenerate tests/golden"
    return name, g, p, r


def test_parser_matches_reference(golden):
    _, g, p, r = golden
    t = N.Topology(p, r, N.SEM_REF)
    a = t.arrays
    assert t.n_atoms == int(g["n_atoms"])
    np.testing.assert_array_equal(t.box, g["box"][:3])
    np.testing.assert_array_equal(a["charge"], g["charge"])
    np.testing.assert_array_equal(a["type"] + 1, g["type_index"])
    np.testing.assert_array_equal(a["nb_index"], g["nb_matrix"])
    for k in ["lj_a", "lj_b", "lj14_a", "lj14_b"]:
        np.testing.assert_array_equal(a[k], g[k])
    # the reference's lists (last records dropped, CMAP included) term by term
    np.testing.assert_array_equal(a["bond_atoms"], g["bond_atoms"])
    np.testing.assert_array_equal(a["bond_k"][a["bond_type"]], g["bond_k"])
    np.testing.assert_array_equal(a["bond_r0"][a["bond_type"]], g["bond_r0"])
    np.testing.assert_array_equal(a["ub_atoms"], g["ub_atoms"])
    np.testing.assert_array_equal(a["ub_k"][a["ub_type"]], g["ub_k"])
    np.testing.assert_array_equal(a["angle_atoms"], g["angle_atoms"])
    np.testing.assert_array_equal(a["angle_theta0"][a["angle_type"]], g["angle_theta0"])
    np.testing.assert_array_equal(a["dihedral_atoms"], g["dih_atoms"])
    np.testing.assert_array_equal(t.dihedral_flags, g["dih_flags"])
    np.testing.assert_array_equal(a["dihedral_k"][a["dihedral_type"]], g["dih_k"])
    np.testing.assert_array_equal(a["dihedral_n"][a["dihedral_type"]], g["dih_n"])
    np.testing.assert_array_equal(a["dihedral_phase"][a["dihedral_type"]], g["dih_phase"])
    np.testing.assert_array_equal(a["improper_atoms"], g["imp_atoms"])
    np.testing.assert_array_equal(a["improper_psi0"][a["improper_type"]], g["imp_psi0"])
    np.testing.assert_array_equal(a["cmap_atoms"], g["cmap_atoms"])
    np.testing.assert_array_equal(a["cmap_set"], g["cmap_set"])
    if "cmap_coeff" in g and g["cmap_set"].size:
        np.testing.assert_array_equal(a["cmap_coeff"], g["cmap_coeff"][:a["cmap_coeff"].size])
    # the reference's exclusion lists are empty after parsing
    assert not np.any(g["n_excluded_after_parse"])
    assert a["excl_start"] is None
    # 1-4 pair map
    p14 = g["pairs14_from_dihedrals"].reshape(-1, 2)
    ours = {tuple(x) for x in a["pairs14"].reshape(-1, 2)}
    assert {tuple(sorted(x)) for x in p14} == ours


def test_bonded_energies_match_reference(golden):
    _, g, p, r = golden
    t = N.Topology(p, r, N.SEM_REF)
    prm = N.make_params(N.SEM_REF, N.PREC_FP64, lj_cutoff=min(10.0, t.box.min() / 2), lj_switch=0.0)
    e, _, _ = oracle.evaluate(t.desc, prm, t.coords, t.box, skip_all_pairs=True, want_forces=False)
    for ours, ref in [("bond", "E_bond"), ("angle", "E_angle"), ("dihedral", "E_dihedral"),
                      ("urey_bradley", "E_ub"), ("improper", "E_improper"), ("cmap", "E_cmap")]:
        assert abs(e[ours] - float(g[ref])) <= 1e-12 * abs(float(g[ref])) + 1e-15, (ours, e[ours], g[ref])


def _two_atom_desc(q1, q2, a, b, excluded_14=False):
    arr = {
        "charge": np.array([q1, q2]), "type": np.array([0, 0], np.int32), "n_types": 1,
        "nb_index": np.array([1], np.int32), "lj_a": np.array([a]), "lj_b": np.array([b]),
        "lj14_a": np.array([a]), "lj14_b": np.array([b]),
        "bond_atoms": np.zeros(0, np.int32), "bond_type": np.zeros(0, np.int32), "bond_k": np.zeros(0),
        "bond_r0": np.zeros(0), "ub_atoms": np.zeros(0, np.int32), "ub_type": np.zeros(0, np.int32),
        "ub_k": np.zeros(0), "ub_r0": np.zeros(0), "angle_atoms": np.zeros(0, np.int32),
        "angle_type": np.zeros(0, np.int32), "angle_k": np.zeros(0), "angle_theta0": np.zeros(0),
        "dihedral_atoms": np.zeros(0, np.int32), "dihedral_type": np.zeros(0, np.int32),
        "dihedral_k": np.zeros(0), "dihedral_n": np.zeros(0), "dihedral_phase": np.zeros(0),
        "improper_atoms": np.zeros(0, np.int32), "improper_type": np.zeros(0, np.int32),
        "improper_k": np.zeros(0), "improper_psi0": np.zeros(0),
        "cmap_atoms": np.zeros(0, np.int32), "cmap_set": np.zeros(0, np.int32), "cmap_n_sets": 0,
        "cmap_resolution": 0, "cmap_coeff": np.zeros(0), "cmap_grid": np.zeros(0),
        "excl_start": None, "excl_list": None, "excl_pairs": np.zeros(0, np.int32),
        "pairs14": np.zeros(0, np.int32), "scee14": np.zeros(0),
    }
    if excluded_14:
        arr["excl_start"] = np.array([0, 1, 2], np.int32)
        arr["excl_list"] = np.array([1, 0], np.int32)
        arr["pairs14"] = np.array([0, 1], np.int32)
        arr["scee14"] = np.array([1.2])
    keep = []
    return N.desc_from_arrays(arr, keep), keep


def test_scalar_formulas_match_reference_samples():
    g = dict(np.load(os.path.join(GOLDEN, "ref_polymer_2k.npz")))
    r2s, ab = g["sample_r2"], g["sample_lj_ab"].reshape(-1, 2)
    box = np.array([60.0, 60.0, 60.0])
    for i, r2 in enumerate(r2s):
        r = math.sqrt(r2)
        x = np.array([[10.0, 10.0, 10.0], [10.0 + r, 10.0, 10.0]])
        for j, (a, b) in enumerate(ab):
            d, keep = _two_atom_desc(-0.834, 0.417, a, b)
            # switch from 14.5 so S = 1 at every sample distance (r < 12)
            prm = N.make_params(N.SEM_REF, N.PREC_FP64, lj_cutoff=15.0, ee_cutoff=10.0 if r < 10 else 15.0,
                                lj_switch=14.5, ewald_alpha=0.3123)
            e, f, _ = oracle.evaluate(d, prm, x, box)
            assert e["vdw"] == pytest.approx(g["sample_lj_e"][3 * i + j], rel=1e-13, abs=1e-13)
            # F_x on atom 0 = (G_lj + G_coul) (x0 - x1) with x0 - x1 = -r
            dq, keepq = _two_atom_desc(-0.834, 0.417, 0.0, 0.0)
            fq = oracle.evaluate(dq, prm, x, box)[1]
            assert f[0, 0] - fq[0, 0] == pytest.approx(-g["sample_lj_g"][3 * i + j] * r, rel=1e-9, abs=1e-9)
        assert e["ee_pairlist"] == pytest.approx(g["sample_coul_e"][i], rel=1e-14)
        assert fq[0, 0] == pytest.approx(-g["sample_coul_g"][i] * r, rel=1e-12)
        assert e["ewald_direct"] == pytest.approx(g["sample_ewald_e"][i], rel=1e-13)
        if r < 10:
            assert e["ee_rf"] == pytest.approx(g["sample_coul_cut_e"][i], rel=1e-12, abs=1e-12)
            # the 1-4 Ewald form of the reference (reached only for excluded 1-4 pairs)
            d14, keep14 = _two_atom_desc(-0.834, 0.417, 0.0, 0.0, excluded_14=True)
            e14, _, _ = oracle.evaluate(d14, prm, x, box)
            assert e14["ewald_direct"] == pytest.approx(g["sample_ewald14_e"][i], rel=1e-12)


@pytest.fixture(scope="module")
def poly(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("poly"))
    p, r, _ = _write(systems.small_polymer(seed=5), d, "poly")
    return p, r


def _fd_check(desc, prm, x, box, atoms, h=1e-5, rtol=2e-5):
    e0, f, _ = oracle.evaluate(desc, prm, x, box)
    worst = 0.0
    for i in atoms:
        for c in range(3):
            xp = x.copy()
            xm = x.copy()
            xp[i, c] += h
            xm[i, c] -= h
            ep = oracle.evaluate(desc, prm, xp, box, want_forces=False)[0]["total"]
            em = oracle.evaluate(desc, prm, xm, box, want_forces=False)[0]["total"]
            fd = -(ep - em) / (2 * h)
            scale = max(1.0, abs(f[i]).max())
            worst = max(worst, abs(fd - f[i, c]) / scale)
    assert worst <= rtol, worst


def test_ref_forces_are_minus_gradient_except_torsions(poly):
    """REF forces are -grad E for bonds, UB, angles, LJ and Coulomb.  The
    torsion terms are checked separately: the reference's torsion force is not
    the gradient of its torsion energy."""
    p, r = poly
    t = N.Topology(p, r, N.SEM_REF)
    arr = dict(t.arrays)
    for k in ["dihedral_k", "improper_k", "cmap_coeff"]:
        arr[k] = np.zeros_like(arr[k])
    keep = []
    d = N.desc_from_arrays(arr, keep)
    prm = N.make_params(N.SEM_REF, N.PREC_FP64, lj_cutoff=10.0, lj_switch=9.0)
    atoms = np.random.default_rng(0).choice(t.n_atoms, 12, replace=False)
    _fd_check(d, prm, t.coords, t.box, atoms, h=1e-4, rtol=2e-5)


def _one_torsion(kind):
    x = np.array([[10.1, 10.3, 9.8], [11.0, 11.1, 10.2], [12.3, 10.8, 10.9], [12.9, 11.9, 11.6],
                  [14.2, 11.5, 12.0]])
    arr = dict(_two_atom_desc_arrays(5))
    if kind == "dihedral":
        arr.update(dihedral_atoms=np.array([0, 1, 2, 3], np.int32), dihedral_type=np.array([0], np.int32),
                   dihedral_k=np.array([1.7]), dihedral_n=np.array([3.0]), dihedral_phase=np.array([0.4]))
    elif kind == "improper":
        arr.update(improper_atoms=np.array([0, 1, 2, 3], np.int32), improper_type=np.array([0], np.int32),
                   improper_k=np.array([20.0]), improper_psi0=np.array([0.9]))
    return arr, x


def _two_atom_desc_arrays(n):
    d = {
        "charge": np.zeros(n), "type": np.zeros(n, np.int32), "n_types": 1,
        "nb_index": np.array([1], np.int32), "lj_a": np.zeros(1), "lj_b": np.zeros(1),
        "lj14_a": np.zeros(1), "lj14_b": np.zeros(1),
        "cmap_atoms": np.zeros(0, np.int32), "cmap_set": np.zeros(0, np.int32), "cmap_n_sets": 0,
        "cmap_resolution": 0, "cmap_coeff": np.zeros(0), "cmap_grid": np.zeros(0),
        "excl_start": None, "excl_list": None, "excl_pairs": np.zeros(0, np.int32),
        "pairs14": np.zeros(0, np.int32), "scee14": np.zeros(0),
    }
    for name, tabs in [("bond", ["bond_k", "bond_r0"]), ("ub", ["ub_k", "ub_r0"]),
                       ("angle", ["angle_k", "angle_theta0"]),
                       ("dihedral", ["dihedral_k", "dihedral_n", "dihedral_phase"]),
                       ("improper", ["improper_k", "improper_psi0"])]:
        d[f"{name}_atoms"] = np.zeros(0, np.int32)
        d[f"{name}_type"] = np.zeros(0, np.int32)
        for tb in tabs:
            d[tb] = np.zeros(0)
    return d


@pytest.mark.parametrize("kind", ["dihedral", "improper"])
def test_torsion_force_distributions(kind):
    """AMBER torsion forces are -grad E; REF ones follow the reference's
    distribution: F_a = +g_a, F_d = +g_d, F_b = -g_b - 2 g_a, F_c = -g_c - 2 g_d
    where g = grad E from finite differences."""
    arr, x = _one_torsion(kind)
    box = np.array([40.0, 40.0, 40.0])
    keep = []
    d = N.desc_from_arrays(arr, keep)
    g = np.zeros((4, 3))
    h = 1e-6
    prm_a = N.make_params(N.SEM_AMBER, N.PREC_FP64, lj_cutoff=9.0, lj_switch=0.0, ewald_alpha=0.3)
    for a in range(4):
        for c in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a, c] += h
            xm[a, c] -= h
            g[a, c] = (oracle.evaluate(d, prm_a, xp, box)[0]["total"] -
                       oracle.evaluate(d, prm_a, xm, box)[0]["total"]) / (2 * h)
    fa = oracle.evaluate(d, prm_a, x, box)[1][:4]
    np.testing.assert_allclose(fa, -g, atol=1e-7 * np.abs(g).max())
    prm_r = N.make_params(N.SEM_REF, N.PREC_FP64, lj_cutoff=9.0, lj_switch=0.0, ewald_alpha=0.3)
    fr = oracle.evaluate(d, prm_r, x, box)[1][:4]
    expect = np.array([g[0], -g[1] - 2 * g[0], -g[2] - 2 * g[3], g[3]])
    np.testing.assert_allclose(fr, expect, atol=1e-7 * np.abs(g).max())


def test_amber_forces_are_minus_gradient(poly):
    p, r = poly
    t = N.Topology(p, r, N.SEM_AMBER)
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-5,
                        pme_grid=(32, 32, 32))
    # polymer atoms (every bonded term type) and a few waters
    atoms = list(range(0, 300, 37)) + [300, 301, 1500]
    _fd_check(t.desc, prm, t.coords, t.box, atoms)


def test_ref_pairlist_equals_brute_force(poly):
    p, r = poly
    t = N.Topology(p, r, N.SEM_REF)
    rl = 7.0
    pl = oracle.ref_pairlist(t.desc, t.coords, t.box, rl)
    x, L = t.coords, t.box
    i, j = np.triu_indices(t.n_atoms, 1)
    d = x[i] - x[j]
    d -= L * np.round(d / L)
    r2 = np.sum(d * d, axis=1)
    keep = (r2 <= rl * rl) & (r2 > 1e-12)
    brute = set(zip(i[keep].tolist(), j[keep].tolist()))
    got = {(int(a), int(b)) for a, b, _ in pl}
    assert len(got) == len(pl)  # no duplicates
    assert got == brute


def test_ee_all_equals_pairlist_sum(poly):
    p, r = poly
    t = N.Topology(p, r, N.SEM_REF)
    prm = N.make_params(N.SEM_REF, N.PREC_FP64, lj_cutoff=10.0, lj_switch=9.0)
    e, _, _ = oracle.evaluate(t.desc, prm, t.coords, t.box, want_forces=False)
    assert e["ee_all"] == pytest.approx(e["ee_pairlist"], rel=1e-12)


@pytest.fixture(scope="module")
def small_water(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("sw"))
    top = systems.water_box(seed=3, box=15.0, n_water=100, n_lattice=5, jitter=0.2)
    p, r, _ = _write(top, d, "sw")
    return N.Topology(p, r, N.SEM_AMBER)


def _electrostatics(t, alpha, rc, mmax):
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, lj_cutoff=rc, lj_switch=0.0, ewald_alpha=alpha)
    e, f, _ = oracle.evaluate(t.desc, prm, t.coords, t.box)
    erec, frec, _ = oracle.ewald_recip_dense(t.arrays["charge"], t.coords, t.box, alpha, mmax)
    parts = [e["ewald_direct"], e["ewald_self"], e["ewald_excl"], e["ee14"], erec]
    return sum(parts), f + frec, sum(abs(v) for v in parts)


def test_ewald_total_is_alpha_independent(small_water):
    t = small_water
    rc = 7.4
    e1, f1, s1 = _electrostatics(t, 0.62, rc, 16)
    e2, f2, s2 = _electrostatics(t, 0.70, rc, 18)
    # the alpha-dependent parts change by ~s/10 between the two alphas
    print(e1, e2, s1, s2)
    assert abs(e1 - e2) <= 1e-9 * max(s1, s2)
    # forces: the LJ part is identical in both evaluations, so the difference is
    # the electrostatic force difference
    assert np.abs(f1 - f2).max() <= 1e-7 * np.abs(f1).max()


def test_ewald_total_is_alpha_independent_with_14(poly):
    p, r = poly
    t = N.Topology(p, r, N.SEM_AMBER)
    e1, f1, s1 = _electrostatics(t, 0.47, 10.0, 21)
    e2, f2, s2 = _electrostatics(t, 0.52, 10.0, 23)
    print(e1, e2, s1, s2)
    assert abs(e1 - e2) <= 1e-9 * max(s1, s2)
    assert np.abs(f1 - f2).max() <= 1e-6 * np.abs(f1).max()


def test_pme_converges_to_structure_factor_sum(small_water):
    t = small_water
    q, x, L = t.arrays["charge"], t.coords, t.box
    alpha = 0.35
    e_ref, f_ref, v_ref = oracle.ewald_recip_dense(q, x, L, alpha, 12)
    errs = []
    for k in (16, 24, 32, 48):
        e, f, v = oracle.pme_recip(q, x, L, alpha, (k, k, k))
        errs.append(abs(e - e_ref) / abs(e_ref))
        ferr = np.sqrt(np.sum((f - f_ref) ** 2) / np.sum(f_ref ** 2))
    # monotone, and order-4 B-splines converge like K^-4: halving the spacing
    # (24 -> 48) must gain at least 8x
    assert errs[-1] < 5e-5 and all(a > b for a, b in zip(errs, errs[1:])), errs
    assert errs[1] / errs[3] > 8.0, errs
    assert ferr < 1e-3
    assert np.abs(v - v_ref).max() < 1e-3 * np.abs(v_ref).max()


@pytest.mark.parametrize("kind", ["pme", "dense"])
def test_recip_virial_is_strain_derivative(small_water, kind):
    t = small_water
    q, x, L = t.arrays["charge"], t.coords, t.box
    alpha = 0.35

    def energy(scale):
        if kind == "pme":
            return oracle.pme_recip(q, x * scale, L * scale, alpha, (24, 24, 24))[0]
        return oracle.ewald_recip_dense(q, x * scale, L * scale, alpha, 9)[0]

    fn = oracle.pme_recip if kind == "pme" else oracle.ewald_recip_dense
    v = fn(q, x, L, alpha, (24, 24, 24))[2] if kind == "pme" else fn(q, x, L, alpha, 9)[2]
    h = 1e-5
    # isotropic strain: -dE/d(eps) = trace of the virial
    dE = (energy(1 + h) - energy(1 - h)) / (2 * h)
    assert -dE == pytest.approx(np.trace(v), rel=1e-6)
    # anisotropic strain along x
    if kind == "pme":
        sx = np.array([1 + h, 1, 1])
        ep = oracle.pme_recip(q, x * sx, L * sx, alpha, (24, 24, 24))[0]
        sx = np.array([1 - h, 1, 1])
        em = oracle.pme_recip(q, x * sx, L * sx, alpha, (24, 24, 24))[0]
        assert -(ep - em) / (2 * h) == pytest.approx(v[0, 0], rel=1e-5)


def test_bspline_closed_forms():
    for w in np.linspace(0.0, 0.999, 23):
        th, dth = oracle.bspline(4, w)
        assert th.sum() == pytest.approx(1.0, abs=1e-15)
        assert dth.sum() == pytest.approx(0.0, abs=1e-15)
        # M4(u) = u^3/6 on [0, 1] and (4 - u)^3 / 6 on [3, 4]
        assert th[0] == pytest.approx(w ** 3 / 6, rel=1e-14, abs=1e-16)
        assert th[3] == pytest.approx((1 - w) ** 3 / 6, rel=1e-14, abs=1e-16)
        if 0.0 < w < 0.99:
            h = 1e-6
            thp, _ = oracle.bspline(4, w + h)
            thm, _ = oracle.bspline(4, w - h)
            np.testing.assert_allclose((thp - thm) / (2 * h), dth, atol=1e-8)
    th5, _ = oracle.bspline(5, 0.3)
    assert th5.sum() == pytest.approx(1.0, abs=1e-15)
