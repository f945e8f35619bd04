"""Pinning of the oracle's triclinic cells (CPU only).

The reference's periodic boundary is orthorhombic (PBC.cpp:72-81); the
triclinic cell is new here, so the oracle's handling of it is pinned by what
the geometry fixes:
  * the cell matrix has the requested edge lengths and angles, a along x, b
    in the x-y plane, and R H = I for the fractional map;
  * angles of exactly 90 degrees reproduce the orthorhombic results bit for
    bit;
  * an orthorhombic lattice written in the sheared basis a, a + b, a + b + c
    (every cell-matrix entry non-zero) is the same periodic system: every
    direct-space, bonded and correction energy and every force equals the
    orthorhombic evaluation, and so does the converged structure-factor
    reciprocal sum with its forces and virial;
  * in a genuinely triclinic cell the smooth PME converges to that
    structure-factor sum as the grid is refined, and its forces are minus the
    finite-difference gradient of its energy;
  * the total Ewald energy (direct + reciprocal + self + exclusion
    corrections) of a triclinic cell does not depend on the splitting
    parameter alpha;
  * a triclinic cell doubled along a (same angles, twice the grid along a)
    doubles every energy term, including the smooth-PME reciprocal energy,
    and gives both copies of an atom the force of the original;
  * the diagonal of the full virial is minus the energy's derivative under
    axial strain of coordinates and cell;
  * the rebuild criterion takes triclinic minimum images.
"""
import numpy as np
import pytest

import oracle
from md_amd import native as N
from md_amd import systems

from tri_systems import cell_vectors, doubled, mapped, sheared_box

TRI = [30.0, 30.0, 30.0, 80.0, 85.0, 75.0]  # a genuinely triclinic cell for the 30 A polymer
TERMS = ("bond", "angle", "dihedral", "urey_bradley", "improper", "cmap", "vdw", "ewald_direct", "ewald_self",
         "ewald_excl", "vdw14", "ee14")


def test_cell_matrix_geometry():
    box = [31.0, 29.0, 33.0, 70.0, 80.0, 65.0]
    h, face = oracle.cell_matrix(box)
    a, b, c = h[:, 0], h[:, 1], h[:, 2]
    assert np.allclose([np.linalg.norm(a), np.linalg.norm(b), np.linalg.norm(c)], box[:3], rtol=1e-14)

    def ang(u, v):
        return np.degrees(np.arccos(np.dot(u, v) / np.linalg.norm(u) / np.linalg.norm(v)))

    assert np.allclose([ang(b, c), ang(a, c), ang(a, b)], box[3:], rtol=0, atol=1e-10)
    assert a[1] == a[2] == b[2] == 0.0 and h[2, 2] > 0
    # face distance: volume over the area of the face spanned by the other two
    vol = abs(np.linalg.det(h))
    want = [vol / np.linalg.norm(np.cross(b, c)), vol / np.linalg.norm(np.cross(c, a)),
            vol / np.linalg.norm(np.cross(a, b))]
    assert np.allclose(face, want, rtol=1e-13)
    assert np.allclose(h, cell_vectors(box), rtol=1e-13, atol=1e-12)
    with pytest.raises(ValueError):
        oracle.cell_matrix([30, 30, 30, 30, 30, 100])  # the angles of no cell


def _polymer(polymer_2k, grid=(0, 0, 0), tol=1e-6, rc=10.0):
    prm_path, rst_path = polymer_2k
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, lj_cutoff=rc, lj_switch=rc - 1.0, ewald_tol=tol, pme_grid=grid)
    return top, prm


def test_right_angles_take_the_orthorhombic_path(polymer_2k):
    top, prm = _polymer(polymer_2k, grid=(32, 32, 32))
    L = list(top.box)
    e3, f3, w3 = oracle.evaluate(top.desc, prm, top.coords, L, skip_all_pairs=True, full_virial=True)
    e6, f6, w6 = oracle.evaluate(top.desc, prm, top.coords, L + [90.0, 90.0, 90.0], skip_all_pairs=True,
                                 full_virial=True)
    assert e3 == e6 and np.array_equal(f3, f6) and np.array_equal(w3, w6)


def test_sheared_basis_is_the_same_system(polymer_2k):
    """Direct space, bonded terms and corrections (no reciprocal part)."""
    top, prm = _polymer(polymer_2k)
    L = np.asarray(top.box)
    box6 = sheared_box(L)
    _, face = oracle.cell_matrix(box6)
    assert face.min() >= 2 * 10.0
    e_o, f_o, _ = oracle.evaluate(top.desc, prm, top.coords, L, skip_all_pairs=True)
    e_t, f_t, _ = oracle.evaluate(top.desc, prm, top.coords, box6, skip_all_pairs=True)
    for k in TERMS:
        assert e_t[k] == pytest.approx(e_o[k], rel=1e-11, abs=1e-9), k
    assert np.abs(f_t - f_o).max() <= 1e-9 * np.abs(f_o).max()
    assert e_o["vdw"] != 0 and e_o["ewald_direct"] != 0 and e_o["ewald_excl"] != 0


def _mmax(alpha, lengths):
    """Largest |m_i| of a structure-factor sum converged to 1e-13."""
    return int(np.ceil(np.sqrt(30.0) * alpha / np.pi * max(lengths))) + 1


def _charges(n, box6, seed):
    rng = np.random.default_rng(seed)
    q = rng.normal(0.0, 0.5, n)
    q -= q.mean()
    h, _ = oracle.cell_matrix(box6)
    x = rng.uniform(0.0, 1.0, (n, 3)) @ h.T
    return q, x


def test_sheared_basis_structure_factor_sum():
    L = np.array([20.0, 22.0, 24.0])
    box6 = sheared_box(L)
    q, x = _charges(40, box6, 3)
    alpha = 0.25
    # both sums hold every m with |m| < 0.44 / A (weight exp(-pi^2 m^2 / alpha^2) > 1e-13); along edge
    # vector a_i the integer m_i = a_i . m, so |m_i| <= 0.44 |a_i| (24 A, 38.2 A)
    e_o, f_o, w_o = oracle.ewald_recip_dense(q, x, L, alpha, _mmax(alpha, L))
    e_t, f_t, w_t = oracle.ewald_recip_dense(q, x, box6, alpha, _mmax(alpha, box6[:3]))
    assert e_t == pytest.approx(e_o, rel=1e-11)
    assert np.abs(f_t - f_o).max() <= 1e-10 * np.abs(f_o).max()
    assert np.abs(w_t - w_o).max() <= 1e-10 * np.abs(w_o).max()


def test_triclinic_pme_converges_to_the_structure_factor_sum():
    box6 = [24.0, 25.0, 26.0, 75.0, 82.0, 68.0]
    q, x = _charges(60, box6, 5)
    alpha = 0.3
    e_d, f_d, w_d = oracle.ewald_recip_dense(q, x, box6, alpha, _mmax(alpha, box6[:3]))
    errs = []
    for k in (16, 32, 64):
        e_p, f_p, w_p = oracle.pme_recip(q, x, box6, alpha, (k, k, k))
        errs.append((abs(e_p - e_d) / abs(e_d), np.abs(f_p - f_d).max() / np.abs(f_d).max(),
                     np.abs(w_p - w_d).max() / np.abs(w_d).max()))
    print(errs)
    # order-4 interpolation: each halving of the spacing cuts the error about 16-fold
    for j in range(3):
        assert errs[0][j] > 8 * errs[1][j] and errs[1][j] > 8 * errs[2][j]
    assert errs[2][0] < 1e-5 and errs[2][1] < 2e-4 and errs[2][2] < 1e-4


def test_triclinic_pme_forces_are_the_gradient():
    box6 = [24.0, 25.0, 26.0, 75.0, 82.0, 68.0]
    q, x = _charges(30, box6, 9)
    grid = (24, 27, 32)
    _, f, _ = oracle.pme_recip(q, x, box6, 0.35, grid)
    h = 1e-4  # truncation ~1e-9 and rounding (|E| ~ 400) ~1e-9 both stay small
    for i in (0, 7, 19):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, d] += h
            xm[i, d] -= h
            fd = -(oracle.pme_recip(q, xp, box6, 0.35, grid)[0] - oracle.pme_recip(q, xm, box6, 0.35, grid)[0]) / (2 * h)
            assert fd == pytest.approx(f[i, d], rel=1e-7, abs=1e-7)


def test_triclinic_ewald_total_is_alpha_independent():
    """Direct sum (within 13 A, erfc(alpha rc) < 1e-11 for both alphas) plus
    the structure-factor sum plus the self term, for charges in a
    triclinic cell; no pair is excluded."""
    box6 = [28.0, 29.0, 30.0, 78.0, 84.0, 72.0]
    h, face = oracle.cell_matrix(box6)
    rc = 13.0
    assert face.min() > 2 * rc
    q, x = _charges(50, box6, 13)
    from scipy.special import erfc
    kc = oracle.coulomb_constant()
    rinv = np.linalg.inv(h)
    totals = []
    for alpha in (0.38, 0.48):
        d = x[:, None, :] - x[None, :, :]
        s = d @ rinv.T
        d = (s - np.rint(s)) @ h.T  # the only image within rc (rc < half every face distance)
        r = np.linalg.norm(d, axis=2)
        iu = np.triu_indices(len(q), 1)
        rr = r[iu]
        qq = (q[:, None] * q[None, :])[iu]
        e_dir = kc * np.sum(np.where(rr < rc, qq * erfc(alpha * rr) / rr, 0.0))
        e_rec = oracle.ewald_recip_dense(q, x, box6, alpha, _mmax(alpha, box6[:3]))[0]
        e_self = -kc * alpha / np.sqrt(np.pi) * np.sum(q * q)
        totals.append(e_dir + e_rec + e_self)
    print(totals)
    assert totals[0] == pytest.approx(totals[1], rel=1e-9)


@pytest.fixture(scope="module")
def tri_pair(tmp_path_factory):
    """The 2.1k-atom polymer carried into a triclinic cell, and the same
    cell doubled along a with a second copy of the system."""
    d = tmp_path_factory.mktemp("tri")
    top_g = systems.small_polymer(seed=5)
    edge = float(top_g.box)
    one = copy_with(top_g, mapped(top_g.coords, edge, TRI))
    h = cell_vectors(TRI)
    two = doubled(one, h[:, 0])
    p1 = systems.write_system(one, str(d / "one"))
    p2 = systems.write_system(two, str(d / "two"))
    return p1, p2


def copy_with(top, coords):
    import copy
    t = copy.deepcopy(top)
    t.coords = np.asarray(coords, dtype=np.float64)
    return t


def test_doubled_triclinic_cell_doubles_every_term(tri_pair):
    (p1, r1), (p2, r2) = tri_pair
    t1 = N.Topology(p1, r1, N.SEM_AMBER)
    t2 = N.Topology(p2, r2, N.SEM_AMBER)
    n = t1.n_atoms
    assert t2.n_atoms == 2 * n
    prm1 = N.make_params(N.SEM_AMBER, N.PREC_FP64, lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-6,
                         pme_grid=(24, 24, 24))
    prm2 = N.make_params(N.SEM_AMBER, N.PREC_FP64, lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-6,
                         pme_grid=(48, 24, 24))
    assert prm1.ewald_alpha == prm2.ewald_alpha
    box2 = [2 * TRI[0]] + TRI[1:]
    e1, f1, w1 = oracle.evaluate(t1.desc, prm1, t1.coords, TRI, skip_all_pairs=True, full_virial=True)
    e2, f2, w2 = oracle.evaluate(t2.desc, prm2, t2.coords, box2, skip_all_pairs=True, full_virial=True)
    for k in TERMS + ("ewald_recip", "total"):
        assert e2[k] == pytest.approx(2 * e1[k], rel=1e-10, abs=1e-8), k
    assert e1["ewald_recip"] != 0
    scale = np.abs(f1).max()
    assert np.abs(f2[:n] - f1).max() <= 1e-9 * scale
    assert np.abs(f2[n:] - f1).max() <= 1e-9 * scale
    # the virial is extensive; its bonded part uses absolute positions, and
    # the copy's bonded forces sum to zero per term, so it doubles as well
    assert np.abs(w2 - 2 * w1).max() <= 1e-9 * np.abs(w1).max()


def test_triclinic_diagonal_virial_is_strain_derivative(tri_pair):
    (p1, r1), _ = tri_pair
    t1 = N.Topology(p1, r1, N.SEM_AMBER)
    prm = N.make_params(N.SEM_AMBER, N.PREC_FP64, lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-12,
                        pme_grid=(32, 32, 32))
    x = t1.coords
    _, _, w = oracle.evaluate(t1.desc, prm, x, TRI, skip_all_pairs=True, full_virial=True)
    h0 = cell_vectors(TRI)

    def box_of(hm):
        a, b, c = hm[:, 0], hm[:, 1], hm[:, 2]

        def ang(u, v):
            return float(np.degrees(np.arccos(np.dot(u, v) / np.linalg.norm(u) / np.linalg.norm(v))))
        return [np.linalg.norm(a), np.linalg.norm(b), np.linalg.norm(c), ang(b, c), ang(a, c), ang(a, b)]

    for axis in range(3):
        def e_at(hh, axis=axis):
            s = np.ones(3)
            s[axis] += hh
            # scaling along x, y or z keeps a along x and b in the x-y plane
            return oracle.evaluate(t1.desc, prm, x * s, box_of(np.diag(s) @ h0), skip_all_pairs=True,
                                   want_forces=False)[0]["total"]
        step = 1e-6
        fd = -(e_at(step) - e_at(-step)) / (2 * step)
        assert fd == pytest.approx(w[axis, axis], rel=1e-6), axis


def test_rebuild_criterion_takes_triclinic_images():
    box6 = [30.0, 31.0, 32.0, 70.0, 80.0, 65.0]
    h, _ = oracle.cell_matrix(box6)
    rng = np.random.default_rng(2)
    xref = rng.uniform(0, 1, (50, 3)) @ h.T
    skin = 2.0
    x = xref.copy()
    # a move of 0.99 skin/2 plus a lattice vector: no rebuild; 1.01 skin/2: rebuild
    u = np.array([0.3, -0.5, 0.81])
    u /= np.linalg.norm(u)
    for frac, want in ((0.99, False), (1.01, True)):
        x = xref.copy()
        x[7] += frac * 0.5 * skin * u + h @ np.array([1, -1, 2])
        assert oracle.pairs_need_rebuild(x, xref, box6, skin) == want
