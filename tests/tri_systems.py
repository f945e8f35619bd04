"""Triclinic test systems built from the cubic synthetic ones.

  * sheared_box(L): the lattice of an L1 x L2 x L3 orthorhombic box written in
    the sheared basis a, a + b, a + b + c, as edge lengths and angles (the
    same periodic system, every cell-matrix entry non-zero);
  * reduced_box(box6): the same lattice in its reduced basis (off-diagonal
    edge components at most half the diagonal ones), as edge lengths and
    angles: the cell the GPU library evaluates a given cell in;
  * mapped(coords, L, box6): coordinates of a cubic box of edge L carried
    into a triclinic cell of the same edge lengths by x' = H x / L (a
    genuinely triclinic system; molecules are sheared with the cell);
  * doubled(top, shift): a synthetic topology with a second copy of every
    atom and term, the copy displaced by shift.
"""
import copy

import numpy as np


def _angle(u, v):
    return float(np.degrees(np.arccos(np.dot(u, v) / (np.linalg.norm(u) * np.linalg.norm(v)))))


def sheared_box(lengths):
    """[|a|, |a+b|, |a+b+c|, alpha, beta, gamma] of the orthorhombic lattice
    with edge lengths `lengths` in the basis a, a + b, a + b + c."""
    l1, l2, l3 = (float(v) for v in lengths)
    a = np.array([l1, 0.0, 0.0])
    b = np.array([l1, l2, 0.0])
    c = np.array([l1, l2, l3])
    return [np.linalg.norm(a), np.linalg.norm(b), np.linalg.norm(c), _angle(b, c), _angle(a, c), _angle(a, b)]


def cell_vectors(box6):
    """H (edge vectors as columns) for edge lengths and angles in degrees:
    a along x, b in the x-y plane, c with a positive z component."""
    la, lb, lc, al, be, ga = (float(v) for v in box6)
    ca, cb, cg = np.cos(np.radians([al, be, ga]))
    sg = np.sin(np.radians(ga))
    cx = lc * cb
    cy = lc * (ca - cb * cg) / sg
    return np.array([[la, lb * cg, cx], [0.0, lb * sg, cy], [0.0, 0.0, np.sqrt(lc * lc - cx * cx - cy * cy)]])


def reduced_box(box6):
    """[|a|, |b'|, |c'|, alpha', beta', gamma'] of the lattice of box6 in the
    basis a, b' = b - m a, c' = c - k b - l a with |c'_y| <= b_y / 2 and
    |c'_x|, |b'_x| <= a_x / 2 (whole m, k, l; rounding to nearest)."""
    h = cell_vectors(box6)
    a, b, c = h[:, 0].copy(), h[:, 1].copy(), h[:, 2].copy()
    k1 = np.rint(c[1] / b[1])
    c -= k1 * b
    k2 = np.rint(c[0] / a[0])
    c -= k2 * a
    k3 = np.rint(b[0] / a[0])
    b -= k3 * a
    if k1 == 0 and k2 == 0 and k3 == 0:
        return [float(v) for v in box6]  # reduced already
    return [float(np.linalg.norm(a)), float(np.linalg.norm(b)), float(np.linalg.norm(c)),
            _angle(b, c), _angle(a, c), _angle(a, b)]


def mapped(coords, edge, box6):
    """x' = H x / edge for coordinates of a cubic box of the given edge."""
    h = cell_vectors(box6)
    return (np.asarray(coords, dtype=np.float64).reshape(-1, 3) / edge) @ h.T


def doubled(top, shift):
    """A copy of a synthetic topology (md_amd.systems.Topology) with every
    atom and term duplicated; the second copy sits at coords + shift."""
    n = top.n_atoms
    t = copy.deepcopy(top)
    t.coords = np.vstack([top.coords, top.coords + np.asarray(shift, dtype=np.float64)])
    for name in ("charge_e", "atom_type", "mass", "atomic_number"):
        v = getattr(top, name)
        setattr(t, name, np.concatenate([v, v]))
    for name in ("names", "amber_types", "res_labels", "mol_sizes"):
        setattr(t, name, list(getattr(top, name)) * 2)
    t.res_starts = list(top.res_starts) + [s + n for s in top.res_starts]

    def shifted(entries, n_idx):
        return list(entries) + [tuple(int(v) + n for v in e[:n_idx]) + tuple(e[n_idx:]) for e in entries]

    t.bonds = shifted(top.bonds, 2)
    t.angles = shifted(top.angles, 3)
    t.ubs = shifted(top.ubs, 2)
    t.dihedrals = shifted(top.dihedrals, 4)
    t.impropers = shifted(top.impropers, 4)
    t.cmaps = shifted(top.cmaps, 5)
    t.exclusions = [list(e) for e in top.exclusions] + [[j + n for j in e] for e in top.exclusions]
    return t
