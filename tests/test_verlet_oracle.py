"""Pinning of the oracle's Verlet rebuild criterion (CPU only).

oracle_pairs_need_rebuild restates System::ljpairs_needs_rebuild
(NonBondedUtils.cpp:87-100) with the limit of
setup_neighbor_rebuild_threshold_lj (:65-69).  The checks fix what the
mathematics fixes, so that a plausible slip fails one of them:
  * the limit is (skin/2)^2 and the test is strict (a displacement of exactly
    skin/2 keeps the list, the next double above it does not), on each axis
    (a dropped or repeated component fails);
  * all three components enter the squared displacement;
  * the displacement is taken at its minimum image (a move across a face);
  * every atom is examined, the last one included;
  * a NaN displacement never becomes the maximum;
  * random displacement sets agree with an exact-rational brute force.
"""
import math
import random
from fractions import Fraction

import numpy as np

import oracle

BOX = np.array([40.0, 37.5, 52.25])


def _ref(n=5, seed=1):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 1.0, (n, 3)) * BOX


def test_strict_limit_on_each_axis():
    x0 = _ref()
    for ax in range(3):
        x = x0.copy()
        x[2, ax] += 1.0  # exactly skin / 2 for skin 2; (1.0)^2 == 1.0 exactly
        assert not oracle.pairs_need_rebuild(x, x0, BOX, 2.0)
        x = x0.copy()
        x[2, ax] = x0[2, ax] + math.nextafter(1.0, 2.0)
        # x0 + d rounds; the stored displacement is what the criterion sees
        d = x[2, ax] - x0[2, ax]
        assert oracle.pairs_need_rebuild(x, x0, BOX, 2.0) == (d * d > 1.0)
        x = x0.copy()
        x[2, ax] += 1.0 + 1e-9
        assert oracle.pairs_need_rebuild(x, x0, BOX, 2.0)


def test_limit_is_half_skin_squared():
    x0 = _ref()
    x = x0.copy()
    x[1, 0] += 1.4  # 1.96 < 2.25 = (3/2)^2: a limit of skin^2 or skin/2 on d^2 would differ
    assert not oracle.pairs_need_rebuild(x, x0, BOX, 3.0)
    x[1, 0] += 0.2  # 1.6^2 = 2.56 > 2.25
    assert oracle.pairs_need_rebuild(x, x0, BOX, 3.0)


def test_all_components_count():
    x0 = _ref()
    x = x0.copy()
    x[3] += np.array([0.6, 0.6, 0.6])  # 1.08 > 1; any two components give 0.72
    assert oracle.pairs_need_rebuild(x, x0, BOX, 2.0)
    x = x0.copy()
    x[3] += np.array([0.5, 0.5, 0.5])  # 0.75 < 1
    assert not oracle.pairs_need_rebuild(x, x0, BOX, 2.0)


def test_minimum_image_across_a_face():
    x0 = _ref()
    x0[0] = [0.1, 0.2, 0.3]
    x = x0.copy()
    x[0] = [BOX[0] - 0.3, BOX[1] - 0.1, 0.3]  # moved by (-0.4, -0.3, 0) through the faces
    assert not oracle.pairs_need_rebuild(x, x0, BOX, 2.0)
    x[0, 0] = 0.5 * BOX[0]  # half a box away
    assert oracle.pairs_need_rebuild(x, x0, BOX, 2.0)
    # a whole box vector is no displacement
    x = x0 + BOX * np.array([1.0, -2.0, 3.0])
    assert not oracle.pairs_need_rebuild(x, x0, BOX, 0.5)


def test_last_atom_is_examined():
    x0 = _ref(n=100)
    x = x0.copy()
    x[-1, 2] += 1.5
    assert oracle.pairs_need_rebuild(x, x0, BOX, 2.0)
    x = x0.copy()
    x[0, 2] += 1.5
    assert oracle.pairs_need_rebuild(x, x0, BOX, 2.0)


def test_nan_never_becomes_the_maximum():
    x0 = _ref()
    x = x0.copy()
    x[0, 0] = np.nan
    x[1, 1] += 0.5
    assert not oracle.pairs_need_rebuild(x, x0, BOX, 2.0)
    x[1, 1] += 1.0
    assert oracle.pairs_need_rebuild(x, x0, BOX, 2.0)


def _exact_decision(x, x0, box, skin):
    lim = (Fraction(skin) / 2) ** 2
    best = Fraction(0)
    for a, b in zip(x, x0):
        s = Fraction(0)
        for d in range(3):
            v, L = Fraction(float(a[d])) - Fraction(float(b[d])), Fraction(float(box[d]))
            q = v / L
            k = math.floor(q + Fraction(1, 2))  # nearest image (no ties here)
            s += (v - k * L) ** 2
        best = max(best, s)
    return best > lim, float(best), float(lim)


def test_random_sets_match_exact_arithmetic():
    rng = random.Random(7)
    checked = 0
    for trial in range(300):
        n = rng.randint(1, 40)
        x0 = _ref(n=n, seed=trial)
        skin = rng.choice([0.5, 1.0, 2.0, 3.0])
        scale = rng.choice([0.05, 0.2, 0.6, 1.5])
        x = x0 + np.array([[rng.gauss(0.0, scale) for _ in range(3)] for _ in range(n)])
        # some atoms jump by whole boxes (re-wrapped coordinates)
        for i in range(n):
            if rng.random() < 0.2:
                x[i] += BOX * np.array([rng.randint(-1, 1) for _ in range(3)])
        want, best, lim = _exact_decision(x, x0, BOX, skin)
        if abs(best - lim) <= 1e-9 * lim:
            continue  # rounding may decide: not a pinning case
        assert oracle.pairs_need_rebuild(x, x0, BOX, skin) == want, (trial, best, lim)
        checked += 1
    assert checked > 250
