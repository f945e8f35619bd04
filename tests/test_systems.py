"""Generators (CPU): the vectorised water-box arrays used by the batched
workload describe the same system as the prmtop written by ``water_box`` and
read back by the C++ reader (AMBER semantics)."""
import math

import numpy as np

from md_amd import native as N
from md_amd import systems


def test_water_arrays_match_the_prmtop_reader(tmp_path):
    n_water = 60
    top = systems.water_box(seed=9, box=14.0, n_water=n_water, n_lattice=4, jitter=0.2)
    p, r = systems.write_system(top, str(tmp_path / "w"))
    ref = N.Topology(p, r, N.SEM_AMBER).arrays
    arr, coords, box = systems.water_arrays(seed=9, n_atoms=3 * n_water)
    for k in ["type", "nb_index", "bond_atoms", "bond_type", "angle_atoms", "angle_type", "excl_start",
              "excl_list", "excl_pairs", "pairs14"]:
        np.testing.assert_array_equal(np.asarray(arr[k]), np.asarray(ref[k]), err_msg=k)
    for k in ["charge", "lj_a", "lj_b", "lj14_a", "lj14_b", "bond_k", "bond_r0", "angle_k", "angle_theta0"]:
        np.testing.assert_allclose(arr[k], ref[k], rtol=1e-12, atol=0, err_msg=k)
    assert arr["n_types"] == ref["n_types"]
    # geometry: rigid waters, box at 0.1 atoms/A^3
    x = coords.reshape(-1, 3, 3)
    d1 = np.linalg.norm(x[:, 1] - x[:, 0], axis=1)
    d2 = np.linalg.norm(x[:, 2] - x[:, 0], axis=1)
    cosang = np.einsum("ij,ij->i", x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]) / (d1 * d2)
    np.testing.assert_allclose(d1, 0.9572, atol=1e-12)
    np.testing.assert_allclose(d2, 0.9572, atol=1e-12)
    np.testing.assert_allclose(np.degrees(np.arccos(cosang)), 104.52, atol=1e-9)
    assert abs(box - (3 * n_water / 0.1) ** (1 / 3)) < 1e-12


def test_batch_workload_sizes():
    w = systems.batch_workload(16)
    for arrays, coords, box, k in w:
        assert 3000 - 2 <= coords.shape[0] <= 30000
        assert k >= box and k == systems.smallest_2a3b(int(math.ceil(box)))
        assert all(q in (2, 3) for q in _factors(k))


def _factors(n):
    out = []
    for p in (2, 3):
        while n % p == 0:
            out.append(p)
            n //= p
    if n > 1:
        out.append(n)
    return out


def test_with_hydrogen_lj(sysdir):
    """The hydrogen type gets Lorentz-Berthelot entries with every type; the
    other entries are untouched."""
    from md_amd import native as N
    prm, rst = systems.cached_system("water_box_6k", sysdir)
    top = N.Topology(prm, rst, N.SEM_AMBER)
    a = top.arrays
    h = int(a["type"][1])
    b = systems.with_hydrogen_lj(a, h, sigma=0.4, eps=0.046)
    t = int(a["n_types"])
    nb = np.asarray(a["nb_index"]).reshape(t, t)
    o = int(a["type"][0])
    so, eo = systems.O_SIGMA, systems.O_EPS
    s, e = 0.5 * (0.4 + so), math.sqrt(0.046 * eo)
    assert math.isclose(b["lj_a"][nb[h, o] - 1], 4 * e * s ** 12, rel_tol=1e-12)
    assert math.isclose(b["lj_b"][nb[h, o] - 1], 4 * e * s ** 6, rel_tol=1e-12)
    assert math.isclose(b["lj_a"][nb[h, h] - 1], 4 * 0.046 * 0.4 ** 12, rel_tol=1e-12)
    assert b["lj_a"][nb[o, o] - 1] == a["lj_a"][nb[o, o] - 1]
