"""Amber restart files and the fixed-column prmtop reader (CPU only).

The reference reads an rst7 by splitting on blanks and taking every value
after the 3N coordinates as the box (AmberTopology.cpp:388-438); it has no
writer.  Checks:
  * write -> read round trips coordinates, velocities (A/ps), box and time to
    the file precision (F12.7);
  * fields that touch ("-123.4567890-234.5678901") are read from their columns;
  * each value-count layout (3N, 3N+3, 3N+6, 6N, 6N+6) is recognised and an
    inconsistent count is an error;
  * AMBER semantics read the box of a restart with velocities, while REF
    semantics reproduce the reference's reading (the first six velocities
    become the box);
  * the reference's own two rst7 files (data only, read where the reference
    tree is present) give the values of an independent blank split;
  * a prmtop whose CHARGE section is rewritten in touching 15-column fields
    loads in AMBER semantics with exactly the charges written.
"""
import os

import numpy as np
import pytest

from md_amd import native as N
from md_amd import systems

REF_FILES = {"step3_input.rst7": 26722, "CB1_apo_assym_fah_sub_100.rst7": 98308}


def _rand(n, seed=1):
    rng = np.random.default_rng(seed)
    return rng.uniform(-99.0, 999.0, (n, 3)), rng.normal(0.0, 5.0, (n, 3))


def test_round_trip(tmp_path):
    x, v = _rand(101)
    box = [41.5, 37.25, 52.0, 90.0, 90.0, 90.0]
    p = str(tmp_path / "a.rst7")
    N.write_rst7(p, x, box=box, velocities=v, title="round trip", time=12.5)
    r = N.read_rst7(p)
    assert r["n_atoms"] == 101 and r["title"] == "round trip" and r["time"] == 12.5
    assert np.abs(r["coords"] - x).max() <= 5.01e-8
    assert np.abs(r["velocities"] - v).max() <= 5.01e-8 * 20.455
    assert np.allclose(r["box"], box, atol=1e-12)
    N.write_rst7(p, x)
    r = N.read_rst7(p)
    assert r["velocities"] is None and r["box"] is None and r["time"] is None


def test_touching_fields(tmp_path):
    n = 4
    x = -np.arange(100.0, 100.0 + 3 * n).reshape(n, 3) - 0.4567890
    p = str(tmp_path / "t.rst7")
    with open(p, "w") as f:
        f.write("touching\n%6d\n" % n)
        flat = x.reshape(-1)
        for k in range(0, flat.size, 6):
            f.write("".join("%12.7f" % t for t in flat[k:k + 6]) + "\n")
        f.write("".join("%12.7f" % t for t in (30.0, 30.0, 30.0, 90.0, 90.0, 90.0)) + "\n")
    line = open(p).read().splitlines()[2]
    assert " " not in line.strip()  # the fields really touch
    r = N.read_rst7(p)
    assert np.array_equal(r["coords"], np.round(x, 7))
    assert np.allclose(r["box"], [30, 30, 30, 90, 90, 90])


@pytest.mark.parametrize("extra,vel,box", [(0, False, None), (3, False, 3), (6, False, 6), ("v", True, None),
                                            ("v6", True, 6)])
def test_layouts(tmp_path, extra, vel, box):
    n = 7
    x, v = _rand(n, 3)
    vals = list(x.reshape(-1))
    if vel:
        vals += list(v.reshape(-1) / 20.455)
    nb = 6 if extra in (6, "v6") else (3 if extra == 3 else 0)
    vals += [20.0, 21.0, 22.0, 90.0, 90.0, 90.0][:nb]
    p = str(tmp_path / "l.rst7")
    with open(p, "w") as f:
        f.write("layout\n%6d\n" % n)
        for k in range(0, len(vals), 6):
            f.write("".join("%12.7f" % t for t in vals[k:k + 6]) + "\n")
    r = N.read_rst7(p)
    assert (r["velocities"] is not None) == vel
    assert (r["box"] is not None) == (box is not None)
    if box:
        assert np.allclose(r["box"][:3], [20, 21, 22])
    with open(p, "a") as f:
        f.write("%12.7f\n" % 1.0)
    with pytest.raises(ValueError):
        N.read_rst7(p)


def test_velocity_box_semantics(tmp_path, polymer_2k):
    """AMBER reads the box after the velocities; REF keeps the reference's
    reading, in which the first six velocity values become the box."""
    prm, rst = polymer_2k
    top = N.Topology(prm, rst, N.SEM_AMBER)
    n = top.n_atoms
    v = np.random.default_rng(4).normal(0.0, 3.0, (n, 3))
    p = str(tmp_path / "v.rst7")
    b = float(top.box[0])
    N.write_rst7(p, top.coords, box=[b, b, b], velocities=v)
    ta = N.Topology(prm, p, N.SEM_AMBER)
    tr = N.Topology(prm, p, N.SEM_REF)
    assert np.allclose(ta.box, [b, b, b])
    vf = np.round(v.reshape(-1)[:3] / 20.455, 7)
    assert np.allclose(tr.box, vf, atol=1e-12)
    assert np.abs(ta.coords - top.coords).max() <= 5.01e-8


@pytest.mark.parametrize("name", sorted(REF_FILES))
def test_reference_rst7_files(name):
    path = os.path.join("/root/reference", name)
    if not os.path.exists(path):
        pytest.skip("the reference tree is not present here")
    r = N.read_rst7(path)
    assert r["n_atoms"] == REF_FILES[name]
    lines = open(path).read().splitlines()
    vals = np.array([float(t) for ln in lines[2:] for t in ln.split()])
    assert vals.size == 3 * r["n_atoms"] + 6
    assert np.array_equal(r["coords"].reshape(-1), vals[:-6])
    assert np.array_equal(r["box"], vals[-6:])
    assert r["velocities"] is None


def test_prmtop_touching_charges(tmp_path, polymer_2k):
    prm, rst = polymer_2k
    text = open(prm).read().splitlines()
    i = text.index("%FLAG CHARGE")
    j = i + 2
    while not text[j].startswith("%"):
        j += 1
    q = np.array([float(t) for ln in text[i + 2:j] for t in ln.split()])
    qn = -np.abs(q) - 1.0  # every value negative with one integer digit: 15 columns, no blank
    lines = ["".join("%15.8E" % t for t in qn[k:k + 5]) for k in range(0, qn.size, 5)]
    assert " " not in lines[0]
    p = str(tmp_path / "touch.prmtop")
    with open(p, "w") as f:
        f.write("\n".join(text[:i + 1] + ["%FORMAT(5E15.8)"] + lines + text[j:]) + "\n")
    t = N.Topology(p, rst, N.SEM_AMBER)
    # the file holds 9 significant digits; charges are divided by 18.2206899283247
    want = np.array([float("%15.8E" % v) for v in qn]) / 18.2206899283247
    assert np.allclose(t.arrays["charge"], want, rtol=1e-13, atol=0)
