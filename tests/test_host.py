"""Host-side checks (CPU): the C-ABI libraries load and export every entry
point their headers declare; the C++ reader's error behaviour."""
import os
import re

import numpy as np
import pytest

from md_amd import native as N

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(md(?:gpu|host)_\w+)\s*\(", txt, re.M)))


@pytest.mark.parametrize("header,lib", [("mdgpu.h", "libmdgpu.so"), ("mdhost.h", "libmdhost.so")])
def test_library_exports_every_declared_symbol(header, lib):
    import ctypes
    names = declared(header)
    assert len(names) >= 6
    so = ctypes.CDLL(os.path.join(ROOT, "md_amd", lib))
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_create_without_gpu_or_bad_device_fails_cleanly():
    import ctypes
    lib = N.gpu_lib()
    ctx = ctypes.c_void_p()
    assert lib.mdgpu_create(10 ** 6, ctypes.byref(ctx)) == 1
    assert not ctx.value


def test_missing_file_reports_error(tmp_path):
    with pytest.raises(RuntimeError, match="Error opening file"):
        N.Topology(str(tmp_path / "nope.prmtop"), None)


def test_bad_version_line(tmp_path):
    p = tmp_path / "bad.prmtop"
    p.write_text("%FLAG POINTERS\n")
    with pytest.raises(RuntimeError, match="VERSION"):
        N.Topology(str(p), None)


def test_rst7_without_box_disables_pbc(polymer_2k, tmp_path):
    prm, rst = polymer_2k
    lines = open(rst).read().splitlines()
    (tmp_path / "nobox.rst7").write_text("\n".join(lines[:-1]) + "\n")
    t = N.Topology(prm, str(tmp_path / "nobox.rst7"))
    assert np.all(t.box == 0.0)


def test_no_coordinates_gives_zeros(polymer_2k):
    prm, _ = polymer_2k
    t = N.Topology(prm, None)
    assert t.coords.shape == (t.n_atoms, 3) and not t.coords.any()


def test_amber_semantics_fills_exclusions(polymer_2k):
    prm, rst = polymer_2k
    ref = N.Topology(prm, rst, N.SEM_REF)
    amb = N.Topology(prm, rst, N.SEM_AMBER)
    assert ref.arrays["excl_start"] is None
    es, el = amb.arrays["excl_start"], amb.arrays["excl_list"]
    assert es[-1] == el.size > 0
    # symmetric: j in excl(i) <=> i in excl(j)
    pairs = {(i, int(j)) for i in range(amb.n_atoms) for j in el[es[i]:es[i + 1]]}
    assert all((j, i) in pairs for (i, j) in pairs)
    # complete lists: one more record per list than the reference keeps
    assert amb.arrays["bond_type"].size == ref.arrays["bond_type"].size + 2   # H and non-H lists
    assert amb.arrays["cmap_set"].size == ref.arrays["cmap_set"].size + 1
    assert amb.arrays["improper_type"].size == ref.arrays["improper_type"].size


def test_noalloc_query_rejects_bad_arguments_before_any_launch():
    """mdgpu_pairs_need_rebuild_noalloc validates on the host (no device
    access): sizes beyond MDGPU_NOALLOC_MAX_ATOMS, negative sizes, NULL
    outputs, bad boxes and skins give MDGPU_ERR_ARG."""
    import ctypes
    lib = N.gpu_lib()
    box = np.array([30.0, 30.0, 30.0])
    bp = box.ctypes.data_as(N.P_F64)
    fake = ctypes.c_void_p(0x1000)   # never dereferenced: the checks come first
    f = lib.mdgpu_pairs_need_rebuild_noalloc
    assert f(0, fake, fake, N.NOALLOC_MAX_ATOMS + 1, bp, 1.0, None, fake, None) == 1
    assert f(0, fake, fake, -1, bp, 1.0, None, fake, None) == 1
    assert f(0, fake, fake, 10, bp, 1.0, None, None, None) == 1
    assert f(0, fake, fake, 10, bp, -1.0, None, fake, None) == 1
    bad = np.array([30.0, 0.0, 30.0])
    assert f(0, fake, fake, 10, bad.ctypes.data_as(N.P_F64), 1.0, None, fake, None) == 1
    assert f(10 ** 6, fake, fake, 10, bp, 1.0, None, fake, None) == 1


def test_shape_entry_points_reject_a_null_context():
    """mdgpu_tune / mdgpu_set_shapes / mdgpu_get_shapes check their context
    before any device work: a null one gives MDGPU_ERR_ARG."""
    import ctypes
    lib = N.gpu_lib()
    box = np.array([30.0, 30.0, 30.0])
    bp = box.ctypes.data_as(N.P_F64)
    fake = ctypes.c_void_p(0x1000)   # never dereferenced: the context check comes first
    sp, fl, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    assert lib.mdgpu_tune(None, fake, bp, None, 3, ctypes.byref(sp), ctypes.byref(fl), ctypes.byref(ms)) == 1
    assert lib.mdgpu_set_shapes(None, 2, 8) == 1
    assert lib.mdgpu_get_shapes(None, ctypes.byref(sp), ctypes.byref(fl)) == 1
