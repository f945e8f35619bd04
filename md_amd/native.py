"""ctypes binding of the two native libraries built in-tree.

* ``libmdhost.so`` - the C++ prmtop/rst7 reader (reference semantics) exposed
  through include/mdhost.h;
* ``libmdgpu.so``  - the HIP kernels behind the mdgpu C ABI (include/mdgpu.h).

No computation happens in Python.  A missing library raises ImportError: there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

SEM_REF, SEM_AMBER = 0, 1
PREC_FP64, PREC_MIXED = 0, 1

ERRORS = {1: "invalid argument", 2: "HIP runtime error", 3: "call out of order",
          4: "fixed-point force overflow", 5: "unsupported parameters"}

KERNELS = ["sort", "pairlist", "nonbonded", "bonded", "corrections", "spread", "fft", "conv",
           "gather", "reduce"]

P_I32 = C.POINTER(C.c_int32)
P_F64 = C.POINTER(C.c_double)


class SystemDesc(C.Structure):
    _fields_ = [
        ("n_atoms", C.c_int32), ("n_types", C.c_int32),
        ("charge", P_F64), ("type", P_I32), ("nb_index", P_I32),
        ("n_lj", C.c_int32), ("lj_a", P_F64), ("lj_b", P_F64), ("lj14_a", P_F64), ("lj14_b", P_F64),
        ("n_bonds", C.c_int32), ("bond_atoms", P_I32), ("bond_type", P_I32), ("n_bond_types", C.c_int32),
        ("bond_k", P_F64), ("bond_r0", P_F64),
        ("n_ubs", C.c_int32), ("ub_atoms", P_I32), ("ub_type", P_I32), ("n_ub_types", C.c_int32),
        ("ub_k", P_F64), ("ub_r0", P_F64),
        ("n_angles", C.c_int32), ("angle_atoms", P_I32), ("angle_type", P_I32), ("n_angle_types", C.c_int32),
        ("angle_k", P_F64), ("angle_theta0", P_F64),
        ("n_dihedrals", C.c_int32), ("dihedral_atoms", P_I32), ("dihedral_type", P_I32),
        ("n_dihedral_types", C.c_int32), ("dihedral_k", P_F64), ("dihedral_n", P_F64),
        ("dihedral_phase", P_F64),
        ("n_impropers", C.c_int32), ("improper_atoms", P_I32), ("improper_type", P_I32),
        ("n_improper_types", C.c_int32), ("improper_k", P_F64), ("improper_psi0", P_F64),
        ("n_cmaps", C.c_int32), ("cmap_atoms", P_I32), ("cmap_set", P_I32), ("cmap_n_sets", C.c_int32),
        ("cmap_resolution", C.c_int32), ("cmap_coeff", P_F64), ("cmap_grid", P_F64),
        ("excl_start", P_I32), ("excl_list", P_I32),
        ("n_excl_pairs", C.c_int32), ("excl_pairs", P_I32),
        ("n_pairs14", C.c_int32), ("pairs14", P_I32), ("scee14", P_F64),
    ]


class Params(C.Structure):
    _fields_ = [
        ("semantics", C.c_int32), ("precision", C.c_int32),
        ("lj_cutoff", C.c_double), ("ee_cutoff", C.c_double), ("lj_switch", C.c_double),
        ("ewald_alpha", C.c_double), ("pme_grid", C.c_int32 * 3), ("pme_order", C.c_int32),
    ]


ENERGY_TERMS = ["bond", "angle", "dihedral", "urey_bradley", "improper", "cmap", "vdw", "ee_all",
                "ee_pairlist", "ee_rf", "ewald_direct", "ewald_self", "ewald_recip", "ewald_excl",
                "vdw14", "ee14", "total"]


class Energies(C.Structure):
    _fields_ = [(n, C.c_double) for n in ENERGY_TERMS]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n in ENERGY_TERMS}


# parameter-gradient blocks of mdgpu_param_grads and the desc count giving each length
GRAD_FIELDS = [("charge", "n_atoms"), ("lj_a", "n_lj"), ("lj_b", "n_lj"), ("lj14_a", "n_lj"),
               ("lj14_b", "n_lj"), ("bond_k", "n_bond_types"), ("bond_r0", "n_bond_types"),
               ("ub_k", "n_ub_types"), ("ub_r0", "n_ub_types"), ("angle_k", "n_angle_types"),
               ("angle_theta0", "n_angle_types"), ("dihedral_k", "n_dihedral_types"),
               ("dihedral_phase", "n_dihedral_types"), ("improper_k", "n_improper_types"),
               ("improper_psi0", "n_improper_types"), ("cmap_coeff", "cmap_coeff")]


class ParamGrads(C.Structure):
    _fields_ = [(name, P_F64) for name, _ in GRAD_FIELDS]


def grad_lengths(desc: SystemDesc) -> dict:
    out = {}
    for name, cnt in GRAD_FIELDS:
        if cnt == "cmap_coeff":
            out[name] = desc.cmap_n_sets * desc.cmap_resolution ** 2 * 16 if desc.n_cmaps > 0 else 0
        else:
            out[name] = int(getattr(desc, cnt))
    return out


def alloc_grads(desc: SystemDesc):
    """numpy arrays for every gradient block and a ParamGrads pointing at them."""
    arrs = {k: np.zeros(max(n, 0)) for k, n in grad_lengths(desc).items()}
    g = ParamGrads()
    for k, a in arrs.items():
        setattr(g, k, a.ctypes.data_as(P_F64) if a.size else None)
    return arrs, g


def energies_struct(d: dict) -> Energies:
    e = Energies()
    for k in ENERGY_TERMS:
        setattr(e, k, float(d.get(k, 0.0)))
    return e


def _load(name: str) -> C.CDLL:
    if name == "libmdhost.so" and os.environ.get("MD_AMD_SANITIZE") == "1":
        name = "libmdhost_asan.so"   # the CPU sanitizer build (make -C md_amd/csrc sanitize)
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built; run __graft_entry__.build() (make -C md_amd/csrc)")
    return C.CDLL(path)


_gpu = None
_host = None


def gpu_lib() -> C.CDLL:
    global _gpu
    if _gpu is None:
        lib = _load("libmdgpu.so")
        lib.mdgpu_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
        lib.mdgpu_destroy.argtypes = [C.c_void_p]
        lib.mdgpu_last_error.argtypes = [C.c_void_p]
        lib.mdgpu_last_error.restype = C.c_char_p
        lib.mdgpu_upload_system.argtypes = [C.c_void_p, C.POINTER(SystemDesc)]
        lib.mdgpu_set_params.argtypes = [C.c_void_p, C.POINTER(Params)]
        lib.mdgpu_eval.argtypes = [C.c_void_p, P_F64, P_F64, C.POINTER(Energies), P_F64, P_F64]
        lib.mdgpu_eval_device.argtypes = [C.c_void_p, C.c_void_p, P_F64, C.POINTER(Energies), C.c_void_p,
                                          P_F64]
        lib.mdgpu_set_timing.argtypes = [C.c_void_p, C.c_int]
        lib.mdgpu_get_timing.argtypes = [C.c_void_p, C.c_int, P_F64, C.POINTER(C.c_int64)]
        lib.mdgpu_get_counts.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        lib.mdgpu_set_counting.argtypes = [C.c_void_p, C.c_int]
        lib.mdgpu_set_graph.argtypes = [C.c_void_p, C.c_int]
        lib.mdgpu_graph_state.argtypes = [C.c_void_p]
        lib.mdgpu_set_shapes.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.mdgpu_get_shapes.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        lib.mdgpu_tune.argtypes = [C.c_void_p, C.c_void_p, P_F64, C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                   C.POINTER(C.c_int), P_F64]
        lib.mdgpu_set_verlet.argtypes = [C.c_void_p, C.c_double]
        lib.mdgpu_set_box_angles.argtypes = [C.c_void_p, P_F64]
        lib.mdgpu_get_box_angles.argtypes = [C.c_void_p, P_F64]
        lib.mdgpu_get_search_mode.argtypes = [C.c_void_p]
        lib.mdgpu_set_virial.argtypes = [C.c_void_p, C.c_int]
        lib.mdgpu_get_virial.argtypes = [C.c_void_p, P_F64]
        lib.mdgpu_md_set_masses.argtypes = [C.c_void_p, P_F64]
        lib.mdgpu_minimize.argtypes = [C.c_void_p, C.c_void_p, P_F64, C.POINTER(MinOpts), C.POINTER(MinResult)]
        lib.mdgpu_md_run.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, P_F64, C.c_double, C.c_int32,
                                     C.c_void_p, C.POINTER(Energies)]
        lib.mdgpu_get_list_builds.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        lib.mdgpu_get_list_prunes.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        lib.mdgpu_pairs_need_rebuild.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
        lib.mdgpu_pairs_need_rebuild_noalloc.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int32, P_F64,
                                                         C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.mdgpu_adjoint.argtypes = [C.c_void_p, P_F64, P_F64, C.POINTER(Energies), C.POINTER(ParamGrads)]
        lib.mdgpu_batch_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
        lib.mdgpu_batch_destroy.argtypes = [C.c_void_p]
        lib.mdgpu_batch_last_error.argtypes = [C.c_void_p]
        lib.mdgpu_batch_last_error.restype = C.c_char_p
        lib.mdgpu_batch_add.argtypes = [C.c_void_p, C.POINTER(SystemDesc), C.POINTER(Params), P_F64,
                                        C.POINTER(C.c_int)]
        lib.mdgpu_batch_eval_device.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(Energies), C.c_void_p]
        lib.mdgpu_batch_eval.argtypes = [C.c_void_p, P_F64, C.POINTER(Energies), P_F64]
        lib.mdgpu_batch_configure.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        lib.mdgpu_batch_tune.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                         C.POINTER(C.c_int), C.POINTER(C.c_int), P_F64]
        lib.mdgpu_batch_system_config.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        lib.mdgpu_batch_mode.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        lib.mdgpu_batch_size.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int64)]
        lib.mdgpu_batch_graph_state.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        _gpu = lib
    return _gpu


def host_lib() -> C.CDLL:
    global _host
    if _host is None:
        gpu_lib()  # libmdhost.so links it
        lib = _load("libmdhost.so")
        lib.mdhost_load.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.POINTER(C.c_void_p)]
        lib.mdhost_desc.argtypes = [C.c_void_p, C.POINTER(SystemDesc)]
        lib.mdhost_coords.argtypes = [C.c_void_p]
        lib.mdhost_coords.restype = P_F64
        lib.mdhost_masses.argtypes = [C.c_void_p]
        lib.mdhost_masses.restype = P_F64
        lib.mdhost_box.argtypes = [C.c_void_p, P_F64]
        lib.mdhost_dihedral_flags.argtypes = [C.c_void_p]
        lib.mdhost_dihedral_flags.restype = P_I32
        lib.mdhost_free.argtypes = [C.c_void_p]
        lib.mdhost_last_error.restype = C.c_char_p
        lib.mdhost_cmap_coefficients.argtypes = [C.c_int, C.c_int, P_F64, P_F64]
        lib.mdhost_cmap_grid_grad.argtypes = [C.c_int, C.c_int, P_F64, P_F64]
        lib.mdhost_rst7_read.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
        lib.mdhost_rst7_info.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int32), P_F64]
        lib.mdhost_rst7_coords.argtypes = [C.c_void_p]
        lib.mdhost_rst7_coords.restype = P_F64
        lib.mdhost_rst7_velocities.argtypes = [C.c_void_p]
        lib.mdhost_rst7_velocities.restype = P_F64
        lib.mdhost_rst7_box.argtypes = [C.c_void_p, P_F64]
        lib.mdhost_rst7_title.argtypes = [C.c_void_p]
        lib.mdhost_rst7_title.restype = C.c_char_p
        lib.mdhost_rst7_free.argtypes = [C.c_void_p]
        lib.mdhost_rst7_write.argtypes = [C.c_char_p, C.c_char_p, C.c_int32, P_F64, P_F64, P_F64, C.c_double,
                                          C.c_int32]
        _host = lib
    return _host


def cmap_coefficients(grid, res: int, n_sets: int) -> np.ndarray:
    """Patch coefficients [n_sets*res*res*16] of CMAP grids [n_sets*res*res]."""
    g = np.ascontiguousarray(grid, dtype=np.float64).reshape(-1)
    out = np.zeros(n_sets * res * res * 16)
    if n_sets:
        st = host_lib().mdhost_cmap_coefficients(res, n_sets, g.ctypes.data_as(P_F64), out.ctypes.data_as(P_F64))
        if st != 0:
            raise ValueError("mdhost_cmap_coefficients: bad arguments")
    return out


def cmap_grid_grad(d_coeff, res: int, n_sets: int) -> np.ndarray:
    """Gradient with respect to the grid values from that of the coefficients."""
    dc = np.ascontiguousarray(d_coeff, dtype=np.float64).reshape(-1)
    out = np.zeros(n_sets * res * res)
    if n_sets:
        st = host_lib().mdhost_cmap_grid_grad(res, n_sets, dc.ctypes.data_as(P_F64), out.ctypes.data_as(P_F64))
        if st != 0:
            raise ValueError("mdhost_cmap_grid_grad: bad arguments")
    return out


def read_rst7(path: str) -> dict:
    """An Amber restart: title, n_atoms, time (or None), coords (N, 3), velocities
    (N, 3) in A/ps or None, box (6,) lengths and angles or None."""
    lib = host_lib()
    h = C.c_void_p()
    if lib.mdhost_rst7_read(path.encode(), C.byref(h)) != 0:
        raise ValueError(f"mdhost_rst7_read: {lib.mdhost_last_error().decode()}")
    try:
        n, hv, hb = C.c_int32(), C.c_int32(), C.c_int32()
        t = C.c_double()
        lib.mdhost_rst7_info(h, C.byref(n), C.byref(hv), C.byref(hb), C.byref(t))
        box = np.zeros(6)
        lib.mdhost_rst7_box(h, box.ctypes.data_as(P_F64))
        out = {"title": lib.mdhost_rst7_title(h).decode(), "n_atoms": n.value,
               "coords": _arr(lib.mdhost_rst7_coords(h), 3 * n.value, np.float64).reshape(-1, 3),
               "velocities": (_arr(lib.mdhost_rst7_velocities(h), 3 * n.value, np.float64).reshape(-1, 3)
                              if hv.value else None),
               "box": box if hb.value else None}
        out["time"] = None if np.isnan(t.value) else t.value
        return out
    finally:
        lib.mdhost_rst7_free(h)


def write_rst7(path: str, coords, box=None, velocities=None, title: str = "", time=None):
    """Write an Amber restart (velocities in A/ps; box: 3 lengths or lengths and angles)."""
    x = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1)
    v = None if velocities is None else np.ascontiguousarray(velocities, dtype=np.float64).reshape(-1)
    b = None
    if box is not None:
        b = np.asarray(box, dtype=np.float64).reshape(-1)
        b = np.ascontiguousarray(np.concatenate([b, [90.0, 90.0, 90.0]])[:6] if b.size == 3 else b[:6])
    lib = host_lib()
    st = lib.mdhost_rst7_write(path.encode(), title.encode(), x.size // 3, x.ctypes.data_as(P_F64),
                               v.ctypes.data_as(P_F64) if v is not None else None,
                               b.ctypes.data_as(P_F64) if b is not None else None,
                               float(time or 0.0), 1 if time is not None else 0)
    if st != 0:
        raise ValueError(f"mdhost_rst7_write: {lib.mdhost_last_error().decode()}")


def _arr(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Topology:
    """A parsed prmtop/rst7: host handle plus numpy copies of the arrays.

    ``semantics`` is SEM_REF (the reference's term lists: last entries dropped,
    no exclusions) or SEM_AMBER (complete lists and exclusions).
    """

    def __init__(self, prmtop: str, rst7: str | None = None, semantics: int = SEM_REF):
        lib = host_lib()
        h = C.c_void_p()
        st = lib.mdhost_load(prmtop.encode(), (rst7 or "").encode(), semantics, C.byref(h))
        if st != 0:
            raise RuntimeError(f"mdhost_load: {lib.mdhost_last_error().decode()}")
        self._h = h
        self.semantics = semantics
        d = SystemDesc()
        lib.mdhost_desc(h, C.byref(d))
        self._desc = d
        n = d.n_atoms
        self.n_atoms = n
        self.coords = _arr(lib.mdhost_coords(h), 3 * n, np.float64).reshape(n, 3)
        self.masses = _arr(lib.mdhost_masses(h), n, np.float64)
        box = (C.c_double * 3)()
        lib.mdhost_box(h, box)
        self.box = np.array(box[:], dtype=np.float64)
        self.dihedral_flags = _arr(lib.mdhost_dihedral_flags(h), d.n_dihedrals, np.int32)
        self.arrays = self._copy_arrays(d)

    @staticmethod
    def _copy_arrays(d: SystemDesc) -> dict:
        n, t = d.n_atoms, d.n_types
        a = {
            "charge": _arr(d.charge, n, np.float64), "type": _arr(d.type, n, np.int32),
            "nb_index": _arr(d.nb_index, t * t, np.int32),
            "lj_a": _arr(d.lj_a, d.n_lj, np.float64), "lj_b": _arr(d.lj_b, d.n_lj, np.float64),
            "lj14_a": _arr(d.lj14_a, d.n_lj, np.float64), "lj14_b": _arr(d.lj14_b, d.n_lj, np.float64),
            "bond_atoms": _arr(d.bond_atoms, 2 * d.n_bonds, np.int32),
            "bond_type": _arr(d.bond_type, d.n_bonds, np.int32),
            "bond_k": _arr(d.bond_k, d.n_bond_types, np.float64),
            "bond_r0": _arr(d.bond_r0, d.n_bond_types, np.float64),
            "ub_atoms": _arr(d.ub_atoms, 2 * d.n_ubs, np.int32), "ub_type": _arr(d.ub_type, d.n_ubs, np.int32),
            "ub_k": _arr(d.ub_k, d.n_ub_types, np.float64), "ub_r0": _arr(d.ub_r0, d.n_ub_types, np.float64),
            "angle_atoms": _arr(d.angle_atoms, 3 * d.n_angles, np.int32),
            "angle_type": _arr(d.angle_type, d.n_angles, np.int32),
            "angle_k": _arr(d.angle_k, d.n_angle_types, np.float64),
            "angle_theta0": _arr(d.angle_theta0, d.n_angle_types, np.float64),
            "dihedral_atoms": _arr(d.dihedral_atoms, 4 * d.n_dihedrals, np.int32),
            "dihedral_type": _arr(d.dihedral_type, d.n_dihedrals, np.int32),
            "dihedral_k": _arr(d.dihedral_k, d.n_dihedral_types, np.float64),
            "dihedral_n": _arr(d.dihedral_n, d.n_dihedral_types, np.float64),
            "dihedral_phase": _arr(d.dihedral_phase, d.n_dihedral_types, np.float64),
            "improper_atoms": _arr(d.improper_atoms, 4 * d.n_impropers, np.int32),
            "improper_type": _arr(d.improper_type, d.n_impropers, np.int32),
            "improper_k": _arr(d.improper_k, d.n_improper_types, np.float64),
            "improper_psi0": _arr(d.improper_psi0, d.n_improper_types, np.float64),
            "cmap_atoms": _arr(d.cmap_atoms, 5 * d.n_cmaps, np.int32),
            "cmap_set": _arr(d.cmap_set, d.n_cmaps, np.int32),
            "cmap_coeff": _arr(d.cmap_coeff, d.cmap_n_sets * d.cmap_resolution ** 2 * 16, np.float64),
            "cmap_grid": _arr(d.cmap_grid, d.cmap_n_sets * d.cmap_resolution ** 2, np.float64),
            "excl_start": _arr(d.excl_start, n + 1, np.int32) if d.excl_start else None,
            "excl_pairs": _arr(d.excl_pairs, 2 * d.n_excl_pairs, np.int32),
            "pairs14": _arr(d.pairs14, 2 * d.n_pairs14, np.int32),
            "scee14": _arr(d.scee14, d.n_pairs14, np.float64),
        }
        a["excl_list"] = (_arr(d.excl_list, int(a["excl_start"][-1]), np.int32)
                          if a["excl_start"] is not None else None)
        a["n_types"], a["cmap_n_sets"], a["cmap_resolution"] = t, d.cmap_n_sets, d.cmap_resolution
        return a

    @property
    def desc(self) -> SystemDesc:
        """The C description (pointers into the host handle)."""
        return self._desc

    def close(self):
        if getattr(self, "_h", None):
            host_lib().mdhost_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def desc_from_arrays(arr: dict, keep: list) -> SystemDesc:
    """Build a SystemDesc over numpy arrays (kept alive through ``keep``)."""

    def p(name, ctype):
        v = arr.get(name)
        if v is None:
            return None
        v = np.ascontiguousarray(v, dtype=np.int32 if ctype is C.c_int32 else np.float64)
        keep.append(v)
        return v.ctypes.data_as(C.POINTER(ctype))

    d = SystemDesc()
    n = int(arr["charge"].size)
    d.n_atoms, d.n_types = n, int(arr["n_types"])
    d.charge, d.type, d.nb_index = p("charge", C.c_double), p("type", C.c_int32), p("nb_index", C.c_int32)
    d.n_lj = int(arr["lj_a"].size)
    d.lj_a, d.lj_b = p("lj_a", C.c_double), p("lj_b", C.c_double)
    d.lj14_a, d.lj14_b = p("lj14_a", C.c_double), p("lj14_b", C.c_double)
    for name, width, ntab in [("bond", 2, ["bond_k", "bond_r0"]), ("ub", 2, ["ub_k", "ub_r0"]),
                              ("angle", 3, ["angle_k", "angle_theta0"]),
                              ("dihedral", 4, ["dihedral_k", "dihedral_n", "dihedral_phase"]),
                              ("improper", 4, ["improper_k", "improper_psi0"])]:
        cnt = int(arr[f"{name}_type"].size)
        setattr(d, f"n_{name}s", cnt)
        setattr(d, f"{name}_atoms", p(f"{name}_atoms", C.c_int32))
        setattr(d, f"{name}_type", p(f"{name}_type", C.c_int32))
        setattr(d, f"n_{name}_types", int(arr[ntab[0]].size))
        for tname in ntab:
            setattr(d, tname, p(tname, C.c_double))
    d.n_cmaps = int(arr["cmap_set"].size)
    d.cmap_atoms, d.cmap_set = p("cmap_atoms", C.c_int32), p("cmap_set", C.c_int32)
    d.cmap_n_sets, d.cmap_resolution = int(arr["cmap_n_sets"]), int(arr["cmap_resolution"])
    d.cmap_coeff, d.cmap_grid = p("cmap_coeff", C.c_double), p("cmap_grid", C.c_double)
    d.excl_start, d.excl_list = p("excl_start", C.c_int32), p("excl_list", C.c_int32)
    d.n_excl_pairs = int(arr["excl_pairs"].size // 2)
    d.excl_pairs = p("excl_pairs", C.c_int32)
    d.n_pairs14 = int(arr["scee14"].size)
    d.pairs14, d.scee14 = p("pairs14", C.c_int32), p("scee14", C.c_double)
    return d


def make_params(semantics=SEM_REF, precision=PREC_FP64, lj_cutoff=10.0, ee_cutoff=None,
                lj_switch=9.0, ewald_alpha=None, ewald_tol=1e-6, pme_grid=(0, 0, 0),
                pme_order=4) -> Params:
    """Parameters; alpha defaults to sqrt(-ln(2 tol))/rc_ee (PBC.cpp:88-91)."""
    ee = lj_cutoff if ee_cutoff is None else ee_cutoff
    if ewald_alpha is None:
        ewald_alpha = float(np.sqrt(-np.log(2.0 * ewald_tol)) / ee)
    p = Params()
    p.semantics, p.precision = semantics, precision
    p.lj_cutoff, p.ee_cutoff, p.lj_switch, p.ewald_alpha = lj_cutoff, ee, lj_switch, ewald_alpha
    p.pme_grid[0], p.pme_grid[1], p.pme_grid[2] = pme_grid
    p.pme_order = pme_order
    return p


class MinOpts(C.Structure):
    _fields_ = [("max_iter", C.c_int32), ("rms_force_tol", C.c_double), ("max_step", C.c_double),
                ("history", C.c_int32)]


class MinResult(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("evaluations", C.c_int32), ("converged", C.c_int32),
                ("energy", C.c_double), ("rms_force", C.c_double), ("max_force", C.c_double)]


class GpuError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{ERRORS.get(status, status)}: {msg}")
        self.status = status


NOALLOC_MAX_ATOMS = 1048576  # MDGPU_NOALLOC_MAX_ATOMS


def pairs_need_rebuild_noalloc(device: int, x_ptr: int, xref_ptr: int, n_atoms: int, box, skin: float,
                               stream_ptr: int, needs_ptr: int, max_d2_ptr: int = 0) -> None:
    """mdgpu_pairs_need_rebuild_noalloc: enqueue the rebuild test of device
    coordinates x against xref (3 n_atoms doubles each) on the stream; the
    int32 decision and (if max_d2_ptr) the largest squared displacement are
    written to device memory.  Allocates nothing; no context."""
    b = np.ascontiguousarray(box, dtype=np.float64).reshape(-1)[:3]
    st = gpu_lib().mdgpu_pairs_need_rebuild_noalloc(int(device), C.c_void_p(x_ptr), C.c_void_p(xref_ptr),
                                                   int(n_atoms), b.ctypes.data_as(P_F64), float(skin),
                                                   C.c_void_p(stream_ptr) if stream_ptr else None,
                                                   C.c_void_p(needs_ptr),
                                                   C.c_void_p(max_d2_ptr) if max_d2_ptr else None)
    if st != 0:
        raise GpuError(st, "mdgpu_pairs_need_rebuild_noalloc")


class Evaluator:
    """One mdgpu context (device + stream) holding one uploaded system."""

    def __init__(self, device: int = 0):
        self.lib = gpu_lib()
        self.ctx = C.c_void_p()
        st = self.lib.mdgpu_create(device, C.byref(self.ctx))
        if st != 0:
            raise GpuError(st, f"mdgpu_create(device={device}) failed")
        self.n_atoms = 0

    def _check(self, st):
        if st != 0:
            raise GpuError(st, self.lib.mdgpu_last_error(self.ctx).decode())

    def upload(self, desc: SystemDesc):
        self._check(self.lib.mdgpu_upload_system(self.ctx, C.byref(desc)))
        self.n_atoms = desc.n_atoms

    def set_params(self, params: Params):
        self._check(self.lib.mdgpu_set_params(self.ctx, C.byref(params)))

    def evaluate(self, coords, box):
        """Host-buffer evaluation: returns (energies dict, forces (N,3), virial (3,3))."""
        x = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1)
        b = np.ascontiguousarray(box, dtype=np.float64)
        f = np.zeros(x.size, dtype=np.float64)
        v = np.zeros(9, dtype=np.float64)
        e = Energies()
        self._check(self.lib.mdgpu_eval(self.ctx, x.ctypes.data_as(P_F64), b.ctypes.data_as(P_F64),
                                        C.byref(e), f.ctypes.data_as(P_F64), v.ctypes.data_as(P_F64)))
        return e.as_dict(), f.reshape(-1, 3), v.reshape(3, 3)

    def evaluate_device(self, coords_ptr: int, box, forces_ptr: int):
        """Device-buffer evaluation (pointers to 3N doubles on the device)."""
        b = np.ascontiguousarray(box, dtype=np.float64)
        e = Energies()
        v = np.zeros(9, dtype=np.float64)
        self._check(self.lib.mdgpu_eval_device(self.ctx, C.c_void_p(coords_ptr), b.ctypes.data_as(P_F64),
                                               C.byref(e), C.c_void_p(forces_ptr),
                                               v.ctypes.data_as(P_F64)))
        return e.as_dict(), v.reshape(3, 3)

    def set_box(self, box):
        """Box for step_device (kept with its ctypes pointer between calls)."""
        self._step_box = np.array(box, dtype=np.float64).reshape(3)
        self._step_box_p = self._step_box.ctypes.data_as(P_F64)
        self._step_e = Energies()
        self._step_e_p = C.byref(self._step_e)
        self._step_v = np.zeros(9, dtype=np.float64)
        self._step_v_p = self._step_v.ctypes.data_as(P_F64)

    def step_device(self, coords_ptr: int, forces_ptr: int):
        """The device-buffer evaluation of evaluate_device in the box of set_box,
        for time-step loops: it builds no Python objects; the energies and
        virial of the call are read with last_energies()."""
        st = self.lib.mdgpu_eval_device(self.ctx, coords_ptr, self._step_box_p, self._step_e_p, forces_ptr,
                                        self._step_v_p)
        if st != 0:
            self._check(st)

    def last_energies(self):
        """(energies dict, virial) of the last step_device call."""
        return self._step_e.as_dict(), self._step_v.copy().reshape(3, 3)

    def adjoint(self, coords, box, weights: dict, desc: SystemDesc) -> dict:
        """sum_t w_t dE_t/dtheta at one configuration (AMBER semantics).

        ``weights`` maps energy-term names to weights (``total`` adds to every
        term of the total); returns a dict of gradient arrays (GRAD_FIELDS).
        """
        x = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1)
        b = np.ascontiguousarray(box, dtype=np.float64)
        w = energies_struct(weights)
        arrs, g = alloc_grads(desc)
        self._check(self.lib.mdgpu_adjoint(self.ctx, x.ctypes.data_as(P_F64), b.ctypes.data_as(P_F64),
                                           C.byref(w), C.byref(g)))
        return arrs

    def set_timing(self, mode):
        """0/False off, 1/True every kernel group, 2 the pair kernel only, 3 the pair
        kernel in one evaluation of every 16."""
        self._check(self.lib.mdgpu_set_timing(self.ctx, int(mode)))

    def timing(self) -> dict:
        out = {}
        for k, name in enumerate(KERNELS):
            ms = C.c_double()
            n = C.c_int64()
            self._check(self.lib.mdgpu_get_timing(self.ctx, k, C.byref(ms), C.byref(n)))
            out[name] = (ms.value, n.value)
        return out

    def set_graph(self, on: bool):
        """Allow (default) or forbid the captured-graph replay of evaluations."""
        self._check(self.lib.mdgpu_set_graph(self.ctx, 1 if on else 0))

    def graph_state(self) -> int:
        """0 no graph yet, 1 replaying a captured evaluation, 2 direct launches for good."""
        return int(self.lib.mdgpu_graph_state(self.ctx))

    def set_shapes(self, split: int, fft_lines: int):
        """Pair-kernel split (1..8) and FFT lines per block (0 = 16, 8, 4, ...)."""
        self._check(self.lib.mdgpu_set_shapes(self.ctx, int(split), int(fft_lines)))

    def shapes(self):
        sp, fl = C.c_int(), C.c_int()
        self._check(self.lib.mdgpu_get_shapes(self.ctx, C.byref(sp), C.byref(fl)))
        return sp.value, fl.value

    def tune(self, d_coords: int, box, d_forces: int = 0, reps: int = 10):
        """Time every pair-kernel split and FFT block shape on these device
        buffers (before the first graph evaluation) and keep the fastest;
        returns (split, fft_lines, median ms per evaluation)."""
        b = np.ascontiguousarray(box, dtype=np.float64)
        sp, fl, ms = C.c_int(), C.c_int(), C.c_double()
        self._check(self.lib.mdgpu_tune(self.ctx, C.c_void_p(d_coords), b.ctypes.data_as(P_F64),
                                        C.c_void_p(d_forces or None), int(reps), C.byref(sp), C.byref(fl),
                                        C.byref(ms)))
        return sp.value, fl.value, ms.value

    def set_verlet(self, skin: float):
        """Keep the pair list across evaluations with this skin (A) until an atom
        has moved more than skin / 2 since the build; 0 builds it every time."""
        self._check(self.lib.mdgpu_set_verlet(self.ctx, float(skin)))

    def set_box_angles(self, angles):
        """Cell angles alpha, beta, gamma in degrees for every later call (the
        box arguments then give the edge lengths); 90, 90, 90 is orthorhombic."""
        a = np.ascontiguousarray(angles, dtype=np.float64).reshape(3)
        self._check(self.lib.mdgpu_set_box_angles(self.ctx, a.ctypes.data_as(P_F64)))

    def search_mode(self) -> int:
        """0: minimum-image cluster search; 1: the general per-shift search."""
        return int(self.lib.mdgpu_get_search_mode(self.ctx))

    def box_angles(self) -> np.ndarray:
        a = np.zeros(3)
        self._check(self.lib.mdgpu_get_box_angles(self.ctx, a.ctypes.data_as(P_F64)))
        return a

    def list_prunes(self) -> int:
        """Prunes of the kept Verlet list since the upload."""
        p = C.c_int64()
        self._check(self.lib.mdgpu_get_list_prunes(self.ctx, C.byref(p)))
        return p.value

    def list_builds(self) -> tuple:
        """(list builds, successful evaluations) since the system upload."""
        b, e = C.c_int64(), C.c_int64()
        self._check(self.lib.mdgpu_get_list_builds(self.ctx, C.byref(b), C.byref(e)))
        return b.value, e.value

    def needs_rebuild(self, x_ptr: int) -> bool:
        """The rebuild criterion for device coordinates at x_ptr (3N doubles)."""
        out = C.c_int()
        self._check(self.lib.mdgpu_pairs_need_rebuild(self.ctx, C.c_void_p(x_ptr), C.byref(out)))
        return bool(out.value)

    def minimize(self, x_ptr: int, box, max_iter=500, rms_force_tol=1e-2, max_step=0.2, history=8) -> dict:
        """L-BFGS on device coordinates at x_ptr (3N doubles, minimised in place)."""
        b = np.ascontiguousarray(box, dtype=np.float64).reshape(-1)[:3]
        o = MinOpts(int(max_iter), float(rms_force_tol), float(max_step), int(history))
        r = MinResult()
        self._check(self.lib.mdgpu_minimize(self.ctx, C.c_void_p(x_ptr), b.ctypes.data_as(P_F64), C.byref(o),
                                            C.byref(r)))
        return {"iterations": r.iterations, "evaluations": r.evaluations, "converged": bool(r.converged),
                "energy": r.energy, "rms_force": r.rms_force, "max_force": r.max_force}

    def md_set_masses(self, masses):
        """Atom masses (amu) for molecular dynamics."""
        m = np.ascontiguousarray(masses, dtype=np.float64)
        self._check(self.lib.mdgpu_md_set_masses(self.ctx, m.ctypes.data_as(P_F64)))

    def md_run(self, x_ptr: int, v_ptr: int, f_ptr: int, box, dt: float, nsteps: int):
        """Velocity Verlet on the device state (3N doubles each; f holds the forces
        at x on entry).  Returns (trace (nsteps, 3): potential, kinetic, pressure (bar,
        NaN unless the virial is on); last energies)."""
        b = np.ascontiguousarray(box, dtype=np.float64).reshape(-1)[:3]
        tr = np.zeros((max(int(nsteps), 1), 3))
        e = Energies()
        self._check(self.lib.mdgpu_md_run(self.ctx, C.c_void_p(x_ptr), C.c_void_p(v_ptr), C.c_void_p(f_ptr),
                                          b.ctypes.data_as(P_F64), float(dt), int(nsteps),
                                          tr.ctypes.data_as(C.c_void_p), C.byref(e)))
        return tr[:int(nsteps)], e.as_dict()

    def set_counting(self, on: bool):
        self._check(self.lib.mdgpu_set_counting(self.ctx, 1 if on else 0))

    def set_virial(self, on: bool):
        """Form the full virial (pair, bonded, correction and reciprocal) in each evaluation."""
        self._check(self.lib.mdgpu_set_virial(self.ctx, 1 if on else 0))

    def virial(self) -> np.ndarray:
        """Full virial W = -dE/d(strain) of the last evaluation (3, 3), kcal/mol."""
        w = np.zeros(9)
        self._check(self.lib.mdgpu_get_virial(self.ctx, w.ctypes.data_as(P_F64)))
        return w.reshape(3, 3)

    def counts(self) -> dict:
        c = (C.c_int64 * 4)()
        self._check(self.lib.mdgpu_get_counts(self.ctx, c))
        return {"pairs": c[0], "tiles": c[1], "clusters": c[2], "atoms": c[3]}

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.mdgpu_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many independent systems evaluated by one graph launch (mdgpu_batch_*)."""

    def __init__(self, device: int = 0):
        self.lib = gpu_lib()
        self.b = C.c_void_p()
        st = self.lib.mdgpu_batch_create(device, C.byref(self.b))
        if st != 0:
            raise GpuError(st, f"mdgpu_batch_create(device={device}) failed")
        self.n_atoms = []
        self._settings = {"streams": 4, "split": 4, "fft_lines": 8}  # mdgpu_batch_create's defaults

    def _check(self, st):
        if st != 0:
            raise GpuError(st, self.lib.mdgpu_batch_last_error(self.b).decode())

    def add(self, desc: SystemDesc, params: Params, box) -> int:
        bx = np.ascontiguousarray(np.broadcast_to(np.asarray(box, dtype=np.float64), (3,)))
        idx = C.c_int()
        self._check(self.lib.mdgpu_batch_add(self.b, C.byref(desc), C.byref(params), bx.ctypes.data_as(P_F64),
                                             C.byref(idx)))
        self.n_atoms.append(int(desc.n_atoms))
        return idx.value

    def __len__(self):
        return len(self.n_atoms)

    def evaluate_device(self, coords_ptr: int, forces_ptr: int = 0) -> list:
        e = (Energies * len(self))()
        self._check(self.lib.mdgpu_batch_eval_device(self.b, C.c_void_p(coords_ptr), e,
                                                     C.c_void_p(forces_ptr) if forces_ptr else None))
        return [x.as_dict() for x in e]

    def evaluate(self, coords, want_forces: bool = True):
        x = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1)
        f = np.zeros_like(x) if want_forces else None
        e = (Energies * len(self))()
        self._check(self.lib.mdgpu_batch_eval(self.b, x.ctypes.data_as(P_F64), e,
                                              f.ctypes.data_as(P_F64) if f is not None else None))
        return [y.as_dict() for y in e], (f.reshape(-1, 3) if f is not None else None)

    def configure(self, mode=0, streams=None, split=None, fft_lines=None):
        """Launch settings; an argument left out keeps its current value."""
        cur = self._settings
        s = cur["streams"] if streams is None else streams
        p = cur["split"] if split is None else split
        f = cur["fft_lines"] if fft_lines is None else fft_lines
        self._check(self.lib.mdgpu_batch_configure(self.b, mode, s, p, f))
        self._settings = {"streams": s, "split": p, "fft_lines": f}

    def tune(self, coords_ptr: int, forces_ptr: int = 0, reps: int = 3) -> dict:
        s, p, fl = C.c_int(), C.c_int(), C.c_int()
        ms = C.c_double()
        self._check(self.lib.mdgpu_batch_tune(self.b, C.c_void_p(coords_ptr),
                                              C.c_void_p(forces_ptr) if forces_ptr else None, reps,
                                              C.byref(s), C.byref(p), C.byref(fl), C.byref(ms)))
        self._settings = {"streams": s.value, "split": p.value, "fft_lines": fl.value}
        return {"streams": s.value, "split": p.value, "fft_lines": fl.value, "ms_pass": ms.value,
                "mode": self.mode()}

    def mode(self) -> int:
        """0 graph, 1 serial, 2 the graph's launches issued directly."""
        m = C.c_int()
        self._check(self.lib.mdgpu_batch_mode(self.b, C.byref(m)))
        return m.value

    def system_config(self, index: int) -> tuple:
        """(pair-tile split, FFT lines) of system `index`."""
        p, f = C.c_int(), C.c_int()
        self._check(self.lib.mdgpu_batch_system_config(self.b, index, C.byref(p), C.byref(f)))
        return p.value, f.value

    def graph_state(self) -> int:
        """0 no graph yet, 1 graph in use, 2 graph stale (serial evaluation)."""
        st = C.c_int()
        self._check(self.lib.mdgpu_batch_graph_state(self.b, C.byref(st)))
        return st.value

    def close(self):
        if getattr(self, "b", None) and self.b.value:
            self.lib.mdgpu_batch_destroy(self.b)
            self.b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
