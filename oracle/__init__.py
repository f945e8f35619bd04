"""ctypes binding of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product never imports it.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os

import numpy as np

from md_amd.native import Energies, ParamGrads, Params, SystemDesc, P_F64, alloc_grads, energies_struct

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_fast = None
FAST_FLAGS = "gcc -std=gnu11 -O3 -march=x86-64-v3 -mtune=generic -fno-fast-math"


class Opts(C.Structure):
    _fields_ = [("list_skin", C.c_double), ("skip_all_pairs", C.c_int), ("full_virial", C.c_int)]


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        name = "liboracle_asan.so" if os.environ.get("MD_AMD_SANITIZE") == "1" else "liboracle.so"
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            raise ImportError(f"{path} is not built (make -C oracle)")
        L = C.CDLL(path)
        L.oracle_eval.argtypes = [C.POINTER(SystemDesc), C.POINTER(Params), P_F64, P_F64, C.POINTER(Opts),
                                  C.POINTER(Energies), P_F64, P_F64]
        L.oracle_ref_pairlist.argtypes = [C.POINTER(SystemDesc), P_F64, P_F64, C.c_double,
                                          C.POINTER(C.c_int), C.c_long]
        L.oracle_ref_pairlist.restype = C.c_long
        L.oracle_ewald_recip_dense.argtypes = [C.c_int, P_F64, P_F64, P_F64, C.c_double, C.c_int, P_F64, P_F64]
        L.oracle_ewald_recip_dense.restype = C.c_double
        L.oracle_pme_recip.argtypes = [C.c_int, P_F64, P_F64, P_F64, C.c_double, C.POINTER(C.c_int), C.c_int,
                                       P_F64, P_F64]
        L.oracle_pme_recip.restype = C.c_double
        L.oracle_pme_potential.argtypes = [C.c_int, P_F64, P_F64, P_F64, C.c_double, C.POINTER(C.c_int),
                                           C.c_int, P_F64]
        L.oracle_pme_potential.restype = C.c_double
        L.oracle_param_grads.argtypes = [C.POINTER(SystemDesc), C.POINTER(Params), P_F64, P_F64,
                                         C.POINTER(Energies), C.POINTER(ParamGrads)]
        L.oracle_bspline.argtypes = [C.c_int, C.c_double, P_F64, P_F64]
        L.oracle_pairs_need_rebuild.argtypes = [C.c_int, P_F64, P_F64, P_F64, C.c_double]
        L.oracle_pairs_need_rebuild.restype = C.c_int
        L.oracle_minimize.argtypes = [C.POINTER(SystemDesc), C.POINTER(Params), P_F64, P_F64, C.c_int, C.c_double,
                                      C.c_double, C.c_int, P_F64, P_F64]
        L.oracle_md_run.argtypes = [C.POINTER(SystemDesc), C.POINTER(Params), P_F64, P_F64, P_F64, P_F64, P_F64,
                                    C.c_double, C.c_int, P_F64]
        L.oracle_coulomb_constant.restype = C.c_double
        L.oracle_set_box_angles.argtypes = [P_F64]
        L.oracle_set_box_angles.restype = C.c_int
        L.oracle_cell_matrix.argtypes = [P_F64, P_F64, P_F64]
        L.oracle_cell_matrix.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(P_F64)


def fast_lib() -> C.CDLL:
    """liboracle_fast.so: the same restatement built for timing (FAST_FLAGS),
    used by the CPU baseline only."""
    global _fast
    if _fast is None:
        path = os.path.join(_HERE, "liboracle_fast.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} is not built (make -C oracle)")
        L = C.CDLL(path)
        L.oracle_eval.argtypes = [C.POINTER(SystemDesc), C.POINTER(Params), P_F64, P_F64, C.POINTER(Opts),
                                  C.POINTER(Energies), P_F64, P_F64]
        _fast = L
    return _fast


def timed_evaluate(desc: SystemDesc, params: Params, coords, box, skip_all_pairs=False, list_skin=2.0) -> float:
    """One full evaluation (energies and forces) of an orthorhombic box by the
    timing build; returns the wall time in seconds."""
    import time
    x = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1)
    b = np.ascontiguousarray(box, dtype=np.float64).reshape(-1)[:3]
    f = np.zeros(x.size)
    v = np.zeros(9)
    e = Energies()
    o = Opts(list_skin, 1 if skip_all_pairs else 0, 0)
    t0 = time.perf_counter()
    st = fast_lib().oracle_eval(C.byref(desc), C.byref(params), _p(x), _p(b), C.byref(o), C.byref(e), _p(f), _p(v))
    dt = time.perf_counter() - t0
    if st != 0:
        raise RuntimeError("oracle_eval failed")
    return dt


@contextlib.contextmanager
def _cell(box):
    """Edge lengths of box (3 values, or 6 with the angles alpha, beta, gamma
    in degrees, as on an rst7 box line); the angles hold for the duration."""
    b = np.ascontiguousarray(box, dtype=np.float64).reshape(-1)
    ang = b[3:6] if b.size >= 6 else np.array([90.0, 90.0, 90.0])
    ang = np.ascontiguousarray(ang, dtype=np.float64)
    if lib().oracle_set_box_angles(_p(ang)) != 0:
        raise ValueError(f"bad cell angles {ang}")
    try:
        yield np.ascontiguousarray(b[:3])
    finally:
        lib().oracle_set_box_angles(_p(np.array([90.0, 90.0, 90.0])))


def evaluate(desc: SystemDesc, params: Params, coords, box, list_skin=2.0, skip_all_pairs=False,
             want_forces=True, full_virial=False):
    """Returns (energies dict, forces (N,3) or None, virial (3,3))."""
    x = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1)
    f = np.zeros(x.size) if want_forces else None
    v = np.zeros(9)
    e = Energies()
    o = Opts(list_skin, 1 if skip_all_pairs else 0, 1 if full_virial else 0)
    if full_virial and f is None:
        f = np.zeros(x.size)
    with _cell(box) as b:
        st = lib().oracle_eval(C.byref(desc), C.byref(params), _p(x), _p(b), C.byref(o), C.byref(e),
                               _p(f) if f is not None else None, _p(v))
    if st != 0:
        raise RuntimeError("oracle_eval failed")
    return e.as_dict(), (f.reshape(-1, 3) if f is not None else None), v.reshape(3, 3)


def cell_matrix(box):
    """(H with the edge vectors a, b, c as columns, face distances) of a box
    of 3 lengths or 6 values with the angles."""
    h = np.zeros(9)
    face = np.zeros(3)
    with _cell(box) as b:
        if lib().oracle_cell_matrix(_p(b), _p(h), _p(face)) != 0:
            raise ValueError("the box forms no cell")
    return h.reshape(3, 3), face


def ref_pairlist(desc: SystemDesc, coords, box, list_cutoff):
    x = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1)
    b = np.ascontiguousarray(box, dtype=np.float64)
    n = lib().oracle_ref_pairlist(C.byref(desc), _p(x), _p(b), list_cutoff, None, 0)
    out = np.zeros(3 * max(n, 1), dtype=np.int32)
    lib().oracle_ref_pairlist(C.byref(desc), _p(x), _p(b), list_cutoff,
                              out.ctypes.data_as(C.POINTER(C.c_int)), n)
    return out[:3 * n].reshape(-1, 3)


def ewald_recip_dense(q, coords, box, alpha, mmax):
    q = np.ascontiguousarray(q, dtype=np.float64)
    x = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1)
    f = np.zeros(x.size)
    v = np.zeros(9)
    with _cell(box) as b:
        e = lib().oracle_ewald_recip_dense(q.size, _p(q), _p(x), _p(b), alpha, mmax, _p(f), _p(v))
    return e, f.reshape(-1, 3), v.reshape(3, 3)


def pme_recip(q, coords, box, alpha, grid, order=4):
    q = np.ascontiguousarray(q, dtype=np.float64)
    x = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1)
    g = (C.c_int * 3)(*grid)
    f = np.zeros(x.size)
    v = np.zeros(9)
    with _cell(box) as b:
        e = lib().oracle_pme_recip(q.size, _p(q), _p(x), _p(b), alpha, g, order, _p(f), _p(v))
    return e, f.reshape(-1, 3), v.reshape(3, 3)


def bspline(order, w):
    th = np.zeros(order)
    dth = np.zeros(order)
    lib().oracle_bspline(order, w, _p(th), _p(dth))
    return th, dth


def coulomb_constant() -> float:
    return lib().oracle_coulomb_constant()


def pme_potential(q, coords, box, alpha, grid, order=4):
    """(E_rec, dE_rec/dq_i)."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    x = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1)
    b = np.ascontiguousarray(box, dtype=np.float64)
    g = (C.c_int * 3)(*grid)
    pot = np.zeros(q.size)
    e = lib().oracle_pme_potential(q.size, _p(q), _p(x), _p(b), alpha, g, order, _p(pot))
    return e, pot


def param_grads(desc: SystemDesc, params: Params, coords, box, weights: dict) -> dict:
    """sum_t w_t dE_t/dtheta (AMBER semantics) as a dict of arrays."""
    x = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1)
    b = np.ascontiguousarray(box, dtype=np.float64)
    arrs, g = alloc_grads(desc)
    w = energies_struct(weights)
    st = lib().oracle_param_grads(C.byref(desc), C.byref(params), _p(x), _p(b), C.byref(w), C.byref(g))
    if st != 0:
        raise RuntimeError("oracle_param_grads failed")
    return arrs


def pairs_need_rebuild(coords, coords_ref, box, skin) -> bool:
    """The Verlet rebuild criterion (NonBondedUtils.cpp:87-100)."""
    x = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1)
    r = np.ascontiguousarray(coords_ref, dtype=np.float64).reshape(-1)
    with _cell(box) as b:
        return bool(lib().oracle_pairs_need_rebuild(x.size // 3, _p(x), _p(r), _p(b), float(skin)))


def md_run(desc: SystemDesc, params: Params, masses, x, v, f, box, dt, nsteps):
    """Velocity Verlet; returns (x, v, f, trace (nsteps, 2): potential, kinetic)."""
    m = np.ascontiguousarray(masses, dtype=np.float64)
    x = np.array(x, dtype=np.float64).reshape(-1)
    v = np.array(v, dtype=np.float64).reshape(-1)
    f = np.array(f, dtype=np.float64).reshape(-1)
    tr = np.zeros(2 * max(nsteps, 1))
    with _cell(box) as b:
        st = lib().oracle_md_run(C.byref(desc), C.byref(params), _p(m), _p(x), _p(v), _p(f), _p(b), float(dt),
                                 int(nsteps), _p(tr))
    if st != 0:
        raise RuntimeError("oracle_md_run failed")
    return x.reshape(-1, 3), v.reshape(-1, 3), f.reshape(-1, 3), tr[:2 * nsteps].reshape(-1, 2)


def minimize(desc: SystemDesc, params: Params, x, box, max_iter=500, rms_tol=1e-2, max_step=0.2, history=8):
    """L-BFGS; returns (x, energies per accepted iteration, result dict)."""
    x = np.array(x, dtype=np.float64).reshape(-1)
    en = np.zeros(max_iter + 2)
    out = np.zeros(5)
    with _cell(box) as b:
        st = lib().oracle_minimize(C.byref(desc), C.byref(params), _p(x), _p(b), int(max_iter), float(rms_tol),
                                   float(max_step), int(history), _p(en), _p(out))
    if st != 0:
        raise RuntimeError("oracle_minimize failed")
    it = int(out[0])
    res = {"iterations": it, "evaluations": int(out[1]), "converged": bool(out[2]), "rms_force": out[3],
           "max_force": out[4], "energy": en[it]}
    return x.reshape(-1, 3), en[:it + 1], res
