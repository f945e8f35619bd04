"""Seeded synthetic systems and an Amber prmtop/rst7 writer.

The reference ships no topology files (its two rst7 files have no matching
prmtop), so every workload is generated here and written in the Amber text
format that the reference parser accepts:

* every numeric field is separated by whitespace, because the reference splits
  numeric sections on blanks (src/util/IO.cpp:109-136);
* POINTERS precedes ATOM_NAME, which precedes NONBONDED_PARM_INDEX
  (src/AmberTopology/AmberTopology.cpp:213, TopologyParser.cpp:133-136);
* five CMAP parameter sets of one resolution are always present, also for
  systems without CMAP terms (CreateNonBondedFromParsed.cpp:69-72);
* the rst7 has no velocities, so the six values after the 3N coordinates are
  the box (AmberTopology.cpp:405-438).

Workloads (sizes as named by the benchmark definitions):

* ``water_box``: 40 A cube, 2,150 three-site waters on a jittered 13^3
  lattice (6,450 atoms);
* ``solvated_polymer``: 66 A cube, a 4,800-atom random-walk heteropolymer
  with 30 atom types, angles + Urey-Bradley, cosine dihedrals, impropers and
  300 CMAP terms, solvated in ~8,000 waters (~28.8 k atoms).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

# Amber's charge unit: prmtop CHARGE = q[e] * 18.2223
AMBER_CHARGE_SCALE = 18.2223

WATER_R_OH = 0.9572
WATER_THETA = 104.52 * math.pi / 180.0
O_SIGMA, O_EPS = 3.15061, 0.1521


@dataclass
class Topology:
    """Plain arrays describing one system (all indices 0-based)."""
    title: str
    coords: np.ndarray                    # (N,3) float64
    box: float
    charge_e: np.ndarray                  # (N,) elementary charges
    atom_type: np.ndarray                 # (N,) 1-based LJ type
    mass: np.ndarray
    atomic_number: np.ndarray
    names: list
    amber_types: list
    type_sigma: np.ndarray                # (T,)
    type_eps: np.ndarray                  # (T,)
    res_labels: list
    res_starts: list                      # 0-based first atom of each residue
    mol_sizes: list
    bonds: list = field(default_factory=list)        # (i, j, type0, hasH)
    bond_k: list = field(default_factory=list)
    bond_r0: list = field(default_factory=list)
    angles: list = field(default_factory=list)       # (i, j, k, type0, hasH)
    angle_k: list = field(default_factory=list)
    angle_t0: list = field(default_factory=list)
    ubs: list = field(default_factory=list)          # (i, j, type0)
    ub_k: list = field(default_factory=list)
    ub_r0: list = field(default_factory=list)
    dihedrals: list = field(default_factory=list)    # (i, j, k, l, type0, hasH, no14, improper)
    dih_k: list = field(default_factory=list)
    dih_n: list = field(default_factory=list)
    dih_phase: list = field(default_factory=list)
    dih_scee: list = field(default_factory=list)
    dih_scnb: list = field(default_factory=list)
    impropers: list = field(default_factory=list)    # (i, j, k, l, type0)
    imp_k: list = field(default_factory=list)
    imp_psi0: list = field(default_factory=list)
    cmaps: list = field(default_factory=list)        # (a, b, c, d, e, set1)
    cmap_res: int = 24
    cmap_grids: list = field(default_factory=list)   # 5 x (res*res) float64, phi-major
    exclusions: list = field(default_factory=list)   # per atom: sorted partners j > i

    @property
    def n_atoms(self) -> int:
        return int(self.coords.shape[0])


# --------------------------------------------------------------------------
# geometry helpers
# --------------------------------------------------------------------------

def _random_rotation(rng: np.random.Generator) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def _water_hydrogens(rng, o):
    s, c = math.sin(WATER_THETA / 2), math.cos(WATER_THETA / 2)
    local = np.array([[WATER_R_OH * s, 0.0, WATER_R_OH * c],
                      [-WATER_R_OH * s, 0.0, WATER_R_OH * c]])
    rot = _random_rotation(rng)
    return o + local @ rot.T


def _nerf(a, b, c, bond, theta, torsion):
    """Place d so that |cd|=bond, angle(b,c,d)=theta, dihedral(a,b,c,d)=torsion."""
    bc = c - b
    bc /= np.linalg.norm(bc)
    n = np.cross(b - a, bc)
    n /= np.linalg.norm(n)
    m = np.cross(n, bc)
    d2 = np.array([-bond * math.cos(theta),
                   bond * math.sin(theta) * math.cos(torsion),
                   bond * math.sin(theta) * math.sin(torsion)])
    return c + d2[0] * bc + d2[1] * m + d2[2] * n


def _lj_pair_tables(sigma, eps):
    """Amber LJ tables indexed by idx(i<=j) = j(j+1)/2 + i (0-based)."""
    t = len(sigma)
    n = t * (t + 1) // 2
    a = np.zeros(n)
    b = np.zeros(n)
    for j in range(t):
        for i in range(j + 1):
            s = 0.5 * (sigma[i] + sigma[j])
            e = math.sqrt(eps[i] * eps[j])
            a[j * (j + 1) // 2 + i] = 4.0 * e * s ** 12
            b[j * (j + 1) // 2 + i] = 4.0 * e * s ** 6
    return a, b


def with_hydrogen_lj(arrays: dict, h_type: int, sigma: float = 0.4, eps: float = 0.046) -> dict:
    """A copy of system arrays (md_amd.native.Topology.arrays layout) whose LJ
    type ``h_type`` (0-based; a type without LJ, such as the TIP3P hydrogens)
    gets ``sigma`` / ``eps``, combined with every type by Lorentz-Berthelot.
    The defaults are the hydrogen parameters of the CHARMM variant of TIP3P.
    Without them a hydrogen and an unlike charge have no repulsion, so the
    energy has no lower bound and a minimiser can follow a hydrogen into
    another atom; energies and forces are unchanged in form.  The other
    types' sigma and eps are recovered from the diagonal table entries
    (A = 4 eps sigma^12, B = 4 eps sigma^6)."""
    out = dict(arrays)
    t = int(arrays["n_types"])
    nb = np.asarray(arrays["nb_index"]).reshape(t, t)
    lj_a, lj_b = np.array(arrays["lj_a"], dtype=np.float64), np.array(arrays["lj_b"], dtype=np.float64)
    sig, ep = np.zeros(t), np.zeros(t)
    for k in range(t):
        a, b = lj_a[nb[k, k] - 1], lj_b[nb[k, k] - 1]
        if a > 0 and b > 0:
            sig[k], ep[k] = (a / b) ** (1.0 / 6.0), b * b / (4.0 * a)
    sig[h_type], ep[h_type] = sigma, eps
    for k in range(t):
        s, e = 0.5 * (sig[h_type] + sig[k]), math.sqrt(ep[h_type] * ep[k])
        lj_a[nb[h_type, k] - 1] = 4.0 * e * s ** 12
        lj_b[nb[h_type, k] - 1] = 4.0 * e * s ** 6
    out["lj_a"], out["lj_b"] = lj_a, lj_b
    return out


# --------------------------------------------------------------------------
# water
# --------------------------------------------------------------------------

def _add_water_topology(top: Topology, first: int, o_type: int, h_type: int):
    o, h1, h2 = first, first + 1, first + 2
    top.bonds += [(o, h1, 0, True), (o, h2, 0, True)]
    top.angles += [(h1, o, h2, 0, True)]
    top.exclusions[o] = [h1, h2]
    top.exclusions[h1] = [h2]
    top.exclusions[h2] = []


def _water_sites(rng, box, n_lattice, jitter):
    sp = box / n_lattice
    g = (np.arange(n_lattice) + 0.5) * sp
    sites = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    return sites + rng.normal(scale=jitter, size=sites.shape)


def _five_cmap_grids(rng, res):
    grids = []
    ang = -math.pi + np.arange(res) * (2 * math.pi / res)
    phi, psi = np.meshgrid(ang, ang, indexing="ij")
    for _ in range(5):
        e = np.zeros((res, res))
        for m in range(4):
            for n in range(4):
                amp = rng.normal(scale=0.6 / (1 + m + n))
                ph = rng.uniform(0, 2 * math.pi)
                e += amp * np.cos(m * phi + n * psi + ph)
        grids.append(e.reshape(-1))
    return grids


def water_box(seed: int = 11, box: float = 40.0, n_water: int = 2150,
              n_lattice: int = 13, jitter: float = 0.3) -> Topology:
    """2,150 waters on a jittered 13^3 lattice in a 40 A cube (6,450 atoms)."""
    rng = np.random.default_rng(seed)
    sites = _water_sites(rng, box, n_lattice, jitter)
    pick = np.sort(rng.permutation(sites.shape[0])[:n_water])
    n = 3 * n_water
    coords = np.zeros((n, 3))
    for w, s in enumerate(pick):
        o = sites[s]
        coords[3 * w] = o
        coords[3 * w + 1:3 * w + 3] = _water_hydrogens(rng, o)
    top = Topology(
        title=f"water box seed {seed}", coords=coords, box=box,
        charge_e=np.tile([-0.834, 0.417, 0.417], n_water),
        atom_type=np.tile([1, 2, 2], n_water),
        mass=np.tile([15.999, 1.008, 1.008], n_water),
        atomic_number=np.tile([8, 1, 1], n_water),
        names=["O", "H1", "H2"] * n_water, amber_types=["OW", "HW", "HW"] * n_water,
        type_sigma=np.array([O_SIGMA, 0.0]), type_eps=np.array([O_EPS, 0.0]),
        res_labels=["WAT"] * n_water, res_starts=list(range(0, n, 3)),
        mol_sizes=[3] * n_water)
    top.exclusions = [[] for _ in range(n)]
    top.bond_k, top.bond_r0 = [553.0], [WATER_R_OH]
    top.angle_k, top.angle_t0 = [100.0], [WATER_THETA]
    for w in range(n_water):
        _add_water_topology(top, 3 * w, 1, 2)
    top.cmap_res = 24
    top.cmap_grids = _five_cmap_grids(np.random.default_rng(seed + 1000), 24)
    return top


# --------------------------------------------------------------------------
# solvated heteropolymer
# --------------------------------------------------------------------------

class _Grid:
    """Uniform hash grid for the self-avoidance and solvent-overlap checks."""

    def __init__(self, box, cell):
        self.n = max(1, int(box // cell))
        self.cell = box / self.n
        self.cells: dict = {}

    def key(self, p):
        return tuple(int(math.floor(v / self.cell)) for v in p)

    def add(self, idx, p):
        self.cells.setdefault(self.key(p), []).append((idx, p))

    def remove(self, idx, p):
        lst = self.cells[self.key(p)]
        for k, (j, _q) in enumerate(lst):
            if j == idx:
                del lst[k]
                return

    def near(self, p, rad):
        k = self.key(p)
        r = int(math.ceil(rad / self.cell))
        out = []
        for dx in range(-r, r + 1):
            for dy in range(-r, r + 1):
                for dz in range(-r, r + 1):
                    out.extend(self.cells.get((k[0] + dx, k[1] + dy, k[2] + dz), ()))
        return out


def _grow_chain(rng, n, box, bond_r, theta, min_dist, margin=1.5, tries=60,
                max_backtracks=200000):
    """Self-avoiding chain confined to the box, grown with backtracking."""
    pos = np.zeros((n, 3))
    grid = _Grid(box, 3.0)
    c = np.full(3, box / 2)
    pos[0] = c
    pos[1] = c + np.array([bond_r[0], 0, 0])
    pos[2] = _nerf(c + np.array([0, 1.0, 0]), pos[0], pos[1], bond_r[1], theta[0], 1.0)
    for i in range(3):
        grid.add(i, pos[i])
    i = 3
    backtracks = 0
    md2 = min_dist * min_dist
    while i < n:
        placed = False
        for _ in range(tries):
            tors = rng.uniform(-math.pi, math.pi)
            p = _nerf(pos[i - 3], pos[i - 2], pos[i - 1], bond_r[i - 1], theta[i - 2], tors)
            if np.any(p < margin) or np.any(p > box - margin):
                continue
            if any(j < i - 3 and float(np.sum((q - p) ** 2)) < md2
                   for j, q in grid.near(p, min_dist)):
                continue
            placed = True
            break
        if placed:
            pos[i] = p
            grid.add(i, p)
            i += 1
            continue
        backtracks += 1
        if backtracks > max_backtracks:
            raise RuntimeError(f"polymer growth stuck at atom {i}")
        back = min(i - 3, int(rng.integers(2, 12)))
        for j in range(i - back, i):
            grid.remove(j, pos[j])
        i -= back
    return pos


def solvated_polymer(seed: int = 22, box: float = 66.0, n_poly: int = 4800,
                     n_water: int = 8000, n_cmap: int = 300, n_types: int = 30,
                     water_spacing: float = 2.9, jitter: float = 0.25,
                     solvent_gap: float = 2.4) -> Topology:
    """Random-walk heteropolymer solvated in water (66 A cube, ~28.8 k atoms)."""
    rng = np.random.default_rng(seed)
    # parameter tables
    nbt, nat, ndt, nit = 20, 20, 24, 8
    bond_k = rng.uniform(200.0, 600.0, nbt)
    bond_r0 = rng.uniform(1.0, 1.5, nbt)
    ang_k = rng.uniform(40.0, 120.0, nat)
    ang_t0 = np.deg2rad(rng.uniform(105.0, 125.0, nat))
    dih_k = rng.uniform(0.1, 2.0, ndt)
    dih_n = rng.integers(1, 7, ndt).astype(float)
    dih_phase = np.where(rng.random(ndt) < 0.7, np.where(rng.random(ndt) < 0.5, 0.0, math.pi),
                         rng.uniform(0, 2 * math.pi, ndt))
    imp_k = rng.uniform(5.0, 50.0, nit)

    bt = rng.integers(0, nbt, n_poly - 1)
    at = rng.integers(0, nat, n_poly - 2)
    r_act = bond_r0[bt] + rng.normal(scale=0.03, size=n_poly - 1)
    t_act = ang_t0[at] + np.deg2rad(rng.normal(scale=3.0, size=n_poly - 2))
    poly = _grow_chain(rng, n_poly, box, r_act, t_act, min_dist=2.6)

    # solvent
    n_lat = int(round(box / water_spacing))
    sites = _water_sites(rng, box, n_lat, jitter)
    pgrid = _Grid(box, 3.0)
    for i, p in enumerate(poly):
        pgrid.add(i, p)
    keep = []
    for s_i, s in enumerate(sites):
        if np.any(s < 0.2) or np.any(s > box - 0.2):
            continue
        if any(np.sum((q - s) ** 2) < solvent_gap ** 2 for _, q in pgrid.near(s, solvent_gap)):
            continue
        keep.append(s_i)
    keep = np.array(keep)
    if keep.size > n_water:
        keep = np.sort(rng.permutation(keep)[:n_water])
    nw = int(keep.size)

    n = n_poly + 3 * nw
    coords = np.zeros((n, 3))
    coords[:n_poly] = poly
    for w, s in enumerate(keep):
        o = sites[s]
        coords[n_poly + 3 * w] = o
        coords[n_poly + 3 * w + 1:n_poly + 3 * w + 3] = _water_hydrogens(rng, o)

    sig = rng.uniform(1.8, 4.0, n_types)
    eps = rng.uniform(0.02, 0.2, n_types)
    ptype = rng.integers(1, n_types + 1, n_poly)
    q = rng.normal(scale=0.3, size=n_poly)
    q -= q.mean()
    o_t, h_t = n_types + 1, n_types + 2
    charge = np.concatenate([q, np.tile([-0.834, 0.417, 0.417], nw)])
    atom_type = np.concatenate([ptype, np.tile([o_t, h_t, h_t], nw)])
    mass = np.concatenate([rng.uniform(10.0, 16.0, n_poly), np.tile([15.999, 1.008, 1.008], nw)])
    anum = np.concatenate([np.full(n_poly, 6), np.tile([8, 1, 1], nw)])
    n_res_poly = (n_poly + 9) // 10
    top = Topology(
        title=f"solvated polymer seed {seed}", coords=coords, box=box,
        charge_e=charge, atom_type=atom_type, mass=mass, atomic_number=anum,
        names=[f"C{i % 10}" for i in range(n_poly)] + ["O", "H1", "H2"] * nw,
        amber_types=[f"T{t}" for t in ptype] + ["OW", "HW", "HW"] * nw,
        type_sigma=np.concatenate([sig, [O_SIGMA, 0.0]]),
        type_eps=np.concatenate([eps, [O_EPS, 0.0]]),
        res_labels=["PLY"] * n_res_poly + ["WAT"] * nw,
        res_starts=list(range(0, n_poly, 10)) + list(range(n_poly, n, 3)),
        mol_sizes=[n_poly] + [3] * nw)
    top.exclusions = [[] for _ in range(n)]

    # water parameters are the last bond/angle types
    top.bond_k = list(bond_k) + [553.0]
    top.bond_r0 = list(bond_r0) + [WATER_R_OH]
    top.angle_k = list(ang_k) + [100.0]
    top.angle_t0 = list(ang_t0) + [WATER_THETA]
    for i in range(n_poly - 1):
        top.bonds.append((i, i + 1, int(bt[i]), False))
    for i in range(n_poly - 2):
        top.angles.append((i, i + 1, i + 2, int(at[i]), False))
    # Urey-Bradley on every angle, equilibrium near the generated 1-3 distance
    for i in range(n_poly - 2):
        d13 = float(np.linalg.norm(poly[i + 2] - poly[i]))
        top.ubs.append((i, i + 2, len(top.ub_k)))
        top.ub_k.append(float(rng.uniform(5.0, 40.0)))
        top.ub_r0.append(d13 + float(rng.normal(scale=0.05)))
    # cosine dihedrals: one to three terms per quadruple; repeats carry no 1-4
    top.dih_k, top.dih_n, top.dih_phase = list(dih_k), list(dih_n), list(dih_phase)
    top.dih_scee = [1.2] * ndt
    top.dih_scnb = [2.0] * ndt
    for i in range(n_poly - 3):
        nterm = int(rng.integers(1, 4))
        types = rng.choice(ndt, size=nterm, replace=False)
        for t_i, t in enumerate(types):
            top.dihedrals.append((i, i + 1, i + 2, i + 3, int(t), False, t_i > 0, False))
    # Amber-style impropers stored as cosine dihedrals (negative 4th index)
    for i in range(1, n_poly - 3, 24):
        top.dihedrals.append((i, i + 1, i + 2, i + 3, int(rng.integers(ndt)), False, True, True))
    # CHARMM harmonic impropers near the generated geometry
    for i in range(0, n_poly - 3, 4):
        a, b, c, d = poly[i], poly[i + 1], poly[i + 2], poly[i + 3]
        psi = _dihedral(a, b, c, d)
        top.impropers.append((i, i + 1, i + 2, i + 3, len(top.imp_k)))
        top.imp_k.append(float(imp_k[len(top.imp_k) % nit]))
        top.imp_psi0.append(psi + float(np.deg2rad(rng.normal(scale=5.0))))
    # CMAP on phi/psi pairs spread along the chain
    starts = np.linspace(0, n_poly - 5, n_cmap).astype(int)
    for s in starts:
        top.cmaps.append((int(s), int(s) + 1, int(s) + 2, int(s) + 3, int(s) + 4,
                          int(rng.integers(1, 6))))
    top.cmap_res = 24
    top.cmap_grids = _five_cmap_grids(rng, 24)
    # exclusions: 1-2, 1-3, 1-4 along the chain
    for i in range(n_poly):
        top.exclusions[i] = [j for j in (i + 1, i + 2, i + 3) if j < n_poly]
    wb, wa = len(top.bond_k) - 1, len(top.angle_k) - 1
    for w in range(nw):
        f = n_poly + 3 * w
        top.bonds += [(f, f + 1, wb, True), (f, f + 2, wb, True)]
        top.angles += [(f + 1, f, f + 2, wa, True)]
        top.exclusions[f] = [f + 1, f + 2]
        top.exclusions[f + 1] = [f + 2]
        top.exclusions[f + 2] = []
    return top


def _dihedral(a, b, c, d):
    b1, b2, b3 = b - a, c - b, d - c
    n1, n2 = np.cross(b1, b2), np.cross(b2, b3)
    x = np.dot(n1, n2)
    y = np.dot(b1, n2) * np.linalg.norm(b2)
    return math.atan2(y, x)


def small_polymer(seed: int = 5, box: float = 30.0, n_poly: int = 300,
                  n_water: int = 600, n_cmap: int = 20, n_types: int = 8) -> Topology:
    """A reduced solvated polymer for fast parity tests (~2 k atoms)."""
    return solvated_polymer(seed=seed, box=box, n_poly=n_poly, n_water=n_water,
                            n_cmap=n_cmap, n_types=n_types)


def scale_up(seed: int = 33) -> Topology:
    """100 A cube, 16,000-atom polymer plus ~28,000 waters (~100 k atoms)."""
    return solvated_polymer(seed=seed, box=100.0, n_poly=16000, n_water=28000, n_cmap=1000)


# --------------------------------------------------------------------------
# prmtop / rst7 writer
# --------------------------------------------------------------------------

def _fmt_ints(vals, per=10, width=8):
    vals = [int(v) for v in vals]
    lines = []
    for k in range(0, len(vals), per):
        lines.append("".join(f"{v:{width}d}" for v in vals[k:k + per]))
    return lines if lines else [""]


def _fmt_reals(vals, per=3, width=26):
    vals = [float(v) for v in vals]
    lines = []
    for k in range(0, len(vals), per):
        lines.append("".join(f"{v:{width}.17E}" for v in vals[k:k + per]))
    return lines if lines else [""]


def _fmt_a4(vals, per=20):
    lines = []
    for k in range(0, len(vals), per):
        lines.append("".join(f"{str(v)[:4]:<4s}" for v in vals[k:k + per]))
    return lines if lines else [""]


def _ico(t):
    """NONBONDED_PARM_INDEX (1-based), row-major over (i, j) types."""
    out = []
    for i in range(t):
        for j in range(t):
            a, b = min(i, j), max(i, j)
            out.append(b * (b + 1) // 2 + a + 1)
    return out


def prmtop_text(top: Topology) -> str:
    n = top.n_atoms
    t = len(top.type_sigma)
    lja, ljb = _lj_pair_tables(top.type_sigma, top.type_eps)
    bonds_h = [b for b in top.bonds if b[3]]
    bonds_a = [b for b in top.bonds if not b[3]]
    ang_h = [a for a in top.angles if a[4]]
    ang_a = [a for a in top.angles if not a[4]]
    dih_h = [d for d in top.dihedrals if d[5]]
    dih_a = [d for d in top.dihedrals if not d[5]]
    excl_counts, excl_list = [], []
    for i in range(n):
        e = top.exclusions[i]
        if e:
            excl_counts.append(len(e))
            excl_list += [j + 1 for j in e]
        else:
            excl_counts.append(1)
            excl_list.append(0)
    nres = len(top.res_labels)
    pointers = [n, t, len(bonds_h), len(bonds_a), len(ang_h), len(ang_a), len(dih_h), len(dih_a),
                0, 0, len(excl_list), nres, len(bonds_a), len(ang_a), len(dih_a),
                len(top.bond_k), len(top.angle_k), len(top.dih_k), t, 0, 0, 0, 0, 0, 0, 0, 0,
                1, max(np.diff(top.res_starts + [n])) if nres else 0, 0, 0]

    def sec(name, fmt, lines):
        return [f"%FLAG {name}", f"%FORMAT({fmt})"] + lines

    def dih_entries(ds):
        v = []
        for (i, j, k, l, ty, _h, no14, imp) in ds:
            v += [3 * i, 3 * j, -3 * k if no14 else 3 * k, -3 * l if imp else 3 * l, ty + 1]
        return v

    out = ["%VERSION  VERSION_STAMP = V0001.000  DATE = 01/01/26  00:00:00"]
    out += sec("TITLE", "20a4", [top.title[:80]])
    out += sec("POINTERS", "10I8", _fmt_ints(pointers))
    out += sec("ATOM_NAME", "20a4", _fmt_a4(top.names))
    out += sec("CHARGE", "3E26.17", _fmt_reals(np.asarray(top.charge_e) * AMBER_CHARGE_SCALE))
    out += sec("ATOMIC_NUMBER", "10I8", _fmt_ints(top.atomic_number))
    out += sec("MASS", "3E26.17", _fmt_reals(top.mass))
    out += sec("ATOM_TYPE_INDEX", "10I8", _fmt_ints(top.atom_type))
    out += sec("NUMBER_EXCLUDED_ATOMS", "10I8", _fmt_ints(excl_counts))
    out += sec("NONBONDED_PARM_INDEX", "10I8", _fmt_ints(_ico(t)))
    out += sec("RESIDUE_LABEL", "20a4", _fmt_a4(top.res_labels))
    out += sec("RESIDUE_POINTER", "10I8", _fmt_ints([s + 1 for s in top.res_starts]))
    out += sec("BOND_FORCE_CONSTANT", "3E26.17", _fmt_reals(top.bond_k))
    out += sec("BOND_EQUIL_VALUE", "3E26.17", _fmt_reals(top.bond_r0))
    out += sec("ANGLE_FORCE_CONSTANT", "3E26.17", _fmt_reals(top.angle_k))
    out += sec("ANGLE_EQUIL_VALUE", "3E26.17", _fmt_reals(top.angle_t0))
    out += sec("CHARMM_UREY_BRADLEY_COUNT", "2I8", _fmt_ints([len(top.ubs), len(top.ub_k)]))
    out += sec("CHARMM_UREY_BRADLEY", "10I8",
               _fmt_ints([v for (i, j, ty) in top.ubs for v in (i + 1, j + 1, ty + 1)]))
    out += sec("CHARMM_UREY_BRADLEY_FORCE_CONSTANT", "3E26.17", _fmt_reals(top.ub_k))
    out += sec("CHARMM_UREY_BRADLEY_EQUIL_VALUE", "3E26.17", _fmt_reals(top.ub_r0))
    out += sec("DIHEDRAL_FORCE_CONSTANT", "3E26.17", _fmt_reals(top.dih_k))
    out += sec("DIHEDRAL_PERIODICITY", "3E26.17", _fmt_reals(top.dih_n))
    out += sec("DIHEDRAL_PHASE", "3E26.17", _fmt_reals(top.dih_phase))
    out += sec("SCEE_SCALE_FACTOR", "3E26.17", _fmt_reals(top.dih_scee))
    out += sec("SCNB_SCALE_FACTOR", "3E26.17", _fmt_reals(top.dih_scnb))
    out += sec("CHARMM_NUM_IMPROPERS", "10I8", _fmt_ints([len(top.impropers)]))
    out += sec("CHARMM_IMPROPERS", "10I8",
               _fmt_ints([v for (i, j, k, l, ty) in top.impropers
                          for v in (i + 1, j + 1, k + 1, l + 1, ty + 1)]))
    out += sec("CHARMM_NUM_IMPR_TYPES", "I8", _fmt_ints([len(top.imp_k)]))
    out += sec("CHARMM_IMPROPER_FORCE_CONSTANT", "3E26.17", _fmt_reals(top.imp_k))
    out += sec("CHARMM_IMPROPER_PHASE", "3E26.17", _fmt_reals(top.imp_psi0))
    out += sec("LENNARD_JONES_ACOEF", "3E26.17", _fmt_reals(lja))
    out += sec("LENNARD_JONES_BCOEF", "3E26.17", _fmt_reals(ljb))
    out += sec("LENNARD_JONES_14_ACOEF", "3E26.17", _fmt_reals(lja * 0.5))
    out += sec("LENNARD_JONES_14_BCOEF", "3E26.17", _fmt_reals(ljb * 0.5))
    out += sec("BONDS_INC_HYDROGEN", "10I8",
               _fmt_ints([v for (i, j, ty, _h) in bonds_h for v in (3 * i, 3 * j, ty + 1)]))
    out += sec("BONDS_WITHOUT_HYDROGEN", "10I8",
               _fmt_ints([v for (i, j, ty, _h) in bonds_a for v in (3 * i, 3 * j, ty + 1)]))
    out += sec("ANGLES_INC_HYDROGEN", "10I8",
               _fmt_ints([v for (i, j, k, ty, _h) in ang_h for v in (3 * i, 3 * j, 3 * k, ty + 1)]))
    out += sec("ANGLES_WITHOUT_HYDROGEN", "10I8",
               _fmt_ints([v for (i, j, k, ty, _h) in ang_a for v in (3 * i, 3 * j, 3 * k, ty + 1)]))
    out += sec("DIHEDRALS_INC_HYDROGEN", "10I8", _fmt_ints(dih_entries(dih_h)))
    out += sec("DIHEDRALS_WITHOUT_HYDROGEN", "10I8", _fmt_ints(dih_entries(dih_a)))
    out += sec("EXCLUDED_ATOMS_LIST", "10I8", _fmt_ints(excl_list))
    out += sec("AMBER_ATOM_TYPE", "20a4", _fmt_a4(top.amber_types))
    out += sec("SOLVENT_POINTERS", "3I8", _fmt_ints([nres, len(top.mol_sizes), 1]))
    out += sec("ATOMS_PER_MOLECULE", "10I8", _fmt_ints(top.mol_sizes))
    out += sec("BOX_DIMENSIONS", "4E26.17", _fmt_reals([90.0, top.box, top.box, top.box], per=4))
    out += sec("CHARMM_CMAP_COUNT", "2I8", _fmt_ints([len(top.cmaps), 5]))
    out += sec("CHARMM_CMAP_RESOLUTION", "20I4", _fmt_ints([top.cmap_res] * 5, per=20, width=4))
    for s in range(5):
        out += sec(f"CHARMM_CMAP_PARAMETER_{s + 1:02d}", "8(F9.5)",
                   _fmt_reals(top.cmap_grids[s]))
    out += sec("CHARMM_CMAP_INDEX", "6I8",
               _fmt_ints([v for c in top.cmaps for v in (c[0] + 1, c[1] + 1, c[2] + 1, c[3] + 1,
                                                          c[4] + 1, c[5])], per=6))
    return "\n".join(out) + "\n"


def rst7_text(top: Topology) -> str:
    c = top.coords.reshape(-1)
    if np.any(c <= -100.0) or np.any(c >= 1000.0):
        raise ValueError("coordinates outside the range that keeps 12.7 fields separated")
    lines = [top.title[:80], f"{top.n_atoms:6d}"]
    for k in range(0, c.size, 6):
        lines.append("".join(f"{v:12.7f}" for v in c[k:k + 6]))
    b = top.box
    lines.append("".join(f"{v:12.7f}" for v in (b, b, b, 90.0, 90.0, 90.0)))
    return "\n".join(lines) + "\n"


def write_system(top: Topology, prefix: str) -> tuple:
    """Write ``prefix.prmtop`` and ``prefix.rst7``; returns both paths."""
    d = os.path.dirname(prefix)
    if d:
        os.makedirs(d, exist_ok=True)
    p, r = prefix + ".prmtop", prefix + ".rst7"
    with open(p, "w") as f:
        f.write(prmtop_text(top))
    with open(r, "w") as f:
        f.write(rst7_text(top))
    return p, r


WORKLOADS = {
    "water_box_6k": lambda: water_box(seed=11),
    "polymer_29k": lambda: solvated_polymer(seed=22),
    "polymer_2k": lambda: small_polymer(seed=5),
    "scale_100k": lambda: scale_up(seed=33),
}


def cached_system(name: str, cache_dir: str | None = None) -> tuple:
    """Generate (once) and return the prmtop/rst7 paths of a named workload."""
    cache_dir = cache_dir or os.environ.get("MD_AMD_CACHE",
                                            os.path.join(os.path.dirname(__file__), "..", "_cache"))
    prefix = os.path.join(cache_dir, name)
    if not (os.path.exists(prefix + ".prmtop") and os.path.exists(prefix + ".rst7")):
        write_system(WORKLOADS[name](), prefix)
    return prefix + ".prmtop", prefix + ".rst7"


# --------------------------------------------------------------------------
# many water boxes as arrays (batched workloads)
# --------------------------------------------------------------------------

def smallest_2a3b(n: int) -> int:
    """Smallest 2^a 3^b >= n."""
    best = None
    p2 = 1
    while p2 < 2 * n:
        p = p2
        while p < 3 * n:
            if p >= n and (best is None or p < best):
                best = p
            p *= 3
        p2 *= 2
    return best


def _rotations(rng, m):
    """m uniformly random rotation matrices (from unit quaternions)."""
    q = rng.normal(size=(m, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], 1)


def water_arrays(seed: int, n_atoms: int, density: float = 0.1, jitter: float = 0.3):
    """A water box of about ``n_atoms`` atoms as system arrays (AMBER
    semantics, the layout md_amd.native.desc_from_arrays takes), without a
    prmtop round trip: the water model, parameters and placement rules of
    ``water_box`` (jittered lattice sites, H at 0.9572 A and 104.52 deg with a
    random orientation), vectorised.  Returns (arrays, coords (N,3), box L)."""
    rng = np.random.default_rng(seed)
    n_water = max(1, n_atoms // 3)
    n = 3 * n_water
    box = (n / density) ** (1.0 / 3.0)
    n_lat = int(math.ceil(n_water ** (1.0 / 3.0)))
    sites = _water_sites(rng, box, n_lat, jitter)
    pick = np.sort(rng.permutation(sites.shape[0])[:n_water])
    o = sites[pick]
    s, c = math.sin(WATER_THETA / 2), math.cos(WATER_THETA / 2)
    local = np.array([[WATER_R_OH * s, 0.0, WATER_R_OH * c], [-WATER_R_OH * s, 0.0, WATER_R_OH * c]])
    rot = _rotations(rng, n_water)
    h = o[:, None, :] + np.einsum("wij,kj->wki", rot, local)
    coords = np.concatenate([o[:, None, :], h], axis=1).reshape(n, 3)
    qscale = AMBER_CHARGE_SCALE / 18.2206899283247
    a_oo = 4.0 * O_EPS * O_SIGMA ** 12
    b_oo = 4.0 * O_EPS * O_SIGMA ** 6
    base = 3 * np.arange(n_water, dtype=np.int32)
    oh1 = np.stack([base, base + 1], 1)
    oh2 = np.stack([base, base + 2], 1)
    bond_atoms = np.stack([oh1, oh2], 1).reshape(-1)
    angle_atoms = np.stack([base + 1, base, base + 2], 1).reshape(-1)
    excl_list = np.stack([base + 1, base + 2, base, base + 2, base, base + 1], 1).reshape(-1).astype(np.int32)
    excl_pairs = np.stack([base, base + 1, base, base + 2, base + 1, base + 2], 1).reshape(-1).astype(np.int32)
    empty_i, empty_d = np.zeros(0, np.int32), np.zeros(0)
    arrays = {
        "charge": np.tile([-0.834 * qscale, 0.417 * qscale, 0.417 * qscale], n_water),
        "type": np.tile(np.array([0, 1, 1], np.int32), n_water), "n_types": 2,
        "nb_index": np.array([1, 2, 2, 3], np.int32),
        "lj_a": np.array([a_oo, 0.0, 0.0]), "lj_b": np.array([b_oo, 0.0, 0.0]),
        "lj14_a": np.array([a_oo / 2, 0.0, 0.0]), "lj14_b": np.array([b_oo / 2, 0.0, 0.0]),
        "bond_atoms": bond_atoms.astype(np.int32), "bond_type": np.zeros(2 * n_water, np.int32),
        "bond_k": np.array([553.0]), "bond_r0": np.array([WATER_R_OH]),
        "ub_atoms": empty_i, "ub_type": empty_i, "ub_k": empty_d, "ub_r0": empty_d,
        "angle_atoms": angle_atoms.astype(np.int32), "angle_type": np.zeros(n_water, np.int32),
        "angle_k": np.array([100.0]), "angle_theta0": np.array([WATER_THETA]),
        "dihedral_atoms": empty_i, "dihedral_type": empty_i, "dihedral_k": empty_d, "dihedral_n": empty_d,
        "dihedral_phase": empty_d,
        "improper_atoms": empty_i, "improper_type": empty_i, "improper_k": empty_d, "improper_psi0": empty_d,
        "cmap_atoms": empty_i, "cmap_set": empty_i, "cmap_coeff": empty_d, "cmap_grid": empty_d,
        "cmap_n_sets": 0, "cmap_resolution": 0,
        "excl_start": (2 * np.arange(n + 1)).astype(np.int32), "excl_list": excl_list,
        "excl_pairs": excl_pairs, "pairs14": empty_i, "scee14": empty_d,
    }
    return arrays, coords, box


def batch_workload(n_systems: int = 256, seed0: int = 4000, lo: int = 3000, hi: int = 30000):
    """Uneven water boxes: atom counts uniform in [lo, hi] (seed 4000 + i for
    system i), cubic boxes at 0.1 atoms/A^3, PME grid = smallest 2^a 3^b >= L/1 A.
    Returns a list of (arrays, coords, box, grid)."""
    out = []
    for i in range(n_systems):
        rng = np.random.default_rng(seed0 + i)
        n = int(rng.integers(lo, hi + 1))
        arrays, coords, box = water_arrays(seed0 + i, n)
        k = smallest_2a3b(int(math.ceil(box / 1.0)))
        out.append((arrays, coords, box, k))
    return out
