"""Small systems for the molecular-dynamics tests."""
import numpy as np

from md_amd import native as N


def _empty():
    ei, ed = np.zeros(0, np.int32), np.zeros(0)
    return {
        "ub_atoms": ei, "ub_type": ei, "ub_k": ed, "ub_r0": ed,
        "angle_atoms": ei, "angle_type": ei, "angle_k": ed, "angle_theta0": ed,
        "dihedral_atoms": ei, "dihedral_type": ei, "dihedral_k": ed, "dihedral_n": ed, "dihedral_phase": ed,
        "improper_atoms": ei, "improper_type": ei, "improper_k": ed, "improper_psi0": ed,
        "cmap_atoms": ei, "cmap_set": ei, "cmap_coeff": ed, "cmap_grid": ed, "cmap_n_sets": 0,
        "cmap_resolution": 0, "pairs14": ei, "scee14": ed,
    }


def harmonic_pair(k=300.0, r0=1.5, stretch=0.1, box=30.0, m=(12.0, 16.0)):
    """Two uncharged atoms without LJ, joined by one bond E = k (r - r0)^2,
    stretched by `stretch` along x.  Returns (desc, keep, coords, box, masses,
    omega (1/ps)) where omega^2 = 2 k 418.4 / mu."""
    arr = _empty()
    arr.update({
        "charge": np.zeros(2), "type": np.zeros(2, np.int32), "n_types": 1, "nb_index": np.array([1], np.int32),
        "lj_a": np.zeros(1), "lj_b": np.zeros(1), "lj14_a": np.zeros(1), "lj14_b": np.zeros(1),
        "bond_atoms": np.array([0, 1], np.int32), "bond_type": np.zeros(1, np.int32),
        "bond_k": np.array([k]), "bond_r0": np.array([r0]),
        "excl_start": np.array([0, 1, 2], np.int32), "excl_list": np.array([1, 0], np.int32),
        "excl_pairs": np.array([0, 1], np.int32),
    })
    keep = []
    desc = N.desc_from_arrays(arr, keep)
    keep.append(arr)
    c = 0.5 * box
    x = np.array([[c, c, c], [c + r0 + stretch, c, c]])
    masses = np.array(m, dtype=np.float64)
    mu = masses[0] * masses[1] / masses.sum()
    omega = np.sqrt(2.0 * k * 418.4 / mu)
    return desc, keep, x, np.array([box, box, box]), masses, omega


def oscillator_params(precision=N.PREC_FP64):
    return N.make_params(N.SEM_AMBER, precision, lj_cutoff=9.0, lj_switch=8.0, ewald_tol=1e-5, pme_grid=(0, 0, 0))


def discrete_stretch(omega, dt, nsteps, amp):
    """Velocity Verlet's exact solution for a harmonic coordinate released at
    rest: r_n - r0 = A cos(n theta), cos theta = 1 - (omega dt)^2 / 2."""
    theta = np.arccos(1.0 - 0.5 * (omega * dt) ** 2)
    return amp * np.cos(theta * np.arange(1, nsteps + 1))


def thermal_velocities(masses, seed, temperature=300.0):
    """Maxwell-Boltzmann velocities (A/ps) at `temperature` K."""
    kT = 0.0019872041 * temperature  # kcal/mol
    sig = np.sqrt(kT * 418.4 / np.asarray(masses, dtype=np.float64))
    rng = np.random.default_rng(seed)
    return rng.normal(size=(sig.size, 3)) * sig[:, None]


def bonded_chain(n=12, seed=2, box=40.0):
    """An uncharged chain without LJ: bonds k 300, r0 1.5 and angles k 50,
    theta0 1.9 rad, from a perturbed zig-zag.  Its minimum energy is 0 (every
    bond and angle at rest).  Returns (desc, keep, coords, box)."""
    rng = np.random.default_rng(seed)
    arr = _empty()
    bonds = np.array([[i, i + 1] for i in range(n - 1)], np.int32).reshape(-1)
    angles = np.array([[i, i + 1, i + 2] for i in range(n - 2)], np.int32).reshape(-1)
    arr.update({
        "charge": np.zeros(n), "type": np.zeros(n, np.int32), "n_types": 1, "nb_index": np.array([1], np.int32),
        "lj_a": np.zeros(1), "lj_b": np.zeros(1), "lj14_a": np.zeros(1), "lj14_b": np.zeros(1),
        "bond_atoms": bonds, "bond_type": np.zeros(n - 1, np.int32), "bond_k": np.array([300.0]),
        "bond_r0": np.array([1.5]),
        "angle_atoms": angles, "angle_type": np.zeros(n - 2, np.int32), "angle_k": np.array([50.0]),
        "angle_theta0": np.array([1.9]),
        "excl_start": np.zeros(n + 1, np.int32), "excl_list": np.zeros(0, np.int32),
        "excl_pairs": np.zeros(0, np.int32),
    })
    keep = []
    desc = N.desc_from_arrays(arr, keep)
    keep.append(arr)
    x = np.zeros((n, 3))
    for i in range(n):
        x[i] = [0.5 * box - 8.0 + 1.25 * i, 0.5 * box + (0.8 if i % 2 else 0.0), 0.5 * box]
    x += rng.normal(0.0, 0.15, x.shape)
    return desc, keep, x, np.array([box, box, box])
