"""Benchmark: energy + force evaluations per second at ~29k atoms.

Workload (one GPU): the 66 A solvated heteropolymer (4,800-atom polymer in
~8,000 three-site waters, 28,800 atoms, 30 atom types, angles + Urey-Bradley,
cosine dihedrals, impropers, 300 CMAP terms), AMBER semantics, 12 A cutoff
(LJ switched from 11 A), Ewald tolerance 1e-5, smooth PME on a 72^3 grid with
order 4.  A step is one full evaluation (spatial sort, cluster-pair list,
nonbonded, bonded, corrections, PME, reductions) with coordinates and forces
resident in device memory.

Reported: the mixed form's evaluations/s (value), the FP64 form's, their
ratio, the force relative RMS error of the mixed form against the FP64 form,
the roofline of the dominant (pair) kernel from HIP-event timings, and the
CPU oracle timed on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic work per unit
PAIR_FLOP = 77            # FP32 flops per atom pair inside the cutoff
FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak
HBM_PEAK_GBS = 8000.0

WORKLOAD = {
    "workload": "polymer_29k",
    "description": "66 A cube, 4,800-atom random-walk heteropolymer + ~8,000 TIP3P-like waters "
                   "(28,800 atoms), AMBER semantics, rc 12 A (LJ switch 11 A), Ewald tol 1e-5, "
                   "PME 72^3 order 4",
    "atoms": 28800,
    "cutoff_A": 12.0,
    "pme_grid": 72,
}


def read_traffic(profile_dir):
    """HBM bytes per launch of the pair kernel from the committed PMC summary."""
    path = os.path.join(profile_dir, "nonbonded_polymer29k_pmc.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def copy_calibration():
    import torch
    n = 64 * 1024 * 1024  # float32 elements: 256 MiB read + 256 MiB written per copy
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 3}), flush=True)


def pme_roofline(kernels_ms, n_at, kg):
    """HBM roofline of the PME kernel groups: algorithmic bytes per launch
    (FP32 form) / HIP-event time of the group."""
    G, Gc = kg ** 3, kg * kg * (kg // 2 + 1)
    pme_bytes = {
        "spread_z_r2c": (16 * n_at + 8 * Gc, "16N (x, y, z, q) read + 8 Gc complex z-transformed grid written"),
        "fft_y_y_inverse_z_c2r": (2 * 16 * Gc + 8 * Gc + 4 * G,
                                  "y pass and inverse y pass 16 Gc each, z c2r 8 Gc read + 4 G written"),
        "x_conv": (16 * Gc, "x pass + influence function + inverse x pass: 8 Gc read + 8 Gc written"),
        "gather": (16 * n_at + 4 * G + 12 * n_at, "16N + 4G read, 12N written"),
    }
    group = {"spread_z_r2c": "spread", "fft_y_y_inverse_z_c2r": "fft", "x_conv": "conv", "gather": "gather"}
    out = {}
    for k, (nbytes, what) in pme_bytes.items():
        ms = kernels_ms.get(group[k])
        if ms:
            gbs = nbytes / (ms * 1e-3) / 1e9
            out[k] = {"bound": "hbm", "bytes": nbytes, "ms": ms, "achieved": gbs, "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS, "algorithmic": what}
    return out


def kernel_breakdown(torch, ev, xp, fp, box, n):
    """Per-kernel-group HIP-event times (ms per launch) over n direct-launch
    evaluations (timing mode 1: twenty events per evaluation, so a pass of its
    own, apart from the throughput figure)."""
    ev.set_timing(1)
    for _ in range(n):
        ev.evaluate_device(xp, box, fp)
    torch.cuda.synchronize()
    tim = ev.timing()
    ev.set_timing(0)
    return {k: round(v[0] / max(v[1], 1), 5) for k, v in tim.items() if v[1] > 0}


def timed_steps(torch, ev, xp, fp, box, steps, warmup):
    """Seconds for `steps` evaluations (the time-step form of the device-buffer
    call: no Python objects per step) after `warmup` untimed ones."""
    ev.set_box(box)
    for _ in range(warmup):
        ev.step_device(xp, fp)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ev.step_device(xp, fp)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def host_description():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def cpu_baseline(N, prm, rst, kw):
    """The reference's CPU evaluation restated in C (oracle/oracle.c, REF
    semantics: the reference's term lists, empty exclusion lists, pair list at
    cutoff + 2 A, the six pair-list passes of System::calculate_energies /
    calculate_forces and, as written, the O(N^2) plain-Coulomb energy pass of
    calculate_EE_energy), single-threaded like the reference, built
    -O3 -march=x86-64-v3 (oracle/liboracle_fast.so), timed on this host on
    the benchmark's system and cutoffs (rc 12 A, switch 11 A; REF forms no
    reciprocal space).  value: the pair-list-only evaluation rate (median of
    2 after one warm-up); as written: one evaluation after the warm-ups; the
    AMBER-semantics evaluation (the GPU's work, PME 72^3 by a direct DFT in
    the restatement) as a labelled extra.  About 20 s on a current server
    core."""
    import oracle
    top_ref = N.Topology(prm, rst, N.SEM_REF)
    p_ref = N.make_params(N.SEM_REF, N.PREC_FP64, lj_cutoff=kw["lj_cutoff"], lj_switch=kw["lj_switch"],
                          ewald_tol=kw["ewald_tol"])
    oracle.timed_evaluate(top_ref.desc, p_ref, top_ref.coords, top_ref.box, skip_all_pairs=True)  # warm-up
    t_list = sorted(oracle.timed_evaluate(top_ref.desc, p_ref, top_ref.coords, top_ref.box, skip_all_pairs=True)
                    for _ in range(2))
    t_all = oracle.timed_evaluate(top_ref.desc, p_ref, top_ref.coords, top_ref.box, skip_all_pairs=False)
    top_a = N.Topology(prm, rst, N.SEM_AMBER)
    p_a = N.make_params(N.SEM_AMBER, N.PREC_FP64, **kw)
    t_amb = sorted(oracle.timed_evaluate(top_a.desc, p_a, top_a.coords, top_a.box, skip_all_pairs=True)
                   for _ in range(2))
    top_ref.close()
    top_a.close()
    med = 0.5 * (t_list[0] + t_list[1])
    out = {"value": 1.0 / med, "unit": "evals/s", "cores": 1, "kind": "port",
           "sample": "REF-semantics evaluations of polymer_29k (rc 12 A, switch 11 A, list at rc + 2 A): 1 warm-up "
                     "+ 2 timed pair-list-only, 1 as written (with the O(N^2) EE pass), 2 AMBER (PME 72^3)",
           "ref_pair_list_only_evals_per_s": 1.0 / med,
           "ref_as_written_evals_per_s": 1.0 / t_all,
           "amber_evals_per_s": 2.0 / (t_amb[0] + t_amb[1]),
           "threads": 1, "build": oracle.FAST_FLAGS}
    out.update(host_description())
    return out


def bench_batched(torch, N, systems, n_systems=256, reps=3):
    """256 uneven water boxes (3k-30k atoms, AMBER, rc 9 A, switch 8 A, Ewald tol
    1e-5, PME grid = smallest 2^a 3^b >= L / 1 A), mixed form: systems/s of one
    graph launch, of the same launches on the same streams issued directly,
    and of one synchronous evaluation per system."""
    t0 = time.perf_counter()
    work = systems.batch_workload(n_systems)
    keep = []
    batch = N.Batch(0)
    for arrays, coords, box, k in work:
        prm = N.make_params(N.SEM_AMBER, N.PREC_MIXED, lj_cutoff=9.0, lj_switch=8.0, ewald_tol=1e-5,
                            pme_grid=(k, k, k))
        batch.add(N.desc_from_arrays(arrays, keep), prm, box)
    dev = torch.device("cuda:0")
    x = torch.tensor(np.concatenate([w[1] for w in work]).reshape(-1), dtype=torch.float64, device=dev)
    f = torch.zeros_like(x)
    setup_s = time.perf_counter() - t0
    cfg = batch.tune(x.data_ptr(), f.data_ptr(), reps=1)
    batch.configure(mode=1, streams=cfg["streams"], split=cfg["split"], fft_lines=cfg["fft_lines"])
    batch.evaluate_device(x.data_ptr(), f.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.evaluate_device(x.data_ptr(), f.data_ptr())
    torch.cuda.synchronize()
    serial_s = (time.perf_counter() - t0) / reps
    batch.configure(mode=2, streams=cfg["streams"], split=cfg["split"], fft_lines=cfg["fft_lines"])
    batch.evaluate_device(x.data_ptr(), f.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.evaluate_device(x.data_ptr(), f.data_ptr())
    torch.cuda.synchronize()
    streams_s = (time.perf_counter() - t0) / reps
    batch.configure(mode=0, streams=cfg["streams"], split=cfg["split"], fft_lines=cfg["fft_lines"])
    batch.evaluate_device(x.data_ptr(), f.data_ptr())  # sizes lists, captures the graph, first launch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.evaluate_device(x.data_ptr(), f.data_ptr())
    torch.cuda.synchronize()
    graph_s = (time.perf_counter() - t0) / reps
    state = batch.graph_state()
    n_atoms = int(sum(w[1].shape[0] for w in work))
    # the tuner's choice per size class (systems by atom count, four classes of equal count)
    sizes = [w[1].shape[0] for w in work]
    order = sorted(range(n_systems), key=lambda i: sizes[i])
    classes = []
    for k in range(4):
        members = order[k * n_systems // 4:(k + 1) * n_systems // 4]
        if members:
            classes.append({"atoms": [int(sizes[members[0]]), int(sizes[members[-1]])],
                            "split_fft_lines": sorted({batch.system_config(i) for i in members})})
    batch.close()
    return {"systems": n_systems, "atoms": n_atoms, "setup_s": setup_s, "tuned": cfg, "tuned_classes": classes,
            "form": "one hipGraph launch: every system's kernel chain (sort, search, pair + bonded, PME, "
                    "reduction) captured once, dealt largest first over 8 streams, then one gathering kernel",
            "graph_state": state, "tuned_mode": {0: "graph", 2: "direct launches"}.get(cfg["mode"], "serial"),
            "tuned_mode_systems_per_s": n_systems / (graph_s if cfg["mode"] == 0 else streams_s),
            "graph_systems_per_s": n_systems / graph_s, "serial_systems_per_s": n_systems / serial_s,
            "streams_systems_per_s": n_systems / streams_s,
            "graph_ms_per_batch": 1e3 * graph_s, "serial_ms_per_batch": 1e3 * serial_s,
            "streams_ms_per_batch": 1e3 * streams_s,
            "speedup_graph_vs_serial": serial_s / graph_s,
            "graph_atoms_per_s": n_atoms / graph_s,
            "workload": "256 water boxes, atoms U[3000, 30000] (seeds 4000+i), 0.1 atoms/A^3, AMBER, "
                        "rc 9 A, switch 8 A, tol 1e-5, PME 2^a3^b >= L/1A (32^3..72^3), mixed form"}


def bench_md(torch, N, systems, top, p32, box, steps=200, warmup=20, dt=0.0005):
    """Energy minimisation, then NVE molecular dynamics of the same geometry
    (mixed form, velocity Verlet, dt 0.5 fs, 300 K Maxwell-Boltzmann start,
    seed 31), with the pair list rebuilt every step and as Verlet lists of
    skin 0.5 and 1.0 A, and with the full virial (pressure) on a skin of 0.5 A.
    The water hydrogens get the CHARMM TIP3P LJ radius
    (sigma 0.4 A, eps 0.046 kcal/mol): without it a hydrogen and an unlike
    charge have no repulsion and a minimiser can follow a hydrogen into
    another atom.  Each MD leg starts from the minimised coordinates;
    steps/s over `steps` timed steps after `warmup` untimed ones."""
    dev = torch.device("cuda:0")
    arr = systems.with_hydrogen_lj(top.arrays, int(top.arrays["type"][-1]))
    keep = []
    desc = N.desc_from_arrays(arr, keep)
    ev = N.Evaluator(0)
    ev.upload(desc)
    ev.set_params(p32)
    ev.set_verlet(1.0)
    x = torch.tensor(top.coords.reshape(-1), dtype=torch.float64, device=dev)
    f = torch.zeros_like(x)
    e0, _ = ev.evaluate_device(x.data_ptr(), box, f.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = ev.minimize(x.data_ptr(), box, max_iter=300, rms_force_tol=1.0, max_step=0.1)
    torch.cuda.synchronize()
    t_min = time.perf_counter() - t0
    xm = x.clone()
    m = np.asarray(top.masses, dtype=np.float64)
    ev.md_set_masses(m)
    kT = 0.0019872041 * 300.0
    v0 = torch.tensor((np.random.default_rng(31).normal(size=(m.size, 3)) *
                       np.sqrt(kT * 418.4 / m)[:, None]).reshape(-1), dtype=torch.float64, device=dev)
    legs = {}
    for skin, vir in ((0.0, False), (0.5, False), (1.0, False), (0.5, True)):
        ev.set_verlet(skin)
        ev.set_virial(vir)
        x.copy_(xm)
        v = v0.clone()
        ev.evaluate_device(x.data_ptr(), box, f.data_ptr())
        ev.md_run(x.data_ptr(), v.data_ptr(), f.data_ptr(), box, dt, warmup)
        b0 = ev.list_builds()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tr, _ = ev.md_run(x.data_ptr(), v.data_ptr(), f.data_ptr(), box, dt, steps)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        b1 = ev.list_builds()
        et = tr[:, 0] + tr[:, 1]
        legs[f"skin_{skin:.1f}" + ("_virial" if vir else "")] = {
            "steps_per_s": steps / t, "ns_per_day": steps / t * dt * 86400.0 / 1000.0,
            "list_builds_per_step": (b1[0] - b0[0]) / steps,
            "total_energy_spread_over_mean_kinetic": float((et.max() - et.min()) / tr[:, 1].mean()),
            "temperature_K": float(2.0 * tr[-1, 1] / (3 * m.size * 0.0019872041)),
        }
        if vir:
            legs[f"skin_{skin:.1f}_virial"]["mean_pressure_bar"] = float(np.mean(tr[:, 2]))
    ev.close()
    return {"workload": "polymer_29k with CHARMM TIP3P hydrogen LJ, mixed form, AMBER, rc 12 A, PME 72^3; "
                        "L-BFGS (max 300 iterations, "
                        "RMS force 1 kcal/mol/A, 0.1 A step cap, skin 1 A), then velocity Verlet NVE, "
                        f"dt {dt * 1000:.1f} fs, 300 K start",
            "minimize": {"energy_before": e0["total"], "energy_after": res["energy"],
                         "iterations": res["iterations"], "evaluations": res["evaluations"],
                         "rms_force": res["rms_force"], "max_force": res["max_force"], "seconds": t_min,
                         "evaluations_per_s": res["evaluations"] / t_min},
            "dynamics": legs}


def bench_triclinic(torch, N, top, p32, box, reps=200, warmup=20):
    """Two triclinic cells of the polymer_29k (mixed form, AMBER, rc 12 A,
    PME 72^3), evaluations/s over `reps` after `warmup` each:
      * `sheared`: its own cubic lattice written with edges a, a + b, a + b + c,
        the same periodic system through the triclinic path, with the relative
        differences of the direct-space energies from the orthorhombic cell;
        the library evaluates it in the reduced basis of the lattice (the
        cube), so its columns stay narrow and the minimum-image search runs;
      * `angles_80_85_75`: the system carried into a 66 A cell of those angles
        (x' = H x / 66), which keeps the minimum-image search."""
    l1, l2, l3 = (float(v) for v in box)
    a, b, c = np.array([l1, 0.0, 0.0]), np.array([l1, l2, 0.0]), np.array([l1, l2, l3])

    def ang(u, v):
        return float(np.degrees(np.arccos(np.dot(u, v) / (np.linalg.norm(u) * np.linalg.norm(v)))))

    lengths = [float(np.linalg.norm(u)) for u in (a, b, c)]
    angles = [ang(b, c), ang(a, c), ang(a, b)]
    dev = torch.device("cuda:0")

    def run(coords, lengths, angles):
        x = torch.tensor(np.asarray(coords).reshape(-1), dtype=torch.float64, device=dev)
        f = torch.zeros_like(x)
        ev = N.Evaluator(0)
        try:
            ev.upload(top.desc)
            ev.set_params(p32)
            e_o, _ = ev.evaluate_device(x.data_ptr(), box, f.data_ptr())
            ev.set_box_angles(angles)
            for _ in range(warmup):
                ev.evaluate_device(x.data_ptr(), lengths, f.data_ptr())
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                e_t, _ = ev.evaluate_device(x.data_ptr(), lengths, f.data_ptr())
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            mode = ev.search_mode()
        finally:
            ev.close()
        return e_o, e_t, {"evals_per_s": reps / t, "cluster_search": "general" if mode else "minimum image"}

    e_o, e_t, sheared = run(top.coords, lengths, angles)
    sheared["cell"] = "edges %.2f, %.2f, %.2f A, angles %.2f, %.2f, %.2f deg" % tuple(lengths + angles)
    sheared["vdw_rel_diff_vs_orthorhombic"] = abs(e_t["vdw"] - e_o["vdw"]) / abs(e_o["vdw"])
    sheared["ewald_direct_rel_diff_vs_orthorhombic"] = \
        abs(e_t["ewald_direct"] - e_o["ewald_direct"]) / abs(e_o["ewald_direct"])
    # x' = H x / L for the cell of edges L and angles 80, 85, 75
    ca, cb, cg = np.cos(np.radians([80.0, 85.0, 75.0]))
    sg = np.sin(np.radians(75.0))
    cy = (ca - cb * cg) / sg
    h = l1 * np.array([[1.0, cg, cb], [0.0, sg, cy], [0.0, 0.0, np.sqrt(1.0 - cb * cb - cy * cy)]])
    xm = (top.coords.reshape(-1, 3) / l1) @ h.T
    _, _, tilted = run(xm, [l1, l1, l1], [80.0, 85.0, 75.0])
    tilted["cell"] = "edges %.2f A, angles 80, 85, 75 deg (x' = H x / %.0f)" % (l1, l1)
    return {"workload": "polymer_29k in two triclinic cells; mixed form, AMBER, rc 12 A, PME 72^3",
            "sheared": sheared, "angles_80_85_75": tilted}


def bench_scale(torch, N, systems, reps=100, warmup=20):
    """The ~100k-atom system (16,000-atom polymer in ~28,000 waters, 100 A
    box; AMBER, rc 10 A, switch 9 A, Ewald tol 1e-5, PME 108^3), mixed form:
    evaluations/s over `reps` graph replays after `warmup`, the pair kernel's
    launch time sampled by event nodes inside that timed region (one
    evaluation in 16), its VALU roofline, and the PME groups' HBM rooflines
    from a per-kernel-group pass."""
    prm, rst = systems.cached_system("scale_100k")
    top = N.Topology(prm, rst, N.SEM_AMBER)
    kg = 108
    p32 = N.make_params(N.SEM_AMBER, N.PREC_MIXED, lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-5,
                        pme_grid=(kg, kg, kg))
    x = torch.tensor(top.coords.reshape(-1), dtype=torch.float64, device="cuda:0")
    f = torch.zeros_like(x)
    ev = N.Evaluator(0)
    try:
        ev.upload(top.desc)
        ev.set_params(p32)
        ev.set_counting(True)
        ev.evaluate_device(x.data_ptr(), top.box, f.data_ptr())
        counts = ev.counts()
        ev.set_counting(False)
        tuned = ev.tune(x.data_ptr(), top.box, f.data_ptr(), reps=6)
        ev.set_timing(3)
        t = timed_steps(torch, ev, x.data_ptr(), f.data_ptr(), top.box, reps, warmup)
        nb_total, nb_launches = ev.timing()["nonbonded"]
        ev.set_timing(0)
        kernels_ms = kernel_breakdown(torch, ev, x.data_ptr(), f.data_ptr(), top.box, 20)
    finally:
        ev.close()
    nb_ms = nb_total / max(nb_launches, 1)
    achieved = PAIR_FLOP * counts["pairs"] / (nb_ms * 1e-3) / 1e12
    return {"workload": "scale_100k: %d atoms, 100 A box, AMBER, rc 10 A (switch 9 A), tol 1e-5, PME 108^3, "
                        "mixed form" % top.n_atoms,
            "evals_per_s": reps / t, "atom_evals_per_s": top.n_atoms * reps / t, "work": counts,
            "tuned_shapes": {"split": tuned[0], "fft_lines": tuned[1] or 16, "tune_median_ms_direct": tuned[2]},
            "kernel_ms": kernels_ms,
            "roofline": {"bound": "valu", "kernel": "pair kernel (mixed, AMBER, switched LJ + erfc Ewald)",
                         "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": achieved / FP32_PEAK_TFLOPS, "avg_launch_ms": nb_ms,
                         "timed_launches": int(nb_launches),
                         "algorithmic": f"{PAIR_FLOP} flop x {counts['pairs']} pairs within 10 A per launch"},
            "pme_roofline": pme_roofline(kernels_ms, top.n_atoms, kg)}


def profile_workload(name, steps, warmup, shapes=None):
    """Only `steps` mixed-form evaluations of one workload after `warmup`
    (graph replay), for profiler runs: polymer_29k or scale_100k, with the
    given (split, fft_lines) launch shapes (no tuning pass, whose candidate
    launches would mix into the profile)."""
    import torch
    from md_amd import native as N
    from md_amd import systems
    if name == "polymer_29k":
        kw = dict(lj_cutoff=12.0, lj_switch=11.0, ewald_tol=1e-5, pme_grid=(72, 72, 72))
    else:
        kw = dict(lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-5, pme_grid=(108, 108, 108))
    prm, rst = systems.cached_system(name)
    top = N.Topology(prm, rst, N.SEM_AMBER)
    x = torch.tensor(top.coords.reshape(-1), dtype=torch.float64, device="cuda:0")
    f = torch.zeros_like(x)
    ev = N.Evaluator(0)
    ev.upload(top.desc)
    ev.set_params(N.make_params(N.SEM_AMBER, N.PREC_MIXED, **kw))
    if shapes:
        ev.set_shapes(*shapes)
    t = timed_steps(torch, ev, x.data_ptr(), f.data_ptr(), top.box, steps, warmup)
    print(json.dumps({"profile_workload": name, "atoms": top.n_atoms, "steps": steps, "warmup": warmup,
                      "shapes": ev.shapes(), "evals_per_s": steps / t}), flush=True)
    ev.close()


def bench_io(N, top, reps=5):
    """Amber restart write and read (coordinates, velocities and box of the
    28,800-atom workload, 6F12.7) through the host library, in MB/s of file."""
    import tempfile
    v = np.random.default_rng(8).normal(0.0, 5.0, top.coords.shape)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "w.rst7")
        t0 = time.perf_counter()
        for _ in range(reps):
            N.write_rst7(p, top.coords, box=top.box, velocities=v, title="bench", time=1.0)
        tw = (time.perf_counter() - t0) / reps
        size = os.path.getsize(p)
        t0 = time.perf_counter()
        for _ in range(reps):
            r = N.read_rst7(p)
        tr = (time.perf_counter() - t0) / reps
    ok = bool(np.abs(r["coords"] - top.coords).max() <= 5.01e-8)
    return {"file_bytes": size, "write_MB_per_s": size / tw / 1e6, "read_MB_per_s": size / tr / 1e6,
            "round_trip_within_F12.7": ok, "workload": "polymer_29k coordinates + velocities + box"}


def bench_adjoint(torch, N, systems, n_frames=64):
    """dLoss/dtheta, loss = sum_f (E_f - E_target,f)^2 over 64 frames of the
    6,450-atom water box (0.1 A Gaussian perturbations, seeds 5000+f), targets
    from parameters perturbed by 5 % log-normal noise (seed 5999).  Two error
    figures of the mixed form against the FP64 form: the adjoint itself (both
    forms' Model.adjoint under the same loss weights, 2 (E_f - E_target,f) of
    the FP64 energies), and the fully mixed loss gradient, whose weights also
    carry the FP32 forward energies' error."""
    from md_amd import autograd as AG
    prm_path, rst_path = systems.cached_system("water_box_6k")
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    frames = torch.tensor(np.stack([top.coords + np.random.default_rng(5000 + i).normal(0, 0.1, top.coords.shape)
                                    for i in range(n_frames)]))
    out = {}
    grads = {}
    models = {}
    target64 = None
    keys = ("charge", "lj_a", "lj_b", "bond_k", "bond_r0", "angle_k", "angle_theta0")
    for name, prec in (("mixed", N.PREC_MIXED), ("fp64", N.PREC_FP64)):
        prm = N.make_params(N.SEM_AMBER, prec, lj_cutoff=9.0, lj_switch=8.0, ewald_tol=1e-5, pme_grid=(48, 48, 48))
        model = AG.Model(top, prm)
        models[name] = model
        base = model.initial()
        rng = np.random.default_rng(5999)
        target_vals = {k: v * torch.tensor(np.exp(rng.normal(0, 0.05, v.shape))) for k, v in base.items()}
        with torch.no_grad():
            target = AG.energy(model, target_vals, frames, top.box)
        if name == "fp64":
            target64 = target
        params = {k: v.clone().requires_grad_(True) for k, v in base.items()}
        t0 = time.perf_counter()
        energies = AG.energy(model, params, frames, top.box)
        loss = ((energies - target) ** 2).sum()
        t1 = time.perf_counter()
        loss.backward()
        t2 = time.perf_counter()
        grads[name] = {k: v.grad.numpy().copy() for k, v in params.items()}
        out[name] = {"forward_frames_per_s": n_frames / (t1 - t0), "backward_frames_per_s": n_frames / (t2 - t1),
                     "loss": float(loss.detach())}
        if name == "fp64":
            e64 = energies.detach().numpy()
    # the adjoint alone: both forms under the FP64 loss weights (each model
    # still holds the base parameters of its last forward pass)
    w = 2.0 * (e64 - target64.numpy())
    fr = frames.numpy()
    same = {name: models[name].adjoint(fr, top.box, w) for name in ("mixed", "fp64")}
    for m in models.values():
        m.close()

    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    out["grad_rel_err_mixed_vs_fp64"] = {k: rel(same["mixed"][k], same["fp64"][k]) for k in keys}
    out["loss_grad_rel_err_fully_mixed_vs_fp64"] = {k: rel(grads["mixed"][k], grads["fp64"][k]) for k in keys}
    out["workload"] = ("64 frames of the 6,450-atom water box (N(0, 0.1 A) perturbations), loss sum (E - E_t)^2, "
                       "E_t from parameters x exp(N(0, 0.05)); host buffers per frame; PME 48^3; "
                       "grad_rel_err_mixed_vs_fp64: the adjoints under the FP64 loss weights; "
                       "loss_grad_rel_err_fully_mixed_vs_fp64: mixed forward energies in the weights")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-system batched run")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the 64-frame parameter-gradient run")
    ap.add_argument("--no-md", action="store_true", help="skip the minimisation + dynamics run")
    ap.add_argument("--no-scale", action="store_true", help="skip the ~100k-atom system")
    ap.add_argument("--tune-reps", type=int, default=10,
                    help="evaluations per candidate in the launch-shape autotuner")
    ap.add_argument("--profile-workload", choices=["polymer_29k", "scale_100k"], default=None,
                    help="only time --steps mixed evaluations of this workload (for profiler runs)")
    ap.add_argument("--shapes", type=int, nargs=2, metavar=("SPLIT", "FFT_LINES"), default=None,
                    help="with --profile-workload: these launch shapes instead of the defaults (4 8)")
    ap.add_argument("--triclinic", action="store_true",
                    help="also time the same system written as a triclinic cell (sheared basis)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device-to-device copy three times (the reference "
                         "kernel for scaling the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.profile_workload:
        profile_workload(args.profile_workload, args.steps, args.warmup, args.shapes)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X; "
              "a multi-GPU run is not measured here.", file=sys.stderr)
        sys.exit(2)

    import torch
    from md_amd import native as N
    from md_amd import systems

    prm, rst = systems.cached_system("polymer_29k")
    top = N.Topology(prm, rst, N.SEM_AMBER)
    kw = dict(lj_cutoff=12.0, lj_switch=11.0, ewald_tol=1e-5, pme_grid=(72, 72, 72))
    p32 = N.make_params(N.SEM_AMBER, N.PREC_MIXED, **kw)
    p64 = N.make_params(N.SEM_AMBER, N.PREC_FP64, **kw)

    dev = torch.device("cuda:0")
    x = torch.tensor(top.coords.reshape(-1), dtype=torch.float64, device=dev)
    f32 = torch.zeros_like(x)
    f64 = torch.zeros_like(x)
    box = top.box

    ev32 = N.Evaluator(0)
    ev32.upload(top.desc)
    ev32.set_params(p32)
    ev64 = N.Evaluator(0)
    ev64.upload(top.desc)
    ev64.set_params(p64)

    # accuracy of the mixed form against the FP64 form (same inputs)
    e32, _ = ev32.evaluate_device(x.data_ptr(), box, f32.data_ptr())
    e64, _ = ev64.evaluate_device(x.data_ptr(), box, f64.data_ptr())
    torch.cuda.synchronize()
    a32, a64 = f32.cpu().numpy(), f64.cpu().numpy()
    rms = float(np.sqrt(np.sum((a32 - a64) ** 2) / np.sum(a64 ** 2)))
    term_err = {k: (abs(e32[k] - e64[k]) / abs(e64[k]) if e64[k] != 0 else abs(e32[k]))
                for k in N.ENERGY_TERMS if k not in ("ee_all", "ee_pairlist", "ee_rf")}

    # work counts of this workload (pairs inside the cutoff, tiles)
    ev32.set_counting(True)
    ev32.evaluate_device(x.data_ptr(), box, f32.data_ptr())
    counts = ev32.counts()
    ev32.set_counting(False)

    # launch shapes picked by the context's autotuner (every pair-kernel
    # split and FFT block shape timed on these buffers, the fastest kept)
    tuned32 = ev32.tune(x.data_ptr(), box, f32.data_ptr(), reps=args.tune_reps)
    tuned64 = ev64.tune(x.data_ptr(), box, f64.data_ptr(), reps=max(3, args.tune_reps // 2))

    # the timed region: in one step of every 16 the pair kernel is bracketed by
    # two HIP events on the evaluator's stream, so its launch duration is
    # measured live; the steps replay a captured graph of the evaluation (the
    # sampled steps the instance with the two event nodes)
    ev32.set_timing(3)
    t32 = timed_steps(torch, ev32, x.data_ptr(), f32.data_ptr(), box, args.steps, args.warmup)
    nb_total, nb_launches = ev32.timing()["nonbonded"]
    graph_replay = ev32.graph_state() == 1
    if nb_launches < (args.warmup + args.steps) // 16:
        raise SystemExit(f"pair-kernel events timed {nb_launches} of {args.warmup + args.steps} steps")
    ev32.set_timing(0)
    nb_ms = nb_total / max(nb_launches, 1)
    steps64 = args.fp64_steps or max(20, args.steps // 2)
    t64 = timed_steps(torch, ev64, x.data_ptr(), f64.data_ptr(), box, steps64, max(5, args.warmup // 4))
    rate32 = args.steps / t32
    rate64 = steps64 / t64

    # host-buffer path (mdgpu_eval: coordinates in and forces out over PCIe)
    xh = top.coords.copy()
    for _ in range(3):
        ev32.evaluate(xh, box)
    nh = max(10, args.steps // 5)
    t0 = time.perf_counter()
    for _ in range(nh):
        ev32.evaluate(xh, box)
    rate_host = nh / (time.perf_counter() - t0)

    kernels_ms = kernel_breakdown(torch, ev32, x.data_ptr(), f32.data_ptr(), box, max(20, args.steps // 4))
    flops = PAIR_FLOP * counts["pairs"]
    achieved = flops / (nb_ms * 1e-3) / 1e12
    traffic = read_traffic(os.path.join(ROOT, "profiles"))
    pme_rl = pme_roofline(kernels_ms, WORKLOAD["atoms"], WORKLOAD["pme_grid"])

    md = None if args.no_md else bench_md(torch, N, systems, top, p32, box)
    io = bench_io(N, top)
    tri = bench_triclinic(torch, N, top, p32, box) if args.triclinic else None
    scale = None if args.no_scale else bench_scale(torch, N, systems)
    batched = None if args.no_batched else bench_batched(torch, N, systems)
    adjoint = None if args.no_adjoint else bench_adjoint(torch, N, systems)

    cpu = None if args.no_cpu_baseline else cpu_baseline(N, prm, rst, kw)

    line = {
        "metric": "energy+force evals/s @29k atoms: mixed vs FP64 GPU form; force rel-RMS err",
        "value": rate32,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 / rate32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, md_amd/systems.py)",
        "config": WORKLOAD,
        "fp64_evals_per_s": rate64,
        "speedup_mixed_vs_fp64": rate32 / rate64,
        "host_buffer_evals_per_s": rate_host,
        "eval_path": "captured-graph replay" if graph_replay else "direct launches",
        "tuned_shapes": {"mixed": {"split": tuned32[0], "fft_lines": tuned32[1] or 16,
                                   "tune_median_ms_direct": tuned32[2]},
                         "fp64": {"split": tuned64[0], "fft_lines": tuned64[1] or 16,
                                  "tune_median_ms_direct": tuned64[2]}},
        "force_rel_rms_mixed_vs_fp64": rms,
        "energy_rel_err_mixed_vs_fp64": term_err,
        "work": counts,
        "kernel_ms": kernels_ms,
        "pme_roofline": pme_rl,
        "roofline": {
            "bound": "valu",
            "engine": "VALU FP32 (no MFMA: the pair kernel is not a dense contraction)",
            "kernel": "k_nonbonded_c (mixed, AMBER, switched LJ + erfc Ewald, in-range pairs packed)",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": traffic,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
            "avg_launch_ms": nb_ms,
            "timed_launches": int(nb_launches),
            "algorithmic": f"{PAIR_FLOP} flop x {counts['pairs']} pairs within 12 A per launch",
        },
        "cpu_baseline": cpu,
        "md": md,
        "io": io,
        "triclinic": tri,
        "scale_100k": scale,
        "batched": batched,
        "adjoint": adjoint,
    }
    print(json.dumps(line))
    ev32.close()
    ev64.close()


if __name__ == "__main__":
    main()
