"""Captured-graph replay of single-context evaluations (mdgpu_set_graph).

A context captures its evaluation once two consecutive evaluations had the
same pointers, box, parameters and buffer sizes, and replays the graph from
then on.  These tests evaluate a sequence of perturbed frames written into
the same device buffer with replay allowed and with direct launches only:
the mixed form must agree bit for bit (fixed-point forces, fixed summation
orders), the FP64 form to 1e-12 (its FP64 atomics sum in any order).  A box
change after capture must switch the context back to direct launches and
still give the direct result.
"""
import numpy as np
import pytest
import torch

from md_amd import native as N

pytestmark = pytest.mark.gpu

KEYS = ["bond", "angle", "dihedral", "urey_bradley", "improper", "cmap", "vdw", "ewald_direct", "ewald_self",
        "ewald_recip", "ewald_excl", "vdw14", "ee14", "total"]


def _frames(top, n, seed=7):
    rng = np.random.default_rng(seed)
    x0 = np.asarray(top.coords, dtype=np.float64).reshape(-1)
    return [x0 + rng.normal(0.0, 0.05, x0.shape) for _ in range(n)]


def _run(top, prm, frames, box_seq, graph):
    ev = N.Evaluator(0)
    ev.upload(top.desc)
    ev.set_params(prm)
    ev.set_graph(graph)
    dev = torch.device("cuda", 0)
    xd = torch.zeros(frames[0].size, dtype=torch.float64, device=dev)
    fd = torch.zeros_like(xd)
    out, states = [], []
    try:
        for fr, box in zip(frames, box_seq):
            xd.copy_(torch.from_numpy(fr))
            e, v = ev.evaluate_device(xd.data_ptr(), box, fd.data_ptr())
            torch.cuda.synchronize()
            out.append((e, fd.cpu().numpy().copy(), v.copy()))
            states.append(ev.graph_state())
    finally:
        ev.close()
    print(f"  graph states {states}")
    return out, states


@pytest.mark.parametrize("prec", [N.PREC_MIXED, N.PREC_FP64], ids=["mixed", "fp64"])
def test_graph_replay_matches_direct_launches(prec, polymer_2k):
    prm_path, rst_path = polymer_2k
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    print(f"case polymer_2k graph replay prec={prec}", flush=True)
    prm = N.make_params(N.SEM_AMBER, prec, lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-5, pme_grid=(32, 32, 32))
    box = list(np.asarray(top.box, dtype=np.float64).reshape(-1)[:3])
    frames = _frames(top, 7)
    # frames 0..4 in the original box (capture at frame 1, replays after),
    # then a slightly larger box: the context returns to direct launches
    boxes = [box] * 5 + [[b * 1.01 for b in box]] * 2
    g, sg = _run(top, prm, frames, boxes, True)
    d, sd = _run(top, prm, frames, boxes, False)
    # captured at the second evaluation (same key as the first), replayed up
    # to the box change, then off for good
    assert sg == [0, 1, 1, 1, 1, 2, 2], sg
    assert all(st == 0 for st in sd), sd
    for k, ((eg, fg, vg), (ed, fdd, vd)) in enumerate(zip(g, d)):
        for key in KEYS:
            if prec == N.PREC_MIXED:
                assert eg[key] == ed[key], (k, key, eg[key], ed[key])
            else:
                assert abs(eg[key] - ed[key]) <= 1e-12 * max(abs(ed[key]), 1.0), (k, key)
        if prec == N.PREC_MIXED:
            assert np.array_equal(fg, fdd), k
            assert np.array_equal(vg, vd), k
        else:
            assert np.abs(fg - fdd).max() <= 1e-11 * np.abs(fdd).max(), k
    # the frames differ, so a replay that ignored new coordinates would repeat forces
    assert not np.array_equal(g[3][1], g[4][1])


def test_sampled_pair_timing_with_graph_replay(polymer_2k):
    """Timing mode 3: the pair kernel's events in one evaluation of every 16,
    the others replaying the graph without event nodes."""
    prm_path, rst_path = polymer_2k
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    print("case polymer_2k sampled timing", flush=True)
    prm = N.make_params(N.SEM_AMBER, N.PREC_MIXED, lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-5,
                        pme_grid=(32, 32, 32))
    box = list(np.asarray(top.box, dtype=np.float64).reshape(-1)[:3])
    ev = N.Evaluator(0)
    try:
        ev.upload(top.desc)
        ev.set_params(prm)
        ev.set_timing(3)
        dev = torch.device("cuda", 0)
        xd = torch.tensor(np.asarray(top.coords, dtype=np.float64).reshape(-1), device=dev)
        fd = torch.zeros_like(xd)
        ev.set_box(box)
        for _ in range(40):
            ev.step_device(xd.data_ptr(), fd.data_ptr())
        torch.cuda.synchronize()
        ms, n = ev.timing()["nonbonded"]
        print(f"  timed launches {n}, mean {ms / max(n, 1):.4f} ms, graph state {ev.graph_state()}")
        assert ev.graph_state() == 1
        assert n == 3  # evaluations 0, 16 and 32
        assert 0.0 < ms / n < 10.0
    finally:
        ev.close()


@pytest.mark.parametrize("prec", [N.PREC_MIXED, N.PREC_FP64], ids=["mixed", "fp64"])
def test_tune_keeps_results_and_graph(prec, polymer_2k):
    """mdgpu_tune times every split / FFT block shape, keeps the fastest and
    leaves the context as it was: no graph yet, timing totals unchanged, and
    evaluations that agree with an untuned context (mixed forces bit for bit,
    energies and FP64 forces to the rounding of their partial sums)."""
    prm_path, rst_path = polymer_2k
    top = N.Topology(prm_path, rst_path, N.SEM_AMBER)
    print(f"case polymer_2k tune prec={prec}", flush=True)
    prm = N.make_params(N.SEM_AMBER, prec, lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-5, pme_grid=(32, 32, 32))
    box = list(np.asarray(top.box, dtype=np.float64).reshape(-1)[:3])
    dev = torch.device("cuda", 0)
    xd = torch.tensor(_frames(top, 1)[0], device=dev)
    ref, out = N.Evaluator(0), N.Evaluator(0)
    try:
        ref.upload(top.desc)
        ref.set_params(prm)
        fr = torch.zeros_like(xd)
        er, _ = ref.evaluate_device(xd.data_ptr(), box, fr.data_ptr())
        out.upload(top.desc)
        out.set_params(prm)
        out.set_timing(3)
        fd = torch.zeros_like(xd)
        split, lines, ms = out.tune(xd.data_ptr(), box, fd.data_ptr(), reps=3)
        print(f"  tuned split {split} fft_lines {lines} median {ms:.4f} ms")
        assert split in (1, 2, 4, 8) and lines in (0, 8, 4) and 0.0 < ms < 100.0
        assert out.graph_state() == 0
        assert out.timing()["nonbonded"] == (0.0, 0)
        for _ in range(3):  # the second captures, the third replays
            e, _ = out.evaluate_device(xd.data_ptr(), box, fd.data_ptr())
        assert out.graph_state() == 1
        torch.cuda.synchronize()
        f_ref, f_out = fr.cpu().numpy(), fd.cpu().numpy()
        # energies: summed in a fixed order for a given split, but the split
        # decides which wave sums which pairs (FP32 partial sums in the mixed
        # form, FP64 in the FP64 form)
        ebar = 1e-7 if prec == N.PREC_MIXED else 1e-12
        for key in KEYS:
            assert abs(e[key] - er[key]) <= ebar * max(abs(er[key]), 1.0), (key, e[key], er[key])
        if prec == N.PREC_MIXED:
            assert np.array_equal(f_out, f_ref)
        else:
            assert np.abs(f_out - f_ref).max() <= 1e-11 * np.abs(f_ref).max()
        # once a graph exists the shapes are fixed
        assert out.shapes() == (split, lines)
        with pytest.raises(N.GpuError):
            out.tune(xd.data_ptr(), box, fd.data_ptr(), reps=1)
        out.set_shapes(split, lines)  # unchanged: accepted
        with pytest.raises(N.GpuError):
            out.set_shapes(2 if split != 2 else 4, lines)
        with pytest.raises(N.GpuError):
            ref.set_shapes(3, 5)  # FFT lines must be a power of two
    finally:
        ref.close()
        out.close()
