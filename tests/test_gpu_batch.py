"""Batched evaluation (mdgpu_batch_*, one hipGraph launch for many systems)
against one context per system, on the GPU.

Mixed form: forces are 64-bit fixed point, so the graph, the serial mode and
separate contexts must agree bit for bit; energies too while the pair-kernel
split (which groups the energy partial sums) is unchanged.  FP64 form: FP64
atomics make the last bits order-dependent, bar 1e-12.  Graph replays are
interleaved with unrelated eager evaluations and synchronisations, and every
replay is checked.
"""
import numpy as np
import pytest
import torch

from md_amd import native as N
from md_amd import systems

pytestmark = pytest.mark.gpu

KEYS = ["bond", "angle", "vdw", "ewald_direct", "ewald_self", "ewald_recip", "ewald_excl", "total"]


def _params(prec, k):
    return N.make_params(N.SEM_AMBER, prec, lj_cutoff=9.0, lj_switch=8.0, ewald_tol=1e-5, pme_grid=(k, k, k))


def _setup(prec, n_sys=6, seed0=7000):
    work = systems.batch_workload(n_sys, seed0=seed0, lo=3000, hi=9000)
    keep, descs = [], []
    batch = N.Batch(0)
    for arrays, coords, box, k in work:
        d = N.desc_from_arrays(arrays, keep)
        descs.append((d, _params(prec, k), box))
        batch.add(d, _params(prec, k), box)
    x = np.concatenate([w[1] for w in work]).reshape(-1)
    return work, descs, keep, batch, x


def _single(descs, work, dev):
    """Each system on its own context."""
    out_e, out_f = [], []
    for (d, prm, box), (_, coords, _, _) in zip(descs, work):
        ev = N.Evaluator(0)
        ev.upload(d)
        ev.set_params(prm)
        xs = torch.tensor(coords.reshape(-1), dtype=torch.float64, device=dev)
        fs = torch.zeros_like(xs)
        e, _ = ev.evaluate_device(xs.data_ptr(), [box] * 3, fs.data_ptr())
        torch.cuda.synchronize()
        out_e.append(e)
        out_f.append(fs.cpu().numpy())
        ev.close()
    return out_e, np.concatenate(out_f)


def _same(ea, eb, fa, fb, exact):
    for a, b in zip(ea, eb):
        for k in KEYS:
            if exact:
                assert a[k] == b[k], (k, a[k], b[k])
            else:
                assert abs(a[k] - b[k]) <= 1e-12 * max(abs(b[k]), 1.0), (k, a[k], b[k])
    if exact:
        assert np.array_equal(fa, fb)
    else:
        assert np.abs(fa - fb).max() <= 1e-11 * np.abs(fb).max()


@pytest.mark.parametrize("prec", [N.PREC_MIXED, N.PREC_FP64], ids=["mixed", "fp64"])
def test_graph_batch_matches_separate_contexts(prec):
    dev = torch.device("cuda:0")
    work, descs, keep, batch, x = _setup(prec)
    print(f"case batch of {len(batch)} systems, {x.size // 3} atoms, prec {prec}", flush=True)
    xd = torch.tensor(x, dtype=torch.float64, device=dev)
    fd = torch.zeros_like(xd)
    e_graph = batch.evaluate_device(xd.data_ptr(), fd.data_ptr())
    f_graph = fd.cpu().numpy()
    e_ref, f_ref = _single(descs, work, dev)
    exact = prec == N.PREC_MIXED
    _same(e_graph, e_ref, f_graph, f_ref, exact)
    # replays, each after unrelated eager work and a synchronisation
    for r in range(3):
        _single(descs[:1], work[:1], dev)
        fd.zero_()
        torch.cuda.synchronize()
        e = batch.evaluate_device(xd.data_ptr(), fd.data_ptr())
        _same(e, e_ref, fd.cpu().numpy(), f_ref, exact)
    # the serial mode (one synchronous evaluation per system)
    batch.configure(mode=1)
    fd.zero_()
    e = batch.evaluate_device(xd.data_ptr(), fd.data_ptr())
    _same(e, e_ref, fd.cpu().numpy(), f_ref, exact)
    batch.close()


@pytest.mark.skip(reason="re-capturing the graph of a batch (a former graph-timing tuner did it) ended the "
                         "process with a host segmentation fault on the MI355X twice; cause not found, the "
                         "re-capture path is switched off (a batch captures once) and the tuner times directly "
                         "issued passes")
def test_tuned_batch_and_host_path():
    dev = torch.device("cuda:0")
    work, descs, keep, batch, x = _setup(N.PREC_MIXED, n_sys=5, seed0=7100)
    xd = torch.tensor(x, dtype=torch.float64, device=dev)
    fd = torch.zeros_like(xd)
    e0 = batch.evaluate_device(xd.data_ptr(), fd.data_ptr())
    f0 = fd.cpu().numpy()
    cfg = batch.tune(xd.data_ptr(), fd.data_ptr(), reps=2)
    print(f"  tuned {cfg}")
    assert cfg["streams"] in (2, 4, 8) and cfg["split"] in (2, 4) and cfg["ms_per_launch"] > 0
    fd.zero_()
    e1 = batch.evaluate_device(xd.data_ptr(), fd.data_ptr())
    assert np.array_equal(fd.cpu().numpy(), f0)  # fixed-point forces do not depend on the launch shape
    _same(e1, e0, f0, f0, exact=False)
    e2, f2 = batch.evaluate(x)
    assert np.array_equal(f2.reshape(-1), f0)
    _same(e2, e1, f0, f0, exact=False)
    batch.close()


def test_class_tuner_then_single_capture_host_path():
    """Tune per size class: the tuner captures the graph once on the tuning
    buffers and keeps the faster of graph and direct launches.  With the tuned
    per-system launch shapes the serial mode, the directly issued passes
    (mode 2) and the graph agree bit for bit, on the host path too; a call
    with other buffers makes the graph stale and the batch issues the same
    launches directly (mode 2's passes), bit for bit as before.
    Against separate contexts (default shape) the FP32 partial sums of the
    i-forces are grouped differently, so the bar there is 2e-6."""
    dev = torch.device("cuda:0")
    work, descs, keep, batch, x = _setup(N.PREC_MIXED, n_sys=5, seed0=7100)
    xd = torch.tensor(x, dtype=torch.float64, device=dev)
    fd = torch.zeros_like(xd)
    print("case class tuner, 5 systems", flush=True)
    cfg = batch.tune(xd.data_ptr(), fd.data_ptr(), reps=1)
    per_sys = [batch.system_config(i) for i in range(len(batch))]
    print(f"  tuned {cfg}, per system {per_sys}")
    assert cfg["split"] in (1, 2, 4, 8) and cfg["fft_lines"] in (0, 4, 8) and cfg["streams"] == 5
    assert cfg["mode"] in (0, 2) and batch.mode() == cfg["mode"]
    assert all(p in (1, 2, 4, 8) and f in (0, 4, 8) for p, f in per_sys)
    assert batch.graph_state() == 1  # captured once, on xd / fd
    batch.configure(mode=1, streams=cfg["streams"], split=cfg["split"], fft_lines=cfg["fft_lines"])
    assert [batch.system_config(i) for i in range(len(batch))] == per_sys  # unchanged settings keep the classes'
    e_ser, f_ser = batch.evaluate(x)
    batch.configure(mode=2, streams=cfg["streams"], split=cfg["split"], fft_lines=cfg["fft_lines"])
    e_str, f_str = batch.evaluate(x)
    assert np.array_equal(f_str, f_ser)
    _same(e_str, e_ser, f_ser, f_ser, exact=True)
    # the graph, on the buffers it was captured with, twice
    batch.configure(mode=0, streams=cfg["streams"], split=cfg["split"], fft_lines=cfg["fft_lines"])
    for _ in range(2):
        fd.zero_()
        e1 = batch.evaluate_device(xd.data_ptr(), fd.data_ptr())
        assert batch.graph_state() == 1
        f1 = fd.cpu().numpy().reshape(-1, 3)
        assert np.array_equal(f1, f_ser)
        _same(e1, e_ser, f_ser, f_ser, exact=True)
    e_ref, f_ref = _single(descs, work, dev)
    f1 = f1.reshape(-1)
    rms = np.sqrt(np.sum((f1 - f_ref) ** 2) / np.sum(f_ref ** 2))
    print(f"  tuned shape vs default shape: force rel-RMS {rms:.2e}")
    assert rms <= 2e-6
    # other buffers (the host path's staging copies): the graph goes stale
    e3, f3 = batch.evaluate(x)
    assert batch.graph_state() == 2
    assert np.array_equal(f3.reshape(-1), f1)
    _same(e3, e1, f_ser, f_ser, exact=True)
    batch.close()
