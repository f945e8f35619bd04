"""CPU sanitizer build (AddressSanitizer + UndefinedBehaviorSanitizer) of the
host library and the oracle: make -C md_amd/csrc sanitize, make -C oracle
sanitize.  A subprocess with the sanitizer runtimes preloaded and
MD_AMD_SANITIZE=1 loads the instrumented libraries (checked in its memory
map), reads a prmtop/rst7 in both semantics, writes and reads a restart file
and runs oracle evaluations; any sanitizer report ends it with an error.
The whole CPU suite runs the same way:

    LD_PRELOAD=$(gcc -print-file-name=libasan.so):$(gcc -print-file-name=libubsan.so) \\
    ASAN_OPTIONS=detect_leaks=0:halt_on_error=1 UBSAN_OPTIONS=halt_on_error=1 \\
    MD_AMD_SANITIZE=1 python -m pytest tests -m "not gpu"

(leak detection is off: the Python interpreter and the HIP runtime are not
instrumented).
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))

SCRIPT = r"""
import os, sys, tempfile
sys.path.insert(0, sys.argv[1])
import numpy as np
from md_amd import native as N, systems
import oracle
with tempfile.TemporaryDirectory() as d:
    prm, rst = systems.write_system(systems.small_polymer(seed=5), os.path.join(d, "s"))
    for sem in (N.SEM_REF, N.SEM_AMBER):
        top = N.Topology(prm, rst, sem)
        grid = (0, 0, 0) if sem == N.SEM_REF else (32, 32, 32)
        p = N.make_params(sem, N.PREC_FP64, lj_cutoff=10.0, lj_switch=9.0, ewald_tol=1e-5, pme_grid=grid)
        e, f, _ = oracle.evaluate(top.desc, p, top.coords, top.box)
        assert np.all(np.isfinite(f))
        top.close()
    v = np.random.default_rng(1).normal(0, 1, (2100, 3))
    path = os.path.join(d, "w.rst7")
    top = N.Topology(prm, rst, N.SEM_AMBER)
    N.write_rst7(path, top.coords, box=top.box, velocities=v, title="t", time=1.0)
    r = N.read_rst7(path)
    assert np.abs(r["coords"] - top.coords).max() < 1e-6
maps = open("/proc/self/maps").read()
assert "libmdhost_asan.so" in maps and "liboracle_asan.so" in maps, "sanitizer builds not loaded"
print("SANITIZED_OK")
"""


def _libs():
    out = {}
    for name in ("libasan.so", "libubsan.so"):
        p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
        out[name] = p
    return out


def test_sanitized_host_library_and_oracle(tmp_path):
    host = os.path.join(ROOT, "md_amd", "libmdhost_asan.so")
    orc = os.path.join(ROOT, "oracle", "liboracle_asan.so")
    if not (os.path.exists(host) and os.path.exists(orc)):
        pytest.skip("sanitizer builds absent (make -C md_amd/csrc sanitize; make -C oracle sanitize)")
    libs = _libs()
    if not all(os.path.isabs(p) and os.path.exists(p) for p in libs.values()):
        pytest.skip("gcc sanitizer runtimes not found")
    env = dict(os.environ)
    env.update({"LD_PRELOAD": libs["libasan.so"] + ":" + libs["libubsan.so"],
                "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1",
                "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1", "MD_AMD_SANITIZE": "1"})
    script = tmp_path / "san.py"
    script.write_text(SCRIPT)
    r = subprocess.run([sys.executable, str(script), ROOT], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "SANITIZED_OK" in r.stdout, (r.stdout[-2000:], r.stderr[-4000:])
