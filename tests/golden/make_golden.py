"""Write the golden fixtures of tests/golden from the partial reference build.

Runs oracle/_ref/ref_harness (the reference's own parser and term formulas,
compiled from its sources by `make -C oracle ref`) on seeded synthetic systems
and stores what it reports: term lists, parameters, charges, CMAP patch
coefficients and every per-term energy, plus scalar formula samples.  Run it in
the container that holds the reference:

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)

from md_amd import systems  # noqa: E402

CASES = {
    "ref_polymer_2k": lambda: systems.small_polymer(seed=5),
    "ref_water_6k": lambda: systems.water_box(seed=11),
}


def main():
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.path.exists(harness):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
    for name, make in CASES.items():
        top = make()
        with tempfile.TemporaryDirectory() as d:
            p, r = systems.write_system(top, os.path.join(d, name))
            out = os.path.join(d, "out.json")
            subprocess.run([harness, p, r, out], check=True, stdout=subprocess.DEVNULL)
            g = json.load(open(out))
            digest = hashlib.sha256(open(p, "rb").read() + open(r, "rb").read()).hexdigest()
        arrays = {"files_sha256": np.array(digest)}
        for k, v in g.items():
            if k == "cmap_coeff" and not g["cmap_set"]:
                continue  # no CMAP terms: the five placeholder tables are not needed
            arrays[k] = np.asarray(v)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print(name, "written;", {k: g[k] for k in g if k.startswith("E_")})


if __name__ == "__main__":
    main()
