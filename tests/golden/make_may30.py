"""Writes the May30 Ti-Al-V fixtures from the reference's data/may30 (read as
data; nothing of the reference is run):

* may30_resolved.npz: data/may30/resolved.toml (schema-v1 resolved model:
  real solid harmonics to l = 6, PolyEnvelope1sR pair envelope, 3 species,
  2,155 invariants, 5,652 products) in the package's binary form
  (potentialframework.jl_amd/ace.py: save_resolved_npz), checked to round-trip.
* may30_{bcc-128,hcp-150,omega-240}.toml: the three structures with their
  ACEpotentials 0.10.3 predictions (energy eV, forces eV/Å, virial eV), copied.

Run from the repository root: python tests/golden/make_may30.py
"""
import os
import shutil
import sys

import numpy as np

try:
    import tomllib as toml
except ImportError:
    import tomli as toml

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
SRC = "/root/reference/data/may30"
OUT = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)

from tests.conftest import load_package  # noqa: E402


def main():
    pfw = load_package()
    with open(os.path.join(SRC, "resolved.toml"), "rb") as f:
        D = toml.load(f)
    path = os.path.join(OUT, "may30_resolved.npz")
    pfw.ace.save_resolved_npz(D, path)
    back = pfw.ace.read_resolved_npz(path)
    for fam in ("radial", "pair_radial"):
        for a, b in zip(D[fam]["channels"], back[fam]["channels"]):
            assert np.array_equal(np.asarray(a["coefficients"]), b["coefficients"])
            assert np.array_equal(np.asarray(a["nodes"]), b["nodes"])
    assert back["products"] == D["products"]
    assert np.array_equal(np.asarray(D["parameters"]["flat"]), back["parameters"]["flat"])
    assert np.array_equal(np.asarray(D["coupling"]["values"]), back["coupling"]["values"])
    for name in ("bcc-128", "hcp-150", "omega-240"):
        shutil.copyfile(os.path.join(SRC, name + ".toml"), os.path.join(OUT, f"may30_{name}.toml"))
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
