"""The oracle against the reference's May30 Ti-Al-V golden vectors (CPU).

data/may30 holds a fitted real model (real solid harmonics to l = 6, the
PolyEnvelope1sR pair envelope, 3 species, 2,155 invariants, 5,652 products,
1,000-node splines) and three nonorthogonal periodic structures with their
ACEpotentials 0.10.3 predictions. The reference checks its own evaluators
against them at E 2e-8 eV, F 2e-9 eV/Å + rtol 2e-11, W 2e-8 eV + rtol 2e-11
(test/ace_may30.jl:3-7, 62-79); so does this test, for the oracle. The fixtures
are written by tests/golden/make_may30.py.
"""
import os

import numpy as np
import pytest

from oracle.ace import ace_evaluate as oracle_ace
from oracle.ace import load_resolved as oracle_load

try:
    import tomllib as toml
except ImportError:
    import tomli as toml

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ("bcc-128", "hcp-150", "omega-240")


def may30_structure(name):
    """(X (n, 3), H (columns = lattice vectors), pbc, atomic numbers, E, F (n, 3), W)."""
    with open(os.path.join(GOLD, f"may30_{name}.toml"), "rb") as f:
        d = toml.load(f)
    n = len(d["species"])
    X = np.asarray(d["positions"], float).reshape(n, 3)
    H = np.asarray(d["cell"], float).reshape(3, 3).T
    F = np.asarray(d["forces"], float).reshape(n, 3)
    W = np.asarray(d["virial"], float).reshape(3, 3).T
    return X, H, (True, True, True), [int(z) for z in d["species"]], float(d["energy"]), F, W


def may30_dict(pfw):
    return pfw.ace.read_resolved_npz(os.path.join(GOLD, "may30_resolved.npz"))


@pytest.fixture(scope="module")
def may30(pfw):
    return oracle_load(may30_dict(pfw))


def test_may30_model_shape(pfw, may30):
    m = may30
    assert m["solid"] and m["lmax"] == 6 and m["atomic_numbers"] == [22, 13, 23]
    assert m["WB"].shape == (2155, 3) and m["Wpair"].shape == (10, 3)
    assert len(m["products"]) == 5652 and len(m["density"]) == 156
    assert m["radial"]["nnodes"] == 1000 and m["radial"]["nr"] == 114 and m["pair"]["nr"] == 10


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matches_may30_predictions(pfw, may30, name):
    X, H, pbc, z, E, F, W = may30_structure(name)
    sp = np.array([may30["atomic_numbers"].index(v) for v in z], np.int32)
    e, f, w, _ = oracle_ace(X, sp, H, pbc, may30, threads=8)
    assert abs(e - E) <= 2e-8
    np.testing.assert_allclose(f, F, atol=2e-9, rtol=2e-11)
    np.testing.assert_allclose(w, W, atol=2e-8, rtol=2e-11)
