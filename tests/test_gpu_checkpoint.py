"""GPU side of calculator checkpoints and basis policies.

* A calculator saved with save_calculator and loaded back evaluates bit for
  bit like the original (the reference's round trip, test/ace_interface.jl:211-217).
* A basis policy's invariant mask and disabled pair term (effective_WB and
  pair_allowed, src/submodules/ACE/interface.jl:122-124) give the oracle's
  result with the masked readout, and the readout gradient vanishes on the
  masked rows while the other rows equal the unmasked model's.
"""
import os

import numpy as np
import pytest

from oracle.ace import ace_evaluate as oracle_ace
from oracle.ace import load_resolved as oracle_load
from oracle.ace import read_frames

try:
    import tomllib as toml
except ImportError:
    import tomli as toml

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def tial():
    with open(os.path.join(GOLD, "resolved_model_v1.toml"), "rb") as f:
        D = toml.load(f)
    frames = read_frames(os.path.join(GOLD, "reference_results.xyz"))
    return D, max(frames, key=lambda f: len(f["X"]))


def test_checkpoint_round_trip_on_device(pfw, tial, tmp_path):
    D, f = tial
    model, ps, _ = pfw.load_resolved(D)
    calc = pfw.ACECalculator(model, ps, precision="float64")
    sys_ = pfw.System(f["X"], f["H"], f["pbc"], f["labels"])
    r0 = calc.evaluate(sys_)
    path = str(tmp_path / "calc.toml")
    pfw.ace.save_calculator(path, calc)
    m2, ps2, _ = pfw.load_resolved(path)
    r1 = pfw.ACECalculator(m2, ps2, precision="float64").evaluate(sys_)
    assert r1["energy"] == r0["energy"]
    assert np.array_equal(r1["forces"], r0["forces"]) and np.array_equal(r1["virial"], r0["virial"])


def test_mask_and_pair_policy(pfw, tial):
    D, f = tial
    model, ps, _ = pfw.load_resolved(D)
    mask = np.ones(model.nbasis, bool)
    mask[1::4] = False
    D2 = dict(D, interface_policy={"version": 1, "mask": mask.tolist(), "species": [], "orders": [],
                                   "pair": False, "reference": True})
    m2, ps2, _ = pfw.load_resolved(D2)
    calc = pfw.ACECalculator(m2, ps2, precision="float64")
    sys_ = pfw.System(f["X"], f["H"], f["pbc"], f["labels"])
    r = calc.evaluate(sys_)
    eWB, eWp = m2.effective(ps["WB"], ps["Wpair"])
    om = oracle_load(D)
    sp = np.array(model.channels(f["labels"]), np.int32)
    E, F, W, _ = oracle_ace(f["X"], sp, f["H"], f["pbc"], dict(om, WB=eWB, Wpair=eWp))
    assert abs(r["energy"] - E) <= 1e-9 * max(1, abs(E))
    np.testing.assert_allclose(r["forces"], F.T, atol=1e-9, rtol=1e-9)
    dF = np.sin(np.arange(3 * len(f["X"]))).reshape(3, -1)
    g = pfw.ace_parameter_vjp(calc, sys_, dE=[1.0], dF=[dF])
    g0 = pfw.ace_parameter_vjp(pfw.ACECalculator(model, ps, precision="float64"), sys_, dE=[1.0], dF=[dF])
    assert not g["WB"][~mask].any() and not g["Wpair"].any()
    np.testing.assert_allclose(g["WB"][mask], g0["WB"][mask], rtol=1e-12, atol=1e-12)
