"""compare_precision on the device (src/submodules/ACE/precision.jl:25-121) with
the cases of test/ace_adaptation.jl:62-90 on the TiAl resolved model: a
two-atom periodic structure (descriptive report, six force components, finite
pooled RMS, the maximum equal to the difference of direct evaluations: FP64
on the structure, mixed on its Float32-rounded geometry), an empty collection,
zero tolerances (failed), and a NaN readout coefficient (failed, infinite
pooled RMS)."""
import math
import os

import numpy as np

import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_compare_precision_device(pfw):
    model, ps, _ = pfw.load_resolved(os.path.join(GOLD, "resolved_model_v1.toml"))
    s = pfw.System(np.array([[0.0, 0.0, 0.0], [2.8, 0.3, 0.2]]), 16 * np.eye(3), (1, 1, 1), ["Ti", "Al"])
    r = pfw.compare_precision(model, ps, s)
    assert "ACE precision comparison" in repr(r)
    assert len(r.structures) == 1 and r.aggregate["status"] == "descriptive"
    assert r.aggregate["forces"]["components"] == 6 and math.isfinite(r.aggregate["forces"]["pooled_rms"])
    f64 = pfw.ACECalculator(model, ps, precision="float64").evaluate(s)
    f32 = pfw.ACECalculator(model, ps, precision="mixed").evaluate(pfw.precision.precision_geometry(s))
    assert r.aggregate["forces"]["maximum"] == pytest.approx(np.abs(f32["forces"] - f64["forces"]).max(), rel=1e-12,
                                                             abs=1e-300)
    assert r.aggregate["forces"]["maximum"] < 1e-4
    assert r.structures[0]["metrics"]["energy"]["absolute"] == pytest.approx(abs(f32["energy"] - f64["energy"]),
                                                                            rel=1e-12, abs=1e-300)
    assert pfw.compare_precision(model, ps, []).aggregate["status"] == "empty"
    zero = pfw.compare_precision(model, ps, [s, s], tolerances={"energy": 0.0, "forces": 0.0, "virial": 0.0})
    assert zero.aggregate["status"] == "failed"
    bad = {"WB": ps["WB"].copy(), "Wpair": ps["Wpair"]}
    bad["WB"].flat[0] = np.nan
    failed = pfw.compare_precision(model, bad, [s, s])
    assert failed.aggregate["status"] == "failed" and math.isinf(failed.aggregate["forces"]["pooled_rms"])
