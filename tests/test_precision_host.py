"""Host logic of compare_precision / PrecisionReport (src/submodules/ACE/precision.jl:
25-121) with stand-in evaluators of known outputs: per-structure maximum, RMS,
component location and atom-vector errors, pooled RMS over all components,
tolerance statuses, empty input, preparation and evaluation failures and
nonfinite outputs propagating infinite aggregate errors, invalid tolerances;
the cases of test/ace_adaptation.jl:62-90."""
import math

import numpy as np
import pytest


def system(pfw, n=2):
    X = np.array([[0.0, 0.0, 0.0], [2.8, 0.3, 0.2], [1.0, 1.0, 1.0]])[:n]
    return pfw.System(X, 16 * np.eye(3), (1, 1, 1), ["Ti", "Al", "Ti"][:n])


def fake(dF=None, dE=1e-6, dW=None, nonfinite=False):
    def f64(s):
        n = len(s)
        return {"energy": -3.0, "forces": np.arange(3 * n, dtype=float).reshape(3, n, order="F"),
                "virial": np.eye(3)}

    def f32(s):
        r = f64(s)
        r["energy"] += dE
        if dF is not None:
            r["forces"] = r["forces"] + dF[:, :len(s)]
        if dW is not None:
            r["virial"] = r["virial"] + dW
        if nonfinite:
            r["forces"][0, 0] = np.nan
        return r
    return {"Float64": f64, "Float32": f32}


def test_metrics_and_pooling(pfw):
    dF = np.zeros((3, 3))
    dF[1, 1] = 3e-5   # atom 1, axis y
    dF[2, 1] = 4e-5   # atom 1, axis z: atom vector 5e-5
    dW = np.zeros((3, 3))
    dW[0, 2] = -2e-4  # row 0, column 2
    ev = fake(dF=dF, dW=dW)
    r = pfw.compare_precision(None, {}, system(pfw), evaluators=ev)
    assert "ACE precision comparison: 1 structures; descriptive" in repr(r)
    assert len(r.structures) == 1 and r.aggregate["status"] == "descriptive"
    m = r.structures[0]["metrics"]
    assert m["forces"]["maximum"] == pytest.approx(4e-5) and m["forces"]["location"] == dict(structure=0, atom=1, axis=2)
    assert m["forces"]["max_atom_vector"] == pytest.approx(5e-5)
    assert m["forces"]["rms"] == pytest.approx(math.sqrt((9e-10 + 16e-10) / 6))
    assert m["virial"]["maximum"] == pytest.approx(2e-4) and m["virial"]["location"] == dict(structure=0, row=0, column=2)
    assert m["energy"]["absolute"] == pytest.approx(1e-6) and m["energy"]["per_atom"] == pytest.approx(5e-7)
    assert r.aggregate["forces"]["components"] == 6
    two = pfw.compare_precision(None, {}, [system(pfw), system(pfw, 3)], evaluators=ev)
    agg = two.aggregate["forces"]
    assert agg["components"] == 15
    assert agg["pooled_rms"] == pytest.approx(math.sqrt(2 * (9e-10 + 16e-10) / 15))
    assert agg["maximum_atom_vector"] == pytest.approx(5e-5)
    assert two.aggregate["energy"]["maximum_per_atom"] == pytest.approx(5e-7)
    assert two.structures[1]["atoms"] == 3 and two.structures[0]["identity"] != two.structures[1]["identity"]
    assert two.structures[0]["position_rounding_max_A"] > 0  # 2.8 and 0.3 are not float32 numbers


def test_tolerances_and_failures(pfw):
    s = system(pfw)
    ok = pfw.compare_precision(None, {}, [s], tolerances={"energy": 1e-5, "forces": 0.0}, evaluators=fake())
    assert ok.aggregate["status"] == "passed" and ok.structures[0]["metrics"]["forces"]["status"] == "passed"
    assert ok.structures[0]["metrics"]["virial"]["status"] == "descriptive"
    tight = pfw.compare_precision(None, {}, [s, s], tolerances={"energy": 0.0}, evaluators=fake())
    assert tight.aggregate["status"] == "failed" and tight.structures[0]["status"] == "failed"
    assert pfw.compare_precision(None, {}, [], evaluators=fake()).aggregate["status"] == "empty"
    prep = pfw.compare_precision(None, {}, s, evaluators={"Float64": "no device", "Float32": "no device"})
    assert prep.aggregate["status"] == "failed"
    assert all("preparation failed" in e for e in prep.structures[0]["errors"])
    assert prep.metadata["preparation_errors"] == ("no device", "no device")
    bad = pfw.compare_precision(None, {}, [s, s], evaluators=fake(nonfinite=True))
    assert bad.aggregate["status"] == "failed" and math.isinf(bad.aggregate["forces"]["pooled_rms"])
    assert "nonfinite output" in bad.structures[0]["errors"][0]

    def broken(sys_):
        raise ValueError("invalid geometry")
    mixed = pfw.compare_precision(None, {}, [s, s], evaluators={"Float64": broken, "Float32": fake()["Float32"]})
    assert [r["status"] for r in mixed.structures] == ["failed", "failed"]
    for t in ({"forces": -1.0}, {"forces": math.inf}, {"stress": 1.0}, {"forces": True}):
        with pytest.raises(pfw.ArgumentError):
            pfw.compare_precision(None, {}, [s], tolerances=t, evaluators=fake())
    with pytest.raises(pfw.ArgumentError):
        pfw.compare_precision(None, {}, [s], properties=("energy", "stress"), evaluators=fake())
    only = pfw.compare_precision(None, {}, [s], properties=("forces",), evaluators=fake())
    assert set(only.aggregate) == {"status", "forces"}


def test_low_precision_leg_receives_float32_geometry(pfw):
    """The Float32 leg evaluates the structure with positions and cell rounded to
    Float32 (precision_geometry, adaptation.jl:122-127); the Float64 leg the
    structure as given. Stand-in evaluators record what they receive."""
    seen = {}

    def record(name):
        def f(s):
            seen[name] = (np.array(s.positions), np.array(s.cell), tuple(s.pbc), list(s.labels))
            return {"energy": 0.0, "forces": np.zeros((3, len(s))), "virial": np.zeros((3, 3))}
        return f

    X = np.array([[0.1, 0.2, 0.3], [1.0 / 3.0, 2.0 / 3.0, np.pi]])
    H = np.array([[5.0 + 1e-9, 0.1, 0.0], [0.0, 5.3, 0.0], [0.0, 0.0, np.e]])
    s = pfw.System(X, H, (1, 0, 1), ["Ti", "Al"])
    r = pfw.compare_precision(None, {}, [s], evaluators={"Float64": record("Float64"), "Float32": record("Float32")})
    X64, H64, pb64, lab64 = seen["Float64"]
    X32, H32, pb32, lab32 = seen["Float32"]
    assert np.array_equal(X64, X) and np.array_equal(H64, H)
    assert np.array_equal(X32, X.astype(np.float32).astype(np.float64))
    assert np.array_equal(H32, H.astype(np.float32).astype(np.float64))
    assert not np.array_equal(X32, X)
    assert pb32 == pb64 == (True, False, True) and lab32 == lab64 == ["Ti", "Al"]
    assert "rounded to Float32" in r.metadata["low_precision_leg"]
    assert r.structures[0]["position_rounding_max_A"] == pytest.approx(np.abs(X32 - X).max())
