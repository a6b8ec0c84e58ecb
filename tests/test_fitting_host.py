"""Host logic of the fitting helpers (src/submodules/ACE/fitting.jl:1-11,103-126):
flat readout vectors and their inverse, and training-row selection with
baseline subtraction and residual weights, on a hand-made design."""
import numpy as np
import pytest


def test_linear_parameters_round_trip(pfw):
    rng = np.random.default_rng(1)
    ps = {"WB": rng.normal(size=(5, 2)), "Wpair": rng.normal(size=(3, 2))}
    theta = pfw.linear_parameters(ps)
    # vcat(vec(WB), vec(Wpair)), column-major
    assert np.array_equal(theta[:10], ps["WB"].T.reshape(-1))
    assert theta[1] == ps["WB"][1, 0] and theta[5] == ps["WB"][0, 1]
    assert np.array_equal(theta[10:], ps["Wpair"].T.reshape(-1))
    back = pfw.install_readout(ps, theta)
    assert np.array_equal(back["WB"], ps["WB"]) and np.array_equal(back["Wpair"], ps["Wpair"])
    only = pfw.install_readout(ps, theta[:10] * 2, fit_pair=False)
    assert np.array_equal(only["WB"], 2 * ps["WB"]) and np.array_equal(only["Wpair"], ps["Wpair"])
    with pytest.raises(pfw.DimensionMismatch):
        pfw.install_readout(ps, theta[:9])
    with pytest.raises(pfw.DimensionMismatch):
        pfw.install_readout(ps, theta, fit_pair=False)


def test_training_rows(pfw):
    n = 2
    rows = 1 + 3 * n + 9
    A = np.arange(rows * 3, dtype=float).reshape(rows, 3)
    base = np.arange(rows, dtype=float) * 0.5
    design = {"matrix": A, "baseline": base}
    F = np.array([[1.0, 2.0], [3.0, 4.0], [5.0, 6.0]])  # 3 x n
    W = np.arange(9.0).reshape(3, 3)
    t = pfw.training_rows(design, {"energy": 10.0, "forces": F, "virial": W},
                          weights={"energy": 1.0, "forces": 2.0, "virial": 0.25})
    assert np.array_equal(t["matrix"], A)
    want = np.concatenate([[10.0], F.T.reshape(-1), W.T.reshape(-1)]) - base
    assert np.allclose(t["labels"], want)
    assert np.array_equal(t["row_weights"], [1.0] + [2.0] * 6 + [0.25] * 9)
    # forces as per-atom three-vectors; a missing or None label omits its rows
    t2 = pfw.training_rows(design, {"forces": [F[:, 0], F[:, 1]], "virial": None})
    assert np.array_equal(t2["matrix"], A[1:7])
    assert np.allclose(t2["labels"], F.T.reshape(-1) - base[1:7])
    with pytest.raises(pfw.DimensionMismatch):
        pfw.training_rows(design, {"forces": np.zeros((3, 3))})
    # an (n, 3) array (one row per atom, as calculators.forces returns) has the
    # right length but the wrong layout: rejected, not silently reordered
    with pytest.raises(pfw.DimensionMismatch):
        pfw.training_rows(design, {"forces": F.T.copy()})
    t3 = pfw.training_rows(design, {"forces": list(F.T)})
    assert np.allclose(t3["labels"], F.T.reshape(-1) - base[1:7])
    with pytest.raises(pfw.DimensionMismatch):
        pfw.training_rows(design, {"forces": [np.zeros(2), np.zeros(2)]})
    with pytest.raises(pfw.ArgumentError):
        pfw.training_rows(design, {"energy": 1.0}, weights={"energy": -1.0})
