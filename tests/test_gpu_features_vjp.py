"""GPU parity of the feature pullback (pfw_features_vjp) against the oracle
pullback (oracle/features.py feature_vjp, pinned by central differences in
tests/test_oracle_features_vjp.py), FP64 to 1e-11 relative to each gradient's
scale; the cases of the reference's pullback tests (test/ace_features.jl:141-200):
the fixture layer open and periodic with all blocks and single-block
cotangents, a correlation-3 layer with L up to 3, a 54-atom periodic cell,
translation invariance of the position gradient, the rrule's saved inputs,
an empty system, the periodic self-image atom (zero position gradient,
nonzero mixing gradient), bad cotangent shapes and a NaN cotangent."""
import numpy as np
import pytest

from oracle.features import ace_features as oracle_features
from oracle.features import feature_vjp as oracle_vjp

from test_oracle_features import X4, channels, fixture, params, tables

pytestmark = pytest.mark.gpu

H3 = np.array([[3.8, 0.2, 0.0], [0.0, 4.1, 0.1], [0.0, 0.0, 4.3]]).T
X3 = np.array([[0.1, 0.2, 0.3], [1.2, 0.4, 1.1], [0.7, 1.8, 0.5]])


def jl_sin(shape):
    return np.sin(np.arange(1, int(np.prod(shape)) + 1)).reshape(shape, order="F")


def close(got, want, tol=1e-11):
    got, want = np.asarray(got, float), np.asarray(want, float)
    assert got.shape == want.shape
    np.testing.assert_allclose(got, want, atol=tol * max(1.0, np.abs(want).max()), rtol=0)


def compare(pfw, layer, ps, X, H, pbc, labels, cot):
    sys_ = pfw.System(X, H, pbc, labels)
    got = pfw.feature_vjp(layer, sys_, ps, cot)
    T, P = tables(layer), params(layer, ps)
    dx, gp = oracle_vjp(X, channels(layer, labels), H, pbc, T, P, cot)
    close(got["positions"], dx)
    for b in range(len(layer.irreps)):
        close(got["parameters"]["mixing"][f"block_{b + 1}"], gp["mixing"][b])
    close(got["parameters"]["radial"]["scale"], gp["scale"])
    close(got["parameters"]["custom_radial"]["rate"], [gp["rate"]])
    close(got["parameters"]["custom_radial"]["amplitude"], [gp["amplitude"]])
    np.testing.assert_allclose(got["positions"].sum(axis=0), 0, atol=1e-11 * max(1.0, np.abs(dx).max()))
    return got


@pytest.mark.parametrize("pbc", [(0, 0, 0), (1, 1, 1)])
def test_fixture_layer(pfw, pbc):
    layer, ps = fixture(pfw)
    rng = np.random.default_rng(3)
    ps["radial"]["scale"][...] = rng.uniform(0.7, 1.3, ps["radial"]["scale"].shape)
    labels = ["Si", "O", "Si"]
    y = oracle_features(X3, channels(layer, labels), H3, pbc, tables(layer), params(layer, ps))
    cot = [jl_sin(b.shape) for b in y]
    for active in (None, 1, 2):
        c = [cot[b] if active is None or active == b else np.zeros_like(cot[b]) for b in range(len(y))]
        compare(pfw, layer, ps, X3, H3, pbc, labels, c)


def test_correlation_three_higher_L(pfw):
    layer, ps = fixture(pfw, correlation=3, lmax=2, extra=True)
    labels = ["Si", "O", "Si", "O"]
    y = oracle_features(X4, channels(layer, labels), 9 * np.eye(3), (0, 0, 0), tables(layer), params(layer, ps))
    compare(pfw, layer, ps, X4, 9 * np.eye(3), (0, 0, 0), labels, [jl_sin(b.shape) for b in y])


def test_periodic_54_atoms(pfw):
    layer, ps = fixture(pfw, correlation=3, lmax=2, extra=True)
    X, H, pbc, sp = pfw.workloads.bcc_binary(3, 3, 3, a=3.0, seed=11)
    labels = ["Si" if s == 0 else "O" for s in sp]
    rng = np.random.default_rng(8)
    cot = [rng.normal(size=(2 * s.L + 1, s.channels, len(X))) for s in layer.irreps]
    compare(pfw, layer, ps, X, H, pbc, labels, cot)


def test_rrule_saves_inputs(pfw):
    layer, ps = fixture(pfw)
    X = X3.copy()
    sys_ = pfw.System(X, H3, (1, 1, 1), ["Si", "O", "Si"])
    y, pb = pfw.ace_features_rrule(layer, sys_, ps)
    cot = [jl_sin(b.shape) for b in y]
    expected = pb(cot)
    ps["mixing"]["block_1"] *= 2
    ps["radial"]["scale"] *= 0.8
    ps["custom_radial"]["rate"][0] += 0.1
    again = pb(cot)
    close(again["positions"], expected["positions"], tol=1e-14)
    close(again["parameters"]["mixing"]["block_2"], expected["parameters"]["mixing"]["block_2"], tol=1e-14)
    assert pb(None) is None
    zero = pb([None] * len(y))
    assert not zero["positions"].any()


def test_empty_self_image_and_shapes(pfw):
    layer, ps = fixture(pfw)
    empty = pfw.System(np.zeros((0, 3)), H3, (0, 0, 0), [])
    g = pfw.feature_vjp(layer, empty, ps, [np.ones((2 * s.L + 1, s.channels, 0)) for s in layer.irreps])
    assert g["positions"].shape == (0, 3)
    assert all(not m.any() for m in g["parameters"]["mixing"].values())
    one = pfw.System(np.zeros((1, 3)), 3.2 * np.eye(3), (1, 1, 1), ["Si"])
    y = pfw.ace_features(layer, one, ps)
    g = pfw.feature_vjp(layer, one, ps, [np.ones(b.shape) for b in y])
    assert np.linalg.norm(y[0]) > 0
    assert np.abs(g["positions"]).max() < 1e-12
    assert np.linalg.norm(g["parameters"]["mixing"]["block_1"]) > 0
    sys_ = pfw.System(X3, H3, (0, 0, 0), ["Si", "O", "Si"])
    bad = [np.ones((2 * s.L + 1, s.channels, 3)) for s in layer.irreps]
    bad[1][0, 0, 2] = np.nan
    with pytest.raises(pfw.ArgumentError):
        pfw.feature_vjp(layer, sys_, ps, bad)
    with pytest.raises(pfw.DimensionMismatch):
        pfw.feature_vjp(layer, sys_, ps, [np.ones((1, 3, 2))] + [None] * (len(layer.irreps) - 1))
    with pytest.raises(pfw.DimensionMismatch):
        pfw.feature_vjp(layer, sys_, ps, [None])
