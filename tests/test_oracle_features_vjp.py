"""The oracle feature pullback (oracle/features.py feature_vjp) pinned by the
reference's pullback tests (test/ace_features.jl:141-200) with its own
tolerances: central differences of the oracle forward (h = 2e-6, atol 2e-8,
rtol 3e-6) for the positions and for one cos-direction per parameter group
(mixing, radial scale, rate and amplitude), translation invariance (the
position gradient sums to zero), single-block cotangents, an empty system,
and the periodic self-image atom whose position gradient vanishes while its
mixing gradient does not; plus a correlation-3 layer with L up to 3."""
import numpy as np
import pytest

from oracle.features import ace_features as oracle_features
from oracle.features import feature_vjp

from test_oracle_features import X4, channels, fixture, params, tables

H3 = np.array([[3.8, 0.2, 0.0], [0.0, 4.1, 0.1], [0.0, 0.0, 4.3]]).T
X3 = np.array([[0.1, 0.2, 0.3], [1.2, 0.4, 1.1], [0.7, 1.8, 0.5]])


def jl_sin(shape):
    return np.sin(np.arange(1, int(np.prod(shape)) + 1)).reshape(shape, order="F")


def loss(X, ch, H, pbc, T, P, cot):
    return sum(float(np.sum(a * b)) for a, b in zip(oracle_features(X, ch, H, pbc, T, P), cot))


def central(f, x, h=2e-6):
    g = np.zeros_like(x)
    for k in np.ndindex(x.shape):
        p, m = x.copy(), x.copy()
        p[k] += h
        m[k] -= h
        g[k] = (f(p) - f(m)) / (2 * h)
    return g


def shifted(P, key, t):
    Q = dict(P)
    Q["mixing"] = [W.copy() for W in P["mixing"]]
    Q["scale"] = P["scale"].copy()
    if key == "mixing":
        for b, W in enumerate(Q["mixing"]):
            W += t * np.cos(np.arange(1, W.size + 1)).reshape(W.shape, order="F")
    elif key == "scale":
        Q["scale"] = Q["scale"] + t * np.cos(np.arange(1, Q["scale"].size + 1)).reshape(Q["scale"].shape, order="F")
    else:
        Q[key] = P[key] + t * np.cos(1.0)
    return Q


def directional(gp, key, P):
    if key == "mixing":
        return sum(float(np.sum(g * np.cos(np.arange(1, g.size + 1)).reshape(g.shape, order="F"))) for g in gp["mixing"])
    if key == "scale":
        g = gp["scale"]
        return float(np.sum(g * np.cos(np.arange(1, g.size + 1)).reshape(g.shape, order="F")))
    return gp[key] * np.cos(1.0)


def check(X, labels, H, pbc, layer, ps, cot):
    T, P = tables(layer), params(layer, ps)
    ch = channels(layer, labels)
    dx, gp = feature_vjp(X, ch, H, pbc, T, P, cot)
    fd = central(lambda x: loss(x, ch, H, pbc, T, P, cot), X)
    np.testing.assert_allclose(dx, fd, atol=2e-8, rtol=3e-6)
    np.testing.assert_allclose(dx.sum(axis=0), 0, atol=1e-12)
    h = 2e-6
    for key in ("mixing", "scale", "rate", "amplitude"):
        f = (loss(X, ch, H, pbc, T, shifted(P, key, h), cot) - loss(X, ch, H, pbc, T, shifted(P, key, -h), cot)) / (2 * h)
        assert np.isclose(directional(gp, key, P), f, atol=2e-8, rtol=3e-6), key
    return dx, gp


@pytest.mark.parametrize("pbc", [(0, 0, 0), (1, 1, 1)])
def test_pullback_positions_and_parameters(pfw, pbc):
    layer, ps = fixture(pfw)
    y = oracle_features(X3, channels(layer, ["Si", "O", "Si"]), H3, pbc, tables(layer), params(layer, ps))
    cot = [jl_sin(b.shape) for b in y]
    for active in (None, 1, 2):
        c = [cot[b] if active is None or active == b else np.zeros_like(cot[b]) for b in range(len(y))]
        check(X3, ["Si", "O", "Si"], H3, pbc, layer, ps, c)


def test_pullback_correlation_three(pfw):
    layer, ps = fixture(pfw, correlation=3, lmax=2, extra=True)
    rng = np.random.default_rng(5)
    ps["radial"]["scale"][...] = rng.uniform(0.7, 1.3, ps["radial"]["scale"].shape)
    labels = ["Si", "O", "Si", "O"]
    y = oracle_features(X4, channels(layer, labels), 9 * np.eye(3), (0, 0, 0), tables(layer), params(layer, ps))
    check(X4, labels, 9 * np.eye(3), (0, 0, 0), layer, ps, [jl_sin(b.shape) for b in y])


def test_pullback_empty_and_self_image(pfw):
    layer, ps = fixture(pfw)
    T, P = tables(layer), params(layer, ps)
    dx, gp = feature_vjp(np.zeros((0, 3)), np.zeros(0, int), H3, (0, 0, 0), T, P,
                         [np.ones((2 * s.L + 1, s.channels, 0)) for s in layer.irreps])
    assert dx.shape == (0, 3) and all(not g.any() for g in gp["mixing"]) and not gp["scale"].any()
    X = np.zeros((1, 3))
    H = 3.2 * np.eye(3)
    y = oracle_features(X, np.zeros(1, int), H, (1, 1, 1), T, P)
    dx, gp = feature_vjp(X, np.zeros(1, int), H, (1, 1, 1), T, P, [np.ones(b.shape) for b in y])
    assert np.linalg.norm(y[0]) > 0
    assert not dx.any()
    assert np.linalg.norm(gp["mixing"][0]) > 0
