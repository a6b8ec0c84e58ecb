"""The equivariant feature tables (host generator) and the oracle forward, pinned
by the reference's own feature tests (test/ace_features.jl):

- basis multiplicities of the fixture, "symmetric" and "wide" layers, and the
  constructor's error cases (:29-58);
- the resolved cross-product reference (:60-99): with explicit tables the
  oracle gives 0e = A_0, 1o = A_1 and 1e = -cross(A_1, A_2)/sqrt(2) in the
  (y, z, x) order, computed here from the radial formula and the harmonics'
  closed forms; the generated L = 1 even layer reproduces the same cross
  product, which fixes the generator's Clebsch-Gordan phases and norms;
- rotation (D_L from sampled harmonics), inversion parity, translation and
  permutation of the generated correlation-3 layer with L up to 3 (:101-137).
"""
import numpy as np
import pytest

from oracle.ace import ylm
from oracle.features import ace_features as oracle_features
from oracle.features import smooth_exponential

SPEC = [(1, 0), (1, 1), (2, 1)]


def radial(lmax=1):
    import importlib
    f = importlib.import_module("pfw_amd.features")
    spec = SPEC if lmax == 1 else SPEC + [(1, 2)]
    return f.SmoothExponential(spec, 4.0, [0, 0, 1] if lmax == 1 else [0, 0, 1, 0])


def fixture(pfw, correlation=2, lmax=1, species=("Si", "O"), extra=False):
    irreps = [dict(L=0, parity="even", channels=3), dict(L=1, parity="odd", channels=2),
              dict(L=1, parity="even", channels=2)]
    if extra:
        irreps += [dict(L=3, parity="odd", channels=2), dict(L=2, parity="even", channels=2)]
    F = pfw.features
    layer = F.ACEFeatureLayer(list(species), radial(lmax), correlation, lmax, irreps)
    ps = F.feature_parameters(layer, custom_radial={"rate": [0.4], "amplitude": [1.3]})
    for b, key in enumerate(sorted(ps["mixing"], key=lambda k: int(k.split("_")[1]))):
        W = ps["mixing"][key]
        W[...] = np.sin(b + 1 + np.arange(1, W.size + 1)).reshape(W.shape, order="F")
    return layer, ps


def tables(layer):
    return dict(density=layer.density, products=layer.products, coupling=layer.coupling,
                irreps=[(s.L, s.channels) for s in layer.irreps], lmax=layer.lmax, cutoff=layer.cutoff,
                radial_cutoff=layer.radial.cutoff, powers=layer.radial.powers)


def params(layer, ps):
    return dict(mixing=[ps["mixing"][f"block_{b + 1}"] for b in range(len(layer.irreps))],
                scale=ps["radial"]["scale"], rate=ps["custom_radial"]["rate"][0],
                amplitude=ps["custom_radial"]["amplitude"][0])


def channels(layer, labels):
    from pfw_amd.ace import ELEMENTS
    return np.array([layer.atomic_numbers.index(ELEMENTS.index(s) + 1) for s in labels])


X4 = np.array([[0.1, 1.2, 0.7, 2.1], [0.2, 0.4, 1.8, 1.1], [0.3, 1.1, 0.5, 1.9]]).T  # (4, 3)


def test_basis_multiplicities_and_errors(pfw):
    F = pfw.features
    layer, _ = fixture(pfw)
    summary = F.feature_basis_summary(layer)
    assert [(s["L"], s["parity"]) for s in summary] == [(0, "even"), (1, "odd"), (1, "even")]
    assert all(s["basis_channels"] > 0 for s in summary)
    assert max(len(p) for p in layer.products) == 2
    assert any(d[0] == 2 for d in layer.density)  # explicit neighbour-species channels
    again, _ = fixture(pfw)
    assert all(np.array_equal(a, b) for a, b in zip(again.coupling, layer.coupling))
    base = dict(species=["Si"], radial=radial(), correlation=2, lmax=1)
    for spec in ([dict(L=-1, parity="even", channels=1)], [dict(L=1.0, parity="odd", channels=1)],
                 [dict(L=True, parity="odd", channels=1)], [dict(L=0, parity="even", channels=0)],
                 [dict(L=0, parity="bad", channels=1)],
                 [dict(L=0, parity="even", channels=1), dict(L=0, parity="even", channels=2)],
                 [dict(L=0, parity="odd", channels=1)], [dict(L=5, parity="even", channels=1)], []):
        with pytest.raises(pfw.ArgumentError):
            F.ACEFeatureLayer(output_irreps=spec, **base)
    one = [dict(L=0, parity="even", channels=1)]
    with pytest.raises(pfw.ArgumentError):
        F.ACEFeatureLayer(["Si"], radial(), 0, 1, one)
    with pytest.raises(pfw.ArgumentError):
        F.ACEFeatureLayer(["Si"], radial(), 2, 1, one, cutoff=1.0)
    with pytest.raises(pfw.ArgumentError):
        F.ACEFeatureLayer(["Si"], radial(), 2, 1, [dict(L=1, parity="odd", channels=1)], totaldegree=1)
    with pytest.raises(pfw.ArgumentError):
        F.ACEFeatureLayer(["Si", "Si"], radial(), 2, 1, one)
    # a single 1o density cannot produce an axial 1e by self coupling
    ch = F.SmoothExponential([(1, 1)], 4.0, [0])
    with pytest.raises(pfw.ArgumentError):
        F.ACEFeatureLayer(["Si"], ch, 2, 1, [dict(L=1, parity="even", channels=1)])
    symmetric = F.ACEFeatureLayer(["Si"], ch, 3, 1, [dict(L=1, parity="odd", channels=1),
                                                      dict(L=3, parity="odd", channels=1),
                                                      dict(L=2, parity="even", channels=1)])
    assert [s["basis_channels"] for s in F.feature_basis_summary(symmetric)] == [2, 1, 1]
    wide = F.ACEFeatureLayer(["Si"], radial(), 2, 1, [dict(L=1, parity="even", channels=8)])
    s = F.feature_basis_summary(wide)[0]
    assert s["basis_channels"] == 1 and s["rank_limit"] == 1


def cross_tables(pfw):
    F = pfw.features
    density = [(1, 1, 1)] + [(1, n, m) for n in (2, 3) for m in (2, 3, 4)]
    products = [[a] for a in range(1, 5)] + [[2, 6], [3, 7], [4, 5], [2, 7], [3, 5], [4, 6]]
    C0 = np.zeros((1, 10))
    C0[0, 0] = 1
    C1 = np.zeros((3, 10))
    C1[[0, 1, 2], [1, 2, 3]] = 1
    C2 = np.zeros((3, 10))
    for r, c, v in zip([1, 1, 2, 2, 3, 3], [10, 6, 8, 7, 9, 5], [1, -1, 1, -1, 1, -1]):
        C2[r - 1, c - 1] = v / np.sqrt(2)
    irreps = [dict(L=0, parity="even", channels=1), dict(L=1, parity="odd", channels=1),
              dict(L=1, parity="even", channels=1)]
    return F.ACEFeatureLayer.from_tables(radial(), density, products, (C0, C1, C2), species=["Si"],
                                         radial_spec=SPEC, lmax=1, output_irreps=irreps)


def cross_expected(i):
    scalar, a, b = 0.0, np.zeros(3), np.zeros(3)
    for j in range(4):
        if i == j:
            continue
        d = X4[j] - X4[i]
        r = np.linalg.norm(d)
        v = smooth_exponential(r, 4.0, [0, 0, 1], 0.4, 1.3)
        scalar += v[0] / np.sqrt(4 * np.pi)
        a += v[1] * np.sqrt(3 / (4 * np.pi)) * d / r
        b += v[2] * np.sqrt(3 / (4 * np.pi)) * d / r
    return scalar, a, b


def test_cross_product_reference(pfw):
    F = pfw.features
    layer = cross_tables(pfw)
    ps = F.feature_parameters(layer, custom_radial={"rate": [0.4], "amplitude": [1.3]})
    H = 9 * np.eye(3)
    ch = np.zeros(4, int)
    y = oracle_features(X4, ch, H, (0, 0, 0), tables(layer), params(layer, ps))
    gen = F.ACEFeatureLayer(["Si"], radial(), 2, 1, [dict(L=1, parity="even", channels=1)])
    pg = F.feature_parameters(gen, custom_radial={"rate": [0.4], "amplitude": [1.3]})
    yg = oracle_features(X4, ch, H, (0, 0, 0), tables(gen), params(gen, pg))[0]
    for i in range(4):
        scalar, a, b = cross_expected(i)
        expected = -np.cross(a, b)[[1, 2, 0]] / np.sqrt(2)
        assert np.isclose(y[0][0, 0, i], scalar, rtol=1e-12)
        np.testing.assert_allclose(y[1][:, 0, i], a[[1, 2, 0]], rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(y[2][:, 0, i], expected, atol=1e-15)
        np.testing.assert_allclose(yg[:, 0, i], expected, atol=1e-15)
        assert np.linalg.norm(expected) > 1e-5


def test_harmonics_closed_forms():
    d = np.array([0.3, -1.1, 0.7])
    u = d / np.linalg.norm(d)
    Y = ylm(1, d)
    assert np.isclose(Y[0], 1 / np.sqrt(4 * np.pi))
    np.testing.assert_allclose(Y[1:4], np.sqrt(3 / (4 * np.pi)) * u[[1, 2, 0]], rtol=1e-14)


def rotation_D(L, R):
    pts = [np.array([np.sin(0.7 * i), np.cos(1.3 * i), np.sin(2.1 * i + 0.2)]) for i in range(1, 2 * L + 8)]
    pts = [p / np.linalg.norm(p) for p in pts]
    Y = np.array([ylm(L, p)[L * L:(L + 1) ** 2] for p in pts])
    Yr = np.array([ylm(L, R @ p)[L * L:(L + 1) ** 2] for p in pts])
    return np.linalg.lstsq(Y, Yr, rcond=None)[0].T


def test_generated_symmetries(pfw):
    layer, ps = fixture(pfw, correlation=3, lmax=2, extra=True)
    assert max(len(p) for p in layer.products) == 3
    labels = ["Si", "O", "Si", "O"]
    ch = channels(layer, labels)
    H = 9 * np.eye(3)
    T, P = tables(layer), params(layer, ps)
    y = oracle_features(X4, ch, H, (0, 0, 0), T, P)
    for s, block in zip(layer.irreps, y):
        assert block.shape == (2 * s.L + 1, s.channels, 4)
        assert np.linalg.norm(block) > 1e-5
    axis = np.array([1.0, 2.0, -0.7])
    axis /= np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    R = np.eye(3) + np.sin(0.73) * K + (1 - np.cos(0.73)) * K @ K
    rotated = oracle_features(X4 @ R.T, ch, R @ H, (0, 0, 0), T, P)
    inverted = oracle_features(-X4, ch, -H, (0, 0, 0), T, P)
    translated = oracle_features(X4 + np.array([0.3, -0.7, 1.4]), ch, H, (0, 0, 0), T, P)
    perm = [2, 0, 3, 1]
    permuted = oracle_features(X4[perm], ch[perm], H, (0, 0, 0), T, P)
    for b, s in enumerate(layer.irreps):
        D = rotation_D(s.L, R)
        for i in range(4):
            np.testing.assert_allclose(rotated[b][:, :, i], D @ y[b][:, :, i], atol=2e-12, rtol=2e-10)
        np.testing.assert_allclose(inverted[b], (1 if s.parity == "even" else -1) * y[b], atol=1e-13)
        np.testing.assert_allclose(translated[b], y[b], atol=1e-13)
        np.testing.assert_allclose(permuted[b], y[b][:, :, perm], atol=1e-13)
