"""GPU parity of the equivariant features (pfw_features) against the oracle
forward (oracle/features.py), whose tables and conventions are pinned by
tests/test_oracle_features.py; FP64 to 1e-12 relative to the block's scale.

Covers the reference's feature fixtures (test/ace_features.jl): the generated
correlation-2 layer and the correlation-3 layer with L up to 3, open and
periodic cells, random radial scales and mixings, the explicit cross-product
tables, a 54-atom periodic cell, the resolved scalar adapter (the features'
sum equals the many-body energy, :200-221), an empty system and coincident
atoms (DomainError).
"""
import os

import numpy as np
import pytest

from oracle.features import ace_features as oracle_features

from test_oracle_features import X4, channels, cross_tables, fixture, params, tables

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def compare(pfw, layer, ps, X, H, pbc, labels, tol=1e-12):
    sys_ = pfw.System(X, H, pbc, labels)
    got = pfw.ace_features(layer, sys_, ps)
    want = oracle_features(X, channels(layer, labels), H, pbc, tables(layer), params(layer, ps))
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g.shape == w.shape
        np.testing.assert_allclose(g, w, atol=tol * max(1.0, np.abs(w).max()), rtol=0)
    return got


@pytest.mark.parametrize("pbc", [(0, 0, 0), (1, 1, 1)])
def test_generated_layer(pfw, pbc):
    layer, ps = fixture(pfw)
    rng = np.random.default_rng(3)
    ps["radial"]["scale"][...] = rng.uniform(0.7, 1.3, ps["radial"]["scale"].shape)
    H = np.array([[3.8, 0.2, 0.0], [0.0, 4.1, 0.1], [0.0, 0.0, 4.3]]).T
    X = np.array([[0.1, 0.2, 0.3], [1.2, 0.4, 1.1], [0.7, 1.8, 0.5]])
    y = compare(pfw, layer, ps, X, H, pbc, ["Si", "O", "Si"])
    assert all(np.linalg.norm(b) > 1e-6 for b in y)


def test_correlation_three_higher_L(pfw):
    layer, ps = fixture(pfw, correlation=3, lmax=2, extra=True)
    compare(pfw, layer, ps, X4, 9 * np.eye(3), (0, 0, 0), ["Si", "O", "Si", "O"])


def test_cross_product_tables(pfw):
    layer = cross_tables(pfw)
    ps = pfw.features.feature_parameters(layer, custom_radial={"rate": [0.4], "amplitude": [1.3]})
    compare(pfw, layer, ps, X4, 9 * np.eye(3), (0, 0, 0), ["Si"] * 4)


def test_periodic_54_atoms(pfw):
    layer, ps = fixture(pfw, correlation=3, lmax=2, extra=True)
    X, H, pbc, sp = pfw.workloads.bcc_binary(3, 3, 3, a=3.0, seed=11)
    labels = ["Si" if s == 0 else "O" for s in sp]
    compare(pfw, layer, ps, X, H, pbc, labels)


def test_resolved_adapter_equals_many_body_energy(pfw):
    model, readout, _ = pfw.load_resolved(os.path.join(GOLD, "resolved_model_v1.toml"))
    layer = pfw.ACEFeatureLayer.from_resolved(model)
    p = pfw.feature_parameters(layer)
    p["mixing"]["block_1"][:, 0, :] = readout["WB"]
    X = np.array([[0.1, 0.2, 0.3], [2.3, 0.7, 1.2], [1.1, 2.4, 1.5]])
    sys_ = pfw.System(X, 7 * np.eye(3), (1, 1, 1), ["Ti", "Al", "Ti"])
    y = pfw.ace_features(layer, sys_, p)
    calc = pfw.ACECalculator(model, readout, precision="float64")
    scalar = calc.evaluate(sys_)
    calc.set_parameters({"WB": np.zeros_like(readout["WB"]), "Wpair": readout["Wpair"]})
    baseline = calc.evaluate(sys_)
    assert abs(y[0].sum() - (scalar["energy"] - baseline["energy"])) <= 1e-10 * max(1, abs(scalar["energy"]))
    ch = model.channels(["Ti", "Al", "Ti"])
    assert sum(model.reference[c] for c in ch) != baseline["energy"]


def test_empty_and_coincident(pfw):
    layer, ps = fixture(pfw)
    y = pfw.ace_features(layer, pfw.System(np.zeros((0, 3)), 9 * np.eye(3), (0, 0, 0), []), ps)
    assert all(b.shape[2] == 0 for b in y)
    X = np.array([[0.1, 0.2, 0.3], [0.1, 0.2, 0.3], [1.0, 1.0, 1.0]])
    with pytest.raises(pfw.DomainError):
        pfw.ace_features(layer, pfw.System(X, 9 * np.eye(3), (0, 0, 0), ["Si", "O", "Si"]), ps)


def test_results_own_their_storage(pfw):
    """Feature outputs sit in page-locked blocks of the library's host cache:
    each call's arrays keep their values through later calls, and a freed
    block is handed out again."""
    import gc
    from pfw_amd import _capi
    layer, ps = fixture(pfw)
    X = np.array([[0.1, 0.2, 0.3], [1.2, 0.4, 1.1], [0.7, 1.8, 0.5]])
    H = 9 * np.eye(3)
    first = pfw.ace_features(layer, pfw.System(X, H, (0, 0, 0), ["Si", "O", "Si"]), ps)
    keep = [b.copy() for b in first]
    second = pfw.ace_features(layer, pfw.System(X * 1.1, H, (0, 0, 0), ["Si", "O", "Si"]), ps)
    for a, b, c in zip(first, keep, second):
        assert np.array_equal(a, b) and not np.array_equal(a, c)
    lib = _capi.load()
    p = lib.pfw_host_alloc(1 << 20)
    assert p
    lib.pfw_host_free(p)
    assert lib.pfw_host_alloc(1 << 20) == p  # the cached block
    lib.pfw_host_free(p)
    lib.pfw_host_free(None)
    assert not lib.pfw_host_alloc(0)
    a = _capi.pinned_empty(1000)
    a[:] = 3.0
    addr = a.ctypes.data
    del a
    gc.collect()
    b = _capi.pinned_empty(1000)
    assert b.ctypes.data == addr


def test_pinned_results_are_capped(pfw, monkeypatch):
    """Live page-locked result arrays are capped (PINNED_LIVE_CAP); past the cap
    results are ordinary arrays, and freeing a page-locked one releases its share."""
    import gc
    from pfw_amd import _capi
    gc.collect()
    base = _capi._pinned_live[0]
    monkeypatch.setattr(_capi, "PINNED_LIVE_CAP", base + 2 * 65536)
    a = _capi.pinned_empty(8192)  # 64 KiB each
    b = _capi.pinned_empty(8192)
    assert _capi._pinned_live[0] == base + 2 * 65536
    c = _capi.pinned_empty(8192)  # over the cap: pageable
    assert _capi._pinned_live[0] == base + 2 * 65536 and c.flags.owndata
    a[:] = 1.0
    b[:] = 2.0
    c[:] = 3.0
    del a
    gc.collect()
    assert _capi._pinned_live[0] == base + 65536
    assert np.all(b == 2.0) and np.all(c == 3.0)
