"""Captured steps (pfw_set_graphs) and launch knobs (pfw_set_tuning, the
autotuner): the knobs of GPUConfig (src/submodules/NeighbourList/gpu.jl:61-73)
and the ACE strategy (src/submodules/ACE/gpu.jl:301-303) change launch shapes
only, so every setting must reproduce the oracle; a replayed graph must give
bit for bit what the eager step gives at the same positions, on every replay,
with other eager work in between, and must be dropped when anything but the
positions changes."""
import ctypes
import json

import numpy as np
import pytest

from oracle import KIND_LJ, KIND_SMOOTH, nl_pairs, pair_evaluate
from oracle.ace import ace_evaluate as oracle_ace
from oracle.ace import load_resolved as oracle_load

pytestmark = pytest.mark.gpu

A = np.array([[1.0, 0.7], [0.7, 1.3]])
PRECS = ["float64", "mixed"]


@pytest.fixture(scope="module")
def synthetic(pfw):
    D = pfw.ace_basis.synthetic_model()
    model, ps, _ = pfw.load_resolved(D)
    return model, ps, oracle_load(D)


def labels_ab(sp):
    return ["A" if s == 0 else "B" for s in sp]


def ace_labels(pfw, model, sp):
    from pfw_amd.ace import ELEMENTS
    return [ELEMENTS[model.atomic_numbers[s] - 1] for s in sp]


def same(r, q):
    assert r["energy"] == q["energy"]
    assert np.array_equal(r["forces"], q["forces"])
    assert np.array_equal(r["virial"], q["virial"])


def trajectory(X, steps=5, amp=0.02, seed=3):
    rng = np.random.default_rng(seed)
    out = [X]
    for _ in range(steps - 1):
        out.append(out[-1] + rng.normal(0, amp, X.shape))
    return out


@pytest.mark.parametrize("prec", PRECS)
def test_pair_graph_replays_equal_eager(pfw, prec):
    X0, H, pbc, sp = pfw.workloads.b2_triclinic(6, 6, 8, seed=5)
    p = pfw.SmoothPair(["A", "B"], A, 6.0)
    wg = pfw.Workspace(precision=prec)
    we = pfw.Workspace(precision=prec)
    wg.set_graphs(True)
    for k, X in enumerate(trajectory(X0)):
        s = pfw.System(X, H, pbc, labels_ab(sp))
        rg = pfw.evaluate_(wg, s, p)
        # eager work on another context between replays
        re = pfw.evaluate_(we, s, p)
        same(rg, re)
        if k == 2:
            E, F, W = pair_evaluate(X, sp, H, pbc, KIND_SMOOTH, A, 6.0)
            tol = 1e-10 if prec == "float64" else 3e-5
            np.testing.assert_allclose(rg["forces"], F.T, atol=tol * max(1, np.abs(F).max()))
    info = wg.graph_info()
    print(f"pair {prec}: {info}")
    assert info["captures"] == 1 and info["replays"] == 3 and info["held"]
    # a model change drops the graph; the next step is eager and correct
    p2 = pfw.SmoothPair(["A", "B"], A * 1.5, 6.0)
    r2 = pfw.evaluate_(wg, s, p2)
    same(r2, pfw.evaluate_(we, s, p2))
    assert not wg.graph_info()["held"]


def test_argon_graph_and_device_positions(pfw):
    """The 4,000-atom argon cell (launch-bound): replays through
    pfw_update_positions_device copy into the captured positions buffer."""
    from pfw_amd import _capi

    _capi.load()
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    X0, H, pbc, sp = pfw.workloads.argon_fcc_rattled(10)
    lj = pfw.lennard_jones(epsilon=0.010323, sigma=3.405, radius=8.5125)
    w = pfw.Workspace(precision="mixed")
    w.set_graphs(True)
    pfw.evaluate_(w, pfw.System(X0, H, pbc, ["Ar"] * len(X0)), lj)
    dev = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dev), X0.nbytes) == 0
    try:
        E, F, W = np.zeros(1), np.zeros((len(X0), 3)), np.zeros(9)
        for X in trajectory(X0, steps=4, amp=0.01)[1:]:
            Xc = np.ascontiguousarray(X)
            assert hip.hipMemcpy(dev, Xc.ctypes.data, Xc.nbytes, 1) == 0
            w.ctx.call("pfw_update_positions_device", dev)
            w.ctx.call("pfw_pair_evaluate", 7, _capi.PFW_MIXED, _capi.ptr(E), _capi.ptr(F), _capi.ptr(W))
            Er, Fr, Wr = pair_evaluate(Xc, sp, H, pbc, KIND_LJ, (0.010323, 3.405), 8.5125)
            np.testing.assert_allclose(F, Fr, atol=3e-5 * max(1, np.abs(Fr).max()))
            assert abs(E[0] - Er) <= 1e-5 * len(X0)
    finally:
        hip.hipFree(dev)
    assert w.graph_info()["replays"] >= 2


@pytest.mark.parametrize("prec", PRECS)
def test_ace_graph_replays_and_overflow(pfw, synthetic, prec):
    model, ps, om = synthetic
    X0, H, pbc, sp = pfw.workloads.bcc_binary(4, 4, 4, seed=8)
    lab = ace_labels(pfw, model, sp)
    cg = pfw.ACECalculator(model, ps, precision=prec)
    ce = pfw.ACECalculator(model, ps, precision=prec)
    cg.workspace.set_graphs(True)
    for X in trajectory(X0, amp=0.03):
        s = pfw.System(X, H, pbc, lab)
        same(cg.evaluate(s), ce.evaluate(s))
    assert cg.workspace.graph_info()["replays"] == 3
    # squeeze the atoms: environments outgrow their rows during a replay; the
    # step is repeated eagerly with wider rows and the graph recaptured later
    for scale in (0.8, 0.8):
        Xs = X0 * scale
        r = cg.evaluate(pfw.System(Xs, H, pbc, lab))
        E, F, W, _ = oracle_ace(Xs, sp, H, pbc, om)
        tol = 1e-9 if prec == "float64" else 1e-4
        assert abs(r["energy"] - E) <= tol * max(1, abs(E))
        np.testing.assert_allclose(r["forces"], F.T, atol=tol * max(1, np.abs(F).max()))


def test_batch_step_graph(pfw, synthetic):
    """The packed fitting-set step replayed with pfw_pair_step / pfw_ace_step."""
    from pfw_amd import _capi

    model, ps, _ = synthetic
    cfgs, _ = pfw.workloads.fitting_set(ncfg=24, nmin=8, nmax=200, seed=17)
    sysp = [pfw.System(X, H, pb, labels_ab(s)) for X, H, pb, s in cfgs]
    p = pfw.SmoothPair(["A", "B"], A, 6.0)
    lib = _capi.load()
    for prec in PRECS:
        w = pfw.Workspace(precision=prec)
        ref = pfw.evaluate_batch(w, sysp, p)
        w.set_graphs(True)
        for _ in range(3):
            _capi.check(w.ctx.h, lib.pfw_pair_step(w.ctx.h, 7, w.mode))
        E = np.zeros(len(cfgs))
        w.ctx.call("pfw_read_batch_results", _capi.ptr(E), None, None)
        assert np.array_equal(E, [r["energy"] for r in ref])
        assert w.graph_info()["replays"] >= 1
    sysa = [pfw.System(X, H, pb, ace_labels(pfw, model, s)) for X, H, pb, s in cfgs]
    calc = pfw.ACECalculator(model, ps, precision="mixed")
    ref = calc.evaluate_batch(sysa)
    calc.workspace.set_graphs(True)
    h = calc.workspace.ctx.h
    for _ in range(3):
        _capi.check(h, lib.pfw_ace_step(h, 7, calc.workspace.mode))
    E = np.zeros(len(cfgs))
    calc.workspace.ctx.call("pfw_read_batch_results", _capi.ptr(E), None, None)
    assert np.array_equal(E, [r["energy"] for r in ref])


@pytest.mark.parametrize("prec", PRECS)
def test_every_knob_setting_matches_oracle(pfw, synthetic, prec):
    model, ps, om = synthetic
    tol = 1e-9 if prec == "float64" else 1e-4
    X, H, pbc, sp = pfw.workloads.bcc_binary(4, 4, 3, seed=12)
    E, F, W, _ = oracle_ace(X, sp, H, pbc, om)
    s = pfw.System(X, H, pbc, ace_labels(pfw, model, sp))
    calc = pfw.ACECalculator(model, ps, precision=prec)
    # radial chunks 40 and 13 give 2 and 4 groups of matrix-core density
    # tiles; 7 and 1 exceed the kernel's 4 groups and fall back to the VALU;
    # dc 3 also runs the pullback on the matrix cores
    for spw in (1, 2):
        for rchunk in (40, 13, 7, 1):
            for waves in (16, 3, 1):
                for dc in (1, 2, 3):
                    calc.workspace.set_tuning(ace_sites_per_wave=spw, ace_radial_chunk=rchunk, ace_product_waves=waves,
                                              ace_density_contraction=dc)
                    r = calc.evaluate(s)
                    assert abs(r["energy"] - E) <= tol * max(1, abs(E)), (spw, rchunk, waves, dc)
                    np.testing.assert_allclose(r["forces"], F.T, atol=tol * max(1, np.abs(F).max()))
                    np.testing.assert_allclose(r["virial"], W, atol=tol * max(1, np.abs(W).max()))
    Xb, Hb, pbb, spb = pfw.workloads.b2_triclinic(5, 5, 6, seed=9)
    i, j, S, _, _ = nl_pairs(Xb, Hb, pbb, 6.0)
    ref = pfw.pair_set_digest(len(Xb), i, j, S)
    Ep, Fp, Wp = pair_evaluate(Xb, spb, Hb, pbb, KIND_SMOOTH, A, 6.0)
    w = pfw.Workspace(precision=prec)
    p = pfw.SmoothPair(["A", "B"], A, 6.0)
    sb = pfw.System(Xb, Hb, pbb, labels_ab(spb))
    ptol = 1e-10 if prec == "float64" else 3e-5
    for lanes in (1, 2, 3, 4):
        w.set_tuning(pair_lanes_per_atom=lanes)
        w.set_audit(True)
        r = pfw.evaluate_(w, sb, p)
        got = w.read_audit(len(Xb))
        w.set_audit(False)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), lanes
        np.testing.assert_allclose(r["forces"], Fp.T, atol=ptol * max(1, np.abs(Fp).max()))
    with pytest.raises(pfw.ArgumentError):
        w.set_tuning(pair_lanes_per_atom=5)
    with pytest.raises(pfw.ArgumentError):
        w.set_tuning(ace_sites_per_wave=3)
    with pytest.raises(pfw.ArgumentError):
        w.set_tuning(ace_density_contraction=4)
    with pytest.raises(pfw.ArgumentError):
        w.set_tuning(bogus=1)


def test_autotune_sweeps_and_caches(pfw, synthetic, tmp_path):
    model, ps, om = synthetic
    X, H, pbc, sp = pfw.workloads.bcc_binary(6, 6, 6, seed=4)
    s = pfw.System(X, H, pbc, ace_labels(pfw, model, sp))
    cache = str(tmp_path / "tuning.json")
    calc = pfw.ACECalculator(model, ps, precision="mixed")
    rep = pfw.autotune(calc, s, steps=3, warmup=1, cache=cache)
    assert not rep["cached"] and len(rep["candidates"]) == 2 * 2 * 3 * 3 * 3
    print("ACE autotune", rep["best"], rep["graphs"], f"{rep['ms']:.3f} ms")
    assert json.load(open(cache))[rep["key"]]["knobs"] == rep["best"]
    E, F, W, _ = oracle_ace(X, sp, H, pbc, om)
    r = calc.evaluate(s)
    np.testing.assert_allclose(r["forces"], F.T, atol=1e-4 * max(1, np.abs(F).max()))
    calc2 = pfw.ACECalculator(model, ps, precision="mixed")
    rep2 = pfw.autotune(calc2, s, cache=cache)
    got = calc2.workspace.get_tuning()
    assert rep2["cached"] and all(got[k] == v for k, v in rep["best"].items())
    Xb, Hb, pbb, spb = pfw.workloads.b2_triclinic(8, 8, 8, seed=2)
    w = pfw.Workspace(precision="mixed")
    p = pfw.SmoothPair(["A", "B"], A, 6.0)
    repp = pfw.autotune(w, pfw.System(Xb, Hb, pbb, labels_ab(spb)), p, steps=3, warmup=1, cache=cache)
    print("pair autotune", repp["best"], repp["graphs"], f"{repp['ms']:.3f} ms")
    assert len(repp["candidates"]) == 3 * 2
