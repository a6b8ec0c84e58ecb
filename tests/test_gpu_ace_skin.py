"""GPU parity of ACE environment reuse (ACECalculator(skin=...)) against the oracle.

With a skin the environments are kept at cutoff + skin and compacted every
step to the edges inside the cutoff at the current positions. Every step of
each trajectory is compared with the oracle's fresh evaluation at the same
positions: FP64 to 1e-9 (the reference's own bound for this test,
test/gpu_ace.jl:140-152), mixed to 1e-4. The first case is that reference
test: a three-atom TiAl system, periodic in x and z, whose first atom moves
by 0, 0.02, 0.12 and 0.35 Å with skin 0.6, rebuilding at least twice.
"""
import os

import numpy as np
import pytest

from oracle.ace import ace_evaluate as oracle_ace
from oracle.ace import load_resolved as oracle_load
from oracle.ace import read_frames

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = {"float64": 1e-9, "mixed": 1e-4}
PRECS = ("float64", "mixed")


@pytest.fixture(scope="module")
def tial(pfw):
    path = os.path.join(GOLD, "resolved_model_v1.toml")
    model, ps, _ = pfw.load_resolved(path)
    return model, ps, oracle_load(path)


@pytest.fixture(scope="module")
def synthetic(pfw):
    D = pfw.ace_basis.synthetic_model()
    model, ps, _ = pfw.load_resolved(D)
    return model, ps, oracle_load(D)


def labels_of(model, sp):
    from pfw_amd.ace import ELEMENTS
    return [ELEMENTS[model.atomic_numbers[s] - 1] for s in sp]


def check(pfw, calc, om, X, H, pbc, sp, labels, prec):
    r = calc.evaluate(pfw.System(X, H, pbc, labels), site_energies=True)
    E, F, W, site = oracle_ace(X, sp, H, pbc, om)
    tol = TOL[prec]
    assert abs(r["energy"] - E) <= tol * max(1, abs(E))
    np.testing.assert_allclose(r["site_energies"], site, atol=tol * 10, rtol=0)
    np.testing.assert_allclose(r["forces"], F.T, atol=tol * max(1, np.abs(F).max()), rtol=tol)
    np.testing.assert_allclose(r["virial"], W, atol=tol * max(1, abs(E)), rtol=tol)
    return r


@pytest.mark.parametrize("prec", PRECS)
def test_reference_three_atom_steps(pfw, tial, prec):
    model, ps, om = tial
    X = np.array([[15.8, 0.0, 0.0], [2.1, 0.2, 0.3], [1.1, 2.8, 0.1]])
    H = np.eye(3) * 18
    pbc = (True, False, True)
    labels = ["Ti", "Al", "Ti"]
    sp = np.array(model.channels(labels), np.int32)
    calc = pfw.ACECalculator(model, ps, precision=prec, skin=0.6)
    for step in (0.0, 0.02, 0.12, 0.35):
        Y = X.copy()
        Y[0, 0] += step
        check(pfw, calc, om, Y, H, pbc, sp, labels, prec)
    assert calc.workspace.rebuilds >= 2


@pytest.mark.parametrize("prec", PRECS)
def test_tial_frame_random_walk(pfw, tial, prec):
    model, ps, om = tial
    frames = read_frames(os.path.join(GOLD, "reference_results.xyz"))
    f = max(frames, key=lambda f: len(f["X"]))
    sp = np.array(model.channels(f["labels"]), np.int32)
    calc = pfw.ACECalculator(model, ps, precision=prec, skin=0.4)
    rng = np.random.default_rng(17)
    X = np.array(f["X"], float)
    for step in range(10):
        check(pfw, calc, om, X, f["H"], f["pbc"], sp, f["labels"], prec)
        X = X + rng.normal(0, 0.02, X.shape)
    info = calc.workspace.verlet_info()
    assert info["rebuilds"] >= 1 and info["reuses"] >= 4, info


@pytest.mark.parametrize("prec", PRECS)
def test_synthetic_random_walk_and_jump(pfw, synthetic, prec):
    model, ps, om = synthetic
    X, H, pbc, sp = pfw.workloads.bcc_binary(3, 3, 3, seed=31)
    labels = labels_of(model, sp)
    calc = pfw.ACECalculator(model, ps, precision=prec, skin=0.5)
    rng = np.random.default_rng(5)
    for step in range(8):
        check(pfw, calc, om, X, H, pbc, sp, labels, prec)
        X = X + rng.normal(0, 0.02, X.shape)
    reuses = calc.workspace.verlet_info()["reuses"]
    assert reuses >= 3
    base = calc.workspace.rebuilds
    X[0] += H[:, 0]  # a wrapped jump rebuilds; the result is unchanged
    check(pfw, calc, om, X, H, pbc, sp, labels, prec)
    assert calc.workspace.rebuilds == base + 1


def test_adjoint_between_skin_steps(pfw, synthetic):
    """The parameter adjoint rebuilds the cells; the next skin step must not
    reuse rows indexed in the old order."""
    model, ps, om = synthetic
    X, H, pbc, sp = pfw.workloads.bcc_binary(3, 3, 2, seed=8)
    labels = labels_of(model, sp)
    calc = pfw.ACECalculator(model, ps, precision="float64", skin=0.5)
    check(pfw, calc, om, X, H, pbc, sp, labels, "float64")
    base = calc.workspace.rebuilds
    y, back = pfw.rrule_ace_evaluate(calc, pfw.System(X, H, pbc, labels))
    back({"energy": 1.0, "forces": np.ones((3, len(X))), "virial": np.eye(3)})
    check(pfw, calc, om, X + 0.01, H, pbc, sp, labels, "float64")
    assert calc.workspace.rebuilds == base + 1


@pytest.mark.parametrize("prec", PRECS)
def test_batch_with_skin(pfw, synthetic, prec):
    model, ps, om = synthetic
    cfgs, _ = pfw.workloads.fitting_set(ncfg=12, nmin=8, nmax=200, seed=19)
    calc = pfw.ACECalculator(model, ps, precision=prec, skin=0.5)
    systems = [pfw.System(X, H, pbc, labels_of(model, sp)) for X, H, pbc, sp in cfgs]
    res = calc.evaluate_batch(systems)
    tol = TOL[prec]
    for (X, H, pbc, sp), r in zip(cfgs, res):
        E, F, W, _ = oracle_ace(X, sp, H, pbc, om)
        assert abs(r["energy"] - E) <= tol * max(1, abs(E))
        np.testing.assert_allclose(r["forces"], F.T, atol=tol * max(1, np.abs(F).max()), rtol=tol)
        np.testing.assert_allclose(r["virial"], W, atol=tol * max(1, abs(E)), rtol=tol)


def test_synthetic_bcc_32k_reused_step(pfw, synthetic):
    """The 32,768-atom workload: a fresh step and a reused step after a
    0.04 Å random displacement, FP64 against the threaded oracle, mixed by the
    force RMSE bound."""
    model, ps, om = synthetic
    X, H, pbc, sp = pfw.workloads.bcc_binary()
    labels = labels_of(model, sp)
    c64 = pfw.ACECalculator(model, ps, precision="float64", skin=0.5)
    cmx = pfw.ACECalculator(model, ps, precision="mixed", skin=0.5)
    rng = np.random.default_rng(23)
    for step in range(2):
        sys_ = pfw.System(X, H, pbc, labels)
        E, F, W, _ = oracle_ace(X, sp, H, pbc, om, threads=16)
        r64 = c64.evaluate(sys_)
        assert abs(r64["energy"] - E) <= 1e-9 * abs(E)
        np.testing.assert_allclose(r64["forces"], F.T, atol=1e-9, rtol=1e-9)
        np.testing.assert_allclose(r64["virial"], W, atol=1e-8 * abs(E), rtol=1e-9)
        rmx = cmx.evaluate(sys_)
        d = rmx["forces"] - F.T
        assert float(np.sqrt(np.mean(d * d))) <= 1e-5
        X = X + rng.uniform(-0.04, 0.04, X.shape)
    assert c64.workspace.verlet_info()["reuses"] == 1 and cmx.workspace.verlet_info()["reuses"] == 1
