"""GPU parity of the linear fitting design (design_matrix_impl,
src/submodules/ACE/fitting.jl:18-63) computed on the device.

Pinned three ways, independently of the way the device builds it:
- linearity: A theta + baseline reproduces the oracle's E, F, W for the
  calculator's readout and for a random one (FP64 to 1e-9, mixed to 1e-4);
- columns: sampled columns (invariant and pair columns of both species) equal
  the oracle evaluated with that unit readout less the oracle with a zero one;
- the baseline equals the oracle with a zero readout (the reference energies),
  and without fit_pair the current pair term as well.
Batches stack the per-configuration blocks; a masked invariant (basis policy)
gives a zero column; the calculator's own parameters are unchanged afterwards.
"""
import copy
import os

import numpy as np
import pytest

from oracle.ace import ace_evaluate as oracle_ace
from oracle.ace import load_resolved as oracle_load
from oracle.ace import read_frames

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def tial(pfw):
    path = os.path.join(GOLD, "resolved_model_v1.toml")
    model, ps, _ = pfw.load_resolved(path)
    return model, ps, oracle_load(path)


@pytest.fixture(scope="module")
def frame():
    frames = read_frames(os.path.join(GOLD, "reference_results.xyz"))
    return max(frames, key=lambda f: len(f["X"]))


def oracle_rows(om, X, sp, H, pbc, WB=None, Wpair=None):
    m = copy.copy(om)
    if WB is not None:
        m["WB"] = WB
    if Wpair is not None:
        m["Wpair"] = Wpair
    E, F, W, _ = oracle_ace(X, sp, H, pbc, m)
    return np.concatenate([[E], F.reshape(-1), W.T.reshape(-1)])


def rows_of(r):
    return np.concatenate([[r["energy"]], r["forces"].T.reshape(-1), r["virial"].T.reshape(-1)])


def check_rows(got, want, tol):
    scale = max(1.0, np.abs(want).max())
    np.testing.assert_allclose(got, want, atol=tol * scale, rtol=0)


@pytest.mark.parametrize("prec,tol", [("float64", 1e-9), ("mixed", 1e-4)])
def test_design_reproduces_outputs(pfw, tial, frame, prec, tol):
    model, ps, om = tial
    sp = np.array(model.channels(frame["labels"]), np.int32)
    calc = pfw.ACECalculator(model, ps, precision=prec)
    sys_ = pfw.System(frame["X"], frame["H"], frame["pbc"], frame["labels"])
    d = pfw.design_matrix(calc, sys_)
    n = len(frame["X"])
    assert d["matrix"].shape == (1 + 3 * n + 9, (model.nbasis + model.pair.nr) * model.nz)
    theta = pfw.linear_parameters(ps)
    want = oracle_rows(om, frame["X"], sp, frame["H"], frame["pbc"])
    check_rows(d["matrix"] @ theta + d["baseline"], want, tol)
    rng = np.random.default_rng(4)
    ps2 = pfw.install_readout(ps, theta * (1 + 0.3 * rng.normal(size=theta.shape)))
    want2 = oracle_rows(om, frame["X"], sp, frame["H"], frame["pbc"], ps2["WB"], ps2["Wpair"])
    check_rows(d["matrix"] @ pfw.linear_parameters(ps2) + d["baseline"], want2, tol)
    # the calculator's parameters are still in force
    check_rows(rows_of(calc.evaluate(sys_)), want, tol)


def test_design_columns_and_baseline(pfw, tial, frame):
    model, ps, om = tial
    sp = np.array(model.channels(frame["labels"]), np.int32)
    calc = pfw.ACECalculator(model, ps, precision="float64")
    sys_ = pfw.System(frame["X"], frame["H"], frame["pbc"], frame["labels"])
    d = pfw.design_matrix(calc, sys_)
    nb, nz, nq = model.nbasis, model.nz, model.pair.nr
    zero_wb, zero_wp = np.zeros((nb, nz)), np.zeros((nq, nz))
    base = oracle_rows(om, frame["X"], sp, frame["H"], frame["pbc"], zero_wb, zero_wp)
    check_rows(d["baseline"], base, 1e-10)
    cols = [0, 7, nb - 1, nb, nb + 40, nz * nb, nz * nb + nq - 1, nz * nb + nq + 3]
    for c in cols:
        wb, wp = zero_wb.copy(), zero_wp.copy()
        if c < nz * nb:
            wb[c % nb, c // nb] = 1.0
        else:
            k = c - nz * nb
            wp[k % nq, k // nq] = 1.0
        want = oracle_rows(om, frame["X"], sp, frame["H"], frame["pbc"], wb, wp) - base
        check_rows(d["matrix"][:, c], want, 1e-9)


def test_design_without_pair_columns(pfw, tial, frame):
    model, ps, om = tial
    sp = np.array(model.channels(frame["labels"]), np.int32)
    calc = pfw.ACECalculator(model, ps, precision="float64")
    sys_ = pfw.System(frame["X"], frame["H"], frame["pbc"], frame["labels"])
    d = pfw.design_matrix(calc, sys_, fit_pair=False)
    assert d["matrix"].shape[1] == model.nbasis * model.nz
    base = oracle_rows(om, frame["X"], sp, frame["H"], frame["pbc"], np.zeros_like(ps["WB"]), ps["Wpair"])
    check_rows(d["baseline"], base, 1e-10)
    want = oracle_rows(om, frame["X"], sp, frame["H"], frame["pbc"])
    check_rows(d["matrix"] @ ps["WB"].T.reshape(-1) + d["baseline"], want, 1e-9)
    with pytest.raises(pfw.ArgumentError):
        pfw.design_matrix(calc, sys_, centers=[0])


def test_design_batch_stacks_configurations(pfw):
    D = pfw.ace_basis.synthetic_model()
    model, ps, _ = pfw.load_resolved(D)
    from pfw_amd.ace import ELEMENTS
    cfgs, _ = pfw.workloads.fitting_set(ncfg=5, nmin=8, nmax=60, seed=29)
    systems = [pfw.System(X, H, pbc, [ELEMENTS[model.atomic_numbers[s] - 1] for s in sp]) for X, H, pbc, sp in cfgs]
    calc = pfw.ACECalculator(model, ps, precision="float64")
    whole = pfw.design_matrix(calc, systems)
    parts = [pfw.design_matrix(calc, s) for s in systems]
    np.testing.assert_allclose(whole["matrix"], np.vstack([p["matrix"] for p in parts]), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(whole["baseline"], np.concatenate([p["baseline"] for p in parts]), rtol=1e-12,
                               atol=1e-12)
    # per configuration: training rows with the reference outputs as labels give zero residual at theta
    theta = pfw.linear_parameters(ps)
    res = calc.evaluate_batch(systems)
    for p, r in zip(parts, res):
        t = pfw.training_rows(p, r, weights={"energy": 1.0, "forces": 2.0, "virial": 0.5})
        np.testing.assert_allclose(t["matrix"] @ theta, t["labels"], atol=1e-9 * max(1, abs(r["energy"])))
        assert t["row_weights"][0] == 1.0 and t["row_weights"][1] == 2.0 and t["row_weights"][-1] == 0.5


def test_masked_invariant_gives_zero_column(pfw, tial, frame):
    import tomli
    model, ps, om = tial
    with open(os.path.join(GOLD, "resolved_model_v1.toml"), "rb") as f:
        D = tomli.load(f)
    mask = [True] * model.nbasis
    mask[2] = mask[5] = False
    D["interface_policy"] = {"version": 1, "mask": mask}
    masked, mps, _ = pfw.load_resolved(D)
    calc = pfw.ACECalculator(masked, mps, precision="float64")
    sys_ = pfw.System(frame["X"], frame["H"], frame["pbc"], frame["labels"])
    d = pfw.design_matrix(calc, sys_)
    nb = model.nbasis
    for c in (2, 5, nb + 2, nb + 5):
        assert not np.any(d["matrix"][:, c])
    assert np.any(d["matrix"][:, 3])


@pytest.mark.parametrize("prec", ["float64", "mixed"])
def test_jacobian_design_equals_unit_readout_sets(pfw, prec):
    """pfw_ace_design (one pass of per-site Jacobians) against the columns of
    the unit-readout parameter sets (a full site pass per column, less the zero
    set), on a five-configuration batch of the synthetic basis with two
    invariants masked: the two device formulations share no kernel beyond the
    environments and densities."""
    from pfw_amd import _capi
    from pfw_amd.ace import ELEMENTS

    D = pfw.ace_basis.synthetic_model()
    model, ps, _ = pfw.load_resolved(D)
    lab = [ELEMENTS[z - 1] for z in model.atomic_numbers]
    cfgs, _ = pfw.workloads.fitting_set(ncfg=5, nmin=8, nmax=90, seed=29)
    systems = [pfw.System(X, H, pb, [lab[s] for s in sp]) for X, H, pb, sp in cfgs]
    calc = pfw.ACECalculator(model, ps, precision=prec)
    w = calc.workspace
    w.prepare_batch(systems, model.cutoff, [model.channels(s.labels) for s in systems])
    nb, nz, nq = model.nbasis, model.nz, model.pair.nr
    scale = np.ones(nb)
    scale[[3, 100]] = 0.0
    rows = sum(10 + 3 * len(s) for s in systems)
    got = np.empty((rows, nb * nz))
    w.ctx.call("pfw_ace_design", w.mode, _capi.ptr(scale), _capi.ptr(got), nb * nz)
    nsets = nb * nz + 1
    WB = np.zeros((nsets, nz, nb))
    WB.reshape(nsets, -1)[np.arange(nb * nz), np.arange(nb * nz)] = np.tile(scale, nz)
    WP = np.zeros((nsets, nz, nq))
    out = np.zeros((nsets, rows))
    w.ctx.call("pfw_ace_parameter_sets", w.mode, nsets, _capi.ptr(WB), _capi.ptr(WP), _capi.ptr(out))
    want = (out[:-1] - out[-1][None, :]).T
    tol = 1e-10 if prec == "float64" else 2e-5
    err = np.abs(got - want).max()
    print(f"{prec}: max |jacobian - sets| = {err:.3e} of max {np.abs(want).max():.3e}")
    np.testing.assert_allclose(got, want, atol=tol * max(1.0, np.abs(want).max()), rtol=0)
    assert not got[:, [3, 100, nb + 3, nb + 100]].any()
    with pytest.raises(pfw.ArgumentError):  # projected bases take the parameter-set path
        pc = pfw.ACECalculator(pfw.project_species(model, lab[0], range(1, 3)), ps, precision=prec)
        pc.workspace.prepare(systems[0], model.cutoff, model.channels(systems[0].labels))
        pc.workspace.ctx.call("pfw_ace_design", pc.workspace.mode, None,
                              _capi.ptr(np.empty((10 + 3 * len(systems[0]), nb * nz))), nb * nz)
