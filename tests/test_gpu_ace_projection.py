"""GPU parity of species-projected, order-restricted ACE bases
(project_species, src/submodules/ACE/interface.jl:115-125): U/V density split
by neighbour species, truncated (U + tV) products (projection.jl:7-38), the
pair term as order 0 or 1 and the reference energy as order 0.

* The device against the oracle (itself pinned by the reference's 2^nu
  enumeration, its identities and finite differences in
  tests/test_oracle_ace.py) on the 33 TiAl frames, the May30 structures and a
  synthetic BCC cell, FP64 to 1e-9 (scaled) and mixed to 1e-4.
* The reference's identities on the device (test/ace_interface.jl:64-79): the
  per-order projections k = 0..3 sum to the model, the all-species and
  empty-species 0:3 projections equal it, the empty-species 1:3 projection
  has no forces, orders 9:10 give zero energy.
"""
import os

import numpy as np
import pytest

from oracle.ace import ace_evaluate as oracle_ace
from oracle.ace import channels as oracle_channels
from oracle.ace import load_resolved as oracle_load
from oracle.ace import read_frames

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = {"float64": 1e-9, "mixed": 1e-4}


@pytest.fixture(scope="module")
def tial(pfw):
    path = os.path.join(GOLD, "resolved_model_v1.toml")
    m, ps, _ = pfw.load_resolved(path)
    return m, ps, oracle_load(path), read_frames(os.path.join(GOLD, "reference_results.xyz"))


def oracle_projected(om, m):
    return dict(om, species=tuple(m.species), orders=m.orders)


def compare(r, E, F, W, tol):
    scale = max(1.0, abs(E))
    assert abs(r["energy"] - E) <= tol * scale
    np.testing.assert_allclose(r["forces"], F.T, atol=tol * max(1.0, np.abs(F).max()), rtol=tol)
    np.testing.assert_allclose(r["virial"], W, atol=tol * scale, rtol=tol)


@pytest.mark.parametrize("prec", ["float64", "mixed"])
def test_tial_frames_projected_vs_oracle(pfw, tial, prec):
    m, ps, om, frames = tial
    for species, orders in ((["Ti"], range(0, 1)), (["Ti"], range(1, 2)), (["Al"], range(2, 4)),
                            (["Ti", "Al"], range(1, 3)), ([], range(0, 1))):
        mp = pfw.project_species(m, species, orders)
        calc = pfw.ACECalculator(mp, ps, precision=prec)
        omp = oracle_projected(om, mp)
        for f in frames[::4]:
            r = calc.evaluate(pfw.System(f["X"], f["H"], f["pbc"], f["labels"]))
            E, F, W, _ = oracle_ace(f["X"], oracle_channels(om, f["labels"]), f["H"], f["pbc"], omp)
            compare(r, E, F, W, TOL[prec])


def test_reference_identities_on_device(pfw, tial):
    m, ps, om, frames = tial
    sums = None
    for f in (frames[0], frames[-1]):
        s = pfw.System(f["X"], f["H"], f["pbc"], f["labels"])
        full = pfw.ACECalculator(m, ps, precision="float64").evaluate(s)
        parts = [pfw.ACECalculator(pfw.project_species(m, ["Ti"], range(k, k + 1)), ps).evaluate(s) for k in range(4)]
        scale = max(1.0, abs(full["energy"]))
        assert sum(p["energy"] for p in parts) == pytest.approx(full["energy"], abs=1e-9 * scale)
        np.testing.assert_allclose(sum(p["forces"] for p in parts), full["forces"], atol=1e-9)
        np.testing.assert_allclose(sum(p["virial"] for p in parts), full["virial"], atol=1e-9 * scale)
        for species in (["Ti", "Al"], []):
            r = pfw.ACECalculator(pfw.project_species(m, species, range(0, 4)), ps).evaluate(s)
            assert r["energy"] == pytest.approx(full["energy"], abs=1e-9 * scale)
            np.testing.assert_allclose(r["forces"], full["forces"], atol=1e-9)
        assert not np.any(pfw.ACECalculator(pfw.project_species(m, [], range(1, 4)), ps).evaluate(s)["forces"])
        assert pfw.ACECalculator(pfw.project_species(m, ["Ti"], range(9, 11)), ps).evaluate(s)["energy"] == 0.0
        sums = parts
    assert any(abs(p["energy"]) > 1e-6 for p in sums[1:])


def test_may30_projected_vs_oracle(pfw):
    from tests.test_oracle_may30 import NAMES, may30_structure
    from pfw_amd.ace import ELEMENTS
    path = os.path.join(GOLD, "may30_resolved.npz")
    m, ps, _ = pfw.load_resolved(path)
    om = oracle_load(pfw.ace.read_resolved_npz(path))
    mp = pfw.project_species(m, ["Ti", "V"], range(1, 3))
    calc = pfw.ACECalculator(mp, ps, precision="float64")
    omp = oracle_projected(om, mp)
    for name in NAMES:
        X, H, pbc, z, _, _, _ = may30_structure(name)
        sp = np.array([om["atomic_numbers"].index(v) for v in z], np.int32)
        r = calc.evaluate(pfw.System(X, H, pbc, [ELEMENTS[v - 1] for v in z]))
        E, F, W, _ = oracle_ace(X, sp, H, pbc, omp, threads=8)
        print(f"May30 {name} projected (Ti, V) 1:2: |dE|={abs(r['energy'] - E):.2e} "
              f"|dF|={np.abs(r['forces'] - F.T).max():.2e}")
        compare(r, E, F, W, 1e-9)


@pytest.mark.parametrize("prec", ["float64", "mixed"])
def test_synthetic_projected_with_skin_and_batch(pfw, prec):
    from pfw_amd.ace import ELEMENTS
    D = pfw.ace_basis.synthetic_model()
    m, ps, _ = pfw.load_resolved(D)
    om = oracle_load(D)
    labels = [ELEMENTS[z - 1] for z in m.atomic_numbers]
    mp = pfw.project_species(m, labels[1], range(1, 3))
    omp = oracle_projected(om, mp)
    X, H, pbc, sp = pfw.workloads.bcc_binary(3, 3, 3, seed=7)
    calc = pfw.ACECalculator(mp, ps, precision=prec, skin=0.5)
    rng = np.random.default_rng(2)
    for step in range(3):
        r = calc.evaluate(pfw.System(X, H, pbc, [labels[s] for s in sp]))
        compare(r, *oracle_ace(X, sp, H, pbc, omp)[:3], TOL[prec])
        X = X + rng.normal(0, 0.03, X.shape)
    cfgs, _ = pfw.workloads.fitting_set(ncfg=5, nmin=8, nmax=80, seed=13)
    res = pfw.ACECalculator(mp, ps, precision=prec).evaluate_batch(
        [pfw.System(Xc, Hc, pc, [labels[s] for s in spc]) for Xc, Hc, pc, spc in cfgs])
    for (Xc, Hc, pc, spc), r in zip(cfgs, res):
        compare(r, *oracle_ace(Xc, spc, Hc, pc, omp)[:3], TOL[prec])


def test_projected_adjoint_is_rejected_and_api_errors(pfw, tial):
    m, ps, _, frames = tial
    f = frames[0]
    calc = pfw.ACECalculator(pfw.project_species(m, ["Ti"], range(1, 3)), ps)
    s = pfw.System(f["X"], f["H"], f["pbc"], f["labels"])
    with pytest.raises(pfw.ArgumentError):
        pfw.ace_parameter_vjp(calc, [s], [1.0], [np.zeros((3, len(f["X"])))])
