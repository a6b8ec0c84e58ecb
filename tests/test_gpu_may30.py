"""GPU parity on the reference's May30 Ti-Al-V model (a fitted real model:
real solid harmonics to l = 6, PolyEnvelope1sR pair envelope, 3 species,
2,155 invariants, 5,652 products, 1,000-node splines) against its
ACEpotentials 0.10.3 predictions for three nonorthogonal periodic structures.

FP64 at the reference's own tolerances (test/ace_may30.jl:3-7: E 2e-8 eV,
F 2e-9 eV/Å + rtol 2e-11, W 2e-8 eV + rtol 2e-11); mixed precision at E 1e-3 eV
per structure, F 1e-4 eV/Å, W 2e-3 eV. The reference's own GPU packing rejects
this model (adaptation.jl:72,79).
"""
import os

import numpy as np
import pytest

from tests.test_oracle_may30 import GOLD, NAMES, may30_structure

pytestmark = pytest.mark.gpu

TOL = {"float64": (2e-8, 2e-9, 2e-11, 2e-8, 2e-11), "mixed": (1e-3, 1e-4, 0, 2e-3, 0)}


@pytest.fixture(scope="module")
def model(pfw):
    return pfw.load_resolved(os.path.join(GOLD, "may30_resolved.npz"))


@pytest.mark.parametrize("prec", ["float64", "mixed"])
def test_may30_predictions(pfw, model, prec):
    m, ps, _ = model
    assert m.solid and m.lmax == 6
    calc = pfw.ACECalculator(m, ps, precision=prec)
    from pfw_amd.ace import ELEMENTS
    etol, ftol, frtol, wtol, wrtol = TOL[prec]
    for name in NAMES:
        X, H, pbc, z, E, F, W = may30_structure(name)
        labels = [ELEMENTS[v - 1] for v in z]
        r = calc.evaluate(pfw.System(X, H, pbc, labels))
        print(f"May30 {name} {prec}: |dE|={abs(r['energy'] - E):.2e} |dF|={np.abs(r['forces'] - F.T).max():.2e} "
              f"|dW|={np.abs(r['virial'] - W).max():.2e}")
        assert abs(r["energy"] - E) <= etol
        np.testing.assert_allclose(r["forces"], F.T, atol=ftol, rtol=frtol)
        np.testing.assert_allclose(r["virial"], W, atol=wtol, rtol=wrtol)


def test_may30_skin_step(pfw, model):
    """Environment reuse on the real model: a moved structure against the
    oracle at the moved positions."""
    from oracle.ace import ace_evaluate as oracle_ace
    from oracle.ace import load_resolved as oracle_load
    from pfw_amd.ace import ELEMENTS
    m, ps, _ = model
    om = oracle_load(pfw.ace.read_resolved_npz(os.path.join(GOLD, "may30_resolved.npz")))
    X, H, pbc, z, E, F, W = may30_structure("omega-240")
    labels = [ELEMENTS[v - 1] for v in z]
    sp = np.array([om["atomic_numbers"].index(v) for v in z], np.int32)
    calc = pfw.ACECalculator(m, ps, precision="float64", skin=0.4)
    rng = np.random.default_rng(3)
    for step in range(3):
        r = calc.evaluate(pfw.System(X, H, pbc, labels))
        e, f, w, _ = oracle_ace(X, sp, H, pbc, om, threads=8)
        assert abs(r["energy"] - e) <= 2e-8
        np.testing.assert_allclose(r["forces"], f.T, atol=2e-9, rtol=2e-11)
        X = X + rng.normal(0, 0.02, X.shape)
    assert calc.workspace.verlet_info()["reuses"] >= 1
