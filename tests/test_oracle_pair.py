"""Pins the oracle's neighbour-list and pair restatement (CPU only).

Each check is independent of the oracle's own formulas: the exhaustive image
sum and the O(N^2) minimum-image oracle restated from the reference's tests
(test/periodic_images.jl:3-28, test/fixtures.jl:28-51), closed forms from the
reference's expectations (test/runtests.jl:93-97, periodic_images.jl:96-108),
finite differences and invariances.
"""
import numpy as np
import pytest

from oracle import KIND_LJ, KIND_SMOOTH, OracleError, nl_pairs, pair_evaluate
from oracle.brute import image_oracle, min_image_oracle

A = np.array([[1.0, 0.7], [0.7, 1.3]])
RC = 2.0


def _wl():
    from tests.conftest import load_package
    return load_package().workloads


def directed(pairs):
    out = set()
    for (i, j, s) in pairs:
        out.add((i, j, tuple(s)))
        out.add((j, i, tuple(-k for k in s)))
    return out


def agrees(E, F, W, ref, tol=2e-10):
    assert abs(E - ref["energy"]) <= tol * max(1, abs(ref["energy"]))
    np.testing.assert_allclose(F, ref["forces"], atol=tol, rtol=0)
    np.testing.assert_allclose(W, ref["virial"], atol=tol, rtol=0)


@pytest.mark.parametrize("cell", range(5))
@pytest.mark.parametrize("pbc", [(1, 1, 1), (1, 0, 1), (0, 1, 0)])
def test_general_cells_vs_image_oracle(cell, pbc):
    wl = _wl()
    H = wl.IMAGE_TEST_CELLS[cell]
    X, sp = wl.image_test_system(H)
    ref = image_oracle(X, H, pbc, sp, A, RC)
    assert ref["pairs"] == image_oracle(X, H, pbc, sp, A, RC, shell=12)["pairs"]
    for bf in (0, 32):
        for threads in (1, 3):
            E, F, W = pair_evaluate(X, sp, H, pbc, KIND_SMOOTH, A, RC, bruteforce_threshold=bf, threads=threads)
            agrees(E, F, W, ref)
        i, j, S, R, info = nl_pairs(X, H, pbc, RC, bruteforce_threshold=bf)
        got = {(int(a), int(b), tuple(int(t) for t in s)) for a, b, s in zip(i, j, S)}
        assert len(got) == len(i)  # no duplicates
        assert got == directed(ref["pairs"])
        # R = x_j - x_i + H S for every directed pair
        np.testing.assert_allclose(R, X[j] - X[i] + S @ H.T, atol=1e-9)


def test_one_atom_chain_closed_form():
    H = np.array([[1.2, 0, 0], [0, 0, 0], [0, 0, 0.]])
    E, F, W = pair_evaluate(np.zeros((1, 3)), [0], H, (1, 0, 0), KIND_SMOOTH, np.ones((1, 1)), RC)
    assert E == pytest.approx((1 - (1.2 / 2) ** 2) ** 3, abs=1e-14)
    assert np.all(F == 0)
    assert W[0, 0] > 0
    i, j, S, R, _ = nl_pairs(np.zeros((1, 3)), H, (1, 0, 0), RC)
    assert sorted(S[:, 0].tolist()) == [-1, 1]


def test_slab_with_infinite_open_vector():
    H = np.array([[3., 1., 0], [0, 2.5, 0], [0, 0, np.inf]])
    X = np.array([[0., 0, 0], [0.5, 0.2, 0.8]])
    sp = np.array([0, 1])
    ref = image_oracle(X, np.where(np.isfinite(H), H, 0.0), (1, 1, 0), sp, A, RC)
    agrees(*pair_evaluate(X, sp, H, (1, 1, 0), KIND_SMOOTH, A, RC), ref)


@pytest.mark.parametrize("pbc", [(0, 0, 0), (1, 0, 1), (1, 1, 1)])
def test_min_image_and_open_vs_quadratic_oracle(pbc):
    wl = _wl()
    X, H, _, sp = wl.fixture(17, pbc=pbc)
    ref = min_image_oracle(X, H, pbc, sp, A, RC)
    for bf in (0, 16, 17, 32):
        E, F, W = pair_evaluate(X, sp, H, pbc, KIND_SMOOTH, A, RC, bruteforce_threshold=bf, threads=2)
        agrees(E, F, W, ref, tol=1e-10)
        i, j, S, R, info = nl_pairs(X, H, pbc, RC, bruteforce_threshold=bf)
        assert {(min(a, b), max(a, b)) for a, b in zip(i.tolist(), j.tolist())} == ref["pairs"]


@pytest.mark.parametrize("n", [0, 1, 3, 35])
def test_small_counts_skew_cell(n):
    H = np.array([[1.2, 0.4, 0.], [0, 1.3, 0.3], [0, 0, 1.1]])
    frac = np.array([[np.mod(np.sin((i + 1) * (a + 1)) * 17.3, 1) for a in range(3)] for i in range(n)]).reshape(n, 3)
    X = frac @ H.T
    sp = np.zeros(n, np.int32)
    ref = image_oracle(X, H, (1, 1, 1), sp, A, RC)
    for bf in (0, 32):
        agrees(*pair_evaluate(X, sp, H, (1, 1, 1), KIND_SMOOTH, A, RC, bruteforce_threshold=bf), ref, tol=1e-9)


def test_closed_form_pairs():
    # runtests.jl:93-97: H/He pair 1 Å apart, amplitude 0.7, cutoff 2
    X = np.array([[0., 0, 0], [1., 0, 0]])
    E, F, W = pair_evaluate(X, [0, 1], np.eye(3) * 10, (1, 0, 1), KIND_SMOOTH, A, RC)
    assert E == pytest.approx(0.7 * (1 - 1 / 4) ** 3, abs=1e-15)
    g = -6 * 0.7 * (1 - 0.25) ** 2 / 4.0  # dE/dr / r
    np.testing.assert_allclose(F[0], [g * 1.0, 0, 0], atol=1e-15)
    np.testing.assert_allclose(F[1], [-g * 1.0, 0, 0], atol=1e-15)
    E, _, _ = pair_evaluate(X, [0, 0], np.zeros((3, 3)), (0, 0, 0), KIND_SMOOTH, np.ones((1, 1)), RC)
    assert E == pytest.approx((1 - 1 / 4) ** 3, abs=1e-15)


def test_lennard_jones_closed_form_and_fd():
    eps, sig, rc = 0.0104, 3.40, 8.5

    def lj(r):
        q, qc = (sig / r) ** 6, (sig / rc) ** 6
        return 4 * eps * (q * q - q - qc * qc + qc)

    for r in (3.3, 3.8, 5.0, rc - 1e-5, rc, rc + 1e-5, 12.0):
        X = np.array([[0., 0, 0], [r, 0, 0]])
        E, F, W = pair_evaluate(X, [0, 0], np.zeros((3, 3)), (0, 0, 0), KIND_LJ, (eps, sig), rc)
        assert E == pytest.approx(lj(r) if r < rc else 0.0, rel=1e-13, abs=1e-16)
    h, r = 1e-5, 3.8
    X = np.array([[0., 0, 0], [r, 0, 0]])
    _, F, _ = pair_evaluate(X, [0, 0], np.zeros((3, 3)), (0, 0, 0), KIND_LJ, (eps, sig), rc)
    assert F[1, 0] == pytest.approx(-(lj(r + h) - lj(r - h)) / (2 * h), rel=1e-8)
    with pytest.raises(OracleError) as e:
        pair_evaluate(np.zeros((2, 3)), [0, 0], np.zeros((3, 3)), (0, 0, 0), KIND_LJ, (eps, sig), rc)
    assert e.value.code == 3


def test_forces_and_strain_virial_by_finite_differences():
    H = np.array([[1.7, 0.6, 0.2], [0, 1.8, 0.4], [0, 0, 1.6]])
    X = (H @ np.array([[0.1, 0.4, 0.8], [0.2, 0.6, 0.3], [0.1, 0.7, 0.5]])).T
    sp = np.array([0, 1, 0])
    h = 1e-5
    for bf in (0, 32):
        E, F, W = pair_evaluate(X, sp, H, (1, 1, 1), KIND_SMOOTH, A, RC, bruteforce_threshold=bf)
        for i in range(3):
            for a in range(3):
                xp, xm = X.copy(), X.copy()
                xp[i, a] += h
                xm[i, a] -= h
                ep = pair_evaluate(xp, sp, H, (1, 1, 1), KIND_SMOOTH, A, RC, bruteforce_threshold=bf, mask=1)[0]
                em = pair_evaluate(xm, sp, H, (1, 1, 1), KIND_SMOOTH, A, RC, bruteforce_threshold=bf, mask=1)[0]
                assert F[i, a] == pytest.approx(-(ep - em) / (2 * h), abs=2e-8)
        for strain in (np.eye(3), np.array([[0, 1., 0], [0, 0, 0], [0, 0, 0]]), np.array([[0, 0, 0], [0, 0, 1.], [0, 0, 0]])):
            ap, am = np.eye(3) + h * strain, np.eye(3) - h * strain
            ep = pair_evaluate(X @ ap.T, sp, ap @ H, (1, 1, 1), KIND_SMOOTH, A, RC, bruteforce_threshold=bf, mask=1)[0]
            em = pair_evaluate(X @ am.T, sp, am @ H, (1, 1, 1), KIND_SMOOTH, A, RC, bruteforce_threshold=bf, mask=1)[0]
            assert (ep - em) / (2 * h) == pytest.approx(-np.sum(W * strain), abs=3e-8)


def test_invariances_and_momentum():
    wl = _wl()
    X, H, pbc, sp = wl.fixture(93, pbc=(1, 1, 1))
    E, F, W = pair_evaluate(X, sp, H, pbc, KIND_SMOOTH, A, RC)
    np.testing.assert_allclose(F.sum(0), 0, atol=1e-12)
    E2, F2, _ = pair_evaluate(X + H[:, 0] * 2 + np.array([0.3, -0.2, 0.1]), sp, H, pbc, KIND_SMOOTH, A, RC)
    assert E2 == pytest.approx(E, abs=1e-10)
    perm = np.arange(len(X))[::-1]
    E3, F3, _ = pair_evaluate(X[perm], sp[perm], H, pbc, KIND_SMOOTH, A, RC)
    assert E3 == pytest.approx(E, abs=1e-10)
    np.testing.assert_allclose(F3, F[perm], atol=1e-10)


def test_error_contract():
    wl = _wl()
    X, H, pbc, sp = wl.fixture(5)
    Hd = H.copy()
    Hd[:, 1] = Hd[:, 0]
    with pytest.raises(OracleError):
        pair_evaluate(X, sp, Hd, pbc, KIND_SMOOTH, A, RC)
    Hi = H.copy()
    Hi[0, 0] = np.inf
    with pytest.raises(OracleError):
        pair_evaluate(X, sp, Hi, pbc, KIND_SMOOTH, A, RC)
    Xn = X.copy()
    Xn[0, 0] = np.nan
    with pytest.raises(OracleError):
        pair_evaluate(Xn, sp, H, pbc, KIND_SMOOTH, A, RC)
