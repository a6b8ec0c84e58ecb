"""Pins the oracle's ACE restatement (CPU only).

* End-to-end: the 33 TiAl frames against the ACEpotentials 0.10.2 predictions
  with the reference's tolerances (test/neighbour_saved_reference.jl:38-40),
  plus the SHA-256 chain that ties the model, geometry and predictions together
  (test/neighbour_saved_reference.jl:28-29). Covers Agnesi splines, both
  envelope kinds, Y_lm up to l = 2, products to order 3, coupling, pair term
  and reference energies.
* Y_lm for l <= 6 (beyond the fixture): closed form from scipy's complex
  spherical harmonics converted to the real, L2, no-Condon-Shortley convention
  (feature_coupling.jl:22-33), quadrature orthonormality, gradients by central
  differences.
* The synthetic order-3 / lmax-4 model: forces and strain virial by finite
  differences, rotation and permutation invariance, and the generator's own
  invariants (degree bound, orders, linear independence, CG orthogonality).
"""
import ctypes
import hashlib
import os

import numpy as np
import pytest
import scipy.special as sps

from oracle.ace import ace_evaluate, channels, load_resolved, read_frames
from oracle.core import lib

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

try:
    import tomllib as toml
except ImportError:
    import tomli as toml


def _pfw():
    from tests.conftest import load_package
    return load_package()


@pytest.fixture(scope="module")
def tial():
    return load_resolved(os.path.join(GOLD, "resolved_model_v1.toml"))


@pytest.fixture(scope="module")
def synthetic():
    D = _pfw().ace_basis.synthetic_model()
    return D, load_resolved(D)


def test_sha_chain():
    with open(os.path.join(GOLD, "reloaded_predictions.toml"), "rb") as f:
        ref = toml.load(f)
    sha = lambda p: hashlib.sha256(open(os.path.join(GOLD, p), "rb").read()).hexdigest()
    assert sha("reference_model.json") == ref["model_sha256"]
    assert sha("reference_results.xyz") == ref["geometry_sha256"]
    with open(os.path.join(GOLD, "resolved_model_v1.toml"), "rb") as f:
        assert toml.load(f)["source_sha256"] == ref["model_sha256"]


def test_tial_33_frames_against_acepotentials(tial):
    frames = read_frames(os.path.join(GOLD, "reference_results.xyz"))
    with open(os.path.join(GOLD, "reloaded_predictions.toml"), "rb") as f:
        preds = toml.load(f)["frames"]
    assert len(frames) == len(preds) == 33
    for fr, ref in zip(frames, preds):
        sp = channels(tial, fr["labels"])
        for threads in (1, 3):
            E, F, W, site = ace_evaluate(fr["X"], sp, fr["H"], fr["pbc"], tial, threads=threads)
            assert E == pytest.approx(ref["energy_eV"], abs=1e-8, rel=0)
            np.testing.assert_allclose(F, np.array(ref["forces_eV_per_A"]), atol=1e-9, rtol=1e-10)
            np.testing.assert_allclose(W, np.array(ref["virial_eV_columns"]).T, atol=1e-8, rtol=1e-10)
            assert site.sum() == pytest.approx(E, abs=1e-9)


def oracle_ylm(L, R, solid=0):
    ny = (L + 1) ** 2
    Y = np.zeros(ny)
    dY = np.zeros(3 * ny)
    R = np.ascontiguousarray(R, float)
    rc = lib.oracle_ylm(L, int(solid), R.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                        Y.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                        dY.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0
    return Y, dY.reshape(ny, 3)


def scipy_real_ylm(L, R):
    r = np.linalg.norm(R)
    theta = np.arccos(R[2] / r)
    phi = np.arctan2(R[1], R[0])
    Y = np.zeros((L + 1) ** 2)
    for l in range(L + 1):
        for m in range(0, l + 1):
            c = sps.sph_harm_y(l, m, theta, phi)  # Condon-Shortley phase included
            if m == 0:
                Y[l * l + l] = c.real
            else:
                Y[l * l + l + m] = np.sqrt(2) * (-1) ** m * c.real
                Y[l * l + l - m] = np.sqrt(2) * (-1) ** m * c.imag
    return Y


def test_ylm_closed_form_and_gradients():
    rng = np.random.default_rng(3)
    L = 6
    for _ in range(40):
        R = rng.normal(size=3) * rng.uniform(0.5, 3.0)
        Y, dY = oracle_ylm(L, R)
        np.testing.assert_allclose(Y, scipy_real_ylm(L, R), atol=1e-12)
        h = 1e-6
        for c in range(3):
            Rp, Rm = R.copy(), R.copy()
            Rp[c] += h
            Rm[c] -= h
            fd = (oracle_ylm(L, Rp)[0] - oracle_ylm(L, Rm)[0]) / (2 * h)
            np.testing.assert_allclose(dY[:, c], fd, atol=2e-7)
    # L = 1 convention: sqrt(3/4pi) (y, z, x) / r
    Y, _ = oracle_ylm(1, np.array([0.3, -0.4, 1.2]))
    u = np.array([0.3, -0.4, 1.2]) / np.linalg.norm([0.3, -0.4, 1.2])
    np.testing.assert_allclose(Y[1:4], np.sqrt(3 / (4 * np.pi)) * u[[1, 2, 0]], atol=1e-14)


def test_ylm_orthonormal_by_quadrature():
    L = 4
    x, wx = np.polynomial.legendre.leggauss(24)
    nphi = 48
    G = np.zeros(((L + 1) ** 2, (L + 1) ** 2))
    for ct, w in zip(x, wx):
        st = np.sqrt(1 - ct * ct)
        for k in range(nphi):
            ph = 2 * np.pi * k / nphi
            Y, _ = oracle_ylm(L, np.array([st * np.cos(ph), st * np.sin(ph), ct]))
            G += w * (2 * np.pi / nphi) * np.outer(Y, Y)
    np.testing.assert_allclose(G, np.eye((L + 1) ** 2), atol=1e-12)


def small_bcc(seed=21):
    return _pfw().workloads.bcc_binary(3, 3, 2, seed=seed)


def test_synthetic_forces_and_virial_by_finite_differences(synthetic):
    _, m = synthetic
    X, H, pbc, sp = small_bcc()
    E, F, W, _ = ace_evaluate(X, sp, H, pbc, m)
    h = 1e-5
    rng = np.random.default_rng(0)
    for i in rng.choice(len(X), 6, replace=False):
        for a in range(3):
            xp, xm = X.copy(), X.copy()
            xp[i, a] += h
            xm[i, a] -= h
            ep = ace_evaluate(xp, sp, H, pbc, m, mask=1)[0]
            em = ace_evaluate(xm, sp, H, pbc, m, mask=1)[0]
            assert F[i, a] == pytest.approx(-(ep - em) / (2 * h), abs=2e-7)
    for strain in (np.eye(3), np.array([[0, 1., 0], [0, 0, 0], [0, 0, 0]]), np.array([[0, 0, 0], [0, 0, 1.], [0, 0, 0]])):
        ap, am = np.eye(3) + h * strain, np.eye(3) - h * strain
        ep = ace_evaluate(X @ ap.T, sp, ap @ H, pbc, m, mask=1)[0]
        em = ace_evaluate(X @ am.T, sp, am @ H, pbc, m, mask=1)[0]
        assert (ep - em) / (2 * h) == pytest.approx(-np.sum(W * strain), abs=5e-7)


def test_tial_frame_forces_by_finite_differences(tial):
    fr = read_frames(os.path.join(GOLD, "reference_results.xyz"))[-1]
    sp = channels(tial, fr["labels"])
    E, F, W, _ = ace_evaluate(fr["X"], sp, fr["H"], fr["pbc"], tial)
    h = 1e-4  # |E| ~ 4.6e4 eV: a smaller step is dominated by cancellation
    for i in (0, 7, 31):
        for a in range(3):
            xp, xm = fr["X"].copy(), fr["X"].copy()
            xp[i, a] += h
            xm[i, a] -= h
            ep = ace_evaluate(xp, sp, fr["H"], fr["pbc"], tial, mask=1)[0]
            em = ace_evaluate(xm, sp, fr["H"], fr["pbc"], tial, mask=1)[0]
            assert F[i, a] == pytest.approx(-(ep - em) / (2 * h), abs=1e-6)


def rotation(seed):
    q = np.random.default_rng(seed).normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def test_synthetic_rotation_and_permutation_invariance(synthetic):
    _, m = synthetic
    X, H, pbc, sp = small_bcc(seed=5)
    E, F, W, site = ace_evaluate(X, sp, H, pbc, m)
    for seed in (1, 2):
        Q = rotation(seed)
        Er, Fr, Wr, sr = ace_evaluate(X @ Q.T, sp, Q @ H, pbc, m)
        assert Er == pytest.approx(E, abs=1e-9 * abs(E))
        np.testing.assert_allclose(sr, site, atol=1e-9)
        np.testing.assert_allclose(Fr, F @ Q.T, atol=1e-9)
        np.testing.assert_allclose(Wr, Q @ W @ Q.T, atol=1e-8)
    perm = np.random.default_rng(9).permutation(len(X))
    Ep, Fp, _, sp_ = ace_evaluate(X[perm], sp[perm], H, pbc, m)
    assert Ep == pytest.approx(E, abs=1e-9)
    np.testing.assert_allclose(Fp, F[perm], atol=1e-10)
    np.testing.assert_allclose(F.sum(0), 0, atol=1e-9)


def test_generator_invariants():
    ab = _pfw().ace_basis
    rspec = [(n, l) for l in range(5) for n in range(1, 9)]
    density, products, (rows, cols, vals, nb) = ab.generate_invariant_tables(rspec, 3, 4, 12)
    weight = [n + l for (n, l) in rspec]
    assert max(len(p) for p in products) == 3
    assert all(sum(weight[density[a][0]] for a in p) <= 12 for p in products)
    C = np.zeros((nb, len(products)))
    C[rows, cols] = vals
    assert np.linalg.matrix_rank(C) == nb  # linearly independent invariants
    # every invariant is normalised under the multiplet inner product (one m component)
    np.testing.assert_allclose((C * C).sum(1), 1.0, atol=1e-10)
    # real CG tensors: orthogonality over (m1, m2) for fixed (a, b)
    for a, b in ((1, 1), (2, 1), (2, 2), (3, 2)):
        blocks = [ab.real_cg(a, b, L).reshape(2 * L + 1, -1) for L in range(abs(a - b), a + b + 1)]
        M = np.vstack(blocks)
        np.testing.assert_allclose(M @ M.T, np.eye(M.shape[0]), atol=1e-12)
    c110 = ab.real_cg(1, 1, 0)[0]
    np.testing.assert_allclose(np.abs(c110), np.eye(3) / np.sqrt(3), atol=1e-14)


def test_solid_harmonics_are_r_l_times_spherical():
    """Real solid harmonics ("real_solid", angular_blocks.jl:47-48) are r^l Y_lm,
    with gradient r^l dY_lm + l r^(l-2) R Y_lm: the reference's own check
    (test/ace_may30.jl:22-29) on its three radii of one direction, l <= 6."""
    for r in (0.3, 1.0, 2.7):
        R = r * np.array([0.3, -0.7, 0.9])
        Ys, dYs = oracle_ylm(6, R)
        Y, dY = oracle_ylm(6, R, solid=1)
        rn = np.linalg.norm(R)
        for l in range(7):
            for k in range(l * l, (l + 1) ** 2):
                g = np.zeros(3) if l == 0 else l * rn ** (l - 2) * R
                np.testing.assert_allclose(Y[k], rn ** l * Ys[k], atol=2e-11, rtol=2e-12)
                np.testing.assert_allclose(dY[k], rn ** l * dYs[k] + g * Ys[k], atol=2e-11, rtol=2e-12)


# ---- species projection and order restriction (interface.jl:115-125,
# projection.jl:7-38, serial.jl:76-79), pinned by the reference's own checks

def test_projected_product_against_subset_enumeration():
    """The 2^nu factor-assignment enumeration of test/ace_interface.jl:316-330:
    the order-k part of prod(U + t V) sums the products taking k factors from
    V. The reverse is pinned by central differences (exact for polynomials)."""
    from oracle.ace import projected_product
    products = [[0, 1], [0, 0, 1], [1], [0, 2, 0], [2, 1, 0, 2]]
    for U in ([0., 2., -0.3], [1.2, -0.7, 0.2]):
        for V in ([3., 0., 0.4], [0., -0.2, 1.1]):
            for lo, hi in ((0, 0), (1, 1), (2, 2), (3, 3), (4, 4), (1, 3), (0, 4), (2, 9)):
                for idx in products:
                    expect = 0.0
                    for bits in range(2 ** len(idx)):
                        if lo <= bin(bits).count("1") <= hi:
                            expect += np.prod([V[a] if (bits >> t) & 1 else U[a] for t, a in enumerate(idx)])
                    P, dU, dV = projected_product(idx, U, V, lo, hi, w=0.7)
                    assert P == pytest.approx(expect, abs=1e-13)
                    h = 1e-3
                    for a in range(3):
                        for arr, grad in ((U, dU), (V, dV)):
                            up, um = list(arr), list(arr)
                            up[a] += h
                            um[a] -= h
                            args = (up, V) if arr is U else (U, up)
                            argm = (um, V) if arr is U else (U, um)
                            fd = (projected_product(idx, *args, lo, hi)[0] - projected_product(idx, *argm, lo, hi)[0]) / (2 * h)
                            assert grad[a] == pytest.approx(0.7 * fd, abs=1e-9)


def projected(model, species_channels, lo, hi):
    return dict(model, species=tuple(species_channels), orders=(lo, hi))


def tial_cell(tial):
    """The reference's three-atom skew cell (test/ace_interface.jl:331-333)."""
    H = np.array([[3.7, 0.6, 0.1], [0, 3.8, 0.2], [0, 0, 3.6]])
    X = np.array([[0.1, 2.3, 1.1], [0.2, 0.7, 2.4], [0.3, 1.2, 1.5]]).T
    return X, channels(tial, ["Ti", "Al", "Ti"]), H, (True, True, True)


def test_projection_identities(tial):
    """test/ace_interface.jl:64-79: the per-order projections k = 0..3 on Ti sum
    to the model; Ti+Al over 0:3 and the empty set over 0:3 equal it; the
    empty set over 1:3 has no forces; orders 9:10 give zero energy."""
    frames = read_frames(os.path.join(GOLD, "reference_results.xyz"))
    ti = tial["atomic_numbers"].index(22)
    al = tial["atomic_numbers"].index(13)
    for X, sp, H, pbc in [tial_cell(tial)] + [(f["X"], channels(tial, f["labels"]), f["H"], f["pbc"])
                                              for f in (frames[0], frames[-1])]:
        E, F, W, _ = ace_evaluate(X, sp, H, pbc, tial)
        parts = [ace_evaluate(X, sp, H, pbc, projected(tial, [ti], k, k)) for k in range(4)]
        scale = max(1.0, abs(E))
        assert sum(p[0] for p in parts) == pytest.approx(E, abs=1e-9 * scale)
        np.testing.assert_allclose(sum(p[1] for p in parts), F, atol=1e-9)
        np.testing.assert_allclose(sum(p[2] for p in parts), W, atol=1e-9 * scale)
        for m in (projected(tial, [ti, al], 0, 3), projected(tial, [], 0, 3)):
            E2, F2, W2, _ = ace_evaluate(X, sp, H, pbc, m)
            assert E2 == pytest.approx(E, abs=1e-9 * scale)
            np.testing.assert_allclose(F2, F, atol=1e-9)
        assert not np.any(ace_evaluate(X, sp, H, pbc, projected(tial, [], 1, 3))[1])
        assert ace_evaluate(X, sp, H, pbc, projected(tial, [ti], 9, 10))[0] == 0.0
        # at least one higher-order part is nonzero, so the split is exercised
        assert any(abs(p[0]) > 1e-6 for p in parts[1:])


def test_zero_order_projection_equals_excluded_edges(tial):
    """test/ace_interface.jl:334-343: orders 0:0 on Ti keep only what involves
    no Ti neighbour, i.e. the model with every edge to a Ti neighbour dropped."""
    ti = tial["atomic_numbers"].index(22)
    X, sp, H, pbc = tial_cell(tial)
    z = ace_evaluate(X, sp, H, pbc, projected(tial, [ti], 0, 0))
    x = ace_evaluate(X, sp, H, pbc, dict(tial, exclude=ti))
    assert z[0] == pytest.approx(x[0], abs=1e-10)
    np.testing.assert_allclose(z[1], x[1], atol=1e-10)
    np.testing.assert_allclose(z[2], x[2], atol=1e-10)


def test_projected_forces_by_finite_differences(tial):
    ti = tial["atomic_numbers"].index(22)
    X, sp, H, pbc = tial_cell(tial)
    for lo, hi in ((1, 2), (2, 3)):
        m = projected(tial, [ti], lo, hi)
        E, F, W, _ = ace_evaluate(X, sp, H, pbc, m)
        h = 1e-5
        for i in range(3):
            for a in range(3):
                xp, xm = X.copy(), X.copy()
                xp[i, a] += h
                xm[i, a] -= h
                fd = -(ace_evaluate(xp, sp, H, pbc, m, mask=1)[0] - ace_evaluate(xm, sp, H, pbc, m, mask=1)[0]) / (2 * h)
                assert F[i, a] == pytest.approx(fd, abs=3e-6)


def test_coloured_site_execution_matches_serial(synthetic):
    """The coloured phase executor (execute_sites!, execution.jl:153-167, with
    color_environment!'s greedy footprint colouring) used as the threaded CPU
    baseline gives the serial result to rounding."""
    _, m = synthetic
    X, H, pbc, sp = small_bcc(seed=9)
    E, F, W, site = ace_evaluate(X, sp, H, pbc, m)
    Ec, Fc, Wc, sitec = ace_evaluate(X, sp, H, pbc, m, threads=4, coloured=True)
    assert Ec == pytest.approx(E, abs=1e-10 * abs(E))
    np.testing.assert_allclose(Fc, F, atol=1e-11)
    np.testing.assert_allclose(Wc, W, atol=1e-10 * abs(E))
    np.testing.assert_array_equal(sitec, site)


def test_low_precision_input_study(synthetic):
    """Precision study for MFMA input types: rounding R, R', Y and grad Y to
    binary32 keeps the forces of the synthetic BCC cell within 1e-6 (RMS) of
    FP64; bfloat16 inputs miss the 1e-5 eV/A force target by orders of
    magnitude (and binary16 by two), so a BF16/FP16-input contraction cannot
    meet the accuracy target. The rounding leaves the FP64 path untouched."""
    from oracle.ace import set_input_rounding

    _, m = synthetic
    X, H, pbc, sp = small_bcc(seed=9)
    E, F, W, _ = ace_evaluate(X, sp, H, pbc, m)
    rms = {}
    try:
        for kind in ("float32", "float16", "bfloat16"):
            set_input_rounding(kind)
            _, Fk, _, _ = ace_evaluate(X, sp, H, pbc, m)
            rms[kind] = float(np.sqrt(np.mean((Fk - F) ** 2)))
    finally:
        set_input_rounding("none")
    assert 0 < rms["float32"] < 1e-6
    assert rms["float16"] > 1e-4 and rms["bfloat16"] > 1e-3
    assert rms["bfloat16"] > 4 * rms["float16"]  # 8 against 11 significant bits
    E2, F2, _, _ = ace_evaluate(X, sp, H, pbc, m)
    assert E2 == E and np.array_equal(F2, F)
