"""GPU parity of the cell list and the pair E/F/V kernels against the oracle.

Pair sets (i, j, S) and displacements R must match the oracle bit for bit in
both precisions. FP64 energies/forces/virials agree to 1e-10 (scaled); the
mixed path (float arithmetic, double accumulation) to 3e-5 absolute+relative,
the reference's own FP32 GPU tolerance (test/gpu_pairs.jl:19-37).
"""
import numpy as np
import pytest

from oracle import KIND_LJ, KIND_SMOOTH, nl_pairs, pair_evaluate

pytestmark = pytest.mark.gpu

A = np.array([[1.0, 0.7], [0.7, 1.3]])
RC = 2.0
TOL = {"float64": 1e-10, "mixed": 3e-5}


@pytest.fixture(scope="module")
def ws(pfw):
    return {p: pfw.Workspace(precision=p) for p in ("float64", "mixed")}


def sorted_pairs(i, j, S, R):
    i, j, S = np.asarray(i, np.int64), np.asarray(j, np.int64), np.asarray(S, np.int64)
    o = np.lexsort((S[:, 2], S[:, 1], S[:, 0], j, i))
    return i[o], j[o], S[o], np.asarray(R)[o]


def assert_same_pairs(gpu, ref):
    gi, gj, gS, gR = sorted_pairs(*gpu)
    ri, rj, rS, rR = sorted_pairs(*ref)
    assert len(gi) == len(ri)
    assert np.array_equal(gi, ri) and np.array_equal(gj, rj) and np.array_equal(gS, rS)
    assert np.array_equal(gR, rR)  # bitwise: same FP64 formula, no contraction


def labels_of(sp):
    return ["H" if s == 0 else "He" for s in sp]


def systems(pfw):
    wl = pfw.workloads
    out = []
    for k, H in enumerate(wl.IMAGE_TEST_CELLS):
        X, sp = wl.image_test_system(H)
        for pbc in ((1, 1, 1), (1, 0, 1), (0, 1, 0)):
            out.append((f"cell{k}-{pbc}", X, H, pbc, sp))
    for n in (0, 1, 2, 7, 65, 257, 700):  # 700: a cell of ~350 atoms (the large-cell sort)
        for pbc in ((0, 0, 0), (1, 0, 1), (1, 1, 1)):
            X, H, _, sp = wl.fixture(n, pbc=pbc, cluster=True)
            out.append((f"fixture{n}-{pbc}", X, H, pbc, sp))
    X, H, pbc, sp = wl.fixture(65, L=4.1)  # two cells on a periodic axis
    out.append(("two-cell", X, H, pbc, sp))
    out.append(("two-cell-translated", X + np.array([20.5, -12.3, 8.2]), H, pbc, sp))
    X, H, pbc, sp = wl.fixture(7, pbc=(0, 0, 0))
    X[-1] += 1e6  # sparse cluster with a far atom
    out.append(("sparse", X, H, pbc, sp))
    X, H, pbc, sp = wl.b2_triclinic(3, 3, 6, seed=5)
    out.append(("b2-small", X, H, pbc, sp))
    return out


@pytest.fixture(scope="module")
def cases(pfw):
    return systems(pfw)


def test_pair_sets_bit_exact(pfw, ws, cases):
    for name, X, H, pbc, sp in cases:
        rc = 6.0 if name == "b2-small" else RC
        sys = pfw.System(X, H, pbc, labels_of(sp))
        for prec in ("float64", "mixed"):
            w = ws[prec]
            w.prepare(sys, rc, sp)
            gpu = w.export_pairs()
            i, j, S, R, _ = nl_pairs(X, H, pbc, rc)
            assert_same_pairs(gpu, (i, j, S, R)), name


def compare(res, E, F, W, tol, what=("energy", "forces", "virial")):
    scale = max(1.0, abs(E))
    if "energy" in what:
        assert abs(res["energy"] - E) <= tol * scale
    if "forces" in what:
        np.testing.assert_allclose(res["forces"], F.T, atol=tol, rtol=tol)
    if "virial" in what:
        np.testing.assert_allclose(res["virial"], W, atol=tol * scale, rtol=tol)


@pytest.mark.parametrize("prec", ["float64", "mixed"])
def test_smoothpair_efv_vs_oracle(pfw, ws, cases, prec):
    for name, X, H, pbc, sp in cases:
        rc = 6.0 if name == "b2-small" else RC
        p = pfw.SmoothPair(["H", "He"], A, rc)
        sys = pfw.System(X, H, pbc, labels_of(sp))
        res = pfw.evaluate_(ws[prec], sys, p)
        E, F, W = pair_evaluate(X, sp, H, pbc, KIND_SMOOTH, A, rc)
        compare(res, E, F, W, TOL[prec])


@pytest.mark.parametrize("prec", ["float64", "mixed"])
def test_request_masks_parameters_permutation(pfw, ws, prec):
    X, H, pbc, sp = pfw.workloads.fixture(23)
    sys = pfw.System(X, H, pbc, labels_of(sp))
    p = pfw.SmoothPair(["H", "He"], A, RC)
    E, F, W = pair_evaluate(X, sp, H, pbc, KIND_SMOOTH, A, RC)
    names = ("energy", "forces", "virial")
    for mask in range(8):
        props = tuple(n for b, n in enumerate(names) if (mask >> b) & 1)
        r = pfw.calculate(props, sys, p, neighbourlist=ws[prec])
        assert set(r) == set(props) | {"state"}
        if props:
            res = {k: (v.T if k == "forces" else v) for k, v in r.items() if k != "state"}
            compare(res, E, F, W, TOL[prec], what=props)
    y = pfw.evaluate_(ws[prec], sys, p)
    saved = y["forces"].copy()
    ps = p.get_parameters()
    ps["amplitudes"] = ps["amplitudes"] * 2
    y2 = pfw.evaluate_(ws[prec], sys, p, parameters=ps)
    assert y2["energy"] == pytest.approx(2 * y["energy"], rel=TOL[prec])
    assert np.array_equal(y["forces"], saved)
    perm = np.arange(len(X))[::-1]
    yp = pfw.evaluate_(ws[prec], pfw.System(X[perm], H, pbc, labels_of(sp[perm])), p)
    np.testing.assert_allclose(yp["forces"], y["forces"][:, perm], rtol=TOL[prec], atol=TOL[prec])


def test_coincident_smooth_is_regular_and_lj_singular(pfw, ws):
    labels = ["C", "C13", "CA"]
    p = pfw.SmoothPair(labels, [[1., .4, .7], [.4, 2., .2], [.7, .2, 3.]], 2.0)
    X = np.zeros((3, 3))
    sys = pfw.System(X, np.zeros((3, 3)), (0, 0, 0), labels)
    for prec in ws:
        r = pfw.evaluate_(ws[prec], sys, p)
        assert r["energy"] == pytest.approx(0.4 + 0.7 + 0.2, rel=1e-6)
        assert np.all(r["forces"] == 0)
    lj = pfw.lennard_jones(sigma=0.8, radius=2.0)
    two = pfw.System(np.array([[0., 0, 0], [1.2, .2, 0]]), np.zeros((3, 3)), (0, 0, 0), ["Ar", "Ar"])
    E, F, W = pair_evaluate(two.positions, [0, 0], two.cell, two.pbc, KIND_LJ, (0.0103, 0.8), 2.0)
    for prec in ws:
        compare(pfw.evaluate_(ws[prec], two, lj), E, F, W, TOL[prec])
    same = pfw.System(np.zeros((2, 3)), np.zeros((3, 3)), (0, 0, 0), ["Ar", "Ar"])
    for prec in ws:
        for req in (pfw.EnergyForcesVirial, ("forces",), ("virial",)):
            with pytest.raises(pfw.DomainError):
                pfw.evaluate_(ws[prec], same, lj, request=req)


def test_error_contract(pfw, ws):
    X, H, pbc, sp = pfw.workloads.fixture(5)
    p = pfw.SmoothPair(["H", "He"], A, RC)
    w = ws["float64"]
    with pytest.raises(pfw.ArgumentError):
        pfw.evaluate_(w, pfw.System(X, H, pbc, ["H", "Ne", "H", "He", "H"]), p)
    Hd = H.copy()
    Hd[:, 1] = Hd[:, 0]
    with pytest.raises(pfw.ArgumentError):
        pfw.evaluate_(w, pfw.System(X, Hd, pbc, labels_of(sp)), p)
    Xn = X.copy()
    Xn[0, 0] = np.nan
    with pytest.raises(pfw.ArgumentError):
        pfw.evaluate_(w, pfw.System(Xn, H, pbc, labels_of(sp)), p)
    with pytest.raises(pfw.ArgumentError):
        pfw.SmoothPair(["H", "He"], [[1., 2.], [3., 1.]], 2.0)
    # the workspace still works after errors
    r = pfw.evaluate_(w, pfw.System(X, H, pbc, labels_of(sp)), p)
    assert np.isfinite(r["energy"])


def force_errors(F_gpu, F_ref):
    d = F_gpu - F_ref
    return float(np.abs(d).max()), float(np.sqrt(np.mean(d * d)))


def test_argon_lj_4000_full_size(pfw, ws):
    X, H, pbc, sp = pfw.workloads.argon_fcc(10)
    lj = pfw.lennard_jones(epsilon=0.0104, sigma=3.40, radius=8.5)
    E, F, W = pair_evaluate(X, sp, H, pbc, KIND_LJ, (0.0104, 3.40), 8.5)
    sys = pfw.System(X, H, pbc, ["Ar"] * len(X))
    r64 = pfw.evaluate_(ws["float64"], sys, lj)
    compare(r64, E, F, W, 1e-10)
    rmx = pfw.evaluate_(ws["mixed"], sys, lj)
    fmax, frms = force_errors(rmx["forces"].T, F)
    print(f"LJ 4000 mixed: max|dF|={fmax:.3e} rms={frms:.3e} dE/atom={(rmx['energy'] - E) / len(X):.3e}")
    assert frms <= 1e-5
    assert abs(rmx["energy"] - E) / len(X) <= 1e-6
    gpu = ws["mixed"].export_pairs()
    assert_same_pairs(gpu, nl_pairs(X, H, pbc, 8.5)[:4])


def test_b2_smoothpair_256k_full_size(pfw, ws):
    X, H, pbc, sp = pfw.workloads.b2_triclinic()
    p = pfw.SmoothPair(["A", "B"], A, 6.0)
    sys = pfw.System(X, H, pbc, ["A" if s == 0 else "B" for s in sp])
    E, F, W = pair_evaluate(X, sp, H, pbc, KIND_SMOOTH, A, 6.0, threads=8)
    r64 = pfw.evaluate_(ws["float64"], sys, p)
    compare(r64, E, F, W, 1e-9)
    rmx = pfw.evaluate_(ws["mixed"], sys, p)
    fmax, frms = force_errors(rmx["forces"].T, F)
    print(f"B2 256k mixed: max|dF|={fmax:.3e} rms={frms:.3e} dE/atom={(rmx['energy'] - E) / len(X):.3e}")
    assert frms <= 1e-5
    assert abs(rmx["energy"] - E) / len(X) <= 1e-6
    np.testing.assert_allclose(rmx["virial"], W, rtol=1e-5, atol=1e-3)
    gpu = ws["mixed"].export_pairs()
    assert_same_pairs(gpu, nl_pairs(X, H, pbc, 6.0)[:4])
