"""GPU parity of Verlet-list reuse (Workspace(skin=...)) against the oracle.

Every step of a trajectory is compared with the oracle's fresh evaluation at
the same positions (FP64 to 1e-10, mixed to 3e-5, the reference's tolerances,
test/gpu_pairs.jl:19-37), while the rebuild counter shows when the list was
kept and when it was rebuilt. The motions follow the reference's skin test
(test/gpu_pairs.jl:88-113): an approaching pair that enters the cutoff
without a rebuild, displacements below / at / above half the skin, a wrapped
jump by a lattice vector, cell / species / order / cutoff changes; plus
triclinic and multi-image cells (ghost images), open axes, a batch, LJ, and a
cell build by another entry point between two steps.
"""
import numpy as np
import pytest

from oracle import KIND_LJ, KIND_SMOOTH, nl_pairs, pair_evaluate

pytestmark = pytest.mark.gpu

A = np.array([[1.0, 0.7], [0.7, 1.3]])
TOL = {"float64": 1e-10, "mixed": 3e-5}
PRECS = ("float64", "mixed")


def labels_of(sp):
    return ["H" if s == 0 else "He" for s in sp]


def check(pfw, w, prec, X, H, pbc, sp, rc, kind=KIND_SMOOTH, params=A):
    if kind == KIND_SMOOTH:
        p = pfw.SmoothPair(["H", "He"], params, rc)
        labels = labels_of(sp)
    else:
        p = pfw.lennard_jones(epsilon=params[0], sigma=params[1], radius=rc)
        labels = ["Ar"] * len(X)
    r = pfw.evaluate_(w, pfw.System(X, H, pbc, labels), p)
    E, F, W = pair_evaluate(X, sp, H, pbc, kind, params, rc)
    tol = TOL[prec]
    scale = max(1.0, abs(E))
    assert abs(r["energy"] - E) <= tol * scale
    np.testing.assert_allclose(r["forces"], F.T, atol=tol, rtol=tol)
    np.testing.assert_allclose(r["virial"], W, atol=tol * scale, rtol=tol)
    return r


@pytest.mark.parametrize("prec", PRECS)
def test_skin_trajectory_and_invalidation(pfw, prec):
    w = pfw.Workspace(precision=prec, skin=0.4)
    # two atoms 2.2 Å apart approach inside the 2.0 Å cutoff with no rebuild
    X = np.array([[0.0, 0, 0], [2.2, 0, 0]])
    H = np.zeros((3, 3))
    sp = np.array([0, 1], np.int32)
    check(pfw, w, prec, X, H, (0, 0, 0), sp, 2.0)
    initial = w.rebuilds
    assert initial == 1
    X[0, 0] += 0.15
    X[1, 0] -= 0.15
    r = check(pfw, w, prec, X, H, (0, 0, 0), sp, 2.0)
    assert r["energy"] > 0
    assert w.rebuilds == initial
    X[0, 0] += 0.05  # 2 x 0.2 reaches the 0.4 skin: rebuild
    check(pfw, w, prec, X, H, (0, 0, 0), sp, 2.0)
    assert w.rebuilds > initial
    # a periodic 65-atom fixture: small, threshold and large moves
    X, H, pbc, sp = pfw.workloads.fixture(65)
    check(pfw, w, prec, X, H, pbc, sp, 2.0)
    base = w.rebuilds
    kept = 0
    for dx in (0.01, 0.08, 0.199, 0.2, 0.21, 0.3, -0.01):
        before = w.rebuilds
        X[0, 0] += dx
        check(pfw, w, prec, X, H, pbc, sp, 2.0)
        kept += w.rebuilds == before
    assert w.rebuilds > base and kept > 0
    base = w.rebuilds
    X[0, 0] += H[0, 0]  # a wrapped jump rebuilds; the result is unchanged
    check(pfw, w, prec, X, H, pbc, sp, 2.0)
    assert w.rebuilds > base
    H2 = H.copy()
    H2[0, 0] += 0.1
    check(pfw, w, prec, X, H2, pbc, sp, 2.0)
    check(pfw, w, prec, X, H2, pbc, sp[::-1].copy(), 2.0)
    check(pfw, w, prec, X[::-1].copy(), H2, pbc, sp[::-1].copy(), 2.0)
    check(pfw, w, prec, X, H2, pbc, sp, 1.8)


@pytest.mark.parametrize("prec", PRECS)
def test_skin_random_walk_triclinic(pfw, prec):
    """Random walk in a sheared B2 cell (ghost images on every face): each step
    against the oracle; most steps keep the list."""
    X, H, pbc, sp = pfw.workloads.b2_triclinic(4, 4, 6, seed=7)
    w = pfw.Workspace(precision=prec, skin=0.5)
    rng = np.random.default_rng(3)
    for step in range(16):
        check(pfw, w, prec, X, H, pbc, sp, 6.0)
        if step == 0:  # the list holds every directed pair within cutoff + skin
            assert w.verlet_info()["entries"] == len(nl_pairs(X, H, pbc, 6.5)[0])
        X = X + rng.normal(0, 0.03, X.shape)
    info = w.verlet_info()
    assert info["rebuilds"] >= 2 and info["reuses"] >= 8, info
    assert info["ghosts"] > 0


@pytest.mark.parametrize("prec", PRECS)
def test_skin_multi_image_cells(pfw, prec):
    """The reference's general cells (test/periodic_images.jl:63-69), several
    of them shorter than the cutoff: many images of every atom are ghosts."""
    rng = np.random.default_rng(5)
    for k, H in enumerate(pfw.workloads.IMAGE_TEST_CELLS):
        X, sp = pfw.workloads.image_test_system(H)
        for pbc in ((1, 1, 1), (1, 0, 1)):
            w = pfw.Workspace(precision=prec, skin=0.3)
            Y = X.copy()
            for step in range(4):
                check(pfw, w, prec, Y, H, pbc, sp, 2.0)
                if step == 0:  # every image pair within cutoff + skin, ghosts included
                    assert w.verlet_info()["entries"] == len(nl_pairs(Y, H, pbc, 2.3)[0])
                Y = Y + rng.normal(0, 0.02, Y.shape)
            assert w.rebuilds >= 1


@pytest.mark.parametrize("prec", PRECS)
def test_skin_lennard_jones(pfw, prec):
    X, H, pbc, sp = pfw.workloads.argon_fcc(4)
    w = pfw.Workspace(precision=prec, skin=0.6)
    rng = np.random.default_rng(9)
    for step in range(6):
        check(pfw, w, prec, X, H, pbc, sp, 8.5, kind=KIND_LJ, params=(0.0104, 3.40))
        X = X + rng.normal(0, 0.02, X.shape)
    assert w.verlet_info()["reuses"] >= 3


def test_skin_survives_other_cell_builds(pfw):
    """A cell build by another entry point (pair export) reorders the atoms: the
    next step must rebuild the list, and results stay right."""
    X, H, pbc, sp = pfw.workloads.fixture(65)
    w = pfw.Workspace(precision="float64", skin=0.3)
    check(pfw, w, "float64", X, H, pbc, sp, 2.0)
    base = w.rebuilds
    w.export_pairs()
    check(pfw, w, "float64", X, H, pbc, sp, 2.0)
    assert w.rebuilds == base + 1
    check(pfw, w, "float64", X + 0.01, H, pbc, sp, 2.0)
    assert w.rebuilds == base + 1


def test_skin_zero_and_changes(pfw):
    X, H, pbc, sp = pfw.workloads.fixture(65)
    w = pfw.Workspace(precision="float64", skin=0.3)
    check(pfw, w, "float64", X, H, pbc, sp, 2.0)
    w.ctx.call("pfw_set_skin", __import__("ctypes").c_double(0.0))
    check(pfw, w, "float64", X + 0.05, H, pbc, sp, 2.0)
    w.ctx.call("pfw_set_skin", __import__("ctypes").c_double(0.7))
    check(pfw, w, "float64", X + 0.07, H, pbc, sp, 2.0)
    with pytest.raises(pfw.ArgumentError):
        pfw.Workspace(skin=-0.1)
    with pytest.raises(pfw.ArgumentError):
        pfw.Workspace(skin=float("nan"))


@pytest.mark.parametrize("prec", PRECS)
def test_skin_batch(pfw, prec):
    """A batch evaluated through the list (per-atom records reduced per configuration)."""
    cfgs, _ = pfw.workloads.fitting_set(ncfg=20, nmin=8, nmax=300, seed=13)
    p = pfw.SmoothPair(["A", "B"], A, 6.0)
    w = pfw.Workspace(precision=prec, skin=0.5)
    systems = [pfw.System(X, H, pbc, ["A" if s == 0 else "B" for s in sp]) for X, H, pbc, sp in cfgs]
    res = pfw.evaluate_batch(w, systems, p)
    assert w.rebuilds == 1
    tol = TOL[prec]
    for (X, H, pbc, sp), r in zip(cfgs, res):
        E, F, W = pair_evaluate(X, sp, H, pbc, KIND_SMOOTH, A, 6.0)
        assert abs(r["energy"] - E) <= tol * max(1, abs(E))
        np.testing.assert_allclose(r["forces"], F.T, atol=tol, rtol=tol)
        np.testing.assert_allclose(r["virial"], W, atol=tol * max(1, abs(E)), rtol=tol)


def test_skin_b2_256k_full_size(pfw):
    """The 256k-atom sheared B2 workload: a fresh step and a reused step after a
    0.05 Å random displacement, FP64 against the threaded oracle, mixed by the
    force RMSE bound."""
    X, H, pbc, sp = pfw.workloads.b2_triclinic()
    p = pfw.SmoothPair(["A", "B"], A, 6.0)
    labels = ["A" if s == 0 else "B" for s in sp]
    w64 = pfw.Workspace(precision="float64", skin=0.5)
    wmx = pfw.Workspace(precision="mixed", skin=0.5)
    rng = np.random.default_rng(21)
    for step in range(2):
        sys = pfw.System(X, H, pbc, labels)
        r64 = pfw.evaluate_(w64, sys, p)
        rmx = pfw.evaluate_(wmx, sys, p)
        E, F, W = pair_evaluate(X, sp, H, pbc, KIND_SMOOTH, A, 6.0, threads=8)
        assert abs(r64["energy"] - E) <= 1e-9 * abs(E)
        np.testing.assert_allclose(r64["forces"], F.T, atol=1e-9, rtol=1e-9)
        np.testing.assert_allclose(r64["virial"], W, atol=1e-9 * abs(E), rtol=1e-9)
        d = rmx["forces"].T - F
        assert np.sqrt(np.mean(d * d)) <= 1e-5
        assert abs(rmx["energy"] - E) / len(X) <= 1e-6
        X = X + rng.uniform(-0.05, 0.05, X.shape)
    assert w64.verlet_info()["reuses"] == 1 and wmx.verlet_info()["reuses"] == 1


def test_device_positions_update(pfw):
    """pfw_update_positions_device: positions handed over in device memory (a
    hipMalloc buffer standing in for an integrator's array), checked for
    finiteness; a rejected update keeps the previous positions."""
    import ctypes

    from pfw_amd import _capi

    _capi.load()
    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime the library already uses
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    H2D = 1

    X, H, pbc, sp = pfw.workloads.fixture(65)
    Y = np.ascontiguousarray(X + np.sin(np.arange(X.size)).reshape(X.shape) * 0.03)
    bad = Y.copy()
    bad[3, 1] = np.nan
    dev = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dev), Y.nbytes) == 0
    try:
        for skin in (0.0, 0.3):
            w = pfw.Workspace(precision="float64", skin=skin)
            check(pfw, w, "float64", X, H, pbc, sp, 2.0)
            E, F, W = np.zeros(1), np.zeros((len(X), 3)), np.zeros(9)

            def step():
                w.ctx.call("pfw_pair_evaluate", 7, _capi.PFW_FP64, _capi.ptr(E), _capi.ptr(F), _capi.ptr(W))
                return E[0], F.copy(), W.reshape(3, 3).T.copy()

            assert hip.hipMemcpy(dev, Y.ctypes.data, Y.nbytes, H2D) == 0
            w.ctx.call("pfw_update_positions_device", dev)
            e, f, v = step()
            Er, Fr, Wr = pair_evaluate(Y, sp, H, pbc, KIND_SMOOTH, A, 2.0)
            assert abs(e - Er) <= 1e-10 * max(1, abs(Er))
            np.testing.assert_allclose(f, Fr, atol=1e-10, rtol=1e-10)
            np.testing.assert_allclose(v, Wr, atol=1e-10 * max(1, abs(Er)), rtol=1e-10)
            assert hip.hipMemcpy(dev, bad.ctypes.data, bad.nbytes, H2D) == 0
            with pytest.raises(pfw.ArgumentError):
                w.ctx.call("pfw_update_positions_device", dev)
            e2, f2, _ = step()
            assert e2 == e and np.array_equal(f2, f)
    finally:
        hip.hipFree(dev)
