"""The allocation-free pair step (pfw_pair_evaluate_noalloc) against the
ordinary entry point (pfw_set_system + pfw_set_pair_model +
pfw_pair_evaluate) on the same inputs: the same kernels over the same cell
grid, so energies, forces and virials agree to 1e-12 (they are bitwise equal
in practice). Two calls in a row, the second with a smaller input, share the
static scratch area; the size limits raise PFW_CAPACITY."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A = np.array([[1.0, 0.7], [0.7, 1.3]])


class DeviceBuffer:
    """Device memory from the HIP runtime (test-side only)."""

    _hip = None

    def __init__(self, nbytes):
        if DeviceBuffer._hip is None:
            DeviceBuffer._hip = ctypes.CDLL("libamdhip64.so")
        self.p = ctypes.c_void_p()
        self.nbytes = max(int(nbytes), 8)
        assert self._hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(self.nbytes)) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self._hip.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0
        return self

    def download(self, shape, dtype=np.float64):
        out = np.zeros(shape, dtype)
        assert self._hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def __del__(self):
        if self.p:
            self._hip.hipFree(self.p)


def noalloc(pfw, ctx, kind, params, rc, X, sp, H, pbc, mask=7, precision=1):
    from pfw_amd import _capi
    n = len(X)
    Xd = DeviceBuffer(24 * n).upload(np.asarray(X, np.float64))
    sd = DeviceBuffer(4 * n).upload(np.asarray(sp, np.int32))
    Ed, Fd, Wd = DeviceBuffer(8), DeviceBuffer(24 * n), DeviceBuffer(72)
    params = np.ascontiguousarray(np.asarray(params, np.float64).T)
    nz = params.shape[0] if kind == _capi.PFW_PAIR_SMOOTH else 0
    Hc = _capi.as_cell(H)
    pb = np.array([1 if p else 0 for p in pbc], np.uint8)
    rc_ = _capi.load().pfw_pair_evaluate_noalloc(
        ctx.h, kind, nz, _capi.ptr(params), ctypes.c_double(rc), ctypes.c_int64(n), Xd.p, sd.p, _capi.ptr(Hc),
        _capi.ptr(pb, ctypes.c_uint8), mask, precision, Ed.p, Fd.p, Wd.p)
    _capi.check(ctx.h, rc_)
    return float(Ed.download(1)[0]), Fd.download((n, 3)).T, Wd.download(9).reshape(3, 3).T


def reference(pfw, kind, rc, X, sp, H, pbc, prec):
    if kind == 0:
        p = pfw.SmoothPair(["H", "He"], A, rc)
        labels = ["H" if s == 0 else "He" for s in sp]
    else:
        p = pfw.lennard_jones(epsilon=0.0104, sigma=3.40, radius=rc)
        labels = ["Ar"] * len(X)
    return pfw.evaluate_(pfw.Workspace(precision=prec), pfw.System(X, H, pbc, labels), p)


@pytest.mark.parametrize("prec", ["float64", "mixed"])
def test_noalloc_matches_pair_evaluate(pfw, prec):
    from pfw_amd import _capi
    mode = _capi.PFW_MIXED if prec == "mixed" else _capi.PFW_FP64
    ctx = _capi.Context(0)
    wl = pfw.workloads
    cases = [
        ("B2 256k SmoothPair", 0, A, 6.0, wl.b2_triclinic()),
        ("B2 576 SmoothPair (smaller, same scratch)", 0, A, 6.0, wl.b2_triclinic(4, 4, 6, seed=3)),
        ("argon 4000 LJ", 1, [0.0104, 3.40], 8.5, wl.argon_fcc()),
        ("argon 108 LJ (smaller)", 1, [0.0104, 3.40], 8.5, wl.argon_fcc(m=3)),
        ("open cluster", 0, A, 2.0, wl.fixture(65, pbc=(0, 0, 0), cluster=True)),
        ("slab", 0, A, 2.0, wl.fixture(40, pbc=(1, 0, 1))),
    ]
    for name, kind, params, rc, (X, H, pbc, sp) in cases:
        print(f"noalloc {prec}: {name}, n={len(X)}", flush=True)
        E, F, W = noalloc(pfw, ctx, kind, params, rc, X, sp, H, pbc, precision=mode)
        r = reference(pfw, kind, rc, X, sp, H, pbc, prec)
        scale = max(1.0, abs(r["energy"]))
        assert abs(E - r["energy"]) <= 1e-12 * scale, name
        np.testing.assert_allclose(F, r["forces"], atol=1e-12, rtol=1e-12, err_msg=name)
        np.testing.assert_allclose(W, r["virial"], atol=1e-12 * scale, rtol=1e-12, err_msg=name)
        print(f"  bitwise equal: E {E == r['energy']}, F {np.array_equal(F, r['forces'])}, "
              f"W {np.array_equal(W, r['virial'])}", flush=True)


def test_noalloc_limits_and_errors(pfw):
    from pfw_amd import _capi
    ctx = _capi.Context(0)
    X, H, pbc, sp = pfw.workloads.fixture(9)
    # energy only: no force output needed
    E, _, _ = noalloc(pfw, ctx, 0, A, 2.0, X, sp, H, pbc, mask=1)
    r = reference(pfw, 0, 2.0, X, sp, H, pbc, "mixed")
    assert abs(E - r["energy"]) <= 1e-12 * max(1, abs(r["energy"]))
    lib = _capi.load()
    params = np.ascontiguousarray(A)
    Hc = _capi.as_cell(H)
    pb = np.ones(3, np.uint8)
    big = 262145
    rc_ = lib.pfw_pair_evaluate_noalloc(ctx.h, 0, 2, _capi.ptr(params), ctypes.c_double(2.0), ctypes.c_int64(big),
                                        None, None, _capi.ptr(Hc), _capi.ptr(pb, ctypes.c_uint8), 7, 1, None, None,
                                        None)
    assert rc_ == _capi.PFW_CAPACITY
    assert "262,144" in lib.pfw_last_error(ctx.h).decode()
    # species outside the amplitude table
    bad = np.array(sp, np.int32)
    bad[0] = 5
    with pytest.raises(_capi.ArgumentError):
        noalloc(pfw, ctx, 0, A, 2.0, X, bad, H, pbc)
    Xn = np.array(X)
    Xn[1, 2] = np.nan
    with pytest.raises(_capi.ArgumentError):
        noalloc(pfw, ctx, 0, A, 2.0, Xn, sp, H, pbc)
    # the area is reused afterwards
    E2, F2, _ = noalloc(pfw, ctx, 0, A, 2.0, X, sp, H, pbc)
    assert abs(E2 - r["energy"]) <= 1e-12 * max(1, abs(r["energy"]))
