"""ctypes wrapper of the oracle's C++ FP64 restatement (test infrastructure only)."""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")

KIND_SMOOTH = 0
KIND_LJ = 1


class OracleError(Exception):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def _load():
    if not os.path.exists(_SO):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    so = ctypes.CDLL(_SO)
    d = ctypes.POINTER(ctypes.c_double)
    i32 = ctypes.POINTER(ctypes.c_int32)
    i64 = ctypes.POINTER(ctypes.c_int64)
    u8 = ctypes.POINTER(ctypes.c_uint8)
    so.oracle_pair_evaluate.argtypes = [ctypes.c_int64, d, i32, d, u8, ctypes.c_int, ctypes.c_int, d,
                                        ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_uint,
                                        d, d, d, ctypes.c_char_p, ctypes.c_int]
    so.oracle_pair_evaluate.restype = ctypes.c_int
    so.oracle_nl_build.argtypes = [ctypes.c_int64, d, d, u8, ctypes.c_double, ctypes.c_int,
                                   ctypes.c_char_p, ctypes.c_int]
    so.oracle_nl_build.restype = ctypes.c_void_p
    so.oracle_nl_count.argtypes = [ctypes.c_void_p]
    so.oracle_nl_count.restype = ctypes.c_int64
    so.oracle_nl_copy.argtypes = [ctypes.c_void_p, i64, i64, i64, d]
    so.oracle_nl_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                  i64, i64, d, d]
    so.oracle_nl_free.argtypes = [ctypes.c_void_p]
    return so


lib = _load()


def _ptr(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def _geom(X, H, pbc):
    X = np.ascontiguousarray(np.asarray(X, dtype=np.float64).reshape(-1, 3))
    Hc = np.ascontiguousarray(np.asarray(H, dtype=np.float64).T)  # column-major H
    pb = np.ascontiguousarray(np.asarray(pbc, dtype=np.uint8).reshape(3))
    return X, Hc, pb


def nl_pairs(X, H, pbc, cutoff, bruteforce_threshold=32):
    """Directed neighbour pairs (i, j, S, R) in original order, as
    ``neighbour_list`` (src/submodules/NeighbourList/interface.jl:47-83, through
    ``visit_directed_images!`` :116-143) emits them.
    ``X`` is (n, 3) in Å; ``H`` is 3x3 with lattice vectors in COLUMNS."""
    X, Hc, pb = _geom(X, H, pbc)
    n = X.shape[0]
    err = ctypes.create_string_buffer(512)
    h = lib.oracle_nl_build(n, _ptr(X, ctypes.c_double), _ptr(Hc, ctypes.c_double),
                            _ptr(pb, ctypes.c_uint8), float(cutoff), int(bruteforce_threshold), err, 512)
    if not h:
        raise OracleError(1, err.value.decode())
    try:
        m = lib.oracle_nl_count(h)
        i = np.zeros(m, np.int64)
        j = np.zeros(m, np.int64)
        S = np.zeros((m, 3), np.int64)
        R = np.zeros((m, 3), np.float64)
        if m:
            lib.oracle_nl_copy(h, _ptr(i, ctypes.c_int64), _ptr(j, ctypes.c_int64),
                               _ptr(S, ctypes.c_int64), _ptr(R, ctypes.c_double))
        mi = ctypes.c_int()
        br = ctypes.c_int()
        dims = np.zeros(3, np.int64)
        nc = np.zeros(1, np.int64)
        basis = np.zeros(9)
        recip = np.zeros(9)
        lib.oracle_nl_info(h, ctypes.byref(mi), ctypes.byref(br), _ptr(dims, ctypes.c_int64),
                           _ptr(nc, ctypes.c_int64), _ptr(basis, ctypes.c_double), _ptr(recip, ctypes.c_double))
        info = dict(minimum_image=bool(mi.value), bruteforce=bool(br.value), dims=dims, ncells=int(nc[0]),
                    basis=basis.reshape(3, 3).T.copy(), reciprocal=recip.reshape(3, 3).T.copy())
        return i, j, S, R, info
    finally:
        lib.oracle_nl_free(h)


def pair_evaluate(X, species, H, pbc, kind, params, cutoff, bruteforce_threshold=32, threads=1, mask=7):
    """FP64 E, F (n, 3) original order, W (3x3) of ``evaluate!`` with the given pair
    model (src/submodules/PairPotentials/evaluation.jl:31-52). ``params`` is the nz x nz
    amplitude table for SmoothPair or (epsilon, sigma) for Lennard-Jones."""
    X, Hc, pb = _geom(X, H, pbc)
    n = X.shape[0]
    sp = np.ascontiguousarray(np.asarray(species, dtype=np.int32).reshape(n))
    if kind == KIND_SMOOTH:
        A = np.asarray(params, dtype=np.float64)
        nz = A.shape[0]
        P = np.ascontiguousarray(A.T)
    else:
        nz = max(1, int(sp.max()) + 1) if n else 1
        P = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(2))
    E = np.zeros(1)
    F = np.zeros((n, 3))
    W = np.zeros(9)
    err = ctypes.create_string_buffer(512)
    rc = lib.oracle_pair_evaluate(n, _ptr(X, ctypes.c_double), _ptr(sp, ctypes.c_int32),
                                  _ptr(Hc, ctypes.c_double), _ptr(pb, ctypes.c_uint8), int(kind), int(nz),
                                  _ptr(P, ctypes.c_double), float(cutoff), int(bruteforce_threshold),
                                  int(threads), int(mask), _ptr(E, ctypes.c_double), _ptr(F, ctypes.c_double),
                                  _ptr(W, ctypes.c_double), err, 512)
    if rc != 0:
        raise OracleError(rc, err.value.decode())
    return float(E[0]), F, W.reshape(3, 3).T.copy()
