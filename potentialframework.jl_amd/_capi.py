"""ctypes binding of libpfw_mi355x.so (the C ABI in include/pfw_mi355x.h).

The library is loaded from this package directory. There is no CPU fallback:
if the HIP library is missing or no GPU is visible, every entry point raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpfw_mi355x.so")

PFW_OK, PFW_INVALID_ARGUMENT, PFW_DIMENSION, PFW_DOMAIN, PFW_STALE, PFW_CAPACITY, PFW_DEVICE = range(7)
PFW_ENERGY, PFW_FORCES, PFW_VIRIAL = 1, 2, 4
PFW_FP64, PFW_MIXED = 0, 1
PFW_PAIR_SMOOTH, PFW_PAIR_LJ = 0, 1


class ArgumentError(ValueError):
    """Julia ArgumentError: invalid input."""


class DimensionMismatch(ValueError):
    """Julia DimensionMismatch: inconsistent shapes."""


class DomainError(ArithmeticError):
    """Julia DomainError: singular or nonfinite model output."""


class StaleError(RuntimeError):
    """Julia ErrorException for stale topology / models."""


class CapacityError(RuntimeError):
    pass


class DeviceError(RuntimeError):
    pass


_ERRORS = {PFW_INVALID_ARGUMENT: ArgumentError, PFW_DIMENSION: DimensionMismatch, PFW_DOMAIN: DomainError,
           PFW_STALE: StaleError, PFW_CAPACITY: CapacityError, PFW_DEVICE: DeviceError}

_lib = None

_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.POINTER(ctypes.c_int32)
_I64 = ctypes.POINTER(ctypes.c_int64)
_U8 = ctypes.POINTER(ctypes.c_uint8)

# (name, restype, argtypes): every symbol declared in include/pfw_mi355x.h
SIGNATURES = [
    ("pfw_ctx_create", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_P)]),
    ("pfw_ctx_destroy", None, [_P]),
    ("pfw_last_error", ctypes.c_char_p, [_P]),
    ("pfw_set_system", ctypes.c_int, [_P, ctypes.c_int64, _D, _I32, _D, _U8, ctypes.c_double]),
    ("pfw_set_batch", ctypes.c_int, [_P, ctypes.c_int32, _I64, _D, _I32, _D, _U8, ctypes.c_double]),
    ("pfw_read_batch_results", ctypes.c_int, [_P, _D, _D, _D]),
    ("pfw_update_positions", ctypes.c_int, [_P, _D]),
    ("pfw_update_positions_device", ctypes.c_int, [_P, ctypes.c_void_p]),
    ("pfw_set_pair_model", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _D, ctypes.c_double]),
    ("pfw_pair_step", ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_int]),
    ("pfw_read_results", ctypes.c_int, [_P, _D, _D, _D]),
    ("pfw_pair_evaluate", ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_int, _D, _D, _D]),
    ("pfw_export_pairs", ctypes.c_int, [_P, _I32, _I32, _I32, _D, ctypes.c_int64, _I64]),
    ("pfw_grid_info", ctypes.c_int, [_P, _I64, _I32, _I32]),
    ("pfw_set_skin", ctypes.c_int, [_P, ctypes.c_double]),
    ("pfw_verlet_info", ctypes.c_int, [_P, _I64, _D]),
    ("pfw_set_timing", ctypes.c_int, [_P, ctypes.c_int]),
    ("pfw_get_timings", ctypes.c_int, [_P, _D, ctypes.c_int]),
    ("pfw_pair_evaluate_noalloc", ctypes.c_int,
     [_P, ctypes.c_int, ctypes.c_int, _D, ctypes.c_double, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, _D, _U8,
      ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("pfw_set_audit", ctypes.c_int, [_P, ctypes.c_int]),
    ("pfw_read_audit", ctypes.c_int, [_P, _I32, ctypes.POINTER(ctypes.c_uint64)]),
    ("pfw_set_graphs", ctypes.c_int, [_P, ctypes.c_int]),
    ("pfw_graph_info", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64)]),
    ("pfw_set_tuning", ctypes.c_int, [_P, _I32, ctypes.c_int]),
    ("pfw_get_tuning", ctypes.c_int, [_P, _I32, ctypes.c_int]),
    ("pfw_synchronize", ctypes.c_int, [_P]),
    ("pfw_host_alloc", ctypes.c_void_p, [ctypes.c_size_t]),
    ("pfw_host_free", None, [ctypes.c_void_p]),
    ("pfw_set_ace_model", ctypes.c_int, [_P, ctypes.c_void_p]),
    ("pfw_set_ace_parameters", ctypes.c_int, [_P, _D, _D]),
    ("pfw_ace_step", ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_int]),
    ("pfw_ace_evaluate", ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_int, _D, _D, _D]),
    ("pfw_ace_site_energies", ctypes.c_int, [_P, _D]),
    ("pfw_ace_adjoint", ctypes.c_int, [_P, ctypes.c_int, _D, _D, _D, _D, _D]),
    ("pfw_pair_adjoint", ctypes.c_int, [_P, _D, _D, _D, _D]),
    ("pfw_ace_parameter_sets", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int64, _D, _D, _D]),
    ("pfw_ace_design", ctypes.c_int, [_P, ctypes.c_int, _D, _D, ctypes.c_int64]),
    ("pfw_features", ctypes.c_int, [_P, _P, _D]),
    ("pfw_features_vjp", ctypes.c_int, [_P, _P, _D, _D, _D, _D, _D]),
]


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "or python potentialframework.jl_amd/build.py")
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


# Page-locked bytes that live result arrays may hold at once; past this,
# results are ordinary (pageable) arrays, so a caller that keeps every step's
# forces cannot pin an unbounded amount of host memory.
PINNED_LIVE_CAP = 256 << 20
_pinned_live = [0]


def _release_pinned(p, nbytes):
    _pinned_live[0] -= nbytes
    load().pfw_host_free(p)


def pinned_empty(count, dtype=np.float64):
    """A flat array of ``count`` elements in a page-locked block of the
    library's host cache (pfw_host_alloc), handed back to the cache when the
    last array viewing it is freed; an ordinary array when no block can be had
    or when live page-locked results already hold PINNED_LIVE_CAP bytes. The
    library's device-to-host result copies into a page-locked block run as one
    DMA. Either way the array is C-contiguous and owns nothing else."""
    import weakref
    dt = np.dtype(dtype)
    nbytes = int(count) * dt.itemsize
    if nbytes == 0 or _pinned_live[0] + nbytes > PINNED_LIVE_CAP:
        return np.empty(int(count), dt)
    lib = load()
    p = lib.pfw_host_alloc(nbytes)
    if not p:
        return np.empty(int(count), dt)
    _pinned_live[0] += nbytes
    buf = (ctypes.c_char * nbytes).from_address(p)
    weakref.finalize(buf, _release_pinned, p, nbytes).atexit = False
    return np.frombuffer(buf, dtype=dt, count=int(count))


def ptr(a, t=ctypes.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(t))


def check(ctx, rc):
    if rc == PFW_OK:
        return
    msg = load().pfw_last_error(ctx)
    msg = msg.decode() if msg else ""
    raise _ERRORS.get(rc, DeviceError)(msg or f"pfw status {rc}")


class Context:
    """Owns one pfw_ctx (device memory, stream)."""

    def __init__(self, device=0):
        lib = load()
        h = _P()
        rc = lib.pfw_ctx_create(int(device), ctypes.byref(h))
        if rc != PFW_OK:
            raise _ERRORS.get(rc, DeviceError)(f"pfw_ctx_create failed with status {rc} (is a GPU visible?)")
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            load().pfw_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def call(self, name, *args):
        check(self.h, getattr(load(), name)(self.h, *args))


def as_positions(X):
    return np.ascontiguousarray(np.asarray(X, dtype=np.float64).reshape(-1, 3))


def as_cell(H):
    H = np.asarray(H, dtype=np.float64)
    if H.shape != (3, 3):
        raise DimensionMismatch("cell must be 3x3")
    return np.ascontiguousarray(H.T)  # column-major
