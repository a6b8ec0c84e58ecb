"""Device workspace and directed neighbour snapshots.

* ``Workspace`` stands where the reference's ``NeighbourList.Workspace`` /
  ``GPUWorkspace`` stand (src/submodules/NeighbourList/storage.jl:24-72,
  gpu.jl:61-116): reusable storage passed as ``neighbourlist=`` to the
  calculators. It owns one device context; do not share it between concurrent
  calls (README.md:217-219).
* ``neighbour_list`` / ``neighbours`` mirror src/submodules/NeighbourList/interface.jl:1-151:
  a materialised directed snapshot in original atom order with integer shifts
  relative to the original (unwrapped) coordinates.
"""
import ctypes

import numpy as np

from . import _capi
from ._capi import ArgumentError, DimensionMismatch, Context, as_cell, as_positions, ptr
from .system import System, validate_geometry

STAGES = ("binning", "scan", "scatter_sort", "environments", "interaction", "reduction")
PRECISIONS = {"float64": _capi.PFW_FP64, "fp64": _capi.PFW_FP64, "mixed": _capi.PFW_MIXED,
              "float32": _capi.PFW_MIXED, np.float64: _capi.PFW_FP64, np.float32: _capi.PFW_MIXED,
              float: _capi.PFW_FP64}


def precision_mode(precision):
    try:
        return PRECISIONS[precision]
    except (KeyError, TypeError):
        raise ArgumentError("precision must be 'float64' (plain FP64) or 'mixed' (FP32 arithmetic, FP64 "
                            "accumulation)") from None


class Workspace:
    """Reusable device storage for one system at a time.

    ``precision``: ``"float64"`` or ``"mixed"``. ``device``: HIP device ordinal.
    ``timings`` holds per-stage device milliseconds of the last call when
    ``timing=True`` (the reference's ``w.timings``). ``skin`` (Å, default 0)
    is the reference's ``GPUConfig(skin=...)`` (gpu.jl:61-72, 504-549): with
    ``skin > 0`` pair evaluations keep a neighbour list at cutoff + skin while
    only positions change and no atom has moved skin/2 since its build;
    ``rebuilds`` counts the builds."""

    def __init__(self, precision="float64", device=0, timing=False, skin=0.0):
        self.precision = precision
        self.mode = precision_mode(precision)
        skin = float(skin)
        if not (np.isfinite(skin) and skin >= 0):
            raise ArgumentError("skin must be finite and nonnegative (Å)")
        self.ctx = Context(device)
        self._key = None
        self._model_key = None
        self.timings = {}
        self.timing = bool(timing)
        self.ctx.call("pfw_set_timing", int(self.timing))
        self.skin = skin
        self.ctx.call("pfw_set_skin", ctypes.c_double(skin))

    def verlet_info(self):
        """``rebuilds``, ``reuses``, ``ghosts``, ``slots`` (per atom), ``entries``
        (list entries at the last build) and the last step's largest
        displacement since the build (Å)."""
        info = np.zeros(5, np.int64)
        md = np.zeros(1)
        self.ctx.call("pfw_verlet_info", ptr(info, ctypes.c_int64), ptr(md))
        return dict(rebuilds=int(info[0]), reuses=int(info[1]), ghosts=int(info[2]), slots=int(info[3]),
                    entries=int(info[4]), max_displacement=float(md[0]))

    @property
    def rebuilds(self):
        return self.verlet_info()["rebuilds"]

    # -- geometry ---------------------------------------------------------------------
    def prepare(self, sys, cutoff, channels):
        """Upload geometry; re-uses the cell grid when only positions changed."""
        validate_geometry(sys)
        X = as_positions(sys.positions)
        H = np.asarray(sys.cell, dtype=np.float64)
        pb = np.array([1 if p else 0 for p in sys.pbc], dtype=np.uint8)
        key = (len(sys), H.tobytes(), pb.tobytes(), np.asarray(channels, np.int32).tobytes(), float(cutoff))
        if key == self._key:
            self.ctx.call("pfw_update_positions", ptr(X))
            return
        sp = np.ascontiguousarray(np.asarray(channels, dtype=np.int32))
        self.ctx.call("pfw_set_system", ctypes.c_int64(len(sys)), ptr(X), ptr(sp, ctypes.c_int32), ptr(as_cell(H)),
                      ptr(pb, ctypes.c_uint8), ctypes.c_double(float(cutoff)))
        self._key = key
        self._model_key = None

    def prepare_batch(self, systems, cutoff, channels_list):
        """Upload a batch of independent configurations packed into one cell list."""
        if not systems:
            raise DimensionMismatch("a batch needs at least one configuration")
        for sys in systems:
            validate_geometry(sys)
        offsets = np.zeros(len(systems) + 1, np.int64)
        offsets[1:] = np.cumsum([len(sys) for sys in systems])
        X = np.ascontiguousarray(np.concatenate([as_positions(sys.positions) for sys in systems]))
        sp = np.ascontiguousarray(np.concatenate([np.asarray(c, np.int32) for c in channels_list]))
        H = np.ascontiguousarray(np.concatenate([as_cell(sys.cell).reshape(9) for sys in systems]))
        pb = np.ascontiguousarray(np.array([[1 if p else 0 for p in sys.pbc] for sys in systems], np.uint8).reshape(-1))
        self.ctx.call("pfw_set_batch", ctypes.c_int32(len(systems)), ptr(offsets, ctypes.c_int64), ptr(X),
                      ptr(sp, ctypes.c_int32), ptr(H), ptr(pb, ctypes.c_uint8), ctypes.c_double(float(cutoff)))
        self._key = None
        self._model_key = None
        return offsets

    def read_batch(self, offsets, names):
        ncfg = len(offsets) - 1
        E = np.zeros(ncfg)
        F = np.zeros((int(offsets[-1]), 3))
        W = np.zeros((ncfg, 9))
        self.ctx.call("pfw_read_batch_results", ptr(E), ptr(F), ptr(W))
        out = []
        for c in range(ncfg):
            r = {}
            if "energy" in names:
                r["energy"] = float(E[c])
            if "forces" in names:
                r["forces"] = F[offsets[c]:offsets[c + 1]].T.copy()
            if "virial" in names:
                r["virial"] = W[c].reshape(3, 3).T.copy()
            out.append(r)
        return out

    def grid_info(self):
        nc = np.zeros(1, np.int64)
        dims = np.zeros(3, np.int32)
        mi = np.zeros(1, np.int32)
        self.ctx.call("pfw_grid_info", ptr(nc, ctypes.c_int64), ptr(dims, ctypes.c_int32), ptr(mi, ctypes.c_int32))
        return dict(ncells=int(nc[0]), dims=tuple(int(d) for d in dims), minimum_image=bool(mi[0]))

    def export_pairs(self):
        """Directed pairs (i, j, S, R) of the current system (0-based original indices)."""
        count = np.zeros(1, np.int64)
        rc = _capi.load().pfw_export_pairs(self.ctx.h, None, None, None, None, ctypes.c_int64(0),
                                           ptr(count, ctypes.c_int64))
        if rc not in (_capi.PFW_OK, _capi.PFW_CAPACITY):
            _capi.check(self.ctx.h, rc)
        m = int(count[0])
        i = np.zeros(m, np.int32)
        j = np.zeros(m, np.int32)
        S = np.zeros((m, 3), np.int32)
        R = np.zeros((m, 3), np.float64)
        self.ctx.call("pfw_export_pairs", ptr(i, ctypes.c_int32), ptr(j, ctypes.c_int32), ptr(S, ctypes.c_int32),
                      ptr(R), ctypes.c_int64(m), ptr(count, ctypes.c_int64))
        return i, j, S, R

    # -- captured steps and launch knobs ----------------------------------------------
    def set_graphs(self, enabled=True):
        """Replay fresh pair/ACE steps from a captured HIP graph while only the
        positions change (``pfw_set_graphs``)."""
        self.ctx.call("pfw_set_graphs", int(bool(enabled)))

    def graph_info(self):
        out = np.zeros(3, np.int64)
        self.ctx.call("pfw_graph_info", ptr(out, ctypes.c_int64))
        return dict(captures=int(out[0]), replays=int(out[1]), held=bool(out[2]))

    TUNING_KEYS = ("ace_sites_per_wave", "ace_radial_chunk", "ace_product_waves", "pair_lanes_per_atom",
                   "ace_density_contraction")

    def set_tuning(self, **knobs):
        """Launch-shape knobs (``pfw_set_tuning``); 0 selects the built-in choice."""
        cur = self.get_tuning()
        for k, v in knobs.items():
            if k not in self.TUNING_KEYS:
                raise ArgumentError(f"unknown tuning knob {k!r}")
            cur[k] = int(v)
        vals = np.array([cur[k] for k in self.TUNING_KEYS], np.int32)
        self.ctx.call("pfw_set_tuning", ptr(vals, ctypes.c_int32), len(vals))

    def get_tuning(self):
        vals = np.zeros(len(self.TUNING_KEYS), np.int32)
        self.ctx.call("pfw_get_tuning", ptr(vals, ctypes.c_int32), len(vals))
        return dict(zip(self.TUNING_KEYS, (int(v) for v in vals)))

    # -- pair-set audit ---------------------------------------------------------------
    def set_audit(self, enabled=True):
        """Record, for every later pair or ACE step, per-atom digests of the directed
        pairs the evaluating kernel accumulated (``pfw_set_audit``)."""
        self.ctx.call("pfw_set_audit", int(bool(enabled)))

    def read_audit(self, n):
        """``(count, hash)`` of the last audited step: per centre atom (original
        order) the number of accumulated directed pairs and the wrapping uint64
        sum of ``pair_digest`` over them."""
        count = np.zeros(n, np.int32)
        digest = np.zeros(n, np.uint64)
        self.ctx.call("pfw_read_audit", ptr(count, ctypes.c_int32), ptr(digest, ctypes.c_uint64))
        return count, digest

    def read_timings(self):
        ms = np.zeros(len(STAGES))
        _capi.load().pfw_get_timings(self.ctx.h, ptr(ms), len(STAGES))
        self.timings = dict(zip(STAGES, ms.tolist()))
        return self.timings


_MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def pair_digest(j, S):
    """The 64-bit mix of pair identities ``(j, S)`` used by the pair-set audit
    (``pfw_set_audit``): splitmix64's finaliser of j | (S + 1024 & 2047) << 32 |
    ... << 43 | ... << 54, after adding the golden-ratio increment."""
    j = np.asarray(j, np.int64).astype(np.uint64) & np.uint64(0xFFFFFFFF)
    S = (np.asarray(S, np.int64).reshape(-1, 3) + 1024) & 2047
    z = j | (S[:, 0].astype(np.uint64) << np.uint64(32)) | (S[:, 1].astype(np.uint64) << np.uint64(43)) \
        | (S[:, 2].astype(np.uint64) << np.uint64(54))
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def pair_set_digest(n, i, j, S):
    """Per-atom ``(count, hash)`` of a directed pair set given as arrays
    ``(i, j, S)`` (original indices), in the audit's convention."""
    i = np.asarray(i, np.int64)
    count = np.bincount(i, minlength=n).astype(np.int32)
    digest = np.zeros(n, np.uint64)
    if len(i):
        np.add.at(digest, i, pair_digest(j, S))
    return count, digest


class NeighbourSnapshot:
    """Materialised directed neighbour snapshot (interface.jl:1-8)."""

    def __init__(self, n, i, j, S, R):
        order = np.lexsort((S[:, 2], S[:, 1], S[:, 0], j, i)) if len(i) else np.zeros(0, int)
        i, j, S, R = i[order], j[order], S[order], R[order]
        bounds = np.searchsorted(i, np.arange(n + 1))
        self.indices = [j[bounds[k]:bounds[k + 1]].copy() for k in range(n)]
        self.displacements = [R[bounds[k]:bounds[k + 1]].copy() for k in range(n)]
        self.shifts = [S[bounds[k]:bounds[k + 1]].copy() for k in range(n)]


def neighbours(snapshot, i):
    """``(js, Rs, Ss)`` for atom ``i`` (0-based), as independent arrays (interface.jl:145-151)."""
    if not 0 <= i < len(snapshot.indices):
        raise IndexError(i)
    return snapshot.indices[i].copy(), snapshot.displacements[i].copy(), snapshot.shifts[i].copy()


def neighbour_list(X, cutoff, cell=None, pbc=None, *, workspace=None, device=0):
    """Directed neighbour snapshot on the GPU.

    ``neighbour_list(X, cutoff, cell_rows, pbc)`` takes a 3x3 cell with lattice
    vectors in ROWS, like the reference's raw-array API (interface.jl:28-65);
    ``neighbour_list(system, cutoff)`` takes a ``System``. For raw input
    ``Rs[k] = X[js[k]] - X[i] + cell.T @ Ss[k]``. All images strictly inside the
    cutoff are included, including both +/- self images; only the zero-shift
    self pair is excluded."""
    if isinstance(X, System):
        sys = X
    else:
        cell = np.asarray(cell, dtype=np.float64)
        if cell.shape != (3, 3):
            raise DimensionMismatch("cell must be 3x3 with vectors in rows")
        if pbc is None or len(pbc) != 3:
            raise DimensionMismatch("pbc must contain three flags")
        X = np.asarray(X, dtype=np.float64)
        if X.ndim != 2 or X.shape[1] != 3:
            raise DimensionMismatch("positions must be an (n, 3) array")
        sys = System(X, cell.T, pbc, ["H"] * len(X))
    if not (np.isfinite(cutoff) and cutoff > 0):
        raise ArgumentError("cutoff must be finite and positive (Å)")
    w = workspace or Workspace(device=device)
    w.prepare(sys, cutoff, np.zeros(len(sys), np.int32))
    i, j, S, R = w.export_pairs()
    return NeighbourSnapshot(len(sys), i.astype(np.int64), j.astype(np.int64), S.astype(np.int64), R)
