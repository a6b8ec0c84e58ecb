"""Pair potentials and the ``evaluate!`` boundary on the GPU.

Mirrors src/submodules/PairPotentials/models.jl (SmoothPair), lennard_jones.jl
(energy-shifted LJ), parameters.jl (triangle parameter order) and
evaluation.jl:31-73 (``evaluate!`` / ``evaluate``). Results own their storage:
``energy`` (eV), ``forces`` as a 3 x N array in ORIGINAL atom order (eV/Å) and
``virial`` 3 x 3 (eV), the reference's NamedTuple without the Unitful wrapper.
"""
import numpy as np

from . import _capi
from ._capi import ArgumentError, pinned_empty, ptr
from .neighbourlist import Workspace
from .system import channels, unique_in_order

ENERGY, FORCES, VIRIAL = "energy", "forces", "virial"


class EnergyOnly:
    """Request only the energy (evaluation.jl / models.jl:16)."""


class EnergyForcesVirial:
    """Request energy, forces and virial (models.jl:17)."""


def request_mask(request):
    """Normalise a request to the C mask (normalize_properties, execution.jl:4-12)."""
    if isinstance(request, EnergyOnly) or request is EnergyOnly:
        return _capi.PFW_ENERGY, (ENERGY,)
    if isinstance(request, EnergyForcesVirial) or request is EnergyForcesVirial:
        return 7, (ENERGY, FORCES, VIRIAL)
    props = request if isinstance(request, (tuple, list)) else (request,)
    mask = 0
    for p in props:
        if p == ENERGY:
            mask |= _capi.PFW_ENERGY
        elif p == FORCES:
            mask |= _capi.PFW_FORCES
        elif p == VIRIAL:
            mask |= _capi.PFW_VIRIAL
        else:
            raise ArgumentError("supported properties are energy, forces and virial")
    names = tuple(n for n, b in ((ENERGY, 1), (FORCES, 2), (VIRIAL, 4)) if mask & b)
    return mask, names


class SmoothPair:
    """U_ab(r) = A_ab (1 - (r/c)^2)^3 for r < c (models.jl:24-44). ``labels`` are
    species labels, ``amplitudes`` a symmetric table in eV, ``radius`` the cutoff in Å."""

    def __init__(self, labels, amplitudes, radius):
        labels = list(labels)
        A = np.array(amplitudes, dtype=np.float64, copy=True)
        n = len(labels)
        if A.shape != (n, n):
            raise ArgumentError("parameter table size must match species labels")
        if len(set(labels)) != n:
            raise ArgumentError("species labels must be unique")
        if not (np.all(np.isfinite(A)) and np.array_equal(A, A.T)):
            raise ArgumentError("amplitude table must be finite and symmetric")
        if not (np.isfinite(radius) and radius > 0):
            raise ArgumentError("cutoff must be finite and positive")
        self.labels = labels
        self.amplitudes = A
        self.radius = float(radius)

    def cutoff(self):
        return self.radius

    def get_parameters(self):
        """Upper triangle in column order: A11, A12, A22, ... (parameters.jl:34-37)."""
        n = len(self.labels)
        return {"amplitudes": np.array([self.amplitudes[i, j] for j in range(n) for i in range(j + 1)])}

    def with_parameters(self, parameters):
        """parameters.jl:38-53."""
        if set(parameters) != {"amplitudes"}:
            raise ArgumentError("SmoothPair parameters must contain only amplitudes; cutoff is fixed")
        a = np.asarray(parameters["amplitudes"], dtype=np.float64)
        n = len(self.labels)
        if a.ndim != 1 or len(a) != n * (n + 1) // 2:
            raise ArgumentError("amplitudes must be a real vector of length n*(n+1)/2")
        if not np.all(np.isfinite(a)):
            raise ArgumentError("amplitudes must be finite")
        table = np.zeros((n, n))
        k = 0
        for j in range(n):
            for i in range(j + 1):
                table[i, j] = table[j, i] = a[k]
                k += 1
        return SmoothPair(self.labels, table, self.radius)

    def _device_model(self, labels):
        ch = channels(labels, self.labels)
        return ch, _capi.PFW_PAIR_SMOOTH, len(self.labels), np.ascontiguousarray(self.amplitudes.T)


class LennardJonesRadial:
    """E = 4 eps [(s/r)^12 - (s/r)^6 - (s/c)^12 + (s/c)^6] (lennard_jones.jl:5-17)."""

    def __init__(self, epsilon, sigma, radius):
        self.epsilon, self.sigma, self.radius = float(epsilon), float(sigma), float(radius)


class RadialPotential:
    """Radial adapter (models.jl:72-90) with a certified device radial."""

    def __init__(self, radial, radius):
        self.radial = radial
        self.radius = float(radius)

    def cutoff(self):
        return self.radius

    def get_parameters(self):
        if isinstance(self.radial, LennardJonesRadial):
            return {"epsilon": self.radial.epsilon, "sigma": self.radial.sigma}
        raise ArgumentError("radial has no parameters")

    def with_parameters(self, ps):
        if not isinstance(self.radial, LennardJonesRadial) or set(ps) != {"epsilon", "sigma"}:
            raise ArgumentError("LJ parameters require epsilon and sigma; cutoff is fixed")
        return lennard_jones(epsilon=ps["epsilon"], sigma=ps["sigma"], radius=self.radius)

    def _device_model(self, labels):
        if not isinstance(self.radial, LennardJonesRadial):
            raise ArgumentError("radial callable needs device certification for GPU execution")
        ch = channels(labels, unique_in_order(labels))
        return ch, _capi.PFW_PAIR_LJ, 1, np.array([self.radial.epsilon, self.radial.sigma])


def lennard_jones(epsilon=0.0103, sigma=3.4, radius=8.5):
    """lennard_jones.jl:18-22."""
    e, s, c = float(epsilon), float(sigma), float(radius)
    if not (all(np.isfinite(v) for v in (e, s, c)) and e > 0 and s > 0 and c > 0):
        raise ArgumentError("LJ parameters must be positive and finite")
    return RadialPotential(LennardJonesRadial(e, s, c), c)


def evaluate_(w, sys, p, request=EnergyForcesVirial, parameters=None):
    """``evaluate!(w, sys, p; request, parameters)`` (evaluation.jl:31-52): rebuild
    geometry and topology on the device, evaluate, return owning original-order outputs."""
    if not isinstance(w, Workspace):
        raise ArgumentError("neighbourlist must be a Workspace or None")
    if parameters is not None:
        p = p.with_parameters(parameters)
    mask, names = request_mask(request)
    ch, kind, nz, params = p._device_model(sys.labels)
    w.prepare(sys, p.cutoff(), ch)
    model_key = (kind, nz, params.tobytes(), p.cutoff())
    if w._model_key != model_key:
        w.ctx.call("pfw_set_pair_model", kind, nz, ptr(np.ascontiguousarray(params)), float(p.cutoff()))
        w._model_key = model_key
    n = len(sys)
    E = np.zeros(1)
    F = pinned_empty(3 * n).reshape(n, 3)  # written whole when forces are requested
    W = np.zeros(9)
    w.ctx.call("pfw_pair_evaluate", mask, w.mode, ptr(E), ptr(F), ptr(W))
    if w.timing:
        w.read_timings()
    out = {}
    if ENERGY in names:
        out["energy"] = float(E[0])
    if FORCES in names:
        out["forces"] = F.T  # 3 x n, an owning view (nothing else holds F)
    if VIRIAL in names:
        out["virial"] = W.reshape(3, 3).T.copy()
    return out


def evaluate_batch(w, systems, p, request=EnergyForcesVirial, parameters=None):
    """Evaluate many independent configurations with one launch per stage
    (fitting sets). Returns one result dictionary per configuration."""
    if parameters is not None:
        p = p.with_parameters(parameters)
    mask, names = request_mask(request)
    chans, kind, nz, params = None, None, None, None
    chans = []
    for sys in systems:
        ch, kind, nz, params = p._device_model(sys.labels)
        chans.append(ch)
    offsets = w.prepare_batch(systems, p.cutoff(), chans)
    w.ctx.call("pfw_set_pair_model", kind, nz, ptr(np.ascontiguousarray(params)), float(p.cutoff()))
    w.ctx.call("pfw_pair_step", mask, w.mode)
    if w.timing:
        w.read_timings()
    return w.read_batch(offsets, names)


def evaluate(sys, p, precision="float64", device=0, **kwargs):
    """``evaluate(sys, p; ...)`` with a fresh workspace (evaluation.jl:53-55)."""
    return evaluate_(Workspace(precision=precision, device=device), sys, p, **kwargs)
