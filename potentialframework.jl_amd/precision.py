"""Mixed-precision (FP32) against FP64 inference of an ACE model on supplied
structures: ``compare_precision`` and ``PrecisionReport``
(src/submodules/ACE/precision.jl:1-121).

Both versions run on the device from the same model and parameters: the FP64
kernels on the structure as given, and the mixed kernels (FP32 per-edge and
per-site arithmetic, FP64 reductions) on the structure with its positions and
cell rounded to Float32 first, as the reference's Float32 leg receives them
(precision_geometry, src/submodules/ACE/adaptation.jl:122-127). The low-precision
leg therefore carries the geometry rounding of the reference's Float32 path;
its arithmetic is the device's mixed path, not a pure Float32 evaluation (the
neighbour search and the sums stay FP64), which the metadata states.
Per structure and property the report holds the maximum and RMS
absolute component error, where it occurs, and (with tolerances) a pass/fail
status; the aggregate pools all components. A failure (preparation, an error
in either evaluation, nonfinite output) is recorded with its message, makes
that structure's errors infinite and the aggregate status "failed". Indices
(structure, atom, axis, row, column) are 0-based.
"""
import hashlib
import math

import numpy as np

from ._capi import ArgumentError

PROPERTIES = ("energy", "forces", "virial")


class PrecisionReport:
    """Per-structure records, pooled aggregate and metadata (precision.jl:1-12)."""

    def __init__(self, structures, aggregate, metadata):
        self.structures = structures
        self.aggregate = aggregate
        self.metadata = metadata

    def __repr__(self):
        lines = [f"ACE precision comparison: {len(self.structures)} structures; {self.aggregate['status']}",
                 f"  source {self.metadata['source_precision']}; storage {self.metadata['storage']}"]
        for p in PROPERTIES:
            if p in self.aggregate:
                lines.append(f"  {p}: {self.aggregate[p]}")
        return "\n".join(lines)

    __str__ = __repr__


def _device_evaluators(model, ps, props):
    from .ace import ACECalculator

    out = {}
    for name, prec in (("Float64", "float64"), ("Float32", "mixed")):
        try:
            calc = ACECalculator(model, ps, precision=prec)
            out[name] = (lambda c: (lambda sys: c.evaluate(sys, request=props)))(calc)
        except Exception as e:  # recorded per structure, as the reference does
            out[name] = str(e) or type(e).__name__
    return out


def precision_geometry(sys_):
    """The structure with positions and cell rounded to Float32 (held as FP64
    values), species and periodicity unchanged (adaptation.jl:122-127)."""
    from .system import System
    X = np.asarray(sys_.positions, float).astype(np.float32).astype(np.float64)
    H = np.asarray(sys_.cell, float).astype(np.float32).astype(np.float64)
    return System(X, H, sys_.pbc, sys_.labels)


def _precision_of(a):
    a = np.asarray(a)
    return "Float32" if a.dtype == np.float32 else "Float64"


def compare_precision(model, ps, systems, properties=PROPERTIES, tolerances=None, evaluators=None):
    """``compare_precision(m, ps, systems; properties, tolerances)`` (precision.jl:25-121).

    systems: one System or an iterable of them. tolerances: optional dict with
    finite nonnegative "energy" (absolute total energy), "forces" and "virial"
    (maximum component) errors; without it the status is descriptive.
    evaluators (for tests): {"Float64": f, "Float32": f} with f(system) -> dict
    of the requested properties, or a string (a preparation error)."""
    props = tuple(properties)
    if not props or any(p not in PROPERTIES for p in props) or len(set(props)) != len(props):
        raise ArgumentError("properties must be distinct names among energy, forces and virial")
    tolerances = dict(tolerances or {})
    for k, v in tolerances.items():
        if k not in PROPERTIES or isinstance(v, bool) or not isinstance(v, (int, float, np.floating, np.integer)) \
                or not math.isfinite(v) or v < 0:
            raise ArgumentError("tolerances must be finite nonnegative property errors")
    from .system import System
    systems = [systems] if isinstance(systems, System) else list(systems)
    ev = evaluators if evaluators is not None else _device_evaluators(model, ps, props)
    records, totals, failed = [], {}, False
    for index, sys_ in enumerate(systems):
        n = len(sys_) if hasattr(sys_, "__len__") else 0
        results, errors = [], []
        for name in ("Float64", "Float32"):
            f = ev[name]
            try:
                if isinstance(f, str):
                    raise RuntimeError("preparation failed: " + f)
                r = f(sys_ if name == "Float64" else precision_geometry(sys_))
                if not all(np.all(np.isfinite(np.asarray(r[p], float))) for p in props):
                    raise RuntimeError("nonfinite output")
                results.append(r)
            except Exception as e:
                results.append(None)
                errors.append(f"{name}: {e}")
        bad = bool(errors)
        failed |= bad
        metrics = {}
        for key in props:
            if bad:
                metrics[key] = dict(status="failed", maximum=math.inf, rms=math.inf, component=None)
                totals[key] = dict(maximum=math.inf, squares=math.inf, count=0, worst_structure=index, component=None)
                continue
            x = np.asarray(results[0][key], float)
            y = np.asarray(results[1][key], float)
            # components in the reference's column-major order: forces 3 x n, virial 3 x 3
            x = x.reshape(-1, order="F") if x.ndim else x.reshape(1)
            y = y.reshape(-1, order="F") if y.ndim else y.reshape(1)
            diff = y - x
            absdiff = np.abs(diff)
            maxerr = float(absdiff.max()) if diff.size else 0.0
            component = int(absdiff.argmax()) if diff.size else None
            rms = float(math.sqrt(np.sum(diff * diff) / diff.size)) if diff.size else 0.0
            ok = key not in tolerances or maxerr <= tolerances[key]
            status = ("passed" if ok else "failed") if key in tolerances else "descriptive"
            failed |= not ok
            m = dict(status=status, maximum=maxerr, rms=rms, component=component)
            if key == "energy":
                m.update(location=dict(structure=index), absolute=maxerr, per_atom=maxerr / n if n else 0.0)
            elif key == "forces":
                m.update(location=dict(structure=index, atom=None if component is None else component // 3,
                                       axis=None if component is None else component % 3),
                         max_atom_vector=float(np.sqrt((diff.reshape(-1, 3) ** 2).sum(axis=1)).max()) if diff.size else 0.0)
            else:
                m.update(location=dict(structure=index, row=None if component is None else component % 3,
                                       column=None if component is None else component // 3))
            metrics[key] = m
            old = totals.get(key, dict(maximum=-math.inf, squares=0.0, count=0, worst_structure=index, component=None))
            worse = maxerr > old["maximum"]
            totals[key] = dict(maximum=max(old["maximum"], maxerr), squares=old["squares"] + float(np.sum(diff * diff)),
                               count=old["count"] + diff.size, worst_structure=index if worse else old["worst_structure"],
                               component=component if worse else old["component"])
        rounding = cellrounding = math.nan
        identity = hashlib.sha256(f"{type(sys_).__name__}{index}".encode()).hexdigest()
        try:
            X = np.asarray(sys_.positions, float)
            H = np.asarray(sys_.cell, float)
            Hf = H[np.isfinite(H)]
            rounding = float(np.abs(X.astype(np.float32).astype(float) - X).max()) if X.size else 0.0
            cellrounding = float(np.abs(Hf.astype(np.float32).astype(float) - Hf).max()) if Hf.size else 0.0
            identity = hashlib.sha256(X.tobytes() + H.tobytes() + repr(list(sys_.labels)).encode()).hexdigest()
        except Exception as e:
            errors.append("geometry metadata: " + (str(e) or type(e).__name__))
            bad = failed = True
            for key in metrics:
                metrics[key] = dict(status="failed", maximum=math.inf, rms=math.inf, component=None)
                totals[key] = dict(maximum=math.inf, squares=math.inf, count=0, worst_structure=index, component=None)
        status = "failed" if bad or any(v["status"] == "failed" for v in metrics.values()) else \
            ("descriptive" if not tolerances else "passed")
        records.append(dict(index=index, identity=identity, atoms=n, status=status, position_rounding_max_A=rounding,
                            cell_rounding_max_A=cellrounding, errors=errors, metrics=metrics))
    aggregate = {"status": "empty" if not records else "failed" if failed else
                 "descriptive" if not tolerances else "passed"}
    for key, v in totals.items():
        pooled = (0.0 if v["count"] == 0 else math.sqrt(v["squares"] / v["count"])) if math.isfinite(v["squares"]) \
            else math.inf
        aggregate[key] = dict(maximum=v["maximum"], pooled_rms=pooled, components=v["count"],
                              worst_structure=v["worst_structure"], component=v["component"])
    if "energy" in aggregate:
        aggregate["energy"]["maximum_per_atom"] = max((r["metrics"]["energy"].get("per_atom", math.inf)
                                                       for r in records), default=0.0)
    if "forces" in aggregate:
        vec = [r["metrics"]["forces"].get("max_atom_vector", math.inf) for r in records]
        aggregate["forces"]["maximum_atom_vector"] = max(vec, default=0.0)
        aggregate["forces"]["worst_vector_structure"] = int(np.argmax(vec)) if vec else None
    WB = ps.get("WB") if isinstance(ps, dict) else None
    metadata = dict(source_precision=_precision_of(WB) if WB is not None else "Float64",
                    parameter_precision=_precision_of(WB) if WB is not None else "Float64",
                    stored_precision="Float64", storage="device (MI355X HBM)",
                    low_precision_leg=("mixed device kernels (FP32 per-edge and per-site arithmetic, FP64 neighbour "
                                       "search and reductions) on positions and cell rounded to Float32"),
                    model_sha256=getattr(model, "source_sha256", None), properties=props, tolerances=tolerances,
                    preparation_errors=tuple(v if isinstance(v, str) else None for v in (ev["Float64"], ev["Float32"])),
                    units=dict(energy="eV", forces="eV/Å", virial="eV"),
                    weighting="pooled components; failed cases propagate Inf")
    return PrecisionReport(records, aggregate, metadata)
