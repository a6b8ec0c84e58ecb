"""AtomsCalculators-style functions over the GPU path
(src/submodules/PairPotentials/calculators.jl:38-143; ACE in ace.py).

``forces`` returns an (n, 3) array (the vector of 3-vectors of AtomsCalculators);
``forces_`` ADDS into the caller's (n, 3) array like ``AC.forces!``.
"""
import numpy as np

from ._capi import ArgumentError, DimensionMismatch
from .neighbourlist import Workspace, precision_mode
from .pair import ENERGY, FORCES, VIRIAL, EnergyForcesVirial, EnergyOnly, evaluate_


def select_workspace(neighbourlist=None, precision=None, device=0):
    """calculators.jl:23-36."""
    if neighbourlist is None:
        return Workspace(precision=precision or "float64", device=device)
    if not isinstance(neighbourlist, Workspace):
        raise ArgumentError("neighbourlist must be a Workspace or None")
    if precision is not None and precision_mode(precision) != neighbourlist.mode:
        raise ArgumentError("precision conflicts with the supplied neighbourlist workspace")
    return neighbourlist


def _compute(sys, p, request, neighbourlist=None, precision=None, parameters=None, device=0, **_):
    if parameters not in (None, {}):
        raise ArgumentError("explicit parameters require calculate(energy, ...; parameters)")
    w = select_workspace(neighbourlist, precision, device)
    return evaluate_(w, sys, p, request=request)


def potential_energy(sys, p, **kw):
    return _compute(sys, p, EnergyOnly, **kw)["energy"]


def forces(sys, p, **kw):
    return _compute(sys, p, EnergyForcesVirial, **kw)["forces"].T.copy()


def virial(sys, p, **kw):
    return _compute(sys, p, EnergyForcesVirial, **kw)["virial"]


def energy_forces(sys, p, **kw):
    r = _compute(sys, p, EnergyForcesVirial, **kw)
    return {"energy": r["energy"], "forces": r["forces"].T.copy()}


def energy_forces_virial(sys, p, **kw):
    r = _compute(sys, p, EnergyForcesVirial, **kw)
    return {"energy": r["energy"], "forces": r["forces"].T.copy(), "virial": r["virial"]}


def _check_destination(f, sys):
    if f.shape != (len(sys), 3):
        raise DimensionMismatch("force destination must have one entry per atom")


def forces_(f, sys, p, **kw):
    _check_destination(f, sys)
    f += forces(sys, p, **kw)
    return f


def energy_forces_(f, sys, p, **kw):
    _check_destination(f, sys)
    r = energy_forces(sys, p, **kw)
    f += r["forces"]
    return {"energy": r["energy"], "forces": f}


def energy_forces_virial_(f, sys, p, **kw):
    _check_destination(f, sys)
    r = energy_forces_virial(sys, p, **kw)
    f += r["forces"]
    return {"energy": r["energy"], "forces": f, "virial": r["virial"]}


def calculate(properties, sys, p, parameters=None, state=None, **kw):
    """``AC.calculate`` (calculators.jl:100-143): one fused traversal for a tuple of
    properties; explicit parameters go through ``with_parameters``."""
    props = properties if isinstance(properties, (tuple, list)) else (properties,)
    for q in props:
        if q not in (ENERGY, FORCES, VIRIAL):
            raise ArgumentError("supported properties are energy, forces and virial")
    if not props:
        return {"state": state}
    if parameters not in (None, {}):
        p = p.with_parameters(parameters)
    w = select_workspace(kw.get("neighbourlist"), kw.get("precision"), kw.get("device", 0))
    r = evaluate_(w, sys, p, request=tuple(props))
    out = {}
    for q in props:
        out[q] = r[q].T.copy() if q == FORCES else r[q]
    out["state"] = state
    return out
