"""Linear fitting helpers of the ACE readout (src/submodules/ACE/fitting.jl).

* ``design_matrix``: the dense energy / force / virial design of
  ``design_matrix_impl`` (fitting.jl:18-63), computed on the device.
  E, F and W are linear in the readout (WB, Wpair), so column c of the design
  is the outputs of the unit parameter set e_c less those of the zero set
  (which are the baseline). ``pfw_ace_parameter_sets`` evaluates all of
  those sets over one environment build: the site kernels, force gather and
  reductions of a step, once per column.
* ``training_rows``: label selection and baseline subtraction
  (fitting.jl:103-126).
* ``linear_parameters`` / ``install_readout``: the flat readout vector and its
  inverse (fitting.jl:1-11).

Row layout, per configuration in order: the energy, vec(F) (3 x n, original
atom order), vec(W) (3 x 3, column-major); 1 + 3n + 9 rows. Columns: vec(WB)
(nbasis x nz, column-major), then vec(Wpair) (nq x nz) with ``fit_pair``.
Masked invariants (basis policy) and a disabled pair term give zero columns,
as the reference's masked B and zeroed Q do.
"""
import ctypes

import numpy as np

from ._capi import ArgumentError, DimensionMismatch, ptr


def linear_parameters(ps):
    """``linear_parameters``: vcat(vec(WB), vec(Wpair)) (fitting.jl:1)."""
    return np.concatenate([np.asarray(ps["WB"], float).T.reshape(-1), np.asarray(ps["Wpair"], float).T.reshape(-1)])


def install_readout(ps, coefficients, fit_pair=True):
    """``install_readout`` (fitting.jl:2-11): a parameter dictionary with the
    readout taken from a flat coefficient vector (column-major blocks)."""
    WB = np.asarray(ps["WB"], float)
    Wp = np.asarray(ps["Wpair"], float)
    c = np.asarray(coefficients, float).reshape(-1)
    n = WB.size
    if not fit_pair:
        if len(c) != n:
            raise DimensionMismatch("ACE readout length")
        return {"WB": c.reshape(WB.shape[1], WB.shape[0]).T.copy(), "Wpair": Wp.copy()}
    if len(c) != n + Wp.size:
        raise DimensionMismatch("readout length")
    return {"WB": c[:n].reshape(WB.shape[1], WB.shape[0]).T.copy(),
            "Wpair": c[n:].reshape(Wp.shape[1], Wp.shape[0]).T.copy()}


def design_matrix(calc, systems, fit_pair=True, centers=None):
    """``design_matrix(X, model, ps, env, H; fit_pair)`` for one system, or the
    stacked designs of a list of systems (one batch on the device).

    Returns ``{"matrix": A, "baseline": b}`` with ``A @ theta + b`` equal to
    the stacked [E; vec(F); vec(W)] of every configuration for the readout
    ``theta = linear_parameters(ps)`` (or ``vec(WB)`` alone without
    ``fit_pair``, the current ``Wpair`` then entering the baseline). Computed
    in the calculator's precision."""
    if centers is not None:
        raise ArgumentError("center selection is not supported by the device design")
    single = not isinstance(systems, (list, tuple))
    systems = [systems] if single else list(systems)
    if not systems:
        raise ArgumentError("no configurations")
    m = calc.model
    nb, nz, nq = m.nbasis, m.nz, m.pair.nr
    w = calc.workspace
    chans = [m.channels(s.labels) for s in systems]
    if len(systems) == 1:
        w.prepare(systems[0], m.cutoff, chans[0])
    else:
        w.prepare_batch(systems, m.cutoff, chans)
    rows = sum(10 + 3 * len(s) for s in systems)
    nbasis_cols = nz * nb
    # unprojected bases: the vec(WB) columns in one pass of per-site Jacobians
    # (pfw_ace_design); the parameter sets then give only the Wpair columns and
    # the baseline. Projected bases take every column from parameter sets.
    jac = m.orders is None
    nsets_b = 0 if jac else nbasis_cols
    npair = nz * nq if (fit_pair and nq > 0) else 0
    nsets = nsets_b + npair + 1
    # parameter sets: the unit readouts (masked as the device model masks them), then the baseline set
    WB = np.zeros((nsets, nz, nb))   # set k: WB^T (nz x nbasis), i.e. vec(WB) column-major
    WP = np.zeros((nsets, nz, max(nq, 1)))
    if nsets_b:
        eye = np.arange(nbasis_cols)
        WB.reshape(nsets, -1)[eye, eye] = np.tile(m.mask.astype(float), nz)
    if npair:
        k = nsets_b + np.arange(npair)
        WP[k, np.repeat(np.arange(nz), nq), np.tile(np.arange(nq), nz)] = 1.0 if m.pair_enabled else 0.0
    if not fit_pair and nq > 0 and m.pair_enabled:  # the current pair term in every set, so it cancels in A
        WP[:] = np.asarray(calc.parameters["Wpair"], float).T[None]
    WP = WP[:, :, :nq] if nq > 0 else None
    out = np.zeros((nsets, rows))  # column-major (rows x sets)
    w.ctx.call("pfw_ace_parameter_sets", w.mode, nsets, ptr(np.ascontiguousarray(WB)),
               ptr(np.ascontiguousarray(WP)) if WP is not None else None, ptr(out))
    base = out[nsets - 1].copy()
    A = np.empty((rows, nbasis_cols + npair))
    if jac:  # written straight into the first nb nz columns (row stride = A's width)
        w.ctx.call("pfw_ace_design", w.mode, ptr(np.ascontiguousarray(m.mask, np.float64)), ptr(A),
                   ctypes.c_int64(A.shape[1]))
    else:
        A[:, :nbasis_cols] = (out[:nbasis_cols] - base[None, :]).T
    if npair:
        A[:, nbasis_cols:] = (out[nsets_b:nsets_b + npair] - base[None, :]).T
    return {"matrix": A, "baseline": base}


def training_rows(design, labels, weights=None):
    """``training_rows`` (fitting.jl:103-126) for one configuration's design:
    the rows of the labels present (``energy``, ``forces`` as 3 x n or n
    three-vectors, ``virial`` 3 x 3; a missing key or None omits them), the
    baseline-subtracted labels, and per-row residual weights. Labels are plain
    numbers in eV (energy, virial) and eV/Å (forces)."""
    wts = {"energy": 1.0, "forces": 1.0, "virial": 1.0}
    wts.update(weights or {})
    A = np.asarray(design["matrix"])
    n3 = A.shape[0] - 10
    idx, vals, rw = [], [], []
    for key, rng in (("energy", range(0, 1)), ("forces", range(1, 1 + n3)), ("virial", range(1 + n3, 10 + n3))):
        v = labels.get(key) if isinstance(labels, dict) else getattr(labels, key, None)
        if v is None:
            continue
        if key == "energy":
            v = np.array([float(v)])
        elif key == "forces" and isinstance(v, (list, tuple)) and all(np.ndim(x) == 1 for x in v):
            if not all(len(x) == 3 for x in v):
                raise DimensionMismatch("forces need three components per atom")
            v = np.concatenate([np.asarray(x, float) for x in v])
        else:
            v = np.asarray(v, float)
            # a 2-D force label must be the reference's 3 x n matrix: an (n, 3)
            # array (one row per atom) would be laid out axis by axis here
            if key == "forces" and v.ndim == 2 and v.shape[0] != 3:
                raise DimensionMismatch("2-D force labels must be 3 x n (one column per atom); pass an (n, 3) "
                                        "array as a list of three-vectors or transpose it")
            if key == "virial" and v.ndim == 2 and v.shape != (3, 3):
                raise DimensionMismatch("virial labels must be 3 x 3")
            v = v.T.reshape(-1)  # vec() of a column-major 3 x n / 3 x 3 array
        if len(v) != len(rng):
            raise DimensionMismatch(f"{key} label dimensions")
        wt = wts[key]
        if not (np.isfinite(wt) and wt >= 0):
            raise ArgumentError(f"invalid {key} weight")
        idx.extend(rng)
        vals.extend(v)
        rw.extend([float(wt)] * len(rng))
    idx = np.asarray(idx, int)
    return {"matrix": A[idx, :], "labels": np.asarray(vals, float) - np.asarray(design["baseline"])[idx],
            "row_weights": np.asarray(rw, float)}
