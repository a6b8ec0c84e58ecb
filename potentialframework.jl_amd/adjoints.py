"""Parameter adjoints (reverse mode) of the pair and ACE paths on the device.

Mirrors the reference's ChainRules boundary for parameter gradients:

* ``rrule_pair_evaluate`` / ``pair_parameter_vjp``: the ``pair_evaluate`` rrule
  (ext/PotentialFrameWorkChainRulesCoreExt.jl:26-52) restricted to parameter
  cotangents, with ``pair_vjp`` (ext/PotentialFrameWorkForwardDiffExt.jl:27-54)
  and ``energy_parameter_gradient(::SmoothPair)``
  (src/submodules/PairPotentials/parameters.jl:126-138).
* ``rrule_ace_evaluate`` / ``ace_parameter_vjp``: the ``ace_evaluate`` rrule
  (ext/PotentialFrameWorkChainRulesCoreExt.jl:54-76) with ``ace_vjp_impl``
  (ext/PotentialFrameWorkForwardDiffExt.jl:90-113) and
  ``site_parameter_gradient`` (src/submodules/ACE/serial.jl:103-121).
* ``fitting_loss``: weighted least squares over energies, forces and virials of
  a batch of configurations and its parameter gradient (one forward step and
  one adjoint step on the device, whatever the number of configurations).

Cotangents follow the outputs' layout: one energy cotangent per configuration,
forces ``3 x n`` per configuration, virial ``3 x 3``. Missing cotangents are
zero. Gradients are returned in the parameters' own shapes.
"""
import numpy as np

from ._capi import ArgumentError, DimensionMismatch, ptr
from .pair import SmoothPair


def _as_list(systems):
    single = not isinstance(systems, (list, tuple))
    return ([systems] if single else list(systems)), single


def _cotangents(systems, dE, dF, dW):
    """Pack per-configuration cotangents into the C ABI's flat buffers."""
    ncfg = len(systems)
    E = np.zeros(ncfg)
    if dE is not None:
        E[:] = np.broadcast_to(np.asarray(dE, float), (ncfg,))
    F = np.zeros((sum(len(s) for s in systems), 3))
    if dF is not None:
        if ncfg == 1 and not isinstance(dF, (list, tuple)):
            dF = [dF]
        if len(dF) != ncfg:
            raise DimensionMismatch("one force cotangent per configuration")
        o = 0
        for s, f in zip(systems, dF):
            n = len(s)
            if f is not None:
                f = np.asarray(f, float)
                if f.shape != (3, n):
                    raise DimensionMismatch("force cotangent must be 3 x n")
                F[o:o + n] = f.T
            o += n
    W = None
    if dW is not None:
        if ncfg == 1 and not isinstance(dW, (list, tuple)):
            dW = [dW]
        if len(dW) != ncfg:
            raise DimensionMismatch("one virial cotangent per configuration")
        W = np.zeros((ncfg, 9))
        for c, w in enumerate(dW):
            if w is not None:
                w = np.asarray(w, float)
                if w.shape != (3, 3):
                    raise DimensionMismatch("virial cotangent must be 3 x 3")
                W[c] = w.T.reshape(9)  # column-major
    if not (np.all(np.isfinite(E)) and np.all(np.isfinite(F)) and (W is None or np.all(np.isfinite(W)))):
        raise ArgumentError("cotangents must be finite")
    return E, F, W


def pair_parameter_vjp(w, systems, p, dE=None, dF=None, dW=None, parameters=None):
    """Gradient of sum_c dE_c E_c + <dF, F> + <dW, W> with respect to the
    SmoothPair amplitudes, as ``{"amplitudes": g}`` over the independent upper
    triangle (the order of ``SmoothPair.get_parameters``). FP64 on the device."""
    if parameters is not None:
        p = p.with_parameters(parameters)
    if not isinstance(p, SmoothPair):
        raise ArgumentError("parameter adjoints are implemented for SmoothPair")
    systems, _ = _as_list(systems)
    chans = []
    for sys in systems:
        ch, kind, nz, params = p._device_model(sys.labels)
        chans.append(ch)
    if len(systems) == 1:
        w.prepare(systems[0], p.cutoff(), chans[0])
    else:
        w.prepare_batch(systems, p.cutoff(), chans)
    w.ctx.call("pfw_set_pair_model", kind, nz, ptr(np.ascontiguousarray(params)), float(p.cutoff()))
    w._model_key = None
    E, F, W = _cotangents(systems, dE, dF, dW)
    gA = np.zeros(nz * nz)
    w.ctx.call("pfw_pair_adjoint", ptr(E), ptr(F), ptr(W), ptr(gA))
    G = gA.reshape(nz, nz)  # G[a, b]: entry read by pairs of centre a, neighbour b
    g = np.array([G[i, j] + G[j, i] if i != j else G[i, i] for j in range(nz) for i in range(j + 1)])
    return {"amplitudes": g}


def ace_parameter_vjp(calc, systems, dE=None, dF=None, dW=None):
    """Gradient of sum_c dE_c E_c + <dF, F> + <dW, W> with respect to the ACE
    readout, as ``{"WB": (nbasis, nz), "Wpair": (nq, nz)}``, in the calculator's
    precision (mixed: FP32 site arithmetic, FP64 accumulation)."""
    systems, _ = _as_list(systems)
    w = calc.workspace
    m = calc.model
    chans = [m.channels(s.labels) for s in systems]
    if len(systems) == 1:
        w.prepare(systems[0], m.cutoff, chans[0])
    else:
        w.prepare_batch(systems, m.cutoff, chans)
    E, F, W = _cotangents(systems, dE, dF, dW)
    gWB = np.zeros((m.nz, m.nbasis))
    gWp = np.zeros((m.nz, max(m.pair.nr, 1)))
    w.ctx.call("pfw_ace_adjoint", w.mode, ptr(E), ptr(F), ptr(W), ptr(gWB), ptr(gWp))
    # the device evaluates effective_WB = WB .* mask and Wpair only with the pair term on
    gp = gWp[:, :m.pair.nr].T.copy()
    return {"WB": gWB.T * m.mask[:, None], "Wpair": gp if m.pair_enabled else np.zeros_like(gp)}


def _pullback_factory(vjp, systems, single):
    def pullback(delta):
        deltas = [delta] if single else list(delta)
        if len(deltas) != len(systems):
            raise DimensionMismatch("one cotangent per configuration")
        dE = [d.get("energy", 0.0) if d is not None else 0.0 for d in deltas]
        dF = [d.get("forces") if d is not None else None for d in deltas]
        dW = [d.get("virial") if d is not None else None for d in deltas]
        if all(v is None for v in dW):
            dW = None
        return vjp(dE, dF, dW)
    return pullback


def rrule_pair_evaluate(w, systems, p, parameters=None):
    """``(y, pullback)``: y as from ``evaluate_`` (one system) or ``evaluate_batch``;
    ``pullback(delta)`` takes cotangent dictionaries with any of energy, forces,
    virial and returns the amplitude gradient. The pullback recomputes from the
    geometry and parameters captured now (later caller mutation has no effect)."""
    from .pair import evaluate_, evaluate_batch
    systems_l, single = _as_list(systems)
    if parameters is not None:
        p = p.with_parameters(parameters)
    y = evaluate_(w, systems_l[0], p) if single else evaluate_batch(w, systems_l, p)
    saved = [s.copy() for s in systems_l]
    saved_p = SmoothPair(p.labels, p.amplitudes, p.radius) if isinstance(p, SmoothPair) else p

    def vjp(dE, dF, dW):
        return pair_parameter_vjp(w, saved, saved_p, dE, dF, dW)
    return y, _pullback_factory(vjp, saved, single)


def rrule_ace_evaluate(calc, systems):
    """``(y, pullback)`` for the ACE calculator's current parameters; the pullback
    returns ``{"WB", "Wpair"}`` gradients (parameters and geometry captured now)."""
    systems_l, single = _as_list(systems)
    y = calc.evaluate(systems_l[0]) if single else calc.evaluate_batch(systems_l)
    saved = [s.copy() for s in systems_l]
    ps = {"WB": calc.parameters["WB"].copy(), "Wpair": calc.parameters["Wpair"].copy()}

    def vjp(dE, dF, dW):
        if not (np.array_equal(calc.parameters["WB"], ps["WB"])
                and np.array_equal(calc.parameters["Wpair"], ps["Wpair"])):
            calc.set_parameters(ps)
        return ace_parameter_vjp(calc, saved, dE, dF, dW)
    return y, _pullback_factory(vjp, saved, single)


def fitting_loss(evaluate, vjp, systems, targets, weights=None):
    """Weighted least squares over a batch and its parameter gradient.

    ``L = sum_c [wE (E_c - E*_c)^2 / n_c^2 + wF |F_c - F*_c|^2 + wV |W_c - W*_c|^2 / n_c^2]``
    (energies and virials per atom). ``evaluate(systems)`` returns one result
    dictionary per configuration, ``vjp(dE, dF, dW)`` the parameter gradient for
    those cotangents; ``targets`` holds per-configuration dictionaries with
    ``energy``, ``forces`` (3 x n) and optionally ``virial``. Returns (L, grad)."""
    wts = {"energy": 1.0, "forces": 1.0, "virial": 0.0}
    wts.update(weights or {})
    res = evaluate(systems)
    L = 0.0
    dE, dF, dW = [], [], []
    for s, r, t in zip(systems, res, targets):
        n = len(s)
        e = r["energy"] - t["energy"]
        L += wts["energy"] * e * e / n ** 2
        dE.append(2 * wts["energy"] * e / n ** 2)
        f = r["forces"] - t["forces"]
        L += wts["forces"] * float(np.sum(f * f))
        dF.append(2 * wts["forces"] * f)
        if wts["virial"] and "virial" in t:
            v = r["virial"] - t["virial"]
            L += wts["virial"] * float(np.sum(v * v)) / n ** 2
            dW.append(2 * wts["virial"] * v / n ** 2)
        else:
            dW.append(np.zeros((3, 3)))
    return L, vjp(dE, dF, dW if wts["virial"] else None)
