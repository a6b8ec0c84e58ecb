"""Sweep-based launch autotuner for the pair and ACE steps.

The reference exposes its launch choices as ``GPUConfig(strategy, workgroup,
tile, skin)`` (src/submodules/NeighbourList/gpu.jl:61-73) and the ACE
``strategy`` keyword (src/submodules/ACE/gpu.jl:301-303), picked by hand. Here
the knobs of ``pfw_set_tuning`` and the graph switch of ``pfw_set_graphs`` are
swept on the workload itself: every candidate runs the resident system for a
few warm-up steps and then ``steps`` timed steps (each ends with its status
read, as a caller sees it); the fastest median wins, is applied to the
workspace and is cached in a JSON file keyed by the workload's shape, so a
later ``autotune(..., cache=path)`` on a like workload applies it without a
sweep. Every candidate computes the same pair sets and results to rounding
(the knobs change launch shapes and tile grouping only), so tuning never
changes what is computed.

Cache file: ``cache`` argument, else ``$PFW_TUNING_CACHE``, else
``~/.cache/pfw_mi355x/tuning.json``.
"""
import itertools
import json
import os
import time

import numpy as np

from . import _capi
from ._capi import ArgumentError
from .neighbourlist import Workspace

PAIR_SPACE = {"pair_lanes_per_atom": (4, 3, 2)}
ACE_SPACE = {"ace_sites_per_wave": (2, 1), "ace_radial_chunk": (40, 20, 10), "ace_product_waves": (16, 8, 4),
             "ace_density_contraction": (1, 2, 3)}


def _cache_path(cache):
    if cache:
        return cache
    return os.environ.get("PFW_TUNING_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "pfw_mi355x",
                                                               "tuning.json")


def _bucket(n):
    return int(2 ** round(np.log2(max(n, 1))))


def _target(target, model):
    """(workspace, step entry name, kind, signature) of a Workspace + pair model
    or an ACE calculator."""
    from .ace import ACECalculator

    if isinstance(target, ACECalculator):
        m = target.model
        sig = f"ace|lmax={m.lmax}|nr={m.radial.nr}|dens={len(m.density)}|prod={len(m.products)}"
        return target.workspace, "pfw_ace_step", "ace", sig
    if isinstance(target, Workspace):
        if model is None:
            raise ArgumentError("a pair workspace is tuned with its model")
        kind = type(model).__name__
        return target, "pfw_pair_step", "pair", f"pair|{kind}|cutoff={model.cutoff():.4g}"
    raise ArgumentError("autotune takes a Workspace (with its pair model) or an ACECalculator")


def _key(w, sig, n):
    return f"{sig}|{w.precision}|skin={w.skin:g}|n~{_bucket(n)}"


def _time(w, entry, mode, steps, warmup, repeats=3):
    lib = _capi.load()
    fn = getattr(lib, entry)
    h = w.ctx.h
    for _ in range(warmup):
        _capi.check(h, fn(h, 7, mode))
    meds = []
    for _ in range(repeats):
        ts = []
        for _ in range(steps):
            t0 = time.perf_counter()
            _capi.check(h, fn(h, 7, mode))
            ts.append(time.perf_counter() - t0)
        meds.append(float(np.median(ts)))
    return min(meds)


def _setup(target, system, model):
    from .ace import ACECalculator
    from .pair import evaluate_

    if isinstance(target, ACECalculator):
        target.evaluate(system)
    else:
        evaluate_(target, system, model)


def autotune(target, system, model=None, steps=10, warmup=3, cache=None, use_cache=True, graphs=(False, True)):
    """Sweep the launch knobs (and graph capture) on ``system`` and apply the
    fastest. ``target``: a ``Workspace`` with its pair ``model``, or an
    ``ACECalculator``. Returns ``dict(key, best, ms, candidates, cached)``;
    ``candidates`` lists (knobs, graphs, median ms per step)."""
    w, entry, kind, sig = _target(target, model)
    key = _key(w, sig, len(system))
    path = _cache_path(cache)
    table = {}
    if os.path.exists(path):
        try:
            with open(path) as f:
                table = json.load(f)
        except (OSError, ValueError):
            table = {}
    if use_cache and key in table:
        rec = table[key]
        w.set_tuning(**rec["knobs"])
        w.set_graphs(rec["graphs"])
        _setup(target, system, model)
        return dict(key=key, best=rec["knobs"], graphs=rec["graphs"], ms=rec["ms"], candidates=[], cached=True)
    space = PAIR_SPACE if kind == "pair" else ACE_SPACE
    names = list(space)
    results = []
    for values in itertools.product(*(space[k] for k in names)):
        knobs = dict(zip(names, values))
        for g in graphs:
            w.set_tuning(**knobs)
            w.set_graphs(g)
            _setup(target, system, model)
            ms = 1e3 * _time(w, entry, w.mode, steps, warmup)
            results.append((knobs, bool(g), ms))
    best = min(results, key=lambda r: r[2])
    w.set_tuning(**best[0])
    w.set_graphs(best[1])
    _setup(target, system, model)
    table[key] = {"knobs": best[0], "graphs": best[1], "ms": best[2]}
    try:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        with open(path, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
    except OSError:
        pass
    return dict(key=key, best=best[0], graphs=best[1], ms=best[2], candidates=results, cached=False)
