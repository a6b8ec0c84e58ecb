"""Per-atom equivariant ACE features (src/submodules/ACE/features.jl,
feature_coupling.jl) with the site computation on the device.

``ACEFeatureLayer`` holds numerical tables:

* ``density``: (neighbour species channel, radial index, harmonic index);
  channel 0 sums every species, 1..nz selects one;
* ``products``: tuples of density indices;
* ``coupling``: one sparse block per output irrep (L, parity, channels) with
  (2L+1) x basis_channels rows, m fastest.

The tables come from the generated construction (``generate_feature_tables``,
feature_coupling.jl:96-143: species-resolved density multiplets coupled with
real Clebsch-Gordan tensors along paths of at most ``correlation`` factors
with sum(n + l) <= ``totaldegree``, then orthonormalised per multiplet), from
explicit tables (``ACEFeatureLayer.from_tables``), or from a resolved scalar
model (``ACEFeatureLayer.from_resolved``, the L = 0 adapter of features.jl:84-94).

``ace_features(layer, system, ps)`` returns a tuple with one
(2L+1, channels, n) array per irrep, atoms in input order
(features.jl:164-178): B_b = coupling_b P reshaped (2L+1, basis), times the
mixing (basis, channels) of the centre's species.

Radials: the registered smooth exponential family
(``PotentialFrameWork.smooth_exponential.v1``, radials.jl:187-206),
R_k(r) = amplitude (r/rc)^p_k (1 - r/rc)^4 exp(-rate r), times the radial scale
parameter (radial, centre species, neighbour species). The resolved adapter
keeps the resolved model's own radial tables and runs on its ACE kernels.
"""
import ctypes
import math

import numpy as np

from ._capi import ArgumentError, DimensionMismatch, pinned_empty, ptr
from .ace_basis import _couple, _independent, real_cg

PFW_RADIAL_SMOOTH_EXPONENTIAL = 1


class FeatureIrrep:
    """One output block: angular momentum L, parity, channels (feature_coupling.jl:3-7)."""

    def __init__(self, L, parity, channels):
        self.L, self.parity, self.channels = L, parity, channels

    def __repr__(self):
        return f"FeatureIrrep(L={self.L}, parity={self.parity}, channels={self.channels})"


def feature_irreps(spec):
    """Validated irreps of ((L=, parity=, channels=), ...) (feature_coupling.jl:8-21)."""
    spec = tuple(spec)
    if not spec:
        raise ArgumentError("output_irreps must be nonempty")
    out = []
    for s in spec:
        if not isinstance(s, dict) or set(s) != {"L", "parity", "channels"}:
            raise ArgumentError("each irrep must be (L=..., parity=..., channels=...)")
        L, par, ch = s["L"], s["parity"], s["channels"]
        if isinstance(L, bool) or not isinstance(L, (int, np.integer)) or L < 0:
            raise ArgumentError("L must be a nonnegative integer")
        if isinstance(ch, bool) or not isinstance(ch, (int, np.integer)) or ch <= 0:
            raise ArgumentError("channels must be a positive integer")
        if par not in ("even", "odd"):
            raise ArgumentError("parity must be :even or :odd")
        out.append(FeatureIrrep(int(L), par, int(ch)))
    if len({(s.L, s.parity) for s in out}) != len(out):
        raise ArgumentError("duplicate (L, parity) block")
    return out


def generate_feature_tables(nz, rspec, correlation, lmax, totaldegree, irreps):
    """Density, products and coupling blocks of the generated layer
    (generate_feature_tables, feature_coupling.jl:96-143). Density entries are
    (channel 1..nz, radial index 0-based, harmonic index 0-based); products are
    tuples of 0-based density indices; couplings are dense (rows x nprod)."""
    density, factors, vectors = [], [], []
    for z in range(1, nz + 1):
        for r, (n, l) in enumerate(rspec):
            if not 0 <= l <= lmax:
                continue
            factors.append((z, r, l, n + l))
            idx = []
            for m in range(-l, l + 1):
                density.append((z, r, l * l + l + m))
                idx.append(len(density) - 1)
            vectors.append([{(i,): 1.0} for i in idx])
    blocks = [[] for _ in irreps]

    def visit(first, degree, order, parity, states):
        if order >= correlation:
            return
        for f in range(first, len(factors)):
            _, _, l, weight = factors[f]
            if degree + weight > totaldegree:
                continue
            nxt = {}
            for J in sorted(states):
                for x in states[J]:
                    for L in range(abs(J - l), J + l + 1):
                        y = _couple(x, vectors[f], real_cg(J, l, L))
                        _independent(nxt.setdefault(L, []), y)
            p = parity * (-1) ** l
            for b, s in enumerate(irreps):
                if (1 if s.parity == "even" else -1) == p:
                    blocks[b].extend(nxt.get(s.L, []))
            visit(f, degree + weight, order + 1, p, nxt)

    visit(0, 0, 0, 1, {0: [[{(): 1.0}]]})
    for b, s in enumerate(irreps):
        if not blocks[b]:
            raise ArgumentError(f"no basis channels survive for L={s.L}, parity={s.parity}")
    keys = set()
    for block in blocks:
        for x in block:
            for p in x:
                keys.update(p)
    products = sorted(keys, key=lambda k: (len(k), k))
    lookup = {p: i for i, p in enumerate(products)}
    coupling = []
    for b, s in enumerate(irreps):
        dim = 2 * s.L + 1
        C = np.zeros((dim * len(blocks[b]), len(products)))
        for i, x in enumerate(blocks[b]):
            for m in range(dim):
                for k, v in sorted(x[m].items()):
                    C[i * dim + m, lookup[k]] += v
        coupling.append(C)
    return density, [tuple(p) for p in products], coupling


class SmoothExponential:
    """The registered smooth exponential radial family (radials.jl:187-206):
    labels ``spec`` [(n, l)], cutoff and one power per label; the rate and
    amplitude are parameters (``custom_radial``)."""

    def __init__(self, spec, cutoff, powers):
        spec = [tuple(int(v) for v in s) for s in spec]
        if not spec or any(n <= 0 or l < 0 for n, l in spec) or len(set(spec)) != len(spec):
            raise ArgumentError("radial spec must have unique positive n, nonnegative l labels")
        if len(powers) != len(spec) or any(int(p) < 0 for p in powers):
            raise DimensionMismatch("one nonnegative power per radial label")
        if not (np.isfinite(cutoff) and cutoff > 0):
            raise ArgumentError("radial cutoff must be positive")
        self.spec, self.cutoff, self.powers = spec, float(cutoff), [int(p) for p in powers]

    def values(self, r, rate, amplitude):
        """R(r) for every label (FP64, host; the device evaluates the same formula)."""
        x = r / self.cutoff
        q = max(1.0 - x, 0.0)
        ex = math.exp(-rate * r)
        return np.array([amplitude * x ** p * q ** 4 * ex for p in self.powers])


class ACEFeatureLayer:
    """``ACEFeatureLayer(; species, radial, correlation, lmax, output_irreps,
    cutoff, totaldegree)`` (features.jl:21-47), generated tables."""

    def __init__(self, species, radial, correlation, lmax, output_irreps, cutoff=None, totaldegree=None):
        if isinstance(correlation, bool) or not isinstance(correlation, (int, np.integer)) or correlation <= 0:
            raise ArgumentError("correlation must be a positive integer")
        if isinstance(lmax, bool) or not isinstance(lmax, (int, np.integer)) or lmax < 0:
            raise ArgumentError("lmax must be a nonnegative integer")
        irreps = feature_irreps(output_irreps)
        z = _atomic_numbers(species)
        if not isinstance(radial, SmoothExponential):
            raise ArgumentError("the device feature layer supports the smooth exponential radial family")
        rc = radial.cutoff
        if cutoff is not None:
            if not (np.isfinite(cutoff) and cutoff >= rc):
                raise ArgumentError("cutoff must cover every radial cutoff")
            rc = float(cutoff)
        td = correlation * max(n + l for n, l in radial.spec) if totaldegree is None else totaldegree
        if isinstance(td, bool) or not isinstance(td, (int, np.integer)) or td <= 0:
            raise ArgumentError("totaldegree must be a positive integer")
        density, products, coupling = generate_feature_tables(len(z), radial.spec, int(correlation), int(lmax),
                                                              int(td), irreps)
        self._set(z, radial, int(lmax), density, products, coupling, irreps, rc)

    def _set(self, z, radial, lmax, density, products, coupling, irreps, cutoff, resolved=None):
        self.atomic_numbers, self.radial, self.radial_spec = list(z), radial, list(radial.spec) if radial else []
        self.lmax, self.density, self.products = lmax, list(density), [tuple(p) for p in products]
        self.coupling, self.irreps, self.cutoff = [np.asarray(c, float) for c in coupling], irreps, float(cutoff)
        self.resolved = resolved

    @classmethod
    def from_tables(cls, radial, density, products, coupling, species, radial_spec, lmax, output_irreps, cutoff=None):
        """Explicit numerical tables (features.jl:55-82); density entries are
        (channel 0..nz, 1-based radial, 1-based harmonic) as in the reference,
        products 1-based density indices, one coupling block per irrep."""
        self = cls.__new__(cls)
        ir = feature_irreps(output_irreps)
        z = _atomic_numbers(species)
        if isinstance(lmax, bool) or lmax < 0:
            raise ArgumentError("lmax must be nonnegative")
        rs = [tuple(int(v) for v in s) for s in radial_spec]
        if not rs or any(n <= 0 or not 0 <= l <= lmax for n, l in rs) or rs != list(radial.spec):
            raise ArgumentError("invalid radial labels")
        d = [tuple(int(v) for v in e) for e in density]
        if not all(0 <= s <= len(z) and 1 <= n <= len(rs) and 1 <= m <= (lmax + 1) ** 2 for s, n, m in d):
            raise ArgumentError("invalid density indices")
        p = [tuple(int(v) for v in q) for q in products]
        if not all(q and all(1 <= a <= len(d) for a in q) for q in p):
            raise ArgumentError("invalid products")
        if len(coupling) != len(ir):
            raise DimensionMismatch("coupling block count")
        C = []
        for b, s in enumerate(ir):
            c = np.asarray(coupling[b].toarray() if hasattr(coupling[b], "toarray") else coupling[b], float)
            dim = 2 * s.L + 1
            if c.shape[1] != len(p) or c.shape[0] == 0 or c.shape[0] % dim:
                raise DimensionMismatch("incomplete coupling multiplets")
            if not np.all(np.isfinite(c)):
                raise ArgumentError("nonfinite coupling coefficients")
            if not np.any(c):
                raise ArgumentError("no basis channels survive")
            C.append(c)
        rc = radial.cutoff
        if cutoff is not None:
            if not (np.isfinite(cutoff) and cutoff >= rc):
                raise ArgumentError("cutoff must cover radial cutoffs")
            rc = float(cutoff)
        self._set(z, radial, int(lmax), [(s, n - 1, m - 1) for s, n, m in d], [tuple(a - 1 for a in q) for q in p],
                  C, ir, rc)
        return self

    @classmethod
    def from_resolved(cls, model, channels=1):
        """The resolved scalar adapter (features.jl:84-94): one L = 0 even block
        with the model's masked coupling; evaluated with the model's kernels."""
        if model.solid:
            raise ArgumentError("resolved feature adapter requires spherical harmonics")
        self = cls.__new__(cls)
        ir = feature_irreps(({"L": 0, "parity": "even", "channels": int(channels)},))
        self._set(model.atomic_numbers, None, model.lmax, [], [], [np.zeros((model.nbasis, 0))], ir, model.cutoff,
                  resolved=model)
        return self

    def basis_channels(self, b):
        return self.coupling[b].shape[0] // (2 * self.irreps[b].L + 1)


def _atomic_numbers(species):
    from .ace import ELEMENTS
    z = []
    for s in species:
        if isinstance(s, (int, np.integer)) and not isinstance(s, bool):
            z.append(int(s))
        elif isinstance(s, str) and s in ELEMENTS:
            z.append(ELEMENTS.index(s) + 1)
        else:
            raise ArgumentError("feature species must be elements")
    if not z or len(set(z)) != len(z):
        raise ArgumentError("species must be nonempty and unique")
    return z


def feature_basis_summary(layer):
    """Coupled basis multiplicities and readout rank limits (features.jl:97-100)."""
    out = []
    for b, s in enumerate(layer.irreps):
        n = layer.basis_channels(b)
        out.append(dict(L=s.L, parity=s.parity, channels=s.channels, basis_channels=n, rank_limit=min(n, s.channels)))
    return out


def feature_parameters(layer, custom_radial=None):
    """Owning parameters (features.jl:106-116): mixing per block (basis,
    channels, centre species) starting from the leading identity rows, radial
    scales ones (radial, centre, neighbour), and the radial family's own
    parameters (rate, amplitude)."""
    nz = len(layer.atomic_numbers)
    mixing = {}
    for b, s in enumerate(layer.irreps):
        n = layer.basis_channels(b)
        W = np.zeros((n, s.channels, nz))
        for z in range(nz):
            for c in range(s.channels):
                W[c % n, c, z] = 1.0
        mixing[f"block_{b + 1}"] = W
    return {"mixing": mixing, "radial": {"scale": np.ones((max(len(layer.radial_spec), 1), nz, nz))},
            "custom_radial": dict(custom_radial or {})}


def _check(layer, ps):
    if set(ps) != {"mixing", "radial", "custom_radial"}:
        raise ArgumentError("feature parameters must contain mixing, radial and custom_radial")
    if set(ps["mixing"]) != {f"block_{b + 1}" for b in range(len(layer.irreps))}:
        raise ArgumentError("mixing block names")
    nz = len(layer.atomic_numbers)
    for b, s in enumerate(layer.irreps):
        if np.shape(ps["mixing"][f"block_{b + 1}"]) != (layer.basis_channels(b), s.channels, nz):
            raise DimensionMismatch(f"mixing block {b + 1}")


def _prepared(layer, system, ps, workspace):
    """Checks the parameters, sets the system on the workspace at the layer's
    cutoff and packs the layer's device tables."""
    from .neighbourlist import Workspace
    w = workspace or Workspace(precision="float64")
    nz = len(layer.atomic_numbers)
    ch = _channels(layer, system)
    w.prepare(system, layer.cutoff, ch)
    rate = np.asarray(ps["custom_radial"].get("rate", [np.nan]), float).reshape(-1)
    amp = np.asarray(ps["custom_radial"].get("amplitude", [np.nan]), float).reshape(-1)
    if rate.size != 1 or amp.size != 1 or not (np.isfinite(rate[0]) and np.isfinite(amp[0])):
        raise ArgumentError("smooth exponential radials need finite custom_radial rate and amplitude")
    scale = np.asarray(ps["radial"]["scale"], float)
    nr = len(layer.radial_spec)
    if scale.shape != (nr, nz, nz):
        raise DimensionMismatch("radial scale shape")
    st, keep = _struct(layer, ps, rate[0], amp[0], scale)
    return w, st, keep


def ace_features(layer, system, ps, workspace=None):
    """``ace_features(X, layer, ps, environment, H)`` (features.jl:164-178) on
    the device: one (2L+1, channels, n) array per irrep, input atom order
    (column-major views of one buffer, as the reference's Julia arrays)."""
    _check(layer, ps)
    if layer.resolved is not None:
        return _resolved_features(layer, system, ps)
    w, st, keep = _prepared(layer, system, ps, workspace)
    n = len(system)
    sizes = [(2 * s.L + 1) * s.channels * n for s in layer.irreps]
    out = pinned_empty(sum(sizes))  # filled completely by the call
    w.ctx.call("pfw_features", ctypes.byref(st), ptr(out))
    del keep
    res, o = [], 0
    for s, k in zip(layer.irreps, sizes):
        res.append(out[o:o + k].reshape((2 * s.L + 1, s.channels, n), order="F"))  # column-major, no copy
        o += k
    return tuple(res)


def feature_vjp(layer, system, ps, cotangents, workspace=None):
    """``feature_vjp(X, layer, ps, environment, H, cotangents)`` (features.jl:204-225)
    on the device: the gradient of sum_b <cotangents[b], features[b]> with
    respect to the positions ((n, 3), input order) and the parameters (the
    structure of ``ps``: mixing blocks, radial scale, custom_radial rate and
    amplitude). A cotangent may be None (a zero block)."""
    _check(layer, ps)
    if layer.resolved is not None:
        raise ArgumentError("the feature pullback needs a layer with smooth exponential radials")
    n = len(system)
    blocks = []
    for s, c in zip(layer.irreps, cotangents):
        shape = (2 * s.L + 1, s.channels, n)
        c = np.zeros(shape) if c is None else np.asarray(c, float)
        if c.shape != shape:
            raise DimensionMismatch(f"cotangent of block L={s.L} has shape {c.shape}, expected {shape}")
        blocks.append(c.reshape(-1, order="F"))
    if len(blocks) != len(layer.irreps):
        raise DimensionMismatch("one cotangent per feature block")
    cot = pinned_empty(sum(b.size for b in blocks))  # the upload is then one DMA
    if cot.size:
        np.concatenate(blocks, out=cot)
    w, st, keep = _prepared(layer, system, ps, workspace)
    nz, nr = len(layer.atomic_numbers), len(layer.radial_spec)
    dx = pinned_empty(3 * n)  # filled completely by the call
    dmix = np.zeros(sum(layer.basis_channels(b) * s.channels * nz for b, s in enumerate(layer.irreps)))
    dscale = np.zeros(nr * nz * nz)
    drad = np.zeros(2)
    w.ctx.call("pfw_features_vjp", ctypes.byref(st), ptr(cot), ptr(dx), ptr(dmix), ptr(dscale), ptr(drad))
    del keep
    mixing, o = {}, 0
    for b, s in enumerate(layer.irreps):
        nb = layer.basis_channels(b)
        k = nb * s.channels * nz
        mixing[f"block_{b + 1}"] = dmix[o:o + k].reshape(nz, s.channels, nb).transpose(2, 1, 0).copy()
        o += k
    custom = {}
    for key, val in zip(("rate", "amplitude"), drad):
        if key in ps["custom_radial"]:
            custom[key] = np.full(np.shape(ps["custom_radial"][key]), val)
    return {"positions": dx.reshape(n, 3),
            "parameters": {"mixing": mixing, "radial": {"scale": dscale.reshape(nz, nz, nr).transpose(2, 1, 0).copy()},
                           "custom_radial": custom}}


def ace_features_rrule(layer, system, ps, workspace=None):
    """``rrule(ace_features, ...)`` (ext/PotentialFrameWorkChainRulesCoreExt.jl:77-98): the features and a
    pullback over the saved positions and parameters (copies, so that later
    changes by the caller do not alter it). The pullback takes one cotangent
    per block (None for a zero block) and returns the ``feature_vjp`` dict."""
    import copy
    y = ace_features(layer, system, ps, workspace)
    saved_sys, saved_ps = copy.deepcopy(system), copy.deepcopy(ps)

    def pullback(cotangents):
        if cotangents is None:
            return None
        return feature_vjp(layer, saved_sys, saved_ps, cotangents, workspace)

    return y, pullback


def _channels(layer, system):
    from .ace import ELEMENTS
    from .system import map_labels
    lut = {ELEMENTS[z - 1]: c for c, z in enumerate(layer.atomic_numbers)}
    out, bad = map_labels(system.labels, lambda s: lut.get(s) if isinstance(s, str) else None)
    if out is None:
        raise ArgumentError(f"species {bad} is not in the feature layer")
    return out


class _Layer(ctypes.Structure):
    _fields_ = [("nz", ctypes.c_int32), ("nr", ctypes.c_int32), ("lmax", ctypes.c_int32),
                ("ndens", ctypes.c_int32), ("nprod", ctypes.c_int32), ("nblocks", ctypes.c_int32),
                ("dens_z", ctypes.c_void_p), ("dens_n", ctypes.c_void_p), ("dens_h", ctypes.c_void_p),
                ("prod_ptr", ctypes.c_void_p), ("prod_idx", ctypes.c_void_p),
                ("block_L", ctypes.c_void_p), ("block_channels", ctypes.c_void_p),
                ("block_nbasis", ctypes.c_void_p), ("row_ptr", ctypes.c_void_p), ("col", ctypes.c_void_p),
                ("val", ctypes.c_void_p), ("mixing", ctypes.c_void_p), ("cutoff", ctypes.c_double),
                ("radial_cutoff", ctypes.c_double), ("radial_kind", ctypes.c_int32), ("powers", ctypes.c_void_p), ("rate", ctypes.c_double),
                ("amplitude", ctypes.c_double), ("scale", ctypes.c_void_p)]


def _tables(layer):
    """The layer's flat device tables (0-based density columns, product and
    coupling CSR over all blocks), built once per layer."""
    t = getattr(layer, "_flat", None)
    if t is not None:
        return t
    d = np.asarray(layer.density, np.int32).reshape(-1, 3)
    pptr = np.zeros(len(layer.products) + 1, np.int32)
    pptr[1:] = np.cumsum([len(p) for p in layer.products])
    pidx = np.concatenate([np.asarray(p, np.int32) for p in layer.products]) if layer.products else np.zeros(0, np.int32)
    C = np.vstack(layer.coupling)
    r, c = np.nonzero(C)  # row-major: rows in turn, columns ascending
    rows = np.zeros(C.shape[0] + 1, np.int32)
    rows[1:] = np.cumsum(np.bincount(r, minlength=C.shape[0]))
    t = dict(dz=np.ascontiguousarray(d[:, 0]), dn=np.ascontiguousarray(d[:, 1]), dh=np.ascontiguousarray(d[:, 2]),
             pptr=pptr, pidx=pidx.astype(np.int32), rows=rows, cols=c.astype(np.int32),
             vals=np.ascontiguousarray(C[r, c], np.float64),
             L=np.array([s.L for s in layer.irreps], np.int32),
             ch=np.array([s.channels for s in layer.irreps], np.int32),
             nb=np.array([layer.basis_channels(b) for b in range(len(layer.irreps))], np.int32),
             powers=np.asarray(layer.radial.powers, np.int32))
    layer._flat = t
    return t


def _struct(layer, ps, rate, amplitude, scale):
    t = _tables(layer)
    mix = np.concatenate([np.asarray(ps["mixing"][f"block_{b + 1}"], float).transpose(2, 1, 0).reshape(-1)
                          for b in range(len(layer.irreps))])  # column-major blocks in turn
    sc = np.ascontiguousarray(scale.transpose(2, 1, 0).reshape(-1), np.float64)
    keep = [t, mix, sc]
    st = _Layer(len(layer.atomic_numbers), len(layer.radial_spec), layer.lmax, len(t["dz"]), len(layer.products),
                len(layer.irreps), t["dz"].ctypes.data, t["dn"].ctypes.data, t["dh"].ctypes.data,
                t["pptr"].ctypes.data, t["pidx"].ctypes.data, t["L"].ctypes.data, t["ch"].ctypes.data,
                t["nb"].ctypes.data, t["rows"].ctypes.data, t["cols"].ctypes.data, t["vals"].ctypes.data,
                mix.ctypes.data, layer.cutoff, layer.radial.cutoff, PFW_RADIAL_SMOOTH_EXPONENTIAL,
                t["powers"].ctypes.data, float(rate), float(amplitude), sc.ctypes.data)
    return st, keep


def _resolved_features(layer, system, ps):
    """The L = 0 adapter: channel c of atom i is sum_b B_ib W[b, c, z_i], the
    site energy of the resolved model with readout W[:, c, :], no pair term
    and no reference energy; one device step per channel."""
    from .ace import ACECalculator
    m = layer.resolved
    W = np.asarray(ps["mixing"]["block_1"], float)
    n = len(system)
    out = np.zeros((1, W.shape[1], n))
    calc = ACECalculator(m, precision="float64")
    ch = m.channels(system.labels)
    ref = np.asarray(m.reference, float)[ch] if m.reference_enabled else np.zeros(n)
    for c in range(W.shape[1]):
        calc.set_parameters({"WB": W[:, c, :], "Wpair": np.zeros_like(m.Wpair)})
        r = calc.evaluate(system, request=("energy",), site_energies=True)
        out[0, c] = r["site_energies"] - ref
    return (out,)
