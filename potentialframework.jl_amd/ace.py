"""ACE on the GPU: resolved-model loading and the calculator boundary.

Mirrors src/submodules/ACE/resolved.jl:57-142 (``load_resolved`` / ``read_resolved``
for schema v1, with the basis policy of src/submodules/ACE/interface.jl:1-125),
src/submodules/ACE/environments.jl:1-10
(strict element channels) and src/submodules/ACE/calculators.jl (``ACECalculator``
with ``energy``/``forces``/``virial`` results in original atom order).
"""
import ctypes

import numpy as np

from . import _capi
from ._capi import ArgumentError, DimensionMismatch, pinned_empty, ptr
from .neighbourlist import Workspace
from .pair import ENERGY, FORCES, VIRIAL, request_mask

try:
    import tomllib as _toml
except ImportError:  # Python 3.10
    import tomli as _toml

ENVELOPES = {"PolyEnvelope2sX": 0, "PolyEnvelope1sR": 1, "ACE1_PolyEnvelope1sR": 2}
ELEMENTS = ["H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar",
            "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr",
            "Rb", "Sr", "Y", "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I", "Xe",
            "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu",
            "Hf", "Ta", "W", "Re", "Os", "Ir", "Pt", "Au", "Hg"]


def atomic_number(label):
    """Element labels only: isotopes or custom names are rejected (environments.jl:1-10)."""
    if isinstance(label, (int, np.integer)):
        return int(label)
    try:
        return ELEMENTS.index(label) + 1
    except ValueError:
        raise ArgumentError(f"ACE element channels do not accept isotopes or named species: {label}") from None


class RadialFamily:
    def __init__(self, D, nz):
        if D.get("kind") != "piecewise_cubic_agnesi":
            raise ArgumentError("unsupported resolved radial kind")
        self.spec = [(int(b["n"]), int(b["l"])) for b in D["spec"]]
        self.nr = len(self.spec)
        seen = {}
        for ch in D["channels"]:
            key = (int(ch["centre"]) - 1, int(ch["neighbour"]) - 1)
            if not (0 <= key[0] < nz and 0 <= key[1] < nz) or key in seen:
                raise ArgumentError("invalid/duplicate radial species channel")
            seen[key] = ch
        if len(seen) != nz * nz:
            raise ArgumentError("missing radial species channels")
        nn = len(next(iter(seen.values()))["nodes"]) if seen else 2
        self.nnodes = nn
        self.transform = np.zeros((nz * nz, 8))
        self.envelope_kind = np.zeros(nz * nz, np.int32)
        self.envelope = np.zeros((nz * nz, 5))
        self.nodes = np.zeros((nz * nz, nn))
        self.coefficients = np.zeros((nz * nz, 4 * self.nr * (nn - 1)))
        for (i, j), ch in seen.items():
            c = i + nz * j
            t = ch["transform"]
            if t.get("normalization_rin", t["rin"]) != t["rin"]:
                raise ArgumentError("unsupported differing transform origins")
            self.transform[c] = [t["p"], t["q"], t["a"], t["rin"], t["r0"], t["yin"], t["ycut"], t["rcut"]]
            kind = ch["envelope_kind"]
            if kind not in ENVELOPES:
                raise ArgumentError(f"unsupported resolved envelope {kind}")
            e = ch["envelope"]
            self.envelope_kind[c] = ENVELOPES[kind]
            if kind == "PolyEnvelope2sX":
                self.envelope[c] = [e["x1"], e["x2"], e["p1"], e["p2"], e["s"]]
            elif kind == "PolyEnvelope1sR":
                self.envelope[c, :2] = [e["rcut"], e["p"]]
            else:
                self.envelope[c, :3] = [e["rcut"], e["r0"], e["p"]]
            nodes = np.asarray(ch["nodes"], float)
            if len(nodes) != nn or nodes[0] != -1 or nodes[-1] != 1 or np.any(np.diff(nodes) <= 0):
                raise ArgumentError("invalid spline nodes")
            if list(ch["coefficient_shape"]) != [4, self.nr, nn - 1]:
                raise DimensionMismatch("invalid spline coefficients")
            coef = np.asarray(ch["coefficients"], float)
            if not np.all(np.isfinite(coef)):
                raise ArgumentError("nonfinite spline coefficients")
            self.nodes[c] = nodes
            self.coefficients[c] = coef
        self.cutoff = float(self.transform[:, 7].max()) if nz else 0.0


class ResolvedModel:
    """Device-ready resolved ACE model (schema v1 with its basis policy)."""

    def __init__(self, D):
        if D.get("schema") != "PotentialFrameWork.resolved_ACE" or D.get("schema_version") != 1:
            raise ArgumentError("unsupported resolved ACE schema")
        self.atomic_numbers = [int(v) for v in D["species_atomic_numbers"]]
        nz = len(self.atomic_numbers)
        if len(set(self.atomic_numbers)) != nz or min(self.atomic_numbers) <= 0:
            raise ArgumentError("invalid species channels")
        ang = D["angular"]
        if ang["kind"] not in ("real_spherical", "real_solid") or ang["normalization"] != "L2" \
                or ang["ordering"] != "l*l+l+m+1":
            raise ArgumentError("unsupported angular convention")  # resolved.jl:109-110
        self.solid = ang["kind"] == "real_solid"
        self.nz = nz
        self.lmax = int(ang["lmax"])
        self.radial = RadialFamily(D["radial"], nz)
        self.pair = RadialFamily(D["pair_radial"], nz)
        self.density = np.array([[int(a) - 1, int(b) - 1] for a, b in D["density_indices"]], np.int32).reshape(-1, 2)
        self.products = [[int(v) - 1 for v in p] for p in D["products"]]
        c = D["coupling"]
        self.nbasis, nprod = [int(v) for v in c["shape"]]
        if nprod != len(self.products):
            raise DimensionMismatch("coupling/product dimensions")
        self.c_rows = np.asarray(c["rows"], np.int32) - 1
        self.c_cols = np.asarray(c["columns"], np.int32) - 1
        self.c_vals = np.asarray(c["values"], float)
        p = D["parameters"]
        if p["ordering"] != "vec(WB), vec(Wpair); Julia column-major":
            raise ArgumentError("unsupported parameter ordering")
        flat = np.asarray(p["flat"], float)
        nq = self.pair.nr
        if len(flat) != (self.nbasis + nq) * nz:
            raise DimensionMismatch("parameter length")
        self.WB = flat[: self.nbasis * nz].reshape(nz, self.nbasis).T.copy()
        self.Wpair = flat[self.nbasis * nz:].reshape(nz, nq).T.copy()
        self.reference = np.asarray(D["reference_eV"], float)
        if len(self.reference) != nz:
            raise DimensionMismatch("reference energy channels")
        self.cutoff = max(self.radial.cutoff, self.pair.cutoff)
        self.source_sha256 = D.get("source_sha256")
        self._read_policy(D.get("interface_policy"))
        self.source = {k: v for k, v in D.items() if k != "parameters"}  # kept for save_calculator

    def _read_policy(self, P):
        """The basis policy (interface.jl:2-18, interface_persistence.jl:3-25): the
        invariant mask (effective_WB = WB .* mask), the pair and reference flags
        (pair_allowed / reference_value, interface.jl:122-124), and the species
        projection with its orders (project_species, interface.jl:115-121):
        ``species`` holds 0-based channels, ``orders`` is None or ``(lo, hi)``."""
        self.mask = np.ones(self.nbasis, bool)
        self.pair_enabled = True
        self.reference_enabled = True
        self.species = ()
        self.orders = None
        if P is None:
            return
        if P.get("version") != 1:
            raise ArgumentError("unsupported basis policy version")
        mask = P.get("mask", [True] * self.nbasis)
        if len(mask) != self.nbasis or not all(isinstance(v, bool) for v in mask):
            raise ArgumentError("invalid invariant mask")
        sp = [int(v) for v in P.get("species", [])]
        if not all(1 <= v <= self.nz for v in sp) or len(set(sp)) != len(sp):
            raise ArgumentError("invalid projected species")
        o = [int(v) for v in P.get("orders", [])]
        if o and not (len(o) == 2 and 0 <= o[0] <= o[1]):
            raise ArgumentError("invalid projected orders")
        self.mask = np.asarray(mask, bool)
        self.pair_enabled = bool(P.get("pair", True))
        self.reference_enabled = bool(P.get("reference", True))
        self.species = tuple(sorted(v - 1 for v in sp))
        self.orders = (o[0], o[1]) if o else None

    def order_allowed(self, k):
        """order_allowed (interface.jl:122)."""
        return self.orders is None or self.orders[0] <= k <= self.orders[1]

    def effective(self, WB, Wpair):
        """The readout the device evaluates: WB .* mask, and Wpair only when the
        pair term is enabled."""
        return WB * self.mask[:, None], (Wpair if self.pair_enabled else np.zeros_like(Wpair))

    def parameters(self):
        return {"WB": self.WB.copy(), "Wpair": self.Wpair.copy()}

    def channels(self, labels):
        from .system import map_labels

        def lookup(s):
            z = atomic_number(s)
            return self.atomic_numbers.index(z) if z in self.atomic_numbers else None

        out, bad = map_labels(labels, lookup)
        if out is None:
            raise ArgumentError("ACE model has no element channel for a species in the system")
        return out

    def _struct(self):
        keep = []

        def a(x, dt):
            x = np.ascontiguousarray(np.asarray(x, dtype=dt))
            keep.append(x)
            return x.ctypes.data

        def fam(f):
            return _Radial(f.nr, f.nnodes, a(f.transform, np.float64), a(f.envelope_kind, np.int32),
                           a(f.envelope, np.float64), a(f.nodes, np.float64), a(f.coefficients, np.float64))

        pp = np.zeros(len(self.products) + 1, np.int32)
        pp[1:] = np.cumsum([len(p) for p in self.products])
        pidx = np.concatenate([np.asarray(p, np.int32) for p in self.products]) if self.products else np.zeros(0)
        WB, Wpair = self.effective(self.WB, self.Wpair)
        # reference_value (interface.jl:124): the reference energy is order 0
        ref = self.reference if (self.reference_enabled and self.order_allowed(0)) else np.zeros_like(self.reference)
        sel = np.zeros(self.nz, np.int32)
        sel[list(self.species)] = 1
        lo, hi = self.orders if self.orders is not None else (-1, -1)
        m = _Model(self.nz, self.lmax, len(self.density), len(self.products), self.nbasis,
                   a(ref, np.float64), fam(self.radial), fam(self.pair), a(self.density, np.int32),
                   a(pp, np.int32), a(pidx, np.int32), len(self.c_vals), a(self.c_rows, np.int32),
                   a(self.c_cols, np.int32), a(self.c_vals, np.float64), a(WB.T, np.float64),
                   a(Wpair.T, np.float64), 1 if self.solid else 0, a(sel, np.int32), lo, hi)
        return m, keep


class _Radial(ctypes.Structure):
    _fields_ = [("nr", ctypes.c_int32), ("nnodes", ctypes.c_int32), ("transform", ctypes.c_void_p),
                ("envelope_kind", ctypes.c_void_p), ("envelope", ctypes.c_void_p), ("nodes", ctypes.c_void_p),
                ("coefficients", ctypes.c_void_p)]


class _Model(ctypes.Structure):
    _fields_ = [("nz", ctypes.c_int32), ("lmax", ctypes.c_int32), ("ndens", ctypes.c_int32),
                ("nprod", ctypes.c_int32), ("nbasis", ctypes.c_int32), ("reference", ctypes.c_void_p),
                ("radial", _Radial), ("pair", _Radial), ("density", ctypes.c_void_p),
                ("product_ptr", ctypes.c_void_p), ("product_idx", ctypes.c_void_p), ("coupling_nnz", ctypes.c_int64),
                ("coupling_rows", ctypes.c_void_p), ("coupling_cols", ctypes.c_void_p),
                ("coupling_vals", ctypes.c_void_p), ("WB", ctypes.c_void_p), ("Wpair", ctypes.c_void_p),
                ("angular", ctypes.c_int32), ("selected", ctypes.c_void_p), ("order_lo", ctypes.c_int32),
                ("order_hi", ctypes.c_int32)]


def project_species(model, species, orders):
    """``project_species(m, species, orders)`` (interface.jl:115-121): a copy of
    the model whose products keep only the orders ``orders`` (a ``range`` with
    step 1, nonnegative, nonempty) of prod(U + t V), V the densities of
    neighbours of the given species (an element label, atomic number, or a
    list of them) and U those of the others; the pair term of a projected
    neighbour counts as order 1, otherwise 0, and the reference energy as
    order 0 (order_allowed / pair_allowed / reference_value, :122-124)."""
    import copy
    if not isinstance(orders, range) or orders.step != 1:
        raise ArgumentError("orders must be a nonnegative inclusive unit range")
    if len(orders) == 0 or orders.start < 0:
        raise ArgumentError("orders must be a nonnegative inclusive range")
    group = (species,) if isinstance(species, (str, int, np.integer)) else tuple(species)
    chans = []
    for s_ in group:
        z = atomic_number(s_)
        if z not in model.atomic_numbers:
            raise ArgumentError(f"ACE model has no element channel for species {s_}")
        chans.append(model.atomic_numbers.index(z))
    out = copy.copy(model)
    out.species = tuple(sorted(set(chans)))
    out.orders = (orders.start, orders[-1])
    out.source = dict(model.source)
    pol = dict(out.source.get("interface_policy") or {}, version=1)
    pol["species"] = [c + 1 for c in out.species]
    pol["orders"] = [out.orders[0], out.orders[1]]
    out.source["interface_policy"] = pol
    return out


def load_resolved(source):
    """``load_resolved(path)`` (resolved.jl:94-102); also accepts the parsed dictionary
    or the binary form written by ``save_resolved_npz``. Returns ``(model, parameters, state)``."""
    if isinstance(source, dict):
        D = source
    elif str(source).endswith(".npz"):
        D = read_resolved_npz(source)
    else:
        with open(source, "rb") as f:
            D = _toml.load(f)
    if D.get("schema") == CHECKPOINT_SCHEMA:
        return _read_checkpoint(D)
    m = ResolvedModel(D)
    return m, m.parameters(), {}


# ---- calculator checkpoints (src/submodules/ACE/persistence.jl): a versioned
# TOML tree {schema, schema_version, model (the resolved dictionary without its
# parameters), parameters and state in the checkpoint encoding}.
CHECKPOINT_SCHEMA = "PotentialFrameWork.ACECalculator"
_CK_TYPES = {"Float64": np.float64, "Float32": np.float32, "Int64": np.int64, "Int32": np.int32,
             "Int16": np.int16, "Int8": np.int8, "UInt8": np.uint8, "UInt16": np.uint16, "UInt32": np.uint32,
             "Bool": np.bool_}


def checkpoint_decode(d):
    """persistence.jl:24-52: nested named tuples, tuples, strings, nothing, and
    column-major numeric arrays / scalars of the listed element types."""
    kind = d["kind"]
    if kind == "namedtuple":
        names, items = d["names"], d["items"]
        if len(names) != len(items) or len(set(names)) != len(names):
            raise ArgumentError("invalid checkpoint field names")
        return {n: checkpoint_decode(v) for n, v in zip(names, items)}
    if kind == "tuple":
        return tuple(checkpoint_decode(v) for v in d["items"])
    if kind == "nothing":
        return None
    if kind == "string":
        return str(d["value"])
    if kind in ("array", "scalar"):
        T = _CK_TYPES.get(d["eltype"])
        if T is None:
            raise ArgumentError("unsupported checkpoint numeric type")
        if kind == "scalar":
            v = T(d["value"])
            if not np.all(np.isfinite(v)):
                raise ArgumentError("nonfinite checkpoint value")
            return v
        shape = tuple(int(v) for v in d["shape"])
        vals = np.asarray(d["values"], T)
        if any(v < 0 for v in shape) or int(np.prod(shape)) != vals.size:
            raise DimensionMismatch("checkpoint array shape")
        if not np.all(np.isfinite(vals)):
            raise ArgumentError("nonfinite checkpoint array")
        return vals.reshape(shape[::-1]).T if len(shape) > 1 else vals.reshape(shape)
    raise ArgumentError(f"unsupported checkpoint value kind {kind}")


def checkpoint_encode(x):
    """persistence.jl:6-22 (Julia type names, column-major values)."""
    if isinstance(x, dict):
        return {"kind": "namedtuple", "names": list(x), "items": [checkpoint_encode(v) for v in x.values()]}
    if isinstance(x, tuple):
        return {"kind": "tuple", "items": [checkpoint_encode(v) for v in x]}
    if x is None:
        return {"kind": "nothing"}
    if isinstance(x, str):
        return {"kind": "string", "value": x}
    a = np.asarray(x)
    name = {np.dtype(v): k for k, v in _CK_TYPES.items()}.get(a.dtype)
    if name is None:
        raise ArgumentError(f"unsupported checkpoint numeric type {a.dtype}")
    if not np.all(np.isfinite(a)):
        raise ArgumentError("nonfinite checkpoint values")
    if a.ndim == 0:
        return {"kind": "scalar", "eltype": name, "value": a.item()}
    return {"kind": "array", "eltype": name, "shape": list(a.shape), "values": a.T.reshape(-1).tolist()}


def _read_checkpoint(d):
    """ACECalculator(path) for a checkpoint (persistence.jl:83-102, 116-129)."""
    if d.get("schema_version") not in (1, 2):
        raise ArgumentError("unsupported ACECalculator checkpoint version")
    ps = checkpoint_decode(d["parameters"])
    st = checkpoint_decode(d["state"])
    if not (isinstance(ps, dict) and "WB" in ps and "Wpair" in ps):
        raise ArgumentError("checkpoint parameters require WB and Wpair")
    if set(ps) - {"WB", "Wpair"}:
        raise ArgumentError("checkpoints with radial scale parameters are not supported on the device")
    WB, Wpair = np.asarray(ps["WB"], float), np.asarray(ps["Wpair"], float)
    D = dict(d["model"])
    nz = len(D["species_atomic_numbers"])
    if WB.shape != (int(D["coupling"]["shape"][0]), nz) or Wpair.shape != (len(D["pair_radial"]["spec"]), nz):
        raise DimensionMismatch("checkpoint readout dimensions")
    D["parameters"] = {"WB_shape": list(WB.shape), "Wpair_shape": list(Wpair.shape),
                       "ordering": "vec(WB), vec(Wpair); Julia column-major",
                       "flat": np.concatenate([WB.T.reshape(-1), Wpair.T.reshape(-1)])}
    m = ResolvedModel(D)
    return m, m.parameters(), st if st is not None else {}


def _toml_value(v):
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        if not np.isfinite(v):
            raise ArgumentError("nonfinite value in a checkpoint")
        return repr(float(v))
    if isinstance(v, str):
        return '"' + v.replace("\\", "\\\\").replace('"', '\\"') + '"'
    if isinstance(v, np.ndarray):
        v = v.tolist()
    if isinstance(v, (list, tuple)):
        return "[" + ", ".join(_toml_value(x) for x in v) + "]"
    if isinstance(v, dict):
        return "{" + ", ".join(f"{_toml_key(k)} = {_toml_value(x)}" for k, x in v.items()) + "}"
    raise ArgumentError(f"unsupported checkpoint value {type(v)}")


def _toml_key(k):
    k = str(k)
    return k if k.replace("_", "").replace("-", "").isalnum() else '"' + k + '"'


def _toml_table(d, path, out):
    tables = []
    for k, v in d.items():
        if isinstance(v, dict):
            tables.append((k, v, False))
        elif isinstance(v, list) and v and all(isinstance(x, dict) for x in v):
            tables.append((k, v, True))
        else:
            out.append(f"{_toml_key(k)} = {_toml_value(v)}")
    for k, v, many in tables:
        sub = path + [_toml_key(k)]
        if many:
            for item in v:
                out.append("")
                out.append("[[" + ".".join(sub) + "]]")
                _toml_table(item, sub, out)
        else:
            out.append("")
            out.append("[" + ".".join(sub) + "]")
            _toml_table(v, sub, out)


def save_calculator(path, calc, parameters=None, state=None):
    """``save_calculator(path, calc; parameters, state)`` (persistence.jl:147-167):
    a self-contained versioned TOML checkpoint of the resolved basis, explicit
    parameters and state; written to a temporary file and moved into place."""
    import os
    import tempfile
    m = calc.model
    ps = parameters if parameters is not None else calc.parameters
    WB, Wpair = np.asarray(ps["WB"], np.float64), np.asarray(ps["Wpair"], np.float64)
    if WB.shape != m.WB.shape or Wpair.shape != m.Wpair.shape:
        raise DimensionMismatch("checkpoint readout dimensions")
    model = dict(m.source)
    model["interface_policy"] = dict(model.get("interface_policy") or {}, version=1,
                                     mask=[bool(v) for v in m.mask], pair=m.pair_enabled,
                                     reference=m.reference_enabled)
    pol = model["interface_policy"]
    pol["species"] = [c + 1 for c in m.species]
    pol["orders"] = list(m.orders) if m.orders is not None else []
    # provenance maps of a default policy (interface_persistence.jl:8-25 validates them)
    pol.setdefault("rows", list(range(1, m.nbasis + 1)))
    pol.setdefault("products", list(range(1, len(m.products) + 1)))
    pol.setdefault("density", list(range(1, len(m.density) + 1)))
    pol.setdefault("radial", list(range(1, m.radial.nr + 1)))
    if len(str(pol.get("origin", ""))) != 64:
        import hashlib
        pol["origin"] = hashlib.sha256(repr((m.source_sha256, m.radial.spec, m.density.tolist(), m.products,
                                             m.c_vals.tolist())).encode()).hexdigest()
    d = {"schema": CHECKPOINT_SCHEMA, "schema_version": 2, "writer": "potentialframework.jl_amd", "model": model,
         "parameters": checkpoint_encode({"WB": WB, "Wpair": Wpair}),
         "state": checkpoint_encode(state if state is not None else {})}
    out = []
    _toml_table(d, [], out)
    dest = os.path.abspath(path)
    fd, tmp = tempfile.mkstemp(dir=os.path.dirname(dest))
    try:
        with os.fdopen(fd, "w") as f:
            f.write("\n".join(out) + "\n")
        os.replace(tmp, dest)
    finally:
        if os.path.exists(tmp):
            os.remove(tmp)
    return path


# Binary form of a schema-v1 resolved model: the same dictionary, with every
# large numeric list stored as a float64 / int64 array and the rest as JSON,
# in one compressed .npz (no pickled objects). A 48 MB TOML export of a
# 114-radial, 5,652-product model is ~3.4 MB here and loads in well under a second.
_FAMILY_ARRAYS = ("nodes", "coefficients")


def save_resolved_npz(D, path):
    import json
    arrays, meta = {}, {}
    for key, val in D.items():
        if key in ("radial", "pair_radial"):
            fam = {k: v for k, v in val.items() if k != "channels"}
            chans = []
            for k, ch in enumerate(val["channels"]):
                chans.append({a: b for a, b in ch.items() if a not in _FAMILY_ARRAYS})
                for a in _FAMILY_ARRAYS:
                    arrays[f"{key}.{k}.{a}"] = np.asarray(ch[a], np.float64)
            fam["channels"] = chans
            meta[key] = fam
        elif key == "products":
            arrays["products.ptr"] = np.cumsum([0] + [len(p) for p in val]).astype(np.int64)
            arrays["products.idx"] = np.asarray([v for p in val for v in p], np.int64)
        elif key == "density_indices":
            arrays[key] = np.asarray(val, np.int64).reshape(-1, 2)
        elif key == "coupling":
            meta[key] = {k: v for k, v in val.items() if k not in ("rows", "columns", "values")}
            arrays["coupling.rows"] = np.asarray(val["rows"], np.int64)
            arrays["coupling.columns"] = np.asarray(val["columns"], np.int64)
            arrays["coupling.values"] = np.asarray(val["values"], np.float64)
        elif key == "parameters":
            meta[key] = {k: v for k, v in val.items() if k != "flat"}
            arrays["parameters.flat"] = np.asarray(val["flat"], np.float64)
        else:
            meta[key] = val
    arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), np.uint8)
    np.savez_compressed(path, **arrays)


def read_resolved_npz(path):
    import json
    with np.load(path, allow_pickle=False) as z:
        D = json.loads(bytes(z["meta"]).decode())
        for key in ("radial", "pair_radial"):
            for k, ch in enumerate(D[key]["channels"]):
                for a in _FAMILY_ARRAYS:
                    ch[a] = z[f"{key}.{k}.{a}"]
        ptr, idx = z["products.ptr"], z["products.idx"]
        D["products"] = [idx[ptr[k]:ptr[k + 1]].tolist() for k in range(len(ptr) - 1)]
        D["density_indices"] = z["density_indices"].tolist()
        D["coupling"].update(rows=z["coupling.rows"], columns=z["coupling.columns"], values=z["coupling.values"])
        D["parameters"]["flat"] = z["parameters.flat"]
    return D


class ACECalculator:
    """``ACECalculator(model, ps; precision)`` (calculators.jl:6-29) on one device context.
    ``skin`` (Å) keeps the environments at cutoff + skin between evaluations of
    the same system with moved atoms (GPUExecutor with ``GPUConfig(skin)``,
    src/submodules/ACE/gpu.jl:301-343)."""

    def __init__(self, model, parameters=None, precision="float64", device=0, workspace=None, skin=0.0):
        self.model = model
        self.workspace = workspace or Workspace(precision=precision, device=device, skin=skin)
        self.precision = self.workspace.precision
        st, keep = model._struct()
        self.workspace.ctx.call("pfw_set_ace_model", ctypes.byref(st))
        del keep
        self.parameters = model.parameters()
        if parameters is not None:
            self.set_parameters(parameters)

    def cutoff(self):
        return self.model.cutoff

    def set_parameters(self, ps):
        """``update_prepared!``: new WB/Wpair, precontraction refreshed on the host."""
        WB = np.asarray(ps["WB"], float)
        Wp = np.asarray(ps.get("Wpair", self.model.Wpair), float)
        if WB.shape != self.model.WB.shape or Wp.shape != self.model.Wpair.shape:
            raise DimensionMismatch("readout shape")
        eWB, eWp = self.model.effective(WB, Wp)
        self.workspace.ctx.call("pfw_set_ace_parameters", ptr(np.ascontiguousarray(eWB.T)),
                                ptr(np.ascontiguousarray(eWp.T)))
        self.parameters = {"WB": WB.copy(), "Wpair": Wp.copy()}

    def evaluate(self, sys, request=("energy", "forces", "virial"), site_energies=False):
        mask, names = request_mask(request)
        ch = self.model.channels(sys.labels)
        w = self.workspace
        w.prepare(sys, self.model.cutoff, ch)
        n = len(sys)
        E = np.zeros(1)
        F = pinned_empty(3 * n).reshape(n, 3)  # written whole when forces are requested
        W = np.zeros(9)
        w.ctx.call("pfw_ace_evaluate", mask, w.mode, ptr(E), ptr(F), ptr(W))
        if w.timing:
            w.read_timings()
        out = {}
        if ENERGY in names:
            out["energy"] = float(E[0])
        if FORCES in names:
            out["forces"] = F.T  # 3 x n, an owning view (nothing else holds F)
        if VIRIAL in names:
            out["virial"] = W.reshape(3, 3).T.copy()
        if site_energies:
            s = np.zeros(n)
            w.ctx.call("pfw_ace_site_energies", ptr(s))
            out["site_energies"] = s
        return out


def _evaluate_batch(self, systems, request=("energy", "forces", "virial")):
    """All configurations of a fitting set in one packed cell list, one launch
    per stage; one result dictionary per configuration."""
    mask, names = request_mask(request)
    w = self.workspace
    offsets = w.prepare_batch(systems, self.model.cutoff, [self.model.channels(s.labels) for s in systems])
    w.ctx.call("pfw_ace_step", mask, w.mode)
    if w.timing:
        w.read_timings()
    return w.read_batch(offsets, names)


ACECalculator.evaluate_batch = _evaluate_batch


def ace_evaluate(X, model, ps, species, H, pbc, precision="float64", calculator=None):
    """Numeric ``ace_evaluate(X, model, ps, env, H)`` (serial.jl:132-161) on the GPU:
    X (n, 3) Å, species labels, H columns = lattice vectors. Returns E, F (3 x n), W."""
    from .system import System
    calc = calculator or ACECalculator(model, ps, precision=precision)
    if calculator is not None and ps is not None:
        calc.set_parameters(ps)
    r = calc.evaluate(System(X, H, pbc, species))
    return r
