"""ACE part of the oracle (test infrastructure only).

* ``load_resolved``: reader of the resolved schema v1 TOML
  (src/submodules/ACE/resolved.jl:57-142), default basis policy only.
* ``read_frames``: the 33-frame extxyz fixture reader of
  test/neighbour_saved_reference.jl:9-26 (Lattice columns = lattice vectors).
* ``ace_evaluate``: FP64 site pipeline in ace_oracle.cpp via ctypes.

Pinned in tests/test_oracle_ace.py against ACEpotentials 0.10.2 predictions
(data/ace_ref/reloaded_predictions.toml, tolerances of
test/neighbour_saved_reference.jl:38-40), finite differences, closed-form real
spherical harmonics (scipy) and brute-force product enumeration.
"""
import ctypes

import numpy as np

from .core import OracleError, lib

try:
    import tomllib as _toml
except ImportError:  # Python 3.10
    import tomli as _toml

ENVELOPES = {"PolyEnvelope2sX": 0, "PolyEnvelope1sR": 1, "ACE1_PolyEnvelope1sR": 2}


def _family(D, nz):
    if D["kind"] != "piecewise_cubic_agnesi":
        raise ValueError("unsupported resolved radial kind")
    spec = [(int(b["n"]), int(b["l"])) for b in D["spec"]]
    nr = len(spec)
    chans = {}
    for ch in D["channels"]:
        chans[(int(ch["centre"]) - 1, int(ch["neighbour"]) - 1)] = ch
    if len(chans) != nz * nz:
        raise ValueError("missing radial species channels")
    nnodes = len(next(iter(chans.values()))["nodes"])
    tf = np.zeros((nz * nz, 8))
    env_kind = np.zeros(nz * nz, np.int32)
    env = np.zeros((nz * nz, 5))
    nodes = np.zeros((nz * nz, nnodes))
    coef = np.zeros((nz * nz, (nnodes - 1) * nr * 4))
    for (i, j), ch in chans.items():
        c = i + nz * j
        t = ch["transform"]
        tf[c] = [t["p"], t["q"], t["a"], t["rin"], t["r0"], t["yin"], t["ycut"], t["rcut"]]
        kind = ch["envelope_kind"]
        e = ch["envelope"]
        env_kind[c] = ENVELOPES[kind]
        if kind == "PolyEnvelope2sX":
            env[c] = [e["x1"], e["x2"], e["p1"], e["p2"], e["s"]]
        elif kind == "PolyEnvelope1sR":
            env[c, :2] = [e["rcut"], e["p"]]
        else:
            env[c, :3] = [e["rcut"], e["r0"], e["p"]]
        if len(ch["nodes"]) != nnodes or list(ch["coefficient_shape"]) != [4, nr, nnodes - 1]:
            raise ValueError("ragged spline tables")
        nodes[c] = ch["nodes"]
        coef[c] = ch["coefficients"]  # Julia column-major [4, nr, nint] flattened
    return dict(spec=spec, nr=nr, nnodes=nnodes, tf=tf, env_kind=env_kind, env=env, nodes=nodes, coef=coef,
                rcut=float(tf[:, 7].max()))


def load_resolved(path):
    if isinstance(path, dict):
        D = path
    else:
        with open(path, "rb") as f:
            D = _toml.load(f)
    if D.get("schema") != "PotentialFrameWork.resolved_ACE" or D.get("schema_version") != 1:
        raise ValueError("unsupported resolved ACE schema")
    z = [int(v) for v in D["species_atomic_numbers"]]
    nz = len(z)
    ang = D["angular"]
    if ang["kind"] not in ("real_spherical", "real_solid") or ang["normalization"] != "L2" \
            or ang["ordering"] != "l*l+l+m+1":
        raise ValueError("unsupported angular convention")  # resolved.jl:109-110
    lmax = int(ang["lmax"])
    radial = _family(D["radial"], nz)
    pair = _family(D["pair_radial"], nz)
    dens = [(int(a) - 1, int(b) - 1) for a, b in D["density_indices"]]
    products = [[int(v) - 1 for v in p] for p in D["products"]]
    c = D["coupling"]
    m, nprod = [int(v) for v in c["shape"]]
    C = np.zeros((m, nprod))
    for r, col, v in zip(c["rows"], c["columns"], c["values"]):
        C[int(r) - 1, int(col) - 1] += float(v)
    flat = np.asarray(D["parameters"]["flat"], float)
    nq = pair["nr"]
    WB = flat[: m * nz].reshape(nz, m).T.copy()
    Wpair = flat[m * nz:].reshape(nz, nq).T.copy()
    out = dict(atomic_numbers=z, nz=nz, lmax=lmax, solid=ang["kind"] == "real_solid",
               reference=np.asarray(D["reference_eV"], float), radial=radial,
               pair=pair, density=dens, products=products, coupling=C, WB=WB, Wpair=Wpair,
               cutoff=max(radial["rcut"], pair["rcut"]), source_sha256=D.get("source_sha256"))
    return with_policy(out, D.get("interface_policy"))


def with_policy(model, P):
    """The basis policy of a resolved file (interface_persistence.jl:3-25) applied
    the way the evaluation uses it: WB .* mask (effective_WB), no pair term when
    disabled (pair_allowed), no reference energies when disabled
    (reference_value), and the projected species (0-based channels) and orders
    (interface.jl:115-125)."""
    if not P:
        return model
    m = dict(model)
    mask = np.asarray(P.get("mask", [True] * m["WB"].shape[0]), bool)
    m["WB"] = m["WB"] * mask[:, None]
    if not P.get("pair", True):
        m["Wpair"] = np.zeros_like(m["Wpair"])
    if not P.get("reference", True):
        m["reference"] = np.zeros_like(m["reference"])
    orders = list(P.get("orders", []))
    if orders:
        m["orders"] = (int(orders[0]), int(orders[1]))
        m["species"] = tuple(int(v) - 1 for v in P.get("species", []))
    return m


def read_frames(path):
    frames = []
    with open(path) as f:
        lines = f.read().splitlines()
    k = 0
    while k < len(lines):
        if not lines[k].strip():
            k += 1
            continue
        n = int(lines[k])
        header = lines[k + 1]
        lat = header.split('Lattice="')[1].split('"')[0]
        H = np.array([float(v) for v in lat.split()]).reshape(3, 3).T  # columns = lattice vectors
        labels, X = [], np.zeros((n, 3))
        for i in range(n):
            fields = lines[k + 2 + i].split()
            labels.append(fields[0])
            X[i] = [float(v) for v in fields[1:4]]
        frames.append(dict(X=X, H=H, labels=labels, pbc=(True, True, True)))
        k += 2 + n
    return frames


ELEMENTS = {"H": 1, "He": 2, "Li": 3, "C": 6, "N": 7, "O": 8, "Al": 13, "Si": 14, "Ti": 22, "V": 23, "Fe": 26,
            "Ni": 28, "Cu": 29, "Nb": 41, "Mo": 42, "W": 74}


def channels(model, labels):
    z = model["atomic_numbers"]
    return np.array([z.index(ELEMENTS[s]) for s in labels], np.int32)


class _Family(ctypes.Structure):
    _fields_ = [("nr", ctypes.c_int), ("nnodes", ctypes.c_int), ("tf", ctypes.c_void_p),
                ("env_kind", ctypes.c_void_p), ("env", ctypes.c_void_p), ("nodes", ctypes.c_void_p),
                ("coef", ctypes.c_void_p)]


class _Model(ctypes.Structure):
    _fields_ = [("nz", ctypes.c_int), ("lmax", ctypes.c_int), ("ndens", ctypes.c_int), ("nprod", ctypes.c_int),
                ("reference", ctypes.c_void_p), ("radial", _Family), ("pair", _Family),
                ("dens_n", ctypes.c_void_p), ("dens_m", ctypes.c_void_p), ("prod_ptr", ctypes.c_void_p),
                ("prod_idx", ctypes.c_void_p), ("weights", ctypes.c_void_p), ("wpair", ctypes.c_void_p),
                ("solid", ctypes.c_int), ("selected", ctypes.c_void_p), ("order_lo", ctypes.c_int),
                ("order_hi", ctypes.c_int), ("exclude", ctypes.c_int)]


def _pack(model):
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(np.asarray(a, dtype=dt))
        keep.append(a)
        return a.ctypes.data

    def fam(f):
        return _Family(f["nr"], f["nnodes"], arr(f["tf"], np.float64), arr(f["env_kind"], np.int32),
                       arr(f["env"], np.float64), arr(f["nodes"], np.float64), arr(f["coef"], np.float64))

    prods = model["products"]
    ptr = np.zeros(len(prods) + 1, np.int32)
    ptr[1:] = np.cumsum([len(p) for p in prods])
    idx = np.concatenate([np.asarray(p, np.int32) for p in prods]) if prods else np.zeros(0, np.int32)
    weights = model["coupling"].T @ model["WB"]  # C^T (WB .* mask), default mask
    orders = model.get("orders")  # None, or (lo, hi) of project_species
    sel = np.zeros(model["nz"], np.int32)
    for z in model.get("species", ()):
        sel[z] = 1
    m = _Model(model["nz"], model["lmax"], len(model["density"]), len(prods), arr(model["reference"], np.float64),
               fam(model["radial"]), fam(model["pair"]), arr([d[0] for d in model["density"]], np.int32),
               arr([d[1] for d in model["density"]], np.int32), arr(ptr, np.int32), arr(idx, np.int32),
               arr(weights.T.copy(), np.float64), arr(model["Wpair"].T.copy(), np.float64),
               int(model.get("solid", False)), arr(sel, np.int32), -1 if orders is None else int(orders[0]),
               -1 if orders is None else int(orders[1]), int(model.get("exclude", -1)))
    return m, keep


lib.oracle_projected_product.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                                         ctypes.c_void_p]
lib.oracle_projected_product.restype = ctypes.c_double


def projected_product(indices, U, V, lo, hi, w=1.0):
    """(P, dU, dV) of one product under orders lo:hi (projection.jl:7-38)."""
    idx = np.ascontiguousarray(np.asarray(indices, np.int32))
    U = np.ascontiguousarray(np.asarray(U, float))
    V = np.ascontiguousarray(np.asarray(V, float))
    dU = np.zeros(len(U))
    dV = np.zeros(len(V))
    P = lib.oracle_projected_product(idx.ctypes.data, len(idx), U.ctypes.data, V.ctypes.data, int(lo), int(hi),
                                     float(w), dU.ctypes.data, dV.ctypes.data)
    return P, dU, dV


lib.oracle_ace_evaluate.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(_Model), ctypes.c_int,
                                    ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
lib.oracle_ace_evaluate.restype = ctypes.c_int


lib.oracle_ace_evaluate_scheduled.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(_Model), ctypes.c_int,
                                              ctypes.c_int, ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
lib.oracle_ace_evaluate_scheduled.restype = ctypes.c_int


lib.oracle_ace_set_input_rounding.argtypes = [ctypes.c_int]
lib.oracle_ace_set_input_rounding.restype = None
ROUNDING = {"none": 0, "bfloat16": 1, "float16": 2, "float32": 3}


def set_input_rounding(kind):
    """Precision study: round R, R', Y and grad Y to ``kind`` ("none",
    "bfloat16", "float16", "float32") before the density contraction and the
    pullback in every later evaluation (FP64 otherwise)."""
    lib.oracle_ace_set_input_rounding(ROUNDING[kind])


def ace_evaluate(X, species, H, pbc, model, threads=1, mask=7, coloured=False):
    """FP64 (E, F (n,3), W (3x3), site energies (n)) of the resolved model. With
    threads > 1 and ``coloured`` the sites run phase by phase over the greedy
    footprint colouring with one shared force array (execute_sites!);
    otherwise each thread keeps private buffers."""
    X = np.ascontiguousarray(np.asarray(X, float).reshape(-1, 3))
    n = len(X)
    sp = np.ascontiguousarray(np.asarray(species, np.int32))
    Hc = np.ascontiguousarray(np.asarray(H, float).T)
    pb = np.ascontiguousarray(np.asarray(pbc, np.uint8))
    m, keep = _pack(model)
    E = np.zeros(1)
    F = np.zeros((n, 3))
    W = np.zeros(9)
    site = np.zeros(n)
    err = ctypes.create_string_buffer(512)
    rc = lib.oracle_ace_evaluate_scheduled(n, X.ctypes.data, sp.ctypes.data, Hc.ctypes.data, pb.ctypes.data,
                                           float(model["cutoff"]), ctypes.byref(m), int(threads), int(bool(coloured)),
                                           int(mask), E.ctypes.data, F.ctypes.data, W.ctypes.data,
                                           site.ctypes.data, err, 512)
    del keep
    if rc != 0:
        raise OracleError(rc, err.value.decode())
    return float(E[0]), F, W.reshape(3, 3).T.copy(), site


lib.oracle_ylm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
lib.oracle_ylm.restype = ctypes.c_int


def ylm(L, R, solid=False, grad=False):
    """Real L2-normalised Y_lm (index l*l+l+m) of the direction of R (r^l Y_lm
    with ``solid``), from the oracle's angular basis; with ``grad`` also the
    Cartesian gradients ((L+1)^2 x 3)."""
    R = np.ascontiguousarray(np.asarray(R, float).reshape(3))
    Y = np.zeros((L + 1) ** 2)
    dY = np.zeros(3 * (L + 1) ** 2)
    rc = lib.oracle_ylm(int(L), int(solid), R.ctypes.data, Y.ctypes.data, dY.ctypes.data)
    if rc != 0:
        raise OracleError(rc, "harmonics of a zero vector")
    return (Y, dY.reshape(-1, 3)) if grad else Y
