"""Construction of a synthetic resolved ACE model (construction only, host side).

Generates the invariant (L = 0, even parity) coupled basis the way
``generate_feature_tables`` does (src/submodules/ACE/feature_coupling.jl:96-143):
density factors are (radial channel r = (n, l), harmonic l*l+l+m) multiplets;
paths of at most ``correlation`` factors with sum(n + l) <= ``totaldegree`` are
coupled with real Clebsch-Gordan tensors (``real_cg``, feature_coupling.jl:22-60)
and orthonormalised per multiplet (``independent_multiplet!``, :68-83).

The result is written in the resolved schema v1 (src/submodules/ACE/resolved.jl:103-142)
with piecewise-cubic Agnesi radial splines, so the same loader and kernels run
it and the TiAl fixture. Radial functions: on x in [-1, 1] (the Agnesi
coordinate), R_(n,l)(x) = T_(n-1)(x) * (1 + l/10) with the PolyEnvelope2sX
envelope (1 - x)^2 (1 + x)^2, tabulated as cubic Hermite pieces on ``nnodes``
uniform nodes, then scaled per (centre, neighbour) species by seeded factors in
[0.8, 1.2]. Readout WB[i, z] ~ N(0, 1) / (1 + degree_i)^2 with degree_i the
largest sum(n + l) over the products of invariant i (fitting.jl:77-80). The
pair term uses 4 functions with the ACE1_PolyEnvelope1sR envelope.
"""
import itertools
import math
from fractions import Fraction

import numpy as np


# ---------------------------------------------------------------- Clebsch-Gordan

def complex_cg(a, m, b, n, L, M):
    """<a m b n | L M> (Condon-Shortley), exact rational radicands (feature_coupling.jl:34-46)."""
    if m + n != M or not (abs(a - b) <= L <= a + b):
        return 0.0
    f = math.factorial
    r1 = Fraction((2 * L + 1) * f(L + a - b) * f(L - a + b) * f(a + b - L), f(a + b + L + 1))
    r2 = Fraction(f(L + M) * f(L - M) * f(a - m) * f(a + m) * f(b - n) * f(b + n))
    lo = max(0, b - L - m, a + n - L)
    hi = min(a + b - L, a - m, b + n)
    s = Fraction(0)
    for k in range(lo, hi + 1):
        s += Fraction((-1) ** k, f(k) * f(a + b - L - k) * f(a - m - k) * f(b + n - k) * f(L - b + m + k)
                      * f(L - a - n + k))
    return math.sqrt(r1 * r2) * float(s) if s != 0 else 0.0


def real_harmonic_map(l):
    """Condon-Shortley complex Y -> real L2 Y without phase (feature_coupling.jl:22-33)."""
    U = np.zeros((2 * l + 1, 2 * l + 1), complex)
    U[l, l] = 1
    for m in range(1, l + 1):
        U[l + m, l + m] = (-1) ** m / math.sqrt(2)
        U[l + m, l - m] = 1 / math.sqrt(2)
        U[l - m, l + m] = -1j * (-1) ** m / math.sqrt(2)
        U[l - m, l - m] = 1j / math.sqrt(2)
    return U


_CG = {}


def real_cg(a, b, L):
    key = (a, b, L)
    if key in _CG:
        return _CG[key]
    Ua, Ub, U = real_harmonic_map(a), real_harmonic_map(b), real_harmonic_map(L)
    C = np.zeros((2 * L + 1, 2 * a + 1, 2 * b + 1), complex)
    for m in range(-a, a + 1):
        for n in range(-b, b + 1):
            M = m + n
            if abs(M) > L:
                continue
            c = complex_cg(a, m, b, n, L, M) * (1j ** (a + b - L))
            if c == 0:
                continue
            C += c * np.einsum("r,s,t->rst", U[:, L + M], np.conj(Ua[:, a + m]), np.conj(Ub[:, b + n]))
    if np.abs(C.imag).max() > 1e-12:
        raise RuntimeError("nonreal harmonic coupling")
    R = C.real.copy()
    R[np.abs(R) < 1e-14] = 0
    _CG[key] = R
    return R


# ---------------------------------------------------------------- polynomial multiplets

def _axpy(a, x, y):
    for k in sorted(x):
        y[k] = y.get(k, 0.0) + a * x[k]


def _mdot(x, y):
    return sum(sum(v * y[m].get(k, 0.0) for k, v in sorted(x[m].items())) for m in range(len(x))) / len(x)


def _independent(basis, x):
    for _ in range(2):
        for y in basis:
            c = _mdot(x, y)
            for m in range(len(x)):
                _axpy(-c, y[m], x[m])
    scale = math.sqrt(max(0.0, _mdot(x, x)))
    if not scale > 1e-10:
        return False
    for p in x:
        for k in list(p):
            v = p[k] / scale
            if abs(v) < 1e-13:
                del p[k]
            else:
                p[k] = v
    basis.append(x)
    return True


def _couple(x, y, C):
    out = [dict() for _ in range(C.shape[0])]
    nz = np.argwhere(C != 0)
    for r, s, t in nz:
        c = C[r, s, t]
        p = out[r]
        for i, v in sorted(x[s].items()):
            for j, w in sorted(y[t].items()):
                key = tuple(sorted(i + j))
                p[key] = p.get(key, 0.0) + c * v * w
    return out


def generate_invariant_tables(rspec, correlation, lmax, totaldegree):
    """Density (r, lm) list, products (tuples of 0-based density indices) and the
    sparse coupling (rows, cols, vals) of the L = 0 even basis, species-summed (nz = 1)."""
    density, factors, vectors = [], [], []
    for r, (n, l) in enumerate(rspec):
        if not 0 <= l <= lmax:
            continue
        factors.append((r, l, n + l))
        idx = []
        for m in range(-l, l + 1):
            density.append((r, l * l + l + m))
            idx.append(len(density) - 1)
        vectors.append([{(i,): 1.0} for i in idx])
    block = []

    def visit(first, degree, order, parity, states):
        if order >= correlation:
            return
        for f in range(first, len(factors)):
            _, l, weight = factors[f]
            if degree + weight > totaldegree:
                continue
            nxt = {}
            for J in sorted(states):
                for x in states[J]:
                    for L in range(abs(J - l), J + l + 1):
                        y = _couple(x, vectors[f], real_cg(J, l, L))
                        _independent(nxt.setdefault(L, []), y)
            p = parity * (-1) ** l
            if p == 1:
                block.extend(nxt.get(0, []))
            visit(f, degree + weight, order + 1, p, nxt)

    visit(0, 0, 0, 1, {0: [[{(): 1.0}]]})
    keys = set()
    for x in block:
        for p in x:
            keys.update(p)
    products = sorted(keys, key=lambda k: (len(k), k))
    lookup = {p: i for i, p in enumerate(products)}
    rows, cols, vals = [], [], []
    for i, x in enumerate(block):
        for k, v in sorted(x[0].items()):
            rows.append(i)
            cols.append(lookup[k])
            vals.append(v)
    return density, products, (rows, cols, vals, len(block))


# ---------------------------------------------------------------- radial tables

def _agnesi(p, q, a, rin, r0, rcut):
    def y_of(r):
        s = (r - rin) / (r0 - rin)
        return 1.0 if r <= rin else 1.0 / (1 + a * s ** q / (1 + s ** (q - p)))
    return dict(p=p, q=q, a=a, rin=rin, r0=r0, rcut=rcut, normalization_rin=rin, yin=y_of(rin), ycut=y_of(rcut))


def _hermite(f, df, nodes):
    """Cubic Hermite coefficients [4, nfun, nint] (power fastest) of f on nodes."""
    nfun = f(nodes[0]).shape[0]
    C = np.zeros((4, nfun, len(nodes) - 1))
    for k in range(len(nodes) - 1):
        h = nodes[k + 1] - nodes[k]
        y0, y1 = f(nodes[k]), f(nodes[k + 1])
        d0, d1 = df(nodes[k]) * h, df(nodes[k + 1]) * h
        C[0, :, k] = y0
        C[1, :, k] = d0
        C[2, :, k] = 3 * (y1 - y0) - 2 * d0 - d1
        C[3, :, k] = 2 * (y0 - y1) + d0 + d1
    return C


def _chebyshev(nmax, x):
    T = np.zeros(nmax)
    dT = np.zeros(nmax)
    T[0], dT[0] = 1.0, 0.0
    if nmax > 1:
        T[1], dT[1] = x, 1.0
    for k in range(2, nmax):
        T[k] = 2 * x * T[k - 1] - T[k - 2]
        dT[k] = 2 * T[k - 1] + 2 * x * dT[k - 1] - dT[k - 2]
    return T, dT


def synthetic_model(nz=2, atomic_numbers=(26, 28), nmax=8, lmax=4, correlation=3, totaldegree=12, rcut=5.0,
                    r0=2.5, nnodes=100, npair=4, seed=7):
    """The resolved-schema-v1 dictionary of the synthetic model (TOML-ready)."""
    rng = np.random.default_rng(seed)
    rspec = [(n, l) for l in range(lmax + 1) for n in range(1, nmax + 1)]
    density, products, (rows, cols, vals, nbasis) = generate_invariant_tables(rspec, correlation, lmax, totaldegree)
    nodes = np.linspace(-1.0, 1.0, nnodes)

    def radial_f(x):
        T, _ = _chebyshev(nmax, x)
        return np.array([T[n - 1] * (1 + 0.1 * l) for (n, l) in rspec])

    def radial_df(x):
        _, dT = _chebyshev(nmax, x)
        return np.array([dT[n - 1] * (1 + 0.1 * l) for (n, l) in rspec])

    base = _hermite(radial_f, radial_df, nodes)
    tf = _agnesi(2, 4, 1.0, 0.0, r0, rcut)
    ptf = _agnesi(1, 3, 0.4, 0.0, r0, rcut)

    def pair_f(x):
        T, _ = _chebyshev(npair, x)
        return T.copy()

    def pair_df(x):
        _, dT = _chebyshev(npair, x)
        return dT.copy()

    pbase = _hermite(pair_f, pair_df, nodes)
    channels, pchannels = [], []
    for i in range(nz):
        for j in range(nz):
            scale = rng.uniform(0.8, 1.2, size=len(rspec))
            channels.append(dict(centre=i + 1, neighbour=j + 1, transform=dict(tf), envelope_kind="PolyEnvelope2sX",
                                 envelope=dict(x1=-1.0, x2=1.0, p1=2, p2=2, s=1.0), nodes=nodes.tolist(),
                                 coefficient_shape=[4, len(rspec), nnodes - 1],
                                 coefficients=(base * scale[None, :, None]).flatten(order="F").tolist()))
            pchannels.append(dict(centre=i + 1, neighbour=j + 1, transform=dict(ptf),
                                  envelope_kind="ACE1_PolyEnvelope1sR", envelope=dict(rcut=rcut, r0=r0, p=2),
                                  nodes=nodes.tolist(), coefficient_shape=[4, npair, nnodes - 1],
                                  coefficients=pbase.flatten(order="F").tolist()))
    weight_of = [n + l for (n, l) in rspec]
    deg = np.zeros(nbasis, int)
    for r_, c_ in zip(rows, cols):
        d = sum(weight_of[density[a][0]] for a in products[c_])
        deg[r_] = max(deg[r_], d)
    WB = rng.normal(size=(nbasis, nz)) / (1.0 + deg[:, None]) ** 2
    Wpair = 0.01 * rng.normal(size=(npair, nz))
    return {
        "schema": "PotentialFrameWork.resolved_ACE",
        "schema_version": 1,
        "source_sha256": "synthetic-invariant-basis-nmax%d-lmax%d-corr%d-deg%d-seed%d" % (
            nmax, lmax, correlation, totaldegree, seed),
        "species_atomic_numbers": list(atomic_numbers),
        "reference_eV": [-3.0, -4.0][:nz] + [-3.5] * max(0, nz - 2),
        "angular": {"kind": "real_spherical", "lmax": lmax, "normalization": "L2", "ordering": "l*l+l+m+1"},
        "radial": {"kind": "piecewise_cubic_agnesi", "spec": [{"n": n, "l": l} for (n, l) in rspec],
                   "channels": channels},
        "pair_radial": {"kind": "piecewise_cubic_agnesi", "spec": [{"n": n, "l": 0} for n in range(1, npair + 1)],
                        "channels": pchannels},
        "density_indices": [[r + 1, m + 1] for (r, m) in density],
        "products": [[a + 1 for a in p] for p in products],
        "coupling": {"shape": [nbasis, len(products)], "rows": [r + 1 for r in rows], "columns": [c + 1 for c in cols],
                     "values": vals},
        "parameters": {"ordering": "vec(WB), vec(Wpair); Julia column-major", "WB_shape": [nbasis, nz],
                       "Wpair_shape": [npair, nz],
                       "flat": np.concatenate([WB.flatten(order="F"), Wpair.flatten(order="F")]).tolist()},
    }


def write_toml(model, path):
    """Minimal TOML writer for the resolved schema (tables and arrays of tables)."""
    def val(v):
        if isinstance(v, bool):
            return "true" if v else "false"
        if isinstance(v, (int, np.integer)):
            return str(int(v))
        if isinstance(v, (float, np.floating)):
            return repr(float(v))
        if isinstance(v, str):
            return '"' + v.replace("\\", "\\\\").replace('"', '\\"') + '"'
        if isinstance(v, (list, tuple)):
            return "[" + ", ".join(val(x) for x in v) + "]"
        raise TypeError(type(v))

    lines = []

    def table(prefix, d):
        simple = {k: v for k, v in d.items() if not isinstance(v, dict) and not (
            isinstance(v, list) and v and isinstance(v[0], dict) and k == "channels")}
        for k, v in simple.items():
            if isinstance(v, list) and v and isinstance(v[0], dict):
                lines.append(f"{k} = [" + ", ".join("{" + ", ".join(f"{a} = {val(b)}" for a, b in x.items()) + "}"
                                                     for x in v) + "]")
            else:
                lines.append(f"{k} = {val(v)}")
        for k, v in d.items():
            if isinstance(v, dict):
                lines.append("")
                lines.append(f"[{prefix}{k}]")
                table(f"{prefix}{k}.", v)
        for k, v in d.items():
            if k == "channels" and isinstance(v, list):
                for ch in v:
                    lines.append("")
                    lines.append(f"[[{prefix}{k}]]")
                    table(f"{prefix}{k}.", ch)

    table("", model)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
