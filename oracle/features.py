"""CPU restatement of the equivariant feature forward (test infrastructure only).

``ace_features`` follows src/submodules/ACE/features.jl:141-178: for every
atom, the species-resolved densities A_a = sum_j [channel_a in {0, z_j}]
R_(n_a)(r_ij) Y_(h_a)(r_ij) over the directed neighbours within the layer
cutoff (oracle neighbour list), products P_k = prod_a A_a, B_b = coupling_b P
reshaped (2L+1, basis) with m fastest, and B_b W_b[:, :, z_i]. Radials are
the smooth exponential family of radials.jl:193-206 times the radial scale
(radial, centre species, neighbour species). Pure Python loops: small cases.
"""
import numpy as np

from .ace import ylm
from .core import nl_pairs


def smooth_exponential(r, cutoff, powers, rate, amplitude):
    """radials.jl:193-206: amplitude (r/rc)^p (1 - r/rc)^4 exp(-rate r), clamped at rc."""
    x = r / cutoff
    q = max(1.0 - x, 0.0)
    return np.array([amplitude * x ** p * q ** 4 * np.exp(-rate * r) for p in powers])


def ace_features(X, channels, H, pbc, tables, params):
    """tables: dict(density=[(channel 0..nz, radial 0-based, harmonic 0-based)],
    products=[tuple of 0-based density indices], coupling=[dense rows x nprod],
    irreps=[(L, channels)], lmax, cutoff, radial_cutoff, powers);
    params: dict(mixing=[(basis, channels, nz)], scale (nr, nz, nz), rate, amplitude).
    Returns one (2L+1, channels, n) array per block."""
    X = np.asarray(X, float).reshape(-1, 3)
    n = len(X)
    out = [np.zeros((2 * L + 1, ch, n)) for L, ch in tables["irreps"]]
    if n == 0:
        return out
    i, j, _, R, _ = nl_pairs(X, H, pbc, tables["cutoff"])
    dens = tables["density"]
    for a in range(n):
        A = np.zeros(len(dens))
        za = channels[a]
        for e in np.nonzero(i == a)[0]:
            r = float(np.linalg.norm(R[e]))
            if not r > 0:
                raise ValueError("coincident physical atoms have undefined harmonics")
            zj = channels[j[e]]
            Rv = smooth_exponential(r, tables["radial_cutoff"], tables["powers"], params["rate"], params["amplitude"])
            Rv = Rv * params["scale"][:, za, zj]
            Y = ylm(tables["lmax"], R[e])
            for k, (zc, rn, h) in enumerate(dens):
                if zc == 0 or zc == zj + 1:
                    A[k] += Rv[rn] * Y[h]
        P = np.array([np.prod([A[q] for q in p]) for p in tables["products"]])
        for b, (L, ch) in enumerate(tables["irreps"]):
            dim = 2 * L + 1
            Bb = (tables["coupling"][b] @ P).reshape(-1, dim).T  # (dim, basis), m fastest
            out[b][:, :, a] = Bb @ params["mixing"][b][:, :, za]
    return out


def feature_vjp(X, channels, H, pbc, tables, params, cot):
    """The reverse of ``ace_features`` for cotangents ``cot`` (one
    (2L+1, channels, n) array per block), following feature_vjp and
    feature_local_reverse (features.jl:204-255): mixing B^T G, dB = G W^T,
    product weights coupling^T vec(dB), dA by the product rule, then per edge
    g = sum_a dA_a (R'_n Y_m r^ + R_n grad Y_m) with dx_i -= g, dx_j += g, the
    radial scale gradient dA_a R_n Y_m, and the smooth exponential's rate and
    amplitude gradients. Returns dx (n, 3) and a dict of parameter gradients."""
    X = np.asarray(X, float).reshape(-1, 3)
    n = len(X)
    dx = np.zeros((n, 3))
    gmix = [np.zeros_like(W) for W in params["mixing"]]
    gscale = np.zeros_like(params["scale"])
    grate = gamp = 0.0
    if n == 0:
        return dx, dict(mixing=gmix, scale=gscale, rate=grate, amplitude=gamp)
    i, j, _, R, _ = nl_pairs(X, H, pbc, tables["cutoff"])
    dens, prods = tables["density"], tables["products"]
    rc, powers, rate, amp = tables["radial_cutoff"], tables["powers"], params["rate"], params["amplitude"]
    for a in range(n):
        za = channels[a]
        edges = np.nonzero(i == a)[0]
        data = []
        A = np.zeros(len(dens))
        for e in edges:
            d = R[e]
            r = float(np.linalg.norm(d))
            zj = channels[j[e]]
            x = r / rc
            q = max(1.0 - x, 0.0)
            ex = np.exp(-rate * r)
            raw = np.array([amp * x ** p * q ** 4 * ex for p in powers])
            draw = np.array([amp * ex * ((p * x ** (p - 1) / rc if p else 0.0) * q ** 4 - 4 * x ** p * q ** 3 / rc
                                         - rate * x ** p * q ** 4) for p in powers])
            sc = params["scale"][:, za, zj]
            Y, dY = ylm(tables["lmax"], d, grad=True)
            data.append((e, d, r, zj, raw, draw, sc, Y, dY))
            for k, (zc, rn, h) in enumerate(dens):
                if zc == 0 or zc == zj + 1:
                    A[k] += raw[rn] * sc[rn] * Y[h]
        P = np.array([np.prod([A[t] for t in p]) for p in prods])
        weights = np.zeros(len(prods))
        for b, (L, ch) in enumerate(tables["irreps"]):
            dim = 2 * L + 1
            Bb = (tables["coupling"][b] @ P).reshape(-1, dim).T  # (dim, basis)
            G = cot[b][:, :, a]
            W = params["mixing"][b][:, :, za]
            gmix[b][:, :, za] += Bb.T @ G
            dB = G @ W.T  # (dim, basis)
            weights += tables["coupling"][b].T @ dB.T.reshape(-1)  # vec(dB), m fastest
        dA = np.zeros(len(dens))
        for k, p in enumerate(prods):
            for q in range(len(p)):
                others = np.prod([A[p[t]] for t in range(len(p)) if t != q])
                dA[p[q]] += weights[k] * others
        for (e, d, r, zj, raw, draw, sc, Y, dY) in data:
            u = d / r
            g = np.zeros(3)
            for k, (zc, rn, h) in enumerate(dens):
                if not (zc == 0 or zc == zj + 1):
                    continue
                g += dA[k] * sc[rn] * (draw[rn] * Y[h] * u + raw[rn] * dY[h])
                gscale[rn, za, zj] += dA[k] * raw[rn] * Y[h]
                grate += dA[k] * sc[rn] * Y[h] * (-r * raw[rn])
                gamp += dA[k] * sc[rn] * Y[h] * (raw[rn] / amp)
            dx[a] -= g
            dx[j[e]] += g
    return dx, dict(mixing=gmix, scale=gscale, rate=grate, amplitude=gamp)
