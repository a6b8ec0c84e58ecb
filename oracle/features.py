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
