"""Independent brute-force oracles restated from the reference's own tests
(test infrastructure only).

* ``min_image_oracle``: O(N^2) minimum-image pair sum with a deliberately different
  wrapping implementation (test/fixtures.jl:28-51).
* ``image_oracle``: exhaustive image sum over a fixed oversized lattice cube, no
  reciprocal bounds, sorting or cell stencil (test/periodic_images.jl:3-28).

They are written as plain loops over numpy rows and are meant for small systems.
"""
import itertools

import numpy as np


def smooth_pair(a, r2, radius):
    q = 1 - r2 / radius ** 2
    return a * q ** 3, (-6 * a * q ** 2 / radius ** 2)


def min_image_oracle(X, H, pbc, species, amplitudes, radius):
    X = np.asarray(X, float)
    n = len(X)
    f = np.zeros((n, 3))
    v = np.zeros((3, 3))
    energy = 0.0
    pairs = set()
    for i in range(n):
        for j in range(i + 1, n):
            d = X[j] - X[i]
            for a in range(3):
                if pbc[a]:
                    L = H[a, a]
                    while d[a] > L / 2:
                        d[a] -= L
                    while d[a] < -L / 2:
                        d[a] += L
            r2 = float(d @ d)
            if not r2 < radius ** 2:
                continue
            pairs.add((i, j))
            e, coef = smooth_pair(amplitudes[species[i], species[j]], r2, radius)
            energy += e
            force = coef * d
            f[i] += force
            f[j] -= force
            v -= np.outer(d, force)
    return dict(energy=energy, forces=f, virial=v, pairs=pairs)


def image_oracle(X, H, pbc, species, amplitudes, radius, shell=10, model=None):
    """``model(zi, zj, r2) -> (energy, coef)`` overrides the SmoothPair callback; the
    force on i is ``coef * d`` with ``d = x_j - x_i + H s``."""
    X = np.asarray(X, float)
    H = np.asarray(H, float)
    n = len(X)
    f = np.zeros((n, 3))
    v = np.zeros((3, 3))
    energy = 0.0
    pairs = set()
    ax = [range(-shell, shell + 1) if pbc[a] else range(0, 1) for a in range(3)]
    shifts = []
    for z, y, x in itertools.product(ax[2], ax[1], ax[0]):
        s = (x, y, z)
        off = np.zeros(3)
        for a in range(3):
            if pbc[a]:
                off = off + s[a] * H[:, a]
        shifts.append((s, off))
    for i in range(n):
        for j in range(i, n):
            base = X[j] - X[i]
            for s, off in shifts:
                if i == j and s <= (0, 0, 0):
                    continue
                d = base + off
                r2 = float(d @ d)
                if not r2 < radius ** 2:
                    continue
                pairs.add((i, j, s))
                if model is None:
                    e, coef = smooth_pair(amplitudes[species[i], species[j]], r2, radius)
                else:
                    e, coef = model(species[i], species[j], r2)
                energy += e
                force = coef * d
                f[i] += force
                f[j] -= force
                v -= np.outer(d, force)
    return dict(energy=energy, forces=f, virial=v, pairs=pairs)


def directed_from_image_pairs(pairs, X, H, pbc):
    """Directed (i, j, S) set implied by unordered image pairs whose shift is relative to
    the given (unwrapped) coordinates: both (i, j, s) and (j, i, -s)."""
    out = set()
    for (i, j, s) in pairs:
        out.add((i, j, tuple(s)))
        out.add((j, i, tuple(-k for k in s)))
    return out
