"""Minimal AtomsBase-like system used at the host boundary.

Mirrors what the reference reads from an ``AtomsBase.AbstractSystem{3}``
(src/submodules/NeighbourList/storage.jl:135-170): positions in Å, cell vectors
(lattice vectors in COLUMNS of ``cell``), periodicity and species labels.
Units are plain Å and eV; there is no Unitful layer.
"""
import numpy as np

from ._capi import ArgumentError, DimensionMismatch


class Labels(tuple):
    """Immutable species labels of a system. The distinct labels (in order of
    first appearance) and each atom's index among them are computed once, so
    mapping the labels onto a model's channels costs one lookup per distinct
    label and a gather, not a Python loop over the atoms, on every call."""

    def codes(self):
        c = self.__dict__.get("_codes")
        if c is None:
            index = {}
            codes = np.fromiter((index.setdefault(s, len(index)) for s in self), np.int32, count=len(self))
            c = self.__dict__["_codes"] = (list(index), codes)
        return c


def map_labels(labels, lookup):
    """Channel of every label: ``lookup(label)`` (None for an unknown label)
    evaluated once per distinct label. Returns (int32 channels, first unknown
    label or None)."""
    labels = labels if isinstance(labels, Labels) else Labels(labels)
    try:
        distinct, codes = labels.codes()
    except TypeError:
        raise ArgumentError("species labels must be hashable") from None
    lut = np.empty(len(distinct), np.int32)
    for u, s in enumerate(distinct):
        k = lookup(s)
        if k is None:
            return None, s
        lut[u] = k
    return lut[codes], None


class System:
    def __init__(self, positions, cell, pbc, species):
        X = np.array(positions, dtype=np.float64, copy=True)
        if X.ndim != 2 or X.shape[1] != 3:
            raise DimensionMismatch("positions must be an (n, 3) array")
        H = np.array(cell, dtype=np.float64, copy=True)
        if H.shape != (3, 3):
            raise DimensionMismatch("cell must be 3x3 with lattice vectors in columns")
        if len(pbc) != 3:
            raise DimensionMismatch("pbc must contain three flags")
        species = species if isinstance(species, Labels) else Labels(species)
        if len(species) != len(X):
            raise DimensionMismatch("one species label per atom")
        self.positions = X
        self.cell = H
        self.pbc = tuple(bool(p) for p in pbc)
        self.labels = species

    def __len__(self):
        return len(self.positions)

    def copy(self):
        return System(self.positions, self.cell, self.pbc, self.labels)

    def position(self, i):
        return self.positions[i]

    def cell_vectors(self):
        return tuple(self.cell[:, a] for a in range(3))

    def periodicity(self):
        return self.pbc

    def species(self, i):
        return self.labels[i]


def channels(labels, known):
    """Map species labels onto model channels (bind_species,
    src/submodules/PairPotentials/models.jl:50-58). Unknown labels raise ArgumentError."""
    index = {s: k for k, s in enumerate(known)}
    out, bad = map_labels(labels, index.get)
    if out is None:
        raise ArgumentError(f"potential has no parameters for species {bad}")
    return out


def unique_in_order(labels):
    labels = labels if isinstance(labels, Labels) else Labels(labels)
    return list(labels.codes()[0])


def validate_geometry(sys):
    if not np.all(np.isfinite(sys.positions)):
        raise ArgumentError("positions must be finite")
    for a in range(3):
        if sys.pbc[a] and not np.all(np.isfinite(sys.cell[:, a])):
            raise ArgumentError("periodic cell vectors must be finite")
