"""Seeded synthetic workloads (geometry only; no files are read).

* ``argon_fcc``: FCC lattice a = 5.26 Å, m^3 conventional cells, 4 m^3 atoms in
  a cubic periodic box, Gaussian displacements (the 4,000-atom Lennard-Jones case
  at m = 10: eps 0.0104 eV, sigma 3.40 Å, cutoff 8.5 Å, energy shifted).
* ``argon_fcc_rattled``: the reference's deterministic argon benchmark cell
  (4,000 atoms at m = 10; its CPU timings are quoted for this geometry).
* ``b2_triclinic``: two-species B2 lattice a = 3.0 Å (species 0 at the origin,
  1 at the body centre) on nx x ny x nz conventional cells, lattice vectors
  sheared by seeded uniform +/-5 % off-diagonal strains, Gaussian displacements
  0.08 Å (the 256,000-atom SmoothPair case at 40 x 40 x 80, cutoff 6.0 Å,
  amplitudes [1.0 0.7; 0.7 1.3] eV).
* ``bcc_binary``: BCC a = 3.2 Å on nx x ny x nz cells with iid 50/50 species and
  Gaussian displacements 0.05 Å (the 32,768-atom ACE case at 32 x 32 x 16).
* ``fitting_set``: ncfg independent periodic configurations with atom counts
  log-uniform in [nmin, nmax], each a randomly sheared cubic cell at
  0.085 atoms/Å^3 with uniformly placed atoms rejected below 1.8 Å from any
  periodic image of an accepted atom, 2 species 50/50 (the 1,024-configuration
  fitting-set case).
* ``fixture``: the reference tests' deterministic scatter (test/fixtures.jl:17-27).

Positions are (n, 3) arrays in Å; cells have lattice vectors in columns.
"""
import numpy as np


def argon_fcc(m=10, a=5.26, sigma=0.05, seed=1):
    rng = np.random.default_rng(seed)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 3)
    frac = (g[:, None, :] + basis[None, :, :]).reshape(-1, 3)
    X = frac * a + rng.normal(0.0, sigma, size=frac.shape)
    H = np.eye(3) * (m * a)
    return X, H, (True, True, True), np.zeros(len(X), np.int32)


def argon_fcc_rattled(m=10, amplitude=0.04, a=5.26):
    """The reference's deterministic argon benchmark fixture
    (benchmark/lennard_jones/common.jl:65-82): FCC a = 5.26 Å with ix fastest,
    then iy, iz, and the basis (0,0,0), (0,1/2,1/2), (1/2,0,1/2), (1/2,1/2,0);
    atom i (1-based) is displaced by amplitude * sin(i (0.37 + k)) along axis
    k = 1, 2, 3 and wrapped into the periodic box of side m a."""
    basis = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]])
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1)
    g = g.transpose(2, 1, 0, 3).reshape(-1, 3)  # (ix, iy, iz) with ix fastest
    frac = (g[:, None, :] + basis[None, :, :]).reshape(-1, 3)
    i = np.arange(1, len(frac) + 1, dtype=float)[:, None]
    k = np.arange(1, 4, dtype=float)[None, :]
    L = m * a
    X = np.mod(a * frac + amplitude * np.sin(i * (0.37 + k)), L)
    return X, np.eye(3) * L, (True, True, True), np.zeros(len(X), np.int32)


def b2_triclinic(nx=40, ny=40, nz=80, a=3.0, shear=0.05, sigma=0.08, seed=2):
    rng = np.random.default_rng(seed)
    L = np.array([nx, ny, nz], float) * a
    S = rng.uniform(-shear, shear, size=(3, 3))
    np.fill_diagonal(S, 0.0)
    H = (np.eye(3) + S) * L[None, :]  # column b = L_b (e_b + sum_a S_ab e_a)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0.5]])
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    cells = (g[:, None, :] + basis[None, :, :]).reshape(-1, 3)
    frac = cells / np.array([nx, ny, nz], float)
    X = frac @ H.T + rng.normal(0.0, sigma, size=frac.shape)
    species = np.tile(np.array([0, 1], np.int32), len(g))
    return X, H, (True, True, True), species


def bcc_binary(nx=32, ny=32, nz=16, a=3.2, sigma=0.05, seed=3):
    rng = np.random.default_rng(seed)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0.5]])
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    frac = (g[:, None, :] + basis[None, :, :]).reshape(-1, 3)
    X = frac * a + rng.normal(0.0, sigma, size=frac.shape)
    H = np.diag(np.array([nx, ny, nz], float) * a)
    species = rng.integers(0, 2, size=len(X)).astype(np.int32)
    return X, H, (True, True, True), species


def fitting_set(ncfg=1024, nmin=8, nmax=2048, density=0.085, rmin=1.8, shear=0.05, seed=4, max_attempts=500):
    """List of (X, H, pbc, species) configurations and the total placement attempts.

    Generated by libpfw_workloads.so (csrc/workloads.cpp): random sequential
    addition with a bin grid; raises when a configuration needs more than
    max_attempts placement attempts per atom (saturation)."""
    import ctypes
    import os

    so = ctypes.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpfw_workloads.so"))
    f = so.pfw_wl_fitting_set
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                  ctypes.c_uint64, ctypes.c_int64] + [ctypes.c_void_p] * 5
    off = np.zeros(ncfg + 1, np.int64)
    rc = f(ncfg, nmin, nmax, density, rmin, shear, seed, max_attempts, off.ctypes.data, None, None, None, None)
    if rc:
        raise ValueError("invalid fitting-set parameters")
    n = int(off[-1])
    X = np.zeros((n, 3))
    sp = np.zeros(n, np.int32)
    H = np.zeros((ncfg, 3, 3))
    att = np.zeros(1, np.int64)
    rc = f(ncfg, nmin, nmax, density, rmin, shear, seed, max_attempts, off.ctypes.data, X.ctypes.data, sp.ctypes.data,
           H.ctypes.data, att.ctypes.data)
    if rc == 2:
        raise RuntimeError("random placement saturated")
    if rc:
        raise ValueError("invalid fitting-set parameters")
    out = [(X[off[c]:off[c + 1]].copy(), H[c].T.copy(), (True, True, True), sp[off[c]:off[c + 1]].copy())
           for c in range(ncfg)]
    return out, int(att[0])


def fixture(n, L=9.0, pbc=(True, True, True), cluster=False, mixed=True):
    i = np.arange(1, n + 1)[:, None]
    a = np.arange(1, 4)[None, :]
    X = L * np.mod(np.sin(i * (a + 0.71)) * 437.13, 1.0)
    if cluster:
        X[: (n + 1) // 2] *= 0.18
    H = np.eye(3) * L
    species = np.array([1 if (mixed and k % 2 == 1) else 0 for k in range(1, n + 1)], np.int32)
    return X.astype(float), H, tuple(pbc), species


# The reference's general-cell test matrices (test/periodic_images.jl:63-69), columns = lattice vectors.
IMAGE_TEST_CELLS = [
    np.array([[7., 2.5, -1.], [0., 6., 1.5], [0., 0., 5.]]),
    np.array([[4., 3.8, 0.], [0., 0.6, 0.], [0., 0., 4.]]),
    np.array([[1.2, 0.4, 0.], [0., 1.3, 0.3], [0., 0., 1.1]]),
    np.array([[2., 0, 0], [0, 2., 0], [0, 0, 2.]]),
    np.array([[0., -5., 0.], [6., 0., 1.], [0., 0., 7.]]),
]


def image_test_system(H, n=7):
    frac = np.array([[np.mod(np.sin((i + 1) * (a + 1 + 0.37)) * 31.3, 1.0) for a in range(3)] for i in range(n)])
    X = frac @ np.asarray(H).T
    species = np.array([k % 2 for k in range(n)], np.int32)
    return X, species
