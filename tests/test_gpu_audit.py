"""The neighbour sets the evaluating kernels actually accumulate, bit for bit.

The export path (``pfw_export_pairs``) is a separate FP64 enumeration; the
kernels that sum energies and forces decide membership their own way: the
fresh-cell pair kernel by an FP32 (mixed) or FP64 test with an exact FP64
re-decision inside a band around the cutoff, the Verlet list kernel by the
current FP64 displacement, the ACE environment builder by an FP32 prefilter
and the exact test, the ACE skin path by compaction of the skin rows. With
the pair-set audit on (``pfw_set_audit``) each of them records, per centre
atom, the count and an order-independent 64-bit digest of the (j, S) it
accumulated; those must equal the digests of the oracle's directed pair set
(``nl_pairs``, the restatement of visit_directed_images!,
src/submodules/NeighbourList/interface.jl:116-143) exactly.

SmoothPair and every ACE envelope vanish at the cutoff, so a pair dropped or
added at r ~ rc passes any E/F/V tolerance; only this comparison sees it. Each
case also reports the smallest | |R| - rc | over all images within rc + 0.05,
so that an exact match is not a matter of luck.
"""
import os

import numpy as np
import pytest

from oracle import nl_pairs
from oracle.ace import read_frames
from tests.test_gpu_pair import systems as reference_cells

pytestmark = pytest.mark.gpu

A = np.array([[1.0, 0.7], [0.7, 1.3]])
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PRECS = ("float64", "mixed")


def oracle_digest(pfw, X, H, pbc, rc, offset=0):
    """Per-atom digests of the oracle's directed pairs of one configuration;
    within a batch the neighbour index j is global (``offset`` + local)."""
    i, j, S, R, _ = nl_pairs(X, H, pbc, rc)
    return pfw.pair_set_digest(len(X), np.asarray(i, np.int64), np.asarray(j, np.int64) + offset, S)


def margin(X, H, pbc, rc, delta=0.05):
    """Smallest | |R| - rc | over every directed image within rc + delta (Å)."""
    _, _, _, R, _ = nl_pairs(X, H, pbc, rc + delta)
    if len(R) == 0:
        return float("inf")
    return float(np.abs(np.sqrt((np.asarray(R) ** 2).sum(1)) - rc).min())


def assert_same(got, ref, what):
    c, h = got
    rc_, rh = ref
    bad = np.flatnonzero((c != rc_) | (h != rh))
    assert bad.size == 0, (f"{what}: {bad.size} atoms differ, first {bad[:5].tolist()}: counts "
                           f"{c[bad[:5]].tolist()} vs oracle {rc_[bad[:5]].tolist()}")
    assert int(c.sum()) == int(rc_.sum())


def pair_step(pfw, w, X, H, pbc, sp, rc, kind="smooth"):
    if kind == "smooth":
        p = pfw.SmoothPair(["H", "He"], A, rc)
        labels = ["H" if s == 0 else "He" for s in sp]
    else:
        p = pfw.lennard_jones(epsilon=0.0104, sigma=3.40, radius=rc)
        labels = ["Ar"] * len(X)
    pfw.evaluate_(w, pfw.System(X, H, pbc, labels), p)
    return w.read_audit(len(X))


@pytest.mark.parametrize("prec", PRECS)
def test_pair_kernel_reference_cells(pfw, prec):
    """The reference's five general cells x three pbc settings, the clustered
    fixtures (open, slab, periodic), a two-cell axis and a sparse cluster."""
    w = pfw.Workspace(precision=prec)
    w.set_audit(True)
    for name, X, H, pbc, sp in reference_cells(pfw):
        rc = 6.0 if name == "b2-small" else 2.0
        got = pair_step(pfw, w, X, H, pbc, sp, rc)
        assert_same(got, oracle_digest(pfw, X, H, pbc, rc), f"{name} {prec}")


@pytest.mark.parametrize("prec", PRECS)
def test_pair_kernel_argon_4000_lj(pfw, prec):
    X, H, pbc, sp = pfw.workloads.argon_fcc()
    w = pfw.Workspace(precision=prec)
    w.set_audit(True)
    got = pair_step(pfw, w, X, H, pbc, sp, 8.5, kind="lj")
    assert_same(got, oracle_digest(pfw, X, H, pbc, 8.5), f"argon 4000 {prec}")
    print(f"argon FCC 4,000 LJ {prec}: {int(got[0].sum())} directed pairs, min | |R| - rc | = "
          f"{margin(X, H, pbc, 8.5):.3e} A")


@pytest.mark.parametrize("prec", PRECS)
def test_pair_kernel_b2_256k(pfw, prec):
    X, H, pbc, sp = pfw.workloads.b2_triclinic()
    w = pfw.Workspace(precision=prec)
    w.set_audit(True)
    got = pair_step(pfw, w, X, H, pbc, sp, 6.0)
    assert_same(got, oracle_digest(pfw, X, H, pbc, 6.0), f"B2 256k {prec}")
    print(f"B2 256,000 SmoothPair {prec}: {int(got[0].sum())} directed pairs, min | |R| - rc | = "
          f"{margin(X, H, pbc, 6.0):.3e} A")


@pytest.mark.parametrize("prec", PRECS)
def test_verlet_list_kernel_trajectories(pfw, prec):
    """k_pair_list on a random walk in a sheared B2 cell and in the reference's
    general cells (ghosts of ghosts), with list reuse and rebuilds."""
    rng = np.random.default_rng(7)
    cases = [("b2", *pfw.workloads.b2_triclinic(5, 5, 6, seed=9), 6.0)]
    for k, H in enumerate(pfw.workloads.IMAGE_TEST_CELLS):
        X, sp = pfw.workloads.image_test_system(H)
        cases.append((f"cell{k}", X, H, (True, True, True), sp, 2.0))
    for name, X0, H, pbc, sp, rc in cases:
        w = pfw.Workspace(precision=prec, skin=0.4)
        w.set_audit(True)
        X = X0.copy()
        for step in range(8):
            got = pair_step(pfw, w, X, H, pbc, sp, rc)
            assert_same(got, oracle_digest(pfw, X, H, pbc, rc), f"verlet {name} step {step} {prec}")
            X = X + rng.normal(0, 0.03, X.shape)
        assert w.verlet_info()["reuses"] > 0


@pytest.mark.parametrize("prec", PRECS)
def test_batched_pair_kernel_fitting_set_sample(pfw, prec):
    cfgs, _ = pfw.workloads.fitting_set()
    pick = np.sort(np.random.default_rng(5).choice(len(cfgs), 48, replace=False))
    sample = [cfgs[c] for c in pick]
    p = pfw.SmoothPair(["A", "B"], A, 6.0)
    w = pfw.Workspace(precision=prec)
    w.set_audit(True)
    pfw.evaluate_batch(w, [pfw.System(X, H, pbc, ["A" if s == 0 else "B" for s in sp]) for X, H, pbc, sp in sample], p)
    n = sum(len(c[0]) for c in sample)
    got = w.read_audit(n)
    off = np.cumsum([0] + [len(c[0]) for c in sample])
    ref_c, ref_h = np.zeros(n, np.int32), np.zeros(n, np.uint64)
    for k, (X, H, pbc, sp) in enumerate(sample):
        c, h = oracle_digest(pfw, X, H, pbc, 6.0, offset=off[k])
        ref_c[off[k]:off[k + 1]] = c
        ref_h[off[k]:off[k + 1]] = h
    assert_same(got, (ref_c, ref_h), f"fitting-set sample {prec}")


# ---------------------------------------------------------------- ACE

@pytest.fixture(scope="module")
def synthetic(pfw):
    model, ps, _ = pfw.load_resolved(pfw.ace_basis.synthetic_model())
    return model, ps


def ace_labels(model, sp):
    from pfw_amd.ace import ELEMENTS
    return [ELEMENTS[model.atomic_numbers[s] - 1] for s in sp]


@pytest.mark.parametrize("prec", PRECS)
def test_ace_environments_bcc_32k(pfw, synthetic, prec):
    model, ps = synthetic
    X, H, pbc, sp = pfw.workloads.bcc_binary()
    calc = pfw.ACECalculator(model, ps, precision=prec)
    calc.workspace.set_audit(True)
    calc.evaluate(pfw.System(X, H, pbc, ace_labels(model, sp)))
    got = calc.workspace.read_audit(len(X))
    rc = calc.cutoff()
    assert_same(got, oracle_digest(pfw, X, H, pbc, rc), f"ACE BCC 32k {prec}")
    print(f"ACE BCC 32,768 {prec}: {int(got[0].sum())} edges, min | |R| - rc | = {margin(X, H, pbc, rc):.3e} A")


def test_ace_environments_tial_frames(pfw):
    """The 33 TiAl frames: 2-atom cells of 2.65 Å with rc 5.5 Å (heavy
    multi-image and self-image enumeration) and 54-atom skew cells."""
    model, ps, _ = pfw.load_resolved(os.path.join(GOLD, "resolved_model_v1.toml"))
    calc = pfw.ACECalculator(model, ps, precision="mixed")
    calc.workspace.set_audit(True)
    for k, f in enumerate(read_frames(os.path.join(GOLD, "reference_results.xyz"))):
        calc.evaluate(pfw.System(f["X"], f["H"], f["pbc"], f["labels"]))
        got = calc.workspace.read_audit(len(f["X"]))
        assert_same(got, oracle_digest(pfw, f["X"], f["H"], f["pbc"], calc.cutoff()), f"TiAl frame {k}")


@pytest.mark.parametrize("prec", PRECS)
def test_ace_skin_compaction_walk(pfw, synthetic, prec):
    """ACE environments kept at cutoff + skin and compacted every step."""
    model, ps = synthetic
    X, H, pbc, sp = pfw.workloads.bcc_binary(4, 4, 3, seed=8)
    calc = pfw.ACECalculator(model, ps, precision=prec, skin=0.6)
    calc.workspace.set_audit(True)
    rng = np.random.default_rng(3)
    for step in range(8):
        calc.evaluate(pfw.System(X, H, pbc, ace_labels(model, sp)))
        got = calc.workspace.read_audit(len(X))
        assert_same(got, oracle_digest(pfw, X, H, pbc, calc.cutoff()), f"ACE skin step {step} {prec}")
        X = X + rng.normal(0, 0.05, X.shape)
    assert calc.workspace.verlet_info()["reuses"] > 0


def test_ace_batched_fitting_set_sample(pfw, synthetic):
    model, ps = synthetic
    cfgs, _ = pfw.workloads.fitting_set()
    pick = np.sort(np.random.default_rng(6).choice(len(cfgs), 48, replace=False))
    sample = [cfgs[c] for c in pick]
    calc = pfw.ACECalculator(model, ps, precision="mixed")
    calc.workspace.set_audit(True)
    calc.evaluate_batch([pfw.System(X, H, pbc, ace_labels(model, sp)) for X, H, pbc, sp in sample])
    n = sum(len(c[0]) for c in sample)
    got = calc.workspace.read_audit(n)
    off = np.cumsum([0] + [len(c[0]) for c in sample])
    ref_c, ref_h = np.zeros(n, np.int32), np.zeros(n, np.uint64)
    for k, (X, H, pbc, sp) in enumerate(sample):
        c, h = oracle_digest(pfw, X, H, pbc, calc.cutoff(), offset=off[k])
        ref_c[off[k]:off[k + 1]] = c
        ref_h[off[k]:off[k + 1]] = h
    assert_same(got, (ref_c, ref_h), "ACE fitting-set sample")


def test_audit_is_stale_until_a_step(pfw):
    from pfw_amd._capi import StaleError
    X, H, pbc, sp = pfw.workloads.fixture(9)
    w = pfw.Workspace(precision="mixed")
    w.set_audit(True)
    with pytest.raises(StaleError):
        w.read_audit(len(X))
    got = pair_step(pfw, w, X, H, pbc, sp, 2.0)
    assert_same(got, oracle_digest(pfw, X, H, pbc, 2.0), "fixture")
