"""Throughput benchmark of the MI355X force-evaluation hot path.

Headline workload (one GPU): the 256,000-atom two-species SmoothPair B2 lattice
in a sheared triclinic cell (cutoff 6.0 Å, amplitudes [1.0 0.7; 0.7 1.3] eV).
One "step" = fresh cell build + E+F+V + FP64 reductions + nonfinite status
read, positions already resident in HBM. ``value`` is the mixed-precision
throughput in Matom-evals/s. Reported beside it: the plain-FP64 throughput of
the same step, the 32,768-atom ACE workload (order-3, lmax-4, degree-12
synthetic basis on a BCC binary, cutoff 5.0 Å) in both precisions, per-stage
device times, the roofline of each dominant kernel, accuracy against the FP64
oracle, a bit-exact audit of the neighbour sets the evaluating kernels
accumulated, and the CPU baseline (the oracle's threaded FP64 restatement, on
all physical cores of the process's affinity up to 16 and on one core, with
the host described). The 4,000-atom argon Lennard-Jones cell, the
1,024-configuration fitting set (batched pair and ACE steps and the parameter
adjoints of a least-squares loss over it), the fitted May30 model, the
equivariant features and the Verlet-reuse trajectories follow.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-ace] [--no-batch]
"""
import argparse
import ctypes
import importlib.util
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "potentialframework.jl_amd")

METRIC = "Matom-evals/s (E+F+V), 256k-atom pair & 32k-atom ACE; mixed vs FP64 at tol"
PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 (vector/matrix) peak
PEAK_FP64_TFLOPS = 78.6
# FP32 VALU issue rate without packed (two-wide) FP32, which these kernels are
# built without: one FMA per lane per cycle, the FP64 vector rate
PEAK_FP32_UNPACKED_TFLOPS = 78.6
CPU_THREAD_CAP = 16  # the CPU share of one GPU on the measurement host
PEAK_HBM_GBS = 8000.0  # MI355X HBM3E
STAGES = ("binning", "scan", "scatter_sort", "environments", "interaction", "reduction")
# Algorithmic flops per directed in-cutoff pair (FMA = 2): d + shift 6, r^2 5,
# compare 1, q 2, a q^3 3, g 2, force 6, 1/2 E 2, 1/2 W 18.
FLOPS_PER_PAIR = {"smooth": 45.0, "lj": 50.0}


def load_package():
    spec = importlib.util.spec_from_file_location("pfw_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pfw_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def ace_flops(model, edges, sites):
    """Algorithmic ACE flops of one E+F+V evaluation (FMA = 2).
    Per edge: geometry 9, radial values+derivatives 12 per function (many-body
    and pair), Y_lm with Cartesian gradients 32 per harmonic, density 2 per
    density, pullback 4 per density + 8 per harmonic. Per site: each product
    of order nu costs 2 nu (value and readout) + 2 nu^2 (reverse)."""
    ny = (model.lmax + 1) ** 2
    nd = len(model.density)
    per_edge = 9 + 12 * model.radial.nr + 12 * model.pair.nr + 32 * ny + 2 * nd + 4 * nd + 8 * ny
    per_site = sum(2 * len(p) + 2 * len(p) ** 2 for p in model.products)
    return per_edge * edges + per_site * sites, per_edge, per_site


def time_steps(lib, h, step, steps, warmup, mode):
    """Wall time of `steps` device steps (each ends with the status read), with
    the per-stage event timing off; then the mean stage ms of a second, shorter
    pass with HIP-event timing on."""
    from pfw_amd import _capi

    for _ in range(warmup):
        _capi.check(h, step(h, 7, mode))
    _capi.check(h, lib.pfw_synchronize(h))
    t0 = time.perf_counter()
    for _ in range(steps):
        _capi.check(h, step(h, 7, mode))
    _capi.check(h, lib.pfw_synchronize(h))
    t1 = time.perf_counter()
    _capi.check(h, lib.pfw_set_timing(h, 1))
    stage = np.zeros(len(STAGES))
    buf = np.zeros(len(STAGES))
    reps = max(1, min(steps, 10))
    for _ in range(reps):
        _capi.check(h, step(h, 7, mode))
        lib.pfw_get_timings(h, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(STAGES))
        stage += buf
    _capi.check(h, lib.pfw_set_timing(h, 0))
    return t1 - t0, stage / reps


def graph_steps(lib, _capi, w, step, steps, warmup, mode):
    """Seconds of `steps` steps replayed from a captured graph (one eager step
    sizes the buffers, the next captures); graphs are switched off after."""
    h = w.ctx.h
    w.set_graphs(True)
    for _ in range(2 + warmup):
        _capi.check(h, step(h, 7, mode))
    _capi.check(h, lib.pfw_synchronize(h))
    t0 = time.perf_counter()
    for _ in range(steps):
        _capi.check(h, step(h, 7, mode))
    _capi.check(h, lib.pfw_synchronize(h))
    secs = time.perf_counter() - t0
    info = w.graph_info()
    w.set_graphs(False)
    return secs, info


def read_traffic(name):
    """HBM bytes per launch of a kernel from its committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, calibrated against a copy kernel), if present."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    except Exception:
        return None


def host_info():
    """The CPU the baseline ran on: model, sockets, SMT, the logical CPUs and
    physical cores of this process's affinity mask."""
    info = {"model": None, "sockets": None, "threads_per_core": None}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        info["model"] = kv.get("Model name")
        info["sockets"] = int(kv["Socket(s)"]) if kv.get("Socket(s)", "").isdigit() else None
        tpc = kv.get("Thread(s) per core", "")
        info["threads_per_core"] = int(tpc) if tpc.isdigit() else None
    except Exception:
        pass
    if info["model"] is None:
        try:
            with open("/proc/cpuinfo") as f:
                info["model"] = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
        except Exception:
            pass
    aff = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in aff:
        try:
            base = f"/sys/devices/system/cpu/cpu{c}/topology/"
            with open(base + "physical_package_id") as f1, open(base + "core_id") as f2:
                cores.add((f1.read().strip(), f2.read().strip()))
        except OSError:
            cores.add(("?", str(c)))
    ranges, start = [], None
    for k, c in enumerate(aff):
        if start is None:
            start = c
        if k + 1 == len(aff) or aff[k + 1] != c + 1:
            ranges.append(f"{start}-{c}" if c > start else str(start))
            start = None
    info.update(affinity_cpus=len(aff), affinity_physical_cores=len(cores), affinity=",".join(ranges),
                smt_in_affinity=len(aff) > len(cores), pinning="none (threads float within the affinity mask)")
    return info


def baseline_threads(host):
    return max(1, min(CPU_THREAD_CAP, host["affinity_physical_cores"]))


def cpu_timing(fn, warmups=3, samples=7, budget=20.0):
    """Wall times of a CPU evaluation: warm-ups, then up to `samples` timed
    calls while the budget (seconds, warm-ups included) lasts; at least one."""
    t_start = time.perf_counter()
    out = None
    for _ in range(warmups):
        out = fn()
        if time.perf_counter() - t_start > budget / 3:
            break
    times = []
    while len(times) < samples and (not times or time.perf_counter() - t_start < budget):
        t0 = time.perf_counter()
        out = fn()
        times.append(time.perf_counter() - t0)
    return np.array(times), out


def cpu_baseline(make, atoms, what, threads, host, all_kind="", single_kind=""):
    """The oracle's FP64 restatement timed on `threads` cores (3 warm-ups,
    median of 7) and on one core (1 warm-up, median of up to 3 within 20 s).
    `make(t)` returns a callable evaluating with t threads. Returns the
    cpu_baseline object and the all-core result."""
    times, out = cpu_timing(make(threads))
    one, _ = cpu_timing(make(1), warmups=1, samples=3, budget=20.0)
    med, med1 = float(np.median(times)), float(np.median(one))
    return {"value": round(atoms / med / 1e6, 5), "unit": "Matom-evals/s", "cores": threads, "kind": "port",
            "sample": f"{what}; {threads} threads{all_kind}: median of {len(times)} after 3 warm-ups",
            "median_ms": round(1e3 * med, 3), "min_ms": round(1e3 * float(times.min()), 3),
            "max_ms": round(1e3 * float(times.max()), 3),
            "single_core": {"value": round(atoms / med1 / 1e6, 5), "median_ms": round(1e3 * med1, 3),
                            "samples": len(one), "schedule": single_kind or "serial"},
            "host": host}, out


def neighbour_audit(pfw, w, run, X, H, pbc, rc, delta=0.05):
    """One audited step: per-atom count and digest of the directed pairs (j, S)
    the evaluating kernel accumulated, compared bit-exactly with the oracle's
    neighbour list; and the smallest | |R| - rc | over all images within
    rc + delta (how close the workload comes to the membership boundary)."""
    from oracle.core import nl_pairs

    n = len(X)
    w.set_audit(True)
    run()
    got = w.read_audit(n)
    w.set_audit(False)
    i, j, S, _, _ = nl_pairs(X, H, pbc, rc)
    ref = pfw.pair_set_digest(n, i, j, S)
    bad = int(np.count_nonzero((got[0] != ref[0]) | (got[1] != ref[1])))
    _, _, _, R, _ = nl_pairs(X, H, pbc, rc + delta)
    margin = float(np.abs(np.sqrt((R * R).sum(1)) - rc).min()) if len(R) else None
    return {"exact": bad == 0, "atoms_differing": bad, "directed_pairs": int(got[0].sum()),
            "min_abs_r_minus_rc_A": margin}


def accuracy(res, E, F, W, n):
    d = res["forces"].T - F
    return {"force_max": float(np.abs(d).max()), "force_rms": float(np.sqrt(np.mean(d * d))),
            "energy_per_atom": float(abs(res["energy"] - E) / n), "virial_max": float(np.abs(res["virial"] - W).max())}


def pooled_accuracy(pairs):
    """Accuracy over several structures: maxima, the force RMS pooled over all
    components, and the largest per-structure energy-per-atom error."""
    fmax = emax = wmax = 0.0
    ss = cnt = 0
    for res, (E, F, W) in pairs:
        d = res["forces"].T - F
        fmax = max(fmax, float(np.abs(d).max()) if d.size else 0.0)
        ss += float((d * d).sum())
        cnt += d.size
        emax = max(emax, abs(res["energy"] - E) / max(1, F.shape[0]))
        wmax = max(wmax, float(np.abs(res["virial"] - W).max()))
    return {"force_max": fmax, "force_rms": float(np.sqrt(ss / max(1, cnt))), "energy_per_atom": emax,
            "virial_max": wmax, "structures": len(pairs)}


def bench_fitting_set(pfw, lib, _capi, model, ps, pair_model, args):
    """1,024 configurations (8..2,048 atoms, log-uniform) packed in one cell list:
    batched SmoothPair and ACE steps in both precisions, and the parameter
    adjoints of the loss sum (E - E*)^2 / N^2 + 0.1 sum |F - F*|^2 with targets
    from WB* = WB (1 + 0.01 N(0, 1))."""
    from pfw_amd.ace import ELEMENTS

    t0 = time.perf_counter()
    cfgs, attempts = pfw.workloads.fitting_set()
    gen_s = time.perf_counter() - t0
    natoms = sum(len(c[0]) for c in cfgs)
    lab = [ELEMENTS[z - 1] for z in model.atomic_numbers]
    sys_pair = [pfw.System(X, H, pbc, ["A" if s == 0 else "B" for s in sp]) for X, H, pbc, sp in cfgs]
    sys_ace = [pfw.System(X, H, pbc, [lab[s] for s in sp]) for X, H, pbc, sp in cfgs]
    K, Wm = args.batch_steps, max(1, args.warmup // 2)
    out = {"workload": "fitting_set_1024", "configurations": len(cfgs), "atoms": natoms,
           "placement_attempts": attempts, "generation_s": round(gen_s, 3), "unit": "Matom-evals/s"}
    results = {}
    for prec, mode in (("mixed", _capi.PFW_MIXED), ("float64", _capi.PFW_FP64)):
        w = pfw.Workspace(precision=prec)
        results[("pair", prec)] = pfw.evaluate_batch(w, sys_pair, pair_model)
        secs, stage = time_steps(lib, w.ctx.h, lib.pfw_pair_step, K, Wm, mode)
        out[f"pair_{prec}"] = {"value": round(natoms * K / secs / 1e6, 3), "ms_per_step": round(1e3 * secs / K, 4),
                               "stages_ms": {k: round(float(v), 5) for k, v in zip(STAGES, stage)}}
        gsecs, _ = graph_steps(lib, _capi, w, lib.pfw_pair_step, K, Wm, mode)
        out[f"pair_{prec}"]["graph"] = {"value": round(natoms * K / gsecs / 1e6, 3),
                                        "ms_per_step": round(1e3 * gsecs / K, 4)}
        calc = pfw.ACECalculator(model, ps, precision=prec)
        results[("ace", prec)] = calc.evaluate_batch(sys_ace)
        secs, stage = time_steps(lib, calc.workspace.ctx.h, lib.pfw_ace_step, K, Wm, mode)
        out[f"ace_{prec}"] = {"value": round(natoms * K / secs / 1e6, 4), "ms_per_step": round(1e3 * secs / K, 4),
                              "stages_ms": {k: round(float(v), 5) for k, v in zip(STAGES, stage)}}
        gsecs, _ = graph_steps(lib, _capi, calc.workspace, lib.pfw_ace_step, K, Wm, mode)
        out[f"ace_{prec}"]["graph"] = {"value": round(natoms * K / gsecs / 1e6, 4),
                                       "ms_per_step": round(1e3 * gsecs / K, 4)}
    if not args.no_cpu_baseline:
        # every 21st configuration (49 of 1,024) against the oracle, one at a time
        from oracle.ace import ace_evaluate as oracle_ace
        from oracle.ace import load_resolved as oracle_load
        from oracle.core import KIND_SMOOTH, pair_evaluate

        om = oracle_load(pfw.ace_basis.synthetic_model())
        sample = list(range(0, len(cfgs), 21))
        ref_pair = [pair_evaluate(cfgs[c][0], cfgs[c][3], cfgs[c][1], cfgs[c][2], KIND_SMOOTH,
                                  pair_model.amplitudes, pair_model.cutoff()) for c in sample]
        ref_ace = [oracle_ace(cfgs[c][0], cfgs[c][3], cfgs[c][1], cfgs[c][2], om)[:3] for c in sample]
        out["accuracy_vs_fp64_oracle"] = {
            "sample": f"{len(sample)} configurations (every 21st), {sum(len(cfgs[c][0]) for c in sample)} atoms",
            **{f"{kind}_{prec}": pooled_accuracy([(results[(kind, prec)][c], ref[k]) for k, c in enumerate(sample)])
               for kind, ref in (("pair", ref_pair), ("ace", ref_ace)) for prec in ("mixed", "float64")}}
    # adjoints: loss and gradient with respect to all ACE readout coefficients
    rng = np.random.default_rng(5)
    WBt = model.WB * (1 + 0.01 * rng.normal(size=model.WB.shape))
    calc = pfw.ACECalculator(model, {"WB": WBt, "Wpair": model.Wpair}, precision="mixed")
    targets = calc.evaluate_batch(sys_ace)
    calc.set_parameters(ps)
    weights = {"energy": 1.0, "forces": 0.1, "virial": 0.0}

    def vjp(dE, dF, dW):
        return pfw.ace_parameter_vjp(calc, sys_ace, dE, dF, dW)

    L, g = pfw.fitting_loss(calc.evaluate_batch, vjp, sys_ace, targets, weights)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        pfw.fitting_loss(calc.evaluate_batch, vjp, sys_ace, targets, weights)
        times.append(time.perf_counter() - t0)
    # the adjoint entry point alone (cotangent upload, environments, adjoint kernel, gradient download)
    dE = rng.normal(size=len(cfgs))
    dF = rng.normal(size=(natoms, 3))
    gWB = np.zeros((model.nz, model.nbasis))
    gWp = np.zeros((model.nz, max(model.pair.nr, 1)))
    h = calc.workspace.ctx.h
    P = ctypes.POINTER(ctypes.c_double)
    adj = {}
    for prec, mode in (("mixed", _capi.PFW_MIXED), ("float64", _capi.PFW_FP64)):
        args_ = (h, mode, dE.ctypes.data_as(P), dF.ctypes.data_as(P), None, gWB.ctypes.data_as(P),
                 gWp.ctypes.data_as(P))
        _capi.check(h, lib.pfw_ace_adjoint(*args_))
        t0 = time.perf_counter()
        for _ in range(K):
            _capi.check(h, lib.pfw_ace_adjoint(*args_))
        adj[prec] = (time.perf_counter() - t0) / K
    wp = pfw.Workspace(precision="float64")
    pfw.pair_parameter_vjp(wp, sys_pair, pair_model, dE, [f.T for f in np.split(dF, np.cumsum(
        [len(c[0]) for c in cfgs])[:-1])])
    gA = np.zeros(4)
    hp = wp.ctx.h
    t0 = time.perf_counter()
    for _ in range(K):
        _capi.check(hp, lib.pfw_pair_adjoint(hp, dE.ctypes.data_as(P), dF.ctypes.data_as(P), None,
                                             gA.ctypes.data_as(P)))
    pair_adj = (time.perf_counter() - t0) / K
    out["ace_adjoint"] = {
        "loss": "sum (E - E*)^2 / N^2 + 0.1 sum |F - F*|^2, WB* = WB (1 + 0.01 N(0,1))",
        "parameters": int(model.WB.size + model.Wpair.size), "loss_value": L,
        "loss_and_gradient_ms": round(1e3 * float(np.median(times)), 3),
        "adjoint_ms": {k: round(1e3 * v, 3) for k, v in adj.items()},
        "adjoint_value": {k: round(natoms / v / 1e6, 4) for k, v in adj.items()},
        "gradient_norm": float(np.linalg.norm(g["WB"])),
    }
    out["pair_adjoint"] = {"precision": "float64", "adjoint_ms": round(1e3 * pair_adj, 3),
                           "adjoint_value": round(natoms / pair_adj / 1e6, 3)}
    # the linear design (energy, force and virial rows per readout coefficient)
    # of the first 64 configurations, one batch on the device
    sub = sys_ace[:64]
    design = {"configurations": len(sub), "atoms": sum(len(x) for x in sub)}
    for prec in ("mixed", "float64"):
        dcalc = pfw.ACECalculator(model, ps, precision=prec)
        pfw.design_matrix(dcalc, sub)
        t0 = time.perf_counter()
        d = pfw.design_matrix(dcalc, sub)
        secs = time.perf_counter() - t0
        design["rows"], design["columns"] = d["matrix"].shape
        design[prec] = {"ms": round(1e3 * secs, 2), "ms_per_column": round(1e3 * secs / d["matrix"].shape[1], 4)}
        del d
    out["design_matrix"] = design
    return out


def bench_verlet(pfw, lib, _capi, X, H, pbc, sp, model, labels, skin, prec, mode, steps=40, amp=0.005, ace=False,
                 check=None):
    """The reference's evolving-trajectory measurement (benchmark/gpu_pairs.jl:118-142):
    raw coordinates x_s = x_0 + s v with v[a, i] = amp sin(a + 0.73 i) (1-based a, i),
    40 steps on a list at cutoff + skin built at x_0; each step's positions are
    uploaded outside the timed region and every step is timed from the call to
    its synchronised status read. Also the stage times of a reused step. With
    `check(X) -> (E, F, W)` (the FP64 oracle), one more step at the last
    positions, on the kept list, is audited and compared with it."""
    n = len(X)
    i1 = np.arange(1, n + 1)[:, None]
    a1 = np.arange(1, 4)[None, :]
    vel = amp * np.sin(a1 + 0.73 * i1)
    sys0 = pfw.System(X, H, pbc, labels)
    if ace:  # model is (resolved model, parameters)
        calc = pfw.ACECalculator(model[0], model[1], precision=prec, skin=skin)
        w = calc.workspace
        calc.evaluate(sys0)
        calc.evaluate(sys0)
        stepfn = lib.pfw_ace_step
    else:
        w = pfw.Workspace(precision=prec, skin=skin)
        pfw.evaluate_(w, sys0, model)
        pfw.evaluate_(w, sys0, model)
        stepfn = lib.pfw_pair_step
    h = w.ctx.h
    P = ctypes.POINTER(ctypes.c_double)
    times, rebuilt = [], []
    for s in range(1, steps + 1):
        Xs = np.ascontiguousarray(X + s * vel)
        _capi.check(h, lib.pfw_update_positions(h, Xs.ctypes.data_as(P)))
        before = w.rebuilds
        t0 = time.perf_counter()
        _capi.check(h, stepfn(h, 7, mode))
        times.append(time.perf_counter() - t0)
        rebuilt.append(w.rebuilds > before)
    times = np.array(times)
    rebuilt = np.array(rebuilt)
    reuse = times[~rebuilt]
    info = w.verlet_info()
    # stage times of reused steps at the last positions
    _capi.check(h, lib.pfw_set_timing(h, 1))
    stage = np.zeros(len(STAGES))
    buf = np.zeros(len(STAGES))
    for _ in range(10):
        _capi.check(h, stepfn(h, 7, mode))
        lib.pfw_get_timings(h, buf.ctypes.data_as(P), len(STAGES))
        stage += buf
    _capi.check(h, lib.pfw_set_timing(h, 0))
    stage /= 10
    checked = {}
    if check is not None:
        sys_s = pfw.System(Xs, H, pbc, labels)
        before = w.rebuilds
        held = {}

        def run():
            held["r"] = calc.evaluate(sys_s) if ace else pfw.evaluate_(w, sys_s, model)

        audit = neighbour_audit(pfw, w, run, Xs, H, pbc, calc.cutoff() if ace else model.cutoff())
        audit["list_kept"] = w.rebuilds == before
        E, F, W = check(Xs)
        checked = {"neighbour_audit": audit, "accuracy_vs_fp64_oracle": accuracy(held["r"], E, F, W, n)}
    return {
        **checked,
        "atoms": n, "skin_A": skin, "precision": prec, "steps": steps, "step_amplitude_A": amp,
        "rebuild_steps": int(rebuilt.sum()), "amortized_ms": round(1e3 * float(times.mean()), 4),
        "value": round(n / float(times.mean()) / 1e6, 3), "unit": "Matom-evals/s",
        "reuse_median_ms": round(1e3 * float(np.median(reuse)), 4) if len(reuse) else None,
        "rebuild_ms": [round(1e3 * float(t), 4) for t in times[rebuilt]],
        **({} if ace else {"ghosts": info["ghosts"], "slots_per_atom": info["slots"], "list_entries": info["entries"]}),
        "reuse_stages_ms": {k: round(float(v), 5) for k, v in zip(STAGES, stage)},
    }


def bench_may30(pfw, lib, _capi, args, reps=5):
    """The reference's fitted May30 Ti-Al-V model (real solid harmonics to l = 6,
    5,652 products, 1,000-node splines) on its omega-240 fixture structure
    replicated 5 x 5 x 5 (30,000 atoms, nonorthogonal periodic cell), one
    fresh step per evaluation as for the synthetic ACE workload."""
    try:
        import tomllib as toml
    except ImportError:
        import tomli as toml
    from pfw_amd.ace import ELEMENTS
    gold = os.path.join(ROOT, "tests", "golden")
    model, ps, _ = pfw.load_resolved(os.path.join(gold, "may30_resolved.npz"))
    with open(os.path.join(gold, "may30_omega-240.toml"), "rb") as f:
        d = toml.load(f)
    n0 = len(d["species"])
    X0 = np.asarray(d["positions"], float).reshape(n0, 3)
    H0 = np.asarray(d["cell"], float).reshape(3, 3).T
    k = 5
    shifts = np.array([[a, b, c] for a in range(k) for b in range(k) for c in range(k)], float) @ H0.T
    X = (X0[None, :, :] + shifts[:, None, :]).reshape(-1, 3)
    H = H0 * k
    labels = [ELEMENTS[z - 1] for z in d["species"]] * len(shifts)
    sys_ = pfw.System(X, H, (True, True, True), labels)
    out = {"workload": "may30 omega-240 x 5^3", "atoms": len(X), "cutoff_A": model.cutoff,
           "basis": {"radial": model.radial.nr, "lmax": model.lmax, "densities": len(model.density),
                     "products": len(model.products), "invariants": model.nbasis, "angular": "real_solid"},
           "unit": "Matom-evals/s",
           "reference_cpu_context": {"atoms": 1024, "prepared_ms": 74.629, "threads": 1,
                                     "source": "benchmark/cpu_ace/MAY30-REPORT.md:8-10,133-139 (DAG, bcc-1024)"}}
    results = {}
    for prec, mode in (("mixed", _capi.PFW_MIXED), ("float64", _capi.PFW_FP64)):
        calc = pfw.ACECalculator(model, ps, precision=prec)
        results[prec] = calc.evaluate(sys_)
        secs, stage = time_steps(lib, calc.workspace.ctx.h, lib.pfw_ace_step, max(3, args.steps // 10),
                                 min(args.warmup, 2), mode)
        nsteps = max(3, args.steps // 10)
        out[prec] = {"value": round(len(X) * nsteps / secs / 1e6, 4), "ms_per_step": round(1e3 * secs / nsteps, 4),
                     "stages_ms": {k2: round(float(v), 5) for k2, v in zip(STAGES, stage)}}
        if prec == "mixed":
            edges = len(calc.workspace.export_pairs()[0])
            flops, per_edge, per_site = ace_flops(model, edges, len(X))
            kms = float(stage[4])
            out["edges"] = edges
            out["roofline"] = {"kernel": "k_ace_dens+k_ace_prod+k_ace_pull<float,6>", "bound": "valu",
                               "compute_unit": "VALU fp32", "achieved": round(flops / (kms * 1e-3) / 1e12, 3),
                               "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                               "frac": round(flops / (kms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
                               "frac_of_unpacked_issue": round(flops / (kms * 1e-3) / 1e12 / PEAK_FP32_UNPACKED_TFLOPS,
                                                               4),
                               "flops_per_edge": per_edge, "flops_per_site": per_site, "kernel_ms": round(kms, 5)}
            out["mixed_calc"] = calc
    if not args.no_cpu_baseline:
        from oracle.ace import ace_evaluate as oracle_ace
        from oracle.ace import load_resolved as oracle_load

        om = oracle_load(pfw.ace.read_resolved_npz(os.path.join(gold, "may30_resolved.npz")))
        sp = np.array([om["atomic_numbers"].index(z) for z in d["species"]] * len(shifts), np.int32)
        run = lambda: out["mixed_calc"].evaluate(sys_)  # noqa: E731
        out["neighbour_audit"] = neighbour_audit(pfw, out["mixed_calc"].workspace, run, X, H, (True,) * 3,
                                                 model.cutoff)
        E, F, W, _ = oracle_ace(X, sp, H, (True, True, True), om, threads=args.cpu_threads_used)
        out["accuracy_vs_fp64_oracle"] = {prec: accuracy(results[prec], E, F, W, len(X)) for prec in results}
    out.pop("mixed_calc", None)
    return out


def bench_features(pfw, args):
    """Per-atom equivariant features of the reference's feature fixture layer
    (test/ace_features.jl:16-27: Si/O, smooth exponential radials
    [(1,0),(1,1),(2,1),(1,2)] with cutoff 4 A, correlation 3, lmax 2, blocks
    0e x3, 1o x2, 1e x2, 3o x2, 2e x2) on a 32,768-atom periodic BCC cell
    (a = 3.2 A), FP64. The CPU baseline is the oracle's Python restatement on
    a 54-atom cell of the same lattice."""
    F = pfw.features
    spec = [(1, 0), (1, 1), (2, 1), (1, 2)]
    irreps = [dict(L=0, parity="even", channels=3), dict(L=1, parity="odd", channels=2),
              dict(L=1, parity="even", channels=2), dict(L=3, parity="odd", channels=2),
              dict(L=2, parity="even", channels=2)]
    layer = F.ACEFeatureLayer(["Si", "O"], F.SmoothExponential(spec, 4.0, [0, 0, 1, 0]), 3, 2, irreps)
    ps = F.feature_parameters(layer, custom_radial={"rate": [0.4], "amplitude": [1.3]})
    for b, key in enumerate(sorted(ps["mixing"], key=lambda k: int(k.split("_")[1]))):
        W = ps["mixing"][key]
        W[...] = np.sin(b + 1 + np.arange(1, W.size + 1)).reshape(W.shape, order="F")
    X, H, pbc, sp = pfw.workloads.bcc_binary()
    sys_ = pfw.System(X, H, pbc, ["Si" if s == 0 else "O" for s in sp])
    w = pfw.Workspace(precision="float64")
    pfw.ace_features(layer, sys_, ps, workspace=w)
    reps = max(3, args.steps // 10)
    t0 = time.perf_counter()
    for _ in range(reps):
        pfw.ace_features(layer, sys_, ps, workspace=w)
    secs = (time.perf_counter() - t0) / reps
    nnz = sum(int(np.count_nonzero(c)) for c in layer.coupling)
    out = {"workload": "feature fixture layer (correlation 3, lmax 2, L <= 3) on bcc_binary 32k",
           "atoms": len(X), "densities": len(layer.density), "products": len(layer.products),
           "coupling_nonzeros": nnz, "blocks": [(s.L, s.parity, s.channels) for s in layer.irreps],
           "value": round(len(X) / secs / 1e6, 4), "unit": "Matom-features/s", "ms_per_call": round(1e3 * secs, 3),
           "call": "fresh cell list + pair export on the device, feature kernel, outputs to host"}
    y = pfw.ace_features(layer, sys_, ps, workspace=w)
    cot = [np.sin(np.arange(1, b.size + 1)).reshape(b.shape, order="F") for b in y]
    pfw.feature_vjp(layer, sys_, ps, cot, workspace=w)
    t0 = time.perf_counter()
    for _ in range(reps):
        pfw.feature_vjp(layer, sys_, ps, cot, workspace=w)
    vsecs = (time.perf_counter() - t0) / reps
    out["pullback"] = {"value": round(len(X) / vsecs / 1e6, 4), "unit": "Matom-pullbacks/s",
                       "ms_per_call": round(1e3 * vsecs, 3),
                       "call": "cotangents to the device, fresh cell list + pair export, per-atom reverse kernel, "
                               "position gather, species sums of the parameter gradients, gradients to host"}
    # No CPU baseline: the feature oracle is a Python checker, not a compiled
    # restatement whose speed would mean anything.
    out["cpu_baseline"] = None
    if not args.no_cpu_baseline:
        from oracle.features import ace_features as oracle_features
        Xs, Hs, pbs, sps = pfw.workloads.bcc_binary(3, 3, 3, seed=3)
        T = dict(density=layer.density, products=layer.products, coupling=layer.coupling,
                 irreps=[(s.L, s.channels) for s in layer.irreps], lmax=layer.lmax, cutoff=layer.cutoff,
                 radial_cutoff=layer.radial.cutoff, powers=layer.radial.powers)
        P = dict(mixing=[ps["mixing"][f"block_{b + 1}"] for b in range(len(layer.irreps))],
                 scale=ps["radial"]["scale"], rate=0.4, amplitude=1.3)
        from pfw_amd.ace import ELEMENTS
        labs = ["Si" if v == 0 else "O" for v in sps]
        chans = np.array([layer.atomic_numbers.index(ELEMENTS.index(x) + 1) for x in labs])
        ref = oracle_features(Xs, chans, Hs, pbs, T, P)
        got = pfw.ace_features(layer, pfw.System(Xs, Hs, pbs, labs), ps)
        err = max(float(np.abs(np.asarray(g) - np.asarray(r)).max()) for g, r in zip(got, ref))
        scale = max(float(np.abs(np.asarray(r)).max()) for r in ref)
        out["accuracy_vs_fp64_oracle"] = {"sample": "54-atom cell of the same lattice (seed 3)",
                                          "feature_max_abs": err, "feature_max_magnitude": scale}
    return out


def bench_argon_4000(pfw, lib, _capi, args, host):
    """The reference's CPU comparison cell (benchmark/lennard_jones/common.jl:53-82,
    RESULTS.md:25-28): 4,000 argon atoms, FCC a = 5.26 Å with the deterministic
    0.04 Å rattle, cutoff 2.5 sigma = 8.5125 Å, eps 0.010323 eV, sigma 3.405 Å,
    here with the framework's energy-shifted Lennard-Jones (lennard_jones.jl:5-22;
    the benchmark's inverse-distance tilt costs the same per pair). One step =
    fresh cell build + E+F+V + FP64 reductions + status read."""
    X, H, pbc, sp = pfw.workloads.argon_fcc_rattled(10)
    n = len(X)
    eps, sigma = 0.010323, 3.405
    rc = 2.5 * sigma
    lj = pfw.lennard_jones(epsilon=eps, sigma=sigma, radius=rc)
    sys_ = pfw.System(X, H, pbc, ["Ar"] * n)
    steps = max(4 * args.steps, 200)
    out = {"workload": "argon_fcc_4000_lj", "atoms": n, "cutoff_A": rc, "epsilon_eV": eps, "sigma_A": sigma,
           "steps": steps, "unit": "Matom-evals/s",
           "reference_cpu": {"threads_1_ms": 5.270, "threads_8_ms": 1.551, "cpu": "AMD Ryzen 9 9950X",
                             "call": "complete public call rebuilding topology (W=4 phased executor)",
                             "source": "benchmark/lennard_jones/RESULTS.md:25-28"}}
    results, ws = {}, {}
    for prec, mode in (("mixed", _capi.PFW_MIXED), ("float64", _capi.PFW_FP64)):
        w = pfw.Workspace(precision=prec)
        results[prec] = pfw.evaluate_(w, sys_, lj)
        secs, stage = time_steps(lib, w.ctx.h, lib.pfw_pair_step, steps, args.warmup, mode)
        ws[prec] = w
        out[prec] = {"value": round(n * steps / secs / 1e6, 3), "ms_per_step": round(1e3 * secs / steps, 4),
                     "stages_ms": {k: round(float(v), 5) for k, v in zip(STAGES, stage)}}
        gsecs, _ = graph_steps(lib, _capi, w, lib.pfw_pair_step, steps, args.warmup, mode)
        out[prec]["graph"] = {"value": round(n * steps / gsecs / 1e6, 3), "ms_per_step": round(1e3 * gsecs / steps, 4)}
    npairs = len(ws["mixed"].export_pairs()[0])
    kms = out["mixed"]["stages_ms"]["interaction"]
    flops = npairs * FLOPS_PER_PAIR["lj"]
    out["roofline"] = {"kernel": "k_pair<float,LJ>", "bound": "valu", "compute_unit": "VALU fp32",
                       "achieved": round(flops / (kms * 1e-3) / 1e12, 3), "peak": PEAK_FP32_TFLOPS,
                       "unit": "TFLOP/s", "frac": round(flops / (kms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
                       "directed_pairs": npairs, "flops_per_pair": FLOPS_PER_PAIR["lj"], "kernel_ms": kms,
                       "cells": ws["mixed"].grid_info()["ncells"]}
    if not args.no_cpu_baseline:
        from oracle.core import KIND_LJ, pair_evaluate

        def make(t):
            return lambda: pair_evaluate(X, sp, H, pbc, KIND_LJ, (eps, sigma), rc, threads=t)

        out["cpu_baseline"], (E, F, W) = cpu_baseline(
            make, n, "full 4,000-atom E+F+V evaluations (cell build + phase-coloured half-list, FP64)",
            args.cpu_threads_used, host, " over phase-coloured cell tasks")
        out["accuracy_vs_fp64_oracle"] = {prec: accuracy(results[prec], E, F, W, n) for prec in results}
        out["neighbour_audit"] = {prec: neighbour_audit(pfw, ws[prec], lambda: pfw.evaluate_(ws[prec], sys_, lj),
                                                        X, H, pbc, rc) for prec in ws}
    return out


def profile_run(pfw, lib, _capi, which, prec, steps, density=0):
    """A short run for profilers: `steps` device steps of one workload after
    one set-up evaluation (pair: the 256k B2 SmoothPair step; ace: the 32k BCC
    ACE step; may30: the 30,000-atom May30 step; argon: the 4,000-atom LJ
    step), then one line naming what ran. `density` selects the ACE density
    contraction (the ace_density_contraction knob)."""
    mode = _capi.PFW_MIXED if prec == "mixed" else _capi.PFW_FP64
    if which == "design":  # the linear design of the first 64 fitting-set configurations
        from pfw_amd.ace import ELEMENTS
        model, ps, _ = pfw.load_resolved(pfw.ace_basis.synthetic_model())
        lab = [ELEMENTS[z - 1] for z in model.atomic_numbers]
        cfgs, _ = pfw.workloads.fitting_set()
        sub = [pfw.System(X, H, pb, [lab[v] for v in sp]) for X, H, pb, sp in cfgs[:64]]
        calc = pfw.ACECalculator(model, ps, precision=prec)
        for _ in range(max(1, steps)):
            d = pfw.design_matrix(calc, sub)
        print(json.dumps({"profile": which, "precision": prec, "rows": d["matrix"].shape[0],
                          "columns": d["matrix"].shape[1], "calls": max(1, steps)}), flush=True)
        return
    if which in ("pair", "argon"):
        if which == "pair":
            X, H, pbc, sp = pfw.workloads.b2_triclinic()
            model = pfw.SmoothPair(["A", "B"], np.array([[1.0, 0.7], [0.7, 1.3]]), 6.0)
            labels = ["A" if v == 0 else "B" for v in sp]
        else:
            X, H, pbc, sp = pfw.workloads.argon_fcc_rattled(10)
            model = pfw.lennard_jones(epsilon=0.010323, sigma=3.405, radius=2.5 * 3.405)
            labels = ["Ar"] * len(X)
        w = pfw.Workspace(precision=prec)
        pfw.evaluate_(w, pfw.System(X, H, pbc, labels), model)
        h, step = w.ctx.h, lib.pfw_pair_step
    else:
        from pfw_amd.ace import ELEMENTS
        if which == "ace":
            model, ps, _ = pfw.load_resolved(pfw.ace_basis.synthetic_model())
            X, H, pbc, sp = pfw.workloads.bcc_binary()
            labels = [ELEMENTS[model.atomic_numbers[v] - 1] for v in sp]
        else:
            try:
                import tomllib as toml
            except ImportError:
                import tomli as toml
            gold = os.path.join(ROOT, "tests", "golden")
            model, ps, _ = pfw.load_resolved(os.path.join(gold, "may30_resolved.npz"))
            with open(os.path.join(gold, "may30_omega-240.toml"), "rb") as f:
                d = toml.load(f)
            n0 = len(d["species"])
            H0 = np.asarray(d["cell"], float).reshape(3, 3).T
            shifts = np.array([[a, b, c] for a in range(5) for b in range(5) for c in range(5)], float) @ H0.T
            X = (np.asarray(d["positions"], float).reshape(n0, 3)[None] + shifts[:, None, :]).reshape(-1, 3)
            H, pbc = 5 * H0, (True, True, True)
            labels = [ELEMENTS[z - 1] for z in d["species"]] * len(shifts)
        calc = pfw.ACECalculator(model, ps, precision=prec)
        calc.workspace.set_tuning(ace_density_contraction=density)
        calc.evaluate(pfw.System(X, H, pbc, labels))
        h, step = calc.workspace.ctx.h, lib.pfw_ace_step
    for _ in range(steps):
        _capi.check(h, step(h, 7, mode))
    _capi.check(h, lib.pfw_synchronize(h))
    print(json.dumps({"profile": which, "precision": prec, "atoms": len(X), "steps": steps,
                      "ace_density_contraction": density}), flush=True)


def calibrate_copy(nbytes=256 << 20, reps=5):
    """A plain device copy of known size: the reference point that scales the
    FETCH_SIZE / WRITE_SIZE counters into HBM bytes."""
    import torch

    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ace", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-verlet", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--cpu-threads", type=int, default=None,
                    help="CPU baseline threads (default: physical cores of the affinity mask, at most 16)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a device copy of 256 MiB five times (HBM counter calibration)")
    ap.add_argument("--profile", choices=("pair", "ace", "may30", "argon", "design"), default=None,
                    help="only run --steps device steps of one workload (for rocprofv3)")
    ap.add_argument("--precision", choices=("mixed", "float64"), default="mixed", help="precision of --profile")
    ap.add_argument("--ace-density-contraction", type=int, choices=(0, 1, 2, 3), default=0,
                    help="ACE contractions of --profile: 0 built-in, 1 VALU, 2 densities on matrix cores, "
                         "3 densities and pullback on matrix cores")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (each GPU box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    pfw = load_package()
    from pfw_amd import _capi

    if args.profile:
        profile_run(pfw, _capi.load(), _capi, args.profile, args.precision, args.steps, args.ace_density_contraction)
        return
    if not args.no_cpu_baseline:
        sys.path.insert(0, ROOT)  # the oracle: CPU baseline and accuracy checks only

    lib = _capi.load()
    host = host_info()
    threads = args.cpu_threads or baseline_threads(host)
    args.cpu_threads_used = threads

    # ---------------------------------------------------------------- pair (headline)
    X, H, pbc, sp = pfw.workloads.b2_triclinic()
    n = len(X)
    A = np.array([[1.0, 0.7], [0.7, 1.3]])
    rc = 6.0
    p = pfw.SmoothPair(["A", "B"], A, rc)
    sys_ = pfw.System(X, H, pbc, ["A" if s == 0 else "B" for s in sp])
    pair = {}
    for prec, mode in (("mixed", _capi.PFW_MIXED), ("float64", _capi.PFW_FP64)):
        w = pfw.Workspace(precision=prec)
        res = pfw.evaluate_(w, sys_, p)  # uploads geometry + model outside the timed region
        secs, stage = time_steps(lib, w.ctx.h, lib.pfw_pair_step, args.steps, args.warmup, mode)
        gsecs, _ = graph_steps(lib, _capi, w, lib.pfw_pair_step, args.steps, args.warmup, mode)
        pair[prec] = dict(seconds=secs, stage_ms=stage, result=res, ws=w, graph_seconds=gsecs)
    # host boundary: positions uploaded from and forces returned to pageable host memory every step
    wh = pair["mixed"]["ws"]
    Xh = np.ascontiguousarray(X)
    Fh = np.zeros((n, 3))
    Eh, Wh = np.zeros(1), np.zeros(9)
    P = ctypes.POINTER(ctypes.c_double)
    hh = wh.ctx.h
    _capi.check(hh, lib.pfw_synchronize(hh))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _capi.check(hh, lib.pfw_update_positions(hh, Xh.ctypes.data_as(P)))
        _capi.check(hh, lib.pfw_pair_evaluate(hh, 7, _capi.PFW_MIXED, Eh.ctypes.data_as(P), Fh.ctypes.data_as(P),
                                              Wh.ctypes.data_as(P)))
    host_secs = time.perf_counter() - t0
    # the same with page-locked host buffers from the library's cache (one DMA each way)
    Xp = _capi.pinned_empty(3 * n).reshape(n, 3)
    Xp[...] = Xh
    Fp = _capi.pinned_empty(3 * n).reshape(n, 3)
    _capi.check(hh, lib.pfw_synchronize(hh))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _capi.check(hh, lib.pfw_update_positions(hh, Xp.ctypes.data_as(P)))
        _capi.check(hh, lib.pfw_pair_evaluate(hh, 7, _capi.PFW_MIXED, Eh.ctypes.data_as(P), Fp.ctypes.data_as(P),
                                              Wh.ctypes.data_as(P)))
    pinned_secs = time.perf_counter() - t0
    npairs = len(pair["mixed"]["ws"].export_pairs()[0])
    pflops = npairs * FLOPS_PER_PAIR["smooth"]
    mixed = pair["mixed"]
    value = n * args.steps / mixed["seconds"] / 1e6
    fp64_value = n * args.steps / pair["float64"]["seconds"] / 1e6
    k_ms = float(mixed["stage_ms"][4])
    k64_ms = float(pair["float64"]["stage_ms"][4])
    achieved = pflops / (k_ms * 1e-3) / 1e12

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Matom-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * mixed["seconds"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "b2_smoothpair_256k_triclinic", "atoms": n, "cutoff_A": rc, "species": 2,
                   "cell": "B2 a=3.0 A, 40x40x80 cells, +/-5% seeded shear", "precision": "mixed",
                   "step": "cell build + E+F+V + FP64 reductions + status read"},
        "roofline": {"kernel": "k_pair<float,SMOOTH>", "bound": "valu", "compute_unit": "VALU fp32",
                     "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "frac_of_unpacked_issue": round(achieved / PEAK_FP32_UNPACKED_TFLOPS, 4),
                     "traffic": read_traffic("pair_b2_256k_traffic.json"),
                     "directed_pairs": npairs, "flops_per_pair": FLOPS_PER_PAIR["smooth"],
                     "kernel_ms": round(k_ms, 5)},
        "fp64": {"value": round(fp64_value, 3), "ms_per_step": round(1e3 * pair["float64"]["seconds"] / args.steps, 4),
                 "kernel_ms": round(k64_ms, 5),
                 "roofline_frac": round(pflops / (k64_ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4)},
        "mixed_over_fp64": round(value / fp64_value, 3),
        "graph": {prec: {"value": round(n * args.steps / pair[prec]["graph_seconds"] / 1e6, 3),
                         "ms_per_step": round(1e3 * pair[prec]["graph_seconds"] / args.steps, 4)} for prec in pair},
        "host_boundary": {"value": round(n * args.steps / host_secs / 1e6, 3),
                          "ms_per_step": round(1e3 * host_secs / args.steps, 4),
                          "note": "positions host->device and forces device->host (pageable) inside every step",
                          "pinned": {"value": round(n * args.steps / pinned_secs / 1e6, 3),
                                     "ms_per_step": round(1e3 * pinned_secs / args.steps, 4),
                                     "note": "the same with page-locked buffers from pfw_host_alloc"}},
        "stages_ms": {k: round(float(v), 5) for k, v in zip(STAGES, mixed["stage_ms"])},
        "cpu_baseline": None,
    }

    # ---------------------------------------------------------------- argon 4,000 LJ (CPU reference cell)
    line["argon_4000"] = bench_argon_4000(pfw, lib, _capi, args, host)

    # ---------------------------------------------------------------- Verlet reuse (skin 0.5 Å)
    checks = not args.no_cpu_baseline
    if checks:
        from oracle.core import KIND_LJ, KIND_SMOOTH, pair_evaluate
    if not args.no_verlet:
        labels_b2 = ["A" if s == 0 else "B" for s in sp]
        chk = (lambda Xs: pair_evaluate(Xs, sp, H, pbc, KIND_SMOOTH, A, rc, threads=threads)) if checks else None
        vb = {prec: bench_verlet(pfw, lib, _capi, X, H, pbc, sp, p, labels_b2, 0.5, prec, mode, check=chk)
              for prec, mode in (("mixed", _capi.PFW_MIXED), ("float64", _capi.PFW_FP64))}
        kl = vb["mixed"]["reuse_stages_ms"]["interaction"]
        # HBM is the nearer roof of the list kernel: per list entry a 4 B index;
        # per extended entry (atoms, ghosts, the sentinel) its 32 B position
        # record; per atom its row length, permutation slot and FP64 force
        # (4 + 4 + 24 B); per 16-atom wavefront a 56 B energy / virial record.
        ne = vb["mixed"]["list_entries"]
        next_ = n + vb["mixed"]["ghosts"] + 1
        lbytes = 4 * ne + 32 * next_ + 32 * n + 56 * ((n + 15) // 16)
        # the FLOP side: the same directed in-cutoff pairs as the fresh step (to within the motion)
        vb["mixed"]["roofline"] = {"kernel": "k_pair_list<float,SMOOTH>", "bound": "hbm",
                                   "achieved": round(lbytes / (kl * 1e-3) / 1e9, 1), "peak": PEAK_HBM_GBS,
                                   "unit": "GB/s", "frac": round(lbytes / (kl * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                                   "algorithmic_bytes": lbytes,
                                   "traffic": read_traffic("pair_list_b2_256k_traffic.json"),
                                   "valu_fp32_frac": round(pflops / (kl * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
                                   "kernel_ms": kl}
        line["verlet"] = {"workload": "b2_smoothpair_256k_triclinic, trajectory x_s = x_0 + s v", **vb}
        Xr, Hr, pbr, spr = pfw.workloads.argon_fcc(14, sigma=0.03)
        lj = pfw.lennard_jones()
        line["verlet_argon_11k"] = {
            "workload": "argon FCC 14^3 (10,976 atoms), LJ eps 0.0103 sigma 3.4 rc 8.5, the reference's "
                        "skin trajectory (benchmark/gpu_pairs.jl:118-142)",
            "mixed": bench_verlet(pfw, lib, _capi, Xr, Hr, pbr, spr, lj, ["Ar"] * len(Xr), 0.5, "mixed",
                                  _capi.PFW_MIXED,
                                  check=(lambda Xs: pair_evaluate(Xs, spr, Hr, pbr, KIND_LJ, (0.0103, 3.4), 8.5,
                                                                  threads=threads)) if checks else None),
            "reference_rx6800xt_float32": {"amortized_ms": 0.256, "reuse_median_ms": 0.169, "rebuild_steps": 1,
                                           "source": "docs/implementation/GPU-PAIR-RESULTS.md:87-104"},
        }

    # ---------------------------------------------------------------- ACE
    if not args.no_ace:
        D = pfw.ace_basis.synthetic_model()
        model, ps, _ = pfw.load_resolved(D)
        Xa, Ha, pba, spa = pfw.workloads.bcc_binary()
        from pfw_amd.ace import ELEMENTS
        labels = [ELEMENTS[model.atomic_numbers[s] - 1] for s in spa]
        sys_a = pfw.System(Xa, Ha, pba, labels)
        ace = {}
        for prec, mode in (("mixed", _capi.PFW_MIXED), ("float64", _capi.PFW_FP64)):
            calc = pfw.ACECalculator(model, ps, precision=prec)
            res = calc.evaluate(sys_a)
            h = calc.workspace.ctx.h
            secs, stage = time_steps(lib, h, lib.pfw_ace_step, args.steps, args.warmup, mode)
            gsecs, _ = graph_steps(lib, _capi, calc.workspace, lib.pfw_ace_step, args.steps, args.warmup, mode)
            ace[prec] = dict(seconds=secs, stage_ms=stage, result=res, calc=calc, graph_seconds=gsecs)
        # the density contraction on the VALU and on the matrix cores, same calculator
        dcw = ace["mixed"]["calc"].workspace
        dens_ms = {}
        for name, dc in (("valu", 1), ("mfma_densities", 2), ("mfma_densities_pullback", 3)):
            dcw.set_tuning(ace_density_contraction=dc)
            dsecs, dstage = time_steps(lib, dcw.ctx.h, lib.pfw_ace_step, args.steps, args.warmup, _capi.PFW_MIXED)
            dens_ms[name] = {"ms_per_step": round(1e3 * dsecs / args.steps, 4), "interaction_ms": round(float(dstage[4]), 5)}
        dcw.set_tuning(ace_density_contraction=0)
        edges = len(ace["mixed"]["calc"].workspace.export_pairs()[0])
        na = len(Xa)
        aflops, per_edge, per_site = ace_flops(model, edges, na)
        ka = float(ace["mixed"]["stage_ms"][4])
        ka64 = float(ace["float64"]["stage_ms"][4])
        line["ace"] = {
            "workload": "bcc_binary_ace_32k", "atoms": na, "cutoff_A": model.cutoff, "edges": edges,
            "basis": {"radial": model.radial.nr, "lmax": model.lmax, "densities": len(model.density),
                      "products": len(model.products), "invariants": model.nbasis, "correlation": 3, "degree": 12},
            "value": round(na * args.steps / ace["mixed"]["seconds"] / 1e6, 4),
            "fp64_value": round(na * args.steps / ace["float64"]["seconds"] / 1e6, 4),
            "unit": "Matom-evals/s",
            "ms_per_step": round(1e3 * ace["mixed"]["seconds"] / args.steps, 4),
            "fp64_ms_per_step": round(1e3 * ace["float64"]["seconds"] / args.steps, 4),
            "roofline": {"kernel": "k_ace_dens_mfma<4,2>+k_ace_prod+k_ace_pull<float,4>", "bound": "valu",
                         "compute_unit": "VALU fp32 (the density contraction on FP32 MFMA)",
                         "achieved": round(aflops / (ka * 1e-3) / 1e12, 3), "peak": PEAK_FP32_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(aflops / (ka * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
                         "frac_of_unpacked_issue": round(aflops / (ka * 1e-3) / 1e12 / PEAK_FP32_UNPACKED_TFLOPS, 4),
                         "flops_per_edge": per_edge, "flops_per_site": per_site, "kernel_ms": round(ka, 5),
                         "traffic": read_traffic("ace_bcc_32k_traffic.json")},
            "graph": {prec: {"value": round(na * args.steps / ace[prec]["graph_seconds"] / 1e6, 4),
                             "ms_per_step": round(1e3 * ace[prec]["graph_seconds"] / args.steps, 4)} for prec in ace},
            "fp64_roofline_frac": round(aflops / (ka64 * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4),
            "fp64_kernel_ms": round(ka64, 5),
            "stages_ms": {k: round(float(v), 5) for k, v in zip(STAGES, ace["mixed"]["stage_ms"])},
            "density_contraction": dens_ms,
        }
        # the autotuner's sweep on this workload (a fresh calculator; knob and graph choices)
        import tempfile
        with tempfile.TemporaryDirectory() as td:
            tcalc = pfw.ACECalculator(model, ps, precision="mixed")
            rep = pfw.autotune(tcalc, sys_a, steps=10, warmup=2, cache=os.path.join(td, "tuning.json"))
            tp = pfw.Workspace(precision="mixed")
            repp = pfw.autotune(tp, sys_, p, steps=10, warmup=2, cache=os.path.join(td, "tuning.json"))
        default = [c for c in rep["candidates"] if c[0] == {"ace_sites_per_wave": 2, "ace_radial_chunk": 20,
                                                             "ace_product_waves": 16, "ace_density_contraction": 2}
                   and not c[1]]
        line["autotune"] = {
            "ace_bcc_32k_mixed": {"best": rep["best"], "graphs": rep["graphs"], "ms_per_step": round(rep["ms"], 4),
                                  "built_in_ms": round(default[0][2], 4) if default else None,
                                  "candidates": len(rep["candidates"])},
            "b2_256k_mixed": {"best": repp["best"], "graphs": repp["graphs"], "ms_per_step": round(repp["ms"], 4),
                              "candidates": [[c[0]["pair_lanes_per_atom"], c[1], round(c[2], 4)]
                                             for c in repp["candidates"]]},
            "timing": "median per-step wall time (each step ends with its status read), best of 3 runs of 10"}
        if checks:
            from oracle.ace import ace_evaluate as oracle_ace
            from oracle.ace import load_resolved as oracle_load

            om = oracle_load(D)
        if not args.no_verlet:
            chk = (lambda Xs: oracle_ace(Xs, spa, Ha, pba, om, threads=threads)[:3]) if checks else None
            line["ace"]["verlet"] = {
                "workload": "bcc_binary_ace_32k, trajectory x_s = x_0 + s v (environments at cutoff + skin)",
                **{prec: bench_verlet(pfw, lib, _capi, Xa, Ha, pba, spa, (model, ps), labels, 0.5, prec, mode,
                                      ace=True, check=chk)
                   for prec, mode in (("mixed", _capi.PFW_MIXED), ("float64", _capi.PFW_FP64))}}

    # ---------------------------------------------------------------- May30 real model
    if not args.no_ace:
        line["ace_may30"] = bench_may30(pfw, lib, _capi, args)

    # ---------------------------------------------------------------- equivariant features
    if not args.no_ace:
        line["features"] = bench_features(pfw, args)

    # ---------------------------------------------------------------- fitting set (batch + adjoints)
    if not args.no_batch and not args.no_ace:
        line["fitting_set"] = bench_fitting_set(pfw, lib, _capi, model, ps, p, args)

    # ---------------------------------------------------------------- CPU baseline + accuracy
    if checks:
        line["cpu_baseline"], (E, F, W) = cpu_baseline(
            lambda t: lambda: pair_evaluate(X, sp, H, pbc, KIND_SMOOTH, A, rc, threads=t), n,
            "full 256,000-atom E+F+V evaluations (cell build + phase-coloured half-list, FP64)", threads, host,
            " over phase-coloured cell tasks")
        line["accuracy_vs_fp64_oracle"] = {prec: accuracy(pair[prec]["result"], E, F, W, n) for prec in pair}
        line["neighbour_audit"] = {
            prec: neighbour_audit(pfw, pair[prec]["ws"], lambda: pfw.evaluate_(pair[prec]["ws"], sys_, p), X, H, pbc,
                                  rc) for prec in pair}
        if "ace" in line:
            line["ace"]["cpu_baseline"], (Ea, Fa, Wa, _) = cpu_baseline(
                lambda t: lambda: oracle_ace(Xa, spa, Ha, pba, om, threads=t, coloured=True), na,
                "full 32,768-atom E+F+V evaluations (FP64 site pipeline)", threads, host,
                " over greedily coloured site phases with one shared force array")
            line["ace"]["accuracy_vs_fp64_oracle"] = {prec: accuracy(ace[prec]["result"], Ea, Fa, Wa, na)
                                                      for prec in ace}
            line["ace"]["neighbour_audit"] = {
                prec: neighbour_audit(pfw, ace[prec]["calc"].workspace, lambda: ace[prec]["calc"].evaluate(sys_a), Xa,
                                      Ha, pba, model.cutoff) for prec in ace}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
