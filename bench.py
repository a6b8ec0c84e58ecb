"""Throughput benchmark of the MI355X force-evaluation hot path.

Default workload (one GPU): the 256,000-atom two-species SmoothPair B2 lattice
in a sheared triclinic cell (cutoff 6.0 Å), one "step" = fresh cell build +
E+F+V + FP64 reductions + nonfinite status check, positions already resident
in HBM. ``value`` is the mixed-precision throughput in Matom-evals/s; the
plain-FP64 throughput of the same step, per-stage device times, the roofline
of the dominant kernel and a CPU baseline are reported beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "potentialframework.jl_amd")

METRIC = "Matom-evals/s (E+F+V), 256k-atom pair & 32k-atom ACE; mixed vs FP64 at tol"
# Published MI355X peaks used for roofline fractions.
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
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


def time_pair_steps(pfw, w, sys_, p, steps, warmup, precision_mode):
    """Time `steps` device steps of the pair path; returns (seconds, mean stage ms)."""
    import ctypes
    from pfw_amd import _capi

    pfw.evaluate_(w, sys_, p)  # upload geometry + model (outside the timed region)
    lib = _capi.load()
    h = w.ctx.h
    for _ in range(warmup):
        _capi.check(h, lib.pfw_pair_step(h, 7, precision_mode))
    _capi.check(h, lib.pfw_set_timing(h, 1))
    stage = np.zeros(5)
    buf = np.zeros(5)
    _capi.check(h, lib.pfw_synchronize(h))
    t0 = time.perf_counter()
    for _ in range(steps):
        _capi.check(h, lib.pfw_pair_step(h, 7, precision_mode))
        lib.pfw_get_timings(h, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 5)
        stage += buf
    _capi.check(h, lib.pfw_synchronize(h))
    t1 = time.perf_counter()
    _capi.check(h, lib.pfw_set_timing(h, 0))
    return t1 - t0, stage / steps


def read_traffic(name):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE, calibrated), if present."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    except Exception:
        return None


def cpu_baseline_pair(X, sp, H, pbc, A, rc, threads):
    """The oracle's threaded FP64 restatement of the reference CPU pair path,
    timed on the host on the full workload (bounded number of repetitions).
    Also returns its FP64 outputs, used as the accuracy checker."""
    sys.path.insert(0, ROOT)
    from oracle import KIND_SMOOTH, pair_evaluate

    pair_evaluate(X, sp, H, pbc, KIND_SMOOTH, A, rc, threads=threads)  # warm-up
    times = []
    out = None
    budget = 20.0
    t_start = time.perf_counter()
    while len(times) < 7 and (time.perf_counter() - t_start) < budget:
        t0 = time.perf_counter()
        out = pair_evaluate(X, sp, H, pbc, KIND_SMOOTH, A, rc, threads=threads)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return med, len(times), out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (each GPU box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    pfw = load_package()
    from pfw_amd import _capi

    X, H, pbc, sp = pfw.workloads.b2_triclinic()
    n = len(X)
    A = np.array([[1.0, 0.7], [0.7, 1.3]])
    rc = 6.0
    p = pfw.SmoothPair(["A", "B"], A, rc)
    sys_ = pfw.System(X, H, pbc, ["A" if s == 0 else "B" for s in sp])

    results = {}
    for prec, mode in (("mixed", _capi.PFW_MIXED), ("float64", _capi.PFW_FP64)):
        w = pfw.Workspace(precision=prec)
        secs, stage = time_pair_steps(pfw, w, sys_, p, args.steps, args.warmup, mode)
        res = pfw.evaluate_(w, sys_, p)
        results[prec] = dict(seconds=secs, stage_ms=stage, result=res, ws=w)

    npairs = len(results["mixed"]["ws"].export_pairs()[0])
    mixed = results["mixed"]
    ms_per_step = 1e3 * mixed["seconds"] / args.steps
    value = n * args.steps / mixed["seconds"] / 1e6
    fp64_value = n * args.steps / results["float64"]["seconds"] / 1e6
    k_ms = float(mixed["stage_ms"][3])
    flops = npairs * FLOPS_PER_PAIR["smooth"]
    achieved = flops / (k_ms * 1e-3) / 1e12
    k64_ms = float(results["float64"]["stage_ms"][3])
    traffic = read_traffic("pair_b2_256k_traffic.json")

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Matom-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "b2_smoothpair_256k_triclinic", "atoms": n, "cutoff_A": rc,
                   "species": 2, "cell": "B2 a=3.0 40x40x80, +/-5% shear", "precision": "mixed",
                   "step": "cell build + E+F+V + FP64 reductions + status"},
        "roofline": {"kernel": "k_pair<float,SMOOTH>", "bound": "mfma", "compute_unit": "VALU fp32",
                     "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": traffic,
                     "directed_pairs": npairs, "flops_per_pair": FLOPS_PER_PAIR["smooth"],
                     "kernel_ms": round(k_ms, 5)},
        "fp64": {"value": round(fp64_value, 3), "ms_per_step": round(1e3 * results["float64"]["seconds"] / args.steps, 4),
                 "kernel_ms": round(k64_ms, 5),
                 "roofline_frac_fp64": round(flops / (k64_ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4)},
        "mixed_over_fp64": round(value / fp64_value, 3),
        "stages_ms": {k: round(float(v), 5) for k, v in zip(("binning", "scan", "scatter_sort", "interaction",
                                                              "reduction"), mixed["stage_ms"])},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        threads = max(1, min(args.cpu_threads, os.cpu_count() or 1))
        med, reps, (E, F, W) = cpu_baseline_pair(X, sp, H, pbc, A, rc, threads)
        line["cpu_baseline"] = {"value": round(n / med / 1e6, 4), "unit": "Matom-evals/s", "cores": threads,
                                "kind": "port",
                                "sample": f"median of {reps} full 256,000-atom E+F+V evaluations (cell build + "
                                          f"phase-coloured half-list FP64, {threads} threads)"}
        acc = {}
        for prec in ("mixed", "float64"):
            r = results[prec]["result"]
            d = r["forces"].T - F
            acc[prec] = {"force_max": float(np.abs(d).max()), "force_rms": float(np.sqrt(np.mean(d * d))),
                         "energy_per_atom": float(abs(r["energy"] - E) / n),
                         "virial_max": float(np.abs(r["virial"] - W).max())}
        line["accuracy_vs_fp64_oracle"] = acc
    print(json.dumps(line))


if __name__ == "__main__":
    main()
