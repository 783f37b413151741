#!/usr/bin/env python3
"""Headline benchmark: trial-move Delta-U evaluations per second.

Workload: the Ewald LJ + Coulomb electrolyte -- N = 20,000 ions (+-1e, 50/50)
in a periodic cube of side 120 A, Lorentz-Berthelot LJ (sigma 4 A, eps 0.1 kT,
uncut over all pairs as the reference's NonbondedEwald<LennardJonesLB>
evaluates it), screened Ewald real space cut at 12 A (alpha 0.25 / A,
kappa -> 0), spherical k-space |n| <= 10 (2,242 stored k-vectors), tinfoil
boundary, AtomicTranslation dp = 1 A, S(k) refreshed every 20,000 accepted
moves.  Synthetic seeded configuration (jittered lattice).

One step = one Propagator::move() per replica = one sweep of N trial moves
for each of R independent replicas, all advanced by one persistent launch
(one workgroup per replica).  Throughput counts every trial (Delta-U
computed and Metropolis decision taken).

The CPU baseline is the reference's own hot path compiled from its sources
(oracle/_ref/ref_harness, single thread) on a bounded sample of the same
system, or the C restatement (oracle) where that binary is absent.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp32": 157.3, "fp64": 78.6}           # MI355X dense vector TFLOP/s
HBM_PEAK_GBS = 8000.0
FLOPS_PAIR_EWALD_LJ = 31                       # EwaldReal(kappa->0) + LJ-LB per pair
FLOPS_K_PER_TRIAL = 32                         # exp(i k.r) update per k-vector


def cpu_baseline(w, seconds_target=12.0):
    """Time the reference (or its C restatement) on a bounded sample: single
    trials (peratom off) of the same system, one thread."""
    from beofaunus_amd import workloads
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    with tempfile.TemporaryDirectory() as td:
        cfg, state = workloads.write_reference_inputs(w, td, seed=7, peratom=False)
        if os.path.exists(harness):
            # calibrate: a short run, then a sample sized for ~seconds_target
            probe = json.loads(subprocess.run([harness, "mctime", w["ref_kind"], cfg, state, "200"],
                                              check=True, capture_output=True, text=True).stdout)
            rate = probe["trials_per_s"]
            ncalls = int(max(200, min(200000, rate * seconds_target)))
            out = json.loads(subprocess.run([harness, "mctime", w["ref_kind"], cfg, state,
                                             str(ncalls)], check=True, capture_output=True,
                                            text=True).stdout)
            return {"value": out["trials_per_s"], "unit": "trial dU evals/s", "cores": 1,
                    "kind": "reference",
                    "sample": "%d single-particle trials of the same N=%d system "
                              "(reference C++ hot path, g++ -O2, 1 thread; timed loop only)"
                              % (int(out["trials"]), len(w["charge"]))}
    # C restatement fallback
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_py import OracleSystem, lib as olib
    from fixture_systems import ELEC, SR, NB  # noqa: F401
    o = OracleSystem("ewaldlj", w["box"], w["xyz"], w["charge"], w["radius"], w["mw"], w["types"],
                     w["groups"], w["ham"]["lB"], type_sigma=w["ham"]["type_sigma"],
                     type_eps=w["ham"]["type_eps"], type_charge=w["ham"]["type_charge"],
                     type_radius=w["ham"]["type_radius"], type_dp=w["ham"]["type_dp"],
                     ewald=w["ham"]["ewald"])
    olib().or_add_move(o.h, 1, 0, 1.0, 0, 0, None, None, 0.0, 0.0, 1)
    o.seed(7)
    olib().or_propagator_init(o.h)
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < seconds_target:
        du, acc, _, _, _ = o.propagate(50)
        n += len(du)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "trial dU evals/s", "cores": 1, "kind": "port",
            "sample": "%d single-particle trials of the same system (C restatement, 1 thread)" % n}


def du_error_vs_fp64(w, seed=11, ntrials=4000):
    """Mixed-path Delta-U error against the FP64 device path on the same RNG
    stream: compare trial by trial while the two chains agree."""
    from beofaunus_amd import fgx, workloads
    out = {}
    rec = {}
    for prec in (fgx.FP64, fgx.MIXED):
        s = workloads.make_system(w, precision=prec)
        s.rng_seed(seed)
        s.propagator_init()
        st, du, acc = s.propagate(1, record=ntrials)
        rec[prec] = (du[:ntrials], acc[:ntrials])
        s.close()
    d64, a64 = rec[fgx.FP64]
    dmx, amx = rec[fgx.MIXED]
    n = min(len(d64), len(dmx))
    mism = np.nonzero(a64[:n] != amx[:n])[0]
    upto = int(mism[0]) if len(mism) else n
    fin = np.isfinite(d64[:upto]) & np.isfinite(dmx[:upto])
    err = np.abs(dmx[:upto][fin] - d64[:upto][fin])
    out["compared_trials"] = int(fin.sum())
    out["max_abs_kT"] = float(err.max()) if len(err) else None
    out["p99_abs_kT"] = float(np.percentile(err, 99)) if len(err) else None
    out["first_accept_mismatch"] = int(mism[0]) if len(mism) else None
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--n", type=int, default=20000)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error-check", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; replicas of the Markov chain are batched "
              "on that GPU (--replicas).  --gpus %d is not supported." % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    from beofaunus_amd import fgx, workloads

    prec = fgx.MIXED if args.precision == "mixed" else fgx.FP64
    L = 120.0 * (args.n / 20000.0) ** (1.0 / 3.0)
    w = workloads.ewald_electrolyte(n=args.n, L=L, precision=prec)
    R = args.replicas
    t_build = time.perf_counter()
    systems = []
    for r in range(R):
        s = workloads.make_system(w)
        s.rng_seed(1000 + r)
        s.propagator_init()
        systems.append(s)
    t_build = time.perf_counter() - t_build
    # warm-up doubles as the schedule autotuner: every (block size, unroll)
    # candidate advances the replicas by one sweep; the fastest is kept
    tuned = fgx.autotune_batch(systems, 1) if args.warmup > 0 else None
    for _ in range(max(0, args.warmup - 1)):
        fgx.propagate_batch(systems, 1)
    torch.cuda.synchronize()
    dev = []
    trials = 0
    acc = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = fgx.propagate_batch(systems, 1)
        dev.append(st["seconds"])
        trials += st["trials"]
        acc += st["accepted"]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    n = args.n
    value = trials / wall
    # algorithmic work per trial: 2 (N-1) pair evaluations (old and new, as the
    # reference counts them) + K k-vector updates
    from beofaunus_amd.fgx import lib  # noqa: F401
    nk = 2242 if n == 20000 else None
    flops_trial = 2 * (n - 1) * FLOPS_PAIR_EWALD_LJ + (nk or 0) * FLOPS_K_PER_TRIAL
    trials_per_launch = trials / args.steps
    dev_mean = float(np.mean(dev))
    achieved = flops_trial * trials_per_launch / dev_mean / 1e12
    peak = PEAK["fp32" if prec == fgx.MIXED else "fp64"]
    line = {
        "metric": "trial-move ΔU evals/s at stated ΔU error vs FP64; p50/p99 sweep latency",
        "value": value,
        "unit": "trial dU evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pairs + f64 accumulate" if prec == fgx.MIXED else "f64",
        "data": "synthetic (seeded jittered-lattice configuration)",
        "config": {
            "workload": "Ewald LJ+Coulomb electrolyte: N=%d, L=%.1f A, +-1e 50/50, LJ-LB sigma 4 A "
                        "eps 0.1 kT uncut, Ewald alpha 0.25/A rc 12 A |n|<=10 spherical, tinfoil, "
                        "AtomicTranslation dp 1 A, S(k) refresh every 20000 accepts" % (n, L),
            "replicas": R,
            "trials_per_step": int(trials_per_launch),
            "precision": args.precision,
        },
        "pairs_per_s": value * 2 * (n - 1),
        "acceptance": acc / max(1, trials),
        "sweep_latency_ms": {"p50": 1e3 * float(np.percentile(dev, 50)),
                             "p99": 1e3 * float(np.percentile(dev, 99))},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None,
                     "kernel": "k_sweep_at<%s>" % ("mixed" if prec == fgx.MIXED else "fp64"),
                     "flops_per_trial": flops_trial},
        "build_seconds": t_build,
        "schedule": tuned,
    }
    for s in systems:
        s.close()
    if not args.no_error_check:
        try:
            line["du_error_vs_fp64"] = du_error_vs_fp64(w)
        except Exception as e:  # reported, never hidden
            line["du_error_vs_fp64"] = {"error": str(e)}
    if not args.no_cpu_baseline:
        try:
            line["cpu_baseline"] = cpu_baseline(w)
        except Exception as e:
            line["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
