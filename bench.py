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
FLOPS_PAIR_DHHS = 27                           # Debye-Hueckel + hard sphere per pair
# Algorithmic bytes per trial: one partner record per partner (20 B FP32 SoA,
# 40 B FP64), the moved particle twice, and per k-vector A_k + S(k) read
# (24 B) plus S(k) read+write on an accepted move (32 B).
BYTES_PARTNER = {"mixed": 20, "fp64": 40}
BYTES_K_READ, BYTES_K_ACCEPT = 24, 32
# HBM-side traffic per launch (256 replicas x N=20,000, mixed path), from
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of one timed step, scaled by the
# counters' reading for a 1 GiB copy (FETCH x2.0, WRITE x1.0).  Valid only for
# the kernel variant and schedule it was measured on; other schedules report
# no traffic.
MEASURED_TRAFFIC = {
    ("ewald", 20000, 256, "mixed"): {
        "bytes": 1024940990 * 2 * 1024 + 432348 * 1024,
        "kernel": "k_sweep_at<3,2,true,4,true,false>", "threads": 512, "unroll": 4,
        "source": "profiles/k_sweep_at_traffic_ewald_electrolyte_n20000.md"},
    ("ewald-cells", 20000, 256, "mixed"): {
        "bytes": 64251548 * 2 * 1024 + 564152 * 1024,
        "kernel": "k_sweep_at<3,2,true,4,true,true>", "threads": 512, "unroll": 4,
        "source": "profiles/k_sweep_at_traffic_ewald_electrolyte_n20000.md"},
}


def traffic_for(key, schedule):
    """measured traffic of a launch, when its schedule is the measured one"""
    m = MEASURED_TRAFFIC.get(key)
    if not m or not schedule or schedule.get("threads") != m["threads"] or \
            schedule.get("unroll") != m["unroll"]:
        return None, None
    return m["bytes"], {k: m[k] for k in ("kernel", "threads", "unroll", "source")}


def host_cores():
    """host cores to use: the process's CPU set, at most 16 (the box's CPU
    share per GPU)"""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(w, seconds_target=8.0, all_cores=True):
    """Time the reference (or its C restatement) on a bounded sample: single
    trials (one trial per move() call) of the same system, one thread; then
    C independent chains of the same sample, one per host core (pinned with
    taskset where available), as their aggregate rate."""
    from beofaunus_amd import workloads
    import shutil
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    npart = len(w["charge"])
    what = "single-group TranslateRotate" if w["move"]["kind"] == 2 else "single-particle"
    with tempfile.TemporaryDirectory() as td:
        cfg, state = workloads.write_reference_inputs(w, td, seed=7, peratom=False)
        # the reference's NonbondedEwald cannot take a cut LJ: the C
        # restatement stands in for it there
        if os.path.exists(harness) and w["ref_kind"] != "ewaldljcut":
            # calibrate on a short run, then a sample sized for ~seconds_target
            probe_n = 200 if what == "single-particle" else 2
            probe = json.loads(subprocess.run([harness, "mctime", w["ref_kind"], cfg, state,
                                               str(probe_n)], check=True, capture_output=True,
                                              text=True).stdout)
            rate = probe["trials_per_s"]
            ncalls = int(max(probe_n, min(200000, rate * seconds_target)))
            out = json.loads(subprocess.run([harness, "mctime", w["ref_kind"], cfg, state,
                                             str(ncalls)], check=True, capture_output=True,
                                            text=True).stdout)
            res = {"value": out["trials_per_s"], "unit": "trial dU evals/s", "cores": 1,
                   "kind": "reference",
                   "sample": "%d %s trials of the same N=%d system (reference C++ hot path "
                             "compiled with its Release flags -O3 -DNDEBUG -funroll-loops, 1 "
                             "thread; timed loop only)" % (int(out["trials"]), what, npart)}
            if all_cores:
                C = host_cores()
                ts = shutil.which("taskset")
                cores = sorted(os.sched_getaffinity(0))[:C] if hasattr(os, "sched_getaffinity") \
                    else list(range(C))
                t0 = time.perf_counter()
                procs = [subprocess.Popen(([ts, "-c", str(c)] if ts else []) +
                                          [harness, "mctime", w["ref_kind"], cfg, state,
                                           str(ncalls)], stdout=subprocess.PIPE, text=True)
                         for c in cores]
                outs = [json.loads(p.communicate()[0]) for p in procs]
                wall = time.perf_counter() - t0
                tr = sum(o["trials"] for o in outs)
                res["all_cores"] = {
                    "value": tr / max(o["seconds"] for o in outs), "cores": len(cores),
                    "pinned": bool(ts), "wall_s": wall,
                    "sample": "%d independent chains of the same sample, one per core"
                              % len(cores)}
            return res
    if w["move"]["kind"] != 1:
        return {"error": "reference binary absent; the C restatement baseline covers "
                         "AtomicTranslation workloads only"}
    # C restatement fallback
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_py import OracleSystem, lib as olib
    kind = w["ref_kind"]
    h = w["ham"]
    o = OracleSystem(kind, w["box"], w["xyz"], w["charge"], w["radius"], w["mw"], w["types"],
                     w["groups"], h["lB"], kappa=h.get("kappa", 0.0),
                     type_sigma=h["type_sigma"], type_eps=h["type_eps"],
                     type_charge=h["type_charge"], type_radius=h["type_radius"],
                     type_dp=h["type_dp"], ewald=h.get("ewald"), lj_cut=h.get("lj_cut", 0.0))
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


def du_error_vs_fp64(w, seed=11, ntrials=4000, ncalls=1):
    """Mixed-path Delta-U error against the FP64 device path on the same RNG
    stream: compare trial by trial while the two chains agree."""
    from beofaunus_amd import fgx, workloads
    out = {}
    rec = {}
    for prec in (fgx.FP64, fgx.MIXED):
        s = workloads.make_system(w, precision=prec)
        s.rng_seed(seed)
        s.propagator_init()
        st, du, acc = s.propagate(ncalls, record=ntrials)
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


def du_error_batched(w, chains=64, ncalls=2, cap=256, seed=11):
    """Mixed-path Delta-U error against the FP64 device path for many chains
    at once (one batched launch per precision, identical RNG streams per
    chain): every chain compared trial by trial up to its first accept/reject
    mismatch."""
    from beofaunus_amd import fgx, workloads
    rec = {}
    for prec in (fgx.FP64, fgx.MIXED):
        systems = []
        for c in range(chains):
            sy = workloads.make_system(w, precision=prec)
            sy.rng_seed(seed + c)
            sy.propagator_init()
            systems.append(sy)
        rec[prec] = fgx.propagate_batch_record(systems, ncalls, cap)[1:]
        for sy in systems:
            sy.close()
    errs, compared, finite, diverged, overlap_mismatch = [], 0, 0, 0, 0
    for c in range(chains):
        d64, a64 = rec[fgx.FP64][0][c], rec[fgx.FP64][1][c]
        dmx, amx = rec[fgx.MIXED][0][c], rec[fgx.MIXED][1][c]
        n = min(len(d64), len(dmx))
        mism = np.nonzero(a64[:n] != amx[:n])[0]
        upto = int(mism[0]) if len(mism) else n
        diverged += len(mism) > 0
        fin = np.isfinite(d64[:upto]) & np.isfinite(dmx[:upto])
        overlap_mismatch += int((np.isfinite(d64[:upto]) != np.isfinite(dmx[:upto])).sum())
        compared += upto
        finite += int(fin.sum())
        errs.extend(np.abs(dmx[:upto][fin] - d64[:upto][fin]).tolist())
    errs = np.array(errs)
    return {"chains": chains, "compared_trials": compared, "finite_trials": finite,
            "max_abs_kT": float(errs.max()) if len(errs) else None,
            "p99_abs_kT": float(np.percentile(errs, 99)) if len(errs) else None,
            "chains_diverged": diverged, "overlap_mismatch": overlap_mismatch,
            "accept_mismatch_rate": diverged / max(1, compared)}


METRIC = "trial-move ΔU evals/s at stated ΔU error vs FP64; p50/p99 sweep latency"


def timed_batch(fgx, systems, steps, ncalls=1):
    """`steps` launches of `ncalls` Propagator::move() calls per chain; wall
    clock around all of them (synchronised), device time of each from HIP
    events."""
    import torch
    torch.cuda.synchronize()
    dev, trials, acc = [], 0, 0
    t0 = time.perf_counter()
    for _ in range(steps):
        st = fgx.propagate_batch(systems, ncalls)
        dev.append(st["seconds"])
        trials += st["trials"]
        acc += st["accepted"]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return {"wall": wall, "dev": dev, "trials": trials, "accepted": acc,
            "value": trials / wall,
            "sweep_latency_ms": {"p50": 1e3 * float(np.percentile(dev, 50)),
                                 "p99": 1e3 * float(np.percentile(dev, 99))}}


def base_line(args, value, wall, dtype, config):
    return {
        "metric": METRIC,
        "value": value,
        "unit": "trial dU evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded configuration)",
        "config": config,
    }


def cells_vs_screen(w, seed=13, ntrials=4000):
    """The cell-list sweep against the all-pairs screen of the same mixed
    sweep (same Hamiltonian): accept sequences and the largest dU difference."""
    from beofaunus_amd import fgx, workloads
    rec = {}
    for mode in (0, 1):
        s = workloads.make_system(w, precision=fgx.MIXED)
        s.set_cell_list(mode)
        s.rng_seed(seed)
        s.propagator_init()
        st, du, acc = s.propagate(1, record=ntrials)
        rec[mode] = (du[:ntrials], acc[:ntrials], s.cell_list())
        s.close()
    (d0, a0, _), (d1, a1, cl) = rec[0], rec[1]
    return {"trials": int(len(d1)), "accept_sequences_equal": bool(np.array_equal(a0, a1)),
            "max_abs_diff_kT": float(np.max(np.abs(d1 - d0))), "cells_per_axis": cl[0],
            "cell_capacity": cl[1]}


def bench_ewald(args, fgx, workloads, cut=False):
    """Ewald LJ+Coulomb electrolyte, AtomicTranslation sweeps, R replicas.
    cut=True: LJ as CutShift<LennardJonesLB> at rc (every pair term vanishes
    beyond rc) and the cell-list sweep."""
    prec = fgx.MIXED if args.precision == "mixed" else fgx.FP64
    n = args.n or 20000
    L = 120.0 * (n / 20000.0) ** (1.0 / 3.0)
    w = workloads.ewald_electrolyte(n=n, L=L, precision=prec, lj_cut=cut)
    R = args.replicas or 256
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
    t = timed_batch(fgx, systems, args.steps)
    cells = systems[0].cell_list()
    for s in systems:
        s.close()
    # algorithmic work per trial: 2 (N-1) pair evaluations (old and new, as the
    # reference counts them) + K k-vector updates; with a cell list, the
    # candidates of the 27 cells around the moved particle
    nk = 2242 if n == 20000 else None
    partners = n - 1 if not cells[0] else 27.0 * n / cells[0] ** 3
    flops_trial = 2 * partners * FLOPS_PAIR_EWALD_LJ + (nk or 0) * FLOPS_K_PER_TRIAL
    trials_per_launch = t["trials"] / args.steps
    launch_s = float(np.mean(t["dev"]))
    tflops = flops_trial * trials_per_launch / launch_s / 1e12
    peak = PEAK["fp32" if prec == fgx.MIXED else "fp64"]
    acc_rate = t["accepted"] / max(1, t["trials"])
    k = nk or 0
    bytes_trial = (BYTES_PARTNER[args.precision] * (partners + 2) + BYTES_K_READ * k
                   + acc_rate * BYTES_K_ACCEPT * k)
    gbs = bytes_trial * trials_per_launch / launch_s / 1e9
    traffic = traffic_for(("ewald-cells" if cut else "ewald", n, R, args.precision), tuned)
    line = base_line(args, t["value"], t["wall"],
                     "f32 pairs + f64 accumulate" if prec == fgx.MIXED else "f64",
                     {"workload": "Ewald LJ+Coulomb electrolyte: N=%d, L=%.1f A, +-1e 50/50, LJ-LB "
                                  "sigma 4 A eps 0.1 kT %s, Ewald alpha 0.25/A rc 12 A |n|<=10 "
                                  "spherical, tinfoil, AtomicTranslation dp 1 A, S(k) refresh every "
                                  "20000 accepts" % (n, L, "cut and shifted at 12 A (cell list "
                                                     "%d^3)" % cells[0] if cut else "uncut"),
                      "replicas": R, "trials_per_step": int(trials_per_launch),
                      "precision": args.precision})
    line.update({
        "pairs_per_s": t["value"] * 2 * partners,
        "acceptance": t["accepted"] / max(1, t["trials"]),
        "sweep_latency_ms": t["sweep_latency_ms"],
        # one-vs-all streaming of the partner records: HBM roofline on the
        # algorithmic bytes; the flop rate against the vector peak beside it
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS,
                     "traffic": traffic[0], "traffic_measured_on": traffic[1],
                     "kernel": "k_sweep_at<%s>" % ("mixed" if prec == fgx.MIXED else "fp64"),
                     "bytes_per_trial": bytes_trial, "trials_per_launch": int(trials_per_launch),
                     "launch_ms": 1e3 * launch_s,
                     "valu": {"achieved": tflops, "peak": peak, "unit": "TFLOP/s",
                              "frac": tflops / peak, "flops_per_trial": flops_trial}},
        "build_seconds": t_build,
        "schedule": tuned,
    })
    if not args.no_error_check:
        try:
            line["du_error_vs_fp64"] = du_error_vs_fp64(w)
        except Exception as e:  # reported, never hidden
            line["du_error_vs_fp64"] = {"error": str(e)}
        if cut and prec == fgx.MIXED:
            try:
                line["cells_vs_all_pairs"] = cells_vs_screen(w)
            except Exception as e:
                line["cells_vs_all_pairs"] = {"error": str(e)}
    if not args.no_cpu_baseline:
        try:
            line["cpu_baseline"] = cpu_baseline(w)
        except Exception as e:
            line["cpu_baseline"] = {"error": str(e)}
    if not cut and args.config == "ewald" and prec == fgx.MIXED:
        # the same electrolyte with every pair term cut at rc (CutShift LJ),
        # swept with a cell list: a different Hamiltonian from the
        # reference's uncut NonbondedEwald<LennardJonesLB>, reported beside it
        sub = argparse.Namespace(**vars(args))
        sub.no_cpu_baseline = True
        c = bench_ewald(sub, fgx, workloads, cut=True)
        line["cut_lj_cell_list"] = {k: c[k] for k in ("value", "ms_per_step", "sweep_latency_ms",
                                                        "acceptance", "pairs_per_s")
                                    if k in c}
        line["cut_lj_cell_list"]["workload"] = c["config"]["workload"]
        for k in ("du_error_vs_fp64", "cells_vs_all_pairs"):
            if k in c:
                line["cut_lj_cell_list"][k] = c[k]
        line["cut_lj_cell_list"]["launch_ms"] = c["roofline"]["launch_ms"]
    if not cut and args.config == "ewald" and prec == fgx.MIXED and not args.no_sub:
        headline_subconfigs(args, fgx, workloads, line)
    return line


def headline_subconfigs(args, fgx, workloads, line):
    """The other configurations' rates beside the headline, a few steps each:
    the same electrolyte on the FP64 path, the rigid-group system (mixed with
    and without the group-pair cache, FP64), 48-replica parallel tempering
    (mixed, FP64) and a served job stream; then the mixed-over-FP64 target
    under both readings of the configuration numbering."""
    def sub(**kw):
        a = argparse.Namespace(**vars(args))
        a.no_cpu_baseline = True
        a.no_sub = True
        for k, v in kw.items():
            setattr(a, k, v)
        return a

    def guarded(key, fn):
        t0 = time.perf_counter()
        try:
            line[key] = fn()
        except Exception as e:  # reported, never hidden
            line[key] = {"error": str(e)}
        line[key]["bench_seconds"] = time.perf_counter() - t0
        print("[bench] %s done in %.1f s" % (key, line[key]["bench_seconds"]),
              file=sys.stderr, flush=True)

    def fp64_line():
        f = bench_ewald(sub(precision="fp64", steps=2, warmup=0, no_error_check=True), fgx,
                        workloads)
        return {"value": f["value"], "sweep_latency_ms": f["sweep_latency_ms"],
                "launch_ms": f["roofline"]["launch_ms"], "pairs_per_s": f["pairs_per_s"],
                "steps": 2, "mixed_over_fp64": line["value"] / f["value"]}

    def groups_line():
        g = bench_groups(sub(steps=2, warmup=1, replicas=256), fgx, workloads)
        out = {k: g[k] for k in ("value", "sweep_latency_ms", "pairs_per_s", "fp64",
                                 "mixed_uncached", "speedup_vs_fp64", "du_error_vs_fp64")
               if k in g}
        out["workload"] = g["config"]["workload"]
        out["roofline"] = g["roofline"]
        if "mixed_uncached" in g:
            out["uncached_over_fp64"] = g["mixed_uncached"]["value"] / g["fp64"]["value"]
        return out

    def temper_line():
        t = bench_temper(sub(steps=2, warmup=1, exchange_every=10, ladders=6, rungs=8), fgx,
                         workloads)
        out = {k: t[k] for k in ("value", "pairs_per_s", "sweep_latency_ms", "exchange_acceptance",
                                 "fp64", "speedup_vs_fp64", "ms_per_step")}
        out["workload"] = t["config"]["workload"]
        return out

    def serve_line():
        v = bench_serve(sub(jobs=64, job_sweeps=2, rate=50.0, hbm_mb=256.0, tick_ms=50.0,
                            seed=5), fgx, workloads)
        out = {k: v[k] for k in ("value", "pairs_per_s", "job_latency_s", "sweep_latency_ms",
                                 "queue_wait_s", "peak_resident_mb", "failed_jobs", "wall_s")}
        out["workload"] = v["config"]["workload"]
        out["note"] = ("64 jobs x 2 sweeps each, not the 256 x 1,000 of the served-stream "
                       "configuration (that stream is hours of device time)")
        return out

    guarded("fp64_same_workload", fp64_line)
    guarded("rigid_groups", groups_line)
    guarded("parallel_tempering", temper_line)
    guarded("served_stream", serve_line)
    tgt = {}
    g, t = line.get("rigid_groups", {}), line.get("parallel_tempering", {})
    if "speedup_vs_fp64" in g:
        err = (g.get("du_error_vs_fp64") or {}).get("max_abs_kT")
        tgt["rigid_groups (configs counted from 1: config 3)"] = {
            "mixed_over_fp64": g["speedup_vs_fp64"], "meets_5x": g["speedup_vs_fp64"] >= 5.0,
            "uncached_over_fp64": g.get("uncached_over_fp64"), "max_abs_du_err_kT": err,
            "p99_sweep_ms_under_load": (t.get("sweep_latency_ms") or {}).get("p99"),
            "load": "parallel tempering (config 4 counted from 1)"}
    if "speedup_vs_fp64" in t:
        tgt["parallel_tempering (configs counted from 0: config 3)"] = {
            "mixed_over_fp64": t["speedup_vs_fp64"], "meets_5x": t["speedup_vs_fp64"] >= 5.0,
            "p99_sweep_ms_under_load": (line.get("served_stream", {}).get("sweep_latency_ms")
                                        or {}).get("p99"),
            "load": "served stream (config 4 counted from 0)"}
    line["target_5x_vs_fp64"] = tgt


def bench_groups(args, fgx, workloads):
    """Rigid-group g2g: 64 molecules x 400 sites + 8000 ions, TranslateRotate.
    The FP64 device path (the reference's two-pass dU) and the mixed path
    (FP32 pairs, FP64 sums, group-pair energy cache) run the same replicas."""
    R = args.replicas or 256
    w = workloads.rigid_molecules()
    n = len(w["charge"])
    nsite = w["groups"][0]["back"] - w["groups"][0]["front"] + 1
    pairs_ref = 2 * nsite * (n - nsite)          # old + new, as the reference evaluates
    res = {}
    for key, prec, cached in (("fp64", fgx.FP64, False), ("mixed", fgx.MIXED, True)):
        systems = []
        for r in range(R):
            s = workloads.make_system(w, precision=prec)
            s.set_energy_matrix(cached)
            s.rng_seed(1000 + r)
            s.propagator_init()
            systems.append(s)
        for _ in range(max(1, args.warmup)):
            fgx.propagate_batch(systems, 1)        # includes the cache build
        res[key] = timed_batch(fgx, systems, args.steps)
        if key == "mixed" and not args.no_error_check:
            # the same chains without the cache (FP32 pairs for old and new)
            for s in systems:
                s.set_energy_matrix(False)
            res["mixed_uncached"] = timed_batch(fgx, systems, 1)
        for s in systems:
            s.close()
    m, f = res["mixed"], res["fp64"]
    trials_per_launch = m["trials"] / args.steps
    pairs_exec = nsite * (n - nsite)              # cached path: trial state only
    achieved = pairs_exec * FLOPS_PAIR_DHHS * trials_per_launch / float(np.mean(m["dev"])) / 1e12
    # each partner record is read once per trial and reused against all
    # staged sites, so the kernel sits far right of the HBM ridge
    bytes_trial = BYTES_PARTNER["mixed"] * (n - nsite) + 2 * BYTES_PARTNER["mixed"] * nsite
    gbs = bytes_trial * trials_per_launch / float(np.mean(m["dev"])) / 1e9
    line = base_line(args, m["value"], m["wall"], "f32 pairs + f64 accumulate",
                     {"workload": "rigid g2g: 64 molecules x 400 sites (q -1/0/+1 p 0.2/0.6/0.2, "
                                  "radii U[1.5,3] A, 20 A spheres) + 8000 ions, L=250 A, "
                                  "Debye-Hueckel 15 A + hard spheres, TranslateRotate dp 5 A "
                                  "dprot 0.5, sweep = 64 trials",
                      "replicas": R, "trials_per_step": int(trials_per_launch),
                      "precision": "mixed", "energy_matrix": True})
    line.update({
        "pairs_per_s": m["value"] * pairs_ref,
        "acceptance": m["accepted"] / max(1, m["trials"]),
        "sweep_latency_ms": m["sweep_latency_ms"],
        "fp64": {"value": f["value"], "sweep_latency_ms": f["sweep_latency_ms"],
                 "pairs_per_s": f["value"] * pairs_ref},
        "speedup_vs_fp64": m["value"] / f["value"],
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK["fp32"],
                     "unit": "TFLOP/s", "frac": achieved / PEAK["fp32"], "traffic": None,
                     "kernel": "k_sweep_tr<DH,HS,mixed> (cached)",
                     "flops_per_trial": pairs_exec * FLOPS_PAIR_DHHS,
                     "hbm": {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": gbs / HBM_PEAK_GBS, "bytes_per_trial": bytes_trial}},
    })
    if "mixed_uncached" in res:
        u = res["mixed_uncached"]
        line["mixed_uncached"] = {"value": u["value"], "sweep_latency_ms": u["sweep_latency_ms"]}
    if not args.no_error_check:
        try:
            line["du_error_vs_fp64"] = du_error_batched(w, chains=64, ncalls=2, cap=128)
        except Exception as e:
            line["du_error_vs_fp64"] = {"error": str(e)}
    if not args.no_cpu_baseline:
        try:
            line["cpu_baseline"] = cpu_baseline(w)
        except Exception as e:
            line["cpu_baseline"] = {"error": str(e)}
    return line


def temper_run_stats(args, fgx, workloads, precision, rounds, sweeps):
    """48 replicas (6 equal-N ladders of 8) advanced by fgx_temper_run: each
    round is `sweeps` sweeps of every replica in one launch, then one exchange
    attempt per ladder, captured once as a HIP graph and replayed."""
    import torch
    prec = fgx.MIXED if precision == "mixed" else fgx.FP64
    ladders = workloads.pt_ladders(nladder=args.ladders, rungs=args.rungs, precision=prec)
    systems, tempers = [], []
    t_build = time.perf_counter()
    for l, lad in enumerate(ladders):
        reps = []
        for r, w in enumerate(lad["replicas"]):
            s = workloads.make_system(w)
            s.rng_seed(100 * l + r + 1)
            s.propagator_init()
            reps.append(s)
        systems += reps
        tempers.append(fgx.Temper(reps, ctl_seed=11 + l))
    t_build = time.perf_counter() - t_build
    tuned = fgx.autotune_batch(systems, 1) if args.warmup > 0 else None
    fgx.temper_run(tempers, 1, 1)               # warm-up round (graph capture path)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st, dev_s = fgx.temper_run(tempers, rounds, sweeps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat, ex_att, ex_acc = [], 0, 0
    for t in tempers:
        acc, du, slot, ssec = t.records()
        lat.extend((ssec / sweeps).reshape(-1).tolist())
        for row_a, row_d in zip(acc, du):
            ex_att += int(np.isfinite(row_d).sum()) // 2
            ex_acc += int(row_a.sum()) // 2
    pairs = sum(2 * (lad["n"] - 1) * lad["n"] * len(lad["replicas"]) * sweeps * rounds
                for lad in ladders)
    for t in tempers:
        t.close()
    for s in systems:
        s.close()
    return {"value": st["trials"] / wall, "wall": wall, "device_s": dev_s,
            "trials": st["trials"], "accepted": st["accepted"], "pairs_per_s": pairs / wall,
            "sweep_latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)),
                                 "p99": 1e3 * float(np.percentile(lat, 99)),
                                 "max": 1e3 * float(np.max(lat))},
            "round_ms": 1e3 * wall / rounds,
            "exchange_acceptance": ex_acc / max(1, ex_att), "exchange_attempts": ex_att,
            "build_seconds": t_build, "schedule": tuned,
            "ladders": [l["n"] for l in ladders]}


def bench_temper(args, fgx, workloads):
    """Batched parallel tempering: 48 replicas in 6 equal-N ladders, each
    round one launch of `exchange_every` sweeps of all replicas and an
    exchange attempt in every ladder, replayed as one HIP graph; the mixed
    path, and the FP64 path for the same rounds."""
    sweeps = args.exchange_every
    m = temper_run_stats(args, fgx, workloads, "mixed", args.steps, sweeps)
    f = temper_run_stats(args, fgx, workloads, "fp64", max(1, args.steps // 2), sweeps)
    line = base_line(args, m["value"], m["wall"], "f32 pairs + f64 accumulate",
                     {"workload": "batched parallel tempering: %d ladders x %d replicas, N "
                                  "log-uniform 1000-16000 (%s), salt at 500 ions / (60 A)^3, "
                                  "Debye-Hueckel 10 A + hard spheres 2 A, lB 7.1 A at 298.15 K "
                                  "scaled 1/T, T geometric 280-420 K, AtomicTranslation dp 2 A, "
                                  "exchange every %d sweeps, rounds replayed as one HIP graph"
                                  % (args.ladders, args.rungs,
                                     ",".join(str(n) for n in m["ladders"]), sweeps),
                      "replicas": args.ladders * args.rungs, "sweeps_per_step": sweeps,
                      "precision": "mixed"})
    line["ms_per_step"] = m["round_ms"]
    line.update({k: m[k] for k in ("pairs_per_s", "sweep_latency_ms", "exchange_acceptance",
                                   "exchange_attempts", "build_seconds", "schedule")})
    line["acceptance"] = m["accepted"] / max(1, m["trials"])
    line["fp64"] = {k: f[k] for k in ("value", "pairs_per_s", "sweep_latency_ms", "round_ms",
                                      "exchange_acceptance")}
    line["speedup_vs_fp64"] = m["value"] / f["value"]
    return line


def bench_serve(args, fgx, workloads):
    """Served job stream: jobs of N log-uniform 500-50000 (half Ewald
    electrolytes, half Debye-Hueckel salts) arrive as a Poisson process and
    are served by fgx_server under an HBM budget.  Reports job latency
    (arrival to completion), per-job sweep latency and aggregate throughput."""
    import torch
    jobs = workloads.serve_jobs(njobs=args.jobs, seed=args.seed)
    rng = np.random.default_rng(args.seed + 1)
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, size=len(jobs)))
    budget = int(args.hbm_mb * 1024 * 1024)
    sv = fgx.Server(budget, tick_seconds=args.tick_ms * 1e-3)
    ids, arr_of = [], {}
    res = {}
    submitted = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    while len(res) < len(jobs):
        now = time.perf_counter() - t0
        while submitted < len(jobs) and arrivals[submitted] <= now:
            jid = sv.submit(jobs[submitted]["w"], seed=1000 + submitted, sweeps=args.job_sweeps)
            arr_of[jid] = (submitted, arrivals[submitted])
            ids.append(jid)
            submitted += 1
        inf = sv.info()
        if inf["active"] or inf["queued"]:
            sv.step()
        elif submitted < len(jobs):
            time.sleep(max(0.0, min(0.01, arrivals[submitted] - now)))
        for jid in sv.poll():
            r = sv.result(jid)
            r["done_at"] = time.perf_counter() - t0
            res[jid] = r
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    peak = sv.info()["peak"]
    sv.close()
    lat = [r["done_at"] - arr_of[j][1] for j, r in res.items()]
    sweep = [r["mean_sweep_seconds"] for r in res.values()]
    trials = sum(r["trials"] for r in res.values())
    pairs = sum(r["trials"] * 2 * (jobs[arr_of[j][0]]["n"] - 1) for j, r in res.items())
    bad = [j for j, r in res.items() if r["status"] != 0]
    line = base_line(args, trials / wall, wall, "f32 pairs + f64 accumulate",
                     {"workload": "served job stream: %d jobs, Poisson arrivals %.1f jobs/s, N "
                                  "log-uniform 500-50000 at 0.01 A^-3, half Ewald LJ+Coulomb "
                                  "electrolyte (alpha 0.25, rc 12, |n|<=10), half Debye-Hueckel "
                                  "10 A + hard spheres, AtomicTranslation, %d sweeps per job, "
                                  "HBM budget %d MB"
                                  % (len(jobs), args.rate, args.job_sweeps, args.hbm_mb),
                      "jobs": len(jobs), "sweeps_per_job": args.job_sweeps,
                      "precision": "mixed"})
    line["ms_per_step"] = None
    line["steps"] = None
    line.update({
        "pairs_per_s": pairs / wall,
        "job_latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                          "max": float(np.max(lat))},
        "sweep_latency_ms": {"p50": 1e3 * float(np.percentile(sweep, 50)),
                             "p99": 1e3 * float(np.percentile(sweep, 99))},
        "queue_wait_s": {"p50": float(np.percentile([r["queue_seconds"] for r in res.values()], 50)),
                         "p99": float(np.percentile([r["queue_seconds"] for r in res.values()], 99))},
        "peak_resident_mb": peak / 2 ** 20,
        "failed_jobs": len(bad),
        "wall_s": wall,
    })
    return line


def bench_salt(args, fgx, workloads):
    """The 500-ion salt solution (Debye-Hueckel 10 A, hard spheres 2 A, lB 7.1
    A, L = 60 A, AtomicTranslation dp 2 A): batched FP64 chains (the
    reference's own precision for this system) and mixed chains; the
    mixed path's acceptance-mismatch rate against FP64 on identical RNG
    streams, and <U> and g(r) of the two ensembles compared within
    statistical error."""
    from beofaunus_amd import analysis
    import torch
    R = args.replicas or 256
    sweeps = args.sweeps_per_step
    res = {}
    for key, prec in (("fp64", fgx.FP64), ("mixed", fgx.MIXED)):
        systems = []
        for r in range(R):
            w = workloads.salt_dh(seed=1 + r % 16, precision=prec)
            s = workloads.make_system(w)
            s.rng_seed(1000 + r)
            s.propagator_init()
            systems.append(s)
        tuned = fgx.autotune_batch(systems, 1) if args.warmup > 0 else None
        res[key] = timed_batch(fgx, systems, args.steps, sweeps)
        res[key]["schedule"] = tuned
        for s in systems:
            s.close()
    f, m = res["fp64"], res["mixed"]
    line = base_line(args, f["value"], f["wall"], "f64",
                     {"workload": "salt solution: 250 + 250 ions, L=60 A, hard spheres 2 A, "
                                  "Debye-Hueckel 10 A, lB 7.1 A, AtomicTranslation dp 2 A "
                                  "(%d sweeps per step)" % sweeps,
                      "replicas": R, "trials_per_step": int(f["trials"] / args.steps),
                      "precision": "fp64"})
    line.update({
        "pairs_per_s": f["value"] * 2 * 499,
        "acceptance": f["accepted"] / max(1, f["trials"]),
        "sweep_latency_ms": {k: v / sweeps for k, v in f["sweep_latency_ms"].items()},
        "mixed": {"value": m["value"], "pairs_per_s": m["value"] * 2 * 499,
                  "sweep_latency_ms": {k: v / sweeps for k, v in m["sweep_latency_ms"].items()}},
        "schedule": {"fp64": f["schedule"], "mixed": m["schedule"]},
    })
    if not args.no_error_check:
        w = workloads.salt_dh(seed=3)
        errs, mism, compared = [], 0, 0
        for c in range(8):
            e = du_error_vs_fp64(w, seed=50 + c, ntrials=20000, ncalls=40)
            compared += e["compared_trials"]
            mism += e["first_accept_mismatch"] is not None
            if e["max_abs_kT"] is not None:
                errs.append(e["max_abs_kT"])
        line["du_error_vs_fp64"] = {"max_abs_kT": max(errs), "compared_trials": compared,
                                    "chains_diverged": mism,
                                    "accept_mismatch_rate": mism / max(1, compared)}
        out = {}
        for prec in (fgx.FP64, fgx.MIXED):
            systems = []
            for c in range(64):
                w = workloads.salt_dh(seed=300 + c, precision=prec)
                s = workloads.make_system(w)
                s.rng_seed(5000 + c)
                s.propagator_init()
                systems.append(s)
            out[prec] = analysis.sample_chains(systems, 100, 40, 5, [(1, 2), (1, 1)], 0.1)
            for s in systems:
                s.close()
        (ua, ca), (ub, cb) = out[fgx.FP64], out[fgx.MIXED]
        line["ensemble_mixed_vs_fp64"] = analysis.compare(ua, ub, ca, cb, 0.1, 60.0 ** 3)
        line["ensemble_mixed_vs_fp64"]["sample"] = ("64 chains per precision, 100 sweeps "
                                                    "equilibration, 40 samples 5 sweeps apart")
    if not args.no_cpu_baseline:
        try:
            line["cpu_baseline"] = cpu_baseline(workloads.salt_dh(seed=1))
        except Exception as e:
            line["cpu_baseline"] = {"error": str(e)}
    torch.cuda.synchronize()
    return line


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", choices=["ewald", "cells", "salt", "groups", "temper", "serve"],
                    default="ewald",
                    help="ewald: N=20000 Ewald electrolyte (headline); cells: the same with LJ "
                         "cut and shifted at rc and a cell list; salt: 500-ion DH salt; "
                         "groups: 64 x 400-site rigid molecules + 8000 ions; temper: 48-replica "
                         "parallel tempering; serve: 256-job stream through the job server")
    ap.add_argument("--sweeps-per-step", type=int, default=20, help="salt: sweeps per launch")
    ap.add_argument("--jobs", type=int, default=256)
    ap.add_argument("--rate", type=float, default=50.0, help="job arrivals per second")
    ap.add_argument("--job-sweeps", type=int, default=4)
    ap.add_argument("--hbm-mb", type=float, default=256.0)
    ap.add_argument("--tick-ms", type=float, default=50.0)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--ladders", type=int, default=6)
    ap.add_argument("--rungs", type=int, default=8)
    ap.add_argument("--exchange-every", type=int, default=10, help="sweeps between exchanges")
    ap.add_argument("--replicas", type=int, default=0, help="chains per launch (default 256)")
    ap.add_argument("--n", type=int, default=0, help="particles (ewald workload)")
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error-check", action="store_true")
    ap.add_argument("--no-sub", action="store_true",
                    help="headline only: skip the other configurations' sub-results")
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
    if args.config == "ewald":
        line = bench_ewald(args, fgx, workloads)
    elif args.config == "cells":
        line = bench_ewald(args, fgx, workloads, cut=True)
    elif args.config == "salt":
        line = bench_salt(args, fgx, workloads)
    elif args.config == "groups":
        line = bench_groups(args, fgx, workloads)
    elif args.config == "temper":
        line = bench_temper(args, fgx, workloads)
    else:
        line = bench_serve(args, fgx, workloads)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
