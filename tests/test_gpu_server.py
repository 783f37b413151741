"""Job server (fgx_server): continuous batching under an HBM budget.

A job's result must not depend on what else is on the device: the same job
run alone and in a mix (other Hamiltonians, other sizes, other tick splits)
ends in the same configuration and the same accumulated energy.  Chains are
independent workgroups with fixed-order reductions and one fixed launch
schedule, so the stated tolerance is exact equality of positions and
|utot difference| <= 1e-12 kT.  Admission must keep the resident footprint
within the budget while every job still completes.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from beofaunus_amd import fgx
    return fgx


def gpu_fgx():
    from beofaunus_amd import fgx
    return fgx


def drain(sv, positions=True):
    def res(j):
        r = sv.result(j)
        if positions and r["status"] == 0:
            try:
                r = sv.result(j, positions=True)
            except gpu_fgx().FgxError:   # positions not kept for this job
                pass
        return r
    out = {}
    for _ in range(100000):
        inf = sv.info()
        if not inf["active"] and not inf["queued"]:
            break
        sv.step()
        for j in sv.poll():
            out[j] = res(j)
    for j in sv.poll():
        out[j] = res(j)
    return out


def test_alone_equals_in_mix(gpu):
    from beofaunus_amd import workloads
    say("[server] alone vs in mix")
    a = workloads.salt_dh(n=300, L=40.0, seed=21, lattice=True)
    b = workloads.ewald_electrolyte(n=600, L=40.0, seed=22)
    c = workloads.salt_dh(n=900, L=60.0, seed=23, lattice=True)
    alone = gpu.Server(1 << 30, tick_seconds=0.001)
    ja = alone.submit(a, seed=5, sweeps=6, keep_positions=True)
    ra = drain(alone)[ja]
    alone.close()
    mix = gpu.Server(1 << 30, tick_seconds=0.02)
    jb = mix.submit(b, seed=6, sweeps=3, keep_positions=True)
    jc = mix.submit(c, seed=7, sweeps=2, keep_positions=True)
    ja2 = mix.submit(a, seed=5, sweeps=6, keep_positions=True)
    out = drain(mix)
    mix.close()
    rm = out[ja2]
    say("  alone: trials %d utot %.12f; in mix: trials %d utot %.12f"
        % (ra["trials"], ra["utot"], rm["trials"], rm["utot"]))
    assert ra["status"] == 0 and rm["status"] == 0
    assert out[jb]["status"] == 0 and out[jc]["status"] == 0
    assert ra["trials"] == rm["trials"] == 6 * 300
    assert ra["accepted"] == rm["accepted"]
    assert np.array_equal(ra["xyz"], rm["xyz"])
    # the sum over a job's launches groups differently with the tick boundaries
    assert abs(ra["utot"] - rm["utot"]) <= 1e-12 * max(1.0, abs(ra["utot"]))


def test_cell_list_job_alone_equals_in_mix(gpu):
    """A cut-LJ Ewald job (cell-list sweep) gives the same result alone and
    served beside jobs of other kinds, through per-tick trial budgets that
    stop and resume its Propagator::move() calls."""
    from beofaunus_amd import workloads
    say("[server] cell-list job alone vs in mix")
    a = workloads.ewald_electrolyte(n=2000, L=55.7, seed=24, lj_cut=True)
    b = workloads.ewald_electrolyte(n=600, L=40.0, seed=25)
    c = workloads.salt_dh(n=500, L=50.0, seed=26, lattice=True)
    alone = gpu.Server(1 << 30, tick_seconds=0.5)
    ja = alone.submit(a, seed=8, sweeps=3, keep_positions=True)
    ra = drain(alone)[ja]
    alone.close()
    mix = gpu.Server(1 << 30, tick_seconds=0.002)
    jb = mix.submit(b, seed=9, sweeps=2, keep_positions=True)
    ja2 = mix.submit(a, seed=8, sweeps=3, keep_positions=True)
    jc = mix.submit(c, seed=10, sweeps=4, keep_positions=True)
    out = drain(mix)
    mix.close()
    rm = out[ja2]
    say("  alone: trials %d utot %.12f; in mix: trials %d utot %.12f"
        % (ra["trials"], ra["utot"], rm["trials"], rm["utot"]))
    assert ra["status"] == 0 and rm["status"] == 0
    assert out[jb]["status"] == 0 and out[jc]["status"] == 0
    assert ra["trials"] == rm["trials"] == 3 * 2000
    assert ra["accepted"] == rm["accepted"]
    assert np.array_equal(ra["xyz"], rm["xyz"])
    # the sum over a job's launches groups differently with the tick boundaries
    assert abs(ra["utot"] - rm["utot"]) <= 1e-12 * max(1.0, abs(ra["utot"]))


def test_rigid_molecule_job_alone_equals_in_mix(gpu):
    """A TranslateRotate job (rigid molecules beside ions, mixed path with the
    group-pair energy cache) gives the same result alone and beside an
    AtomicTranslation job, through per-tick trial budgets."""
    from beofaunus_amd import workloads
    say("[server] TranslateRotate job alone vs in mix")
    a = workloads.rigid_molecules(nmol=6, nsite=40, nion=300, L=90.0, seed=33)
    b = workloads.salt_dh(n=800, L=60.0, seed=34, lattice=True)
    alone = gpu.Server(1 << 30, tick_seconds=0.5)
    ja = alone.submit(a, seed=11, sweeps=20, keep_positions=True)
    ra = drain(alone)[ja]
    alone.close()
    mix = gpu.Server(1 << 30, tick_seconds=0.002)
    jb = mix.submit(b, seed=12, sweeps=3, keep_positions=True)
    ja2 = mix.submit(a, seed=11, sweeps=20, keep_positions=True)
    out = drain(mix)
    mix.close()
    rm = out[ja2]
    say("  alone: trials %d utot %.12f; in mix: trials %d utot %.12f"
        % (ra["trials"], ra["utot"], rm["trials"], rm["utot"]))
    assert ra["status"] == 0 and rm["status"] == 0 and out[jb]["status"] == 0
    assert ra["trials"] == rm["trials"] > 0
    assert ra["accepted"] == rm["accepted"]
    assert np.array_equal(ra["xyz"], rm["xyz"])
    assert abs(ra["utot"] - rm["utot"]) <= 1e-12 * max(1.0, abs(ra["utot"]))


def test_split_jobs_alone_equal_in_mix(gpu):
    """Jobs large enough for several workgroups (block sweeps spread over
    one XCD) give the same result alone and in a mix whose split jobs do not
    all fit the device at once (some wait for a later tick), beside
    single-workgroup jobs; each job's workgroup count is its own."""
    from beofaunus_amd import workloads
    say("[server] split jobs alone vs in a mix that overfills the XCDs")
    rho = 0.01
    a = workloads.salt_dh(n=6000, L=(6000 / rho) ** (1 / 3), seed=81, lattice=True)
    e = workloads.ewald_electrolyte(n=3000, L=(3000 / rho) ** (1 / 3), seed=82)
    big = [workloads.salt_dh(n=12300, L=(12300 / rho) ** (1 / 3), seed=90 + k, lattice=True)
           for k in range(9)]
    small = [workloads.salt_dh(n=500, L=(500 / rho) ** (1 / 3), seed=83, lattice=True),
             workloads.ewald_electrolyte(n=600, L=(600 / rho) ** (1 / 3), seed=84)]
    alone = {}
    for w, sd, sw in ((a, 5, 2), (e, 6, 1)):
        sv = gpu.Server(1 << 33, tick_seconds=0.01)
        j = sv.submit(w, seed=sd, sweeps=sw, keep_positions=True)
        alone[sd] = drain(sv)[j]
        sv.close()
        say("  alone N=%d: %d workgroups, trials %d, utot %.12f"
            % (len(w["charge"]), alone[sd]["workgroups"], alone[sd]["trials"], alone[sd]["utot"]))
    assert alone[5]["workgroups"] == 11 and alone[6]["workgroups"] == 5
    sv = gpu.Server(1 << 34, tick_seconds=0.01)
    ids = {}
    for k, w in enumerate(big):
        ids[100 + k] = sv.submit(w, seed=100 + k, sweeps=1)
    ids[5] = sv.submit(a, seed=5, sweeps=2, keep_positions=True)
    for k, w in enumerate(small):
        ids[200 + k] = sv.submit(w, seed=200 + k, sweeps=3)
    ids[6] = sv.submit(e, seed=6, sweeps=1, keep_positions=True)
    out = drain(sv)
    sv.close()
    assert all(r["status"] == 0 for r in out.values())
    assert all(out[ids[100 + k]]["workgroups"] == 24 for k in range(len(big)))
    for sd in (5, 6):
        r, al = out[ids[sd]], alone[sd]
        say("  in mix: %d workgroups, trials %d, utot %.12f" % (r["workgroups"], r["trials"], r["utot"]))
        assert r["workgroups"] == al["workgroups"]
        assert r["trials"] == al["trials"] and r["accepted"] == al["accepted"]
        assert np.array_equal(r["xyz"], al["xyz"])
        assert abs(r["utot"] - al["utot"]) <= 1e-12 * max(1.0, abs(al["utot"]))


def test_served_stream_split_jobs_complete(gpu):
    """A stream like the served-stream benchmark's (64 jobs, N 500-50,000,
    half Ewald, half Debye-Hueckel, mixed, 2 sweeps each, arriving while
    earlier ones run): every job completes, every job from N = 1,024 up runs
    split, and the split jobs' launches (run alone each tick) never time out
    waiting for a workgroup."""
    import time
    from beofaunus_amd import workloads
    say("[server] 64-job stream with split jobs")
    jobs = workloads.serve_jobs(njobs=64, seed=5)
    sv = gpu.Server(256 << 20, tick_seconds=0.05)
    ids, out, k = {}, {}, 0
    t0 = time.perf_counter()
    while len(out) < len(jobs):
        # four arrivals per tick until all are in
        for _ in range(4):
            if k < len(jobs):
                ids[sv.submit(jobs[k]["w"], seed=1000 + k, sweeps=2)] = k
                k += 1
        inf = sv.info()
        if inf["active"] or inf["queued"]:
            sv.step()
        for j in sv.poll():
            out[j] = sv.result(j)
        assert time.perf_counter() - t0 < 240, "stream did not finish"
    sv.close()
    bad = [(jobs[ids[j]]["n"], r["status"]) for j, r in out.items() if r["status"] != 0]
    say("  %d jobs, %d failed, %d split" % (len(out), len(bad),
                                           sum(1 for r in out.values() if r["workgroups"] > 1)))
    assert not bad, bad
    for j, r in out.items():
        n = jobs[ids[j]]["n"]
        assert r["sweeps"] == 2 and r["trials"] == 2 * n
        assert r["workgroups"] == min(24, max(1, n // 512)), (n, r["workgroups"])


def test_budget_admission(gpu):
    """A budget that holds about one job at a time: every job completes, the
    peak resident footprint stays within the budget, and later jobs wait."""
    from beofaunus_amd import workloads
    say("[server] admission under a small budget")
    ws = [workloads.salt_dh(n=400 + 100 * k, L=45.0, seed=30 + k, lattice=True) for k in range(4)]
    probe = workloads.make_system(ws[-1])
    one = probe.device_bytes
    probe.close()
    budget = int(1.5 * one)
    sv = gpu.Server(budget, tick_seconds=0.005)
    ids = [sv.submit(w, seed=40 + k, sweeps=3) for k, w in enumerate(ws)]
    out = drain(sv, positions=False)
    peak = sv.info()["peak"]
    sv.close()
    say("  budget %d B, one job ~%d B, peak %d B" % (budget, one, peak))
    assert sorted(out) == sorted(ids)
    assert all(out[j]["status"] == 0 and out[j]["sweeps"] == 3 for j in ids)
    assert peak <= budget
    assert max(out[j]["queue_seconds"] for j in ids) > 0


def test_unsupported_job_fails_alone(gpu):
    """A job the device sweeps cannot run (an AtomicTranslation of molecular
    groups) is retired at admission with its own error; the jobs beside it
    finish as they do alone."""
    from beofaunus_amd import workloads
    say("[server] one unsupported job beside good ones")
    a = workloads.salt_dh(n=300, L=40.0, seed=41, lattice=True)
    b = workloads.salt_dh(n=500, L=50.0, seed=42, lattice=True)
    bad = workloads.rigid_molecules(nmol=3, nsite=20, nion=100, L=60.0, R=8.0, seed=43)
    bad["move"] = dict(kind=gpu.MOVE_AT, molid=0, peratom=True)
    alone = {}
    for w, sd in ((a, 1), (b, 2)):
        sv = gpu.Server(1 << 30, tick_seconds=0.01)
        j = sv.submit(w, seed=sd, sweeps=3, keep_positions=True)
        alone[sd] = drain(sv)[j]
        sv.close()
    sv = gpu.Server(1 << 30, tick_seconds=0.01)
    ja = sv.submit(a, seed=1, sweeps=3, keep_positions=True)
    jx = sv.submit(bad, seed=3, sweeps=3, keep_positions=True)
    jb = sv.submit(b, seed=2, sweeps=3, keep_positions=True)
    out = drain(sv)
    sv.close()
    assert out[jx]["status"] == -5   # FGX_EUNSUPPORTED
    for j, sd in ((ja, 1), (jb, 2)):
        assert out[j]["status"] == 0
        assert out[j]["trials"] == alone[sd]["trials"]
        assert np.array_equal(out[j]["xyz"], alone[sd]["xyz"])


def _oracle_of(w):
    from oracle_py import OracleSystem, lib as olib
    h = w["ham"]
    kw = dict(type_sigma=h["type_sigma"], type_eps=h["type_eps"], type_charge=h["type_charge"],
              type_radius=h["type_radius"], type_dp=h["type_dp"])
    if h.get("ewald"):
        kw["ewald"] = h["ewald"]
    else:
        kw["kappa"] = h["kappa"]
    o = OracleSystem(w["ref_kind"], w["box"], w["xyz"], w["charge"], w["radius"], w["mw"],
                     w["types"], w["groups"], h["lB"], **kw)
    olib().or_add_move(o.h, 1, 0, 1.0, 0, 0, None, None, 0.0, 0.0, 1)
    return o


def test_stream_sizes_big_jobs(gpu):
    """The served stream's largest sizes: a 50,000-ion Debye-Hueckel job and a
    20,000-ion Ewald job (FP64, 2,000 one-trial Propagator calls each) served
    beside small mixed-precision jobs.  Each big job ends as it does alone,
    and as the oracle's run of the same 2,000 trials (same accepts, final
    positions within 1e-9 A)."""
    from beofaunus_amd import workloads
    from oracle_py import lib as olib
    say("[server] N=50000 DH and N=20000 Ewald jobs beside small jobs")
    rho = 0.01
    big_dh = workloads.salt_dh(n=50000, L=(50000 / rho) ** (1 / 3), seed=51, lattice=True)
    big_ew = workloads.ewald_electrolyte(n=20000, L=(20000 / rho) ** (1 / 3), seed=52)
    for w in (big_dh, big_ew):
        w["move"] = dict(kind=gpu.MOVE_AT, molid=0, peratom=False)
    small = [workloads.salt_dh(n=500, L=(500 / rho) ** (1 / 3), seed=53, lattice=True),
             workloads.ewald_electrolyte(n=800, L=(800 / rho) ** (1 / 3), seed=54),
             workloads.salt_dh(n=2000, L=(2000 / rho) ** (1 / 3), seed=55, lattice=True)]
    calls = 2000
    alone = {}
    for w, sd in ((big_dh, 61), (big_ew, 62)):
        sv = gpu.Server(1 << 33, tick_seconds=0.05)
        j = sv.submit(w, seed=sd, sweeps=calls, precision=gpu.FP64, keep_positions=True)
        alone[sd] = drain(sv)[j]
        sv.close()
        say("  alone N=%d: trials %d accepted %d" % (len(w["charge"]), alone[sd]["trials"],
                                                    alone[sd]["accepted"]))
    sv = gpu.Server(1 << 33, tick_seconds=0.02)
    ids = {}
    ids[61] = sv.submit(big_dh, seed=61, sweeps=calls, precision=gpu.FP64, keep_positions=True)
    for k, w in enumerate(small):
        sv.submit(w, seed=70 + k, sweeps=2)
    ids[62] = sv.submit(big_ew, seed=62, sweeps=calls, precision=gpu.FP64, keep_positions=True)
    out = drain(sv)
    sv.close()
    assert all(r["status"] == 0 for r in out.values())
    for sd, w in ((61, big_dh), (62, big_ew)):
        r, a = out[ids[sd]], alone[sd]
        assert r["trials"] == a["trials"] == calls
        assert r["workgroups"] == a["workgroups"] == 24
        assert r["accepted"] == a["accepted"]
        assert np.array_equal(r["xyz"], a["xyz"])
        o = _oracle_of(w)
        o.seed(sd)
        olib().or_propagator_init(o.h)
        du, acc, who, utot, uinit = o.propagate(calls, cap=calls + 1)
        assert int(acc.sum()) == r["accepted"]
        assert np.max(np.abs(o.xyz(0) - r["xyz"])) <= 1e-9
        say("  N=%d served in the mix: matches alone and the oracle (%d accepted)"
            % (len(w["charge"]), r["accepted"]))
