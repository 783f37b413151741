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


def drain(sv, positions=True):
    out = {}
    for _ in range(10000):
        inf = sv.info()
        if not inf["active"] and not inf["queued"]:
            break
        sv.step()
        for j in sv.poll():
            out[j] = sv.result(j, positions=positions)
    for j in sv.poll():
        out[j] = sv.result(j, positions=positions)
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
