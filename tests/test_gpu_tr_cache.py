"""Mixed TranslateRotate sweep with the group-pair energy cache.

The cache (after Energy::EnergyMatrix, energy.h:225-641) keeps the accepted
state's g2g energies per group pair, so a trial evaluates only the moved
group's new energies.  Checked against the FP64 device path (which follows
the reference's two-pass energyChange and is itself pinned to the golden
vectors) trial by trial, against the uncached mixed path, and for energy
drift: the sum of accepted dU must match the change of the FP64 system
energy.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_DU = 1e-5      # kT, mixed path vs FP64


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


def run(s, seed, ncalls, record=20000):
    s.rng_seed(seed)
    s.propagator_init()
    return s.propagate(ncalls, record=record)


def compare(du_a, acc_a, du_b, acc_b):
    n = min(len(du_a), len(du_b))
    mism = np.nonzero(acc_a[:n] != acc_b[:n])[0]
    upto = int(mism[0]) if len(mism) else n
    fa, fb = np.isfinite(du_a[:upto]), np.isfinite(du_b[:upto])
    assert np.array_equal(fa, fb), "finite/infinite dU pattern differs"
    err = np.abs(du_a[:upto][fa] - du_b[:upto][fa])
    return upto, (float(err.max()) if len(err) else 0.0), (int(mism[0]) if len(mism) else None)


@pytest.mark.parametrize("name", ["mol_dhhs", "mol_dhhs_cut"])
def test_cache_vs_fp64_and_uncached(gpu, name):
    from fixture_systems import fgx_system_from_fixture
    from oracle_py import load_fixture
    say("[tr-cache] %s" % name)
    fx = load_fixture(name)
    ncalls = fx["mc"]["ncalls"]
    res = {}
    for key, prec, cached in (("fp64", gpu.FP64, False), ("cached", gpu.MIXED, True),
                              ("plain", gpu.MIXED, False)):
        s = fgx_system_from_fixture(fx, precision=prec)
        s.set_energy_matrix(cached)
        st, du, acc = run(s, fx["rng_seed"], ncalls)
        res[key] = (du, acc)
        s.close()
    upto, err, mism = compare(*res["cached"], *res["fp64"])
    say("  cached vs FP64: compared %d, max |err| %.3e kT, first mismatch %s" % (upto, err, mism))
    assert upto > 0 and err <= TOL_DU
    upto, err2, mism = compare(*res["cached"], *res["plain"])
    say("  cached vs uncached mixed: compared %d, max |diff| %.3e kT" % (upto, err2))
    assert err2 <= 2 * TOL_DU


def test_cache_many_groups_drift(gpu):
    """A small rigid-molecule system (12 molecules + salt) over many sweeps:
    trial dU agrees with FP64 while the chains agree, and the accumulated dU
    of the cached chain equals the FP64 system-energy change (no drift from
    the cache's row/column updates)."""
    from beofaunus_amd import workloads
    w = workloads.rigid_molecules(nmol=12, nsite=60, nion=400, L=70.0, R=7.0, seed=5)
    say("[tr-cache] rigid_molecules 12 x 60 + 400 ions")
    out = {}
    for key, prec in (("fp64", gpu.FP64), ("cached", gpu.MIXED)):
        s = workloads.make_system(w, precision=prec)
        st, du, acc = run(s, 17, 40)
        out[key] = (du, acc, st, s.positions())
        s.close()
    upto, err, mism = compare(out["cached"][0], out["cached"][1], out["fp64"][0], out["fp64"][1])
    say("  trials %d, compared %d, max |err| %.3e kT, first mismatch %s, acceptance %.3f"
        % (len(out["cached"][0]), upto, err, mism, out["cached"][2]["accepted"] / out["cached"][2]["trials"]))
    assert upto > 100 and err <= TOL_DU
    # drift of the cached chain, energies evaluated in FP64 on its end state
    # (inter-group energy: the rigid molecules' internal energy is constant,
    # and infinite here, since their sites overlap)
    du, acc, st, xyz_end = out["cached"]
    s64 = workloads.make_system(w, precision=gpu.FP64)
    ng = len(w["groups"])

    def inter():
        return sum(s64.g2g(a, b) for a in range(ng) for b in range(a + 1, ng))
    e0 = inter()
    s64.set_positions(0, xyz_end, which=2)
    e1 = inter()
    s64.close()
    drift = (e1 - e0) - st["utot"]
    say("  FP64 U_end - U_start = %.6f, sum dU = %.6f, drift %.3e kT" % (e1 - e0, st["utot"], drift))
    assert abs(drift) <= 1e-5 * max(1, st["accepted"]) ** 0.5 + 1e-6 * abs(e1 - e0)
