"""Edge cases of the device path against the oracle (FP64 and mixed):

* one particle: no partners, every energy change zero, every trial accepted;
* two particles in a small box: hard-sphere overlaps give +inf and are
  rejected (Energy::energyChange, energy.h:3597-3607);
* a group list holding an empty group (front = back + 1, as Group leaves
  one without particles): energies and a Propagator run on the other group;
* zero Propagator calls: nothing changes;
* a neutral atom type under Ewald: NaN energy changes, all accepted, as the
  compiled reference does (fixture ewald_neutral);
* arguments the library refuses (indices out of range, a box of zero length,
  a Propagator without moves) raise errors instead of reaching the device.
"""
import sys

import numpy as np
import pytest

from oracle_py import OracleSystem, lib as olib

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


def salt(n, L, seed, groups=None):
    from beofaunus_amd import workloads
    w = workloads.salt_dh(n=n, L=L, seed=seed)
    if groups is not None:
        w["groups"] = groups
    return w


def oracle_of(w):
    h = w["ham"]
    o = OracleSystem(w["ref_kind"], w["box"], w["xyz"], w["charge"], w["radius"], w["mw"],
                     w["types"], w["groups"], h["lB"], kappa=h["kappa"],
                     type_sigma=h["type_sigma"], type_eps=h["type_eps"],
                     type_charge=h["type_charge"], type_radius=h["type_radius"],
                     type_dp=h["type_dp"])
    olib().or_add_move(o.h, 1, 0, 1.0, 1, 0, None, None, 0.0, 0.0, 1)
    return o


def run_both(gpu, w, ncalls, seed, precision):
    from beofaunus_amd import workloads
    o = oracle_of(w)
    o.seed(seed)
    olib().or_propagator_init(o.h)
    odu, oacc, _, _, _ = o.propagate(ncalls)
    oxyz = o.xyz().reshape(-1)
    s = workloads.make_system(w, precision=precision)
    try:
        s.rng_seed(seed)
        s.propagator_init()
        st, du, acc = s.propagate(ncalls, record=max(1, ncalls * len(w["charge"])))
        xyz = s.positions(0).reshape(-1)
    finally:
        s.close()
    return st, du, acc, xyz, odu, oacc, oxyz


def check_same(du, acc, xyz, odu, oacc, oxyz, tol):
    assert len(du) == len(odu)
    assert np.array_equal(acc, oacc)
    fin = np.isfinite(odu)
    assert np.array_equal(np.isfinite(du), fin)
    assert np.all(np.abs(du[fin] - odu[fin]) <= tol * np.maximum(1.0, np.abs(odu[fin])))
    assert np.max(np.abs(xyz - oxyz)) <= 1e-9


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_single_particle(gpu, prec):
    w = salt(1, 20.0, seed=51)
    say("[edges] one particle, %s" % prec)
    precision = gpu.FP64 if prec == "fp64" else gpu.MIXED
    st, du, acc, xyz, odu, oacc, oxyz = run_both(gpu, w, 5, 3, precision)
    assert st["trials"] == 5 and st["accepted"] == 5
    assert np.all(du == 0.0) and np.all(odu == 0.0)
    check_same(du, acc, xyz, odu, oacc, oxyz, 0.0)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_two_particles_with_overlaps(gpu, prec):
    w = salt(2, 9.0, seed=52)
    say("[edges] two particles in a 9 A box, %s" % prec)
    precision = gpu.FP64 if prec == "fp64" else gpu.MIXED
    st, du, acc, xyz, odu, oacc, oxyz = run_both(gpu, w, 60, 4, precision)
    ninf = int(np.isinf(odu).sum())
    say("  %d trials, %d overlaps, %d accepted" % (len(odu), ninf, oacc.sum()))
    assert ninf > 0 and st["rejected_inf"] == ninf
    check_same(du, acc, xyz, odu, oacc, oxyz, 1e-10 if prec == "fp64" else 1e-6)


def test_empty_group(gpu):
    from beofaunus_amd import workloads
    n = 30
    groups = [{"front": 0, "back": n - 1, "molsize": 1, "molid": 0},
              {"front": n, "back": n - 1, "molsize": 1, "molid": 1}]
    w = salt(n, 25.0, seed=53, groups=groups)
    say("[edges] %d ions plus an empty group" % n)
    o = oracle_of(w)
    s = workloads.make_system(w, precision=gpu.FP64)
    try:
        assert s.g2g(0, 1) == 0.0 and s.g_internal(1) == 0.0 and s.g_external(1) == 0.0
        assert s.system_energy() == pytest.approx(o.system_energy(), rel=1e-12)
        assert s.g_internal(0) == pytest.approx(o.g_internal(0), rel=1e-12)
    finally:
        s.close()
    st, du, acc, xyz, odu, oacc, oxyz = run_both(gpu, w, 4, 5, gpu.FP64)
    check_same(du, acc, xyz, odu, oacc, oxyz, 1e-10)


def test_zero_calls(gpu):
    from beofaunus_amd import workloads
    w = salt(20, 25.0, seed=54)
    s = workloads.make_system(w, precision=gpu.FP64)
    try:
        s.rng_seed(6)
        s.propagator_init()
        before = s.positions(0).copy(), s.rng_state()
        st, _, _ = s.propagate(0)
        assert st["trials"] == 0 and st["utot"] == 0.0
        assert np.array_equal(s.positions(0), before[0])
        assert s.rng_state()[1] == before[1][1]
        assert np.array_equal(s.rng_state()[0], before[1][0])
    finally:
        s.close()


def test_refused_arguments(gpu):
    from beofaunus_amd import workloads
    w = salt(10, 25.0, seed=55)
    s = workloads.make_system(w, precision=gpu.FP64)
    try:
        with pytest.raises(gpu.FgxError):
            s.i2i(0, 10)
        with pytest.raises(gpu.FgxError):
            s.g2g(0, 3)
        with pytest.raises(gpu.FgxError):
            s.set_positions(5, np.zeros((6, 3)))
    finally:
        s.close()
    bad = dict(w, box=[25.0, 0.0, 25.0])
    with pytest.raises(gpu.FgxError):
        workloads.make_system(bad, precision=gpu.FP64)
    s = workloads.make_system(dict(w, move=None), precision=gpu.FP64)
    try:
        with pytest.raises(gpu.FgxError):
            s.propagate(1)
    finally:
        s.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_neutral_type_under_ewald(gpu, prec):
    """A particle of a q = 0 atom type under Ewald: the effective-charge ratio
    0/0 = NaN (ewald.h:535-545) makes S(k) and every energy change NaN, and
    Movebase::metropolis accepts NaN (move.h:567-573).  Fixture ewald_neutral
    is the compiled reference's run (102 trials, all NaN, all accepted)."""
    from fixture_systems import fgx_system_from_fixture
    from oracle_py import farr, load_fixture
    fx = load_fixture("ewald_neutral")
    mc = fx["mc"]
    say("[edges] neutral type under Ewald, %s" % prec)
    s = fgx_system_from_fixture(fx, precision=gpu.FP64 if prec == "fp64" else gpu.MIXED)
    try:
        s.rng_seed(fx["rng_seed"])
        s.propagator_init()
        st, du, acc = s.propagate(mc["ncalls"], record=100000)
        x = s.positions().reshape(-1)
    finally:
        s.close()
    assert len(du) == len(mc["du"]) and np.all(np.isnan(farr(mc["du"])))
    assert np.all(np.isnan(du)) and np.all(acc == 1) and st["accepted"] == len(du)
    assert np.allclose(x, farr(mc["final_xyz"]), rtol=0, atol=1e-9)
