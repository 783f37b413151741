"""The oracle against the reference at the benchmark configurations' own
sizes and parameters (tests/golden/baseline): the 500-ion salt's first 1e5
AtomicTranslation trials, the 20,000-ion Ewald electrolyte (alpha 0.25,
rc 12, |n| <= 10) energies, k-vector count and first 2,000 trials, and the
64 x 400-site rigid-molecule system's energies.  Accept sequences exact,
energies within 1e-12 relative (the oracle follows the reference's
summation order)."""
import numpy as np
import pytest

from baseline_fixtures import dec, load
from oracle_py import OracleSystem, fnum, lib as olib


def rel_close(a, b, scale, rel=1e-12):
    if not np.isfinite(b):   # hard-sphere overlaps inside a rigid molecule: +inf
        return a == b or (np.isnan(a) and np.isnan(b))
    return abs(a - b) <= rel * max(abs(b), scale, 1.0)


def test_c1_salt_1e5_trials():
    fx = load("c1_salt_n500")
    mc = fx["mc"]
    o = OracleSystem.from_fixture(fx)
    o.seed(fx["rng_seed"])
    olib().or_propagator_init(o.h)
    du, acc, who, utot, uinit = o.propagate(mc["ncalls"], cap=mc["trials"] + 1)
    assert len(du) == mc["trials"] == 100000
    assert np.array_equal(acc, dec(mc["acc"]))
    idx = dec(mc["du_index"])
    ref = dec(mc["du"])
    assert np.array_equal(du[idx], ref)
    assert np.array_equal(o.xyz(0).reshape(-1), dec(mc["final_xyz"]))
    assert rel_close(uinit, fnum(mc["uinit"]), 1.0)


def test_c2_electrolyte_energies_and_2000_trials():
    fx = load("c2_ewald_n20000")
    e, mc = fx["energy"], fx["mc"]
    o = OracleSystem.from_fixture(fx)
    assert olib().or_ewald_nk(o.h) == e["ewald_nk"] == 2242
    scale = max(abs(float(v)) for v in e["i2all"])
    for k, i in enumerate(e["i2all_i"][:10]):
        assert rel_close(o.i2all(i), fnum(e["i2all"][k]), scale), i
    assert rel_close(o.system_energy(), fnum(e["system_energy"]), 1.0)
    o.seed(fx["rng_seed"])
    olib().or_propagator_init(o.h)
    du, acc, who, utot, uinit = o.propagate(mc["ncalls"], cap=mc["trials"] + 1)
    assert np.array_equal(acc, dec(mc["acc"]))
    assert np.array_equal(who, dec(mc["who"]))
    ref = dec(mc["du"])
    assert np.max(np.abs(du - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    moved = dec(mc["moved"])
    assert np.array_equal(o.xyz(0)[moved].reshape(-1), dec(mc["final_moved_xyz"]).reshape(-1))


def test_c3_rigid_molecules_energies():
    fx = load("c3_rigid_64x400")
    e = fx["energy"]
    o = OracleSystem.from_fixture(fx)
    ng = len(fx["groups"])
    g2g = e["g2g"].reshape(len(e["g2g_rows"]), ng)
    scale = max(1.0, float(np.max(np.abs(g2g[np.isfinite(g2g)]))))
    for r, a in enumerate(e["g2g_rows"][:4]):
        for b in range(ng - 1):
            assert rel_close(o.g2g(a, b), g2g[r, b], scale), (a, b)
    for k, i in enumerate(e["i2all_i"][:6]):
        assert rel_close(o.i2all(i), fnum(e["i2all"][k]), scale), i
