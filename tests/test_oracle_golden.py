"""Pin the CPU oracle against the reference.

Golden vectors come from the reference's own C++ hot path compiled from its
sources (tests/golden/make_golden.py); the closed-form checks come from the
reference's unit tests (src/examples/unittests.cpp) and the C++ standard.
"""
import glob
import math
import os

import numpy as np
import pytest

from oracle_py import init_propagator, fnum, GOLDEN, OracleSystem, farr, lib, load_fixture, dptr

FIXTURES = sorted(os.path.basename(f)[:-5] for f in glob.glob(os.path.join(GOLDEN, "*.json"))
                  if not f.endswith("rng.json"))


def draws(seed, kind, n, a=0, b=0):
    out = np.zeros(n)
    lib().or_rng_draws(seed, kind, a, b, n, dptr(out))
    return out


def test_mt19937_standard_known_answer():
    # C++ [rand.predef]: the 10000th consecutive invocation of a
    # default-constructed mt19937 produces 4123659995.
    assert draws(5489, 0, 10000)[-1] == 4123659995


@pytest.mark.parametrize("key,seed", [("seed_default", 5489), ("seed_7", 7)])
def test_rng_stream_matches_libstdcxx(key, seed):
    fx = load_fixture("rng")[key]
    n = len(fx["u32"])
    assert np.array_equal(draws(seed, 0, n), np.array(fx["u32"], dtype=float))
    # the harness draws u32, canonical, range(0,9), range(0,499), range(0,0),
    # half from one engine in that order; replay the same stream
    from oracle_py import C
    L = lib()

    class MT(C.Structure):
        _fields_ = [("x", C.c_uint32 * 624), ("p", C.c_int)]
    L.or_mt_seed.argtypes = [C.POINTER(MT), C.c_uint32]
    L.or_mt_u32.argtypes = [C.POINTER(MT)]
    L.or_mt_u32.restype = C.c_uint32
    L.or_canonical.argtypes = [C.POINTER(MT)]
    L.or_canonical.restype = C.c_double
    L.or_range.argtypes = [C.POINTER(MT), C.c_int, C.c_int]
    L.or_range.restype = C.c_int
    m = MT()
    L.or_mt_seed(C.byref(m), seed)
    for _ in range(n):
        L.or_mt_u32(C.byref(m))
    can = [L.or_canonical(C.byref(m)) for _ in range(n)]
    assert can == fx["canonical"]
    assert [L.or_range(C.byref(m), 0, 9) for _ in range(n)] == fx["range_0_9"]
    assert [L.or_range(C.byref(m), 0, 499) for _ in range(n)] == fx["range_0_499"]
    assert [L.or_range(C.byref(m), 0, 0) for _ in range(n)] == fx["range_0_0"]
    assert [L.or_canonical(C.byref(m)) - 0.5 for _ in range(n)] == fx["half"]


def test_bjerrum_length_known_answers():
    # unittests.cpp:51 -- lB(epsr=1, 298.15 K) = 560.455786334 A
    e, e0, kB, pi = 1.602177e-19, 8.85419e-12, 1.380658e-23, math.acos(-1.0)
    lb = e * e / (4 * pi * e0 * 1.0 * 1e-10 * kB * 298.15)
    assert lb == pytest.approx(560.455786334, rel=1e-10)
    # LastTest.log:218 -- lB(epsr=80) = 7.0057 A; fixture lj_mix stores the
    # reference's own value for epsr=80
    assert load_fixture("lj_mix")["energy"]["lB"] == pytest.approx(7.0057, abs=5e-5)


@pytest.mark.parametrize("name", FIXTURES)
def test_energies_match_reference(name):
    fx = load_fixture(name)
    e = fx["energy"]
    s = OracleSystem.from_fixture(fx)
    ng = len(fx["groups"])
    rel = 1e-13

    def close(a, b):
        b = float(b) if not isinstance(b, str) else float(b)
        if math.isinf(b) or math.isnan(b):
            return (math.isnan(a) and math.isnan(b)) or a == b
        return abs(a - b) <= rel * max(1.0, abs(b))

    for k, (i, j) in enumerate(zip(e["i2i_i"], e["i2i_j"])):
        assert close(s.i2i(i, j), e["i2i"][k]), (i, j)
    for k, i in enumerate(e["i2all_i"]):
        assert close(s.i2all(i), e["i2all"][k]), i
        for g in range(ng):
            assert close(s.i2g(g, i), e["i2g"][k * ng + g]), (i, g)
    for a in range(ng):
        for b in range(ng):
            assert close(s.g2g(a, b), e["g2g"][a * ng + b]), (a, b)
        assert close(s.g_internal(a), e["g_internal"][a])
        assert close(s.g_external(a), e["g_external"][a])
    assert close(s.system_energy(), e["system_energy"])
    assert close(s.external(), e["external"])


def is_npt(fx):
    return any(m["type"] == "NPT" for m in fx.get("moves", []))


@pytest.mark.parametrize("name", [f for f in FIXTURES if "mc" in load_fixture(f)
                                  and not is_npt(load_fixture(f))])
def test_metropolis_sequence_matches_reference(name):
    fx = load_fixture(name)
    mc = fx["mc"]
    s = OracleSystem.from_fixture(fx)
    s.seed(fx["rng_seed"])
    lib().or_propagator_init(s.h)
    du, acc, who, utot, uinit = s.propagate(mc["ncalls"])
    ref_du = farr(mc["du"])
    assert len(du) == len(ref_du)
    # accept/reject: bit-exact sequence
    assert np.array_equal(acc, np.array(mc["acc"], dtype=np.int32))
    assert np.array_equal(who, np.array(mc["who"], dtype=np.int32))
    # energy changes: the oracle follows the reference's arithmetic order
    # (including Eigen's reduction order), so every value is bit-identical
    # (NaN where the reference gives NaN: a neutral type under Ewald)
    assert np.array_equal(du, ref_du, equal_nan=True)
    # final configuration is bit-identical too
    assert np.array_equal(s.xyz().reshape(-1), farr(mc["final_xyz"]))
    assert np.array_equal(utot, fnum(mc["utot"]), equal_nan=True)
    assert np.array_equal(uinit, fnum(mc["uinit"]), equal_nan=True)


def npt_index(fx):
    """position of the Isobaric move in the Propagator's move list: the
    input's move keys in nlohmann's (sorted) order"""
    keys = sorted(k for k in fx["config"]["moves"] if k in ("atomtranslate", "isobaric",
                                                               "moltransrot"))
    return keys.index("isobaric")


@pytest.mark.parametrize("name", [f for f in FIXTURES if is_npt(load_fixture(f))
                                  and load_fixture(f)["mc"].get("mode") == "mcprop"])
def test_mixed_propagator_matches_reference(name):
    """One Propagator holding AtomicTranslation and Isobaric, picked per call
    (move.h:7877-7891): the picks, every trial's dU and accept flag, the
    final box, coordinates and engine state as the compiled reference left
    them."""
    fx = load_fixture(name)
    mc = fx["mc"]
    mv = [m for m in fx["moves"] if m["type"] == "NPT"][0]
    s = OracleSystem.from_fixture(fx)
    init_propagator(s.seed, lambda: lib().or_propagator_init(s.h), fx)
    du, acc, who, utot = s.propagate_npt(mv["dp"], mv["prob"], npt_index(fx), mc["ncalls"])
    assert (who == -1).sum() > 0 and (who >= 0).sum() > 0
    assert np.array_equal(who, np.array(mc["who"], dtype=np.int32))
    assert np.array_equal(acc, np.array(mc["acc"], dtype=np.int32))
    assert np.array_equal(du, farr(mc["du"]))
    assert np.array_equal(s.box(), farr(mc["final_len"]))
    assert np.array_equal(s.xyz().reshape(-1), farr(mc["final_xyz"]))
    assert utot == mc["utot"]
    assert s.system_energy() == pytest.approx(mc["ufinal"], rel=1e-12)


@pytest.mark.parametrize("name", [f for f in FIXTURES if is_npt(load_fixture(f))
                                  and load_fixture(f)["mc"].get("mode") != "mcprop"])
def test_isobaric_matches_reference(name):
    """Move::Isobaric with ExternalPressure (move.h:5424-5640, energy.h:1817-1894):
    volume trials, accept sequence, final box and coordinates as the compiled
    reference left them."""
    fx = load_fixture(name)
    mc = fx["mc"]
    mv = [m for m in fx["moves"] if m["type"] == "NPT"][0]
    s = OracleSystem.from_fixture(fx)
    s.seed(fx["rng_seed"])
    lib().or_propagator_init(s.h)
    du, acc, utot = s.isobaric_propagate(mv["dp"], mv["prob"], mc["ncalls"])
    nvol = len(du)
    if mc.get("ncalls_at"):
        # the same Propagator goes on with AtomicTranslation calls: with
        # Ewald they start from the structure factors the last accepted
        # volume move left (those of the unscaled coordinates)
        du2, acc2, who2, utot2, _ = s.propagate(mc["ncalls_at"])
        assert nvol == mc["nvol"]
        assert np.array_equal(who2, np.array(mc["who"][nvol:], dtype=np.int32))
        du, acc = np.concatenate([du, du2]), np.concatenate([acc, acc2])
        assert utot == mc["utot_vol"]
        utot = utot + utot2
    ref_du = farr(mc["du"])
    assert len(du) == len(ref_du)
    assert np.array_equal(acc, np.array(mc["acc"], dtype=np.int32))
    assert acc[:nvol].sum() > 0 and (acc[:nvol] == 0).sum() > 0
    # the oracle follows the reference's arithmetic: bit-identical
    assert np.array_equal(du, ref_du)
    assert np.array_equal(s.box(), farr(mc["final_len"]))
    assert np.array_equal(s.xyz().reshape(-1), farr(mc["final_xyz"]))
    if mc.get("ncalls_at"):
        assert utot == pytest.approx(mc["utot"], rel=1e-13, abs=1e-13)
    else:
        assert utot == mc["utot"]
    assert s.system_energy() == pytest.approx(mc["ufinal"], rel=1e-12)
