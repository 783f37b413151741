"""Parity of the HIP engine (FP64 path) with the reference.

Expected values are the golden vectors the compiled reference wrote
(tests/golden/*.json) and the CPU oracle, which reproduces those vectors
bit for bit (tests/test_oracle_golden.py).

Tolerances: energies and energy changes agree to 1e-10 relative, with an
absolute floor of 1e-10 x (sum of |pair terms|) scale -- the reference sums
sequentially, the GPU in a fixed tree order, and for atomic moves it sums
the per-pair differences u(i',j) - u(i,j) directly instead of two full
configuration energies (energy.h:3521-3572), so the reference's own rounding
of the two large sums sets the floor.  Accept/reject sequences (integer
work) are compared exactly.
"""
import glob
import math
import os
import sys

import numpy as np
import pytest

from oracle_py import GOLDEN, KINDS, OracleSystem, farr, fnum, lib as olib, load_fixture

pytestmark = pytest.mark.gpu

FIXTURES = sorted(os.path.basename(f)[:-5] for f in glob.glob(os.path.join(GOLDEN, "*.json"))
                  if not f.endswith("rng.json"))
REL = 1e-10


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from beofaunus_amd import fgx
    info = fgx.device_info()
    say("device %s, %d CUs" % (info["arch"], info["cus"]))
    return fgx


def close(a, b, scale=1.0, rel=REL):
    if isinstance(b, str):
        b = float(b)
    if math.isinf(b) or math.isnan(b):
        return (math.isnan(a) and math.isnan(b)) or a == b
    return abs(a - b) <= rel * max(abs(b), scale, 1.0)


@pytest.mark.parametrize("name", FIXTURES)
def test_energies_fp64(gpu, name):
    from fixture_systems import fgx_system_from_fixture
    say("[energies] %s" % name)
    fx = load_fixture(name)
    e = fx["energy"]
    s = fgx_system_from_fixture(fx)
    ng = len(fx["groups"])
    scale = max(1.0, max(abs(float(v)) for v in e["g2g"] if not isinstance(v, str)))
    for k, (i, j) in enumerate(zip(e["i2i_i"], e["i2i_j"])):
        assert close(s.i2i(i, j), e["i2i"][k]), (i, j)
    for k, i in enumerate(e["i2all_i"][:12]):
        assert close(s.i2all(i), e["i2all"][k], scale), i
        for g in range(ng):
            assert close(s.i2g(g, i), e["i2g"][k * ng + g], scale), (i, g)
    for a in range(ng):
        for b in range(ng):
            assert close(s.g2g(a, b), e["g2g"][a * ng + b], scale), (a, b)
        assert close(s.g_internal(a), e["g_internal"][a], scale)
        assert close(s.g_external(a), e["g_external"][a], scale)
    assert close(s.external(), e["external"], scale)
    assert close(s.system_energy(), e["system_energy"], scale)
    s.close()


@pytest.mark.parametrize("name", [f for f in FIXTURES if "mc" in load_fixture(f) and
                                  all(m["type"] in ("AT", "TR") for m in load_fixture(f)["moves"])])
def test_device_propagator_matches_reference(gpu, name):
    """fgx_propagate replays the reference's Propagator on the device: same
    RNG stream, same accept/reject sequence, same energy changes."""
    from fixture_systems import fgx_system_from_fixture
    say("[propagate] %s" % name)
    fx = load_fixture(name)
    mc = fx["mc"]
    s = fgx_system_from_fixture(fx)
    s.rng_seed(fx["rng_seed"])
    s.propagator_init()
    st, du, acc = s.propagate(mc["ncalls"], record=100000)
    ref_du = farr(mc["du"])
    assert st["trials"] == len(ref_du)
    assert np.array_equal(acc, np.array(mc["acc"], dtype=np.int32)), \
        "first mismatch at %d" % int(np.argmax(acc != np.array(mc["acc"])))
    assert np.array_equal(np.isinf(du), np.isinf(ref_du))
    assert np.array_equal(np.isnan(du), np.isnan(ref_du))   # NaN: accepted, as there
    fin = np.isfinite(ref_du)
    if fin.any():
        scale = max(1.0, np.max(np.abs(ref_du[fin])))
        err = np.abs(du[fin] - ref_du[fin])
        assert np.all(err <= REL * np.maximum(np.abs(ref_du[fin]), scale)), err.max()
    # final configuration: trial coordinates follow the reference arithmetic
    x = s.positions().reshape(-1)
    assert np.allclose(x, farr(mc["final_xyz"]), rtol=0, atol=1e-9)
    ut = fnum(mc["utot"])
    if math.isnan(ut):
        assert math.isnan(st["utot"])
    else:
        assert st["utot"] == pytest.approx(ut, rel=1e-9, abs=1e-9)
    # RNG stream position identical: continue both and compare the next draws
    s.close()


@pytest.mark.parametrize("name", ["salt_dhhs", "salt_coulhs", "lj_mix", "lj_cutshift",
                                  "ewald_tinfoil", "ewald_surface", "mol_dhhs", "mol_dhhs_cut"])
def test_per_call_energy_change(gpu, name):
    """Host-driven drop-in path: the oracle proposes each trial; the GPU
    computes Energy::energyChange for it and commits the oracle's decision."""
    from fixture_systems import fgx_system_from_fixture
    say("[energy_change] %s" % name)
    fx = load_fixture(name)
    o = OracleSystem.from_fixture(fx)
    s = fgx_system_from_fixture(fx, moves=False)
    rng = np.random.default_rng(5)
    box = np.array(fx["box"])
    ng = len(fx["groups"])
    mol = [g for g in range(ng) if fx["groups"][g]["molsize"] > 1]
    for t in range(40):
        if mol and t % 2 == 0:
            # whole-group translation of a molecule (TranslateRotate-like)
            g = int(rng.choice(mol))
            G = fx["groups"][g]
            xyz = o.xyz(0)
            d = (rng.random(3) - 0.5) * 4.0
            new = xyz[G["front"]:G["back"] + 1] + d
            new -= box * np.round(new / box)
            for k, i in enumerate(range(G["front"], G["back"] + 1)):
                o.set_pos(i, new[k])
            s.set_positions(G["front"], new, which=1)
            cm, _ = s.group_cm(g)
            cmt = cm + d
            cmt -= box * np.round(cmt / box)
            s.set_group_cm(g, cm_trial=cmt)
            _set_oracle_cm(o, g, cmt)
            mv = [(g, [])]
            idx = None
        else:
            ga = [g for g in range(ng) if fx["groups"][g]["molsize"] == 1][0]
            G = fx["groups"][ga]
            i = int(rng.integers(G["front"], G["back"] + 1))
            p = o.xyz(0)[i] + (rng.random(3) - 0.5) * 2.0
            p -= box * np.round(p / box)
            o.set_pos(i, p)
            s.set_positions(i, p.reshape(1, 3), which=1)
            g, mv, idx = ga, [(ga, [i])], [i]
        if o.prm.elec == 3:
            _oracle_update_change(o, mv)
        ref = o.energy_change(mv)
        s.update_change(g, idx)
        du = s.energy_change(g, idx)
        scale = 1.0 + abs(ref)
        if math.isinf(ref):
            assert du == ref
        else:
            assert abs(du - ref) <= 1e-9 * scale, (t, du, ref)
        accept = bool(rng.random() < 0.5) and math.isfinite(ref)
        o.commit(mv, accept)
        if o.prm.elec == 3:
            olib().or_ewald_update(o.h, int(accept))
        s.commit(g, idx, accept)
    assert np.allclose(s.positions(0), o.xyz(0), atol=1e-12)
    s.close()


def _set_oracle_cm(o, g, cmt):
    import ctypes as C
    # or_sys keeps cm_trial internally; recompute it the way the test moved it
    lib = olib()
    lib.or_set_cm_trial.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
    a = np.ascontiguousarray(cmt, dtype=np.float64)
    lib.or_set_cm_trial(o.h, g, a.ctypes.data_as(C.POINTER(C.c_double)))


def _oracle_update_change(o, mv):
    import ctypes as C
    g = np.array([m[0] for m in mv], dtype=np.int32)
    off = np.zeros(len(mv) + 1, dtype=np.int32)
    idx = []
    for k, m in enumerate(mv):
        idx += list(m[1])
        off[k + 1] = len(idx)
    ia = np.array(idx + [0], dtype=np.int32)
    P = C.POINTER(C.c_int)
    olib().or_ewald_update_change(o.h, len(mv), g.ctypes.data_as(P), off.ctypes.data_as(P),
                                  ia.ctypes.data_as(P))


@pytest.mark.parametrize("name", ["salt_dhhs", "salt_coulhs", "lj_mix", "lj_cutshift",
                                  "ewald_tinfoil", "ewald_kc12", "mol_dhhs", "mol_dhhs_cut"])
def test_mixed_precision_error(gpu, name):
    """Mixed path (FP32 pairs on fixed-point coordinates, FP64 accumulation,
    FP64 S(k) and Metropolis): |dU - dU_FP64| <= 1e-5 kT while the two
    chains agree, and the acceptance-mismatch count is reported."""
    from fixture_systems import fgx_system_from_fixture
    say("[mixed] %s" % name)
    fx = load_fixture(name)
    mc = fx["mc"]
    s = fgx_system_from_fixture(fx, precision=gpu.MIXED)
    s.rng_seed(fx["rng_seed"])
    s.propagator_init()
    st, du, acc = s.propagate(mc["ncalls"], record=100000)
    ref_du = farr(mc["du"])
    ref_acc = np.array(mc["acc"], dtype=np.int32)
    n = min(len(du), len(ref_du))
    mism = np.nonzero(acc[:n] != ref_acc[:n])[0]
    upto = mism[0] if len(mism) else n
    fin = np.isfinite(ref_du[:upto]) & np.isfinite(du[:upto])
    err = np.abs(du[:upto][fin] - ref_du[:upto][fin])
    say("  trials %d, compared %d, max |dU err| %.3e kT, first accept mismatch %s"
        % (n, upto, err.max() if len(err) else 0.0, mism[0] if len(mism) else None))
    assert len(err) > 0
    assert err.max() <= 1e-5
    s.close()


@pytest.mark.parametrize("name", ["ewald_tinfoil", "ewald_surface", "salt_dhhs"])
def test_i_external(gpu, name):
    """NonbondedEwald::i_external (ewald.h:707-710): each particle's self
    energy; their sum is the whole group's g_external, and the oracle's
    self-energy sum (getSelfEnergy over all particles).  Without Ewald: 0."""
    from fixture_systems import fgx_system_from_fixture
    fx = load_fixture(name)
    s = fgx_system_from_fixture(fx, moves=False)
    o = OracleSystem.from_fixture(fx)
    try:
        n = len(fx["particles"]["charge"])
        ui = np.array([s.i_external(i) for i in range(n)])
        if KINDS[fx["kind"]][0] != "ewald":
            assert np.all(ui == 0.0)
            return
        assert np.all(ui < 0.0)
        total = olib().or_ewald_self(o.h, 0)
        assert ui.sum() == pytest.approx(total, rel=1e-12)
        assert ui.sum() == pytest.approx(s.g_external(0), rel=1e-12)
    finally:
        s.close()


@pytest.mark.parametrize("name", ["ewald_tinfoil", "ewald_surface"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ewald_state_in_device_memory(gpu, name, precision):
    """The Ewald sweep with S(k), dS(k), A_k and n in device memory
    (fgx_set_ewald_state) at k-vector counts that also fit LDS: FP64 against
    the reference's trials (identical accepts, dU to 1e-10), mixed within
    1e-5 kT of them."""
    from fixture_systems import fgx_system_from_fixture
    say("[ewald state in device memory] %s %s" % (name, precision))
    fx = load_fixture(name)
    mc = fx["mc"]
    s = fgx_system_from_fixture(fx, precision=gpu.FP64 if precision == "fp64" else gpu.MIXED)
    s.set_ewald_state(True)
    s.rng_seed(fx["rng_seed"])
    s.propagator_init()
    st, du, acc = s.propagate(mc["ncalls"], record=100000)
    ref_du = farr(mc["du"])
    ref_acc = np.array(mc["acc"], dtype=np.int32)
    fin = np.isfinite(ref_du)
    if precision == "fp64":
        assert np.array_equal(acc, ref_acc)
        scale = max(1.0, np.max(np.abs(ref_du[fin])))
        assert np.all(np.abs(du[fin] - ref_du[fin]) <= REL * np.maximum(np.abs(ref_du[fin]), scale))
        assert np.allclose(s.positions().reshape(-1), farr(mc["final_xyz"]), rtol=0, atol=1e-9)
    else:
        mism = np.nonzero(acc != ref_acc)[0]
        upto = mism[0] if len(mism) else len(acc)
        f = fin[:upto] & np.isfinite(du[:upto])
        assert upto > 0 and np.max(np.abs(du[:upto][f] - ref_du[:upto][f])) <= 1e-5
    s.close()
