"""The device against the reference at the benchmark configurations' own
sizes and parameters (tests/golden/baseline, written by the compiled
reference):

  - 500-ion salt (Debye-Hueckel + hard spheres, L 60): the first 1e5
    AtomicTranslation trials;
  - 20,000-ion Ewald LJ + Coulomb electrolyte (alpha 0.25, rc 12, |n| <= 10,
    K = 2,242): energies, k-vector count and the first 2,000 trials;
  - 64 x 400-site rigid molecules + 8,000 ions (L 250, Debye-Hueckel 15):
    energies and one TranslateRotate sweep of 64 trials.

FP64 path: identical accept sequences, energies and energy changes within
1e-10 relative (absolute floor 1e-10 x the largest |dU| of the run, as the
reference forms dU from two full sums), final coordinates within 1e-9 A, the
global engine's final state identical.  Mixed path: energy changes within
1e-5 kT of the reference on the trials before the first accept mismatch
(none expected), with the mismatch reported.
"""
import sys

import numpy as np
import pytest

from baseline_fixtures import dec, load, mt_text_words
from oracle_py import fnum

pytestmark = pytest.mark.gpu

REL = 1e-10
MIXED_TOL = 1e-5


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


def device(fgx, fx, precision):
    from fixture_systems import fgx_system_from_fixture
    s = fgx_system_from_fixture(fx, precision=precision)
    s.rng_seed(fx["rng_seed"])
    s.propagator_init()
    return s


def rel_close(a, b, scale, rel=REL):
    return abs(a - b) <= rel * max(abs(b), scale, 1.0)


def check_fp64_run(s, fx, du, acc, idx=None):
    mc = fx["mc"]
    ref_acc = dec(mc["acc"])
    assert len(acc) == len(ref_acc)
    bad = np.nonzero(acc != ref_acc)[0]
    assert len(bad) == 0, "first accept mismatch at trial %d" % bad[0]
    ref = dec(mc["du"])
    got = du if idx is None else du[idx]
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(got), fin)
    scale = max(1.0, float(np.max(np.abs(ref[fin]))))
    err = np.abs(got[fin] - ref[fin])
    assert np.all(err <= REL * np.maximum(np.abs(ref[fin]), scale)), float(err.max())
    words, pos = s.rng_state()
    rw, rp = mt_text_words(mc["final_rng"])
    assert pos == rp and np.array_equal(np.asarray(words, dtype=np.uint32), rw)
    return float(err.max()) if err.size else 0.0


def check_mixed_run(fx, du, acc, idx=None, min_prefix=1000):
    mc = fx["mc"]
    ref_acc = dec(mc["acc"])
    bad = np.nonzero(acc != ref_acc)[0]
    stop = int(bad[0]) if len(bad) else len(acc)
    ref = dec(mc["du"])
    if idx is None:
        idx = np.arange(len(ref))
    keep = idx < stop
    got, r = du[idx[keep]], ref[keep]
    fin = np.isfinite(r)
    err = float(np.max(np.abs(got[fin] - r[fin]))) if fin.any() else 0.0
    say("  mixed: %d trials before the first accept mismatch (of %d), max |dU err| %.3g kT"
        % (stop, len(acc), err))
    assert stop >= min(min_prefix, len(acc))
    assert err <= MIXED_TOL
    return stop, err


def test_c1_salt_1e5_trials_fp64(gpu):
    fgx = gpu
    fx = load("c1_salt_n500")
    say("[baseline] C1 salt, 500 ions, 1e5 trials, FP64")
    s = device(fgx, fx, fgx.FP64)
    st, du, acc = s.propagate(fx["mc"]["ncalls"], record=fx["mc"]["trials"])
    assert st["trials"] == 100000
    err = check_fp64_run(s, fx, du, acc, dec(fx["mc"]["du_index"]))
    x = s.positions().reshape(-1)
    assert np.max(np.abs(x - dec(fx["mc"]["final_xyz"]))) <= 1e-9
    say("  max |dU err| %.3g" % err)
    s.close()


def test_c1_salt_1e5_trials_mixed(gpu):
    fgx = gpu
    fx = load("c1_salt_n500")
    say("[baseline] C1 salt, 500 ions, 1e5 trials, mixed")
    s = device(fgx, fx, fgx.MIXED)
    st, du, acc = s.propagate(fx["mc"]["ncalls"], record=fx["mc"]["trials"])
    check_mixed_run(fx, du, acc, dec(fx["mc"]["du_index"]), min_prefix=20000)
    s.close()


def test_c2_electrolyte_energies_fp64(gpu):
    fgx = gpu
    fx = load("c2_ewald_n20000")
    e = fx["energy"]
    say("[baseline] C2 electrolyte, N=20000, energies")
    s = device(fgx, fx, fgx.FP64)
    assert s.ewald_k() == e["ewald_nk"] == 2242
    scale = max(abs(float(v)) for v in e["i2all"])
    for k, i in enumerate(e["i2all_i"]):
        assert rel_close(s.i2all(i), fnum(e["i2all"][k]), scale), i
    for k, (i, j) in enumerate(zip(e["i2i_i"], e["i2i_j"])):
        assert rel_close(s.i2i(i, j), float(e["i2i"][k]), 1.0), (i, j)
    assert rel_close(s.external(), fnum(e["external"]), 1.0)
    assert rel_close(s.system_energy(), fnum(e["system_energy"]), 1.0)
    s.close()


def test_c2_electrolyte_2000_trials_fp64(gpu):
    fgx = gpu
    fx = load("c2_ewald_n20000")
    mc = fx["mc"]
    say("[baseline] C2 electrolyte, N=20000, 2000 trials, FP64")
    s = device(fgx, fx, fgx.FP64)
    st, du, acc = s.propagate(mc["ncalls"], record=mc["trials"])
    assert st["trials"] == 2000
    err = check_fp64_run(s, fx, du, acc)
    moved = dec(mc["moved"])
    x = s.positions()
    assert np.max(np.abs(x[moved].reshape(-1) - dec(mc["final_moved_xyz"]).reshape(-1))) <= 1e-9
    say("  max |dU err| %.3g" % err)
    s.close()


def test_c2_electrolyte_2000_trials_mixed(gpu):
    fgx = gpu
    fx = load("c2_ewald_n20000")
    say("[baseline] C2 electrolyte, N=20000, 2000 trials, mixed")
    s = device(fgx, fx, fgx.MIXED)
    st, du, acc = s.propagate(fx["mc"]["ncalls"], record=fx["mc"]["trials"])
    check_mixed_run(fx, du, acc, min_prefix=2000)
    s.close()


def test_c3_rigid_energies_fp64(gpu):
    fgx = gpu
    fx = load("c3_rigid_64x400")
    e = fx["energy"]
    say("[baseline] C3 rigid molecules, energies")
    s = device(fgx, fx, fgx.FP64)
    ng = len(fx["groups"])
    g2g = e["g2g"].reshape(len(e["g2g_rows"]), ng)
    fin = np.isfinite(g2g)
    scale = max(1.0, float(np.max(np.abs(g2g[fin]))))
    for r, a in enumerate(e["g2g_rows"]):
        for b in range(ng):
            v = s.g2g(a, b)
            if np.isfinite(g2g[r, b]):
                assert rel_close(v, g2g[r, b], scale), (a, b)
            else:
                assert v == g2g[r, b] or (np.isnan(v) and np.isnan(g2g[r, b])), (a, b)
    for k, i in enumerate(e["i2all_i"]):
        ref = fnum(e["i2all"][k])
        v = s.i2all(i)
        assert (v == ref) if not np.isfinite(ref) else rel_close(v, ref, scale), i
    s.close()


def test_c3_rigid_sweep_fp64(gpu):
    fgx = gpu
    fx = load("c3_rigid_64x400")
    mc = fx["mc"]
    say("[baseline] C3 rigid molecules, 64 TranslateRotate trials, FP64")
    s = device(fgx, fx, fgx.FP64)
    st, du, acc = s.propagate(mc["ncalls"], record=mc["trials"])
    assert st["trials"] == mc["trials"] == 64
    err = check_fp64_run(s, fx, du, acc)
    sites = dec(mc["sites"])
    x = s.positions()
    assert np.max(np.abs(x[sites].reshape(-1) - dec(mc["final_sites_xyz"]).reshape(-1))) <= 1e-9
    cm = dec(mc["final_cm"]).reshape(-1, 3)
    for g in range(len(fx["groups"]) - 1):
        assert np.max(np.abs(np.asarray(s.group_cm(g)[0]) - cm[g])) <= 1e-9, g
    say("  max |dU err| %.3g" % err)
    s.close()


def test_c3_rigid_sweep_mixed(gpu):
    fgx = gpu
    fx = load("c3_rigid_64x400")
    say("[baseline] C3 rigid molecules, 64 TranslateRotate trials, mixed (cached)")
    s = device(fgx, fx, fgx.MIXED)
    st, du, acc = s.propagate(fx["mc"]["ncalls"], record=fx["mc"]["trials"])
    check_mixed_run(fx, du, acc, min_prefix=64)
    s.close()
