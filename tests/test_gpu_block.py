"""Block AtomicTranslation sweep (fgx_set_block_sweep): B consecutive trials
of a Propagator call are evaluated against one snapshot of the accepted
configuration and decided in order, with exact corrections for the block's
own moves; a chain may be spread over several workgroups that exchange
partial sums once per block.

Checked against the compiled reference's Metropolis runs (the golden
fixtures and the benchmark configurations at their own sizes): identical
accept sequences, energy changes within 1e-10 relative in FP64 (the
reference forms dU from two full sums, so the absolute floor is 1e-10 x the
largest |dU| of the run) and within 1e-5 kT on the mixed path, final
coordinates within 1e-9 A, identical RNG engine states; and, many chains at
once, against the one-trial sweep of the same chains.
"""
import glob
import math
import os
import sys

import numpy as np
import pytest

from oracle_py import GOLDEN, farr, fnum, load_fixture

pytestmark = pytest.mark.gpu

REL = 1e-10
MIXED_TOL = 1e-5

FIXTURES = sorted(os.path.basename(f)[:-5] for f in glob.glob(os.path.join(GOLDEN, "*.json"))
                  if not f.endswith("rng.json"))
AT_MC = [f for f in FIXTURES if "mc" in load_fixture(f) and
         all(m["type"] == "AT" for m in load_fixture(f)["moves"])]


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


def block_system(fgx, fx, precision, nsplit):
    from fixture_systems import fgx_system_from_fixture
    s = fgx_system_from_fixture(fx, precision=precision)
    s.rng_seed(fx["rng_seed"])
    s.propagator_init()
    s.set_block_sweep(True, nsplit)
    return s


def check_reference(fx, du, acc, st):
    mc = fx["mc"]
    ref_du = farr(mc["du"])
    assert st["trials"] == len(ref_du)
    ref_acc = np.array(mc["acc"], dtype=np.int32)
    bad = np.nonzero(acc != ref_acc)[0]
    assert len(bad) == 0, "first accept mismatch at trial %d" % bad[0]
    assert np.array_equal(np.isinf(du), np.isinf(ref_du))
    assert np.array_equal(np.isnan(du), np.isnan(ref_du))
    fin = np.isfinite(ref_du)
    err = 0.0
    if fin.any():
        scale = max(1.0, np.max(np.abs(ref_du[fin])))
        e = np.abs(du[fin] - ref_du[fin])
        assert np.all(e <= REL * np.maximum(np.abs(ref_du[fin]), scale)), e.max()
        err = float(e.max())
    return err


@pytest.mark.parametrize("nsplit", [1, 3])
@pytest.mark.parametrize("name", AT_MC)
def test_block_sweep_matches_reference_fp64(gpu, name, nsplit):
    fgx = gpu
    fx = load_fixture(name)
    say("[block fp64] %s nsplit %d" % (name, nsplit))
    mc = fx["mc"]
    s = block_system(fgx, fx, fgx.FP64, nsplit)
    st, du, acc = s.propagate(mc["ncalls"], record=100000)
    assert s.block_parts() == nsplit, "the block sweep did not run"
    err = check_reference(fx, du, acc, st)
    x = s.positions().reshape(-1)
    assert np.allclose(x, farr(mc["final_xyz"]), rtol=0, atol=1e-9)
    ut = fnum(mc["utot"])
    if math.isnan(ut):
        assert math.isnan(st["utot"])
    else:
        assert st["utot"] == pytest.approx(ut, rel=1e-9, abs=1e-9)
    say("  max |dU err| %.3g" % err)
    s.close()


@pytest.mark.parametrize("nsplit", [1, 2])
@pytest.mark.parametrize("name", [f for f in AT_MC if f != "ewald_neutral"])
def test_block_sweep_matches_reference_mixed(gpu, name, nsplit):
    fgx = gpu
    fx = load_fixture(name)
    mc = fx["mc"]
    say("[block mixed] %s nsplit %d" % (name, nsplit))
    s = block_system(fgx, fx, fgx.MIXED, nsplit)
    st, du, acc = s.propagate(mc["ncalls"], record=100000)
    parts = s.block_parts()
    if parts != nsplit:
        # no fixed-unit mixed form for this Hamiltonian or box: one-trial sweep
        assert parts == 0
        s.close()
        pytest.skip("no mixed block instance for %s" % name)
    ref_du = farr(mc["du"])
    ref_acc = np.array(mc["acc"], dtype=np.int32)
    bad = np.nonzero(acc != ref_acc)[0]
    stop = int(bad[0]) if len(bad) else len(acc)
    fin = np.isfinite(ref_du[:stop])
    err = float(np.max(np.abs(du[:stop][fin] - ref_du[:stop][fin]))) if fin.any() else 0.0
    say("  %d of %d trials before the first accept mismatch, max |dU err| %.3g kT"
        % (stop, len(acc), err))
    assert stop == len(acc)
    assert err <= MIXED_TOL
    s.close()


def test_block_sweep_c1_salt_1e5_trials(gpu):
    """500-ion salt (Debye-Hueckel + hard spheres), the reference's first
    1e5 trials, FP64 in one and in four workgroups, and mixed"""
    from baseline_fixtures import dec, load, mt_text_words
    fgx = gpu
    fx = load("c1_salt_n500")
    mc = fx["mc"]
    idx = dec(mc["du_index"])
    ref_acc = dec(mc["acc"])
    ref = dec(mc["du"])
    for precision, nsplit in ((fgx.FP64, 1), (fgx.FP64, 4), (fgx.MIXED, 1)):
        say("[block C1] precision %d nsplit %d" % (precision, nsplit))
        s = block_system(fgx, fx, precision, nsplit)
        st, du, acc = s.propagate(mc["ncalls"], record=mc["trials"])
        assert s.block_parts() == nsplit
        assert st["trials"] == 100000
        bad = np.nonzero(acc != ref_acc)[0]
        assert len(bad) == 0, "first accept mismatch at trial %d" % bad[0]
        got = du[idx]
        fin = np.isfinite(ref)
        assert np.array_equal(np.isfinite(got), fin)
        e = np.abs(got[fin] - ref[fin])
        if precision == fgx.FP64:
            scale = max(1.0, float(np.max(np.abs(ref[fin]))))
            assert np.all(e <= REL * np.maximum(np.abs(ref[fin]), scale)), float(e.max())
            x = s.positions().reshape(-1)
            assert np.max(np.abs(x - dec(mc["final_xyz"]))) <= 1e-9
            words, pos = s.rng_state()
            rw, rp = mt_text_words(mc["final_rng"])
            assert pos == rp and np.array_equal(np.asarray(words, dtype=np.uint32), rw)
        else:
            assert float(e.max()) <= MIXED_TOL
        say("  max |dU err| %.3g" % float(e.max()))
        s.close()


@pytest.mark.parametrize("precision,nsplit", [("fp64", 1), ("fp64", 4), ("mixed", 1),
                                              ("mixed", 8)])
def test_block_sweep_c2_electrolyte_2000_trials(gpu, precision, nsplit):
    """20,000-ion Ewald electrolyte at its own parameters (K = 2,242): the
    reference's first 2,000 trials"""
    from baseline_fixtures import dec, load, mt_text_words
    fgx = gpu
    fx = load("c2_ewald_n20000")
    mc = fx["mc"]
    prec = fgx.FP64 if precision == "fp64" else fgx.MIXED
    say("[block C2] %s nsplit %d" % (precision, nsplit))
    s = block_system(fgx, fx, prec, nsplit)
    st, du, acc = s.propagate(mc["ncalls"], record=mc["trials"])
    assert s.block_parts() == nsplit
    assert st["trials"] == 2000
    ref_acc = dec(mc["acc"])
    ref = dec(mc["du"])
    bad = np.nonzero(acc != ref_acc)[0]
    assert len(bad) == 0, "first accept mismatch at trial %d" % bad[0]
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(du), fin)
    e = np.abs(du[fin] - ref[fin])
    if prec == fgx.FP64:
        scale = max(1.0, float(np.max(np.abs(ref[fin]))))
        assert np.all(e <= REL * np.maximum(np.abs(ref[fin]), scale)), float(e.max())
        moved = dec(mc["moved"])
        x = s.positions()
        assert np.max(np.abs(x[moved].reshape(-1) - dec(mc["final_moved_xyz"]).reshape(-1))) <= 1e-9
        words, pos = s.rng_state()
        rw, rp = mt_text_words(mc["final_rng"])
        assert pos == rp and np.array_equal(np.asarray(words, dtype=np.uint32), rw)
    else:
        assert float(e.max()) <= MIXED_TOL
    say("  max |dU err| %.3g, accept %.3f" % (float(e.max()), acc.mean()))
    s.close()


@pytest.mark.parametrize("kind", ["dhhs", "ewald"])
def test_block_sweep_batch_equals_one_trial_sweep(gpu, kind):
    """Many chains of different sizes in one launch, parts chosen
    automatically: every chain's accept sequence and energy changes equal the
    one-trial sweep's (FP64), and the chains' final states agree"""
    from beofaunus_amd import workloads
    fgx = gpu
    sizes = [40, 700, 2600, 5000] if kind == "dhhs" else [400, 1500, 3000]
    systems, ref = [], []
    for k, n in enumerate(sizes):
        if kind == "dhhs":   # the 500-ion salt's density
            w = workloads.salt_dh(n, L=60.0 * (n / 500.0) ** (1 / 3), seed=11 + k)
        else:                # the 20,000-ion electrolyte's density, a shorter cutoff
            w = workloads.ewald_electrolyte(n, L=120.0 * (n / 20000.0) ** (1 / 3), rc=8.0,
                                            alpha=0.35, kc=4.0, seed=11 + k)
        for blk in (True, False):
            s = workloads.make_system(w, precision=fgx.FP64)
            s.rng_seed(100 + k)
            s.propagator_init()
            s.set_block_sweep(blk, 0)
            (systems if blk else ref).append(s)
    say("[block batch] %s sizes %s" % (kind, sizes))
    st, dus, accs = fgx.propagate_batch_record(systems, 2, 20000)
    st2, dus2, accs2 = fgx.propagate_batch_record(ref, 2, 20000)
    parts = [s.block_parts() for s in systems]
    say("  parts %s" % parts)
    assert max(parts) > 1
    for c in range(len(sizes)):
        assert np.array_equal(accs[c], accs2[c]), "chain %d: accept sequences differ" % c
        fin = np.isfinite(dus2[c])
        assert np.array_equal(np.isfinite(dus[c]), fin)
        scale = max(1.0, float(np.max(np.abs(dus2[c][fin]))))
        e = np.abs(dus[c][fin] - dus2[c][fin])
        assert np.all(e <= REL * np.maximum(np.abs(dus2[c][fin]), scale)), float(e.max())
        assert np.max(np.abs(systems[c].positions() - ref[c].positions())) <= 1e-9
        w1, p1 = systems[c].rng_state()
        w2, p2 = ref[c].rng_state()
        assert p1 == p2 and np.array_equal(np.asarray(w1), np.asarray(w2))
    for s in systems + ref:
        s.close()


def test_block_sweep_refuses_splits_that_cannot_be_resident(gpu):
    """A batch whose explicit splits cannot all be resident at once (nine
    chains of 32 workgroups each on 8 XCDs of 32 compute units) is refused
    before anything runs; the same chains then run with one workgroup each."""
    from beofaunus_amd import workloads
    fgx = gpu
    systems = []
    for k in range(9):
        w = workloads.salt_dh(1000, L=60.0 * 2.0 ** (1 / 3), seed=40 + k, lattice=True)
        s = workloads.make_system(w, precision=fgx.MIXED)
        s.rng_seed(k + 1)
        s.propagator_init()
        s.set_block_sweep(True, 32)
        systems.append(s)
    say("[block] nine chains of 32 workgroups")
    with pytest.raises(fgx.FgxError):
        fgx.propagate_batch(systems, 1)
    for s in systems:
        s.set_block_sweep(True, 1)
    st = fgx.propagate_batch(systems, 1)
    assert st["trials"] == 9 * 1000
    assert all(s.block_parts() == 1 for s in systems)
    for s in systems:
        s.close()
