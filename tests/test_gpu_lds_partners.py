"""LDS-resident partners of the mixed AtomicTranslation sweep (no Ewald,
chains of up to 2,048 particles): against the same sweep reading its
partners from device memory, on identical RNG streams.  Identical accept
sequences, final positions and engine states; dU equal to FP64 rounding of
the sum (the two loops visit partners in other batches).  Batches that mix
chains that fit with chains that do not; Coulomb, Debye-Hueckel + hard
spheres and Lennard-Jones; and the mixed sweep against the reference's
own trials through the fixtures is covered by the parity tests, which take
this path by default."""
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


def systems(fgx, specs, lds):
    from beofaunus_amd import workloads
    out = []
    for k, (maker, kw) in enumerate(specs):
        w = getattr(workloads, maker)(precision=fgx.MIXED, **kw)
        s = workloads.make_system(w)
        s.set_lds_partners(lds)
        s.rng_seed(31 + k)
        s.propagator_init()
        out.append(s)
    return out


def run_batch(fgx, specs, lds, ncalls, cap):
    ss = systems(fgx, specs, lds)
    st, dus, accs = fgx.propagate_batch_record(ss, ncalls, cap)
    pos = [s.positions().copy() for s in ss]
    rng = [s.rng_state() for s in ss]
    for s in ss:
        s.close()
    return dus, accs, pos, rng


def compare(a, b, tol=1e-9):
    dus_a, acc_a, pos_a, rng_a = a
    dus_b, acc_b, pos_b, rng_b = b
    worst = 0.0
    for c in range(len(dus_a)):
        assert np.array_equal(acc_a[c], acc_b[c]), "chain %d: accept sequences differ" % c
        fa, fb = np.isfinite(dus_a[c]), np.isfinite(dus_b[c])
        assert np.array_equal(fa, fb)
        assert np.array_equal(dus_a[c][~fa], dus_b[c][~fb])
        if fa.any():
            worst = max(worst, float(np.max(np.abs(dus_a[c][fa] - dus_b[c][fa]))))
        assert np.array_equal(pos_a[c], pos_b[c]), "chain %d: positions differ" % c
        assert rng_a[c][1] == rng_b[c][1] and np.array_equal(rng_a[c][0], rng_b[c][0])
    assert worst <= tol, worst
    return worst


RHO = 500.0 / 60.0 ** 3


def salt(n, **kw):
    return ("salt_dh", dict(n=n, L=(n / RHO) ** (1.0 / 3.0), lattice=n > 1500, **kw))


@pytest.mark.parametrize("case", ["dh_hs_500", "mixed_sizes", "coulomb"])
def test_lds_partners_match_device_memory(gpu, case):
    fgx = gpu
    if case == "dh_hs_500":
        specs = [salt(500)] * 4
    elif case == "mixed_sizes":
        # 2,048 fits (64 bytes each in 128 KiB), 2,050 does not
        specs = [salt(500), salt(2048), salt(2050), salt(1200)]
    else:
        specs = [salt(700, debye=1e9)] * 2
    say("[lds partners] %s" % case)
    a = run_batch(fgx, specs, True, 3, 8000)
    b = run_batch(fgx, specs, False, 3, 8000)
    worst = compare(a, b)
    say("  max |dU(LDS) - dU(device memory)| %.2e kT" % worst)


def test_lds_partners_lennard_jones(gpu):
    """Lennard-Jones pairs (the loop reads partner types) and the
    electrolyte's charges with Debye-Hueckel instead of Ewald."""
    fgx = gpu
    from beofaunus_amd import workloads
    out = {}
    for lds in (True, False):
        w = workloads.ewald_electrolyte(n=1500, L=60.0, seed=9, precision=fgx.MIXED)
        w["ham"] = dict(w["ham"], elec=fgx.ELEC_DH, ewald=None, kappa=1.0 / 12.0)
        s = workloads.make_system(w)
        s.set_lds_partners(lds)
        s.rng_seed(77)
        s.propagator_init()
        st, du, acc = s.propagate(3, record=5000)
        out[lds] = ([du], [acc], [s.positions().copy()], [s.rng_state()])
        s.close()
    worst = compare(out[True], out[False])
    say("[lds partners] LJ + DH, N=1500: max |dU diff| %.2e kT" % worst)
