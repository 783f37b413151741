"""Ewald + cut-and-shifted LJ (every pair term vanishes beyond rc) and the
cell-list form of the mixed AtomicTranslation sweep.

* The FP64 device sweep of this Hamiltonian against the oracle (the
  reference's NonbondedEwald cannot take CutShift<LennardJonesLB>, so the
  oracle composes the two parts pinned separately: EwaldReal by the
  ewald_* fixtures, CutShift<LennardJonesLB> by the lj_cutshift fixture):
  identical accept/reject sequence, dU to 1e-10.
* The mixed sweep with a cell list against the FP64 sweep on the same RNG
  stream: |dU err| <= 1e-5 kT while the chains agree.
* Cell list against the all-pairs screen of the same mixed sweep: the same
  Hamiltonian, so the same accept sequence and dU to rounding.
* A cell capacity that fills up mid-launch: the chain finishes on the
  all-pairs screen and the results do not change.
* At N = 20,000 (the full electrolyte): the accepted dU summed over a sweep
  equals the change of the FP64 system energy.
* |n| <= 12 (more k-vectors than the LDS-resident S(k) holds, so S(k),
  dS(k), A_k and n stay in device memory): FP64 against the oracle, cell
  list against the all-pairs screen and against FP64.
"""
import sys

import numpy as np
import pytest

from oracle_py import OracleSystem, lib as olib

pytestmark = pytest.mark.gpu

REL = 1e-10
TOL_MIXED = 1e-5   # kT


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


def electrolyte(n, seed):
    from beofaunus_amd import workloads
    L = 120.0 * (n / 20000.0) ** (1.0 / 3.0)
    return workloads.ewald_electrolyte(n=n, L=L, seed=seed, lj_cut=True)


def oracle_of(w):
    h = w["ham"]
    o = OracleSystem(w["ref_kind"], w["box"], w["xyz"], w["charge"], w["radius"], w["mw"],
                     w["types"], w["groups"], h["lB"], type_sigma=h["type_sigma"],
                     type_eps=h["type_eps"], type_charge=h["type_charge"],
                     type_radius=h["type_radius"], type_dp=h["type_dp"], ewald=h["ewald"],
                     lj_cut=h["lj_cut"])
    olib().or_add_move(o.h, 1, 0, 1.0, 1, 0, None, None, 0.0, 0.0, 1)
    return o


def run(fgx, w, precision, seed, ncalls, cell_mode=1, record=None):
    from beofaunus_amd import workloads
    s = workloads.make_system(w, precision=precision)
    s.set_cell_list(cell_mode)
    s.rng_seed(seed)
    s.propagator_init()
    n = len(w["charge"])
    st, du, acc = s.propagate(ncalls, record=record or ncalls * n)
    cells = s.cell_list()
    s.close()
    return st, du, acc, cells


def test_fp64_cut_hamiltonian_matches_oracle(gpu):
    w = electrolyte(1000, seed=41)
    say("[cells] FP64 sweep vs oracle, N=1000, L=%.2f" % w["box"][0])
    o = oracle_of(w)
    o.seed(13)
    olib().or_propagator_init(o.h)
    odu, oacc, _, _, _ = o.propagate(2)
    st, du, acc, _ = run(gpu, w, gpu.FP64, 13, 2)
    assert len(du) == len(odu) == 2000
    assert np.array_equal(acc, oacc)
    fin = np.isfinite(odu)
    assert np.array_equal(np.isfinite(du), fin)
    err = np.abs(du[fin] - odu[fin]) / np.maximum(1.0, np.abs(odu[fin]))
    say("  max rel dU err %.2e, accepted %d" % (err.max(), oacc.sum()))
    assert err.max() <= REL


def test_cell_list_matches_fp64_and_all_pairs(gpu):
    w = electrolyte(2000, seed=42)
    say("[cells] mixed cells vs FP64 vs all-pairs screen, N=2000, L=%.2f" % w["box"][0])
    _, d64, a64, _ = run(gpu, w, gpu.FP64, 17, 2)
    _, dc, ac, cells = run(gpu, w, gpu.MIXED, 17, 2, cell_mode=1)
    _, dp, ap, nocells = run(gpu, w, gpu.MIXED, 17, 2, cell_mode=0)
    say("  cell list %s, all-pairs %s" % (cells, nocells))
    assert cells[0] >= 3 and nocells == (0, 0)
    # cell list and all-pairs screen: same Hamiltonian, same decisions
    assert np.array_equal(ac, ap)
    assert np.max(np.abs(dc - dp)) <= 1e-9
    # against FP64 while the chains agree
    mism = np.nonzero(ac != a64)[0]
    upto = int(mism[0]) if len(mism) else len(a64)
    err = np.abs(dc[:upto] - d64[:upto])
    say("  compared %d trials, max |err| %.2e kT, first mismatch %s"
        % (upto, err.max(), mism[0] if len(mism) else None))
    assert upto >= 1000
    assert err.max() <= TOL_MIXED


def test_full_cell_falls_back_to_all_pairs(gpu):
    w = electrolyte(2000, seed=43)
    L = w["box"][0]
    nc = int(np.floor(L / ((12.0 + 0.5) * (1 + 1e-6))))
    u = np.floor((np.asarray(w["xyz"]) / L + 0.5) * nc).astype(int) % nc
    occ = np.bincount(u[:, 0] + nc * (u[:, 1] + nc * u[:, 2]), minlength=nc ** 3)
    cap = int(occ.max())
    say("[cells] capacity %d (largest initial cell), N=2000, %d cells per axis" % (cap, nc))
    _, dp, ap, _ = run(gpu, w, gpu.MIXED, 19, 3, cell_mode=0)
    _, dc, ac, cells = run(gpu, w, gpu.MIXED, 19, 3, cell_mode=cap)
    say("  cell list %s" % (cells,))
    assert np.array_equal(ac, ap)
    assert np.max(np.abs(dc - dp)) <= 1e-9


def test_full_size_energy_drift(gpu):
    from beofaunus_amd import workloads
    w = electrolyte(20000, seed=44)
    say("[cells] N=20000 sweep, energy bookkeeping")
    s = workloads.make_system(w, precision=gpu.MIXED)
    s.rng_seed(23)
    s.propagator_init()
    u0 = s.system_energy()
    st = s.propagate(1)[0]
    u1 = s.system_energy()
    cells = s.cell_list()
    s.close()
    drift = (u1 - u0) - st["utot"]
    say("  cells %s, trials %d, accepted %d, U0 %.6f, U1 %.6f, sum dU %.6f, drift %.2e"
        % (cells, st["trials"], st["accepted"], u0, u1, st["utot"], drift))
    assert cells[0] == 9
    assert st["trials"] == 20000
    assert abs(drift) <= 1e-6 * max(1.0, abs(u0))


@pytest.mark.parametrize("case", ["lj_table", "three_cells"])
def test_cell_list_variants(gpu, case):
    """Per-type LJ tables (the scan reads partner types) and the smallest
    grid of 3 x 3 x 3 cells: cell list and all-pairs screen agree."""
    if case == "lj_table":
        w = electrolyte(2000, seed=45)
        w["ham"] = dict(w["ham"], type_sigma=[0.0, 4.0, 3.2], type_eps=[0.0, 0.1, 0.15])
    else:
        from beofaunus_amd import workloads
        # side 40 A: 3 cells of 13.3 A per axis
        w = workloads.ewald_electrolyte(n=740, L=40.0, seed=46, lj_cut=True)
    say("[cells] %s, N=%d, L=%.1f" % (case, len(w["charge"]), w["box"][0]))
    _, dp, ap, _ = run(gpu, w, gpu.MIXED, 21, 2, cell_mode=0)
    _, dc, ac, cells = run(gpu, w, gpu.MIXED, 21, 2, cell_mode=1)
    _, d64, a64, _ = run(gpu, w, gpu.FP64, 21, 2)
    say("  cell list %s" % (cells,))
    assert cells[0] >= 3
    if case == "three_cells":
        assert cells[0] == 3
    assert np.array_equal(ac, ap)
    assert np.max(np.abs(dc - dp)) <= 1e-9
    mism = np.nonzero(ac != a64)[0]
    upto = int(mism[0]) if len(mism) else len(a64)
    assert upto >= 500
    assert np.max(np.abs(dc[:upto] - d64[:upto])) <= TOL_MIXED


def test_device_memory_sk_with_cells(gpu):
    from beofaunus_amd import workloads
    n = 2000
    L = 120.0 * (n / 20000.0) ** (1.0 / 3.0)
    w = workloads.ewald_electrolyte(n=n, L=L, seed=47, lj_cut=True, kc=12.0)
    o = oracle_of(w)
    K = olib().or_ewald_nk(o.h)
    say("[cells] |n| <= 12, K = %d, N=%d, L=%.2f" % (K, n, L))
    # 44 bytes per k-vector beyond the 110 KiB of dynamic LDS
    assert 44 * K > 110 * 1024
    o.seed(23)
    olib().or_propagator_init(o.h)
    odu, oacc, _, _, _ = o.propagate(2)
    st, d64, a64, _ = run(gpu, w, gpu.FP64, 23, 2)
    assert np.array_equal(a64, oacc)
    fin = np.isfinite(odu)
    assert np.array_equal(np.isfinite(d64), fin)
    err = np.abs(d64[fin] - odu[fin]) / np.maximum(1.0, np.abs(odu[fin]))
    say("  FP64 vs oracle: max rel dU err %.2e, accepted %d" % (err.max(), oacc.sum()))
    assert err.max() <= REL
    _, dp, ap, nocells = run(gpu, w, gpu.MIXED, 23, 2, cell_mode=0)
    _, dc, ac, cells = run(gpu, w, gpu.MIXED, 23, 2, cell_mode=1)
    say("  cell list %s, all-pairs %s" % (cells, nocells))
    assert cells[0] >= 3 and nocells == (0, 0)
    assert np.array_equal(ac, ap)
    assert np.max(np.abs(dc - dp)) <= 1e-9
    mism = np.nonzero(ac != a64)[0]
    upto = int(mism[0]) if len(mism) else len(a64)
    e = float(np.max(np.abs(dc[:upto] - d64[:upto])))
    say("  mixed vs FP64: compared %d trials, max |err| %.2e kT" % (upto, e))
    assert upto >= 1000
    assert e <= TOL_MIXED


def test_cell_list_across_volume_moves(gpu):
    """A Propagator holding AtomicTranslation and Isobaric on the cut-LJ
    electrolyte, then translation sweeps in the box the volume moves left:
    after accepted volume moves the cell grid, the records and the
    fixed-point copies follow the new box.  Cell list and all-pairs screen
    give the same accepts and dU to rounding in both phases; the box did
    change and volume trials did run."""
    from beofaunus_amd import workloads, faunus_io
    w = electrolyte(2000, seed=47)
    L0 = w["box"][0]
    w["ham"] = dict(w["ham"], pressure=faunus_io.mM_to_per_A3(20000.0))
    say("[cells] N=2000 with volume moves, L=%.2f" % L0)
    ncalls, n = 12, len(w["charge"])
    out = {}
    for mode in (1, 0):
        s = workloads.make_system(w, precision=gpu.MIXED)
        try:
            s.set_cell_list(mode)
            s.rng_seed(27)
            s.propagator_init()
            st, du, acc = s.propagate_npt(0.05, 1.0, 1, ncalls, record=100000)
            L1 = s.box()[0]
            st2, du2, acc2 = s.propagate(2, record=2 * n)
            out[mode] = (du, acc, L1, du2, acc2, s.cell_list())
        finally:
            s.close()
    (dc, ac, Lc, dc2, ac2, cells), (dp, ap, Lp, dp2, ap2, _) = out[1], out[0]
    # each call is one volume trial or one sweep of n translation trials
    nvol = (ncalls * n - len(dc)) // (n - 1)
    say("  cell list %s, trials %d (%d volume), L %.4f -> %.4f" % (cells, len(dc), nvol, L0, Lc))
    assert 1 <= nvol < ncalls and len(dc) == (ncalls - nvol) * n + nvol
    assert Lc != L0              # a volume move was accepted
    assert cells[0] >= 3
    assert np.array_equal(ac, ap) and Lc == Lp
    fin = np.isfinite(dp)
    assert np.array_equal(np.isfinite(dc), fin)
    assert np.max(np.abs(dc[fin] - dp[fin])) <= 1e-9
    # translation sweeps in the changed box: the rebuilt list against the screen
    assert len(dc2) == 2 * n and np.array_equal(ac2, ap2)
    assert np.max(np.abs(dc2 - dp2)) <= 1e-9
