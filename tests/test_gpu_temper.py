"""ParallelTempering on one GPU (fgx_temper_*) against the oracle's
restatement of move.h:6520-6772.

  - exchanges alone against the energy-matrix form (U[r][c] from separate
    oracle systems at rank r's lB): same partners, dU within 1e-9, same
    decisions and assignment;
  - sweeps interleaved with exchanges against or_pt_step, whose ranks hold
    configurations and keep their own engines: after every exchange each
    rank's next sweep on the device (the configuration the rank now holds,
    with the rank's lB and engines) gives the oracle rank's trials: identical
    accepts and dU within 1e-10 (FP64) or 1e-5 kT (mixed);
  - the same rounds as one replayed HIP graph (fgx_temper_run);
  - Ewald ranks, whose partner energies use the rank's own S(k) and whose
    structure factors follow the reference's setSpace / update sequence;
  - refusals: replicas differing beyond lB, or with ExternalPressure.
"""
import numpy as np
import pytest

from oracle_py import OracleEngine, OracleSystem, dptr, iptr, lib as olib, pt_step

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from beofaunus_amd import fgx
    return fgx


def lbs(R):
    T = 280.0 * 1.5 ** (np.arange(R) / (R - 1))
    return 7.1 * 298.15 / T


def oracle_rank(w, lB, seed):
    h = w["ham"]
    kw = dict(type_sigma=h["type_sigma"], type_eps=h["type_eps"], type_charge=h["type_charge"],
              type_radius=h["type_radius"], type_dp=h["type_dp"])
    if h.get("ewald"):
        kw["ewald"] = h["ewald"]
    else:
        kw["kappa"] = h["kappa"]
    o = OracleSystem(w["ref_kind"], w["box"], w["xyz"], w["charge"], w["radius"], w["mw"],
                     w["types"], w["groups"], lB, **kw)
    olib().or_add_move(o.h, 1, 0, 1.0, 1, 0, None, None, 0.0, 0.0, 1)
    o.seed(seed)
    olib().or_propagator_init(o.h)
    return o


def device_ranks(fgx, ws, seeds, precision):
    from beofaunus_amd import workloads
    out = []
    for w, sd in zip(ws, seeds):
        s = workloads.make_system(w, precision=precision)
        s.rng_seed(sd)
        s.propagator_init()
        out.append(s)
    return out


def salt_ranks(R, n=60, L=30.0, seed0=100):
    from beofaunus_amd import workloads
    return [workloads.salt_dh(n=n, L=L, seed=seed0 + r, lB=lbs(R)[r]) for r in range(R)]


def test_exchanges_match_matrix_form(gpu):
    fgx = gpu
    print("[temper] R=4 salt replicas, exchanges only", flush=True)
    R, nsteps = 4, 24
    ws = salt_ranks(R)
    # ranks in lock-step: every engine starts from the same seed
    dev = device_ranks(fgx, ws, [22] * R, fgx.FP64)
    t = fgx.Temper(dev, ctl_seed=21)
    g_acc, g_du = [], []
    for _ in range(nsteps):
        a, d = t.exchange()
        g_acc.append(a.copy())
        g_du.append(d.copy())
    U = np.array([[oracle_rank(ws[c], lbs(R)[r], 1).system_energy() for c in range(R)]
                  for r in range(R)])
    slot = np.arange(R, dtype=np.int32)
    acc = np.zeros(R * nsteps, dtype=np.int32)
    du = np.zeros(R * nsteps)
    olib().or_pt_exchange_seeded(R, dptr(np.ascontiguousarray(U.reshape(-1))), iptr(slot), 21,
                                 22, nsteps, iptr(acc), dptr(du))
    acc = acc.reshape(nsteps, R)
    du = du.reshape(nsteps, R)
    g_du = np.array(g_du)
    fin = np.isfinite(g_du)
    assert np.array_equal(np.array(g_acc), acc)
    # a rank without partner: +inf on the device (_energyChange), 0 in the matrix form
    assert np.all(g_du[~fin] == np.inf) and np.all(du[~fin] == 0.0)
    assert np.allclose(g_du[fin], du[fin], rtol=1e-9, atol=1e-9)
    assert np.array_equal(t.slots(), slot)
    assert acc.sum() > 0
    t.close()
    for s in dev:
        s.close()


def interleaved(fgx, ws, precision, rounds, calls, ctl_seed, seeds, tol, graph=False):
    R = len(ws)
    lb = [w["ham"]["lB"] for w in ws]
    n = len(ws[0]["charge"])
    dev = device_ranks(fgx, ws, seeds, precision)
    t = fgx.Temper(dev, ctl_seed=ctl_seed)
    ranks = [oracle_rank(ws[r], lb[r], seeds[r]) for r in range(R)]
    ctl = OracleEngine(ctl_seed)
    slot = np.arange(R)
    n_swaps = 0
    if graph:
        st, sec = fgx.temper_run([t], rounds, calls)
        assert st["trials"] == rounds * calls * n * R
        g_acc, g_du, g_slot, g_sec = t.records()
        assert g_acc.shape == (rounds, R) and np.all(g_sec > 0)
    for rd in range(rounds):
        print("[temper] round %d" % rd, flush=True)
        o_recs = [ranks[r].propagate(calls) for r in range(R)]
        if not graph:
            _, dus, accs = fgx.propagate_batch_record(dev, calls, calls * n)
            for r in range(R):
                c = slot[r]
                odu, oacc = o_recs[r][0], o_recs[r][1]
                assert np.array_equal(accs[c], oacc), (rd, r)
                fin = np.isfinite(odu)
                assert np.array_equal(np.isfinite(dus[c]), fin)
                err = np.max(np.abs(dus[c][fin] - odu[fin]) / np.maximum(1.0, np.abs(odu[fin])))
                assert err < tol, (rd, r, err)
            acc, du = t.exchange()
        else:
            acc, du = g_acc[rd], g_du[rd]
        oacc, odu, oret = pt_step(ranks, ctl)
        assert np.array_equal(acc, oacc), (rd, acc, oacc)
        fin = np.isfinite(odu)
        assert np.array_equal(np.isfinite(du), fin)
        assert np.allclose(du[fin], odu[fin], rtol=1e-9, atol=max(tol, 1e-9))
        for r in range(R):
            p = r + 1
            if p < R and oacc[r] and oacc[p] and np.isfinite(odu[r]) and odu[r] == odu[p]:
                slot[r], slot[p] = slot[p], slot[r]
                n_swaps += 1
        if graph:
            assert np.array_equal(g_slot[rd], slot)
        else:
            assert np.array_equal(t.slots(), slot)
    # each rank's configuration: the handle it holds now
    for r in range(R):
        x = dev[slot[r]].positions()
        assert np.max(np.abs(x - ranks[r].xyz(0))) < (1e-9 if tol < 1e-8 else 1e-4)
    t.close()
    for s in dev:
        s.close()
    return n_swaps


def test_interleaved_fp64_dh(gpu):
    fgx = gpu
    print("[temper] R=4 DH ranks, sweeps + exchanges, FP64", flush=True)
    ws = salt_ranks(4)
    nsw = interleaved(fgx, ws, fgx.FP64, rounds=6, calls=2, ctl_seed=5,
                      seeds=[31, 32, 33, 34], tol=1e-10)
    assert nsw > 0


def test_interleaved_mixed_dh(gpu):
    fgx = gpu
    print("[temper] R=4 DH ranks, sweeps + exchanges, mixed", flush=True)
    ws = salt_ranks(4)
    nsw = interleaved(fgx, ws, fgx.MIXED, rounds=6, calls=2, ctl_seed=5,
                      seeds=[31, 32, 33, 34], tol=1e-5)
    assert nsw > 0


def test_interleaved_graph_fp64_dh(gpu):
    fgx = gpu
    print("[temper] R=4 DH ranks, replayed graph, FP64", flush=True)
    ws = salt_ranks(4)
    nsw = interleaved(fgx, ws, fgx.FP64, rounds=6, calls=2, ctl_seed=5,
                      seeds=[31, 32, 33, 34], tol=1e-10, graph=True)
    assert nsw > 0


def test_interleaved_graph_mixed_dh(gpu):
    fgx = gpu
    print("[temper] R=6 DH ranks, replayed graph, mixed", flush=True)
    ws = salt_ranks(6, n=80, L=34.0, seed0=300)
    interleaved(fgx, ws, fgx.MIXED, rounds=5, calls=2, ctl_seed=8,
                seeds=[41, 42, 43, 44, 45, 46], tol=1e-5, graph=True)


def test_interleaved_fp64_ewald(gpu):
    fgx = gpu
    from beofaunus_amd import workloads
    print("[temper] R=4 Ewald ranks, sweeps + exchanges, FP64", flush=True)
    R = 4
    ws = [workloads.ewald_electrolyte(n=120, L=22.0, seed=200 + r, lB=lbs(R)[r], kc=4.0,
                                      rc=7.0, alpha=0.35, update_frequency=3,
                                      precision=fgx.FP64) for r in range(R)]
    # one composition for every rank (the exchange moves positions between
    # ranks whose particles agree index by index)
    for w in ws[1:]:
        w["types"], w["charge"] = ws[0]["types"].copy(), ws[0]["charge"].copy()
    nsw = interleaved(fgx, ws, fgx.FP64, rounds=5, calls=1, ctl_seed=3,
                      seeds=[51, 52, 53, 54], tol=1e-10)
    assert nsw > 0


def test_batched_ladders_graph(gpu):
    """Two ladders of different N in one replayed graph: trial counts, a
    sweep time per configuration and round, slots that stay permutations."""
    fgx = gpu
    from beofaunus_amd import workloads
    print("[temper] 2 ladders x 4 replicas, graph", flush=True)
    ladders = workloads.pt_ladders(nladder=2, rungs=4, nmin=200, nmax=400, seed=9)
    tempers, systems = [], []
    for l, lad in enumerate(ladders):
        reps = [workloads.make_system(w) for w in lad["replicas"]]
        for r, s in enumerate(reps):
            s.rng_seed(31 + 10 * l + r)
            s.propagator_init()
        systems += reps
        tempers.append(fgx.Temper(reps, ctl_seed=3 + l))
    st, sec = fgx.temper_run(tempers, 4, 2)
    assert st["trials"] == 4 * 2 * sum(l["n"] * len(l["replicas"]) for l in ladders)
    assert sec > 0
    for t, lad in zip(tempers, ladders):
        acc, du, slot, ssec = t.records()
        assert acc.shape == (4, len(lad["replicas"]))
        assert np.all(ssec > 0)
        assert sorted(t.slots().tolist()) == list(range(len(lad["replicas"])))
        t.close()
    for s in systems:
        s.close()


def test_temper_refuses_replicas_differing_beyond_lB(gpu):
    fgx = gpu
    from beofaunus_amd import workloads
    ws = [workloads.ewald_electrolyte(n=200, L=26.0, seed=210 + r, eps_kt=0.1 * (1 + r),
                                      kc=5.0, precision=fgx.FP64) for r in range(2)]
    dev = [workloads.make_system(w) for w in ws]
    try:
        with pytest.raises(fgx.FgxError):
            fgx.Temper(dev)
    finally:
        for s in dev:
            s.close()


def test_temper_refuses_external_pressure(gpu):
    fgx = gpu
    ws = salt_ranks(2)
    dev = []
    for r, w in enumerate(ws):
        h = dict(w["ham"])
        h["pressure"] = 1e-5 * (1 + r)
        dev.append(fgx.System(w["box"], w["xyz"], w["charge"], w["radius"], w["mw"], w["types"],
                              w["groups"], **h))
    try:
        with pytest.raises(fgx.FgxError):
            fgx.Temper(dev)
    finally:
        for s in dev:
            s.close()
