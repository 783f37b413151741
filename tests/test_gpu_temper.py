"""ParallelTempering on one GPU (fgx_temper) against the oracle restatement
of move.h:6520-6772: same partner choices, same energy differences (1e-9
relative), same accept/reject sequence and final replica-to-configuration
assignment."""
import numpy as np
import pytest

from oracle_py import OracleSystem, dptr, iptr, lib as olib

pytestmark = pytest.mark.gpu


def test_temper_matches_oracle():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from beofaunus_amd import fgx, workloads
    print("[temper] R=4 salt replicas", flush=True)
    R, nsteps = 4, 24
    T = 280.0 * 1.5 ** (np.arange(R) / (R - 1))
    lBs = 7.1 * 298.15 / T
    ws = [workloads.salt_dh(n=60, L=30.0, seed=100 + r, lB=lBs[r]) for r in range(R)]
    gpu = [workloads.make_system(w) for w in ws]
    t = fgx.Temper(gpu, ctl_seed=21, glob_seed=22)
    g_acc, g_du = [], []
    for _ in range(nsteps):
        a, d = t.exchange()
        g_acc.append(a.copy())
        g_du.append(d.copy())
    g_slots = t.slots()
    # oracle: U[r][c] = replica r's Hamiltonian (lB_r) on configuration c
    U = np.zeros((R, R))
    for r in range(R):
        for c in range(R):
            w = ws[c]
            h = w["ham"]
            o = OracleSystem("dhhs", w["box"], w["xyz"], w["charge"], w["radius"], w["mw"],
                             w["types"], w["groups"], lBs[r], kappa=h["kappa"],
                             type_sigma=h["type_sigma"], type_eps=h["type_eps"],
                             type_charge=h["type_charge"], type_radius=h["type_radius"],
                             type_dp=h["type_dp"])
            U[r, c] = o.system_energy()
    slot = np.arange(R, dtype=np.int32)
    acc = np.zeros(R * nsteps, dtype=np.int32)
    du = np.zeros(R * nsteps)
    uu = np.ascontiguousarray(U.reshape(-1))
    olib().or_pt_exchange_seeded(R, dptr(uu), iptr(slot), 21, 22, nsteps, iptr(acc), dptr(du))
    acc = acc.reshape(nsteps, R)
    du = du.reshape(nsteps, R)
    assert np.array_equal(np.array(g_acc), acc)
    assert np.allclose(np.array(g_du), du, rtol=1e-9, atol=1e-9)
    assert np.array_equal(g_slots, slot)
    assert acc.sum() > 0
    t.close()
    for s in gpu:
        s.close()


def test_batched_ladders_drive():
    """Several equal-N ladders advanced by one batched launch, then exchanges
    inside each ladder: every chain reports its own device time, the trial
    count is the sum of the sweeps, and each ladder's slots stay a
    permutation of its replicas."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from beofaunus_amd import fgx, workloads
    print("[temper] 2 ladders x 4 replicas, batched", flush=True)
    ladders = workloads.pt_ladders(nladder=2, rungs=4, nmin=200, nmax=400, seed=9)
    systems, tempers = [], []
    for l, lad in enumerate(ladders):
        reps = [workloads.make_system(w) for w in lad["replicas"]]
        for r, s in enumerate(reps):
            s.rng_seed(31 + 10 * l + r)
            s.propagator_init()
        systems += reps
        tempers.append(fgx.Temper(reps, ctl_seed=3 + l, glob_seed=5 + l))
    for step in range(3):
        st = fgx.propagate_batch(systems, 2)
        assert st["trials"] == 2 * sum(l["n"] * len(l["replicas"]) for l in ladders)
        assert all(s.chain_seconds() > 0 for s in systems)
        for t in tempers:
            acc, du = t.exchange()
            assert np.all(np.isfinite(du))
    for t, lad in zip(tempers, ladders):
        assert sorted(t.slots().tolist()) == list(range(len(lad["replicas"])))
        t.close()
    for s in systems:
        s.close()


def test_temper_ewald_matches_oracle():
    """The same exchanges for replicas of an Ewald LJ + Coulomb electrolyte
    whose Hamiltonians differ only in lB (LJ eps given in kT): the oracle's
    U_r(x_c) includes the reciprocal, self and real-space terms at lB_r."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from beofaunus_amd import fgx, workloads
    print("[temper] R=4 Ewald replicas", flush=True)
    R, nsteps = 4, 16
    T = 280.0 * 1.5 ** (np.arange(R) / (R - 1))
    lBs = 7.1 * 298.15 / T
    ws = [workloads.ewald_electrolyte(n=200, L=26.0, seed=200 + r, lB=lBs[r], kc=5.0,
                                      precision=fgx.FP64) for r in range(R)]
    gpu = [workloads.make_system(w) for w in ws]
    t = fgx.Temper(gpu, ctl_seed=31, glob_seed=32)
    g_acc, g_du = [], []
    for _ in range(nsteps):
        a, d = t.exchange()
        g_acc.append(a.copy())
        g_du.append(d.copy())
    g_slots = t.slots()
    U = np.zeros((R, R))
    for r in range(R):
        for c in range(R):
            w = ws[c]
            h = w["ham"]
            o = OracleSystem(w["ref_kind"], w["box"], w["xyz"], w["charge"], w["radius"],
                             w["mw"], w["types"], w["groups"], lBs[r],
                             type_sigma=h["type_sigma"], type_eps=h["type_eps"],
                             type_charge=h["type_charge"], type_radius=h["type_radius"],
                             type_dp=h["type_dp"], ewald=h["ewald"])
            U[r, c] = o.system_energy()
    slot = np.arange(R, dtype=np.int32)
    acc = np.zeros(R * nsteps, dtype=np.int32)
    du = np.zeros(R * nsteps)
    uu = np.ascontiguousarray(U.reshape(-1))
    olib().or_pt_exchange_seeded(R, dptr(uu), iptr(slot), 31, 32, nsteps, iptr(acc), dptr(du))
    acc = acc.reshape(nsteps, R)
    du = du.reshape(nsteps, R)
    assert np.array_equal(np.array(g_acc), acc)
    assert np.allclose(np.array(g_du), du, rtol=1e-9, atol=1e-9)
    assert np.array_equal(g_slots, slot)
    t.close()
    for s in gpu:
        s.close()


def test_temper_refuses_replicas_differing_beyond_lB():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from beofaunus_amd import fgx, workloads
    ws = [workloads.ewald_electrolyte(n=200, L=26.0, seed=210 + r, eps_kt=0.1 * (1 + r),
                                      kc=5.0, precision=fgx.FP64) for r in range(2)]
    gpu = [workloads.make_system(w) for w in ws]
    try:
        with pytest.raises(fgx.FgxError):
            fgx.Temper(gpu)
    finally:
        for s in gpu:
            s.close()
