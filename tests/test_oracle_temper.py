"""ParallelTempering restatement (move.h:6520-6772) pinned without a
reference run (the reference's ParallelTempering needs MPI, absent here).

What fixes it:
  - hand values of _energyChange (move.h:6701-6723): rank r's own term
    U_r(x_partner) - U_r(x_own) plus the partner's, on toy energy matrices
    whose entries are all different, so a wrong sign, index or partner fails;
  - the stationary occupancy of a two-state toy, e^{-dU} / (1 + e^{-dU}),
    which a negated dU turns into 1 / (1 + e^{-dU});
  - partner choice: rank +- 1 by a shared draw and rank parity
    (findPartner, move.h:6621-6633), so pairs are disjoint;
  - identical replicas give dU = 0 and are always swapped;
  - the system form (or_pt_step: ranks holding configurations, each with its
    own engines) takes the same decisions as the matrix form when the ranks
    are in lock-step, moves configurations, and evaluates the Ewald
    partner energy with the rank's own S(k), as Energy::systemEnergy's
    setSpace leaves it (ewald.h:716-757).
"""
import numpy as np

from oracle_py import OracleEngine, OracleSystem, dptr, iptr, lib, pt_step


def exchange(u, slot, ctl, glob, nsteps):
    R = len(slot)
    s = np.array(slot, dtype=np.int32)
    acc = np.zeros(R * nsteps, dtype=np.int32)
    du = np.zeros(R * nsteps)
    uu = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(-1))
    lib().or_pt_exchange_seeded(R, dptr(uu), iptr(s), ctl, glob, nsteps, iptr(acc), dptr(du))
    return s, acc.reshape(nsteps, R), du.reshape(nsteps, R)


def step(u, slot, ctl, glob):
    R = len(slot)
    acc = np.zeros(R, dtype=np.int32)
    du = np.zeros(R)
    uu = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(-1))
    lib().or_pt_exchange(R, dptr(uu), iptr(slot), ctl.h, glob.h, iptr(acc), dptr(du))
    return acc, du


def test_identical_replicas_always_swap():
    R = 6
    u = np.full((R, R), 3.25)
    s, acc, du = exchange(u, range(R), 1, 2, 50)
    assert np.all(du == 0.0)
    # every replica with a partner in range swaps
    assert acc.sum() > 0
    for row in acc:
        assert row.sum() % 2 == 0


def test_pairs_are_disjoint_and_symmetric():
    rng = np.random.default_rng(0)
    R = 8
    u = rng.normal(size=(R, R)) * 3
    s, acc, du = exchange(u, range(R), 3, 4, 200)
    assert sorted(s.tolist()) == list(range(R))
    for t in range(acc.shape[0]):
        # partners share the same dU and the same decision
        pairs = [(r, r + 1) for r in range(R - 1)]
        for a, b in pairs:
            if du[t, a] != 0 and du[t, a] == du[t, b]:
                assert acc[t, a] == acc[t, b]


def test_first_attempt_energy_hand_value():
    # U_r(x_c): swapping x0 <-> x1 costs (U_0(x_1) - U_0(x_0)) + (U_1(x_0) - U_1(x_1))
    # = (1.0 - 0.0) + (0.5 - 0.0) = +1.5
    u = np.array([[0.0, 1.0], [0.5, 0.0]])
    ctl, glob = OracleEngine(11), OracleEngine(12)
    slot = np.array([0, 1], dtype=np.int32)
    for _ in range(20):   # until the shared direction pairs rank 0 with rank 1
        acc, du = step(u, slot, ctl, glob)
        if du[0] != 0.0:
            break
    assert du[0] == 1.5 and du[1] == 1.5


def test_four_ranks_partner_terms_hand_values():
    # all entries distinct and not separable in (r, c), as lB_r E(x_c) is
    R = 4
    u = np.array([[10.0 * r + (r + 1) * (c + c * c / 8.0) for c in range(R)] for r in range(R)])
    ctl, glob = OracleEngine(3), OracleEngine(4)
    seen = set()
    slot = np.arange(R, dtype=np.int32)
    for _ in range(40):
        before = slot.copy()
        acc, du = step(u, slot, ctl, glob)
        pairs = [(r, r + 1) for r in range(R - 1) if du[r] != 0.0 and du[r] == du[r + 1]]
        for r, p in pairs:
            a, b = before[r], before[p]
            want = (u[r, b] - u[r, a]) + (u[p, a] - u[p, b])
            assert du[r] == want, (r, p, du[r], want)
            seen.add((r, p))
    # both pairings ((0,1),(2,3)) and (1,2) occurred
    assert {(0, 1), (2, 3), (1, 2)} <= seen


def test_stationary_occupancy_two_state_toy():
    # state A = identity (E = 0), state B = swapped (E = 1.5): the fraction of
    # steps spent in B is e^{-1.5} / (1 + e^{-1.5}) = 0.1824; with dU negated
    # it would be 0.8176
    u = np.array([[0.0, 1.0], [0.5, 0.0]])
    n = 20000
    s, acc, du = exchange(u, [0, 1], 21, 22, 1)   # warm engines are fine: fresh run below
    ctl, glob = OracleEngine(21), OracleEngine(22)
    slot = np.array([0, 1], dtype=np.int32)
    inb = 0
    for _ in range(n):
        step(u, slot, ctl, glob)
        inb += int(slot[0] == 1)
    frac = inb / n
    pi_b = np.exp(-1.5) / (1 + np.exp(-1.5))
    # A -> B with prob 0.5 e^{-1.5}, B -> A with 0.5: autocorrelation factor
    lam = 1 - 0.5 * np.exp(-1.5) - 0.5
    sigma = np.sqrt(pi_b * (1 - pi_b) * (1 + lam) / (1 - lam) / n)
    assert abs(frac - pi_b) < 4 * sigma, (frac, pi_b, sigma)


def test_metropolis_rate_two_replicas():
    # two replicas, U_r(x_c): swapping costs dU = (1.0 - 0.0) + (0.5 - 0.0) = 1.5
    u = np.array([[0.0, 1.0], [0.5, 0.0]])
    n, hits, tries = 4000, 0, 0
    _, acc, du = exchange(u, [0, 1], 11, 12, n)
    # with the slot updated in place, dU alternates between +1.5 and -1.5;
    # count acceptances of the +1.5 moves
    for t in range(n):
        if du[t, 0] > 0:
            tries += 1
            hits += acc[t, 0]
    assert tries > 500
    assert abs(hits / tries - np.exp(-1.5)) < 0.05


def _salt_rank(xyz, lB, seed):
    n = len(xyz)
    q = np.array([1.0 if i % 2 == 0 else -1.0 for i in range(n)])
    o = OracleSystem("dhhs", [14.0] * 3, xyz, q, np.full(n, 1.0), np.ones(n),
                     np.array([1 if i % 2 == 0 else 2 for i in range(n)]),
                     [{"front": 0, "back": n - 1, "molsize": 1, "molid": 0}], lB, kappa=0.1,
                     type_sigma=[0, 2, 2], type_eps=[0, 0, 0], type_charge=[0, 1, -1],
                     type_radius=[0, 1, 1], type_dp=[0, 1, 1])
    o.seed(seed)
    lib().or_propagator_init(o.h)
    return o


def _configs(R, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(R):
        pts = []
        while len(pts) < n:
            p = rng.uniform(-7.0, 7.0, 3)
            if all(np.sum(((p - q + 7.0) % 14.0 - 7.0) ** 2) > 4.5 for q in pts):
                pts.append(p)
        out.append(np.array(pts))
    return out


def test_system_step_matches_matrix_form_in_lock_step():
    """Ranks seeded alike (the reference's default: every rank's engines
    start from mt19937's default seed) draw the same Metropolis uniform, so
    the per-rank decisions agree; they equal the matrix form's with
    U[r][c] from separate systems, and accepted ranks end up holding the
    partner's configuration."""
    R, n = 4, 12
    lBs = [7.1 * 298.15 / t for t in (280.0, 310.0, 345.0, 385.0)]
    xs = _configs(R, n, 5)
    U = np.array([[_salt_rank(xs[c], lBs[r], 1).system_energy() for c in range(R)]
                  for r in range(R)])
    ranks = [_salt_rank(xs[r], lBs[r], 77) for r in range(R)]
    ctl = OracleEngine(9)
    # matrix form: the same controller seed; glob = any rank's global engine
    slot = np.arange(R, dtype=np.int32)
    mctl, mglob = OracleEngine(9), OracleEngine(77)
    held = list(range(R))
    for t in range(12):
        acc, du, ret = pt_step(ranks, ctl)
        macc, mdu = step(U, slot, mctl, mglob)
        for r in range(R):
            if np.isfinite(du[r]):
                assert du[r] == mdu[r]
        assert np.array_equal(acc, macc)
        # configurations follow the matrix form's assignment
        for r in range(R):
            assert np.array_equal(ranks[r].xyz(0), xs[slot[r]])
        # Movebase::move returns unew - uold on accepted ranks
        for r in range(R):
            if acc[r]:
                own = U[r, slot[r]] - U[r, held[r]]
                assert ret[r] == own
        held = slot.tolist()
    assert (np.array(slot) != np.arange(R)).any()


def _ewald_rank(xyz, lB, seed):
    n = len(xyz)
    q = np.array([1.0 if i % 2 == 0 else -1.0 for i in range(n)])
    ew = {"alpha": 0.35, "cutoff": 6.5, "cutoffK": 3.0, "debyelength": 1e12,
          "eps_surf": 1.0, "spherical_sum": True, "update_frequency": 2}
    o = OracleSystem("ewaldlj", [14.0] * 3, xyz, q, np.full(n, 1.0), np.ones(n),
                     np.array([1 if i % 2 == 0 else 2 for i in range(n)]),
                     [{"front": 0, "back": n - 1, "molsize": 1, "molid": 0}], lB,
                     type_sigma=[0, 2, 2], type_eps=[0, 0.1, 0.1], type_charge=[0, 1, -1],
                     type_radius=[0, 1, 1], type_dp=[0, 1, 1], ewald=ew)
    o.seed(seed)
    lib().or_propagator_init(o.h)
    return o


def test_ewald_partner_energy_uses_own_structure_factor():
    """With NonbondedEwald, systemEnergy(trial) after setSpace takes the
    reciprocal energy from Q_trial = Q, the rank's own S(k) (ewald.h:716-729,
    731-757), so at equal volume the reciprocal terms cancel in each rank's
    unew - uold."""
    R, n = 2, 10
    lBs = [7.0, 6.0]
    xs = _configs(R, n, 13)
    full = [[_ewald_rank(xs[c], lBs[r], 1) for c in range(R)] for r in range(R)]
    U = np.array([[full[r][c].system_energy() for c in range(R)] for r in range(R)])
    Rec = np.array([[lib().or_ewald_reciprocal(full[r][c].h, 0) for c in range(R)]
                    for r in range(R)])
    want = ((U[0, 1] - Rec[0, 1]) - (U[0, 0] - Rec[0, 0])) + \
           ((U[1, 0] - Rec[1, 0]) - (U[1, 1] - Rec[1, 1]))
    full_du = (U[0, 1] - U[0, 0]) + (U[1, 0] - U[1, 1])
    ranks = [_ewald_rank(xs[r], lBs[r], 5) for r in range(R)]
    ctl = OracleEngine(1)
    for _ in range(10):
        acc, du, ret = pt_step(ranks, ctl)
        if np.isfinite(du[0]):
            break
    assert abs(du[0] - want) < 1e-10 * max(1.0, abs(want))
    assert abs(full_du - want) > 1e-6    # the reciprocal terms do matter here
