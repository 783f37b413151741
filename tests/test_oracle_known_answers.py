"""Known answers that pin the oracle independently of the reference run:

  - Cuboid minimum image (geometry.h:507-537): the distance the Nonbonded
    (sqdist) and NonbondedVector (vdist) sums use equals the brute-force
    minimum over the 27 periodic images, in a non-cubic box, read back from
    Coulomb / Debye-Hueckel pair energies (r = lB |q_i q_j| / |u|);
  - Ewald (ewald.h:35-760): the ion pair of the reference's own unit test
    (src/examples/unittests.cpp:131-175, after doi:10.1063/1.481216): two
    unit charges 1 A apart in a 10 A cube, alpha 0.894427191, rc 5, |n| <= 11
    spherical, eps_surf 1 (surface term on), at lB(eps_r = 1, 298 K) =
    560.7378949512486 A.  The reference test at HEAD cannot run (its input
    lacks the required debyelength, ewald.h:521, and its neutral particles'
    effective charges are 0/0); here kappa -> 0 via a huge debyelength and
    the two neutral particles, which add nothing, are left out.  Tolerance:
    the test's own Approx (1.2e-5 relative).
"""
import itertools

import numpy as np

from oracle_py import OracleSystem, lib as olib

LB_298_EPS1 = 560.7378949512486


def brute_min_image(a, b, L):
    best = np.inf
    for n in itertools.product((-1, 0, 1), repeat=3):
        d = a - b + np.array(n) * L
        best = min(best, float(d @ d))
    return best


def pair_system(kind, L, xyz, q, kappa=0.0):
    n = len(q)
    ids = np.where(np.asarray(q) > 0, 1, 2)
    return OracleSystem(kind, L, xyz, q, np.zeros(n), np.ones(n), ids,
                        [{"front": 0, "back": n - 1, "molsize": 1, "molid": 0}], 7.0,
                        kappa=kappa, type_sigma=[0, 0, 0], type_eps=[0, 0, 0],
                        type_charge=[0, 1, -1], type_radius=[0, 0, 0], type_dp=[0, 1, 1])


def test_minimum_image_matches_27_image_brute_force():
    rng = np.random.default_rng(3)
    L = np.array([20.0, 27.0, 33.0])
    n = 48
    xyz = (rng.random((n, 3)) - 0.5) * L
    xyz[:6] = 0.5 * L * np.array([[1, 1, 1], [-1, -1, -0.9], [1, -1, 0.999],
                                  [-0.999, 1, -1], [0.25, -1, 1], [-1, 0.5, 1]])
    q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    for kind, kappa in (("coulhs", 0.0), ("dhhs_vec", 1e-17)):
        o = pair_system(kind, L, xyz, q, kappa)
        worst = 0.0
        for i in range(n):
            for j in range(i + 1, n):
                u = o.i2i(i, j)
                r_engine = 7.0 / abs(u)          # lB |q_i q_j| / |u|, exp(-kappa r) = 1
                r_brute = np.sqrt(brute_min_image(xyz[i], xyz[j], L))
                worst = max(worst, abs(r_engine - r_brute) / r_brute)
        assert worst < 1e-11, (kind, worst)


def test_ewald_ion_pair_known_answer():
    L = [10.0, 10.0, 10.0]
    xyz = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    q = np.array([1.0, -1.0])
    ew = {"alpha": 0.894427190999916, "cutoff": 5.0, "cutoffK": 11.0, "debyelength": 1e12,
          "eps_surf": 1.0, "spherical_sum": True, "update_frequency": 10}
    o = OracleSystem("ewaldlj", L, xyz, q, np.zeros(2), np.ones(2), np.array([1, 2]),
                     [{"front": 0, "back": 1, "molsize": 1, "molid": 0}], LB_298_EPS1,
                     type_sigma=[0, 0, 0], type_eps=[0, 0, 0], type_charge=[0, 1, -1],
                     type_radius=[0, 0, 0], type_dp=[0, 1, 1], ewald=ew)
    lB = LB_298_EPS1
    approx = 1.2e-5
    olib().or_ewald_set_space(o.h)
    ureal = o.i2i(0, 1)
    usurf_reci = o.external()
    uself = o.g_external(0)
    usys = o.system_energy()
    assert abs(ureal / lB - -0.205903210732068) <= approx * 0.205903210732068
    assert abs(usurf_reci / lB - 0.215125076289333) <= approx * 0.215125076289333
    assert abs(uself / lB - -1.009253008808064) <= approx * 1.009253008808064
    assert abs(usys / lB - -1.0000311) <= approx * 1.0000311
