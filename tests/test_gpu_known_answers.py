"""The device's energies against known answers that do not come from the
reference run: the 27-image brute-force minimum image (Cuboid sqdist and
vdist, geometry.h:507-537) in a non-cubic box, and the Ewald ion pair of
the reference's unit test (src/examples/unittests.cpp:131-175, after
doi:10.1063/1.481216) at kappa -> 0; tolerances as the oracle's tests
state them (1e-11 relative; the unit test's Approx, 1.2e-5)."""
import numpy as np
import pytest

from test_oracle_known_answers import LB_298_EPS1, brute_min_image

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from beofaunus_amd import fgx
    return fgx


def test_minimum_image_27_images(gpu):
    fgx = gpu
    print("[known] minimum image vs 27 images", flush=True)
    rng = np.random.default_rng(3)
    L = np.array([20.0, 27.0, 33.0])
    n = 32
    xyz = (rng.random((n, 3)) - 0.5) * L
    q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    ids = np.where(q > 0, 1, 2)
    grp = [{"front": 0, "back": n - 1, "molsize": 1, "molid": 0}]
    for nb, elec, kappa in ((fgx.NB_PLAIN, fgx.ELEC_COULOMB, 0.0),
                            (fgx.NB_VECTOR, fgx.ELEC_DH, 1e-17)):
        s = fgx.System(L, xyz, q, np.zeros(n), np.ones(n), ids, grp, elec=elec, sr=fgx.SR_HS,
                       nbkind=nb, precision=fgx.FP64, lB=7.0, kappa=kappa,
                       type_sigma=[0, 0, 0], type_eps=[0, 0, 0], type_charge=[0, 1, -1],
                       type_radius=[0, 0, 0], type_dp=[0, 1, 1])
        worst = 0.0
        for i in range(n):
            for j in range(i + 1, n):
                r = 7.0 / abs(s.i2i(i, j))
                rb = np.sqrt(brute_min_image(xyz[i], xyz[j], L))
                worst = max(worst, abs(r - rb) / rb)
        s.close()
        assert worst < 1e-11, (nb, worst)


def test_ewald_ion_pair_known_answer(gpu):
    fgx = gpu
    print("[known] Ewald ion pair (doi:10.1063/1.481216)", flush=True)
    ew = {"alpha": 0.894427190999916, "cutoff": 5.0, "cutoffK": 11.0, "debyelength": 1e12,
          "eps_surf": 1.0, "spherical_sum": True, "update_frequency": 10}
    s = fgx.System([10.0] * 3, np.array([[0.0, 0, 0], [1.0, 0, 0]]), np.array([1.0, -1.0]),
                   np.zeros(2), np.ones(2), np.array([1, 2]),
                   [{"front": 0, "back": 1, "molsize": 1, "molid": 0}], elec=fgx.ELEC_EWALD,
                   sr=fgx.SR_LJ, nbkind=fgx.NB_VECTOR, precision=fgx.FP64, lB=LB_298_EPS1,
                   type_sigma=[0, 0, 0], type_eps=[0, 0, 0], type_charge=[0, 1, -1],
                   type_radius=[0, 0, 0], type_dp=[0, 1, 1], ewald=ew)
    lB = LB_298_EPS1
    usys = s.system_energy()       # setSpace: S(k) from the particles
    vals = {"real": (s.i2i(0, 1), -0.205903210732068),
            "surface+reciprocal": (s.external(), 0.215125076289333),
            "self": (s.g_external(0), -1.009253008808064),
            "system": (usys, -1.0000311)}
    s.close()
    for k, (v, ref) in vals.items():
        assert abs(v / lB - ref) <= 1.2e-5 * abs(ref), (k, v / lB, ref)
