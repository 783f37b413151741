"""Move::Isobaric with Energy::ExternalPressure on the engine
(fgx_isobaric_propagate) against the compiled reference's run of the same
system (fixture salt_npt: 60 ions, Debye-Hueckel + hard spheres, 300 mM,
60 volume trials): the same accept/reject sequence, energy changes to 1e-10,
the same final box, coordinates and global engine state."""
import sys

import numpy as np
import pytest

from oracle_py import farr, load_fixture

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


def test_isobaric_matches_reference(gpu):
    from fixture_systems import fgx_system_from_fixture
    fx = load_fixture("salt_npt")
    mc = fx["mc"]
    mv = fx["moves"][0]
    say("[isobaric] salt_npt: %d volume trials" % mc["ncalls"])
    s = fgx_system_from_fixture(fx)
    try:
        s.rng_seed(fx["rng_seed"])
        s.propagator_init()
        st, du, acc = s.isobaric_propagate(mv["dp"], mv["prob"], mc["ncalls"], record=1000)
        box = s.box()
        xyz = s.positions(0).reshape(-1)
        words, pos = s.rng_state()
        u1 = s.system_energy()
    finally:
        s.close()
    ref_du = farr(mc["du"])
    assert len(du) == len(ref_du)
    assert np.array_equal(acc, np.array(mc["acc"], dtype=np.int32))
    fin = np.isfinite(ref_du)
    assert np.array_equal(np.isfinite(du), fin)
    err = np.abs(du[fin] - ref_du[fin]) / np.maximum(1.0, np.abs(ref_du[fin]))
    say("  accepted %d of %d, max rel dU err %.2e, box %.6f" % (acc.sum(), len(acc), err.max(),
                                                               box[0]))
    assert err.max() <= 1e-10
    np.testing.assert_allclose(box, farr(mc["final_len"]), rtol=1e-12)
    assert np.max(np.abs(xyz - farr(mc["final_xyz"]))) <= 1e-9
    assert " ".join(str(w) for w in words) + " %d" % pos == mc["final_rng"].strip()
    assert st["utot"] == pytest.approx(mc["utot"], rel=1e-9, abs=1e-9)
    assert u1 == pytest.approx(mc["ufinal"], rel=1e-10)


def test_isobaric_refuses_ewald_and_missing_pressure(gpu):
    from fixture_systems import fgx_system_from_fixture
    s = fgx_system_from_fixture(load_fixture("salt_dhhs"))
    try:
        with pytest.raises(gpu.FgxError):
            s.isobaric_propagate(0.1, 1.0, 1)
    finally:
        s.close()
    s = fgx_system_from_fixture(load_fixture("ewald_tinfoil"))
    try:
        with pytest.raises(gpu.FgxError):
            s.isobaric_propagate(0.1, 1.0, 1)
    finally:
        s.close()
