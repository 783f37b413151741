"""Reference input files driving the device engine unchanged
(beofaunus_amd.faunus_io): a JSON input and a Space state file are read,
the run continues from the state file's global engine state as the reference
does after Space::load, and the final state is written back in the same
format.  The written positions, mass centres and engine state must equal
what the compiled reference left after the same Propagator::move() calls
(golden fixtures), and the reference's own bulk example must give the
reference's energies."""
import os
import sys
import tempfile

import numpy as np
import pytest

from oracle_py import load_fixture
from test_faunus_io import kind_args, write_inputs

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


@pytest.mark.parametrize("name", ["salt_dhhs", "lj_cutshift", "ewald_tinfoil", "mol_dhhs",
                                  "mol_dhhs_offcenter", "salt_npt", "ewald_npt_mixed",
                                  "mol_npt_mixed", "salt_npt_random"])
def test_run_from_reference_files(gpu, name):
    from beofaunus_amd import faunus_io
    fx = load_fixture(name)
    mc = fx["mc"]
    say("[io] %s: %d Propagator::move() calls from the reference's files" % (name, mc["ncalls"]))
    with tempfile.TemporaryDirectory() as d:
        c, s = write_inputs(fx, d)
        out = os.path.join(d, "state.out")
        st = faunus_io.run_input(c, s, out, mc["ncalls"], **kind_args(fx["kind"]))
        fin = faunus_io.read_state(out)
    assert st["trials"] == len(mc["du"])
    assert st["accepted"] == sum(mc["acc"])
    ref_xyz = np.asarray(mc["final_xyz"], dtype=np.float64).reshape(-1, 3)
    assert np.max(np.abs(fin["xyz"] - ref_xyz)) <= 1e-9
    ref_cm = np.asarray(mc["final_cm"], dtype=np.float64).reshape(-1, 3)
    cm = np.array([g["cm"] for g in fin["groups"]])
    assert np.max(np.abs(cm - ref_cm)) <= 1e-9
    assert fin["rng"] == mc["final_rng"].strip()
    if "final_len" in mc:
        np.testing.assert_allclose(fin["box"], mc["final_len"], rtol=1e-12)


def test_reference_example_energies(gpu):
    from beofaunus_amd import faunus_io, workloads
    fx = load_fixture("example_bulk_lj")
    e = fx["energy"]
    say("[io] bulk example (2304 ions), LJ energies from bulk.json + bulk.state")
    with tempfile.TemporaryDirectory() as d:
        c, s = write_inputs(fx, d)
        w = faunus_io.read_input(c, s, **kind_args("lj"))
    sy = workloads.make_system(w)
    try:
        u = sy.system_energy()
        i = int(e["i2all_i"][3])
        ui = sy.i2all(i)
    finally:
        sy.close()
    assert u == pytest.approx(e["system_energy"], rel=1e-10)
    assert ui == pytest.approx(e["i2all"][3], rel=1e-10, abs=1e-10)
