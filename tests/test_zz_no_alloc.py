"""fgx_system_energy_noalloc (the allocation-free Energy::systemEnergy)
against fgx_system_energy: bit-identical energies, on two calls in a row
(the second on a smaller system), for the plain, NonbondedCutg2g (group-pair
mask in the scratch) and Ewald paths; a system too large for the fixed
scratch is refused with FGX_ENOMEM."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from beofaunus_amd import fgx
    return fgx


def test_noalloc_matches_allocating_entry_point(gpu):
    fgx = gpu
    from beofaunus_amd import workloads
    print("[noalloc] system energies, large then small", flush=True)
    cases = [
        ("salt N=3000", workloads.salt_dh(n=3000, L=80.0, seed=1, lattice=True)),
        ("salt N=400", workloads.salt_dh(n=400, L=40.0, seed=2, lattice=True)),
        ("Ewald N=2000", workloads.ewald_electrolyte(n=2000, L=55.7, seed=3, precision=fgx.FP64)),
        ("Ewald N=300", workloads.ewald_electrolyte(n=300, L=29.0, seed=4, precision=fgx.FP64)),
    ]
    from fixture_systems import fgx_system_from_fixture
    from oracle_py import load_fixture
    # NonbondedCutg2g molecules (group-pair mask in the scratch), between the
    # large and the small systems
    systems = [(n, workloads.make_system(w, precision=fgx.FP64)) for n, w in cases[:1]]
    systems.append(("molecules, NonbondedCutg2g",
                    fgx_system_from_fixture(load_fixture("mol_dhhs_cut"), moves=False)))
    systems += [(n, workloads.make_system(w, precision=fgx.FP64)) for n, w in cases[1:]]
    for name, s in systems:
        for which in (0, 1):
            a = s.system_energy_noalloc(which)
            b = s.system_energy(which)
            c = s.system_energy_noalloc(which)
            print("  %s which=%d: %.17g %.17g" % (name, which, a, b), flush=True)
            assert (a == b or (np.isnan(a) and np.isnan(b))) and (c == b or np.isnan(b)), name
        s.close()


def test_noalloc_refuses_oversized_system(gpu):
    fgx = gpu
    from beofaunus_amd import workloads
    print("[noalloc] oversized system refused", flush=True)
    w = workloads.salt_dh(n=70000, L=300.0, seed=5, lattice=True)
    s = workloads.make_system(w)
    with pytest.raises(fgx.FgxError, match="fixed scratch"):
        s.system_energy_noalloc()
    # the allocating entry point still serves it
    assert np.isfinite(s.system_energy())
    s.close()
