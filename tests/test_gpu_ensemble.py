"""The mixed path samples the same ensemble as the FP64 path (which replays
the reference's Markov chain exactly): <U> and g(r) of many independent
salt chains agree within statistical error (|z| <= 5 for <U>, every g(r)
bin with enough counts within 5.5 sigma), the error taken from the
chain-to-chain spread."""
import pytest

pytestmark = pytest.mark.gpu


def test_mixed_ensemble_matches_fp64():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from beofaunus_amd import analysis, fgx, workloads
    print("[ensemble] 48 chains x 2 precisions, salt n=200", flush=True)
    n = 200
    L = 60.0 * (n / 500.0) ** (1.0 / 3.0)
    out = {}
    for prec in (fgx.FP64, fgx.MIXED):
        systems = []
        for c in range(48):
            w = workloads.salt_dh(n=n, L=L, seed=700 + c)
            s = workloads.make_system(w, precision=prec)
            s.rng_seed(9000 + c)
            s.propagator_init()
            systems.append(s)
        out[prec] = analysis.sample_chains(systems, 30, 20, 3, [(1, 2), (1, 1)], 0.2)
        for s in systems:
            s.close()
    (ua, ca), (ub, cb) = out[fgx.FP64], out[fgx.MIXED]
    cmp = analysis.compare(ua, ub, ca, cb, 0.2, L ** 3)
    print("  <U> fp64 %.3f mixed %.3f (se %.3f, z %.2f); rdf %s"
          % (cmp["U_a"], cmp["U_b"], cmp["U_se"], cmp["U_z"], cmp["rdf"]), flush=True)
    assert cmp["U_z"] <= 5.0
    for p, r in cmp["rdf"].items():
        assert r["bins"] > 20, p
        assert r["max_z"] <= 5.5, p
