"""Device g(r) sampling (fgx_rdf, Analysis::AtomRDF::update,
analysis.h:2119-2134) against a brute-force count over all pairs: bin
int(r/dr + 0.5) of the minimum-image distance, unordered type pairs, every
count identical."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def brute(xyz, types, L, ta, tb, dr, nb):
    n = len(types)
    out = np.zeros(nb, dtype=np.int64)
    for i in range(n - 1):
        d = xyz[i + 1:] - xyz[i]
        d -= L * np.round(d / L)
        r = np.sqrt((d * d).sum(1))
        ok = ((types[i] == ta) & (types[i + 1:] == tb)) | ((types[i] == tb) & (types[i + 1:] == ta))
        b = (r[ok] / dr + 0.5).astype(int)
        np.add.at(out, b, 1)
    return out


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_rdf_counts_exact(prec):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from beofaunus_amd import fgx, workloads
    print("[rdf] salt n=700, %s" % prec, flush=True)
    w = workloads.salt_dh(n=700, L=50.0, seed=12)
    s = workloads.make_system(w, precision=fgx.FP64 if prec == "fp64" else fgx.MIXED)
    L = np.array(w["box"])
    for ta, tb in ((1, 2), (1, 1), (2, 2)):
        got = s.rdf_counts(ta, tb, 0.1)
        ref = brute(w["xyz"], w["types"], L, ta, tb, 0.1, len(got))
        assert np.array_equal(got.astype(np.int64), ref), (ta, tb)
    # the counts cover every pair of the chosen types
    n1 = int((w["types"] == 1).sum())
    assert s.rdf_counts(1, 1, 0.1).sum() == n1 * (n1 - 1) // 2
    s.close()
