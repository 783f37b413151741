"""Ensemble averages of device-resident chains: <U> (Analysis::SystemEnergy,
analysis.h:1978-1997) and the pair distribution g(r) (Analysis::AtomRDF,
analysis.h:2117-2148, normalised as PairFunctionBase::normalize,
src/analysis.cpp:190-206).  Chains advance in batched launches; samples are
taken between launches with FP64 energies and FP64 distances, whatever the
precision of the chains' trial moves.
"""
import math

import numpy as np

from . import fgx


def rdf_normalize(counts, dr, volume):
    """g(r_k) = N_k / sum(N) * <V> / (4 pi r_k^2 dr), r_k = k dr (bin k holds
    distances rounding to k dr).  Bin 0 (r = 0) has no shell volume and is
    returned as NaN."""
    counts = np.asarray(counts, dtype=np.float64)
    k = np.arange(len(counts))
    r = k * dr
    with np.errstate(divide="ignore", invalid="ignore"):
        g = counts / counts.sum() * volume / (4.0 * math.pi * r * r * dr)
    g[0] = np.nan
    return r, g


def sample_chains(systems, equil_sweeps, nsamples, every, pairs, dr):
    """Advance `systems` (one batch) by `equil_sweeps`, then take `nsamples`
    samples `every` sweeps apart.  Returns per-chain arrays: mean U (kT),
    and for each type pair the summed RDF counts."""
    if equil_sweeps:
        fgx.propagate_batch(systems, equil_sweeps)
    R = len(systems)
    usum = np.zeros(R)
    counts = {p: None for p in pairs}
    per = {p: [None] * R for p in pairs}
    for _ in range(nsamples):
        fgx.propagate_batch(systems, every)
        for c, s in enumerate(systems):
            usum[c] += s.system_energy()
            for p in pairs:
                h = s.rdf_counts(p[0], p[1], dr).astype(np.float64)
                per[p][c] = h if per[p][c] is None else per[p][c] + h
    for p in pairs:
        counts[p] = np.array(per[p])
    return usum / nsamples, counts


def compare(ua, ub, ca, cb, dr, volume, min_counts=200):
    """Agreement of two ensembles of chains: z-score of the difference of the
    chain-averaged <U>, and the largest |z| over g(r) bins with enough counts
    (errors from the chain-to-chain spread)."""
    out = {}
    ma, mb = ua.mean(), ub.mean()
    se = math.sqrt(ua.var(ddof=1) / len(ua) + ub.var(ddof=1) / len(ub))
    out["U_a"], out["U_b"], out["U_se"] = float(ma), float(mb), se
    out["U_z"] = float(abs(ma - mb) / se) if se > 0 else 0.0
    zmax = {}
    for p in ca:
        # per-chain g(r), then mean and error over chains
        ga = np.array([rdf_normalize(h, dr, volume)[1] for h in ca[p]])
        gb = np.array([rdf_normalize(h, dr, volume)[1] for h in cb[p]])
        ok = (ca[p].sum(0) >= min_counts) & (cb[p].sum(0) >= min_counts)
        ok[0] = False
        d = ga[:, ok].mean(0) - gb[:, ok].mean(0)
        s = np.sqrt(ga[:, ok].var(0, ddof=1) / len(ga) + gb[:, ok].var(0, ddof=1) / len(gb))
        z = np.abs(d) / np.where(s > 0, s, np.inf)
        zmax["%d-%d" % p] = {"bins": int(ok.sum()), "max_z": float(z.max()) if len(z) else 0.0,
                             "frac_z_gt_3": float((z > 3).mean()) if len(z) else 0.0}
    out["rdf"] = zmax
    return out
