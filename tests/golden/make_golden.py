#!/usr/bin/env python3
"""Write the committed golden fixtures for the Delta-U path.

Each fixture is produced by running the reference's own C++ hot path
(``oracle/_ref/ref_harness``, compiled from the reference sources in place by
``make -C oracle ref``) on a small seeded system.  A fixture holds the inputs
(box, particles, groups, types, moves, RNG seed) and what the reference
computed: energies of every Energybase entry point, and the per-trial energy
change / accept sequence of a short Metropolis run.

Run in the build container (where the reference tree exists):

    make -C oracle ref && python tests/golden/make_golden.py

The GPU box never needs the reference: tests read only the JSON written here.
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
REF = os.environ.get("FAUNUS_REF", "/root/reference")


# --------------------------------------------------------------------------
# mt19937 text state for a seed (std::mt19937::seed, C++ [rand.eng.mers])
# --------------------------------------------------------------------------
def mt_state_text(seed):
    x = [0] * 624
    x[0] = seed & 0xFFFFFFFF
    for i in range(1, 624):
        x[i] = (1812433253 * (x[i - 1] ^ (x[i - 1] >> 30)) + i) & 0xFFFFFFFF
    return " ".join(str(v) for v in x) + " 624"


# --------------------------------------------------------------------------
# seeded configurations
# --------------------------------------------------------------------------
def min_image(d, L):
    return d - L * np.round(d / L)


def rsa(rng, n, L, radii, placed_xyz=None, placed_r=None, max_tries=200000):
    """Random sequential addition of n spheres in a periodic cube of side L,
    rejecting any overlap (|r_ij| < r_i + r_j) with earlier spheres."""
    xyz = [] if placed_xyz is None else list(placed_xyz)
    rr = [] if placed_r is None else list(placed_r)
    out = []
    tries = 0
    for k in range(n):
        while True:
            tries += 1
            if tries > max_tries:
                raise RuntimeError("RSA failed")
            p = (rng.random(3) - 0.5) * L
            if xyz:
                d = min_image(np.asarray(xyz) - p, L)
                r2 = (d * d).sum(1)
                if np.any(r2 < (np.asarray(rr) + radii[k]) ** 2):
                    continue
            xyz.append(p)
            rr.append(radii[k])
            out.append(p)
            break
    return np.array(out)


def fmt(v):
    return repr(float(v))


def write_state(path, L, xyz, charge, radius, mw, ids, groups, rng_text):
    """Space::save text format (space.h:736-764, point.h:544-567,
    group.h:392-398)."""
    with open(path, "w") as f:
        f.write(f"{fmt(L[0])} {fmt(L[1])} {fmt(L[2])}\n{len(xyz)}\n")
        for i in range(len(xyz)):
            f.write(" ".join(fmt(c) for c in xyz[i]) +
                    f" {fmt(charge[i])} {fmt(radius[i])} {fmt(mw[i])} {int(ids[i])} 0\n")
        f.write(f"{len(groups)}\n")
        for g in groups:
            f.write(f"{g['front']} {g['back']} {g['molid']} {g['molsize']} 0 0 0\n")
        f.write("randomstate\n" + rng_text + "\n")


def run_harness(args):
    out = subprocess.run([HARNESS] + args, check=True, capture_output=True, text=True).stdout
    return json.loads(out)


def build_fixture(name, kind, cfg, L, parts, groups, seed, ncalls, moves_meta, energy=True,
                  mc_extra=(), mc_mode="mc"):
    with tempfile.TemporaryDirectory() as td:
        cpath = os.path.join(td, "in.json")
        spath = os.path.join(td, "state")
        with open(cpath, "w") as f:
            json.dump(cfg, f, indent=1)
        write_state(spath, L, parts["xyz"], parts["charge"], parts["radius"], parts["mw"],
                    parts["id"], groups, mt_state_text(seed))
        fx = {
            "name": name,
            "kind": kind,
            "box": list(map(float, L)),
            "config": cfg,
            "particles": {k: (np.asarray(v).reshape(-1).tolist()) for k, v in parts.items()},
            "groups": groups,
            "moves": moves_meta,
            "rng_seed": seed,
        }
        if energy:
            fx["energy"] = run_harness(["energy", kind, cpath, spath])
        if ncalls > 0:
            fx["mc"] = run_harness([mc_mode, kind, cpath, spath, str(ncalls)] +
                                   [str(a) for a in mc_extra])
            fx["mc"]["ncalls"] = ncalls
            fx["mc"]["mode"] = mc_mode
            if mc_extra:
                fx["mc"]["ncalls_at"] = int(mc_extra[0])
    path = os.path.join(HERE, name + ".json")
    with open(path, "w") as f:
        json.dump(fx, f, separators=(",", ":"))
    print(f"wrote {path} ({os.path.getsize(path) // 1024} KiB)")
    return fx


def lB_epsr(lB, T=298.15):
    """relative permittivity giving Bjerrum length lB at T (physconst.h:91-94)"""
    e, e0, kB, pi = 1.602177e-19, 8.85419e-12, 1.380658e-23, np.arccos(-1.0)
    return e * e / (4 * pi * e0 * 1e-10 * kB * T) / lB


def salt(name, kind, n, L, seed, ncalls, elec_cfg, dp=2.0, r=2.0):
    rng = np.random.default_rng(seed)
    radii = np.full(n, r)
    xyz = rsa(rng, n, L, radii)
    half = n // 2
    # atom names sort as insertion order (nlohmann json objects are ordered
    # maps), so ids are 1 = "T1p", 2 = "T2m"; id 0 is the dummy "unk"
    ids = np.array([1] * half + [2] * (n - half))
    charge = np.where(ids == 1, 1.0, -1.0)
    cfg = {
        "system": {"temperature": 298.15, "geometry": {"length": L}},
        "atomlist": {"T1p": {"q": 1.0, "r": r, "dp": dp, "mw": 1.0},
                     "T2m": {"q": -1.0, "r": r, "dp": dp, "mw": 1.0}},
        "moleculelist": {"salt": {"atoms": "T1p T2m", "atomic": True, "Ninit": 0}},
        "moves": {"atomtranslate": {"salt": {"peratom": True}}},
        "energy": {"nonbonded": elec_cfg},
    }
    parts = {"xyz": xyz, "charge": charge, "radius": radii, "mw": np.ones(n), "id": ids}
    groups = [{"front": 0, "back": n - 1, "molid": 0, "molsize": 1}]
    moves = [{"type": "AT", "molid": 0, "prob": 1.0, "peratom": 1, "permol": 0}]
    return build_fixture(name, kind, cfg, [L, L, L], parts, groups, seed, ncalls, moves)


def salt_npt(name, n, L, seed, ncalls, pressure_mM, dp, r=2.0, with_at_and_random=False):
    """Salt with Debye-Hueckel + hard spheres and the ExternalPressure term,
    a Propagator holding one Isobaric move (atomic salt group scaled); or,
    with_at_and_random, a Propagator holding AtomicTranslation and Isobaric
    whose move engine the input's moves.random object reseeds
    (move.h:7783-7787: a default-seeded mt19937 without "hardware")."""
    rng = np.random.default_rng(seed)
    radii = np.full(n, r)
    xyz = rsa(rng, n, L, radii)
    half = n // 2
    ids = np.array([1] * half + [2] * (n - half))
    charge = np.where(ids == 1, 1.0, -1.0)
    cfg = {
        "system": {"temperature": 298.15, "geometry": {"length": L}},
        "atomlist": {"T1p": {"q": 1.0, "r": r, "dp": 2.0, "mw": 1.0},
                     "T2m": {"q": -1.0, "r": r, "dp": 2.0, "mw": 1.0}},
        "moleculelist": {"salt": {"atoms": "T1p T2m", "atomic": True, "Ninit": 0}},
        "moves": {"isobaric": {"dp": dp, "pressure": pressure_mM, "prob": 1.0}},
        "energy": {"nonbonded": {"epsr": lB_epsr(7.1), "debyelength": 10.0}},
    }
    parts = {"xyz": xyz, "charge": charge, "radius": radii, "mw": np.ones(n), "id": ids}
    groups = [{"front": 0, "back": n - 1, "molid": 0, "molsize": 1}]
    moves = [{"type": "NPT", "molid": -1, "prob": 1.0, "peratom": 0, "permol": 0, "dp": dp,
              "pressure_mM": pressure_mM}]
    if with_at_and_random:
        cfg["moves"]["atomtranslate"] = {"salt": {"peratom": True}}
        cfg["moves"]["random"] = {}
        moves = [{"type": "AT", "molid": 0, "prob": 1.0, "peratom": 1, "permol": 0}] + moves
        return build_fixture(name, "dhhs_npt", cfg, [L, L, L], parts, groups, seed, ncalls,
                             moves, mc_mode="mcprop")
    return build_fixture(name, "dhhs_npt", cfg, [L, L, L], parts, groups, seed, ncalls, moves)


def ewald_npt(name, n, L, seed, ncalls, ncalls_at, pressure_mM, dp, eps_surf,
              update_frequency, alpha=0.35, rc=7.0, kc=4.0, mixed=False):
    """Ewald salt (NonbondedEwald<LennardJonesLB>) with the ExternalPressure
    term and a Propagator holding an AtomicTranslation and an Isobaric move:
    `ncalls` volume trials, then `ncalls_at` AtomicTranslation calls; or
    (mixed) `ncalls` calls picking either move as Propagator::move does."""
    rng = np.random.default_rng(seed)
    half = n // 2
    radii = np.full(n, 1.5)
    xyz = rsa(rng, n, L, radii)
    ids = np.array([1] * half + [2] * (n - half))
    charge = np.where(ids == 1, 1.0, -1.0)
    cfg = {
        "system": {"temperature": 298.15, "geometry": {"length": L}},
        "atomlist": {"T1p": {"q": 1.0, "sigma": 3.0, "eps": 0.2479, "dp": 1.0, "mw": 1.0},
                     "T2m": {"q": -1.0, "sigma": 3.0, "eps": 0.2479, "dp": 1.0, "mw": 1.0}},
        "moleculelist": {"salt": {"atoms": "T1p T2m", "atomic": True, "Ninit": 0}},
        "moves": {"isobaric": {"dp": dp, "pressure": pressure_mM, "prob": 1.0},
                  "atomtranslate": {"salt": {"peratom": True}}},
        "energy": {"nonbonded": {"epsr": lB_epsr(7.1), "ewald": {
            "alpha": alpha, "cutoff": rc, "cutoffK": kc, "debyelength": 1e12,
            "eps_surf": eps_surf, "spherical_sum": True, "update_frequency": update_frequency}}},
    }
    parts = {"xyz": xyz, "charge": charge, "radius": radii, "mw": np.ones(n), "id": ids}
    groups = [{"front": 0, "back": n - 1, "molid": 0, "molsize": 1}]
    moves = [{"type": "AT", "molid": 0, "prob": 1.0, "peratom": 1, "permol": 0},
             {"type": "NPT", "molid": -1, "prob": 1.0, "peratom": 0, "permol": 0, "dp": dp,
              "pressure_mM": pressure_mM}]
    if mixed:
        return build_fixture(name, "ewaldlj_npt", cfg, [L, L, L], parts, groups, seed, ncalls,
                             moves, mc_mode="mcprop")
    return build_fixture(name, "ewaldlj_npt", cfg, [L, L, L], parts, groups, seed, ncalls, moves,
                         mc_extra=(ncalls_at,))


def ewald_neutral(name, n, n0, L, seed, ncalls, alpha=0.35, rc=7.0, kc=4.0):
    """Ewald salt with n0 particles of a neutral atom type.  The reference's
    effective-charge ratio of a q = 0 type is 0/0 = NaN (ewald.h:535-545), so
    S(k) and every energy change are NaN, and Movebase::metropolis accepts a
    NaN change (move.h:567-573)."""
    rng = np.random.default_rng(seed)
    half = n // 2
    radii = np.full(n + n0, 1.5)
    xyz = rsa(rng, n + n0, L, radii)
    ids = np.array([1] * half + [2] * (n - half) + [3] * n0)
    charge = np.where(ids == 1, 1.0, np.where(ids == 2, -1.0, 0.0))
    cfg = {
        "system": {"temperature": 298.15, "geometry": {"length": L}},
        "atomlist": {"T1p": {"q": 1.0, "sigma": 3.0, "eps": 0.2479, "dp": 1.0, "mw": 1.0},
                     "T2m": {"q": -1.0, "sigma": 3.0, "eps": 0.2479, "dp": 1.0, "mw": 1.0},
                     "T3n": {"q": 0.0, "sigma": 3.0, "eps": 0.2479, "dp": 1.0, "mw": 1.0}},
        "moleculelist": {"salt": {"atoms": "T1p T2m T3n", "atomic": True, "Ninit": 0}},
        "moves": {"atomtranslate": {"salt": {"peratom": True}}},
        "energy": {"nonbonded": {"epsr": lB_epsr(7.1), "ewald": {
            "alpha": alpha, "cutoff": rc, "cutoffK": kc, "debyelength": 1e12,
            "eps_surf": 0.5, "spherical_sum": True, "update_frequency": 5}}},
    }
    parts = {"xyz": xyz, "charge": charge, "radius": radii, "mw": np.ones(n + n0), "id": ids}
    groups = [{"front": 0, "back": n + n0 - 1, "molid": 0, "molsize": 1}]
    moves = [{"type": "AT", "molid": 0, "prob": 1.0, "peratom": 1, "permol": 0}]
    return build_fixture(name, "ewaldlj", cfg, [L, L, L], parts, groups, seed, ncalls, moves)


def lj_mix(name, n, L, seed, ncalls, cutoff=None):
    rng = np.random.default_rng(seed)
    half = n // 2
    # sigma 3 and 4.5 A, eps 0.15 / 0.25 kJ/mol (Lorentz-Berthelot mixing)
    radii = np.array([1.5] * half + [2.25] * (n - half))
    xyz = rsa(rng, n, L, radii * 0.9)
    ids = np.array([1] * half + [2] * (n - half))
    cfg = {
        "system": {"temperature": 298.15, "geometry": {"length": L}},
        "atomlist": {"T1a": {"eps": 0.15, "sigma": 3.0, "dp": 1.5, "mw": 1.0},
                     "T2b": {"eps": 0.25, "sigma": 4.5, "dp": 1.0, "mw": 1.0}},
        "moleculelist": {"fluid": {"atoms": "T1a T2b", "atomic": True, "Ninit": 0}},
        "moves": {"atomtranslate": {"fluid": {"peratom": True}}},
        "energy": {"nonbonded": {} if cutoff is None else {"cutoff": cutoff}},
    }
    parts = {"xyz": xyz, "charge": np.zeros(n), "radius": radii, "mw": np.ones(n), "id": ids}
    groups = [{"front": 0, "back": n - 1, "molid": 0, "molsize": 1}]
    moves = [{"type": "AT", "molid": 0, "prob": 1.0, "peratom": 1, "permol": 0}]
    kind = "lj" if cutoff is None else "ljcut"
    return build_fixture(name, kind, cfg, [L, L, L], parts, groups, seed, ncalls, moves)


def ewald(name, n, L, seed, ncalls, eps_surf, update_frequency, alpha=0.35, rc=7.0, kc=4.0):
    rng = np.random.default_rng(seed)
    half = n // 2
    radii = np.full(n, 1.5)
    xyz = rsa(rng, n, L, radii)
    ids = np.array([1] * half + [2] * (n - half))
    charge = np.where(ids == 1, 1.0, -1.0)
    cfg = {
        "system": {"temperature": 298.15, "geometry": {"length": L}},
        "atomlist": {"T1p": {"q": 1.0, "sigma": 3.0, "eps": 0.2479, "dp": 1.0, "mw": 1.0},
                     "T2m": {"q": -1.0, "sigma": 3.0, "eps": 0.2479, "dp": 1.0, "mw": 1.0}},
        "moleculelist": {"salt": {"atoms": "T1p T2m", "atomic": True, "Ninit": 0}},
        "moves": {"atomtranslate": {"salt": {"peratom": True}}},
        "energy": {"nonbonded": {"epsr": lB_epsr(7.1), "ewald": {
            "alpha": alpha, "cutoff": rc, "cutoffK": kc, "debyelength": 1e12,
            "eps_surf": eps_surf, "spherical_sum": True, "update_frequency": update_frequency}}},
    }
    parts = {"xyz": xyz, "charge": charge, "radius": radii, "mw": np.ones(n), "id": ids}
    groups = [{"front": 0, "back": n - 1, "molid": 0, "molsize": 1}]
    moves = [{"type": "AT", "molid": 0, "prob": 1.0, "peratom": 1, "permol": 0}]
    return build_fixture(name, "ewaldlj", cfg, [L, L, L], parts, groups, seed, ncalls, moves)


def molecules(name, kind, nmol, nsite, nion, L, seed, ncalls, center_rotation=True, isobaric=None,
              cutoff_g2g=None, R=9.0):
    """Rigid molecules (sites packed in a sphere of radius R, charges from
    {-1,0,+1} with p = 0.2/0.6/0.2, radii U[1.5,3]) plus monovalent ions."""
    rng = np.random.default_rng(seed)
    # molecule centres: RSA of exclusion spheres of radius R + 3
    centres = rsa(rng, nmol, L, np.full(nmol, R + 3.0))
    xyz, charge, radius, ids = [], [], [], []
    tnames = {}
    atomlist = {}
    groups = []
    for m in range(nmol):
        front = len(xyz)
        local, lr = [], []
        for s in range(nsite):
            r = 1.5 + 1.5 * rng.random()
            while True:   # sites do not overlap inside a molecule
                v = (rng.random(3) - 0.5) * 2 * R
                if (v * v).sum() > R * R:
                    continue
                if local and np.any(((np.asarray(local) - v) ** 2).sum(1) < (np.asarray(lr) + r) ** 2):
                    continue
                break
            local.append(v)
            lr.append(r)
            q = rng.choice([-1.0, 0.0, 1.0], p=[0.2, 0.6, 0.2])
            p = min_image(centres[m] + v, L)
            xyz.append(p)
            charge.append(q)
            radius.append(r)
        groups.append({"front": front, "back": len(xyz) - 1, "molid": 0, "molsize": nsite})
    # one atom type per site charge class (ids 1..3 sorted names), ions 4,5
    for q, nm in ((-1.0, "M1n"), (0.0, "M2z"), (1.0, "M3p")):
        atomlist[nm] = {"q": q, "r": 2.0, "mw": 1.0}
    atomlist["M4ip"] = {"q": 1.0, "r": 2.0, "dp": 2.0, "mw": 1.0}
    atomlist["M5im"] = {"q": -1.0, "r": 2.0, "dp": 2.0, "mw": 1.0}
    ids = [1 if q < 0 else (2 if q == 0 else 3) for q in charge]
    ion_r = np.full(nion, 2.0)
    ion_xyz = rsa(rng, nion, L, ion_r, placed_xyz=[np.asarray(x) for x in xyz], placed_r=radius)
    gfront = len(xyz)
    for k in range(nion):
        xyz.append(ion_xyz[k])
        charge.append(1.0 if k % 2 == 0 else -1.0)
        radius.append(2.0)
        ids.append(4 if k % 2 == 0 else 5)
    if nion:
        groups.append({"front": gfront, "back": len(xyz) - 1, "molid": 1, "molsize": 1})
    nb = {"epsr": lB_epsr(7.1), "debyelength": 15.0}
    if cutoff_g2g is not None:
        nb["cutoff_g2g"] = cutoff_g2g
    # molecule ids follow moleculelist key order from 0: "amol" = 0, "bions" = 1
    cfg = {
        "system": {"temperature": 298.15, "geometry": {"length": L}},
        "atomlist": atomlist,
        "moleculelist": {"amol": {"atoms": " ".join(["M2z"] * nsite), "Ninit": 0},
                         "bions": {"atoms": "M4ip M5im", "atomic": True, "Ninit": 0}},
        "moves": {"moltransrot": {"center_rotation": center_rotation,
                                  "amol": {"dp": 5.0, "dprot": 0.5, "permol": True}}},
        "energy": {"nonbonded": nb},
    }
    parts = {"xyz": np.array(xyz), "charge": np.array(charge), "radius": np.array(radius),
             "mw": np.ones(len(xyz)), "id": np.array(ids)}
    moves = [{"type": "TR", "molid": 0, "prob": 1.0, "peratom": 0, "permol": 1,
              "dp": 5.0, "dprot": 0.5, "center_rotation": int(center_rotation)}]
    if isobaric is None:
        return build_fixture(name, kind, cfg, [L, L, L], parts, groups, seed, ncalls, moves)
    # one Propagator holding isobaric and moltransrot (sorted key order):
    # volume moves scale the molecules' mass centres and keep them rigid
    cfg["moves"]["isobaric"] = {"dp": isobaric["dp"], "pressure": isobaric["pressure"],
                                "prob": 1.0}
    moves = [{"type": "NPT", "molid": -1, "prob": 1.0, "peratom": 0, "permol": 0,
              "dp": isobaric["dp"], "pressure_mM": isobaric["pressure"]}] + moves
    return build_fixture(name, kind, cfg, [L, L, L], parts, groups, seed, ncalls, moves,
                         mc_mode="mcprop")


def reference_example(name, kind, example):
    """Energies the reference computes for one of its own example inputs
    (src/examples/<example>.json + .state, read unchanged) under the pair
    potential `kind`; the fixture keeps the input data and the results."""
    ex = os.path.join(REF, "src", "examples", example)
    with open(ex + ".json") as f:
        cfg = json.load(f)
    st = read_state_text(ex + ".state")
    fx = {
        "name": name,
        "kind": kind,
        "source": "src/examples/%s.json, src/examples/%s.state" % (example, example),
        "box": st["box"],
        "config": {k: cfg[k] for k in ("system", "atomlist", "moleculelist", "moves", "energy")
                   if k in cfg},
        "particles": {"xyz": st["xyz"], "charge": st["charge"], "radius": st["radius"],
                      "mw": st["mw"], "id": st["id"]},
        "groups": st["groups"],
        "moves": [],
        "rng": st["rng"],
        "energy": run_harness(["energy", kind, ex + ".json", ex + ".state"]),
    }
    path = os.path.join(HERE, name + ".json")
    with open(path, "w") as f:
        json.dump(fx, f, separators=(",", ":"))
    print(f"wrote {path} ({os.path.getsize(path) // 1024} KiB)")
    return fx


def read_state_text(path):
    """Space::load text (space.h:807-880) as plain lists"""
    tok = open(path).read().split()
    box = [float(v) for v in tok[:3]]
    n = int(tok[3])
    rows = [tok[4 + 8 * i:12 + 8 * i] for i in range(n)]
    pos = 4 + 8 * n
    ng = int(tok[pos])
    pos += 1
    groups = []
    for _ in range(ng):
        g = tok[pos:pos + 7]
        pos += 7
        groups.append({"front": int(g[0]), "back": int(g[1]), "molid": int(g[2]),
                       "molsize": int(g[3])})
    rng = " ".join(tok[pos + 1:pos + 626]) if pos < len(tok) and tok[pos] == "randomstate" else None
    return {"box": box, "xyz": [float(v) for r in rows for v in r[:3]],
            "charge": [float(r[3]) for r in rows], "radius": [float(r[4]) for r in rows],
            "mw": [float(r[5]) for r in rows], "id": [int(r[6]) for r in rows],
            "groups": groups, "rng": rng}


def rng_fixture():
    fx = {"seed_default": run_harness(["rng", "0", "2000"]),
          "seed_7": run_harness(["rng", "7", "500"])}
    path = os.path.join(HERE, "rng.json")
    with open(path, "w") as f:
        json.dump(fx, f, separators=(",", ":"))
    print(f"wrote {path}")


def main(only=()):
    """Writes every fixture, or only those named on the command line."""
    if not os.path.exists(HARNESS):
        sys.exit("build oracle/_ref/ref_harness first: make -C oracle ref")
    if only:
        for f in only:
            ONLY[f]()
        return
    rng_fixture()
    dh = {"epsr": lB_epsr(7.1), "debyelength": 10.0}
    salt("salt_dhhs", "dhhs", 80, 30.0, 11, 25, dh)
    salt("salt_dhhs_vec", "dhhs_vec", 40, 25.0, 12, 5, dh)
    salt("salt_coulhs", "coulhs", 60, 30.0, 13, 10, {"epsr": lB_epsr(7.1)})
    salt("salt_dh", "dh", 40, 20.0, 14, 8, dh)
    lj_mix("lj_mix", 50, 25.0, 15, 10)
    lj_mix("lj_cutshift", 80, 26.0, 21, 10, cutoff=8.0)
    salt_npt("salt_npt", 60, 30.0, 22, 60, pressure_mM=300.0, dp=0.4)
    ewald("ewald_tinfoil", 40, 20.0, 16, 6, eps_surf=0.5, update_frequency=7)
    ewald("ewald_surface", 30, 18.0, 17, 4, eps_surf=1.0, update_frequency=-1, kc=3.0)
    molecules("mol_dhhs", "dhhs", 4, 12, 16, 60.0, 18, 12)
    molecules("mol_dhhs_offcenter", "dhhs", 3, 10, 10, 50.0, 19, 10, center_rotation=False)
    molecules("mol_dhhs_cut", "dhhs_cut", 5, 8, 10, 70.0, 20, 10, cutoff_g2g=25.0)
    # the NaCl melt of the reference's bulk example (2304 ions, 1100 K), LJ part
    reference_example("example_bulk_lj", "lj", "bulk")
    for f in ONLY.values():
        f()


# Ewald with volume moves: tin-foil, and a finite surface dielectric (the
# surface term changes with V); refreshes of S(k) every 3 accepted moves
ONLY = {
    "ewald_npt": lambda: ewald_npt("ewald_npt", 40, 20.0, 23, 30, 4, pressure_mM=6000.0,
                                   dp=0.4, eps_surf=0.5, update_frequency=3),
    "ewald_npt_surface": lambda: ewald_npt("ewald_npt_surface", 30, 18.0, 24, 24, 3,
                                           pressure_mM=6000.0, dp=0.3, eps_surf=1.0,
                                           update_frequency=3, kc=3.0),
    # a neutral atom type under Ewald: NaN energies, every trial accepted
    "ewald_neutral": lambda: ewald_neutral("ewald_neutral", 30, 4, 18.0, 26, 3),
    # rigid molecules and ions: one Propagator picking TranslateRotate or Isobaric
    "mol_npt_mixed": lambda: molecules("mol_npt_mixed", "dhhs_npt", 4, 12, 16, 60.0, 27, 30,
                                       isobaric={"dp": 0.3, "pressure": 150.0}),
    # |n| <= 12: more k-vectors than the LDS-resident device S(k) holds
    "ewald_kc12": lambda: ewald("ewald_kc12", 40, 20.0, 29, 6, eps_surf=0.5, update_frequency=7,
                                kc=12.0),
    # AtomicTranslation + Isobaric with moves.random reseeding the move engine
    "salt_npt_random": lambda: salt_npt("salt_npt_random", 50, 30.0, 28, 30, pressure_mM=300.0,
                                        dp=0.4, with_at_and_random=True),
    # one Propagator picking AtomicTranslation or Isobaric per call
    "ewald_npt_mixed": lambda: ewald_npt("ewald_npt_mixed", 40, 20.0, 25, 40, 0,
                                         pressure_mM=6000.0, dp=0.4, eps_surf=0.5,
                                         update_frequency=5, mixed=True),
}


if __name__ == "__main__":
    main(sys.argv[1:])
