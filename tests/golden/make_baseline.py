#!/usr/bin/env python3
"""Write the reference fixtures at the benchmark configurations' own sizes
and parameters (tests/golden/baseline/*.json).

    make -C oracle ref && python tests/golden/make_baseline.py [c1 c2 c3]

Each system is handed to the reference's own hot path
(``oracle/_ref/ref_harness``, compiled from the reference sources in place)
through its JSON input and Space::save state file; the fixture keeps the
inputs and what the reference computed.  To keep the files small:
  - input coordinates sit on a grid of L / 65536 (exact doubles) and are
    stored as int16 multiples of that unit;
  - atom ids are uint8; per-particle charge, radius and mw follow the ids;
  - per-trial energy changes are float64, accept flags one bit each, both
    base64; a long run keeps its first trials and every k-th one after;
  - final coordinates: every particle (500-ion salt), the particles a run
    moved (electrolyte), or mass centres plus ten sites per molecule (rigid
    molecules, whose other sites follow rigidly).
The GPU box never needs the reference: tests read only the JSON written here.

Configurations:
  c1  salt: 500 ions +-1e, L 60 A, hard spheres 2 A, Debye-Hueckel 10 A,
      lB 7.1 A, AtomicTranslation dp 2 A per atom: 200 sweeps = 1e5 trials
  c2  Ewald LJ + Coulomb electrolyte: 20,000 ions, L 120 A, LJ sigma 4 A,
      eps 0.1 kT (Lorentz-Berthelot), Ewald alpha 0.25 / A, rc 12 A,
      |n| <= 10 spherical, tinfoil; AtomicTranslation dp 1 A, one trial per
      call: 2,000 trials
  c3  rigid molecules: 64 x 400 sites + 8,000 ions, L 250 A, Debye-Hueckel
      15 A, TranslateRotate dp 5 A, dprot 0.5 rad per molecule: one sweep of
      64 trials
"""
import base64
import json
import zlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from make_golden import HARNESS, lB_epsr, mt_state_text, rsa, write_state  # noqa: E402

OUT = os.path.join(HERE, "baseline")


def enc(a, dtype):
    """array -> {dtype, shape, zlib-compressed bytes in base64}"""
    a = np.ascontiguousarray(np.asarray(a, dtype=dtype))
    return {"dtype": np.dtype(dtype).str, "shape": list(a.shape), "zlib": True,
            "b64": base64.b64encode(zlib.compress(a.tobytes(), 9)).decode()}


def enc_bits(flags):
    return {"bits": len(flags), "zlib": True,
            "b64": base64.b64encode(zlib.compress(
                np.packbits(np.asarray(flags, dtype=np.uint8)).tobytes(), 9)).decode()}


def quantize(xyz, L):
    """coordinates on the grid k L / 65536, k int16 (exact doubles)"""
    unit = L / 65536.0
    k = np.round(np.asarray(xyz) / unit).astype(np.int64)
    k = ((k + 32768) % 65536) - 32768
    return k.astype(np.int16), unit


def overlaps(xyz, radius, gid, L, cross_groups_only=True):
    """pairs closer than r_i + r_j (between different groups), via cells"""
    from collections import defaultdict
    rmax = float(np.max(radius))
    nc = max(1, int(L // (2 * rmax)))
    cs = L / nc
    cell = defaultdict(list)
    idx = (((xyz + 0.5 * L) // cs).astype(int)) % nc
    for i, c in enumerate(map(tuple, idx)):
        cell[c].append(i)
    bad = 0
    for i, c in enumerate(map(tuple, idx)):
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in cell.get(((c[0] + dx) % nc, (c[1] + dy) % nc, (c[2] + dz) % nc), ()):
                        if j <= i or (cross_groups_only and gid[i] == gid[j]):
                            continue
                        d = xyz[j] - xyz[i]
                        d -= L * np.round(d / L)
                        if d @ d < (radius[i] + radius[j]) ** 2:
                            bad += 1
    return bad


def harness(args):
    out = subprocess.run([HARNESS] + args, check=True, capture_output=True, text=True).stdout
    return json.loads(out)


def write(name, fx):
    os.makedirs(OUT, exist_ok=True)
    path = os.path.join(OUT, name + ".json")
    with open(path, "w") as f:
        json.dump(fx, f, separators=(",", ":"))
    print(f"wrote {path} ({os.path.getsize(path) // 1024} KiB)", flush=True)


def run(name, kind, cfg, L, q16, unit, ids, tq, tr, tmw, groups, seed, ncalls, moves, keep):
    xyz = q16.astype(np.float64) * unit
    charge, radius, mw = tq[ids], tr[ids], tmw[ids]
    with tempfile.TemporaryDirectory() as td:
        cpath = os.path.join(td, "in.json")
        spath = os.path.join(td, "state")
        with open(cpath, "w") as f:
            json.dump(cfg, f, indent=1)
        write_state(spath, [L, L, L], xyz, charge, radius, mw, ids, groups, mt_state_text(seed))
        print(f"[{name}] energies", flush=True)
        en = harness(["energy", kind, cpath, spath])
        print(f"[{name}] Metropolis run, {ncalls} calls", flush=True)
        mc = harness(["mc", kind, cpath, spath, str(ncalls)])
    return xyz, en, mc, keep(en, mc)


def base_fixture(name, kind, cfg, L, q16, unit, ids, tq, tr, tmw, groups, seed, moves):
    return {
        "name": name, "kind": kind, "box": [L, L, L], "config": cfg,
        "particles": {"xyz_q16": enc(q16, "<i2"), "unit": unit, "id": enc(ids, "u1"),
                      "charge_by_id": list(map(float, tq)), "radius_by_id": list(map(float, tr)),
                      "mw_by_id": list(map(float, tmw))},
        "groups": groups, "moves": moves, "rng_seed": seed,
    }


def energy_subset(en, ng, g2g_rows):
    out = {k: en[k] for k in ("type_charge", "type_radius", "type_sigma", "type_eps",
                              "type_dp", "type_mw", "lB", "system_energy", "external",
                              "ewald_nk", "g_external", "g_internal")}
    g2g = np.array([float(v) if not isinstance(v, str) else float(v) for v in en["g2g"]])
    rows = [a for a in g2g_rows if a < ng]
    out["g2g_rows"] = rows
    out["g2g"] = enc(np.concatenate([g2g[a * ng:(a + 1) * ng] for a in rows]), "<f8")
    m = min(300, len(en["i2i"]))
    out["i2i_i"], out["i2i_j"] = en["i2i_i"][:m], en["i2i_j"][:m]
    out["i2i"] = enc([float(v) for v in en["i2i"][:m]], "<f8")
    out["i2all_i"], out["i2all"] = en["i2all_i"], en["i2all"]
    m = min(400, len(en["i2g"]))
    out["i2g_g"], out["i2g"] = en["i2g_g"][:m], en["i2g"][:m]
    return out


def fnum(v):
    return float(v) if not isinstance(v, str) else float(v)


def c1():
    name, L, n, seed = "c1_salt_n500", 60.0, 500, 101
    rng = np.random.default_rng(seed)
    # contact 4.004 A before rounding to the grid (a pair moves < 0.0017 A)
    xyz = rsa(rng, n, L, np.full(n, 2.002), max_tries=10_000_000)
    q16, unit = quantize(xyz, L)
    ids = np.array([1] * (n // 2) + [2] * (n - n // 2), dtype=np.uint8)
    assert overlaps(q16 * unit, np.full(n, 2.0), np.arange(n), L) == 0
    tq, tr, tmw = np.array([0.0, 1.0, -1.0]), np.array([0.0, 2.0, 2.0]), np.ones(3)
    cfg = {
        "system": {"temperature": 298.15, "geometry": {"length": L}},
        "atomlist": {"T1p": {"q": 1.0, "r": 2.0, "dp": 2.0, "mw": 1.0},
                     "T2m": {"q": -1.0, "r": 2.0, "dp": 2.0, "mw": 1.0}},
        "moleculelist": {"salt": {"atoms": "T1p T2m", "atomic": True, "Ninit": 0}},
        "moves": {"atomtranslate": {"salt": {"peratom": True}}},
        "energy": {"nonbonded": {"epsr": lB_epsr(7.1), "debyelength": 10.0}},
    }
    groups = [{"front": 0, "back": n - 1, "molid": 0, "molsize": 1}]
    moves = [{"type": "AT", "molid": 0, "prob": 1.0, "peratom": 1, "permol": 0}]
    fx = base_fixture(name, "dhhs", cfg, L, q16, unit, ids, tq, tr, tmw, groups, seed, moves)
    ncalls = 200
    xyz, en, mc, _ = run(name, "dhhs", cfg, L, q16, unit, ids, tq, tr, tmw, groups, seed, ncalls,
                         moves, lambda e, m: None)
    du = np.array([fnum(v) for v in mc["du"]])
    head, every = 4000, 25
    keep = np.concatenate([np.arange(min(head, len(du))), np.arange(head, len(du), every)])
    fx["energy"] = energy_subset(en, 1, [0])
    fx["mc"] = {"ncalls": ncalls, "trials": len(du), "uinit": mc["uinit"], "ufinal": mc["ufinal"],
                "utot": mc["utot"], "du_index": enc(keep, "<i4"), "du": enc(du[keep], "<f8"),
                "acc": enc_bits(mc["acc"]),
                "final_xyz": enc(np.array(mc["final_xyz"]), "<f8"),
                "final_rng": mc["final_rng"]}
    write(name, fx)


def c2():
    from beofaunus_amd import workloads
    name, n, L, seed = "c2_ewald_n20000", 20000, 120.0, 102
    w = workloads.ewald_electrolyte(n=n, L=L, seed=2)
    q16, unit = quantize(w["xyz"], L)
    ids = w["types"].astype(np.uint8)
    eps_kj = workloads.kt_to_kjmol(0.1)
    tq, tr, tmw = np.array([0.0, 1.0, -1.0]), np.array([0.0, 2.0, 2.0]), np.ones(3)
    cfg = {
        "system": {"temperature": 298.15, "geometry": {"length": L}},
        "atomlist": {"T1p": {"q": 1.0, "sigma": 4.0, "eps": eps_kj, "dp": 1.0, "mw": 1.0},
                     "T2m": {"q": -1.0, "sigma": 4.0, "eps": eps_kj, "dp": 1.0, "mw": 1.0}},
        "moleculelist": {"salt": {"atoms": "T1p T2m", "atomic": True, "Ninit": 0}},
        "moves": {"atomtranslate": {"salt": {"peratom": False}}},
        "energy": {"nonbonded": {"epsr": lB_epsr(7.1), "ewald": {
            "alpha": 0.25, "cutoff": 12.0, "cutoffK": 10.0, "debyelength": 1e12,
            "eps_surf": 0.5, "spherical_sum": True, "update_frequency": 20000}}},
    }
    groups = [{"front": 0, "back": n - 1, "molid": 0, "molsize": 1}]
    moves = [{"type": "AT", "molid": 0, "prob": 1.0, "peratom": 0, "permol": 0}]
    fx = base_fixture(name, "ewaldlj", cfg, L, q16, unit, ids, tq, tr, tmw, groups, seed, moves)
    ncalls = 2000
    xyz, en, mc, _ = run(name, "ewaldlj", cfg, L, q16, unit, ids, tq, tr, tmw, groups, seed,
                         ncalls, moves, lambda e, m: None)
    du = np.array([fnum(v) for v in mc["du"]])
    who = np.array(mc["who"], dtype=np.int64)
    moved = np.unique(who)
    fxyz = np.array(mc["final_xyz"]).reshape(-1, 3)
    fx["energy"] = energy_subset(en, 1, [0])
    fx["mc"] = {"ncalls": ncalls, "trials": len(du), "uinit": mc["uinit"], "ufinal": mc["ufinal"],
                "utot": mc["utot"], "du": enc(du, "<f8"), "acc": enc_bits(mc["acc"]),
                "who": enc(who, "<i4"), "moved": enc(moved, "<i4"),
                "final_moved_xyz": enc(fxyz[moved], "<f8"), "final_rng": mc["final_rng"]}
    write(name, fx)


def c3():
    from beofaunus_amd import workloads
    name, L, seed = "c3_rigid_64x400", 250.0, 103
    # the first configuration seed whose groups stay clear of each other on
    # the grid (ions are placed clear of every site and of each other)
    for wseed in range(3, 20):
        w = workloads.rigid_molecules(seed=wseed)
        q16, unit = quantize(w["xyz"], L)
        xq = q16.astype(np.float64) * unit
        gid = np.zeros(len(xq), dtype=np.int64)
        for g, G in enumerate(w["groups"]):
            gid[G["front"]:G["back"] + 1] = g
        ions = w["groups"][-1]
        gid[ions["front"]:ions["back"] + 1] = np.arange(ions["front"], ions["back"] + 1) + 10**6
        nbad = overlaps(xq, w["radius"], gid, L)
        print(f"[{name}] configuration seed {wseed}: {nbad} contacts", flush=True)
        if nbad == 0:
            break
    assert nbad == 0
    ids = w["types"].astype(np.uint8)
    # per-particle radius: sites carry their own (the state file holds it);
    # ids fix charges.  Radii on a grid of 2^-7 A above 1.5 A (rounded down,
    # so no contact appears), stored as uint8 steps
    rk = np.floor((np.asarray(w["radius"]) - 1.5) * 128.0).astype(np.int64)
    assert rk.min() >= 0 and rk.max() <= 255
    tq = np.array([0.0, -1.0, 0.0, 1.0, 1.0, -1.0])
    cfg = {
        "system": {"temperature": 298.15, "geometry": {"length": L}},
        "atomlist": w["ref_atoms"],
        "moleculelist": w["ref_molecules"],
        "moves": w["ref_moves"](True),
        "energy": {"nonbonded": dict(w["ref_nonbonded"])},
    }
    groups = [{"front": G["front"], "back": G["back"], "molid": G["molid"],
               "molsize": G["molsize"]} for G in w["groups"]]
    moves = [{"type": "TR", "molid": 0, "prob": 1.0, "peratom": 0, "permol": 1,
              "dp": 5.0, "dprot": 0.5, "center_rotation": 1}]
    radius = 1.5 + rk.astype(np.float64) / 128.0
    n = len(ids)
    with tempfile.TemporaryDirectory() as td:
        cpath = os.path.join(td, "in.json")
        spath = os.path.join(td, "state")
        with open(cpath, "w") as f:
            json.dump(cfg, f, indent=1)
        write_state(spath, [L, L, L], xq, tq[ids], radius, np.ones(n), ids, groups,
                    mt_state_text(seed))
        print(f"[{name}] energies", flush=True)
        en = harness(["energy", "dhhs", cpath, spath])
        print(f"[{name}] Metropolis run, 1 call", flush=True)
        mc = harness(["mc", "dhhs", cpath, spath, "1"])
    fx = {
        "name": name, "kind": "dhhs", "box": [L, L, L], "config": cfg,
        "particles": {"xyz_q16": enc(q16, "<i2"), "unit": unit, "id": enc(ids, "u1"),
                      "charge_by_id": list(map(float, tq)),
                      "radius_steps": enc(rk, "u1"), "radius_base": 1.5, "radius_step": 1.0 / 128,
                      "mw_by_id": [1.0] * 6},
        "groups": groups, "moves": moves, "rng_seed": seed,
    }
    ng = len(groups)
    fx["energy"] = energy_subset(en, ng, [0, 1, 2, 3, ng - 1])
    du = np.array([fnum(v) for v in mc["du"]])
    fxyz = np.array(mc["final_xyz"]).reshape(-1, 3)
    sites = np.concatenate([np.arange(G["front"], G["front"] + 10) for G in groups[:-1]])
    fx["mc"] = {"ncalls": 1, "trials": len(du), "uinit": mc["uinit"], "ufinal": mc["ufinal"],
                "utot": mc["utot"], "du": enc(du, "<f8"), "acc": enc_bits(mc["acc"]),
                "who": enc(mc["who"], "<i4"), "final_cm": enc(mc["final_cm"], "<f8"),
                "sites": enc(sites, "<i4"), "final_sites_xyz": enc(fxyz[sites], "<f8"),
                "final_rng": mc["final_rng"]}
    write(name, fx)


def main(names):
    if not os.path.exists(HARNESS):
        sys.exit("build oracle/_ref/ref_harness first: make -C oracle ref")
    todo = {"c1": c1, "c2": c2, "c3": c3}
    for k in (names or todo):
        todo[k]()


if __name__ == "__main__":
    main(sys.argv[1:])
