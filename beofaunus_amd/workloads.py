"""Seeded synthetic systems for the Metropolis Delta-U path.

Each builder returns a plain dict describing one Space<Cuboid,PointParticle>
(box, particles, groups), its Hamiltonian (fgx.System keyword arguments) and
its registered move, so the same system can be built on the GPU
(`make_system`) and written in the reference's own input formats
(`write_reference_inputs`: JSON config + Space::save state file) for a CPU
run of the reference beside it.

Unit conventions follow the reference: eps in the JSON is kJ/mol and is
converted to kT at the system temperature (species.h:293); lB follows from
epsr (physconst.h:91-94).
"""
import json
import math
import os

import numpy as np

from . import fgx

T_DEFAULT = 298.15
_E, _E0, _KB, _NAV = 1.602177e-19, 8.85419e-12, 1.380658e-23, 6.022137e23


def bjerrum(epsr, T=T_DEFAULT):
    """pc::lB (physconst.h:91-94)"""
    return _E * _E / (4 * math.acos(-1.0) * _E0 * epsr * 1e-10 * _KB * T)


def epsr_for(lB, T=T_DEFAULT):
    return bjerrum(1.0, T) / lB


def kjmol_to_kt(u, T=T_DEFAULT):
    """operator "" _kJmol (physconst.h:101-104)"""
    return u / (_KB * T) / _NAV * 1e3


def kt_to_kjmol(u, T=T_DEFAULT):
    return u * (_KB * T) * _NAV / 1e3


def rsa(rng, n, L, radius, max_tries=50_000_000):
    """Uniform random positions in a periodic cube, rejecting any overlap
    (distance < 2 * radius), via a cell grid."""
    d = 2.0 * radius
    nc = max(1, int(L // d))
    cs = L / nc
    cells = {}
    out = np.zeros((n, 3))
    k = 0
    tries = 0
    while k < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("RSA did not converge")
        p = (rng.random(3) - 0.5) * L
        c = tuple(((p + 0.5 * L) // cs).astype(int) % nc)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in cells.get(((c[0] + dx) % nc, (c[1] + dy) % nc, (c[2] + dz) % nc), ()):
                        r = out[j] - p
                        r -= L * np.round(r / L)
                        if r @ r < d * d:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            out[k] = p
            cells.setdefault(c, []).append(k)
            k += 1
    return out


def jittered_lattice(rng, n, L, jitter):
    m = int(math.ceil(n ** (1.0 / 3.0)))
    a = L / m
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1)
    g = g.reshape(-1, 3)
    pick = rng.choice(len(g), size=n, replace=False)
    p = (g[pick] + 0.5) * a - 0.5 * L + (rng.random((n, 3)) - 0.5) * 2 * jitter
    return p - L * np.round(p / L)


def salt_dh(n=500, L=60.0, lB=7.1, debye=10.0, radius=2.0, dp=2.0, seed=1, T=T_DEFAULT,
            precision=fgx.FP64, lattice=False):
    """Plain NVT salt solution: +-1 ions, hard spheres, Debye-Hueckel."""
    rng = np.random.default_rng(seed)
    xyz = (jittered_lattice(rng, n, L, 0.5 * (L / math.ceil(n ** (1 / 3)) - 2 * radius) * 0.99)
           if lattice else rsa(rng, n, L, radius))
    half = n // 2
    types = np.array([1] * half + [2] * (n - half), dtype=np.int32)
    charge = np.where(types == 1, 1.0, -1.0)
    return {
        "name": "salt_dh",
        "box": [L, L, L], "xyz": xyz, "charge": charge, "radius": np.full(n, radius),
        "mw": np.ones(n), "types": types,
        "groups": [{"front": 0, "back": n - 1, "molsize": 1, "molid": 0}],
        "ham": dict(elec=fgx.ELEC_DH, sr=fgx.SR_HS, nbkind=fgx.NB_PLAIN, precision=precision,
                    lB=lB, kappa=1.0 / debye,
                    type_sigma=[0.0, 2 * radius, 2 * radius], type_eps=[0.0, 0.0, 0.0],
                    type_charge=[0.0, 1.0, -1.0], type_radius=[0.0, radius, radius],
                    type_dp=[0.0, dp, dp]),
        "move": dict(kind=fgx.MOVE_AT, molid=0, peratom=True),
        "ref_kind": "dhhs",
        "ref_nonbonded": {"epsr": epsr_for(lB, T), "debyelength": debye},
        "ref_atoms": {"T1p": {"q": 1.0, "r": radius, "dp": dp, "mw": 1.0},
                      "T2m": {"q": -1.0, "r": radius, "dp": dp, "mw": 1.0}},
        "T": T,
    }


def ewald_electrolyte(n=20000, L=120.0, lB=7.1, sigma=4.0, eps_kt=0.1, rc=12.0, alpha=0.25,
                      kc=10.0, dp=1.0, update_frequency=20000, seed=2, T=T_DEFAULT,
                      precision=fgx.MIXED, lj_cut=False):
    """Ewald LJ + Coulomb electrolyte (NonbondedEwald<LennardJonesLB>): +-1
    charges 50/50, Lorentz-Berthelot LJ, screened Ewald with kappa -> 0, tinfoil
    boundary, spherical k-space cutoff.  lj_cut=True makes the LJ term
    CutShift<LennardJonesLB> at rc, so every pair term vanishes beyond rc
    (the cell-list form; the reference's NonbondedEwald cannot take a cut LJ,
    so this Hamiltonian has no reference run)."""
    rng = np.random.default_rng(seed)
    xyz = jittered_lattice(rng, n, L, 0.3)
    types = rng.permutation(np.array([1] * (n // 2) + [2] * (n - n // 2), dtype=np.int32))
    charge = np.where(types == 1, 1.0, -1.0)
    eps_json = kt_to_kjmol(eps_kt, T)
    eps = kjmol_to_kt(eps_json, T)
    r = 0.5 * sigma
    ewald = {"alpha": alpha, "cutoff": rc, "cutoffK": kc, "debyelength": 1e12,
             "eps_surf": 0.5, "spherical_sum": True, "update_frequency": update_frequency}
    return {
        "name": "ewald_electrolyte",
        "box": [L, L, L], "xyz": xyz, "charge": charge, "radius": np.full(n, r),
        "mw": np.ones(n), "types": types,
        "groups": [{"front": 0, "back": n - 1, "molsize": 1, "molid": 0}],
        "ham": dict(elec=fgx.ELEC_EWALD, sr=fgx.SR_LJ, nbkind=fgx.NB_VECTOR, precision=precision,
                    lB=lB, type_sigma=[0.0, sigma, sigma], type_eps=[0.0, eps, eps],
                    type_charge=[0.0, 1.0, -1.0], type_radius=[0.0, r, r],
                    type_dp=[0.0, dp, dp], ewald=ewald, lj_cut=rc if lj_cut else 0.0),
        "move": dict(kind=fgx.MOVE_AT, molid=0, peratom=True),
        "ref_kind": "ewaldljcut" if lj_cut else "ewaldlj",
        "ref_nonbonded": {"epsr": epsr_for(lB, T), "ewald": ewald},
        "ref_atoms": {"T1p": {"q": 1.0, "sigma": sigma, "eps": eps_json, "dp": dp, "mw": 1.0},
                      "T2m": {"q": -1.0, "sigma": sigma, "eps": eps_json, "dp": dp, "mw": 1.0}},
        "T": T,
    }


def pt_ladders(nladder=6, rungs=8, nmin=1000, nmax=16000, tmin=280.0, tmax=420.0,
               lB_ref=7.1, T_ref=T_DEFAULT, rho=500.0 / 60.0 ** 3, seed=4,
               precision=fgx.MIXED):
    """Batched parallel tempering: `nladder` ladders of `rungs` replicas.
    Each ladder has one N (log-uniform in [nmin, nmax], seeded) at the salt
    density of the 500-ion / 60 A system, and geometric temperatures
    T_r = tmin (tmax/tmin)^(r/(rungs-1)).  The solvent permittivity is fixed,
    so lB_r = lB_ref T_ref / T_r (physconst.h:91-94); Debye-Hueckel 10 A, hard
    spheres 2 A, AtomicTranslation dp 2 A.  Exchanges stay inside a ladder
    (the reference swaps configurations of equal N, move.h:6683-6687)."""
    rng = np.random.default_rng(seed)
    out = []
    for l in range(nladder):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        n += n % 2
        L = (n / rho) ** (1.0 / 3.0)
        reps = []
        for r in range(rungs):
            T = tmin * (tmax / tmin) ** (r / (rungs - 1))
            w = salt_dh(n=n, L=L, lB=lB_ref * T_ref / T, debye=10.0, radius=2.0, dp=2.0,
                        seed=seed * 1000 + 10 * l + r, T=T, precision=precision)
            reps.append(w)
        out.append({"n": n, "L": L, "T": [w["T"] for w in reps], "replicas": reps})
    return out


def serve_jobs(njobs=256, nmin=500, nmax=50000, rho=0.01, seed=5, precision=fgx.MIXED):
    """Served job stream: N log-uniform in [nmin, nmax] at number density rho;
    even-numbered draws are Ewald LJ+Coulomb electrolytes (the N=20000
    electrolyte's parameters), the others Debye-Hueckel + hard-sphere salt
    (the 500-ion salt's parameters).  Starting configurations are jittered
    lattices (hard-sphere free), seeded per job."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(njobs):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        n += n % 2
        L = (n / rho) ** (1.0 / 3.0)
        if k % 2 == 0:
            w = ewald_electrolyte(n=n, L=L, seed=seed * 10007 + k, precision=precision)
        else:
            w = salt_dh(n=n, L=L, seed=seed * 10007 + k, lattice=True, precision=precision)
        out.append({"n": n, "L": L, "kind": "ewald" if k % 2 == 0 else "dh", "w": w})
    return out


def rigid_molecules(nmol=64, nsite=400, nion=8000, L=250.0, lB=7.1, debye=15.0, R=20.0,
                    dp=5.0, dprot=0.5, seed=3, T=T_DEFAULT, precision=fgx.MIXED):
    """Rigid molecules (site charges from {-1,0,+1} with p = 0.2/0.6/0.2,
    radii U[1.5,3] A, sites packed in a sphere of radius R) plus monovalent
    ions, Debye-Hueckel + hard spheres, TranslateRotate moves."""
    rng = np.random.default_rng(seed)
    # molecule centres far enough apart that molecules start separated
    centres = rsa(rng, nmol, L, R + 3.0)
    xyz, charge, radius, types, groups = [], [], [], [], []
    for m in range(nmol):
        front = len(xyz)
        v = rng.normal(size=(nsite, 3))
        v *= (R * rng.random(nsite) ** (1 / 3) / np.linalg.norm(v, axis=1))[:, None]
        p = centres[m] + v
        p -= L * np.round(p / L)
        q = rng.choice([-1.0, 0.0, 1.0], size=nsite, p=[0.2, 0.6, 0.2])
        rr = 1.5 + 1.5 * rng.random(nsite)
        xyz += list(p)
        charge += list(q)
        radius += list(rr)
        types += [1 if qq < 0 else (2 if qq == 0 else 3) for qq in q]
        groups.append({"front": front, "back": len(xyz) - 1, "molsize": nsite, "molid": 0})
    # ions: random sequential addition clear of every site (hard-sphere
    # contact ri + rj) and of each other
    ions = _place_clear(rng, nion, L, 2.0, np.array(xyz), np.array(radius))
    front = len(xyz)
    for k in range(nion):
        xyz.append(ions[k])
        charge.append(1.0 if k % 2 == 0 else -1.0)
        radius.append(2.0)
        types.append(4 if k % 2 == 0 else 5)
    if nion:
        groups.append({"front": front, "back": len(xyz) - 1, "molsize": 1, "molid": 1})
    n = len(xyz)
    return {
        "name": "rigid_molecules",
        "box": [L, L, L], "xyz": np.array(xyz), "charge": np.array(charge),
        "radius": np.array(radius), "mw": np.ones(n), "types": np.array(types, dtype=np.int32),
        "groups": groups,
        "ham": dict(elec=fgx.ELEC_DH, sr=fgx.SR_HS, nbkind=fgx.NB_PLAIN, precision=precision,
                    lB=lB, kappa=1.0 / debye,
                    type_sigma=[0.0, 4.0, 4.0, 4.0, 4.0, 4.0], type_eps=[0.0] * 6,
                    type_charge=[0.0, -1.0, 0.0, 1.0, 1.0, -1.0],
                    type_radius=[0.0, 2.0, 2.0, 2.0, 2.0, 2.0],
                    type_dp=[0.0, 0.0, 0.0, 0.0, 2.0, 2.0]),
        "move": dict(kind=fgx.MOVE_TR, molid=0, permol=True, dp=dp, dprot=dprot,
                     center_rotation=True),
        "ref_kind": "dhhs",
        "ref_nonbonded": {"epsr": epsr_for(lB, T), "debyelength": debye},
        # atom ids follow sorted names: 1 = site -1, 2 = site 0, 3 = site +1,
        # 4 = cation, 5 = anion; per-particle charge and radius come from the
        # state file
        "ref_atoms": {"M1n": {"q": -1.0, "r": 2.0, "mw": 1.0},
                      "M2z": {"q": 0.0, "r": 2.0, "mw": 1.0},
                      "M3p": {"q": 1.0, "r": 2.0, "mw": 1.0},
                      "M4ip": {"q": 1.0, "r": 2.0, "dp": 2.0, "mw": 1.0},
                      "M5im": {"q": -1.0, "r": 2.0, "dp": 2.0, "mw": 1.0}},
        # molecule ids follow key order from 0: "amol" = 0, "bions" = 1
        "ref_molecules": {"amol": {"atoms": " ".join(["M2z"] * nsite), "Ninit": 0},
                          "bions": {"atoms": "M4ip M5im", "atomic": True, "Ninit": 0}},
        "ref_moves": lambda per: {"moltransrot": {"center_rotation": True,
                                                  "amol": {"dp": dp, "dprot": dprot,
                                                           "permol": bool(per)}}},
        "T": T,
    }


def _place_clear(rng, n, L, r_new, placed_xyz, placed_r, max_tries=10_000_000):
    """n points of radius r_new placed uniformly at random in the periodic
    cube, each clear of the placed spheres and of the points before it."""
    from collections import defaultdict
    rmax = max(float(np.max(placed_r)) if len(placed_r) else 0.0, r_new)
    cs = r_new + rmax
    nc = max(1, int(L // cs))
    cs = L / nc
    cell = defaultdict(list)

    def key(x):
        return tuple((((x + 0.5 * L) // cs).astype(int)) % nc)

    for x, r in zip(placed_xyz, placed_r):
        cell[key(x)].append((x, r))
    out = []
    tries = 0
    while len(out) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place %d points" % n)
        x = (rng.random(3) - 0.5) * L
        c = key(x)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for (y, r) in cell.get(((c[0] + dx) % nc, (c[1] + dy) % nc, (c[2] + dz) % nc), ()):
                        d = y - x
                        d -= L * np.round(d / L)
                        if d @ d < (r + r_new) ** 2:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
        if ok:
            out.append(x)
            cell[c].append((x, r_new))
    return np.array(out)


def _separate(p, L, dmin, rng):
    """nudge points apart until no pair is closer than dmin (few iterations)"""
    from collections import defaultdict
    for _ in range(50):
        nc = max(1, int(L // dmin))
        cs = L / nc
        cell = defaultdict(list)
        idx = (((p + 0.5 * L) // cs).astype(int)) % nc
        for i, c in enumerate(map(tuple, idx)):
            cell[c].append(i)
        bad = set()
        for i, c in enumerate(map(tuple, idx)):
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        for j in cell.get(((c[0] + dx) % nc, (c[1] + dy) % nc, (c[2] + dz) % nc), ()):
                            if j <= i:
                                continue
                            r = p[j] - p[i]
                            r -= L * np.round(r / L)
                            if r @ r < dmin * dmin:
                                bad.add(j)
        if not bad:
            return p
        for j in bad:
            p[j] = p[j] + (rng.random(3) - 0.5) * dmin
            p[j] -= L * np.round(p[j] / L)
    return p


def make_system(w, precision=None):
    ham = dict(w["ham"])
    if precision is not None:
        ham["precision"] = precision
    s = fgx.System(w["box"], w["xyz"], w["charge"], w["radius"], w["mw"], w["types"],
                   w["groups"], **ham)
    if w.get("move") is None:   # energies only
        return s
    mv = dict(w["move"])
    kind = mv.pop("kind")
    molid = mv.pop("molid")
    s.add_move(kind, molid, **mv)
    return s


def mt_state_text(seed):
    x = [0] * 624
    x[0] = seed & 0xFFFFFFFF
    for i in range(1, 624):
        x[i] = (1812433253 * (x[i - 1] ^ (x[i - 1] >> 30)) + i) & 0xFFFFFFFF
    return " ".join(str(v) for v in x) + " 624"


def write_reference_inputs(w, directory, seed=1, peratom=True):
    """Write the reference's JSON config and Space::save state file
    (space.h:736-764) for a workload, so the reference (or its CPU
    restatement) can run the identical system.  `peratom` selects a sweep
    (peratom / permol) or a single trial per move() call."""
    os.makedirs(directory, exist_ok=True)
    L = w["box"]
    cfg = {
        "system": {"temperature": w.get("T", T_DEFAULT), "geometry": {"length": L[0]}},
        "atomlist": w["ref_atoms"],
        "moleculelist": w.get("ref_molecules") or
                        {"salt": {"atoms": " ".join(w["ref_atoms"].keys()), "atomic": True,
                                  "Ninit": 0}},
        "moves": (w["ref_moves"](peratom) if "ref_moves" in w else
                  {"atomtranslate": {"salt": {"peratom": bool(peratom)}}}),
        "energy": {"nonbonded": w["ref_nonbonded"]},
    }
    cpath = os.path.join(directory, "config.json")
    spath = os.path.join(directory, "state")
    with open(cpath, "w") as f:
        json.dump(cfg, f, indent=1)
    with open(spath, "w") as f:
        n = len(w["charge"])
        f.write("%r %r %r\n%d\n" % (float(L[0]), float(L[1]), float(L[2]), n))
        for i in range(n):
            x = w["xyz"][i]
            f.write("%r %r %r %r %r %r %d 0\n" % (float(x[0]), float(x[1]), float(x[2]),
                                                 float(w["charge"][i]), float(w["radius"][i]),
                                                 float(w["mw"][i]), int(w["types"][i])))
        f.write("%d\n" % len(w["groups"]))
        for g in w["groups"]:
            f.write("%d %d %d %d 0 0 0\n" % (g["front"], g["back"], g["molid"], g["molsize"]))
        f.write("randomstate\n" + mt_state_text(seed) + "\n")
    return cpath, spath
