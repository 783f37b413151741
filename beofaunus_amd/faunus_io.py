"""Faunus input and state files on the engine.

* ``read_state`` / ``write_state``: the text state of ``Space::load`` /
  ``Space::save`` (space.h:736-764, 807-880): box lengths, particle count, one
  ``PointParticle`` per line (x y z charge radius mw id hydrophobic,
  point.h:544-567), group count, one ``Group`` per line (front back molId
  molsize cm_x cm_y cm_z, group.h:388-406), then ``randomstate`` and the text
  state of the global mt19937 engine (slump.h:24-126).
* ``read_input``: the keys of the reference's JSON input that this path reads
  (system.temperature, atomlist, moleculelist, energy.nonbonded, moves) and a
  state file, turned into the workload dictionary ``workloads.make_system``
  builds an engine system from.  The pair-potential classes are compile-time
  template arguments in the reference, so the caller names them (``elec``,
  ``sr``, ``nbkind`` and whether LJ is CutShift), as a Faunus program does in
  its typedefs.
* ``run_input``: a reference input run on the device for a number of
  Propagator::move() calls, its final state written back in the same format
  (with the global engine's state, as Space::save does), so a run can be
  resumed the way the reference resumes one.

Atom types are numbered as the reference's AtomMap numbers them: the dummy
type 0 ("unk", species.h:494-504), then the atomlist entries in the order the
JSON library iterates them (std::map: sorted by name).  Molecules likewise
(MoleculeMap, molecule.h:545-552, no dummy entry).
"""
import json
import math

import numpy as np

from . import fgx

# physconst.h:13-25
PC_E, PC_E0, PC_KB, PC_NAV = 1.602177e-19, 8.85419e-12, 1.380658e-23, 6.022137e23


def bjerrum_length(epsr, T):
    """pc::lB(e_r) (physconst.h:91-94), A"""
    return PC_E * PC_E / (4 * math.acos(-1.0) * PC_E0 * epsr * 1e-10 * PC_KB * T)


def kjmol_to_kt(u, T):
    """operator "" _kJmol (physconst.h:101-104)"""
    return u / (PC_KB * T) / PC_NAV * 1e3


# --------------------------------------------------------------------------
# Space state file
# --------------------------------------------------------------------------
def read_state(path):
    """Space::load for a Cuboid (space.h:807-880).  Returns a dict with box,
    xyz (N x 3), charge, radius, mw, id, hydrophobic, groups (front, back,
    molid, molsize, cm) and rng (the engine's text state, or None)."""
    with open(path) as f:
        tok = f.read().split()
    pos = 0

    def take(n=1):
        nonlocal pos
        v = tok[pos:pos + n]
        if len(v) < n:
            raise ValueError("%s: state file ends early" % path)
        pos += n
        return v

    box = [float(v) for v in take(3)]
    n = int(take()[0])
    rows = np.array(take(8 * n), dtype=object).reshape(n, 8) if n else np.zeros((0, 8))
    xyz = np.array(rows[:, :3], dtype=np.float64)
    charge = np.array(rows[:, 3], dtype=np.float64)
    radius = np.array(rows[:, 4], dtype=np.float64)
    mw = np.array(rows[:, 5], dtype=np.float64)
    ids = np.array([int(v) for v in rows[:, 6]], dtype=np.int32)
    hyd = np.array([int(v) for v in rows[:, 7]], dtype=np.int32)
    ng = int(take()[0])
    groups = []
    for _ in range(ng):
        g = take(7)
        front, back = int(g[0]), int(g[1])
        if back - front + 1 < 0 or back >= n:
            raise ValueError("%s: group range %d..%d outside the particles" % (path, front, back))
        groups.append({"front": front, "back": back, "molid": int(g[2]), "molsize": int(g[3]),
                       "cm": [float(v) for v in g[4:7]]})
    rng = None
    if pos < len(tok) and tok[pos] == "randomstate":
        pos += 1
        words = tok[pos:pos + 625]
        if len(words) == 624:
            # older files hold the 624 words only; the position then stays
            # where a freshly seeded engine keeps it
            words = words + ["624"]
        if len(words) != 625:
            raise ValueError("%s: random state needs 624 words and a position" % path)
        rng = " ".join(words)
    return {"box": box, "xyz": xyz, "charge": charge, "radius": radius, "mw": mw, "id": ids,
            "hydrophobic": hyd, "groups": groups, "rng": rng}


def write_state(path, box, xyz, charge, radius, mw, ids, groups, rng=None, hydrophobic=None):
    """Space::save for a Cuboid (space.h:736-764).  Doubles are written with
    round-trip precision; `rng` is the engine's text state (624 words and the
    position) or None."""
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = len(xyz)
    hyd = np.zeros(n, dtype=np.int32) if hydrophobic is None else hydrophobic
    out = ["%r %r %r" % tuple(float(v) for v in box), "%d" % n]
    for i in range(n):
        out.append("%r %r %r %r %r %r %d %d" % (float(xyz[i, 0]), float(xyz[i, 1]),
                                                float(xyz[i, 2]), float(charge[i]),
                                                float(radius[i]), float(mw[i]), int(ids[i]),
                                                int(hyd[i])))
    out.append("%d" % len(groups))
    for g in groups:
        cm = g.get("cm", [0.0, 0.0, 0.0])
        out.append("%d %d %d %d %r %r %r" % (g["front"], g["back"], g["molid"], g["molsize"],
                                             float(cm[0]), float(cm[1]), float(cm[2])))
    if rng is not None:
        out.append("randomstate")
        out.append(rng)
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def mt_text(words, pos):
    """libstdc++ text form of an mt19937 state (624 words, then the position)"""
    return " ".join(str(int(w)) for w in words) + " %d" % int(pos)


def mt_parse(text):
    v = [int(x) for x in text.split()]
    if len(v) != 625:
        raise ValueError("mt19937 state needs 624 words and a position")
    return v[:624], v[624]


# --------------------------------------------------------------------------
# JSON input
# --------------------------------------------------------------------------
def atom_table(atomlist, T):
    """AtomMap from the atomlist section (species.h:279-417): names and
    per-type charge, radius, sigma, eps (kT), dp, mw, type 0 the dummy."""
    names = ["unk"] + sorted(atomlist.keys())
    nt = len(names)
    q, r, sig, eps, dp, mw = (np.zeros(nt) for _ in range(6))
    mw[0] = 1.0
    for k, name in enumerate(names[1:], start=1):
        a = atomlist[name]
        q[k] = float(a.get("q", 0.0))
        rad = float(a.get("r", 0.0))
        sig[k] = float(a["sigma"]) if "sigma" in a else 2 * rad
        r[k] = 0.5 * sig[k]
        eps[k] = kjmol_to_kt(float(a.get("eps", 0.0)), T)
        dp[k] = float(a.get("dp", 0.0))
        mw[k] = float(a.get("mw", 1.0))
    return names, {"q": q, "r": r, "sigma": sig, "eps": eps, "dp": dp, "mw": mw}


def _vec3(v, default="1 1 1"):
    s = v if v is not None else default
    if isinstance(s, (list, tuple)):
        return [float(x) for x in s]
    return [float(x) for x in str(s).split()]


def read_moves(moves, molnames):
    """Propagator's move registry for the keys this path runs
    (move.h:7789-7843): atomtranslate (move.h:784-794) and moltransrot
    (move.h:1578-1600) with their per-molecule MolListData (move.h:324-332).
    Returns (moves on the path, names of the other registered moves)."""
    out, other = [], []
    molid = {name: k for k, name in enumerate(molnames)}
    for key in sorted(moves.keys()):
        m = moves[key]
        if key == "atomtranslate":
            for mol in sorted(k for k in m.keys() if isinstance(m[k], dict)):
                d = m[mol]
                out.append({"type": "AT", "molid": molid[mol], "prob": float(d.get("prob", 1.0)),
                            "peratom": int(bool(d.get("peratom", False))),
                            "permol": int(bool(d.get("permol", False))),
                            "dir": _vec3(d.get("dir"))})
        elif key == "moltransrot":
            if m.get("center_rotation") is None:
                raise ValueError("moltransrot needs center_rotation (move.h:1589)")
            for mol in sorted(k for k in m.keys() if isinstance(m[k], dict)):
                d = m[mol]
                out.append({"type": "TR", "molid": molid[mol], "prob": float(d.get("prob", 1.0)),
                            "peratom": int(bool(d.get("peratom", False))),
                            "permol": int(bool(d.get("permol", False))),
                            "dp": float(d.get("dp", 0.0)), "dprot": float(d.get("dprot", 0.0)),
                            "dir": _vec3(d.get("dir")), "dir2": _vec3(d.get("dir2")),
                            "center_rotation": int(bool(m["center_rotation"]))})
        elif key == "isobaric":
            # Move::Isobaric (move.h:5455-5472): dp, pressure (mM), prob
            out.append({"type": "NPT", "molid": -1, "prob": float(m.get("prob", 1.0)),
                        "peratom": 0, "permol": 0, "dp": float(m["dp"]),
                        "pressure_mM": float(m["pressure"])})
        elif key not in ("random", "_jsonfile"):   # moves.random: the engine (below)
            other.append(key)
    return out, other


def move_engine_random(moves_cfg):
    """moves.random (move.h:7783-7787): None without it; else whether it asks
    for a hardware seed.  Without one the Propagator's move engine starts
    from mt19937's default seed instead of a copy of the global engine (the
    supported move keys all sort before "random", so no later key copies the
    global engine back)."""
    r = moves_cfg.get("random")
    if not isinstance(r, dict):
        return None
    return {"hardware": bool(r.get("hardware", False))}


def mM_to_per_A3(c):
    """operator "" _mM (physconst.h:170-173): particles per A^3"""
    return c * (1.0e-3 * PC_NAV) / 1e27


def read_input(json_path, state_path, *, elec, sr, nbkind=fgx.NB_PLAIN, precision=fgx.FP64,
               lj_cutshift=False, external_pressure=False):
    """A reference JSON input plus a state file as a workload dictionary.
    elec/sr/nbkind name the compile-time pair potential (FGX_ELEC_*, FGX_SR_*,
    FGX_NB_*); lj_cutshift makes LJ CutShift<LennardJonesLB> at
    energy.nonbonded.cutoff; external_pressure adds Energy::ExternalPressure
    at moves.isobaric.pressure (energy.h:1867-1873)."""
    with open(json_path) as f:
        cfg = json.load(f)
    T = float(cfg.get("system", {}).get("temperature", 298.15))
    names, at = atom_table(cfg["atomlist"], T)
    molnames = sorted(cfg.get("moleculelist", {}).keys())
    st = read_state(state_path)
    if len(st["id"]) and int(st["id"].max()) >= len(names):
        raise ValueError("state file has more species than the atom list (space.h:840)")
    nb = cfg.get("energy", {}).get("nonbonded", {})
    lB = bjerrum_length(float(nb.get("epsr", 80.0)), T)
    kappa = 0.0
    if elec == fgx.ELEC_DH:
        # Potential::DebyeHuckel (src/potentials.cpp:770-782)
        I = float(nb.get("ionicstrength", 0.0))
        c = 8 * lB * math.acos(-1.0) * PC_NAV / 1e27
        k = math.sqrt(I * c)
        kappa = 1.0 / float(nb["debyelength"]) if k < 1e-10 else k
    g2g_cut = 0.0
    if nbkind == fgx.NB_CUTG2G and "cutoff_g2g" in nb:
        g2g_cut = float(nb["cutoff_g2g"])
    ewald = None
    if elec == fgx.ELEC_EWALD:
        e = nb["ewald"]
        ewald = {"alpha": float(e["alpha"]), "cutoff": float(e["cutoff"]),
                 "cutoffK": float(e["cutoffK"]), "debyelength": float(e["debyelength"]),
                 "eps_surf": float(e["eps_surf"]),
                 "spherical_sum": bool(e.get("spherical_sum", True)),
                 "update_frequency": int(e.get("update_frequency", -1))}
    lj_cut = float(nb["cutoff"]) if lj_cutshift else 0.0
    pressure = (mM_to_per_A3(float(cfg["moves"]["isobaric"]["pressure"]))
                if external_pressure else None)
    moves, other = read_moves(cfg.get("moves", {}), molnames)
    # the device Propagator runs one registered move, alone or beside one
    # isobaric move; with any other move registered (even at prob 0, it takes
    # part in Propagator's pick) the system is built for energies only
    move = None
    path = [m for m in moves if m["type"] != "NPT"]
    keys = sorted(k for k in cfg.get("moves", {}) if k in ("atomtranslate", "isobaric",
                                                            "moltransrot"))
    npt_index = keys.index("isobaric") if "isobaric" in keys else None
    if len(path) == 1 and not other:
        mv = path[0]
        move = {"kind": fgx.MOVE_AT if mv["type"] == "AT" else fgx.MOVE_TR,
                "molid": mv["molid"], "prob": mv["prob"], "peratom": mv["peratom"],
                "permol": mv["permol"], "dir": mv["dir"]}
        if mv["type"] == "TR":
            move.update(dp=mv["dp"], dprot=mv["dprot"], dir2=mv["dir2"],
                        center_rotation=mv["center_rotation"])
    return {
        "name": json_path,
        "box": st["box"], "xyz": st["xyz"], "charge": st["charge"], "radius": st["radius"],
        "mw": st["mw"], "types": st["id"],
        "groups": [{k: g[k] for k in ("front", "back", "molid", "molsize")} for g in st["groups"]],
        "ham": dict(elec=elec, sr=sr, nbkind=nbkind, precision=precision, lB=lB, kappa=kappa,
                    g2g_cut=g2g_cut, type_sigma=at["sigma"], type_eps=at["eps"],
                    type_charge=at["q"], type_radius=at["r"], type_dp=at["dp"], ewald=ewald,
                    lj_cut=lj_cut, pressure=pressure),
        "move": move,
        "moves": moves,
        # position of the isobaric move in the Propagator's move list (its
        # keys in sorted order), None without one
        "npt_index": npt_index,
        "other_moves": other,
        "move_random": move_engine_random(cfg.get("moves", {})),
        "atom_names": names,
        "molecule_names": molnames,
        "rng": st["rng"],
        "T": T,
    }


def run_input(json_path, state_path, out_state, ncalls, *, elec, sr, nbkind=fgx.NB_PLAIN,
              precision=fgx.FP64, lj_cutshift=False, external_pressure=False):
    """Load a reference input and state, seed the global engine from the
    state file (as Space::load does), construct the Propagator (its engine
    copied from the global one, move.h:7781), run `ncalls` Propagator::move()
    calls on the device (its registered moves: atomtranslate or moltransrot,
    isobaric, or one of the first two with isobaric) and write the final
    state (box, positions, mass centres, the global engine) to `out_state`.
    Returns the run statistics."""
    from . import workloads
    w = read_input(json_path, state_path, elec=elec, sr=sr, nbkind=nbkind,
                   precision=precision, lj_cutshift=lj_cutshift,
                   external_pressure=external_pressure)
    npt_only = (len(w["moves"]) == 1 and not w["other_moves"]
                and w["moves"][0]["type"] == "NPT")
    if w["move"] is None and not npt_only:
        raise ValueError("%s: the device Propagator needs one registered move (atomtranslate "
                         "or moltransrot for one molecule), optionally with isobaric; found %s "
                         "and %s" % (json_path, [m["type"] for m in w["moves"]],
                                     w["other_moves"]))
    iso = [m for m in w["moves"] if m["type"] == "NPT"]
    s = workloads.make_system(w)
    try:
        if w["rng"] is not None:
            words, pos = mt_parse(w["rng"])
            s.rng_set_state(words, pos)
        s.propagator_init()
        if w["move_random"] is not None:
            if w["move_random"]["hardware"]:
                raise ValueError("%s: moves.random.hardware asks for a non-deterministic seed "
                                 "(slump.h:50-57); a replayable device run needs a fixed one"
                                 % json_path)
            # the move engine from mt19937's default seed, the global engine
            # as the state file left it
            gw, gp = s.rng_state()
            s.rng_seed(5489)
            s.propagator_init()
            s.rng_set_state(gw, gp)
        if npt_only:   # a Propagator holding one Isobaric move
            st = s.isobaric_propagate(iso[0]["dp"], iso[0]["prob"], ncalls)[0]
        elif iso:      # the registered move and an Isobaric move
            st = s.propagate_npt(iso[0]["dp"], iso[0]["prob"], w["npt_index"], ncalls)[0]
        else:
            st = s.propagate(ncalls)[0]
        xyz = s.positions(0)
        words, pos = s.rng_state()
        cms = [s.group_cm(g)[0] for g in range(len(w["groups"]))]
        groups = [dict(g, cm=list(cm)) for g, cm in zip(w["groups"], cms)]
        write_state(out_state, list(s.box()), xyz, w["charge"], w["radius"], w["mw"],
                    w["types"], groups, rng=mt_text(words, pos))
    finally:
        s.close()
    return st
