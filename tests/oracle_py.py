"""ctypes view of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: it is the checker the HIP engine is
compared against and is never imported by the product package.
"""
import ctypes as C
import json
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")

ELEC = {"none": 0, "coulomb": 1, "dh": 2, "ewald": 3}
SR = {"none": 0, "hs": 1, "lj": 2}
NB = {"plain": 0, "vector": 1, "cutg2g": 2}

# fixture kind -> (elec, short-range, Nonbonded flavour)
KINDS = {
    "dhhs": ("dh", "hs", "plain"),
    "dhhs_vec": ("dh", "hs", "vector"),
    "dhhs_cut": ("dh", "hs", "cutg2g"),
    "coulhs": ("coulomb", "hs", "plain"),
    "dh": ("dh", "none", "plain"),
    "lj": ("none", "lj", "plain"),
    "ewaldlj": ("ewald", "lj", "vector"),
    # LJ as Potential::CutShift<LennardJonesLB> at energy.nonbonded.cutoff
    "ljcut": ("none", "lj", "plain"),
    "ewaldljcut": ("ewald", "lj", "vector"),
    # Nonbonded<DebyeHuckelHS> + ExternalPressure, Isobaric move
    "dhhs_npt": ("dh", "hs", "plain"),
    # NonbondedEwald<LennardJonesLB> + ExternalPressure, Isobaric move (then
    # AtomicTranslation)
    "ewaldlj_npt": ("ewald", "lj", "vector"),
}


def lj_cut_of(kind, nonbonded):
    """CutShift cutoff of a fixture kind (0 = uncut LJ)."""
    return float(nonbonded["cutoff"]) if kind.endswith("ljcut") else 0.0


class OrParams(C.Structure):
    _fields_ = [
        ("elec", C.c_int), ("sr", C.c_int), ("nbkind", C.c_int),
        ("lB", C.c_double), ("kappa", C.c_double), ("g2g_cut", C.c_double),
        ("ntypes", C.c_int),
        ("sigma", C.POINTER(C.c_double)), ("eps", C.POINTER(C.c_double)),
        ("tcharge", C.POINTER(C.c_double)), ("tradius", C.POINTER(C.c_double)),
        ("tdp", C.POINTER(C.c_double)),
        ("alpha", C.c_double), ("rc", C.c_double), ("kc", C.c_double),
        ("debyelength", C.c_double), ("eps_surf", C.c_double),
        ("spherical_sum", C.c_int), ("update_frequency", C.c_int),
        ("lj_cut", C.c_double),
        ("has_pressure", C.c_int), ("pressure", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        # ORACLE_LIB names another build of oracle.c (e.g. one with a planted
        # mistake, to show which tests catch it)
        path = os.environ.get("ORACLE_LIB") or os.path.join(ORACLE_DIR, "liboracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-C", ORACLE_DIR, "oracle"], check=True, capture_output=True)
        L = C.CDLL(path)
        D, I, P = C.c_double, C.c_int, C.c_void_p
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.or_create.restype = P
        L.or_create.argtypes = [C.POINTER(OrParams), dp, I, dp, dp, dp, dp, ip, I, ip, ip, ip, ip]
        L.or_destroy.argtypes = [P]
        for fn in ("or_i2i", "or_pair"):
            getattr(L, fn).restype = D
            getattr(L, fn).argtypes = [P, I, I, I]
        L.or_i2g.restype = D; L.or_i2g.argtypes = [P, I, I, I]
        L.or_i2all.restype = D; L.or_i2all.argtypes = [P, I, I]
        L.or_g2g.restype = D; L.or_g2g.argtypes = [P, I, I, I]
        L.or_g_internal.restype = D; L.or_g_internal.argtypes = [P, I, I]
        L.or_g_external.restype = D; L.or_g_external.argtypes = [P, I, I]
        L.or_external.restype = D; L.or_external.argtypes = [P, I]
        L.or_system_energy.restype = D; L.or_system_energy.argtypes = [P, I]
        L.or_energy_change.restype = D; L.or_energy_change.argtypes = [P, I, ip, ip, ip]
        L.or_commit.argtypes = [P, I, ip, ip, ip, I]
        L.or_get_xyz.argtypes = [P, I, dp]
        L.or_set_pos.argtypes = [P, I, I, dp]
        L.or_get_cm.argtypes = [P, I, dp, dp]
        L.or_ewald_nk.restype = I; L.or_ewald_nk.argtypes = [P]
        L.or_ewald_get.argtypes = [P, I, dp, dp, dp, dp]
        L.or_ewald_update_change.restype = D; L.or_ewald_update_change.argtypes = [P, I, ip, ip, ip]
        L.or_ewald_update.restype = D; L.or_ewald_update.argtypes = [P, I]
        L.or_isobaric_propagate.restype = C.c_long
        L.or_isobaric_propagate.argtypes = [P, D, D, I, dp, ip, C.c_long, dp]
        L.or_propagate_npt.restype = C.c_long
        L.or_propagate_npt.argtypes = [P, D, D, I, I, dp, ip, C.c_long, dp]
        L.or_get_box.argtypes = [P, dp]
        for fn in ("or_ewald_self", "or_ewald_surface", "or_ewald_reciprocal"):
            getattr(L, fn).restype = D
            getattr(L, fn).argtypes = [P, I]
        L.or_ewald_set_space.argtypes = [P]
        L.or_add_move.restype = I
        L.or_add_move.argtypes = [P, I, I, D, I, I, dp, dp, D, D, I]
        L.or_rng_seed.argtypes = [P, C.c_uint32]
        L.or_rng_set_text.restype = I; L.or_rng_set_text.argtypes = [P, C.c_char_p]
        L.or_propagator_init.argtypes = [P]
        L.or_propagator_pick.restype = I; L.or_propagator_pick.argtypes = [P, I]
        L.or_propagator_peek.restype = I; L.or_propagator_peek.argtypes = [P, I]
        L.or_propagate.restype = C.c_long
        L.or_propagate.argtypes = [P, I, dp, ip, C.c_long, dp, dp]
        L.or_last_trials.restype = C.c_long
        L.or_last_trials.argtypes = [P, ip, C.c_long]
        L.or_rng_draws.argtypes = [C.c_uint32, I, I, I, I, dp]
        L.or_pt_exchange_seeded.argtypes = [I, dp, ip, C.c_uint32, C.c_uint32, I, ip, dp]
        L.or_pt_exchange.restype = I
        L.or_pt_exchange.argtypes = [I, dp, ip, P, P, ip, dp]
        L.or_mt_new.restype = P
        L.or_mt_new.argtypes = [C.c_uint32]
        L.or_mt_free.argtypes = [P]
        L.or_pt_step.restype = I
        L.or_pt_step.argtypes = [I, C.POINTER(P), P, D, ip, dp, dp]
        L.or_rng_draws.restype = None
        _lib = L
    return _lib


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def random_reseeds_move_engine(cfg):
    """True when the input's moves.random object reseeds the Propagator's move
    engine (move.h:7783-7787): a default-seeded mt19937 unless "hardware",
    which no fixture uses"""
    r = cfg.get("moves", {}).get("random")
    return isinstance(r, dict) and not r.get("hardware", False)


def init_propagator(seed_fn, init_fn, fx):
    """Propagator construction for a fixture: the move engine copies the
    global one (move.h:7781), or, with moves.random, starts from the default
    seed while the global engine keeps the fixture's seed"""
    seed_fn(fx["rng_seed"])
    init_fn()
    if random_reseeds_move_engine(fx["config"]):
        seed_fn(5489)
        init_fn()
        seed_fn(fx["rng_seed"])


def load_fixture(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        return json.load(f)


def fnum(v):
    """fixture number: floats, or the strings inf / -inf / nan"""
    if isinstance(v, str):
        return float(v)
    return float(v)


def farr(v):
    return np.array([fnum(x) for x in v], dtype=np.float64)


class OracleSystem:
    """One oracle Space + Hamiltonian built from explicit arrays."""

    def __init__(self, kind, box, xyz, charge, radius, mw, ids, groups, lB, kappa=0.0,
                 type_sigma=None, type_eps=None, type_charge=None, type_radius=None,
                 type_dp=None, g2g_cut=0.0, ewald=None, lj_cut=0.0, pressure=None):
        L = lib()
        elec, sr, nb = KINDS[kind]
        self._keep = []
        T = len(type_sigma) if type_sigma is not None else 1

        def arr(v, n=T):
            a = np.ascontiguousarray(np.asarray(v if v is not None else np.zeros(n), dtype=np.float64))
            self._keep.append(a)
            return a

        prm = OrParams()
        prm.elec, prm.sr, prm.nbkind = ELEC[elec], SR[sr], NB[nb]
        prm.lB, prm.kappa, prm.g2g_cut = lB, kappa, g2g_cut
        prm.lj_cut = lj_cut
        if pressure is not None:
            prm.has_pressure, prm.pressure = 1, pressure
        prm.ntypes = T
        prm.sigma = dptr(arr(type_sigma))
        prm.eps = dptr(arr(type_eps))
        prm.tcharge = dptr(arr(type_charge))
        prm.tradius = dptr(arr(type_radius))
        prm.tdp = dptr(arr(type_dp))
        if ewald:
            prm.alpha = ewald["alpha"]
            prm.rc = ewald["cutoff"]
            prm.kc = ewald["cutoffK"]
            prm.debyelength = ewald["debyelength"]
            prm.eps_surf = ewald["eps_surf"]
            prm.spherical_sum = int(ewald.get("spherical_sum", True))
            prm.update_frequency = int(ewald.get("update_frequency", -1))
        self.prm = prm
        n = len(charge)
        self.n = n
        b = arr(box, 3)
        x = arr(np.asarray(xyz).reshape(-1), 3 * n)
        q, r, m = arr(charge, n), arr(radius, n), arr(mw, n)
        idv = np.ascontiguousarray(np.asarray(ids, dtype=np.int32)); self._keep.append(idv)
        gf = np.array([g["front"] for g in groups], dtype=np.int32)
        gb = np.array([g["back"] for g in groups], dtype=np.int32)
        gs = np.array([g["molsize"] for g in groups], dtype=np.int32)
        gm = np.array([g["molid"] for g in groups], dtype=np.int32)
        self._keep += [gf, gb, gs, gm]
        self.ng = len(groups)
        self.h = L.or_create(C.byref(prm), dptr(b), n, dptr(x), dptr(q), dptr(r), dptr(m), iptr(idv),
                             self.ng, iptr(gf), iptr(gb), iptr(gs), iptr(gm))

    @classmethod
    def from_fixture(cls, fx, lB=None):
        e = fx.get("energy") or fx.get("mc")
        p = fx["particles"]
        nb = fx["config"]["energy"]["nonbonded"]
        kappa = 1.0 / nb["debyelength"] if "debyelength" in nb else 0.0
        sys_ = cls(fx["kind"], fx["box"], p["xyz"], p["charge"], p["radius"], p["mw"], p["id"],
                   fx["groups"], e["lB"] if lB is None else lB, kappa=kappa,
                   type_sigma=e["type_sigma"], type_eps=e["type_eps"],
                   type_charge=e["type_charge"], type_radius=e["type_radius"],
                   type_dp=e["type_dp"], g2g_cut=nb.get("cutoff_g2g", 0.0),
                   ewald=nb.get("ewald") if KINDS[fx["kind"]][0] == "ewald" else None,
                   lj_cut=lj_cut_of(fx["kind"], nb), pressure=e.get("pressure"))
        sys_.add_moves(fx)
        return sys_

    def add_moves(self, fx):
        L = lib()
        for mv in fx["moves"]:
            if mv["type"] == "NPT":   # driven by isobaric_propagate
                continue
            typ = 1 if mv["type"] == "AT" else 2
            L.or_add_move(self.h, typ, mv["molid"], mv.get("prob", 1.0), mv.get("peratom", 0),
                          mv.get("permol", 0), None, None, mv.get("dp", 0.0), mv.get("dprot", 0.0),
                          mv.get("center_rotation", 1))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_destroy(self.h)
            self.h = None

    # energies
    def i2i(self, i, j, which=0): return lib().or_i2i(self.h, which, i, j)
    def i2g(self, g, j, which=0): return lib().or_i2g(self.h, which, g, j)
    def i2all(self, i, which=0): return lib().or_i2all(self.h, which, i)
    def g2g(self, a, b, which=0): return lib().or_g2g(self.h, which, a, b)
    def g_internal(self, g, which=0): return lib().or_g_internal(self.h, which, g)
    def g_external(self, g, which=0): return lib().or_g_external(self.h, which, g)
    def external(self, which=0): return lib().or_external(self.h, which)
    def system_energy(self, which=0): return lib().or_system_energy(self.h, which)

    def xyz(self, which=0):
        out = np.zeros(3 * self.n)
        lib().or_get_xyz(self.h, which, dptr(out))
        return out.reshape(-1, 3)

    def set_pos(self, i, xyz, which=1):
        a = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64))
        lib().or_set_pos(self.h, which, i, dptr(a))

    def isobaric_propagate(self, dp, prob, ncalls, cap=1 << 16):
        du = np.zeros(cap)
        acc = np.zeros(cap, dtype=np.int32)
        ut = C.c_double()
        n = lib().or_isobaric_propagate(self.h, dp, prob, ncalls, dptr(du), iptr(acc), cap,
                                        C.byref(ut))
        if n < 0:
            raise ValueError("oracle: isobaric move not supported for this system")
        return du[:n], acc[:n], ut.value

    def propagate_npt(self, dp, prob, npt_index, ncalls, cap=1 << 16):
        """Propagator::move calls of a Propagator holding the registered
        moves and one Isobaric move at position npt_index; the trial record
        includes who (-1 for volume trials)."""
        L = lib()
        du = np.zeros(cap)
        acc = np.zeros(cap, dtype=np.int32)
        who = np.full(cap, -1, dtype=np.int32)
        L.or_last_trials(self.h, iptr(who), cap)
        ut = C.c_double()
        n = L.or_propagate_npt(self.h, dp, prob, npt_index, ncalls, dptr(du), iptr(acc), cap,
                               C.byref(ut))
        L.or_last_trials(self.h, None, 0)
        self._who = who
        if n < 0:
            raise ValueError("oracle: isobaric move not supported for this system")
        return du[:n], acc[:n], who[:n], ut.value

    def box(self):
        b = np.zeros(3)
        lib().or_get_box(self.h, dptr(b))
        return b

    def energy_change(self, mv):
        """mv: list of (group, [indices]) pairs, ordered by group"""
        g = np.array([m[0] for m in mv], dtype=np.int32)
        off = np.zeros(len(mv) + 1, dtype=np.int32)
        idx = []
        for k, m in enumerate(mv):
            idx += list(m[1])
            off[k + 1] = len(idx)
        ia = np.array(idx + [0], dtype=np.int32)
        return lib().or_energy_change(self.h, len(mv), iptr(g), iptr(off), iptr(ia))

    def commit(self, mv, accepted):
        g = np.array([m[0] for m in mv], dtype=np.int32)
        off = np.zeros(len(mv) + 1, dtype=np.int32)
        idx = []
        for k, m in enumerate(mv):
            idx += list(m[1])
            off[k + 1] = len(idx)
        ia = np.array(idx + [0], dtype=np.int32)
        lib().or_commit(self.h, len(mv), iptr(g), iptr(off), iptr(ia), int(accepted))

    def seed(self, s): lib().or_rng_seed(self.h, s)

    def propagate(self, ncalls, cap=1 << 20):
        L = lib()
        du = np.zeros(cap)
        acc = np.zeros(cap, dtype=np.int32)
        who = np.zeros(cap, dtype=np.int32)
        L.or_last_trials(self.h, iptr(who), cap)
        ut, ui = C.c_double(), C.c_double()
        n = L.or_propagate(self.h, ncalls, dptr(du), iptr(acc), cap, C.byref(ut), C.byref(ui))
        L.or_last_trials(self.h, None, 0)
        self._who = who
        return du[:n], acc[:n], who[:n], ut.value, ui.value


class OracleEngine:
    """An or_mt engine (libstdc++ mt19937) owned by Python."""

    def __init__(self, seed=5489):
        self.h = lib().or_mt_new(seed)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_mt_free(self.h)
            self.h = None


def pt_step(ranks, ctl, prob=1.0):
    """One ParallelTempering move on every rank (or_pt_step): returns
    (acc, du, ret) per rank."""
    R = len(ranks)
    arr = (C.c_void_p * R)(*[r.h for r in ranks])
    acc = np.zeros(R, dtype=np.int32)
    du = np.zeros(R)
    ret = np.zeros(R)
    rc = lib().or_pt_step(R, arr, ctl.h, prob, iptr(acc), dptr(du), dptr(ret))
    if rc < 0:
        raise ValueError("or_pt_step refused the ranks (%d)" % rc)
    return acc, du, ret
