"""ctypes binding of libfgx.so, the MI355X Metropolis Delta-U engine.

This is the thin host-side mirror of the C ABI declared in include/fgx.h.
It never computes energies itself: every number comes from the HIP library,
and importing it on a machine without the built library raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libfgx.so")

ELEC_NONE, ELEC_COULOMB, ELEC_DH, ELEC_EWALD = 0, 1, 2, 3
SR_NONE, SR_HS, SR_LJ = 0, 1, 2
NB_PLAIN, NB_VECTOR, NB_CUTG2G = 0, 1, 2
FP64, MIXED = 0, 1
MOVE_AT, MOVE_TR = 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class FgxError(RuntimeError):
    pass


class Hamiltonian(C.Structure):
    _fields_ = [
        ("elec", C.c_int), ("sr", C.c_int), ("nbkind", C.c_int), ("precision", C.c_int),
        ("lB", C.c_double), ("kappa", C.c_double), ("g2g_cut", C.c_double),
        ("ntypes", C.c_int),
        ("type_sigma", _dp), ("type_eps", _dp), ("type_charge", _dp),
        ("type_radius", _dp), ("type_dp", _dp),
        ("ewald_alpha", C.c_double), ("ewald_rc", C.c_double), ("ewald_kc", C.c_double),
        ("ewald_debyelength", C.c_double), ("ewald_eps_surf", C.c_double),
        ("ewald_spherical_sum", C.c_int), ("ewald_update_frequency", C.c_int),
        ("lj_cut", C.c_double),
        ("has_pressure", C.c_int), ("pressure", C.c_double),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("trials", C.c_long), ("accepted", C.c_long), ("utot", C.c_double),
        ("dusum", C.c_double), ("rejected_inf", C.c_long), ("ewald_refresh", C.c_long),
        ("seconds", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Job(C.Structure):
    _fields_ = [
        ("h", C.POINTER(Hamiltonian)), ("box", C.c_double * 3), ("n", C.c_int),
        ("xyz", _dp), ("charge", _dp), ("radius", _dp), ("mw", _dp), ("type", _ip),
        ("ng", C.c_int), ("front", _ip), ("back", _ip), ("molsize", _ip), ("molid", _ip),
        ("move_type", C.c_int), ("move_molid", C.c_int), ("peratom", C.c_int),
        ("permol", C.c_int), ("center_rotation", C.c_int),
        ("prob", C.c_double), ("dp", C.c_double), ("dprot", C.c_double),
        ("dir", _dp), ("dir2", _dp), ("seed", C.c_uint32), ("sweeps", C.c_int),
        ("keep_positions", C.c_int),
    ]


class JobResult(C.Structure):
    _fields_ = [
        ("status", C.c_int), ("sweeps", C.c_int), ("trials", C.c_long), ("accepted", C.c_long),
        ("utot", C.c_double), ("queue_seconds", C.c_double), ("latency_seconds", C.c_double),
        ("mean_sweep_seconds", C.c_double), ("max_sweep_seconds", C.c_double),
        ("device_bytes", C.c_size_t), ("workgroups", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Change(C.Structure):
    _fields_ = [("group", C.c_int), ("count", C.c_int), ("index", _ip)]


_lib = None


def lib():
    """Load libfgx.so; raise if it has not been built (no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FgxError("libfgx.so is not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    P, I, D = C.c_void_p, C.c_int, C.c_double
    pp = C.POINTER(C.c_void_p)
    sig = {
        "fgx_last_error": (C.c_char_p, []),
        "fgx_device_info": (I, [C.c_char_p, I, _ip, C.POINTER(C.c_size_t)]),
        "fgx_create": (I, [C.POINTER(Hamiltonian), _dp, I, _dp, _dp, _dp, _dp, _ip, I, _ip, _ip,
                           _ip, _ip, pp]),
        "fgx_destroy": (I, [P]),
        "fgx_device_bytes": (C.c_size_t, [P]),
        "fgx_set_positions": (I, [P, I, I, I, _dp]),
        "fgx_get_positions": (I, [P, I, _dp]),
        "fgx_get_group_cm": (I, [P, I, _dp, _dp]),
        "fgx_set_group_cm": (I, [P, I, _dp, _dp]),
        "fgx_i2i": (I, [P, I, I, I, _dp]),
        "fgx_i2g": (I, [P, I, I, I, _dp]),
        "fgx_i2all": (I, [P, I, I, _dp]),
        "fgx_g2g": (I, [P, I, I, I, _dp]),
        "fgx_g_internal": (I, [P, I, I, _dp]),
        "fgx_g_external": (I, [P, I, I, _dp]),
        "fgx_external": (I, [P, I, _dp]),
        "fgx_system_energy": (I, [P, I, _dp]),
        "fgx_system_energy_noalloc": (I, [P, I, _dp]),
        "fgx_update_change": (I, [P, C.POINTER(Change)]),
        "fgx_energy_change": (I, [P, C.POINTER(Change), _dp]),
        "fgx_commit": (I, [P, C.POINTER(Change), I, _dp]),
        "fgx_add_move": (I, [P, I, I, D, I, I, _dp, _dp, D, D, I]),
        "fgx_rng_seed": (I, [P, C.c_uint32]),
        "fgx_rng_set_state": (I, [P, C.POINTER(C.c_uint32), I]),
        "fgx_rng_get_state": (I, [P, C.POINTER(C.c_uint32), _ip]),
        "fgx_rng_set_internal": (I, [P, C.POINTER(C.c_uint32), I]),
        "fgx_rng_get_internal": (I, [P, C.POINTER(C.c_uint32), _ip]),
        "fgx_propagator_init": (I, [P]),
        "fgx_propagate": (I, [P, I, C.POINTER(Stats), _dp, _ip, C.c_long]),
        "fgx_propagate_batch": (I, [pp, I, I, C.POINTER(Stats)]),
        "fgx_propagate_batch_record": (I, [pp, I, I, C.POINTER(Stats), _dp, _ip, C.c_long,
                                           C.POINTER(C.c_long)]),
        "fgx_set_launch": (I, [P, I, I]),
        "fgx_set_energy_matrix": (I, [P, I]),
        "fgx_set_cell_list": (I, [P, I]),
        "fgx_set_lds_partners": (I, [P, I]),
        "fgx_set_block_sweep": (I, [P, I, I]),
        "fgx_get_block_parts": (I, [P, C.POINTER(C.c_int)]),
        "fgx_set_ewald_state": (I, [P, I]),
        "fgx_isobaric_propagate": (I, [P, D, D, I, C.POINTER(Stats), _dp, _ip, C.c_long]),
        "fgx_propagate_npt": (I, [P, D, D, I, I, C.POINTER(Stats), _dp, _ip, C.c_long]),
        "fgx_set_space": (I, [P, _dp]),
        "fgx_i_external": (I, [P, I, I, _dp]),
        "fgx_config_energy": (I, [P, I, _dp]),
        "fgx_commit_all": (I, [P, I, _dp]),
        "fgx_get_box": (I, [P, _dp]),
        "fgx_ewald_k": (I, [P, _ip]),
        "fgx_get_cell_list": (I, [P, _ip, _ip]),
        "fgx_chain_seconds": (I, [P, C.POINTER(C.c_double)]),
        "fgx_rdf": (I, [P, I, I, I, C.c_double, I, C.POINTER(C.c_ulonglong)]),
        "fgx_server_create": (I, [C.c_size_t, I, C.c_double, pp]),
        "fgx_server_submit": (I, [P, C.POINTER(Job), _ip]),
        "fgx_server_step": (I, [P, _ip, _ip]),
        "fgx_server_poll": (I, [P, _ip, I, _ip]),
        "fgx_server_result": (I, [P, I, C.POINTER(JobResult), _dp]),
        "fgx_server_info": (I, [P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t), _ip, _ip]),
        "fgx_server_destroy": (I, [P]),
        "fgx_get_launch": (I, [P, _ip, _ip]),
        "fgx_autotune": (I, [P, I, _ip, _ip, _dp]),
        "fgx_autotune_batch": (I, [pp, I, I, _ip, _ip, _dp]),
        "fgx_temper_create": (I, [pp, I, C.c_uint32, D, pp]),
        "fgx_temper_run": (I, [pp, I, I, I, C.POINTER(Stats), _dp]),
        "fgx_temper_records": (I, [P, _ip, _ip, _dp, _ip, _dp]),
        "fgx_temper_exchange": (I, [P, _ip, _dp]),
        "fgx_temper_slots": (I, [P, _ip]),
        "fgx_temper_returns": (I, [P, _dp]),
        "fgx_temper_destroy": (I, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        msg = lib().fgx_last_error()
        raise FgxError("fgx error %d: %s" % (rc, msg.decode() if msg else ""))


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def device_info():
    name = C.create_string_buffer(128)
    cu = C.c_int()
    mem = C.c_size_t()
    _check(lib().fgx_device_info(name, 128, C.byref(cu), C.byref(mem)))
    return {"arch": name.value.decode(), "cus": cu.value, "hbm_bytes": mem.value}


def make_hamiltonian(arr, ntypes, *, elec, sr, nbkind=NB_PLAIN, precision=FP64, lB=7.0,
                     kappa=0.0, g2g_cut=0.0, type_sigma=None, type_eps=None, type_charge=None,
                     type_radius=None, type_dp=None, ewald=None, lj_cut=0.0, pressure=None):
    """fgx_hamiltonian from keyword parameters; `arr(values, n)` must keep the
    per-type arrays alive for as long as the structure is used.  `lj_cut` > 0
    makes the LJ term CutShift<LennardJonesLB> at that cutoff."""
    h = Hamiltonian()
    h.elec, h.sr, h.nbkind, h.precision = elec, sr, nbkind, precision
    h.lB, h.kappa, h.g2g_cut = lB, kappa, g2g_cut
    h.lj_cut = lj_cut or 0.0
    if pressure is not None:   # Energy::ExternalPressure, 1/A^3
        h.has_pressure, h.pressure = 1, float(pressure)
    h.ntypes = ntypes
    h.type_sigma = _d(arr(type_sigma, ntypes))
    h.type_eps = _d(arr(type_eps, ntypes))
    h.type_charge = _d(arr(type_charge, ntypes))
    h.type_radius = _d(arr(type_radius, ntypes))
    h.type_dp = _d(arr(type_dp, ntypes))
    if ewald:
        h.ewald_alpha = ewald["alpha"]
        h.ewald_rc = ewald["cutoff"]
        h.ewald_kc = ewald["cutoffK"]
        h.ewald_debyelength = ewald["debyelength"]
        h.ewald_eps_surf = ewald["eps_surf"]
        h.ewald_spherical_sum = int(ewald.get("spherical_sum", True))
        h.ewald_update_frequency = int(ewald.get("update_frequency", -1))
    return h


class System:
    """One Space<Cuboid,PointParticle> + non-bonded Hamiltonian on the GPU."""

    def __init__(self, box, xyz, charge, radius, mw, types, groups, *, elec, sr,
                 nbkind=NB_PLAIN, precision=FP64, lB=7.0, kappa=0.0, g2g_cut=0.0,
                 type_sigma=None, type_eps=None, type_charge=None, type_radius=None,
                 type_dp=None, ewald=None, lj_cut=0.0, pressure=None):
        L = lib()
        n = len(charge)
        ntypes = len(type_sigma) if type_sigma is not None else int(np.max(types)) + 1
        self._keep = []

        def arr(v, default_len):
            a = np.ascontiguousarray(
                np.asarray(v if v is not None else np.zeros(default_len), dtype=np.float64))
            self._keep.append(a)
            return a

        h = make_hamiltonian(arr, ntypes, elec=elec, sr=sr, nbkind=nbkind, precision=precision,
                             lB=lB, kappa=kappa, g2g_cut=g2g_cut, type_sigma=type_sigma,
                             type_eps=type_eps, type_charge=type_charge,
                             type_radius=type_radius, type_dp=type_dp, ewald=ewald,
                             lj_cut=lj_cut, pressure=pressure)
        self.ham = h
        self.n = n
        self.box_len = [float(v) for v in box]
        self.groups = [dict(g) for g in groups]
        self.ng = len(groups)
        b = arr(box, 3)
        x = arr(np.asarray(xyz, dtype=np.float64).reshape(-1), 3 * n)
        q, r, m = arr(charge, n), arr(radius, n), arr(mw, n)
        t = np.ascontiguousarray(np.asarray(types, dtype=np.int32))
        gf = np.array([g["front"] for g in groups], dtype=np.int32)
        gb = np.array([g["back"] for g in groups], dtype=np.int32)
        gs = np.array([g["molsize"] for g in groups], dtype=np.int32)
        gm = np.array([g["molid"] for g in groups], dtype=np.int32)
        self._keep += [t, gf, gb, gs, gm]
        hp = C.c_void_p()
        _check(L.fgx_create(C.byref(h), _d(b), n, _d(x), _d(q), _d(r), _d(m), _i(t), self.ng,
                            _i(gf), _i(gb), _i(gs), _i(gm), C.byref(hp)))
        self.h = hp

    def close(self):
        if getattr(self, "h", None):
            lib().fgx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- Energybase entry points
    def _e(self, fn, *args):
        u = C.c_double()
        _check(getattr(lib(), fn)(self.h, *args, C.byref(u)))
        return u.value

    def i2i(self, i, j, which=0): return self._e("fgx_i2i", which, i, j)
    def i2g(self, g, j, which=0): return self._e("fgx_i2g", which, g, j)
    def i2all(self, i, which=0): return self._e("fgx_i2all", which, i)
    def g2g(self, a, b, which=0): return self._e("fgx_g2g", which, a, b)
    def g_internal(self, g, which=0): return self._e("fgx_g_internal", which, g)
    def g_external(self, g, which=0): return self._e("fgx_g_external", which, g)
    def i_external(self, i, which=0): return self._e("fgx_i_external", which, i)
    def external(self, which=0): return self._e("fgx_external", which)
    def system_energy(self, which=0): return self._e("fgx_system_energy", which)
    def system_energy_noalloc(self, which=0): return self._e("fgx_system_energy_noalloc", which)

    @property
    def device_bytes(self):
        return lib().fgx_device_bytes(self.h)

    # ---- transfers
    def positions(self, which=0):
        out = np.zeros(3 * self.n)
        _check(lib().fgx_get_positions(self.h, which, _d(out)))
        return out.reshape(-1, 3)

    def set_positions(self, first, xyz, which=1):
        a = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1))
        _check(lib().fgx_set_positions(self.h, which, first, len(a) // 3, _d(a)))

    def group_cm(self, g):
        cm, ct = np.zeros(3), np.zeros(3)
        _check(lib().fgx_get_group_cm(self.h, g, _d(cm), _d(ct)))
        return cm, ct

    def set_group_cm(self, g, cm=None, cm_trial=None):
        a = np.ascontiguousarray(np.asarray(cm, dtype=np.float64)) if cm is not None else None
        b = np.ascontiguousarray(np.asarray(cm_trial, dtype=np.float64)) if cm_trial is not None else None
        _check(lib().fgx_set_group_cm(self.h, g, _d(a) if a is not None else None,
                                      _d(b) if b is not None else None))

    # ---- Delta U of one change
    def _change(self, group, index):
        idx = np.ascontiguousarray(np.asarray(index if index is not None else [], dtype=np.int32))
        ch = Change(group, len(idx), _i(idx) if len(idx) else None)
        return ch, idx

    def update_change(self, group, index=None):
        ch, idx = self._change(group, index)
        _check(lib().fgx_update_change(self.h, C.byref(ch)))

    def energy_change(self, group, index=None):
        ch, idx = self._change(group, index)
        du = C.c_double()
        _check(lib().fgx_energy_change(self.h, C.byref(ch), C.byref(du)))
        return du.value

    def commit(self, group, index, accepted):
        ch, idx = self._change(group, index)
        d = C.c_double()
        _check(lib().fgx_commit(self.h, C.byref(ch), int(accepted), C.byref(d)))
        return d.value

    # ---- device Propagator
    def add_move(self, kind, molid, prob=1.0, peratom=False, permol=False, dir=None, dir2=None,
                 dp=0.0, dprot=0.0, center_rotation=True):
        a = np.ascontiguousarray(np.asarray(dir if dir is not None else [1, 1, 1], dtype=np.float64))
        b = np.ascontiguousarray(np.asarray(dir2 if dir2 is not None else [1, 1, 1], dtype=np.float64))
        _check(lib().fgx_add_move(self.h, kind, molid, prob, int(peratom), int(permol), _d(a),
                                  _d(b), dp, dprot, int(center_rotation)))

    def set_launch(self, threads=0, unroll=0):
        _check(lib().fgx_set_launch(self.h, threads, unroll))

    def rdf_counts(self, type_a, type_b, dr, which=0):
        """Pair counts per bin int(r/dr + 0.5) (Analysis::AtomRDF sampling)."""
        L = np.asarray(self.box_len, dtype=np.float64)
        nb = int(0.5 * np.sqrt((L * L).sum()) / dr + 0.5) + 2
        out = np.zeros(nb, dtype=np.uint64)
        _check(lib().fgx_rdf(self.h, which, type_a, type_b, dr, nb,
                             out.ctypes.data_as(C.POINTER(C.c_ulonglong))))
        return out

    def chain_seconds(self):
        """Wall time this chain spent in the last device launch."""
        v = C.c_double()
        _check(lib().fgx_chain_seconds(self.h, C.byref(v)))
        return v.value

    def set_energy_matrix(self, on):
        """Group-pair energy cache of the mixed TranslateRotate sweep."""
        _check(lib().fgx_set_energy_matrix(self.h, 1 if on else 0))

    def isobaric_propagate(self, dp, prob, ncalls, record=0):
        """Propagator::move() calls of a Propagator holding one Isobaric move
        (volume trials; needs the ExternalPressure term)."""
        st = Stats()
        if record:
            du = np.zeros(record)
            acc = np.zeros(record, dtype=np.int32)
            _check(lib().fgx_isobaric_propagate(self.h, dp, prob, ncalls, C.byref(st), _d(du),
                                                _i(acc), record))
            n = min(record, st.trials)
            return st.as_dict(), du[:n], acc[:n]
        _check(lib().fgx_isobaric_propagate(self.h, dp, prob, ncalls, C.byref(st), None, None, 0))
        return st.as_dict(), None, None

    def propagate_npt(self, dp, prob, npt_index, ncalls, record=0):
        """Propagator::move() calls of a Propagator holding the registered
        move and one Isobaric move at position npt_index of its move list
        (the input's move keys in sorted order)."""
        st = Stats()
        if record:
            du = np.zeros(record)
            acc = np.zeros(record, dtype=np.int32)
            _check(lib().fgx_propagate_npt(self.h, dp, prob, npt_index, ncalls, C.byref(st),
                                           _d(du), _i(acc), record))
            n = min(record, st.trials)
            return st.as_dict(), du[:n], acc[:n]
        _check(lib().fgx_propagate_npt(self.h, dp, prob, npt_index, ncalls, C.byref(st), None,
                                       None, 0))
        return st.as_dict(), None, None

    def box(self):
        b = np.zeros(3)
        _check(lib().fgx_get_box(self.h, _d(b)))
        return b

    def ewald_k(self):
        k = np.zeros(1, dtype=np.int32)
        _check(lib().fgx_ewald_k(self.h, _i(k)))
        return int(k[0])

    def set_cell_list(self, mode):
        """Cell list of the mixed AtomicTranslation sweep: 0 off, 1 automatic,
        n >= 2 records per cell."""
        _check(lib().fgx_set_cell_list(self.h, int(mode)))

    def set_ewald_state(self, hbm):
        """Ewald sweep: S(k), dS(k), A_k and n in LDS when they fit (False,
        default) or in device memory (True)."""
        _check(lib().fgx_set_ewald_state(self.h, 1 if hbm else 0))

    def set_lds_partners(self, on):
        """Mixed AtomicTranslation sweep without Ewald: keep the partners of a
        chain of up to 2,048 particles in LDS (True, default) or read them from
        device memory (False); the same pair arithmetic, summed in another
        order."""
        _check(lib().fgx_set_lds_partners(self.h, 1 if on else 0))

    def set_block_sweep(self, on, nsplit=0):
        """Block AtomicTranslation sweep: B trials per pass over the partners,
        decided in order with exact corrections (True), one trial at a time
        (False), or "auto" (default: block sweeps except for batches of
        unsplit mixed Ewald chains with partners in device memory).  nsplit
        > 1 spreads a chain over that many workgroups; 0 chooses
        automatically."""
        mode = 2 if on == "auto" else (1 if on else 0)
        _check(lib().fgx_set_block_sweep(self.h, mode, int(nsplit)))

    def block_parts(self):
        """workgroups per chain of the last device sweep if it was a block
        sweep, 0 if it was the one-trial sweep"""
        p = C.c_int()
        _check(lib().fgx_get_block_parts(self.h, C.byref(p)))
        return p.value

    def cell_list(self):
        """(cells per axis, records per cell) of the last sweep; (0, 0) when
        the sweep used no cell list."""
        nc, cap = C.c_int(), C.c_int()
        _check(lib().fgx_get_cell_list(self.h, C.byref(nc), C.byref(cap)))
        return nc.value, cap.value

    def launch(self):
        t, u = C.c_int(), C.c_int()
        _check(lib().fgx_get_launch(self.h, C.byref(t), C.byref(u)))
        return t.value, u.value

    def autotune(self, ncalls=1):
        t, u, sec = C.c_int(), C.c_int(), C.c_double()
        _check(lib().fgx_autotune(self.h, ncalls, C.byref(t), C.byref(u), C.byref(sec)))
        return {"threads": t.value, "unroll": u.value, "seconds": sec.value}

    def rng_seed(self, seed):
        _check(lib().fgx_rng_seed(self.h, seed))

    def rng_set_state(self, words, pos):
        """global engine state (624 words and the position)"""
        w = (C.c_uint32 * 624)(*[int(x) & 0xFFFFFFFF for x in words])
        _check(lib().fgx_rng_set_state(self.h, w, int(pos)))

    def rng_state(self):
        w = (C.c_uint32 * 624)()
        p = C.c_int()
        _check(lib().fgx_rng_get_state(self.h, w, C.byref(p)))
        return list(w), p.value

    def rng_internal_state(self):
        """the move-internal engine (Movebase::_slump): (624 words, position)"""
        w = (C.c_uint32 * 624)()
        p = C.c_int()
        _check(lib().fgx_rng_get_internal(self.h, w, C.byref(p)))
        return list(w), p.value

    def rng_set_internal_state(self, words, pos):
        w = (C.c_uint32 * 624)(*[int(v) & 0xFFFFFFFF for v in words])
        _check(lib().fgx_rng_set_internal(self.h, w, int(pos)))

    def propagator_init(self):
        _check(lib().fgx_propagator_init(self.h))

    def propagate(self, ncalls, record=0):
        st = Stats()
        if record:
            du = np.zeros(record)
            acc = np.zeros(record, dtype=np.int32)
            _check(lib().fgx_propagate(self.h, ncalls, C.byref(st), _d(du), _i(acc), record))
            n = min(record, st.trials)
            return st.as_dict(), du[:n], acc[:n]
        _check(lib().fgx_propagate(self.h, ncalls, C.byref(st), None, None, 0))
        return st.as_dict(), None, None


def autotune_batch(systems, ncalls=1):
    arr = (C.c_void_p * len(systems))(*[s.h for s in systems])
    t, u, sec = C.c_int(), C.c_int(), C.c_double()
    _check(lib().fgx_autotune_batch(arr, len(systems), ncalls, C.byref(t), C.byref(u),
                                    C.byref(sec)))
    return {"threads": t.value, "unroll": u.value, "seconds_per_call": sec.value / ncalls}


def propagate_batch(systems, ncalls):
    arr = (C.c_void_p * len(systems))(*[s.h for s in systems])
    st = Stats()
    _check(lib().fgx_propagate_batch(arr, len(systems), ncalls, C.byref(st)))
    return st.as_dict()


def propagate_batch_record(systems, ncalls, cap):
    """propagate_batch with every chain's per-trial record: returns the
    statistics and lists of (dU, accept) arrays, one per chain."""
    arr = (C.c_void_p * len(systems))(*[s.h for s in systems])
    st = Stats()
    R = len(systems)
    du = np.zeros(R * cap)
    acc = np.zeros(R * cap, dtype=np.int32)
    cnt = (C.c_long * R)()
    _check(lib().fgx_propagate_batch_record(arr, R, ncalls, C.byref(st), _d(du), _i(acc), cap,
                                            cnt))
    dus = [du[c * cap:c * cap + cnt[c]].copy() for c in range(R)]
    accs = [acc[c * cap:c * cap + cnt[c]].copy() for c in range(R)]
    return st.as_dict(), dus, accs


class Temper:
    """Move::ParallelTempering over replicas resident on one GPU
    (fgx_temper_*): rank r is replica r's Bjerrum length, and the rank's
    engines and Ewald counter travel with it on a swap."""

    def __init__(self, replicas, ctl_seed=5489, prob=1.0):
        self.replicas = list(replicas)
        self.R = len(self.replicas)
        arr = (C.c_void_p * self.R)(*[r.h for r in self.replicas])
        h = C.c_void_p()
        _check(lib().fgx_temper_create(arr, self.R, ctl_seed, float(prob), C.byref(h)))
        self.h = h

    def exchange(self):
        """One exchange attempt: (acc, du) per rank (du = +inf without partner)."""
        acc = np.zeros(self.R, dtype=np.int32)
        du = np.zeros(self.R)
        _check(lib().fgx_temper_exchange(self.h, _i(acc), _d(du)))
        return acc, du

    def records(self):
        """Records of the last exchange / run: acc, du, slot (rounds x R) and
        the configurations' sweep seconds per round."""
        n = C.c_int()
        _check(lib().fgx_temper_records(self.h, C.byref(n), None, None, None, None))
        k = n.value
        acc = np.zeros((k, self.R), dtype=np.int32)
        du = np.zeros((k, self.R))
        slot = np.zeros((k, self.R), dtype=np.int32)
        sec = np.zeros((k, self.R))
        _check(lib().fgx_temper_records(self.h, C.byref(n), _i(acc), _d(du), _i(slot), _d(sec)))
        return acc, du, slot, sec

    def slots(self):
        s = np.zeros(self.R, dtype=np.int32)
        _check(lib().fgx_temper_slots(self.h, _i(s)))
        return s

    def returns(self):
        """each rank's ParallelTempering::move return value in the last exchange"""
        r = np.zeros(self.R)
        _check(lib().fgx_temper_returns(self.h, _d(r)))
        return r

    def close(self):
        if getattr(self, "h", None):
            lib().fgx_temper_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def temper_run(tempers, rounds, calls):
    """fgx_temper_run: `rounds` x (`calls` Propagator::move() calls of every
    configuration, then one exchange per temper), as one replayed HIP graph.
    Returns (stats, device seconds)."""
    arr = (C.c_void_p * len(tempers))(*[t.h for t in tempers])
    st = Stats()
    sec = C.c_double()
    _check(lib().fgx_temper_run(arr, len(tempers), int(rounds), int(calls), C.byref(st),
                                C.byref(sec)))
    return st.as_dict(), sec.value


class Server:
    """Job server (fgx_server_*): independent simulations admitted by HBM
    footprint and advanced together in batched launches."""

    def __init__(self, hbm_budget, max_active=4096, tick_seconds=0.05):
        h = C.c_void_p()
        _check(lib().fgx_server_create(int(hbm_budget), int(max_active), float(tick_seconds),
                                       C.byref(h)))
        self.h = h
        self.n = {}

    def submit(self, w, seed, sweeps, precision=None, keep_positions=False):
        """Queue a workload dict (beofaunus_amd.workloads) for `sweeps`
        Propagator::move() calls; returns the job id."""
        keep = []

        def arr(v, n, dtype=np.float64):
            a = np.ascontiguousarray(np.asarray(v if v is not None else np.zeros(n), dtype=dtype))
            keep.append(a)
            return a
        ham = dict(w["ham"])
        if precision is not None:
            ham["precision"] = precision
        ntypes = len(ham["type_sigma"])
        h = make_hamiltonian(arr, ntypes, **ham)
        n = len(w["charge"])
        mv = dict(w["move"])
        j = Job()
        j.h = C.pointer(h)
        j.box = (C.c_double * 3)(*w["box"])
        j.n = n
        j.xyz = _d(arr(np.asarray(w["xyz"]).reshape(-1), 3 * n))
        j.charge = _d(arr(w["charge"], n))
        j.radius = _d(arr(w["radius"], n))
        j.mw = _d(arr(w["mw"], n))
        j.type = _i(arr(w["types"], n, np.int32))
        g = w["groups"]
        j.ng = len(g)
        j.front = _i(arr([x["front"] for x in g], len(g), np.int32))
        j.back = _i(arr([x["back"] for x in g], len(g), np.int32))
        j.molsize = _i(arr([x["molsize"] for x in g], len(g), np.int32))
        j.molid = _i(arr([x["molid"] for x in g], len(g), np.int32))
        j.move_type = mv["kind"]
        j.move_molid = mv["molid"]
        j.peratom = int(mv.get("peratom", False))
        j.permol = int(mv.get("permol", False))
        j.center_rotation = int(mv.get("center_rotation", True))
        j.prob = mv.get("prob", 1.0)
        j.dp = mv.get("dp", 0.0)
        j.dprot = mv.get("dprot", 0.0)
        j.seed = seed
        j.sweeps = sweeps
        j.keep_positions = int(keep_positions)
        jid = C.c_int()
        _check(lib().fgx_server_submit(self.h, C.byref(j), C.byref(jid)))
        self.n[jid.value] = n
        return jid.value

    def step(self):
        a, f = C.c_int(), C.c_int()
        _check(lib().fgx_server_step(self.h, C.byref(a), C.byref(f)))
        return a.value, f.value

    def poll(self, cap=4096):
        ids = np.zeros(cap, dtype=np.int32)
        cnt = C.c_int()
        _check(lib().fgx_server_poll(self.h, _i(ids), cap, C.byref(cnt)))
        return ids[:cnt.value].tolist()

    def result(self, jid, positions=False):
        r = JobResult()
        xyz = np.zeros(3 * self.n[jid]) if positions else None
        _check(lib().fgx_server_result(self.h, jid, C.byref(r), _d(xyz) if positions else None))
        out = r.as_dict()
        if positions:
            out["xyz"] = xyz.reshape(-1, 3)
        return out

    def info(self):
        res, peak = C.c_size_t(), C.c_size_t()
        act, q = C.c_int(), C.c_int()
        _check(lib().fgx_server_info(self.h, C.byref(res), C.byref(peak), C.byref(act), C.byref(q)))
        return {"resident": res.value, "peak": peak.value, "active": act.value, "queued": q.value}

    def close(self):
        if getattr(self, "h", None):
            lib().fgx_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
