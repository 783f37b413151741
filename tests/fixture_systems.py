"""Build HIP-engine systems from the golden fixtures (test helper)."""
from beofaunus_amd import fgx
from oracle_py import KINDS, lj_cut_of

ELEC = {"none": fgx.ELEC_NONE, "coulomb": fgx.ELEC_COULOMB, "dh": fgx.ELEC_DH,
        "ewald": fgx.ELEC_EWALD}
SR = {"none": fgx.SR_NONE, "hs": fgx.SR_HS, "lj": fgx.SR_LJ}
NB = {"plain": fgx.NB_PLAIN, "vector": fgx.NB_VECTOR, "cutg2g": fgx.NB_CUTG2G}


def fgx_system_from_fixture(fx, precision=fgx.FP64, moves=True, ewald_override=None):
    e = fx.get("energy") or fx.get("mc")
    p = fx["particles"]
    nb = fx["config"]["energy"]["nonbonded"]
    elec, sr, nbk = KINDS[fx["kind"]]
    kappa = 1.0 / nb["debyelength"] if "debyelength" in nb else 0.0
    ew = dict(nb["ewald"]) if "ewald" in nb and elec == "ewald" else None
    if ew and ewald_override:
        ew.update(ewald_override)
    s = fgx.System(fx["box"], p["xyz"], p["charge"], p["radius"], p["mw"], p["id"],
                   fx["groups"], elec=ELEC[elec], sr=SR[sr], nbkind=NB[nbk],
                   precision=precision, lB=e["lB"], kappa=kappa,
                   g2g_cut=nb.get("cutoff_g2g", 0.0),
                   type_sigma=e["type_sigma"], type_eps=e["type_eps"],
                   type_charge=e["type_charge"], type_radius=e["type_radius"],
                   type_dp=e["type_dp"], ewald=ew, lj_cut=lj_cut_of(fx["kind"], nb),
                   pressure=e.get("pressure"))
    if moves:
        for mv in fx["moves"]:
            if mv["type"] == "NPT":   # driven by isobaric_propagate
                continue
            kind = fgx.MOVE_AT if mv["type"] == "AT" else fgx.MOVE_TR
            s.add_move(kind, mv["molid"], prob=mv.get("prob", 1.0),
                       peratom=mv.get("peratom", 0), permol=mv.get("permol", 0),
                       dp=mv.get("dp", 0.0), dprot=mv.get("dprot", 0.0),
                       center_rotation=mv.get("center_rotation", 1))
    return s
