// fgx_types.hpp -- plain-old-data descriptors shared by host and device code.
#pragma once

#include <stdint.h>

namespace fgx {

enum { ELEC_NONE = 0, ELEC_COULOMB = 1, ELEC_DH = 2, ELEC_EWALD = 3 };
enum { SR_NONE = 0, SR_HS = 1, SR_LJ = 2 };
enum { NB_PLAIN = 0, NB_VECTOR = 1, NB_CUTG2G = 2 };
enum { MOVE_AT = 1, MOVE_TR = 2 };

// per-type pair tables, each ntypes x ntypes, stored back to back
enum { TAB_S2 = 0, TAB_EPS4 = 1, TAB_QEQE = 2, TAB_COUNT = 3 };

// Cuboid side lengths (Cuboid::setlen, src/geometry.cpp:263-272)
struct Box {
    double len[3], half[3], inv[3];
    float lenf[3];
    float fscale[3];  // fixed-point unit: len / 2^32
};

// Hamiltonian scalars (tables live in a separate device array)
struct Pot {
    int elec, sr, nbkind, precision;
    int ntypes, ntypes2;
    double lB, kappa, g2g_rc2;
    double ew_alpha, ew_kappa, ew_rc2;
    float lBf, kappaf, ew_alphaf, ew_kappaf, ew_rc2f, ew_sf;
    int ew_plain;  // kappa*rc below 1e-12: screened real space reduces to erfc(ar)/r
    float exact_f; // mixed path: pairs with |u| above this (kT) are redone in FP64
    // CutShift<LennardJonesLB> cutoff^2 (0: uncut LJ)
    double lj_rc2;
    float lj_rc2f;
};

// One registered move (Movebase::MolListData, move.h:305-330, plus the
// TranslateRotate parameters, move.h:1578-1600)
struct MoveDesc {
    int type, molid, peratom, permol;
    double prob;
    double dir[3], dir2[3];
    double dp, dprot;
    int center_rotation;
};

// group table entry (group.h:21-390)
struct GroupDesc {
    int front, back, molsize, molid;
};

}  // namespace fgx
