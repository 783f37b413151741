// fgx_sweep_kernels.hpp -- device kernels of the device-resident Propagator
// (move.h:7877-7891, Movebase::move move.h:515-555) for AtomicTranslation
// (move.h:818-901) and TranslateRotate (move.h:1620-1713), shared by the
// translation units that instantiate them (one per Hamiltonian family, so
// they compile in parallel) and by the host side in fgx_sweep.hip.
//
// One workgroup owns one Markov chain for the whole launch: it replays the
// reference's two mt19937 engines (global `slump` and the move-internal
// `_slump()`) in LDS, generates each trial with the reference's FP64
// arithmetic, sums the energy change over all partners, takes the Metropolis
// decision and commits -- no host round trip per trial.  Many chains (batch,
// replicas, jobs) run as many workgroups of one launch.
//
// Pair math is FP64, or (mixed) FP32 on fixed-point coordinates whose
// integer difference is the exact minimum image, with FP64 accumulation.
// Ewald S(k) stays FP64 in the LDS of the chain's workgroup.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "fgx_device.hpp"
#include "fgx_internal.hpp"

namespace fgx {

static constexpr int MAXK_AXIS = 64;     // |n| per axis for the Ewald tables
static constexpr int MAX_GROUP_SITES = 512;
static constexpr int MAX_GROUPS_LDS = 4096;

struct ChainDev {
    Box box;
    Pot pot;
    int n, ng;
    double *x, *y, *z;                 // accepted vector (FP64)
    const double *q, *r;
    const int *t, *gid;
    const int4 *pk;                    // packed accepted vector (fixed-point xyz, FP32 charge)
    int4 *pk_w;
    const float *qf, *rf;
    const double *tab;
    const float *tabf;
    const GroupDesc *groups;
    double *cm;                        // 6 * ng (cm, cm_trial)
    const double *type_dp;
    MoveDesc mv;
    int nmoves, repeat;
    const int *gvec;
    int ngvec;
    uint32_t *rng;                     // 2 x 625: global, internal
    // Ewald
    int K, kcc;
    const int *kn;
    const double *aks;
    double *Q;
    const double *eff;
    double V, recip_pref, surf_pref;   // (2pi/V) lB ; const_inf 2pi/((2eps+1)V) lB
    int update_frequency;
    int *ew_cnt;                       // cnt_accepted (device)
    // results
    long *stats_i;                     // trials, accepted, inf-rejects, refreshes
    double *stats_d;                   // utot, dusum
    double *du_rec;
    int *acc_rec;
    long cap;
    int ncalls;                        // Propagator::move() calls for this chain
    long max_trials;                   // trial budget of this launch (<= 0: none)
    int *resume;                       // unfinished call: [in progress, rep, go]
    // group-pair energy cache (ng x ng, FP64) of the mixed TranslateRotate
    // sweep, or null
    double *emat;
    // fixed-unit pair form (cubic box; Coulomb/DH + HS, or Ewald + LJ): pair
    // distances stay in units of s = L / 2^32
    int fx;
    float fx_lbs, fx_rs, fx_c1;   // lB / s, sqrt(1 + 1e-6) / s, -kappa s / ln 2
    float fx_rc2, fx_s2scale;     // Ewald rc^2 / s^2; 1 / s^2 for LJ sigma^2
    // LJ parameters when every type pair present shares them (fixed units)
    int lj_uniform, lj_check;
    float lj_s2f, lj_eps4f;
    // erfcx Chebyshev series for the mixed path's FP64 pass (plain Ewald)
    const double *ew_cheb;
    int ew_ncheb;
    double ew_cheb_t;             // 2 / xmax
    double lj_rc2inv;             // 1 / CutShift cutoff^2 (0: uncut LJ)
    // every pair term vanishes beyond the Ewald cutoff (CutShift LJ at rc):
    // the mixed partner loop only screens for partners inside rc
    int lj_screen;
    // cell list of the mixed AtomicTranslation sweep (cubic box, lj_screen):
    // nc^3 cells of side >= rc + the largest displacement per axis, each with
    // `cell_cap` records {fixed x, y, z, particle}; counts live in LDS during
    // a launch.  cell_nc == 0: no cell list.
    int cell_nc, cell_cap;
    int cell_type;                // partner types must be read (type tables not uniform)
    int *cell_cnt;
    double4 *cell_rec;            // FP64 {x, y, z, charge} per slot
    int *cell_idx;                // particle per slot
    int *cell_slot;               // per particle: cell * cap + slot
    // Ewald state in device memory (more k-vectors than LDS holds): the
    // trial's dS(k) and the packed n-triples; S(k) is Q itself, A_k aks
    double2 *sk_dS;
    int *sk_kid;
    // partners [sp_j0, sp_j1) staged in LDS for the launch (mixed
    // AtomicTranslation without Ewald): packed record, FP64 position, charge
    // and radius, FP32 radius and type, LDS_PART_BYTES per particle
    int ldsp;
    int sp_j0, sp_j1;
    // block sweep (k_sweep_blk): workgroups sharing the chain, and their
    // exchange slots (2 x BLK_NSPLIT_MAX x BLK_NV_MAX doubles), arrival and
    // exit counters (zero between launches) and the launch's error word
    int nsplit;
    double *xbuf;
    int *xcnt;
    int *xerr;
};

static constexpr int LDS_PART_BYTES = 16 + 5 * 8 + 4 + 4;

// the LDS-resident partner arrays of a chain (see ChainDev::ldsp); LDS
// address-space pointers, so accesses compile to LDS (not flat) instructions
#define FGX_LDS __attribute__((address_space(3)))
typedef int lds_v4i __attribute__((ext_vector_type(4)));
__device__ __forceinline__ int4 v4i_to_int4(lds_v4i v) { return make_int4(v.x, v.y, v.z, v.w); }
__device__ __forceinline__ lds_v4i int4_to_v4i(int4 v) { return lds_v4i{v.x, v.y, v.z, v.w}; }
struct LdsPart {
    FGX_LDS lds_v4i *pk;
    FGX_LDS double *x, *y, *z, *q, *r;
    FGX_LDS float *rf;
    FGX_LDS int *t;
    int j0, ns;
    __device__ void carve(FGX_LDS void *base, int n)
    {
        pk = (FGX_LDS lds_v4i *)base;
        x = (FGX_LDS double *)(pk + n);
        y = x + n; z = y + n; q = z + n; r = q + n;
        rf = (FGX_LDS float *)(r + n);
        t = (FGX_LDS int *)(rf + n);
    }
};

// ------------------------------------------------------------------------
// pair-loop helpers
// ------------------------------------------------------------------------
// u(new, j) - u(old, j) for one partner with given FP64 data
template <int ELEC, int SR>
__device__ __forceinline__ double du_pair64_v(const ChainDev &C, const double *__restrict__ tab,
                                              double nx, double ny, double nz, double ox, double oy,
                                              double oz, double aq, double ar, int at, double bx,
                                              double by, double bz, double bq, double br, int bt)
{
    double r2n, r2o;
    if (ELEC == ELEC_EWALD || C.pot.nbkind == NB_VECTOR) {
        double rx, ry, rz;
        vdist_d(C.box, nx, ny, nz, bx, by, bz, rx, ry, rz);
        r2n = rx * rx + (ry * ry + rz * rz);
        vdist_d(C.box, ox, oy, oz, bx, by, bz, rx, ry, rz);
        r2o = rx * rx + (ry * ry + rz * rz);
    } else {
        r2n = sqdist_d(C.box, nx, ny, nz, bx, by, bz);
        r2o = sqdist_d(C.box, ox, oy, oz, bx, by, bz);
    }
    return pair_u64_t<ELEC, SR>(C.pot, tab, r2n, aq, bq, ar, br, at, bt) -
           pair_u64_t<ELEC, SR>(C.pot, tab, r2o, aq, bq, ar, br, at, bt);
}

// u(new, j) - u(old, j) for one partner j, FP64
template <int ELEC, int SR>
__device__ __forceinline__ double du_pair64(const ChainDev &C, const double *__restrict__ tab,
                                            double nx, double ny, double nz, double ox, double oy,
                                            double oz, double aq, double ar, int at, int j)
{
    return du_pair64_v<ELEC, SR>(C, tab, nx, ny, nz, ox, oy, oz, aq, ar, at, C.x[j], C.y[j], C.z[j],
                                 C.q[j], C.r[j], C.t[j]);
}

// FP32 pair energy in fixed-point units: distances in units of s = L / 2^32,
// charge as lB q / s, radius as r sqrt(1 + 1e-6) / s; `bad` flags a pair
// whose value must come from FP64 (strong, non-finite, or within FP32
// rounding of hard-sphere contact).
template <int ELEC, int SR>
__device__ __forceinline__ float pair_fx(const ChainDev &C, float r2, float aq, float bq, float ra,
                                         float rb, float exact, bool &bad)
{
    float u = 0.f;
    if constexpr (ELEC == ELEC_COULOMB) {
        u = aq * bq * __builtin_amdgcn_rsqf(r2);
    } else if constexpr (ELEC == ELEC_DH) {
        const float rinv = __builtin_amdgcn_rsqf(r2);
        u = aq * bq * rinv * __builtin_amdgcn_exp2f(C.fx_c1 * (r2 * rinv));
    }
    bool b = !(fabsf(u) <= exact);
    if constexpr (SR == SR_HS) {
        const float m = ra + rb;
        b = b || (r2 < m * m);
    }
    bad = b;
    return u;
}

// erfc(x) for x in [0, xmax] = exp(-x^2) erfcx(x), erfcx from its Chebyshev
// series (coefficients c[0..n), argument t = 2x/xmax - 1, Clenshaw), at two
// arguments in one pass: the two recurrences share each coefficient load.
__device__ __forceinline__ void erfc_series2(const double *c, int n, double tscale, double xa,
                                             double xb, double &fa, double &fb)
{
    const double ta = fma(xa, tscale, -1.0), tb = fma(xb, tscale, -1.0);
    const double ta2 = ta + ta, tb2 = tb + tb;
    double a1 = 0.0, a2 = 0.0, b1 = 0.0, b2 = 0.0;
#pragma unroll 4
    for (int k = n - 1; k >= 1; --k) {
        const double ck = c[k];
        const double a0 = fma(ta2, a1, ck - a2);
        a2 = a1;
        a1 = a0;
        const double b0 = fma(tb2, b1, ck - b2);
        b2 = b1;
        b1 = b0;
    }
    fa = exp(-xa * xa) * fma(ta, a1, c[0] - a2);
    fb = exp(-xb * xb) * fma(tb, b1, c[0] - b2);
}

// u(r2n) - u(r2o) for the mixed path's FP64 evaluations with plain Ewald
// (kappa -> 0): real-space erfc(alpha r) / r from the staged series inside
// rc (both positions in one pass), plus the short-range term (LJ, or
// CutShift LJ zero beyond its cutoff and shifted by its value there).  One
// division per position: 1/r and the LJ ratio come from 1/r^2.
template <int SR>
__device__ __forceinline__ double du_ewald64(const ChainDev &C, const double *__restrict__ tab,
                                             const double *cheb, double r2n, double r2o,
                                             double qq, double aq, double bq, double ar,
                                             double br, int at, int bt)
{
    const double rc2 = C.pot.ew_rc2;
    const double rin2 = 1.0 / r2n, rio2 = 1.0 / r2o;
    double en = 0.0, eo = 0.0;
    const bool inn = r2n <= rc2, ino = r2o <= rc2;
    if (inn || ino) {
        const double rn = sqrt(r2n), ro = sqrt(r2o);
        double fn, fo;
        erfc_series2(cheb, C.ew_ncheb, C.ew_cheb_t, inn ? C.pot.ew_alpha * rn : 0.0,
                     ino ? C.pot.ew_alpha * ro : 0.0, fn, fo);
        if (inn) en = qq * fn * (rn * rin2);
        if (ino) eo = qq * fo * (ro * rio2);
    }
    if constexpr (SR == SR_LJ) {
        const int k = at * C.pot.ntypes + bt;
        const double s2 = tab[TAB_S2 * C.pot.ntypes2 + k], e4 = tab[TAB_EPS4 * C.pot.ntypes2 + k];
        double xn = s2 * rin2, xo = s2 * rio2;
        xn = xn * xn * xn;
        xo = xo * xo * xo;
        double ljn = e4 * (xn * xn - xn), ljo = e4 * (xo * xo - xo);
        if (C.pot.lj_rc2 > 0.0) {
            double xc = s2 * C.lj_rc2inv;
            xc = xc * xc * xc;
            const double uc = e4 * (xc * xc - xc);
            ljn = (r2n > C.pot.lj_rc2) ? 0.0 : ljn - uc;
            ljo = (r2o > C.pot.lj_rc2) ? 0.0 : ljo - uc;
        }
        return (en + ljn) - (eo + ljo);
    } else {
        return (en + pair_u64_t<ELEC_NONE, SR>(C.pot, tab, r2n, aq, bq, ar, br, at, bt)) -
               (eo + pair_u64_t<ELEC_NONE, SR>(C.pot, tab, r2o, aq, bq, ar, br, at, bt));
    }
}

// u(new, j) - u(old, j) in FP64 for the mixed path's exact pass.  As
// du_pair64, except that a plain real-space Ewald term (kappa -> 0, where the
// screened form of ewald.h:109 reduces to erfc(alpha r) / r) takes erfc from
// the staged series instead of the library call.
template <int ELEC, int SR>
__device__ __forceinline__ double du_pair64_mixed(const ChainDev &C, const double *__restrict__ tab,
                                                  const double *cheb, double nx, double ny,
                                                  double nz, double ox, double oy, double oz,
                                                  double aq, double ar, int at, int j)
{
    if constexpr (ELEC == ELEC_EWALD) {
        if (C.ew_ncheb > 0) {
            const double bx = C.x[j], by = C.y[j], bz = C.z[j], bq = C.q[j], br = C.r[j];
            const int bt = C.t[j];
            double rx, ry, rz;
            vdist_d(C.box, nx, ny, nz, bx, by, bz, rx, ry, rz);
            const double r2n = rx * rx + (ry * ry + rz * rz);
            vdist_d(C.box, ox, oy, oz, bx, by, bz, rx, ry, rz);
            const double r2o = rx * rx + (ry * ry + rz * rz);
            const int k = at * C.pot.ntypes + bt;
            const double qq = aq * bq * tab[TAB_QEQE * C.pot.ntypes2 + k] * C.pot.lB;
            return du_ewald64<SR>(C, tab, cheb, r2n, r2o, qq, aq, bq, ar, br, at, bt);
        }
    }
    return du_pair64<ELEC, SR>(C, tab, nx, ny, nz, ox, oy, oz, aq, ar, at, j);
}

static constexpr int DCAP = 2048;   // deferred-pair queue capacity (LDS), all wavefronts
static constexpr int LDS_TYPES2 = 256;  // type tables staged in LDS up to 16 types

// Mixed-path pair loop of one AtomicTranslation trial.  Partners are read as
// packed 16-byte records (fixed-point x, y, z and FP32 charge), U per thread
// per batch, and evaluated in FP32 for the old and the new position with the
// difference summed in FP64.  Partners inside the Ewald real-space cutoff, and
// pairs whose FP32 energy is strong (|u| > Pot::exact_f), non-finite or a
// hard-sphere contact, are flagged in a per-batch bit mask and appended to
// this wavefront's queue (dq, `qcap` entries; the wavefront's count in *nq),
// to be evaluated in FP64 afterwards, so the erfc/exp work is compacted
// instead of idling divergent lanes.  Queue order is fixed by lane order, so
// results are deterministic.  Called by whole wavefronts.
// Loads through a global-address-space pointer: the ChainDev arrays are
// device memory, and saying so lets the compiler use global (not flat)
// loads, which do not hold back LDS traffic.
template <typename T>
__device__ __forceinline__ T gload(const T *p, int i)
{
    return ((const __attribute__((address_space(1))) T *)p)[i];
}
__device__ __forceinline__ double4 gload(const double4 *p, int i)
{
    typedef double v4d __attribute__((ext_vector_type(4)));
    const v4d v = ((const __attribute__((address_space(1))) v4d *)p)[i];
    return make_double4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ int4 gload(const int4 *p, int i)
{
    typedef int v4i __attribute__((ext_vector_type(4)));
    const v4i v = ((const __attribute__((address_space(1))) v4i *)p)[i];
    return make_int4(v.x, v.y, v.z, v.w);
}

template <int U, bool HS, bool NEED_T>
struct PartnerBatch {
    int4 P[U];
    float R[U];
    int T[U];
    // rows base + u nthr; loads past n read the zero padding (the loop reads
    // at most three batches beyond the end: PARTNER_PAD >= 3 U nthr, checked
    // at launch)
    __device__ __forceinline__ void load(const ChainDev &C, int base, int nthr, int tid)
    {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int row = base + u * nthr;   // wavefront-uniform
            P[u] = gload(C.pk + row, tid);
            R[u] = HS ? gload(C.rf + row, tid) : 0.f;
            T[u] = NEED_T ? gload(C.t + row, tid) : 0;
        }
    }
    // the same rows from the LDS-resident slice (rows past its end read zero)
    __device__ __forceinline__ void load_lds(const LdsPart &L, int base, int nthr, int tid)
    {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int row = base + u * nthr + tid;
            const bool in = row < L.ns;
            P[u] = v4i_to_int4(in ? L.pk[row] : lds_v4i{0, 0, 0, 0});
            R[u] = HS ? (in ? L.rf[row] : 0.f) : 0.f;
            T[u] = NEED_T ? (in ? L.t[row] : 0) : 0;
        }
    }
};

// true where the mixed AtomicTranslation loop has a fixed-unit form
template <int ELEC, int SR>
__host__ __device__ constexpr bool at_fx_ok()
{
    return (ELEC == ELEC_EWALD && SR == SR_LJ) ||
           ((ELEC == ELEC_NONE || ELEC == ELEC_COULOMB || ELEC == ELEC_DH) &&
            (SR == SR_NONE || SR == SR_HS));
}

template <int ELEC, int SR, int U, bool FX, bool SCREEN_ONLY = false, bool LP = false>
__device__ __forceinline__ double pair_loop_mixed(const ChainDev &C, const float *__restrict__ tabf,
                                                  const double *__restrict__ tab, int i, int4 me,
                                                  float ar, int at, double aq64, double ar64,
                                                  double tnx, double tny, double tnz, double tox,
                                                  double toy, double toz, int *dq, int qcap,
                                                  int &nq, const LdsPart L = LdsPart{})
{
    // LP: partners are the LDS-resident slice L; rows and queue entries are
    // slice-local (particle L.j0 + row)
    const int nrow = LP ? L.ns : C.n;
    const int iself = LP ? i - L.j0 : i;
    const int tid = threadIdx.x, nthr = blockDim.x, lane = tid & 63;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const float aq = __int_as_float(me.w);
    const int32_t nx = to_fixed(tnx, C.box.inv[0]), ny = to_fixed(tny, C.box.inv[1]),
                  nz = to_fixed(tnz, C.box.inv[2]);
    const float sx = C.box.fscale[0], sy = C.box.fscale[1], sz = C.box.fscale[2];
    const float exact = C.pot.exact_f;
    constexpr bool NEED_T = (SR == SR_LJ || ELEC == ELEC_EWALD);
    // fixed-unit form: charge lB q / s, radius r c / s (see stage_site)
    // Ewald partners inside rc (with a margin above FP32 rounding) go to the
    // FP64 pass, which applies the exact cut
    const float rc2m = C.pot.ew_rc2f * 1.00001f;
    const float aqx = FX ? aq * C.fx_lbs : aq;
    const float arx = FX ? ar * C.fx_rs : ar;
    int cnt = 0;   // wavefront-uniform
    double acc = 0.0;
    // two batches alternate: one is evaluated while the other's loads are in
    // flight (the loop body is unrolled twice so neither is copied)
    PartnerBatch<U, SR == SR_HS, NEED_T> A, B;
    // mode 1 (FAST): one shared LJ parameter pair and no strong-pair test
    // beyond rc; mode 2 (SCREEN): nothing beyond rc (CutShift LJ at rc), so
    // the loop only finds the partners inside rc.  Decided once per launch,
    // so the unrolled body has no branches.
    auto eval = [&](auto mode, const PartnerBatch<U, SR == SR_HS, NEED_T> &X, int base) {
        constexpr bool FAST = decltype(mode)::value >= 1;
        constexpr bool SCREEN = decltype(mode)::value == 2;
        // FAST Ewald: beyond rc only the LJ tail remains (|u| < 1e-2 kT per
        // pair), so a batch's U differences are summed in FP32 and widened once
        constexpr bool BATCH32 = FX && FAST && !SCREEN && ELEC == ELEC_EWALD;
        float bsum = 0.f;
        unsigned dm = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = base + u * nthr + tid;
            const bool ok = j < nrow && j != iself;
            const int4 P = X.P[u];
            if constexpr (FX) {
                float dx = (float)(int32_t)(nx - P.x), dy = (float)(int32_t)(ny - P.y),
                      dz = (float)(int32_t)(nz - P.z);
                const float r2n = dx * dx + dy * dy + dz * dz;
                dx = (float)(int32_t)(me.x - P.x);
                dy = (float)(int32_t)(me.y - P.y);
                dz = (float)(int32_t)(me.z - P.z);
                const float r2o = dx * dx + dy * dy + dz * dz;
                float d;
                bool bad;
                if constexpr (ELEC == ELEC_EWALD && SCREEN) {
                    bad = fminf(r2n, r2o) <= C.fx_rc2;
                    d = 0.f;
                } else if constexpr (ELEC == ELEC_EWALD) {
                    // LJ only beyond rc (real-space Ewald is zero there);
                    // partners inside rc go to the FP64 pass
                    float s2 = C.lj_s2f, e4 = C.lj_eps4f;
                    if constexpr (!FAST) {
                        const int k = at * C.pot.ntypes + X.T[u];
                        s2 = tabf[TAB_S2 * C.pot.ntypes2 + k];
                        e4 = tabf[TAB_EPS4 * C.pot.ntypes2 + k];
                    }
                    float xn = s2 * __builtin_amdgcn_rcpf(r2n);
                    float xo = s2 * __builtin_amdgcn_rcpf(r2o);
                    xn = xn * xn * xn;
                    xo = xo * xo * xo;
                    const float vn = xn * xn - xn, vo = xo * xo - xo;
                    bad = fminf(r2n, r2o) <= C.fx_rc2;
                    if constexpr (!FAST)
                        bad = bad || !(fabsf(e4 * vn) <= exact) || !(fabsf(e4 * vo) <= exact);
                    d = e4 * (vn - vo);
                } else {
                    bool bn, bo;
                    const float rb = X.R[u] * C.fx_rs;
                    const float un = pair_fx<ELEC, SR>(C, r2n, aqx, __int_as_float(P.w), arx, rb,
                                                       exact, bn);
                    const float uo = pair_fx<ELEC, SR>(C, r2o, aqx, __int_as_float(P.w), arx, rb,
                                                       exact, bo);
                    bad = bn || bo;
                    d = un - uo;
                }
                const bool take = ok && !bad;
                dm |= ((ok && bad) ? 1u : 0u) << u;
                if constexpr (BATCH32) {
                    bsum += take ? d : 0.f;
                } else {
                    // select on the FP32 bits, then widen: one select, not two
                    acc += (double)__int_as_float(__float_as_int(d) & (take ? -1 : 0));
                }
                continue;
            }
            float dx = (float)(int32_t)(nx - P.x) * sx, dy = (float)(int32_t)(ny - P.y) * sy,
                  dz = (float)(int32_t)(nz - P.z) * sz;
            const float r2n = dx * dx + dy * dy + dz * dz;
            dx = (float)(int32_t)(me.x - P.x) * sx;
            dy = (float)(int32_t)(me.y - P.y) * sy;
            dz = (float)(int32_t)(me.z - P.z) * sz;
            const float r2o = dx * dx + dy * dy + dz * dz;
            const float bq = __int_as_float(P.w);
            float un, uo;
            bool bad;
            if constexpr (ELEC == ELEC_EWALD) {
                // beyond rc the real-space term is zero; inside, FP64 decides
                un = pair_u32_t<ELEC_NONE, SR>(C.pot, tabf, r2n, aq, bq, ar, X.R[u], at, X.T[u]);
                uo = pair_u32_t<ELEC_NONE, SR>(C.pot, tabf, r2o, aq, bq, ar, X.R[u], at, X.T[u]);
                bad = (r2n <= rc2m) || (r2o <= rc2m);
            } else {
                un = pair_u32_t<ELEC, SR>(C.pot, tabf, r2n, aq, bq, ar, X.R[u], at, X.T[u]);
                uo = pair_u32_t<ELEC, SR>(C.pot, tabf, r2o, aq, bq, ar, X.R[u], at, X.T[u]);
                bad = false;
            }
            bad = bad || !(fabsf(un) <= exact) || !(fabsf(uo) <= exact);
            dm |= ((ok && bad) ? 1u : 0u) << u;
            acc += (double)((ok && !bad) ? un - uo : 0.f);
        }
        if constexpr (BATCH32) acc += (double)bsum;
        if (__ballot(dm != 0)) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool d = (dm >> u) & 1u;
                const uint64_t m = __ballot(d);
                if (m) {
                    if (d) {
                        const int j = base + u * nthr + tid;
                        const int k = cnt + __popcll(m & below);
                        if (k < qcap)
                            dq[k] = j;
                        else if constexpr (LP)
                            acc += du_pair64_v<ELEC, SR>(C, tab, tnx, tny, tnz, tox, toy, toz,
                                                         aq64, ar64, at, L.x[j], L.y[j], L.z[j],
                                                         L.q[j], L.r[j], L.t[j]);
                        else
                            acc += du_pair64<ELEC, SR>(C, tab, tnx, tny, tnz, tox, toy, toz,
                                                       aq64, ar64, at, j);
                    }
                    cnt += __popcll(m);
                }
            }
        }
    };
    const int step = U * nthr;
    auto run = [&](auto mode) {
        if constexpr (LP) {
            // LDS latency needs no second batch in flight
            for (int base = 0; base < nrow; base += step) {
                A.load_lds(L, base, nthr, tid);
                eval(mode, A, base);
            }
        } else {
            A.load(C, 0, nthr, tid);
            for (int base = 0; base < C.n; base += 2 * step) {
                B.load(C, base + step, nthr, tid);        // clamped: always in bounds
                eval(mode, A, base);
                A.load(C, base + 2 * step, nthr, tid);
                if (base + step < C.n) eval(mode, B, base + step);
            }
        }
    };
    if constexpr (FX && ELEC == ELEC_EWALD && SCREEN_ONLY)
        run(std::integral_constant<int, 2>{});
    else if (FX && ELEC == ELEC_EWALD && C.lj_uniform && !C.lj_check)
        run(std::integral_constant<int, 1>{});
    else
        run(std::integral_constant<int, 0>{});
    nq = min(cnt, qcap);
    return acc;
}

// ------------------------------------------------------------------------
// Cell list of the mixed AtomicTranslation sweep.  Cells split each axis of
// the fixed-point coordinate range into nc equal parts: floor((x/L + 1/2) nc).
// Two coordinates whose minimum-image distance is at most one cell side lie
// in the same or in adjacent cells (periodically), so with a side of at least
// rc plus the largest displacement of one trial per axis, the 27 cells around
// the moved particle's old cell hold every partner inside rc of its old and
// of its new position.
// ------------------------------------------------------------------------
// neighbour cells per half-wavefront: cell-list sweeps run 512 threads, whose
// waves 1..7 hold 14 half-wavefronts for the 27 cells
static constexpr int CELL_NB = 2;

__device__ __forceinline__ int cell_axis(int32_t f, int nc)
{
    const uint32_t u = (uint32_t)f ^ 0x80000000u;
    return (int)(((uint64_t)u * (uint64_t)(uint32_t)nc) >> 32);
}

__device__ __forceinline__ int cell_of(int32_t x, int32_t y, int32_t z, int nc)
{
    return cell_axis(x, nc) + nc * (cell_axis(y, nc) + nc * cell_axis(z, nc));
}

// u(new, j) - u(old, j) in FP64 from a cell record (FP64 position and
// charge of j), as du_pair64_mixed computes it from the particle arrays;
// zero when j lies beyond rc of both positions (every pair term is cut there)
template <int ELEC, int SR>
__device__ __forceinline__ double du_pair64_rec(const ChainDev &C, const double *__restrict__ tab,
                                                const double *cheb, double nx, double ny,
                                                double nz, double ox, double oy, double oz,
                                                double aq, double ar, int at, double4 p, int j)
{
    double rx, ry, rz;
    vdist_d(C.box, nx, ny, nz, p.x, p.y, p.z, rx, ry, rz);
    const double r2n = rx * rx + (ry * ry + rz * rz);
    vdist_d(C.box, ox, oy, oz, p.x, p.y, p.z, rx, ry, rz);
    const double r2o = rx * rx + (ry * ry + rz * rz);
    if (r2n > C.pot.ew_rc2 && r2o > C.pot.ew_rc2) return 0.0;
    const int bt = C.cell_type ? gload(C.t, j) : at;
    const double bq = p.w;
    if (ELEC == ELEC_EWALD && C.ew_ncheb > 0) {
        const int k = at * C.pot.ntypes + bt;
        const double qq = aq * bq * tab[TAB_QEQE * C.pot.ntypes2 + k] * C.pot.lB;
        return du_ewald64<SR>(C, tab, cheb, r2n, r2o, qq, aq, bq, ar, 0.0, at, bt);
    }
    return pair_u64_t<ELEC, SR>(C.pot, tab, r2n, aq, bq, ar, 0.0, at, bt) -
           pair_u64_t<ELEC, SR>(C.pot, tab, r2o, aq, bq, ar, 0.0, at, bt);
}

// Sum over partners inside rc of u(new, j) - u(old, j), FP64 (every pair
// term vanishes beyond rc).  Wavefront 0 is drawing the next trial, so the
// half-wavefronts of the other waves take the 27 neighbour cells, cell
// h, h + H, ... for half-wave h of H, one slot per lane; the first 32 slots
// of all its cells are loaded before any arithmetic.  Fixed order: the sum
// is deterministic.  Called by all threads of the block.
template <int ELEC, int SR>
__device__ __forceinline__ double cell_loop(const ChainDev &C, const double *__restrict__ tab,
                                            const double *cheb, const int *lcnt, int i, int4 me,
                                            double aq64, double ar64, int at, double tnx,
                                            double tny, double tnz, double tox, double toy,
                                            double toz)
{
    const int tid = threadIdx.x, nthr = blockDim.x;
    if (tid < 64) return 0.0;
    const int hw = (tid >> 5) - 2, H = (nthr >> 5) - 2, l = tid & 31;
    const int nc = C.cell_nc, cap = C.cell_cap;
    const int cx = cell_axis(me.x, nc), cy = cell_axis(me.y, nc), cz = cell_axis(me.z, nc);
    int cb[CELL_NB], nb[CELL_NB], jb[CELL_NB];
    double4 pb[CELL_NB];
#pragma unroll
    for (int b = 0; b < CELL_NB; ++b) {
        const int k = hw + b * H;
        cb[b] = 0;
        nb[b] = 0;
        if (k < 27) {
            int ax = cx + k % 3 - 1, ay = cy + (k / 3) % 3 - 1, az = cz + k / 9 - 1;
            ax = ax < 0 ? nc - 1 : (ax >= nc ? 0 : ax);
            ay = ay < 0 ? nc - 1 : (ay >= nc ? 0 : ay);
            az = az < 0 ? nc - 1 : (az >= nc ? 0 : az);
            cb[b] = (ax + nc * (ay + nc * az)) * cap;
            nb[b] = lcnt[ax + nc * (ay + nc * az)];
        }
    }
#pragma unroll
    for (int b = 0; b < CELL_NB; ++b) {
        jb[b] = -1;
        if (l < nb[b]) {
            pb[b] = gload(C.cell_rec, cb[b] + l);
            jb[b] = gload(C.cell_idx, cb[b] + l);
        }
    }
    double acc = 0.0;
#pragma unroll
    for (int b = 0; b < CELL_NB; ++b)
        if (jb[b] >= 0 && jb[b] != i)
            acc += du_pair64_rec<ELEC, SR>(C, tab, cheb, tnx, tny, tnz, tox, toy, toz, aq64, ar64,
                                           at, pb[b], jb[b]);
    // cells holding more than 32 records
#pragma unroll
    for (int b = 0; b < CELL_NB; ++b)
        for (int q = l + 32; q < nb[b]; q += 32) {
            const int j = gload(C.cell_idx, cb[b] + q);
            if (j != i)
                acc += du_pair64_rec<ELEC, SR>(C, tab, cheb, tnx, tny, tnz, tox, toy, toz, aq64,
                                               ar64, at, gload(C.cell_rec, cb[b] + q), j);
        }
    return acc;
}

// ------------------------------------------------------------------------
// Ewald per-trial helpers: exp(i 2 pi n x / L) for n = 0..kcc per axis
// ------------------------------------------------------------------------
__device__ __forceinline__ void ew_factor(const double2 *tab, int kcc1, int axis, int n, double &re,
                                          double &im)
{
    double2 e = tab[axis * kcc1 + (n < 0 ? -n : n)];
    re = e.x;
    im = n < 0 ? -e.y : e.y;
}

__device__ __forceinline__ void cmul(double ar, double ai, double br, double bi, double &cr,
                                     double &ci)
{
    cr = ar * br - ai * bi;
    ci = ar * bi + ai * br;
}

// exp(i k.r) from the per-axis tables
__device__ __forceinline__ void ew_phase(const double2 *tab, int kcc1, int nx, int ny, int nz,
                                         double &re, double &im)
{
    double ar, ai, br, bi, cr, ci;
    ew_factor(tab, kcc1, 0, nx, ar, ai);
    ew_factor(tab, kcc1, 1, ny, br, bi);
    cmul(ar, ai, br, bi, cr, ci);
    ew_factor(tab, kcc1, 2, nz, ar, ai);
    cmul(cr, ci, ar, ai, re, im);
}

// ------------------------------------------------------------------------
// AtomicTranslation sweep
// ------------------------------------------------------------------------
// One AtomicTranslation trial: its random numbers, the moved particle's data
// and the proposed position.  `pf` marks a record whose draws and particle
// data were fetched ahead, during the previous trial of the same call.
struct TrialAT {
    int i;
    int go;
    int acc;
    int collide;
    int pf;
    int ti;                       // type of i
    int slot;                     // cell-list record of i (cell * cap + slot)
    int4 pk;                      // packed record of i (old fixed-point position)
    float rf;                     // FP32 radius of i
    double q64, r64, wq;          // FP64 charge, radius, Ewald weight q eff
    double t3[3], u;              // displacement canonicals, Metropolis uniform
    double ox, oy, oz, nx, ny, nz;
};

static constexpr int AT_LDS_GROUPS = 256;   // group table staged in LDS up to this size
static constexpr int AT_LDS_TYPES = 64;

static constexpr int AT_THREADS = 512;
static constexpr int REFRESH_CHUNK = 16;   // particles per LDS table chunk
static constexpr int REFRESH_KAX = 16;     // kcc + 1 limit for the table refresh

__host__ __device__ __forceinline__ int kpack(int nx, int ny, int nz)
{
    return (nx & 0xff) | ((ny + 128) << 8) | ((nz + 128) << 16);
}
__device__ __forceinline__ void kunpack(int v, int &nx, int &ny, int &nz)
{
    nx = v & 0xff;
    ny = ((v >> 8) & 0xff) - 128;
    nz = ((v >> 16) & 0xff) - 128;
}

// bytes of dynamic LDS a chain needs: S(k) and the trial's dS(k) (double2),
// A_k (double), packed n (int)
__host__ __device__ inline size_t at_dyn_lds(int K) { return (size_t)K * (16 + 16 + 8 + 4); }

// CELL: every pair term vanishes beyond the Ewald rc (CutShift LJ at rc); the
// partner loop screens for partners inside rc, over the chain's cell list
// when it has one and over all partners otherwise.
// SKG: the Ewald state (S(k), the trial's dS(k), A_k, packed n) stays in the
// chain's device memory instead of LDS, for more k-vectors than LDS holds.
template <int ELEC, int SR, bool MIX, int U, bool FX, bool CELL = false, bool SKG = false>
__global__ __launch_bounds__(AT_THREADS) void k_sweep_at(const ChainDev *__restrict__ chains,
                                                         int ncalls)
{
    constexpr bool EW = (ELEC == ELEC_EWALD);
    const ChainDev &C = chains[blockIdx.x];
    const int tid = threadIdx.x, nthr = blockDim.x;
    const bool wave0 = tid < 64;
    __shared__ uint32_t mtg[624], mti[624];
    __shared__ double red[AT_THREADS / 64];
    __shared__ TrialAT T[2];
    __shared__ double2 etab[2][3 * MAXK_AXIS];   // old / new per-axis phase tables
    __shared__ double2 rtab[EW ? REFRESH_CHUNK * 3 * REFRESH_KAX : 1];
    __shared__ double rw[EW ? REFRESH_CHUNK : 1];
    __shared__ int dlist[MIX ? DCAP : 1];   // per-wavefront FP64 queues
    // per-type pair tables in LDS (the partner loop indexes them per lane)
    __shared__ double ltab[TAB_COUNT * LDS_TYPES2];
    __shared__ float ltabf[MIX ? TAB_COUNT * LDS_TYPES2 : 1];
    // Ewald state for the whole launch, LDS-resident: S(k) FP64, A_k, n-triples
    // move's group list, group table and per-type dp for the trial draw
    __shared__ int lgvec[AT_LDS_GROUPS];
    __shared__ GroupDesc lgroups[AT_LDS_GROUPS];
    __shared__ double ltdp[AT_LDS_TYPES], leff[AT_LDS_TYPES];
    __shared__ int nqs[AT_THREADS / 64];          // FP64 queue length per wavefront
    // chain scalars kept once in LDS (written by thread 0 only): surface
    // dipole M, utot, dusum
    __shared__ double sst[5];
    __shared__ double lcheb[EW && MIX ? EW_CHEB_MAX : 1];
    extern __shared__ double2 dyn[];
    double2 *S, *dS;
    double *Ak;
    int *kid, *lcnt;
    if constexpr (SKG) {
        S = (double2 *)C.Q;
        dS = C.sk_dS;
        Ak = (double *)C.aks;
        kid = C.sk_kid;
        lcnt = (int *)dyn;
    } else {
        S = dyn;
        dS = S + (EW ? C.K : 0);
        Ak = (double *)(dS + (EW ? C.K : 0));
        kid = (int *)(Ak + (EW ? C.K : 0));
        // cell counts (cell-list sweeps): LDS-resident for the launch
        lcnt = kid + (EW ? C.K : 0);
    }
    __shared__ int lcell_ok;
    // mixed sweep without Ewald: the partner slice is LDS-resident and the
    // trial loop's barriers order LDS only (its waves exchange nothing else)
    constexpr bool LPOK = !EW && MIX && !CELL;
    const bool ldsp = LPOK && C.ldsp;
    LdsPart L;
    L.j0 = ldsp ? C.sp_j0 : 0;
    L.ns = ldsp ? C.sp_j1 - C.sp_j0 : 0;
    L.carve((FGX_LDS void *)dyn, L.ns);
    // the whole chain in LDS: the moved particle's data comes from there too
    const bool lds_all = ldsp && L.j0 == 0 && L.ns == C.n;
    __shared__ double red2[2][AT_THREADS / 64];
    int rpar = 0;

    for (int k = tid; k < 624; k += nthr) {
        mtg[k] = C.rng[k];
        mti[k] = C.rng[625 + k];
    }
    WaveMT G{mtg, (int)C.rng[624]}, I{mti, (int)C.rng[625 + 624]};
    const bool lds_grp = C.ngvec <= AT_LDS_GROUPS && C.ng <= AT_LDS_GROUPS &&
                         C.pot.ntypes <= AT_LDS_TYPES;
    if (lds_grp) {
        for (int k = tid; k < C.ngvec; k += nthr) lgvec[k] = C.gvec[k];
        for (int k = tid; k < C.ng; k += nthr) lgroups[k] = C.groups[k];
        for (int k = tid; k < C.pot.ntypes; k += nthr) {
            ltdp[k] = C.type_dp[k];
            leff[k] = EW ? C.eff[k] : 0.0;
        }
    }
    if constexpr (EW && MIX)
        for (int k = tid; k < C.ew_ncheb; k += nthr) lcheb[k] = C.ew_cheb[k];
    if (tid == 0) {
        T[0].pf = T[1].pf = 0;
        T[0].acc = T[1].acc = 0;
    }
    if constexpr (CELL) {
        const int ncell = C.cell_nc * C.cell_nc * C.cell_nc;
        for (int c = tid; c < ncell; c += nthr) lcnt[c] = C.cell_cnt[c];
        if (tid == 0) lcell_ok = (C.cell_nc > 0 && C.cell_cnt[ncell] == 0) ? 1 : 0;
    }
    const bool lds_tab = C.pot.ntypes2 <= LDS_TYPES2;
    // one trial's draws from the global stream, in the reference's order:
    // group, particle (move.h:825, group.h:132), three displacement
    // canonicals (move.h:843-845), Metropolis uniform (move.h:570).  Called
    // by all 64 lanes of wave 0.
    auto at_draw = [&](TrialAT &t) {
        const int gk = mt_range(G, 0, C.ngvec - 1);
        const int g = lds_grp ? lgvec[gk] : C.gvec[gk];
        const GroupDesc gd = lds_grp ? lgroups[g] : C.groups[g];
        t.i = gd.front + mt_range(G, 0, gd.back - gd.front);
        t.t3[0] = mt_canonical(G);
        t.t3[1] = mt_canonical(G);
        t.t3[2] = mt_canonical(G);
        t.u = mt_canonical(G);
    };
    // the moved particle's data, loads issued together
    auto at_fetch = [&](TrialAT &t) {
        const int i = t.i;
        if (LPOK && lds_all) {
            t.ti = L.t[i];
            t.ox = L.x[i]; t.oy = L.y[i]; t.oz = L.z[i];
            t.pk = v4i_to_int4(L.pk[i]);
            t.rf = L.rf[i];
            t.q64 = L.q[i];
            t.r64 = L.r[i];
            t.wq = 0.0;
            t.slot = 0;
            return;
        }
        t.ti = gload(C.t, i);
        t.ox = gload(C.x, i); t.oy = gload(C.y, i); t.oz = gload(C.z, i);
        t.pk = gload(C.pk, i);
        t.rf = gload(C.rf, i);
        t.q64 = gload(C.q, i);
        t.r64 = gload(C.r, i);
        t.wq = EW ? t.q64 * (lds_grp ? leff[t.ti] : C.eff[t.ti]) : 0.0;
        t.slot = (CELL && C.cell_nc > 0) ? gload(C.cell_slot, i) : 0;
    };
    if (lds_tab) {
        for (int k = tid; k < TAB_COUNT * C.pot.ntypes2; k += nthr) {
            int t = k / C.pot.ntypes2, e = k % C.pot.ntypes2;
            ltab[t * C.pot.ntypes2 + e] = C.tab[k];
            // the fixed-unit loop takes sigma^2 in units of s^2
            if (MIX) ltabf[t * C.pot.ntypes2 + e] = C.tabf[k] * ((FX && t == TAB_S2) ? C.fx_s2scale : 1.f);
        }
    }
    const double *tab = lds_tab ? ltab : C.tab;
    const float *tabf = (lds_tab && MIX) ? ltabf : C.tabf;
    const int kcc1 = C.kcc + 1;
    if constexpr (LPOK) {
        if (ldsp) {
            for (int k = tid; k < L.ns; k += nthr) {
                const int j = L.j0 + k;
                L.pk[k] = int4_to_v4i(gload(C.pk, j));
                L.x[k] = C.x[j]; L.y[k] = C.y[j]; L.z[k] = C.z[j];
                L.q[k] = gload(C.q, j);
                L.r[k] = gload(C.r, j);
                L.rf[k] = gload(C.rf, j);
                L.t[k] = gload(C.t, j);
            }
        }
    }
    if constexpr (EW && !SKG) {
        for (int k = tid; k < C.K; k += nthr) {
            S[k] = make_double2(C.Q[2 * k], C.Q[2 * k + 1]);
            Ak[k] = C.aks[k];
            kid[k] = kpack(C.kn[3 * k], C.kn[3 * k + 1], C.kn[3 * k + 2]);
        }
    }
    // surface dipole M = sum q eff r (getSurfaceEnergy, ewald.h:278-292)
    int ew_cnt = EW ? *C.ew_cnt : 0;
    long n_trials = 0, n_acc = 0, n_inf = 0, n_ref = 0;
    const long long t_start = wall_clock64();   // chain's own wall time (stats slot 4)
    if (tid == 0)
        for (int q = 0; q < 5; ++q) sst[q] = 0.0;
    long rec = 0;
    if (EW && C.surf_pref != 0.0) {
        double mx = 0, my = 0, mz = 0;
        for (int j = tid; j < C.n; j += nthr) {
            double w = C.q[j] * C.eff[C.t[j]];
            mx += w * C.x[j]; my += w * C.y[j]; mz += w * C.z[j];
        }
        mx = block_sum(mx, red);
        my = block_sum(my, red);
        mz = block_sum(mz, red);
        if (tid == 0) { sst[0] = mx; sst[1] = my; sst[2] = mz; }
    }
    __syncthreads();

    int par = 0;
    // a launch may end inside a Propagator::move() call when its trial
    // budget runs out; the next launch resumes that call where it stopped
    const long budget = C.max_trials > 0 ? C.max_trials : LONG_MAX;
    long used = 0;
    int calls_done = 0;
    bool stop = false;
    int resume_rep = (C.resume && C.resume[0]) ? C.resume[1] : -1;
    for (int call = 0; call < C.ncalls && !stop; ++call) {
        int rep0 = 0;
        if (resume_rep >= 0) {
            if (tid == 0) T[par].go = 1;
            rep0 = resume_rep;
            resume_rep = -1;
        } else if (wave0) {
            // ---- Propagator::move (move.h:7887) + Movebase::move prologue
            (void)mt_range(I, 0, C.nmoves - 1);       // pick the move
            (void)mt_range(I, 0, 0);                  // randomMolId: mollist element
            bool run = mt_canonical(I) < C.mv.prob;   // Movebase::run (move.h:575-581)
            if (tid == 0) T[par].go = run ? 1 : 0;
        }
        __syncthreads();
        if (!T[par].go) { par ^= 1; ++calls_done; continue; }
        for (int rep = rep0; rep < C.repeat; ++rep) {
            if (used == budget) {
                stop = true;
                if (tid == 0 && C.resume) { C.resume[0] = 1; C.resume[1] = rep; }
                break;
            }
            ++used;
            TrialAT &tr = T[par];
            TrialAT &tn = T[par ^ 1];   // previous trial's record; next trial's prefetch
            // ---- AtomicTranslation::_trialMove (move.h:818-864), wave 0
            if (wave0) {
                TrialAT c;
                const bool pf = tr.pf;
                if (pf) {
                    c = tr;
                } else {
                    at_draw(c);
                    at_fetch(c);
                }
                // a prefetched position is stale if the previous trial moved i
                double ox = c.ox, oy = c.oy, oz = c.oz;
                int4 pk = c.pk;
                if (pf && tn.acc && tn.i == c.i) {
                    ox = tn.nx; oy = tn.ny; oz = tn.nz;
                    pk = make_int4(to_fixed(ox, C.box.inv[0]), to_fixed(oy, C.box.inv[1]),
                                   to_fixed(oz, C.box.inv[2]), pk.w);
                }
                const double dp = lds_grp ? ltdp[c.ti] : C.type_dp[c.ti];
                double nx, ny, nz;
                {
#pragma clang fp contract(off)
                    double tx = C.mv.dir[0] * dp, ty = C.mv.dir[1] * dp, tz = C.mv.dir[2] * dp;
                    tx *= c.t3[0] - 0.5;
                    ty *= c.t3[1] - 0.5;
                    tz *= c.t3[2] - 0.5;
                    nx = ox + tx; ny = oy + ty; nz = oz + tz;
                    boundary_d(C.box, nx, ny, nz);
                }
                if (tid == 0) {
                    if (!pf) {
                        tr.i = c.i; tr.ti = c.ti; tr.rf = c.rf; tr.slot = c.slot;
                        tr.q64 = c.q64; tr.r64 = c.r64; tr.wq = c.wq; tr.u = c.u;
                    }
                    tr.pf = 0;
                    tr.pk = pk;
                    tr.ox = ox; tr.oy = oy; tr.oz = oz;
                    tr.nx = nx; tr.ny = ny; tr.nz = nz;
                    tr.collide = collision_d(C.box, nx, ny, nz) ? 1 : 0;
                    tn.pf = 0;
                }
            }
            if (ldsp)
                lds_sync();
            else
                __syncthreads();
            const int i = tr.i;
            // next trial in this call and in this launch
            const bool ahead = rep + 1 < C.repeat && used < budget;
            double acc = 0.0;
            // ---- Ewald phase tables exp(i 2 pi n x / L) for old and new position
            if constexpr (EW) {
                const int ntab = 2 * 3 * kcc1;
                for (int e = tid; e < ntab; e += nthr) {
                    int which = e / (3 * kcc1), rem = e % (3 * kcc1);
                    int axis = rem / kcc1, nn = rem % kcc1;
                    double coord = which ? (axis == 0 ? tr.nx : axis == 1 ? tr.ny : tr.nz)
                                         : (axis == 0 ? tr.ox : axis == 1 ? tr.oy : tr.oz);
                    double sn, cs;
                    sincospi(2.0 * (double)nn * coord * C.box.inv[axis], &sn, &cs);
                    etab[which][axis * kcc1 + nn] = make_double2(cs, sn);
                }
            }
            // ---- draw-ahead: the next trial's random numbers and particle
            // data (wave 0; the reference's draw order is unchanged).  With a
            // cell list it overlaps the other waves' cell scan, otherwise the
            // reciprocal sum and the FP64 pass after the pair loop.
            auto commit_ahead = [&](const TrialAT &nxt) {
                if (tid == 0) {
                    tn.pf = 1;
                    tn.i = nxt.i; tn.ti = nxt.ti; tn.pk = nxt.pk; tn.rf = nxt.rf;
                    tn.slot = nxt.slot;
                    tn.q64 = nxt.q64; tn.r64 = nxt.r64; tn.wq = nxt.wq;
                    tn.t3[0] = nxt.t3[0]; tn.t3[1] = nxt.t3[1]; tn.t3[2] = nxt.t3[2];
                    tn.u = nxt.u;
                    tn.ox = nxt.ox; tn.oy = nxt.oy; tn.oz = nxt.oz;
                }
            };
            auto draw_ahead = [&]() {
                TrialAT nxt;
                at_draw(nxt);
                at_fetch(nxt);
                commit_ahead(nxt);
            };
            const bool use_cells = CELL && lcell_ok;
            if (use_cells && wave0 && ahead) draw_ahead();
            // ---- pair part: sum_{j != i} u(i', j) - u(i, j)
            const int nw = nthr >> 6, wv = tid >> 6;
            const int qcap = DCAP / nw;
            if constexpr (!MIX) {
                const double aq = tr.q64, ar = tr.r64;
                const int at = tr.ti;
                const double nx = tr.nx, ny = tr.ny, nz = tr.nz, ox = tr.ox, oy = tr.oy,
                             oz = tr.oz;
                for (int j = tid; j < C.n; j += nthr) {
                    if (j == i) continue;
                    acc += du_pair64<ELEC, SR>(C, tab, nx, ny, nz, ox, oy, oz, aq, ar, at, j);
                }
            } else {
                int nq = 0;
                int *dq = dlist + wv * qcap;
                bool done = false;
                if constexpr (LPOK) {
                    if (ldsp) {
                        // two rows per thread and batch: the LDS loop then
                        // needs no more registers than the device-memory one
                        acc = pair_loop_mixed<ELEC, SR, 2, FX, false, true>(
                            C, ltabf, ltab, i, tr.pk, tr.rf, tr.ti, tr.q64, tr.r64, tr.nx, tr.ny,
                            tr.nz, tr.ox, tr.oy, tr.oz, dq, qcap, nq, L);
                        done = true;
                    }
                }
                // (the fixed-unit form is only launched with LDS tables)
                if (done)
                    ;
                else if (use_cells)
                    acc = cell_loop<ELEC, SR>(C, ltab, lcheb, lcnt, i, tr.pk, tr.q64, tr.r64, tr.ti,
                                              tr.nx, tr.ny, tr.nz, tr.ox, tr.oy, tr.oz);
                else if (FX || lds_tab)
                    acc = pair_loop_mixed<ELEC, SR, U, FX, CELL>(C, ltabf, ltab, i, tr.pk, tr.rf, tr.ti,
                                                           tr.q64, tr.r64, tr.nx, tr.ny, tr.nz,
                                                           tr.ox, tr.oy, tr.oz, dq, qcap, nq);
                else
                    acc = pair_loop_mixed<ELEC, SR, U, false>(C, C.tabf, C.tab, i, tr.pk, tr.rf,
                                                              tr.ti, tr.q64, tr.r64, tr.nx, tr.ny,
                                                              tr.nz, tr.ox, tr.oy, tr.oz, dq, qcap,
                                                              nq);
                if ((tid & 63) == 0) nqs[wv] = nq;
            }
            if (EW || MIX) {   // etab and FP64 queues complete
                if (ldsp)
                    lds_sync();
                else
                    __syncthreads();
            }
            if (!use_cells && wave0 && ahead) draw_ahead();
            // FP64 work of the mixed path goes to the last F wavefronts, the
            // reciprocal-space sum to the others, so the two run side by side
            int Qn = 0, F = 0;
            if constexpr (MIX) {
                for (int w = 0; w < nw; ++w) Qn += nqs[w];
                F = min((Qn + 63) >> 6, EW ? nw / 2 : nw);
            }
            const int nk = 64 * (nw - F);    // threads on the reciprocal sum
            // ---- reciprocal part: A_k (|S+dS|^2 - |S|^2) per k, FP64
            const double wq = tr.wq;
            if constexpr (EW) {
                double de = 0.0;
                // four k per pass, their LDS reads issued before any arithmetic
                constexpr int KB = 4;
                int k0 = tid;
                for (; tid < nk && k0 + (KB - 1) * nk < C.K; k0 += KB * nk) {
                    int kv[KB];
                    double2 sk[KB];
                    double ak[KB];
#pragma unroll
                    for (int b = 0; b < KB; ++b) {
                        kv[b] = kid[k0 + b * nk];
                        sk[b] = S[k0 + b * nk];
                        ak[b] = Ak[k0 + b * nk];
                    }
                    double2 d[KB];
#pragma unroll
                    for (int b = 0; b < KB; ++b) {
                        int nx, ny, nz;
                        kunpack(kv[b], nx, ny, nz);
                        double er, ei, orr, oi;
                        ew_phase(etab[1], kcc1, nx, ny, nz, er, ei);
                        ew_phase(etab[0], kcc1, nx, ny, nz, orr, oi);
                        const double dr = wq * (er - orr), di = wq * (ei - oi);
                        d[b] = make_double2(dr, di);
                        de += ak[b] * (2.0 * (sk[b].x * dr + sk[b].y * di) + (dr * dr + di * di));
                    }
#pragma unroll
                    for (int b = 0; b < KB; ++b) dS[k0 + b * nk] = d[b];
                }
                for (int k = k0; tid < nk && k < C.K; k += nk) {
                    int nx, ny, nz;
                    kunpack(kid[k], nx, ny, nz);
                    double er, ei, orr, oi;
                    ew_phase(etab[1], kcc1, nx, ny, nz, er, ei);
                    ew_phase(etab[0], kcc1, nx, ny, nz, orr, oi);
                    double dr = wq * (er - orr), di = wq * (ei - oi);
                    double2 sk = S[k];
                    dS[k] = make_double2(dr, di);
                    de += Ak[k] * (2.0 * (sk.x * dr + sk.y * di) + (dr * dr + di * di));
                }
                acc += C.recip_pref * de;
            }
            // ---- FP64 pass of the mixed path over all queued partners
            if constexpr (MIX) {
                for (int e = tid - nk; e >= 0 && e < Qn; e += 64 * F) {
                    int w = 0, k = e;
                    while (k >= nqs[w]) { k -= nqs[w]; ++w; }
                    const int j = dlist[w * qcap + k];
                    if (LPOK && ldsp)
                        acc += du_pair64_v<ELEC, SR>(C, tab, tr.nx, tr.ny, tr.nz, tr.ox, tr.oy,
                                                     tr.oz, tr.q64, tr.r64, tr.ti, L.x[j], L.y[j],
                                                     L.z[j], L.q[j], L.r[j], L.t[j]);
                    else
                        acc += du_pair64_mixed<ELEC, SR>(C, tab, lcheb, tr.nx, tr.ny, tr.nz,
                                                         tr.ox, tr.oy, tr.oz, tr.q64, tr.r64,
                                                         tr.ti, j);
                }
            }
            double du;
            if constexpr (EW) {
                du = block_sum(acc, red);
            } else {
                // one barrier: the two buffers alternate, and the next
                // trial's first barrier separates their uses
                du = block_sum1(acc, red2[rpar], ldsp);
                rpar ^= 1;
            }
            if (EW && C.surf_pref != 0.0) {
                const double Mx = sst[0], My = sst[1], Mz = sst[2];
                double dMx = wq * (tr.nx - tr.ox), dMy = wq * (tr.ny - tr.oy), dMz = wq * (tr.nz - tr.oz);
                double Nx = Mx + dMx, Ny = My + dMy, Nz = Mz + dMz;
                du += C.surf_pref * ((Nx * Nx + (Ny * Ny + Nz * Nz)) - (Mx * Mx + (My * My + Mz * Mz)));
            }
            if (tr.collide) du = __builtin_inf();   // energy.h:3597-3607
            // ---- Movebase::metropolis (move.h:567-573): the uniform was drawn
            // with the trial (it is the next number of the same stream); every
            // thread takes the same decision from the same du
            // (du <= 0: exp(-du) >= 1 > u, so the test accepts without exp)
            const bool accept = du <= 0.0 || !(tr.u > exp(-du));
            if (tid == 0) tr.acc = accept ? 1 : 0;
            n_trials++;
            if (tid == 0 && C.du_rec && rec < C.cap) {
                C.du_rec[rec] = du;
                C.acc_rec[rec] = accept;
            }
            rec++;
            if (accept) {
                n_acc++;
                if (tid == 0) { sst[4] += du; sst[3] += du; }
                if (tid == 0) {   // AtomicTranslation::_acceptMove (move.h:866-877)
                    C.x[i] = tr.nx; C.y[i] = tr.ny; C.z[i] = tr.nz;
                    const int4 np = make_int4(to_fixed(tr.nx, C.box.inv[0]),
                                              to_fixed(tr.ny, C.box.inv[1]),
                                              to_fixed(tr.nz, C.box.inv[2]), tr.pk.w);
                    C.pk_w[i] = np;
                    if constexpr (LPOK) {
                        const int il = i - L.j0;
                        if (ldsp && il >= 0 && il < L.ns) {
                            L.pk[il] = int4_to_v4i(np);
                            L.x[il] = tr.nx; L.y[il] = tr.ny; L.z[il] = tr.nz;
                        }
                    }
                    if (CELL && lcell_ok) {
                        // the moved record changes in place, or moves to its
                        // new cell (the old cell's last record fills the hole)
                        const int nc = C.cell_nc, cap = C.cell_cap;
                        const int co = cell_of(tr.pk.x, tr.pk.y, tr.pk.z, nc);
                        const int cn = cell_of(np.x, np.y, np.z, nc);
                        // (slots are fetched with the trial; a prefetched
                        // next trial's slot is corrected here when it moves)
                        const int so = tr.slot;
                        const double4 rec = make_double4(tr.nx, tr.ny, tr.nz, tr.q64);
                        if (co == cn) {
                            C.cell_rec[so] = rec;
                        } else if (lcnt[cn] >= cap) {
                            // full cell: the chain continues on the all-pairs screen
                            lcell_ok = 0;
                        } else {
                            const int last = co * cap + (--lcnt[co]);
                            if (so != last) {
                                const double4 m = C.cell_rec[last];
                                const int mj = C.cell_idx[last];
                                C.cell_rec[so] = m;
                                C.cell_idx[so] = mj;
                                C.cell_slot[mj] = so;
                                if (tn.pf && tn.i == mj) tn.slot = so;
                            }
                            const int sn = cn * cap + (lcnt[cn]++);
                            C.cell_rec[sn] = rec;
                            C.cell_idx[sn] = i;
                            C.cell_slot[i] = sn;
                            if (tn.pf && tn.i == i) tn.slot = sn;
                        }
                    }
                }
                if constexpr (EW) {
                    if (tid == 0) {
                        sst[0] += wq * (tr.nx - tr.ox);
                        sst[1] += wq * (tr.ny - tr.oy);
                        sst[2] += wq * (tr.nz - tr.oz);
                    }
                    // S(k) += dS(k) (dS complete: block_sum synchronised)
                    for (int k = tid; k < C.K; k += nthr) {
                        const double2 sk = S[k], d = dS[k];
                        S[k] = make_double2(sk.x + d.x, sk.y + d.y);
                    }
                    // NonbondedEwald::update (ewald.h:626-643): full recompute of
                    // S(k) every update_frequency accepted moves; the returned
                    // reciprocal-energy difference enters utot
                    if (++ew_cnt > C.update_frequency - 1) {
                        __syncthreads();   // committed x[i] visible
                        double eb = 0.0;
                        for (int k = tid; k < C.K; k += nthr)
                            eb += Ak[k] * (S[k].x * S[k].x + S[k].y * S[k].y);
                        eb = block_sum(eb, red);
                        for (int k = tid; k < C.K; k += nthr) S[k] = make_double2(0.0, 0.0);
                        const bool tabled = kcc1 <= REFRESH_KAX;
                        const int per = 3 * kcc1;
                        for (int j0 = 0; j0 < C.n; j0 += REFRESH_CHUNK) {
                            const int nj = min(REFRESH_CHUNK, C.n - j0);
                            __syncthreads();
                            if (tabled) {
                                for (int e = tid; e < nj * per; e += nthr) {
                                    int jj = e / per, rem = e % per;
                                    int axis = rem / kcc1, nn = rem % kcc1;
                                    int j = j0 + jj;
                                    double coord = axis == 0 ? C.x[j] : axis == 1 ? C.y[j] : C.z[j];
                                    double sn, cs;
                                    sincospi(2.0 * (double)nn * coord * C.box.inv[axis], &sn, &cs);
                                    rtab[jj * per + axis * kcc1 + nn] = make_double2(cs, sn);
                                }
                            }
                            for (int jj = tid; jj < nj; jj += nthr)
                                rw[jj] = C.q[j0 + jj] * C.eff[C.t[j0 + jj]];
                            __syncthreads();
                            for (int k = tid; k < C.K; k += nthr) {
                                int nx, ny, nz;
                                kunpack(kid[k], nx, ny, nz);
                                double sr = 0.0, si = 0.0;
                                for (int jj = 0; jj < nj; ++jj) {
                                    double er, ei;
                                    if (tabled) {
                                        ew_phase(&rtab[jj * per], kcc1, nx, ny, nz, er, ei);
                                    } else {
                                        int j = j0 + jj;
                                        double ph = 2.0 * ((double)nx * C.x[j] * C.box.inv[0] +
                                                           (double)ny * C.y[j] * C.box.inv[1] +
                                                           (double)nz * C.z[j] * C.box.inv[2]);
                                        sincospi(ph, &ei, &er);
                                    }
                                    sr += rw[jj] * er;
                                    si += rw[jj] * ei;
                                }
                                double2 sk = S[k];
                                S[k] = make_double2(sk.x + sr, sk.y + si);
                            }
                        }
                        double ea = 0.0;
                        for (int k = tid; k < C.K; k += nthr)
                            ea += Ak[k] * (S[k].x * S[k].x + S[k].y * S[k].y);
                        ea = block_sum(ea, red);
                        if (C.surf_pref != 0.0) {
                            double mx = 0, my = 0, mz = 0;
                            for (int j = tid; j < C.n; j += nthr) {
                                double w2 = C.q[j] * C.eff[C.t[j]];
                                mx += w2 * C.x[j]; my += w2 * C.y[j]; mz += w2 * C.z[j];
                            }
                            mx = block_sum(mx, red);
                            my = block_sum(my, red);
                            mz = block_sum(mz, red);
                            if (tid == 0) { sst[0] = mx; sst[1] = my; sst[2] = mz; }
                        }
                        if (tid == 0) sst[3] += C.recip_pref * (ea - eb);
                        ew_cnt = 0;
                        n_ref++;
                    }
                }
            } else if (du == __builtin_inf()) {
                n_inf++;
            }
            par ^= 1;
        }
        if (!stop) ++calls_done;
    }
    if (!stop && tid == 0 && C.resume) C.resume[0] = 0;
    // ---- write back chain state
    __syncthreads();
    for (int k = tid; k < 624; k += nthr) {
        C.rng[k] = mtg[k];
        C.rng[625 + k] = mti[k];
    }
    if (tid == 0) {
        C.rng[624] = (uint32_t)G.p;
        C.rng[625 + 624] = (uint32_t)I.p;
        C.stats_i[0] = n_trials;
        C.stats_i[1] = n_acc;
        C.stats_i[2] = n_inf;
        C.stats_i[3] = n_ref;
        C.stats_i[4] = wall_clock64() - t_start;
        C.stats_i[5] = calls_done;
        C.stats_i[6] = (CELL && C.cell_nc > 0 && !lcell_ok) ? 1 : 0;
        C.stats_d[0] = sst[3];
        C.stats_d[1] = sst[4];
        if (EW) *C.ew_cnt = ew_cnt;
    }
    if constexpr (EW && !SKG) {
        for (int k = tid; k < C.K; k += nthr) { C.Q[2 * k] = S[k].x; C.Q[2 * k + 1] = S[k].y; }
    }
    if constexpr (CELL) {
        if (C.cell_nc > 0) {
            const int ncell = C.cell_nc * C.cell_nc * C.cell_nc;
            for (int c = tid; c < ncell; c += nthr) C.cell_cnt[c] = lcnt[c];
            if (tid == 0) C.cell_cnt[ncell] = lcell_ok ? 0 : 1;
        }
    }
}

// ------------------------------------------------------------------------
// TranslateRotate sweep (one workgroup per chain; moved group staged in LDS)
// ------------------------------------------------------------------------
struct TrialTR {
    int g, go, acc, collide, rot, trans;
    double qw, qx, qy, qz;           // quaternion (Eigen AngleAxis -> Quaternion)
    double ox, oy, oz;               // rotation origin
    double px, py, pz;               // translation
    double cmx, cmy, cmz;            // cm_trial after the move
};

__device__ __forceinline__ void cross3(double a0, double a1, double a2, double b0, double b1,
                                       double b2, double &c0, double &c1, double &c2)
{
#pragma clang fp contract(off)
    c0 = a1 * b2 - a2 * b1;
    c1 = a2 * b0 - a0 * b2;
    c2 = a0 * b1 - a1 * b0;
}

// QuaternionRotate::operator() (geometry.h:130-139) with Eigen 3.2.9's
// _transformVector: uv = 2 vec x v; v' = v + w uv + vec x uv
__device__ __forceinline__ void qrotate(const Box &b, const TrialTR &T, double &x, double &y,
                                        double &z)
{
#pragma clang fp contract(off)
    double ax = x - T.ox, ay = y - T.oy, az = z - T.oz;
    boundary_d(b, ax, ay, az);
    double u0, u1, u2, c0, c1, c2;
    cross3(T.qx, T.qy, T.qz, ax, ay, az, u0, u1, u2);
    u0 += u0; u1 += u1; u2 += u2;
    cross3(T.qx, T.qy, T.qz, u0, u1, u2, c0, c1, c2);
    x = ((ax + T.qw * u0) + c0) + T.ox;
    y = ((ay + T.qw * u1) + c1) + T.oy;
    z = ((az + T.qw * u2) + c2) + T.oz;
    boundary_d(b, x, y, z);
}

__device__ __forceinline__ void translate_pt(const Box &b, const TrialTR &T, double &x, double &y,
                                             double &z)
{
#pragma clang fp contract(off)
    x += T.px; y += T.py; z += T.pz;
    boundary_d(b, x, y, z);
}

// ------------------------------------------------------------------------
// Mixed-path group energies with the group-pair energy cache
// (EnergyMatrix, energy.h:225-641: g2All of the trial state is evaluated and
// kept per partner group; the accepted state's g2g values are read from the
// matrix instead of being recomputed, and update(true) writes the new row).
// ------------------------------------------------------------------------
static constexpr int MAT_MAXG = 256;   // groups per chain for the cached path
static constexpr int TR_DW = 128;      // queued FP64 pairs per wavefront

// Adds e into slot[h] for each distinct h >= 0 among the lanes of the
// wavefront, lowest lane's group first (fixed order, so deterministic).
// Called by all 64 lanes.
__device__ __forceinline__ void seg_add(double e, int h, double *slot)
{
    const int lane = threadIdx.x & 63;
    uint64_t pending = __ballot(h >= 0);
    while (pending) {
        const int lead = __builtin_ctzll(pending);
        const int hl = __shfl(h, lead, 64);
        const bool mine = (h == hl);
        const double v = wave_sum(mine ? e : 0.0);
        if (lane == 0) slot[hl] += v;
        pending &= ~__ballot(mine);
    }
}

// Staged-site record of the cached path.  Generic form: fixed-point x, y, z,
// FP32 charge (sp) and FP32 radius (srf).  Fixed-unit form (ChainDev::fx, cubic
// boxes with Coulomb/Debye-Hueckel and hard spheres): the charge is stored as
// lB q / s and the radius as r c / s, where s = L / 2^32 is the fixed-point unit
// and c = sqrt(1 + 1e-6) widens the hard-sphere test, so pair distances stay
// in fixed-point units and no per-pair rescaling is needed.
__device__ __forceinline__ void stage_site(const ChainDev &C, int i, int32_t fx, int32_t fy,
                                           int32_t fz, int4 *sp, float *srf, int k)
{
    const float q = __int_as_float(C.pk[i].w);
    if (C.fx) {
        sp[k] = make_int4(fx, fy, fz, __float_as_int(q * C.fx_lbs));
        srf[k] = C.rf[i] * C.fx_rs;
    } else {
        sp[k] = make_int4(fx, fy, fz, __float_as_int(q));
        srf[k] = C.rf[i];
    }
}

template <int ELEC, int SR>
__device__ __forceinline__ double pair_exact_new(const ChainDev &C, const double *__restrict__ tab,
                                                 int j, int s, const double *sx, const double *sy,
                                                 const double *sz, const double *sq,
                                                 const double *sr, const int *st)
{
    double r2;
    if (C.pot.nbkind == NB_VECTOR) {
        double rx, ry, rz;
        vdist_d(C.box, sx[s], sy[s], sz[s], C.x[j], C.y[j], C.z[j], rx, ry, rz);
        r2 = rx * rx + (ry * ry + rz * rz);
    } else {
        r2 = sqdist_d(C.box, sx[s], sy[s], sz[s], C.x[j], C.y[j], C.z[j]);
    }
    return pair_u64_t<ELEC, SR>(C.pot, tab, r2, sq[s], C.q[j], sr[s], C.r[j], st[s], C.t[j]);
}

static constexpr int TR_B = 8;   // staged sites per inner-loop batch

// Staged group (see stage_site; FP64 copies sx..sr for the exact pass)
// against every site outside [front, back].  Energies are summed per partner
// group into this wavefront's row of wslot (wslot[w * ng + h]); keep[h] == 0
// drops group h (NonbondedCutg2g).  FP32 pairs are accumulated in FP64, TR_B
// sites per batch with the site records read from LDS up front; a pair that
// must be exact (pair_fx) is queued per wavefront and evaluated in FP64
// afterwards.  Must be called by whole wavefronts; the caller clears wslot
// before and synchronises after.
template <int ELEC, int SR, bool FX>
__device__ __forceinline__ void group_energies_mixed_t(const ChainDev &C, const float *__restrict__ tabf,
                                       const double *__restrict__ tab, int front, int back, int ns,
                                       const int4 *sp, const float *srf, const int *st,
                                       const double *sx, const double *sy, const double *sz,
                                       const double *sq, const double *sr,
                                       const unsigned char *keep, double *wslot, int *dl)
{
    const int tid = threadIdx.x, nthr = blockDim.x, lane = tid & 63, w = tid >> 6;
    double *slot = wslot + w * C.ng;
    int *mydl = dl + w * TR_DW;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const float fsx = C.box.fscale[0], fsy = C.box.fscale[1], fsz = C.box.fscale[2];
    const float exact = C.pot.exact_f;
    const int nfull = ns - ns % TR_B;
    int nd = 0;   // wavefront-uniform queue length
    const int iters = (C.n + nthr - 1) / nthr;

    // queue (j, s) for the exact pass, or evaluate it at once when full
    auto enqueue = [&](bool d, int j, int s, double &e) {
        const uint64_t m = __ballot(d);
        if (m) {
            if (d) {
                const int k = nd + __popcll(m & below);
                if (k < TR_DW)
                    mydl[k] = (j << 9) | s;
                else
                    e += pair_exact_new<ELEC, SR>(C, tab, j, s, sx, sy, sz, sq, sr, st);
            }
            nd += __popcll(m);
        }
    };
    // one pair in FP32 (either form)
    auto pair = [&](const int4 a, float ra, int ta, const int4 pj, float bq, float rb, int bt,
                    bool &bad) -> float {
        if constexpr (FX) {
            const float dx = (float)(int32_t)(a.x - pj.x);
            const float dy = (float)(int32_t)(a.y - pj.y);
            const float dz = (float)(int32_t)(a.z - pj.z);
            return pair_fx<ELEC, SR>(C, dx * dx + dy * dy + dz * dz, __int_as_float(a.w), bq, ra,
                                     rb, exact, bad);
        } else {
            const float dx = (float)(int32_t)(a.x - pj.x) * fsx;
            const float dy = (float)(int32_t)(a.y - pj.y) * fsy;
            const float dz = (float)(int32_t)(a.z - pj.z) * fsz;
            const float u = pair_u32_t<ELEC, SR>(C.pot, tabf, dx * dx + dy * dy + dz * dz,
                                                 __int_as_float(a.w), bq, ra, rb, ta, bt);
            bad = !(fabsf(u) <= exact);
            return u;
        }
    };
    constexpr bool NEED_T = (SR == SR_LJ || ELEC == ELEC_EWALD);
    for (int it = 0; it < iters; ++it) {
        const int j = it * nthr + tid;
        int h = -1;
        if (j < C.n && (j < front || j > back)) {
            h = C.gid[j];
            if (keep && !keep[h]) h = -1;
        }
        if (__ballot(h >= 0) == 0) continue;
        const bool act = h >= 0;
        const int jc = j < C.n ? j : C.n - 1;
        const int4 pj = C.pk[jc];
        const float bq = __int_as_float(pj.w);
        const float br = (SR == SR_HS) ? C.rf[jc] * (FX ? C.fx_rs : 1.f) : 0.f;
        const int bt = NEED_T ? C.t[jc] : 0;
        double e = 0.0;
        for (int s0 = 0; s0 < nfull; s0 += TR_B) {
            int4 a[TR_B];
            float ra[TR_B];
            int ta[TR_B];
#pragma unroll
            for (int b = 0; b < TR_B; ++b) {
                a[b] = sp[s0 + b];
                ra[b] = (SR == SR_HS) ? srf[s0 + b] : 0.f;
                ta[b] = NEED_T ? st[s0 + b] : 0;
            }
            unsigned dm = 0;
#pragma unroll
            for (int b = 0; b < TR_B; ++b) {
                bool bad;
                const float u = pair(a[b], ra[b], ta[b], pj, bq, br, bt, bad);
                // an inactive lane's sum is discarded by seg_add
                dm |= ((bad && act) ? 1u : 0u) << b;
                e += (double)(bad ? 0.f : u);
            }
            if (__ballot(dm != 0)) {
                for (int b = 0; b < TR_B; ++b) enqueue((dm >> b) & 1u, j, s0 + b, e);
            }
        }
        for (int s = nfull; s < ns; ++s) {
            bool bad;
            const float u = pair(sp[s], (SR == SR_HS) ? srf[s] : 0.f, NEED_T ? st[s] : 0, pj, bq,
                                 br, bt, bad);
            e += (double)(bad ? 0.f : u);
            enqueue(bad && act, j, s, e);
        }
        seg_add(e, h, slot);
    }
    // exact pass over this wavefront's queue
    wave_sync_lds();
    const int nq = min(nd, TR_DW);
    for (int k0 = 0; k0 < nq; k0 += 64) {
        const int k = k0 + lane;
        int h = -1;
        double e = 0.0;
        if (k < nq) {
            const int v = mydl[k];
            const int j = v >> 9, s = v & 511;
            e = pair_exact_new<ELEC, SR>(C, tab, j, s, sx, sy, sz, sq, sr, st);
            h = C.gid[j];
        }
        seg_add(e, h, slot);
    }
}

template <int ELEC, int SR>
__device__ __forceinline__ void group_energies_mixed(const ChainDev &C, const float *__restrict__ tabf,
                                     const double *__restrict__ tab, int front, int back, int ns,
                                     const int4 *sp, const float *srf, const int *st,
                                     const double *sx, const double *sy, const double *sz,
                                     const double *sq, const double *sr, const unsigned char *keep,
                                     double *wslot, int *dl)
{
    constexpr bool FX_OK = (SR == SR_NONE || SR == SR_HS) &&
                           (ELEC == ELEC_NONE || ELEC == ELEC_COULOMB || ELEC == ELEC_DH);
    if (FX_OK && C.fx)
        group_energies_mixed_t<ELEC, SR, FX_OK>(C, tabf, tab, front, back, ns, sp, srf, st, sx, sy,
                                                sz, sq, sr, keep, wslot, dl);
    else
        group_energies_mixed_t<ELEC, SR, false>(C, tabf, tab, front, back, ns, sp, srf, st, sx, sy,
                                                sz, sq, sr, keep, wslot, dl);
}

// Fills row g of every chain's cache with the accepted state's g2g energies
// (grid: groups x chains).  Rows of groups larger than the LDS stage are left
// untouched: such groups are never moved by the cached sweep.
template <int ELEC, int SR>
__global__ __launch_bounds__(AT_THREADS) void k_emat_init(const ChainDev *__restrict__ chains)
{
    const ChainDev &C = chains[blockIdx.y];
    const int g = blockIdx.x, tid = threadIdx.x, nthr = blockDim.x;
    if (g >= C.ng || !C.emat) return;
    const GroupDesc gd = C.groups[g];
    const int ns = gd.back - gd.front + 1;
    if (ns > MAX_GROUP_SITES) return;
    __shared__ int4 sp[MAX_GROUP_SITES];
    __shared__ float srf[MAX_GROUP_SITES];
    __shared__ int st[MAX_GROUP_SITES];
    __shared__ double sx[MAX_GROUP_SITES], sy[MAX_GROUP_SITES], sz[MAX_GROUP_SITES];
    __shared__ double sq[MAX_GROUP_SITES], sr[MAX_GROUP_SITES];
    __shared__ unsigned char keep[MAT_MAXG];
    __shared__ double wslot[(AT_THREADS / 64) * MAT_MAXG];
    __shared__ int dl[(AT_THREADS / 64) * TR_DW];
    for (int k = tid; k < ns; k += nthr) {
        const int i = gd.front + k;
        const int4 p4 = C.pk[i];
        stage_site(C, i, p4.x, p4.y, p4.z, sp, srf, k);
        st[k] = C.t[i];
        sx[k] = C.x[i]; sy[k] = C.y[i]; sz[k] = C.z[i];
        sq[k] = C.q[i]; sr[k] = C.r[i];
    }
    for (int h = tid; h < C.ng; h += nthr) {
        unsigned char kn = 1;
        if (C.pot.nbkind == NB_CUTG2G) {
            const GroupDesc hd = C.groups[h];
            if (hd.molsize == hd.back - hd.front + 1 && gd.molsize == ns)
                kn = sqdist_d(C.box, C.cm[6 * g], C.cm[6 * g + 1], C.cm[6 * g + 2], C.cm[6 * h],
                              C.cm[6 * h + 1], C.cm[6 * h + 2]) > C.pot.g2g_rc2 ? 0 : 1;
        }
        keep[h] = kn;
    }
    for (int k = tid; k < (nthr >> 6) * C.ng; k += nthr) wslot[k] = 0.0;
    __syncthreads();
    group_energies_mixed<ELEC, SR>(C, C.tabf, C.tab, gd.front, gd.back, ns, sp, srf, st, sx, sy, sz,
                                   sq, sr, keep, wslot, dl);
    __syncthreads();
    for (int h = tid; h < C.ng; h += nthr) {
        double e = 0.0;
        for (int w = 0; w < (nthr >> 6); ++w) e += wslot[w * C.ng + h];
        C.emat[(size_t)g * C.ng + h] = (h == g) ? 0.0 : e;
    }
}

template <int ELEC, int SR, bool MIX>
__global__ __launch_bounds__(AT_THREADS) void k_sweep_tr(const ChainDev *__restrict__ chains, int ncalls)
{
    const ChainDev &C = chains[blockIdx.x];
    const int tid = threadIdx.x, nthr = blockDim.x;
    const bool wave0 = tid < 64;
    __shared__ uint32_t mtg[624], mti[624];
    __shared__ double red[16];
    __shared__ TrialTR T[2];
    __shared__ double sn[3][MAX_GROUP_SITES], so[3][MAX_GROUP_SITES];
    __shared__ int32_t snf[3][MAX_GROUP_SITES], sof[3][MAX_GROUP_SITES];
    __shared__ double sq[MAX_GROUP_SITES], sr[MAX_GROUP_SITES];
    __shared__ int st[MAX_GROUP_SITES];
    __shared__ unsigned char keep_n[MAX_GROUPS_LDS], keep_o[MAX_GROUPS_LDS];
    __shared__ int s_collide;
    // cached path: packed new sites, per-wavefront group sums, FP64 queue
    __shared__ int4 snp[MIX ? MAX_GROUP_SITES : 1];
    __shared__ float srf[MIX ? MAX_GROUP_SITES : 1];
    __shared__ double wslot[MIX ? (AT_THREADS / 64) * MAT_MAXG : 1];
    __shared__ int dl[MIX ? (AT_THREADS / 64) * TR_DW : 1];
    __shared__ double enew[MIX ? MAT_MAXG : 1];
    const bool cached = MIX && C.emat != nullptr;

    for (int k = tid; k < 624; k += nthr) {
        mtg[k] = C.rng[k];
        mti[k] = C.rng[625 + k];
    }
    WaveMT G{mtg, (int)C.rng[624]}, I{mti, (int)C.rng[625 + 624]};
    long n_trials = 0, n_acc = 0, n_inf = 0;
    const long long t_start = wall_clock64();
    double utot = 0.0, dusum = 0.0;
    long rec = 0;
    __syncthreads();
    int par = 0;
    const long budget = C.max_trials > 0 ? C.max_trials : LONG_MAX;
    long used = 0;
    int calls_done = 0;
    bool stop = false;
    int resume_rep = (C.resume && C.resume[0]) ? C.resume[1] : -1;
    for (int call = 0; call < C.ncalls && !stop; ++call) {
        int rep0 = 0;
        if (resume_rep >= 0) {
            if (tid == 0) T[par].go = 1;
            rep0 = resume_rep;
            resume_rep = -1;
        } else if (wave0) {
            (void)mt_range(I, 0, C.nmoves - 1);
            (void)mt_range(I, 0, 0);
            bool run = mt_canonical(I) < C.mv.prob;
            if (tid == 0) T[par].go = run ? 1 : 0;
        }
        __syncthreads();
        if (!T[par].go) { par ^= 1; ++calls_done; continue; }
        for (int rep = rep0; rep < C.repeat; ++rep) {
            if (used == budget) {
                stop = true;
                if (tid == 0 && C.resume) { C.resume[0] = 1; C.resume[1] = rep; }
                break;
            }
            ++used;
            TrialTR &tr = T[par];
            // ---- TranslateRotate::_trialMove (move.h:1620-1686), wave 0
            if (wave0) {
#pragma clang fp contract(off)
                int g = C.gvec[mt_range(G, 0, C.ngvec - 1)];
                GroupDesc gd = C.groups[g];
                double cmx = C.cm[6 * g + 3], cmy = C.cm[6 * g + 4], cmz = C.cm[6 * g + 5];
                double qw = 1, qx = 0, qy = 0, qz = 0, orx = 0, ory = 0, orz = 0;
                int rot = 0, trans = 0;
                double px = 0, py = 0, pz = 0;
                if (C.mv.dprot > 1e-6) {
                    // PointBase::ranunit (point.h:74-89)
                    double v0, v1, v2, r2;
                    do {
                        v0 = (2 * mt_canonical(G) - 1) * C.mv.dir2[0];
                        v1 = (2 * mt_canonical(G) - 1) * C.mv.dir2[1];
                        v2 = (2 * mt_canonical(G) - 1) * C.mv.dir2[2];
                        r2 = v0 * v0 + (v1 * v1 + v2 * v2);
                    } while (r2 > 1);
                    double nrm = sqrt(r2);
                    v0 = v0 / nrm; v1 = v1 / nrm; v2 = v2 / nrm;
                    double cx = C.cm[6 * g], cy = C.cm[6 * g + 1], cz = C.cm[6 * g + 2];
                    if (!C.mv.center_rotation) {
                        int site = (mt_canonical(G) < 0.5) ? gd.front : gd.back;
                        cx = C.x[site]; cy = C.y[site]; cz = C.z[site];
                    }
                    double ex = cx + v0, ey = cy + v1, ez = cz + v2;
                    double angle = C.mv.dprot * (mt_canonical(G) - 0.5);
                    // QuaternionRotate::setAxis (src/geometry.cpp:753-766)
                    double ux = ex - cx, uy = ey - cy, uz = ez - cz;
                    boundary_d(C.box, ux, uy, uz);
                    double un = sqrt(ux * ux + (uy * uy + uz * uz));
                    ux /= un; uy /= un; uz /= un;
                    double ha = 0.5 * angle;
                    qw = cos(ha);
                    double sh = sin(ha);
                    qx = sh * ux; qy = sh * uy; qz = sh * uz;
                    orx = cx; ory = cy; orz = cz;
                    rot = 1;
                }
                if (C.mv.dp > 1e-6) {
                    px = C.mv.dir[0] * C.mv.dp * (mt_canonical(G) - 0.5);
                    py = C.mv.dir[1] * C.mv.dp * (mt_canonical(G) - 0.5);
                    pz = C.mv.dir[2] * C.mv.dp * (mt_canonical(G) - 0.5);
                    trans = 1;
                }
                if (tid == 0) {
                    tr.g = g; tr.rot = rot; tr.trans = trans;
                    tr.qw = qw; tr.qx = qx; tr.qy = qy; tr.qz = qz;
                    tr.ox = orx; tr.oy = ory; tr.oz = orz;
                    tr.px = px; tr.py = py; tr.pz = pz;
                    // Group::newrotate rotates cm_trial; Group::translate moves it
                    if (rot) qrotate(C.box, tr, cmx, cmy, cmz);
                    if (trans) translate_pt(C.box, tr, cmx, cmy, cmz);
                    tr.cmx = cmx; tr.cmy = cmy; tr.cmz = cmz;
                    s_collide = 0;
                }
            }
            __syncthreads();
            const int g = tr.g;
            const GroupDesc gd = C.groups[g];
            const int ns = gd.back - gd.front + 1;
            // ---- move the sites (Group::newrotate / Group::translate)
            for (int s = tid; s < ns; s += nthr) {
                int i = gd.front + s;
                double x = C.x[i], y = C.y[i], z = C.z[i];
                so[0][s] = x; so[1][s] = y; so[2][s] = z;
                const int4 pi4 = C.pk[i];
                sof[0][s] = pi4.x; sof[1][s] = pi4.y; sof[2][s] = pi4.z;
                sq[s] = C.q[i]; sr[s] = C.r[i]; st[s] = C.t[i];
                if (tr.rot) qrotate(C.box, tr, x, y, z);
                if (tr.trans) translate_pt(C.box, tr, x, y, z);
                sn[0][s] = x; sn[1][s] = y; sn[2][s] = z;
                snf[0][s] = to_fixed(x, C.box.inv[0]);
                snf[1][s] = to_fixed(y, C.box.inv[1]);
                snf[2][s] = to_fixed(z, C.box.inv[2]);
                if (MIX) {
                    stage_site(C, i, snf[0][s], snf[1][s], snf[2][s], snp, srf, s);
                }
                if (collision_d(C.box, x, y, z)) s_collide = 1;
            }
            if (cached)
                for (int k = tid; k < (nthr >> 6) * C.ng; k += nthr) wslot[k] = 0.0;
            // ---- NonbondedCutg2g keep flags per other group (energy.h:1343-1358)
            for (int h = tid; h < C.ng; h += nthr) {
                unsigned char kn = 1, ko = 1;
                if (C.pot.nbkind == NB_CUTG2G) {
                    const GroupDesc hd = C.groups[h];
                    bool mol_h = (hd.molsize == hd.back - hd.front + 1);
                    bool mol_g = (gd.molsize == ns);
                    if (mol_h && mol_g) {
                        double hx = C.cm[6 * h], hy = C.cm[6 * h + 1], hz = C.cm[6 * h + 2];
                        kn = sqdist_d(C.box, tr.cmx, tr.cmy, tr.cmz, hx, hy, hz) > C.pot.g2g_rc2 ? 0 : 1;
                        ko = sqdist_d(C.box, C.cm[6 * g], C.cm[6 * g + 1], C.cm[6 * g + 2], hx, hy,
                                      hz) > C.pot.g2g_rc2 ? 0 : 1;
                    }
                }
                keep_n[h] = kn;
                keep_o[h] = ko;
            }
            __syncthreads();
            double acc = 0.0;
            if (cached) {
                // g2All of the trial state per partner group; the accepted
                // state's values come from the cache
                group_energies_mixed<ELEC, SR>(C, C.tabf, C.tab, gd.front, gd.back, ns, snp, srf, st,
                                               sn[0], sn[1], sn[2], sq, sr, keep_n, wslot, dl);
                __syncthreads();
                for (int h = tid; h < C.ng; h += nthr) {
                    double e = 0.0;
                    for (int w = 0; w < (nthr >> 6); ++w) e += wslot[w * C.ng + h];
                    enew[h] = e;
                    if (h != g) acc += e - C.emat[(size_t)g * C.ng + h];
                }
            }
            for (int j = tid; j < (cached ? 0 : C.n); j += nthr) {
                if (j >= gd.front && j <= gd.back) continue;
                int h = C.gid[j];
                bool kn = keep_n[h], ko = keep_o[h];
                if (!kn && !ko) continue;
                if (!MIX) {
                    double bx = C.x[j], by = C.y[j], bz = C.z[j], bq = C.q[j], br = C.r[j];
                    int bt = C.t[j];
                    const bool vec = (C.pot.nbkind == NB_VECTOR);
                    for (int s = 0; s < ns; ++s) {
                        double un = 0.0, uo = 0.0;
                        if (kn) {
                            double r2;
                            if (vec) {
                                double rx, ry, rz;
                                vdist_d(C.box, sn[0][s], sn[1][s], sn[2][s], bx, by, bz, rx, ry, rz);
                                r2 = rx * rx + (ry * ry + rz * rz);
                            } else {
                                r2 = sqdist_d(C.box, sn[0][s], sn[1][s], sn[2][s], bx, by, bz);
                            }
                            un = pair_u64_t<ELEC, SR>(C.pot, C.tab, r2, sq[s], bq, sr[s], br, st[s], bt);
                        }
                        if (ko) {
                            double r2;
                            if (vec) {
                                double rx, ry, rz;
                                vdist_d(C.box, so[0][s], so[1][s], so[2][s], bx, by, bz, rx, ry, rz);
                                r2 = rx * rx + (ry * ry + rz * rz);
                            } else {
                                r2 = sqdist_d(C.box, so[0][s], so[1][s], so[2][s], bx, by, bz);
                            }
                            uo = pair_u64_t<ELEC, SR>(C.pot, C.tab, r2, sq[s], bq, sr[s], br, st[s], bt);
                        }
                        acc += un - uo;
                    }
                } else {
                    const int4 pj = C.pk[j];
                    int32_t bx = pj.x, by = pj.y, bz = pj.z;
                    float bq = __int_as_float(pj.w), br = C.rf[j];
                    int bt = C.t[j];
                    const float fsx = C.box.fscale[0], fsy = C.box.fscale[1], fsz = C.box.fscale[2];
                    for (int s = 0; s < ns; ++s) {
                        float aq = (float)sq[s], ar = (float)sr[s];
                        int at = st[s];
                        float un = 0.f, uo = 0.f;
                        if (kn) {
                            float dx = (float)(int32_t)(snf[0][s] - bx) * fsx;
                            float dy = (float)(int32_t)(snf[1][s] - by) * fsy;
                            float dz = (float)(int32_t)(snf[2][s] - bz) * fsz;
                            un = pair_u32_t<ELEC, SR>(C.pot, C.tabf, dx * dx + dy * dy + dz * dz, aq, bq, ar, br, at, bt);
                        }
                        if (ko) {
                            float dx = (float)(int32_t)(sof[0][s] - bx) * fsx;
                            float dy = (float)(int32_t)(sof[1][s] - by) * fsy;
                            float dz = (float)(int32_t)(sof[2][s] - bz) * fsz;
                            uo = pair_u32_t<ELEC, SR>(C.pot, C.tabf, dx * dx + dy * dy + dz * dz, aq, bq, ar, br, at, bt);
                        }
                        if (!(fabsf(un) <= C.pot.exact_f) || !(fabsf(uo) <= C.pot.exact_f)) {
                            // strong pair or contact decision: FP64 (same as the FP64 path)
                            double jx = C.x[j], jy = C.y[j], jz = C.z[j], jq = C.q[j], jr = C.r[j];
                            double u64n = 0.0, u64o = 0.0;
                            if (kn) u64n = pair_u64_t<ELEC, SR>(C.pot, C.tab, sqdist_d(C.box, sn[0][s], sn[1][s], sn[2][s], jx, jy, jz), sq[s], jq, sr[s], jr, at, bt);
                            if (ko) u64o = pair_u64_t<ELEC, SR>(C.pot, C.tab, sqdist_d(C.box, so[0][s], so[1][s], so[2][s], jx, jy, jz), sq[s], jq, sr[s], jr, at, bt);
                            acc += u64n - u64o;
                        } else {
                            acc += (double)(un - uo);
                        }
                    }
                }
            }
            double du = block_sum(acc, red);
            if (s_collide) du = __builtin_inf();
            if (!tr.rot && !tr.trans) du = 0.0;   // move.h:1709-1710
            if (wave0) {
                bool accept = !(mt_canonical(G) > exp(-du));
                if (tid == 0) tr.acc = accept ? 1 : 0;
            }
            __syncthreads();
            const bool accept = tr.acc;
            n_trials++;
            if (tid == 0 && C.du_rec && rec < C.cap) {
                C.du_rec[rec] = du;
                C.acc_rec[rec] = accept;
            }
            rec++;
            if (accept) {   // Group::accept (group.h:381-389)
                n_acc++;
                utot += du;
                dusum += du;
                for (int s = tid; s < ns; s += nthr) {
                    int i = gd.front + s;
                    C.x[i] = sn[0][s]; C.y[i] = sn[1][s]; C.z[i] = sn[2][s];
                    C.pk_w[i] = make_int4(snf[0][s], snf[1][s], snf[2][s], C.pk[i].w);
                }
                if (tid == 0) {
                    C.cm[6 * g] = tr.cmx; C.cm[6 * g + 1] = tr.cmy; C.cm[6 * g + 2] = tr.cmz;
                    C.cm[6 * g + 3] = tr.cmx; C.cm[6 * g + 4] = tr.cmy; C.cm[6 * g + 5] = tr.cmz;
                }
                if (cached) {   // EnergyMatrix::update(true): row and column g
                    for (int h = tid; h < C.ng; h += nthr) {
                        if (h == g) continue;
                        C.emat[(size_t)g * C.ng + h] = enew[h];
                        C.emat[(size_t)h * C.ng + g] = enew[h];
                    }
                }
            } else if (du == __builtin_inf()) {
                n_inf++;
            }
            __syncthreads();   // committed sites / cm visible to the next trial
            par ^= 1;
        }
        if (!stop) ++calls_done;
    }
    if (!stop && tid == 0 && C.resume) C.resume[0] = 0;
    __syncthreads();
    for (int k = tid; k < 624; k += nthr) {
        C.rng[k] = mtg[k];
        C.rng[625 + k] = mti[k];
    }
    if (tid == 0) {
        C.rng[624] = (uint32_t)G.p;
        C.rng[625 + 624] = (uint32_t)I.p;
        C.stats_i[0] = n_trials;
        C.stats_i[1] = n_acc;
        C.stats_i[2] = n_inf;
        C.stats_i[3] = 0;
        C.stats_i[4] = wall_clock64() - t_start;
        C.stats_i[5] = calls_done;
        C.stats_d[0] = utot;
        C.stats_d[1] = dusum;
    }
}

template <int ELEC, int SR, bool MIX, int U, bool FX, bool CELL = false, bool SKG = false>
static void launch_at_f(hipStream_t st, const ChainDev *d_chains, int nchains, int nthr,
                        size_t dyn, int ncalls)
{
    if (dyn > 0)
        (void)hipFuncSetAttribute((const void *)k_sweep_at<ELEC, SR, MIX, U, FX, CELL, SKG>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
    hipLaunchKernelGGL((k_sweep_at<ELEC, SR, MIX, U, FX, CELL, SKG>), dim3(nchains), dim3(nthr),
                       dyn, st, d_chains, ncalls);
}

// Ewald chains whose S(k) lives in device memory: batch depth 4, no cell list
template <int SR, bool MIX>
static void launch_at_skg(hipStream_t st, const ChainDev *d_chains, int nchains, int nthr,
                          size_t dyn, int ncalls, bool fx, bool cell)
{
    if constexpr (MIX && SR == SR_LJ && at_fx_ok<ELEC_EWALD, SR>())
        if (fx && cell) {
            // pair terms cut at rc: cell lists where present, else the screen
            launch_at_f<ELEC_EWALD, SR, MIX, 4, true, true, true>(st, d_chains, nchains, nthr, dyn,
                                                                  ncalls);
            return;
        }
    if constexpr (MIX && at_fx_ok<ELEC_EWALD, SR>())
        if (fx) {
            launch_at_f<ELEC_EWALD, SR, MIX, 4, true, false, true>(st, d_chains, nchains, nthr, 0,
                                                                   ncalls);
            return;
        }
    launch_at_f<ELEC_EWALD, SR, MIX, 4, false, false, true>(st, d_chains, nchains, nthr, 0, ncalls);
}

template <int ELEC, int SR, bool MIX, int U>
static void launch_at_u(hipStream_t st, const ChainDev *d_chains, int nchains, int nthr,
                        size_t dyn, int ncalls, bool fx, bool cell)
{
    if constexpr (MIX && at_fx_ok<ELEC, SR>()) {
        if constexpr (ELEC == ELEC_EWALD && SR == SR_LJ && U == 4) {
            // pair terms cut at rc: chains with a cell list scan it, the
            // others screen all partners
            if (fx && cell) {
                launch_at_f<ELEC, SR, MIX, U, true, true>(st, d_chains, nchains, nthr, dyn, ncalls);
                return;
            }
        }
        if (fx) {
            launch_at_f<ELEC, SR, MIX, U, true>(st, d_chains, nchains, nthr, dyn, ncalls);
            return;
        }
    }
    launch_at_f<ELEC, SR, MIX, U, false>(st, d_chains, nchains, nthr, dyn, ncalls);
}

// partner-loop unroll depth: the autotuned knob of the mixed path (the FP64
// path keeps a plain strided loop; cell-list launches use depth 4 for the
// all-pairs screen they fall back to)
template <int ELEC, int SR, bool MIX>
static void launch_at_k(hipStream_t st, const ChainDev *d_chains, int nchains, int nthr,
                        size_t dyn, int ncalls, int unroll, bool fx, bool cell)
{
    if (!MIX || unroll <= 4 || cell)
        launch_at_u<ELEC, SR, MIX, 4>(st, d_chains, nchains, nthr, dyn, ncalls, fx, cell);
    else
        launch_at_u<ELEC, SR, MIX, 8>(st, d_chains, nchains, nthr, dyn, ncalls, fx, cell);
}

// A prepared batched launch (fgx_sweep.hip): the ChainDev array on the device
// and the launch shape of its sweep kernel.
struct BatchPlan {
    ChainDev *d_chains = nullptr;
    int nchains = 0, nthr = 0, ncalls = 0, unroll = 4, move_type = 0;
    size_t dyn = 0;
    bool fx = false, cell = false;
    bool skg = false;   // Ewald S(k) in device memory
    Pot pot{};
    // block sweep: workgroup map (chain, part) and grid size
    bool blk = false;
    const int2 *d_map = nullptr;
    int ngrid = 0;
    int *d_xerr = nullptr;
};
int prepare_batch(std::vector<fgx_system *> &sys, const std::vector<int> &ncalls,
                  const std::vector<long> *max_trials, double *du_rec, int *acc_rec, long cap,
                  BatchPlan &B);
int enqueue_batch(const BatchPlan &B, hipStream_t st);
int blk_check(const std::vector<fgx_system *> &sys, int *d_xerr);
int mirror_trial(std::vector<fgx_system *> &sys, hipStream_t st);

// Per-family launchers (fgx_sweep_at_*.hip, fgx_sweep_tr.hip).  Each returns
// FGX_EUNSUPPORTED without an error message for a Hamiltonian outside its
// family, so the caller can try the next one.
int launch_at_plain(bool mix, const Pot &P, hipStream_t st, const ChainDev *d_chains, int nchains,
                    int nthr, size_t dyn, int ncalls, int unroll, bool fx, bool cell);
int launch_at_dh(bool mix, const Pot &P, hipStream_t st, const ChainDev *d_chains, int nchains,
                 int nthr, size_t dyn, int ncalls, int unroll, bool fx, bool cell);
int launch_at_ewald_mixed(const Pot &P, hipStream_t st, const ChainDev *d_chains, int nchains,
                          int nthr, size_t dyn, int ncalls, int unroll, bool fx, bool cell,
                          bool skg);
int launch_at_ewald_fp64(const Pot &P, hipStream_t st, const ChainDev *d_chains, int nchains,
                         int nthr, size_t dyn, int ncalls, bool skg);
int launch_tr(bool mix, const Pot &P, hipStream_t st, const ChainDev *d_chains, int nchains,
              int nthr, int ncalls);
int launch_emat_init(const Pot &P, hipStream_t st, const ChainDev *d_chains, int nchains,
                     int maxng);

}  // namespace fgx
