/*
 * oracle.c -- CPU restatement of the Faunus Metropolis Delta-U path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Every function cites the
 * reference file:line whose behaviour it restates.  Arithmetic follows the
 * reference's evaluation order (left-to-right C++ expressions, Eigen 3.2.9
 * fixed-size reductions x0 + (x1 + x2)) and is compiled with
 * -ffp-contract=off, so trial coordinates and accept/reject sequences match
 * an x86-64 build of the reference bit for bit.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ===================================================================== */
/* RNG                                                                    */
/* ===================================================================== */

/* std::mt19937 seeding, C++ [rand.eng.mers]; default seed 5489 is what
 * RandomTwister's default constructor gives (include/faunus/slump.h:37). */
void or_mt_seed(or_mt *m, uint32_t s)
{
    m->x[0] = s;
    for (int i = 1; i < 624; ++i) {
        uint32_t prev = m->x[i - 1];
        m->x[i] = 1812433253u * (prev ^ (prev >> 30)) + (uint32_t)i;
    }
    m->p = 624;
}

static void or_mt_twist(or_mt *m)
{
    const uint32_t UPPER = 0x80000000u, LOWER = 0x7fffffffu, MATRIX = 0x9908b0dfu;
    for (int k = 0; k < 624; ++k) {
        uint32_t y = (m->x[k] & UPPER) | (m->x[(k + 1) % 624] & LOWER);
        m->x[k] = m->x[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? MATRIX : 0u);
    }
    m->p = 0;
}

uint32_t or_mt_u32(or_mt *m)
{
    if (m->p >= 624) or_mt_twist(m);
    uint32_t z = m->x[m->p++];
    z ^= (z >> 11);
    z ^= (z << 7) & 0x9d2c5680u;
    z ^= (z << 15) & 0xefc60000u;
    z ^= (z >> 18);
    return z;
}

/* RandomTwister::operator() (slump.h:87-95) = uniform_real_distribution
 * (0,1) = generate_canonical<double,53>: two 32-bit words, low word first,
 * sum = w1 + w2*2^32 rounded once, divided by 2^64, clamped below 1
 * (libstdc++ bits/random.tcc, generate_canonical). */
double or_canonical(or_mt *m)
{
    double sum = 0.0, tmp = 1.0;
    sum += (double)or_mt_u32(m) * tmp;
    tmp *= 4294967296.0;
    sum += (double)or_mt_u32(m) * tmp;
    tmp *= 4294967296.0;
    double r = sum / tmp;
    if (r >= 1.0) r = nextafter(1.0, 0.0);
    return r;
}

/* RandomTwister::range (slump.h:62-66) = uniform_int_distribution<int>:
 * libstdc++ 11 uses Lemire's nearly-divisionless method for a 32-bit
 * engine (bits/uniform_int_dist.h, _S_nd). */
int or_range(or_mt *m, int a, int b)
{
    uint32_t erange = (uint32_t)((uint32_t)b - (uint32_t)a) + 1u;
    uint64_t prod = (uint64_t)or_mt_u32(m) * (uint64_t)erange;
    uint32_t low = (uint32_t)prod;
    if (low < erange) {
        uint32_t thr = (uint32_t)(-erange) % erange;
        while (low < thr) {
            prod = (uint64_t)or_mt_u32(m) * (uint64_t)erange;
            low = (uint32_t)prod;
        }
    }
    return (int)((uint32_t)(prod >> 32) + (uint32_t)a);
}

/* libstdc++ operator<< / >> for mersenne_twister_engine: 624 words, then
 * the position index (the format Space::save writes, space.h:761). */
int or_mt_from_text(or_mt *m, const char *txt)
{
    const char *c = txt;
    char *end;
    for (int i = 0; i < 624; ++i) {
        unsigned long v = strtoul(c, &end, 10);
        if (end == c) return -1;
        m->x[i] = (uint32_t)v;
        c = end;
    }
    long p = strtol(c, &end, 10);
    if (end == c) return -1;
    m->p = (int)p;
    return 0;
}

int or_mt_to_text(const or_mt *m, char *buf, int cap)
{
    int w = 0;
    for (int i = 0; i < 624; ++i) {
        int k = snprintf(buf + w, (size_t)(cap > w ? cap - w : 0), "%u ", m->x[i]);
        if (k < 0 || w + k >= cap) return -1;
        w += k;
    }
    int k = snprintf(buf + w, (size_t)(cap - w), "%d", m->p);
    if (k < 0 || w + k >= cap) return -1;
    return w + k;
}

void or_rng_draws(uint32_t seed, int kind, int a, int b, int n, double *out)
{
    or_mt m;
    or_mt_seed(&m, seed);
    for (int i = 0; i < n; ++i) {
        if (kind == 0) out[i] = (double)or_mt_u32(&m);
        else if (kind == 1) out[i] = or_canonical(&m);
        else out[i] = (double)or_range(&m, a, b);
    }
}

/* ===================================================================== */
/* Geometry::Cuboid (geometry.h:468-555)                                  */
/* ===================================================================== */

typedef struct { double len[3], half[3], inv[3]; } or_box;

static void box_set(or_box *b, const double L[3])
{   /* Cuboid::setlen, src/geometry.cpp:263-272 */
    for (int k = 0; k < 3; ++k) {
        b->len[k] = L[k];
        b->half[k] = L[k] * 0.5;
        b->inv[k] = 1 / L[k];
    }
}

/* Geometrybase::anint, geometry.h:56-59 */
static inline int anint(double x) { return (int)(x > 0. ? x + .5 : x - .5); }

/* Cuboid::sqdist, geometry.h:507-518: |a-b| per axis, fold once, then
 * Eigen squaredNorm = d0^2 + (d1^2 + d2^2). */
static inline double box_sqdist(const or_box *b, const double *a, const double *c)
{
    double d[3];
    for (int k = 0; k < 3; ++k) {
        d[k] = fabs(a[k] - c[k]);
        if (d[k] > b->half[k]) d[k] -= 2 * b->half[k];
    }
    return d[0] * d[0] + (d[1] * d[1] + d[2] * d[2]);
}

/* Cuboid::vdist, geometry.h:521-537 */
static inline void box_vdist(const or_box *b, const double *a, const double *c, double *r)
{
    for (int k = 0; k < 3; ++k) {
        r[k] = a[k] - c[k];
        if (r[k] > b->half[k]) r[k] -= b->len[k];
        else if (r[k] < -b->half[k]) r[k] += b->len[k];
    }
}

/* Cuboid::boundary, geometry.h:539-547 */
static inline void box_boundary(const or_box *b, double *a)
{
    for (int k = 0; k < 3; ++k)
        if (fabs(a[k]) > b->half[k]) a[k] -= b->len[k] * anint(a[k] * b->inv[k]);
}

/* Cuboid::collision, geometry.h:492-501 */
static inline int box_collision(const or_box *b, const double *a)
{
    for (int k = 0; k < 3; ++k)
        if (fabs(a[k]) > b->half[k]) return 1;
    return 0;
}

/* ===================================================================== */
/* System                                                                 */
/* ===================================================================== */

typedef struct {
    double r[3];
    double charge, radius, mw;
    int id;
} or_particle;

typedef struct {
    int front, back, molsize, molid;
    double cm[3], cm_trial[3];
} or_group;

typedef struct {
    int type;           /* OR_MOVE_AT / OR_MOVE_TR */
    int molid;
    double prob;
    int peratom, permol, repeat;
    double dir[3], dir2[3], dp1, dp2;
    int center_rotation;
    long cnt;
    /* AtomicTranslation state carried between trials (move.h:749-752) */
    int iparticle;
    int igroup;
} or_move;

typedef struct {
    int nk;
    double *kv;          /* 3*nk */
    double *aks;
    double *Qre, *Qim, *Qtre, *Qtim;
    double V, V_trial;
    double surf, surf_trial, reci, reci_trial;
    double alpha, alpha2, rc2, kappa, kc, kc2, eps_surf, const_inf;
    int kcc, spherical, update_frequency, cnt_accepted;
    double *eff;         /* per type effective-charge ratio (ewald.h:535-545) */
    double *qeqe;        /* per type pair (ewald.h:541-545) */
    double update_drift;
} or_ewald;

#define OR_MAX_MOVES 8
#define OR_MAX_CHANGE 4096

struct or_sys {
    or_box box;
    int n;
    or_particle *p, *trial;
    int ng;
    or_group *g;
    /* potential parameters */
    int elec, sr, nbkind;
    double lB, kappa, g2g_rc2;
    int ntypes;
    double *s2, *eps4;   /* LennardJonesMixed tables (potentials.h:2010-2023) */
    double lj_rc2;       /* CutShift cutoff^2 (potentials.h:1785-1786); 0 = uncut */
    int has_press;       /* ExternalPressure term present */
    double press;        /* its pressure, 1/A^3 */
    long iso_cnt;        /* Isobaric trial count (Movebase::cnt) */
    double *tdp;
    or_ewald *ew;
    /* RNG: global slump and the Movebase-internal _slump (move.h:356-360) */
    or_mt slump, internal;
    or_move moves[OR_MAX_MOVES];
    int nmoves;
    int propagator_started;
    double uinit, dusum;
    /* last change (for Ewald update and recording) */
    int ch_n, ch_g[16], ch_off[17], ch_idx[OR_MAX_CHANGE];
    int *last_which; long last_cap, last_n;
};

static inline or_particle *vecp(or_sys *s, int which) { return which ? s->trial : s->p; }

/* ---------------- pair potentials ---------------- */

/* erfc-based screened Ewald real-space term, ewald.h:97-110 (early return
 * at :110 makes the tabulated branch dead). */
static inline double pair_ewald_real(const or_sys *s, const or_particle *a, const or_particle *b,
                                     const double *r)
{
    const or_ewald *e = s->ew;
    double r2 = r[0] * r[0] + (r[1] * r[1] + r[2] * r[2]);
    if (r2 > e->rc2) return 0.0;
    double r1 = sqrt(r2);
    double E = 0.0;
    E += a->charge * b->charge *
         (erfc(e->alpha * r1 + e->kappa / 2.0 / e->alpha) * exp(e->kappa * r1) +
          erfc(e->alpha * r1 - e->kappa / 2.0 / e->alpha) * exp(-e->kappa * r1)) /
         (2.0 * r1) * e->qeqe[a->id * s->ntypes + b->id] * s->lB;
    return E;
}

/* CombinedPairPotential::operator() (potentials.h:3768-3771):
 * first(a,b,r) + second(a,b,r).  r2 is the squared distance; r the vector
 * (used by EwaldReal, whose LJ partner squares r itself). */
static inline double pair_u(const or_sys *s, const or_particle *a, const or_particle *b,
                            double r2, const double *r)
{
    double u1 = 0.0, u2 = 0.0;
    switch (s->elec) {
    case OR_ELEC_COULOMB: /* Coulomb, potentials.h:2532-2539 */
        u1 = s->lB * a->charge * b->charge / sqrt(r2);
        break;
    case OR_ELEC_DH: {    /* DebyeHuckel, potentials.h:3314-3324 */
        double rr = sqrt(r2);
        u1 = s->lB * a->charge * b->charge / rr * exp(-s->kappa * rr);
        break;
    }
    case OR_ELEC_EWALD:
        u1 = pair_ewald_real(s, a, b, r);
        break;
    default: break;
    }
    switch (s->sr) {
    case OR_SR_HS: {      /* HardSphere, potentials.h:463-467 */
        double m = a->radius + b->radius;
        u2 = (r2 < m * m) ? INFINITY : 0;
        break;
    }
    case OR_SR_LJ: {      /* LennardJonesMixed, potentials.h:2035-2040 */
        const int k = a->id * s->ntypes + b->id;
        if (s->lj_rc2 > 0) {
            /* CutShift<LennardJonesLB> (potentials.h:1805-1811): zero beyond
             * the cutoff, shifted by the energy at the cutoff inside it */
            if (r2 > s->lj_rc2) {
                u2 = 0;
            } else {
                double x = s->s2[k] / r2, xc = s->s2[k] / s->lj_rc2;
                x = x * x * x;
                xc = xc * xc * xc;
                u2 = s->eps4[k] * (x * x - x) - s->eps4[k] * (xc * xc - xc);
            }
            break;
        }
        double x = s->s2[k] / r2;
        x = x * x * x;
        u2 = s->eps4[k] * (x * x - x);
        break;
    }
    default: break;
    }
    if (s->elec == OR_ELEC_NONE) return u2;
    if (s->sr == OR_SR_NONE) return u1;
    return u1 + u2;
}

/* Nonbonded::i2i uses geo.sqdist (energy.h:842-845); NonbondedVector::i2i
 * uses geo.vdist (energy.h:1119-1122). */
static inline double pot_ij(const or_sys *s, const or_particle *a, const or_particle *b)
{
    double r[3];
    if (s->nbkind == OR_NB_VECTOR) {
        box_vdist(&s->box, a->r, b->r, r);
        double r2 = r[0] * r[0] + (r[1] * r[1] + r[2] * r[2]);
        return pair_u(s, a, b, r2, r);
    }
    double r2 = box_sqdist(&s->box, a->r, b->r);
    if (s->elec == OR_ELEC_EWALD) box_vdist(&s->box, a->r, b->r, r);
    return pair_u(s, a, b, r2, r);
}

double or_pair(or_sys *s, int which, int i, int j)
{
    or_particle *p = vecp(s, which);
    return pot_ij(s, &p[i], &p[j]);
}

/* ---------------- mass centre (geometry.h:713-740, 817-822) ---------------- */

static void mass_center(const or_sys *s, const or_particle *p, const or_group *g, double *cm)
{
    cm[0] = cm[1] = cm[2] = 0;
    if (g->back < g->front) return;
    /* origo = p[front + (back-front)*0.5], double index truncated */
    const double *o = p[(size_t)(g->front + (g->back - g->front) * 0.5)].r;
    double sum = 0;
    for (int i = g->front; i <= g->back; ++i) {
        double t[3] = {p[i].r[0] - o[0], p[i].r[1] - o[1], p[i].r[2] - o[2]};
        box_boundary(&s->box, t);
        double w = p[i].mw;
        for (int k = 0; k < 3; ++k) cm[k] += t[k] * w;
        sum += w;
    }
    if (fabs(sum) < 1e-6) sum = 1;
    for (int k = 0; k < 3; ++k) cm[k] = cm[k] / sum + o[k];
    box_boundary(&s->box, cm);
}

static inline int g_find(const or_group *g, int i) { return i >= g->front && i <= g->back; }
static inline int g_empty(const or_group *g) { return g->back < g->front; }
static inline int g_size(const or_group *g) { return g->back - g->front + 1; }
static inline int g_molecular(const or_group *g) { return g->molsize == g_size(g); }
static inline int g_atomic(const or_group *g) { return g->molsize == 1; }

/* ===================================================================== */
/* Ewald (ewald.h:229-760)                                                */
/* ===================================================================== */

static double complex_norm(double re, double im)
{   /* std::norm on complex<double> (libstdc++ <complex>, _Norm_helper) */
    return re * re + im * im;
}

static void ewald_kvectors(or_sys *s)
{   /* NonbondedEwald::kVectorChange, ewald.h:330-378 (isotropic_pbc=false) */
    or_ewald *e = s->ew;
    double Lx = s->box.len[0], Ly = s->box.len[1], Lz = s->box.len[2];
    double maxL = Lx;
    if (Ly > maxL) maxL = Ly;
    if (Lz > maxL) maxL = Lz;
    const double pi = acos(-1.);
    double check_k2_zero = 0.1 * (4 * pi * pi) / (maxL * maxL);
    int kcc = e->kcc;
    int len = (2 * kcc + 1) * (2 * kcc + 1) * (2 * kcc + 1) - 1;
    free(e->kv); free(e->aks); free(e->Qre); free(e->Qim); free(e->Qtre); free(e->Qtim);
    if (len == 0) {
        e->nk = 1;
        e->kv = calloc(3, sizeof(double)); e->kv[0] = 1.0;
        e->aks = calloc(1, sizeof(double));
    } else {
        e->kv = calloc((size_t)3 * len, sizeof(double));
        e->aks = calloc((size_t)len, sizeof(double));
        int nk = 0;
        for (int kx = 0; kx <= kcc; kx++) {
            double dkx2 = (double)(kx * kx);
            for (int ky = -kcc; ky <= kcc; ky++) {
                double dky2 = (double)(ky * ky);
                for (int kz = -kcc; kz <= kcc; kz++) {
                    double factor = 1.0;
                    if (kx > 0) factor = 2.0;
                    double dkz2 = (double)(kz * kz);
                    double twopi = 2 * pi;
                    double kv[3] = {twopi * (kx / Lx), twopi * (ky / Ly), twopi * (kz / Lz)};
                    double k2 = kv[0] * kv[0] + (kv[1] * kv[1] + kv[2] * kv[2]);
                    if (k2 < check_k2_zero) continue;
                    if (e->spherical)
                        if ((dkx2 / e->kc2) + (dky2 / e->kc2) + (dkz2 / e->kc2) > 1.0) continue;
                    e->kv[3 * nk + 0] = kv[0];
                    e->kv[3 * nk + 1] = kv[1];
                    e->kv[3 * nk + 2] = kv[2];
                    e->aks[nk] = factor * exp(-(k2 + e->kappa * e->kappa) / (4.0 * e->alpha2)) /
                                 (k2 + e->kappa * e->kappa);
                    nk++;
                }
            }
        }
        e->nk = nk;
    }
    e->Qre = calloc((size_t)e->nk, sizeof(double));
    e->Qim = calloc((size_t)e->nk, sizeof(double));
    e->Qtre = calloc((size_t)e->nk, sizeof(double));
    e->Qtim = calloc((size_t)e->nk, sizeof(double));
}

/* updateAllComplexNumbers, ewald.h:388-415 */
static void ewald_all_complex(or_sys *s, const or_particle *p, double *Qre, double *Qim)
{
    or_ewald *e = s->ew;
    for (int k = 0; k < e->nk; k++) {
        const double *kv = &e->kv[3 * k];
        double re = 0.0, im = 0.0;
        for (int i = 0; i < s->n; i++) {
            double dot = kv[0] * p[i].r[0] + (kv[1] * p[i].r[1] + kv[2] * p[i].r[2]);
            double q = p[i].charge, ef = e->eff[p[i].id];
            re += q * cos(dot) * ef;
            im += q * sin(dot) * ef;
        }
        Qre[k] = re;
        Qim[k] = im;
    }
}

/* getSelfEnergy, ewald.h:260-272 */
static double ewald_self_range(const or_sys *s, const or_particle *p, int front, int back)
{
    const or_ewald *e = s->ew;
    double Eq = 0;
    for (int i = front; i <= back; ++i)
        Eq += p[i].charge * p[i].charge * e->eff[p[i].id] * e->eff[p[i].id];
    return (-e->alpha * (Eq + e->alpha2 * (2.0 / 3.0) * 0.0) / sqrt(acos(-1.))) * s->lB;
}

/* getSurfaceEnergy, ewald.h:278-292 */
static double ewald_surface_v(const or_sys *s, const or_particle *p, double V)
{
    const or_ewald *e = s->ew;
    if (e->const_inf < 0.5) return 0.0;
    double qrs[3] = {0, 0, 0};
    for (int i = 0; i < s->n; ++i) {
        double w = e->eff[p[i].id];
        for (int k = 0; k < 3; ++k) qrs[k] = qrs[k] + p[i].charge * p[i].r[k] * w;
    }
    double qq = qrs[0] * qrs[0] + (qrs[1] * qrs[1] + qrs[2] * qrs[2]);
    const double pi = acos(-1.);
    return e->const_inf * (2 * pi / ((2 * e->eps_surf + 1) * V)) * (qq + 2 * 0.0 + 0.0) * s->lB;
}

/* getReciprocalEnergy, ewald.h:298-312 */
static double ewald_reci_v(const or_sys *s, const double *Qre, const double *Qim, double V)
{
    const or_ewald *e = s->ew;
    double E = 0.0;
    for (int k = 0; k < e->nk; k++) {
        double Q2 = 0.0;
        Q2 += complex_norm(Qre[k], Qim[k]);
        E += (e->aks[k] * Q2);
    }
    const double pi = acos(-1.);
    return (2 * pi / V) * E * s->lB;
}

static void ewald_undo(or_sys *s)
{   /* ewald.h:568-586 */
    or_ewald *e = s->ew;
    e->V_trial = e->V;
    e->surf_trial = e->surf;
    e->reci_trial = e->reci;
    memcpy(e->Qtre, e->Qre, sizeof(double) * e->nk);
    memcpy(e->Qtim, e->Qim, sizeof(double) * e->nk);
}

static void ewald_accept(or_sys *s)
{   /* ewald.h:588-607 */
    or_ewald *e = s->ew;
    e->V = e->V_trial;
    e->surf = e->surf_trial;
    e->reci = e->reci_trial;
    memcpy(e->Qre, e->Qtre, sizeof(double) * e->nk);
    memcpy(e->Qim, e->Qtim, sizeof(double) * e->nk);
}

/* setSpace -> setGeometry(geo) -> kVectorChange + updateAllComplexNumbers(p);
 * then surface/reciprocal energies and undo() (ewald.h:731-757) */
void or_ewald_set_space(or_sys *s)
{
    or_ewald *e = s->ew;
    if (!e) return;
    e->V = s->box.len[0] * s->box.len[1] * s->box.len[2];
    ewald_kvectors(s);
    ewald_all_complex(s, s->p, e->Qre, e->Qim);
    e->surf = ewald_surface_v(s, s->p, e->V);
    e->reci = ewald_reci_v(s, e->Qre, e->Qim, e->V);
    ewald_undo(s);
}

/* updateChange, ewald.h:651-705 (no geometry change) */
double or_ewald_update_change(or_sys *s, int nmv, const int *mvg, const int *off, const int *idx)
{
    or_ewald *e = s->ew;
    if (!e) return 0.0;
    (void)mvg;
    for (int k = 0; k < e->nk; k++) {
        double re = e->Qre[k], im = e->Qim[k];
        const double *kv = &e->kv[3 * k];
        for (int m = 0; m < nmv; ++m) {
            for (int t = off[m]; t < off[m + 1]; ++t) {
                int i = idx[t];
                const or_particle *pt = &s->trial[i], *pp = &s->p[i];
                /* kVectors.col(k).dot(r): a dynamic-size Eigen block, so
                 * the reduction runs left to right, (k0 r0 + k1 r1) + k2 r2 */
                double dotT = (kv[0] * pt->r[0] + kv[1] * pt->r[1]) + kv[2] * pt->r[2];
                double dot = (kv[0] * pp->r[0] + kv[1] * pp->r[1]) + kv[2] * pp->r[2];
                double eT = e->eff[pt->id], eP = e->eff[pp->id];
                re += pt->charge * cos(dotT) * eT;
                im += pt->charge * sin(dotT) * eT;
                re -= pp->charge * cos(dot) * eP;
                im -= pp->charge * sin(dot) * eP;
            }
        }
        e->Qtre[k] = re;
        e->Qtim[k] = im;
    }
    return 0.0;
}

/* update(bool), ewald.h:613-648 */
double or_ewald_update(or_sys *s, int accepted)
{
    or_ewald *e = s->ew;
    if (!e) return 0.0;
    if (!accepted) {
        ewald_undo(s);
        return 0.0;
    }
    if (++e->cnt_accepted > e->update_frequency - 1) {
        double duB = ewald_reci_v(s, e->Qtre, e->Qtim, e->V_trial);
        ewald_all_complex(s, s->trial, e->Qtre, e->Qtim);
        double duA = ewald_reci_v(s, e->Qtre, e->Qtim, e->V_trial);
        ewald_accept(s);
        e->cnt_accepted = 0;
        e->update_drift += fabs(duA - duB);
        return duA - duB;
    }
    ewald_accept(s);
    return 0.0;
}

int or_ewald_nk(or_sys *s) { return s->ew ? s->ew->nk : 0; }

void or_ewald_get(or_sys *s, int which, double *kvec, double *aks, double *Qre, double *Qim)
{
    or_ewald *e = s->ew;
    if (!e) return;
    if (kvec) memcpy(kvec, e->kv, sizeof(double) * 3 * e->nk);
    if (aks) memcpy(aks, e->aks, sizeof(double) * e->nk);
    if (Qre) memcpy(Qre, which ? e->Qtre : e->Qre, sizeof(double) * e->nk);
    if (Qim) memcpy(Qim, which ? e->Qtim : e->Qim, sizeof(double) * e->nk);
}

double or_ewald_self(or_sys *s, int which) { return s->ew ? ewald_self_range(s, vecp(s, which), 0, s->n - 1) : 0; }
double or_ewald_surface(or_sys *s, int which)
{
    if (!s->ew) return 0;
    return ewald_surface_v(s, vecp(s, which), which ? s->ew->V_trial : s->ew->V);
}
double or_ewald_reciprocal(or_sys *s, int which)
{
    if (!s->ew) return 0;
    or_ewald *e = s->ew;
    return which ? ewald_reci_v(s, e->Qtre, e->Qtim, e->V_trial) : ewald_reci_v(s, e->Qre, e->Qim, e->V);
}

/* ===================================================================== */
/* Energies (energy.h)                                                     */
/* ===================================================================== */

double or_i2i(or_sys *s, int which, int i, int j)
{   /* Nonbonded::i2i, energy.h:842-845 */
    or_particle *p = vecp(s, which);
    return pot_ij(s, &p[i], &p[j]);
}

static double i2g_raw(or_sys *s, or_particle *p, const or_group *g, int j)
{   /* Nonbonded::i2g, energy.h:847-865 */
    double u = 0;
    if (!g_empty(g)) {
        int len = g->back + 1;
        if (g_find(g, j)) {
            for (int i = g->front; i < j; i++) u += pot_ij(s, &p[i], &p[j]);
            for (int i = j + 1; i < len; i++) u += pot_ij(s, &p[i], &p[j]);
        } else
            for (int i = g->front; i < len; i++) u += pot_ij(s, &p[i], &p[j]);
    }
    return u;
}

static int find_group(or_sys *s, int i)
{   /* Space::findGroup, space.h:420-426 */
    for (int k = 0; k < s->ng; ++k)
        if (g_find(&s->g[k], i)) return k;
    return -1;
}

/* NonbondedCutg2g::cut, energy.h:1343-1358 */
static int cut_g2g(or_sys *s, int which1, const or_group *g1, int which2, const or_group *g2)
{
    if (s->nbkind != OR_NB_CUTG2G) return 0;
    if (g_molecular(g1) && g_molecular(g2)) {
        const double *a = which1 ? g1->cm_trial : g1->cm;
        const double *b = which2 ? g2->cm_trial : g2->cm;
        if (box_sqdist(&s->box, a, b) > s->g2g_rc2) return 1;
    }
    return 0;
}

double or_i2g(or_sys *s, int which, int g, int j)
{
    or_particle *p = vecp(s, which);
    if (s->nbkind == OR_NB_CUTG2G) {  /* energy.h:1378-1382 */
        int gi = find_group(s, j);
        if (gi >= 0 && cut_g2g(s, which, &s->g[g], which, &s->g[gi])) return 0;
    }
    return i2g_raw(s, p, &s->g[g], j);
}

double or_i2all(or_sys *s, int which, int i)
{
    or_particle *p = vecp(s, which);
    if (s->nbkind == OR_NB_VECTOR) {
        /* NonbondedVector::i2all swaps p[0]<->p[i] and loops j=1..n-1
         * (energy.h:1144-1154) */
        or_particle tmp = p[0]; p[0] = p[i]; p[i] = tmp;
        double u = 0;
        for (int j = 1; j < s->n; ++j) u += pot_ij(s, &p[0], &p[j]);
        tmp = p[0]; p[0] = p[i]; p[i] = tmp;
        return u;
    }
    if (s->nbkind == OR_NB_CUTG2G) {
        /* energy.h:1390-1413 with noPairPotentialCutoff=false */
        double u = 0;
        int gi = find_group(s, i);
        for (int k = 0; k < s->ng; ++k)
            if (!cut_g2g(s, which, &s->g[gi], which, &s->g[k])) u += i2g_raw(s, p, &s->g[k], i);
        return u;
    }
    /* Nonbonded::i2all, energy.h:867-877 */
    double u = 0;
    for (int j = 0; j != i; ++j) u += pot_ij(s, &p[i], &p[j]);
    for (int j = i + 1; j < s->n; ++j) u += pot_ij(s, &p[i], &p[j]);
    return u;
}

static double g2g_raw(or_sys *s, or_particle *p, const or_group *g1, const or_group *g2)
{   /* Nonbonded::g2g, energy.h:879-919 */
    double u = 0;
    if (g_empty(g1) || g_empty(g2)) return 0;
    if (g_find(g1, g2->front) && g_find(g1, g2->back)) {
        for (int i = g1->front; i < g2->front; i++)
            for (int j = g2->front; j <= g2->back; ++j) u += pot_ij(s, &p[i], &p[j]);
        for (int i = g2->back + 1; i <= g1->back; i++)
            for (int j = g2->front; j <= g2->back; ++j) u += pot_ij(s, &p[i], &p[j]);
        return u;
    }
    if (g_find(g2, g1->front) && g_find(g2, g1->back)) {
        for (int i = g2->front; i < g1->front; i++)
            for (int j = g1->front; j <= g1->back; ++j) u += pot_ij(s, &p[i], &p[j]);
        for (int i = g1->back + 1; i <= g2->back; i++)
            for (int j = g1->front; j <= g1->back; ++j) u += pot_ij(s, &p[i], &p[j]);
        return u;
    }
    for (int i = g1->front; i <= g1->back; ++i)
        for (int j = g2->front; j <= g2->back; ++j) u += pot_ij(s, &p[i], &p[j]);
    return u;
}

double or_g2g(or_sys *s, int which, int g1, int g2)
{
    if (cut_g2g(s, which, &s->g[g1], which, &s->g[g2])) return 0;
    return g2g_raw(s, vecp(s, which), &s->g[g1], &s->g[g2]);
}

double or_g_internal(or_sys *s, int which, int gi)
{   /* Nonbonded::g_internal, energy.h:930-941 (pairpot.internal = 0) */
    or_particle *p = vecp(s, which);
    const or_group *g = &s->g[gi];
    double u = 0;
    int b = g->back, f = g->front;
    if (!g_empty(g))
        for (int i = f; i < b; ++i)
            for (int j = i + 1; j <= b; ++j) u += pot_ij(s, &p[i], &p[j]);
    return u;
}

static double box_volume(const or_box *b)
{   /* Cuboid::_getVolume (src/geometry.cpp:293-300) */
    return b->len[0] * b->len[1] * b->len[2];
}

static double nonbonded_external(or_sys *s, int which)
{   /* NonbondedEwald::external, ewald.h:716-729; Energybase default 0 */
    or_ewald *e = s->ew;
    if (!e) return 0.0;
    if (which) {
        e->surf_trial = ewald_surface_v(s, s->trial, e->V_trial);
        e->reci_trial = ewald_reci_v(s, e->Qtre, e->Qtim, e->V_trial);
        return e->surf_trial + e->reci_trial;
    }
    e->surf = ewald_surface_v(s, s->p, e->V);
    e->reci = ewald_reci_v(s, e->Qre, e->Qim, e->V);
    return e->surf + e->reci;
}

double or_external(or_sys *s, int which)
{   /* CombinedEnergy (energy.h:646-749): the nonbonded term, then
     * ExternalPressure::external = P V - ln V (energy.h:1878-1883) at the
     * current geometry */
    double u = nonbonded_external(s, which);
    if (s->has_press) {
        const double V = box_volume(&s->box);
        u += s->press * V - log(V);
    }
    return u;
}

double or_g_external(or_sys *s, int which, int g)
{   /* NonbondedEwald::g_external = getSelfEnergy over the group, ewald.h:712-714;
     * ExternalPressure::g_external = -N ln V, N molecules (energy.h:1885-1893) */
    double u = s->ew ? ewald_self_range(s, vecp(s, which), s->g[g].front, s->g[g].back) : 0.0;
    if (s->has_press) {
        const or_group *gr = &s->g[g];
        const int N = g_empty(gr) ? 0 : g_size(gr) / gr->molsize;
        const double V = box_volume(&s->box);
        u += -N * log(V);
    }
    return u;
}

/* Energy::systemEnergy (energy.h:3486-3491) -> setSpace + Energybase::
 * systemEnergy (energy.h:169-180) */
double or_system_energy(or_sys *s, int which)
{
    if (s->ew) or_ewald_set_space(s);
    double u_pair = 0.0;
    double u = or_external(s, which);
    for (int k = 0; k < s->ng; ++k)
        if (!g_empty(&s->g[k])) u += or_g_external(s, which, k) + or_g_internal(s, which, k);
    for (int i = 0; i < s->ng; ++i)
        for (int j = i + 1; j < s->ng; ++j) u_pair += or_g2g(s, which, i, j);
    return u + u_pair;
}

/* Energybase::g2All, energy.h:182-214; the moved-group map is ordered by
 * group index (std::map). */
static double g2all(or_sys *s, int which, int nmv, const int *mvg)
{
    double du = 0;
    for (int m = 0; m < nmv; ++m) {
        int i = mvg[m];
        for (int j = 0; j < s->ng; j++) {
            int moved = 0;
            for (int t = 0; t < nmv; ++t) moved |= (mvg[t] == j);
            if (!moved) {
                du += or_g2g(s, which, i, j);
                if (du == INFINITY) return INFINITY;
            }
        }
    }
    for (int a = 0; a < nmv; ++a)
        for (int b = a; b < nmv; ++b) {
            du += or_g2g(s, which, mvg[a], mvg[b]);
            if (du == INFINITY) return INFINITY;
        }
    return du;
}

/* energyChangeConfiguration, energy.h:3521-3572 */
static double change_config(or_sys *s, int which, int nmv, const int *mvg, const int *off,
                            const int *idx)
{
    double du = 0;
    du += g2all(s, which, nmv, mvg);
    du += or_external(s, which);
    for (int m = 0; m < nmv; ++m) {
        int gi = mvg[m];
        const or_group *g = &s->g[gi];
        du += or_g_external(s, which, gi);
        if (g_atomic(g)) {
            for (int t = off[m]; t < off[m + 1]; ++t) {
                int j = idx[t];
                for (int k = g->front; k <= g->back; ++k) {
                    int inmoved = 0;
                    for (int q = off[m]; q < off[m + 1]; ++q) inmoved |= (idx[q] == k);
                    if (!inmoved || j > k) du += or_i2i(s, which, j, k);
                }
            }
        }
        if (g_molecular(g))
            if (off[m + 1] > off[m]) du += or_g_internal(s, which, gi);
    }
    return du;
}

/* Energy::energyChange, energy.h:3577-3621 */
double or_energy_change(or_sys *s, int nmv, const int *mvg, const int *off, const int *idx)
{
    if (nmv == 0) return 0;
    for (int m = 0; m < nmv; ++m) {
        if (off[m + 1] == off[m]) {
            const or_group *g = &s->g[mvg[m]];
            for (int j = g->front; j <= g->back; ++j)
                if (box_collision(&s->box, s->trial[j].r)) return INFINITY;
        } else
            for (int t = off[m]; t < off[m + 1]; ++t)
                if (box_collision(&s->box, s->trial[idx[t]].r)) return INFINITY;
    }
    double duNew = change_config(s, 1, nmv, mvg, off, idx);
    double duOld = change_config(s, 0, nmv, mvg, off, idx);
    return duNew - duOld;
}

void or_commit(or_sys *s, int nmv, const int *mvg, const int *off, const int *idx, int accepted)
{
    for (int m = 0; m < nmv; ++m) {
        or_group *g = &s->g[mvg[m]];
        if (off[m + 1] == off[m]) {
            for (int i = g->front; i <= g->back; ++i) {
                if (accepted) s->p[i] = s->trial[i]; else s->trial[i] = s->p[i];
            }
            if (accepted) memcpy(g->cm, g->cm_trial, sizeof g->cm);
            else memcpy(g->cm_trial, g->cm, sizeof g->cm);
        } else {
            for (int t = off[m]; t < off[m + 1]; ++t) {
                int i = idx[t];
                if (accepted) s->p[i] = s->trial[i]; else s->trial[i] = s->p[i];
            }
            if (g_molecular(g)) {
                if (accepted) memcpy(g->cm, g->cm_trial, sizeof g->cm);
                else memcpy(g->cm_trial, g->cm, sizeof g->cm);
            }
        }
    }
}

/* ===================================================================== */
/* Construction                                                            */
/* ===================================================================== */

or_sys *or_create(const or_params *prm, const double box[3], int n,
                  const double *xyz, const double *charge, const double *radius,
                  const double *mw, const int *id,
                  int ng, const int *front, const int *back,
                  const int *molsize, const int *molid)
{
    or_sys *s = calloc(1, sizeof(or_sys));
    box_set(&s->box, box);
    s->n = n;
    s->p = calloc((size_t)(n > 0 ? n : 1), sizeof(or_particle));
    s->trial = calloc((size_t)(n > 0 ? n : 1), sizeof(or_particle));
    for (int i = 0; i < n; ++i) {
        or_particle *q = &s->p[i];
        q->r[0] = xyz[3 * i]; q->r[1] = xyz[3 * i + 1]; q->r[2] = xyz[3 * i + 2];
        q->charge = charge[i]; q->radius = radius[i]; q->mw = mw[i]; q->id = id[i];
        s->trial[i] = *q;
    }
    s->lj_rc2 = prm->lj_cut > 0 ? pow(prm->lj_cut, 2) : 0.0;   /* potentials.h:1786 */
    s->has_press = prm->has_pressure;
    s->press = prm->pressure;
    s->ng = ng;
    s->g = calloc((size_t)(ng > 0 ? ng : 1), sizeof(or_group));
    for (int k = 0; k < ng; ++k) {
        s->g[k].front = front[k]; s->g[k].back = back[k];
        s->g[k].molsize = molsize[k]; s->g[k].molid = molid[k];
    }
    for (int k = 0; k < ng; ++k) {   /* Space::load -> setMassCenter (space.h:850) */
        mass_center(s, s->p, &s->g[k], s->g[k].cm);
        memcpy(s->g[k].cm_trial, s->g[k].cm, sizeof s->g[k].cm);
    }
    s->elec = prm->elec; s->sr = prm->sr; s->nbkind = prm->nbkind;
    s->lB = prm->lB; s->kappa = prm->kappa;
    s->g2g_rc2 = prm->g2g_cut > 0 ? pow(prm->g2g_cut, 2) : INFINITY;
    int T = prm->ntypes;
    s->ntypes = T;
    s->s2 = calloc((size_t)(T * T > 0 ? T * T : 1), sizeof(double));
    s->eps4 = calloc((size_t)(T * T > 0 ? T * T : 1), sizeof(double));
    s->tdp = calloc((size_t)(T > 0 ? T : 1), sizeof(double));
    for (int a = 0; a < T; ++a) {
        if (prm->tdp) s->tdp[a] = prm->tdp[a];
        for (int b = 0; b < T; ++b) {
            /* LennardJonesMixed::init with LorentzBerthelot, potentials.h:2010-2019,
             * 1816-1831 */
            double sigma = 0.5 * (prm->sigma ? prm->sigma[a] + prm->sigma[b] : 0.0);
            double epsilon = prm->eps ? sqrt(prm->eps[a] * prm->eps[b]) : 0.0;
            s->s2[a * T + b] = sigma * sigma;
            s->eps4[a * T + b] = 4 * epsilon;
        }
    }
    if (prm->elec == OR_ELEC_EWALD) {
        or_ewald *e = calloc(1, sizeof(or_ewald));
        s->ew = e;
        e->eps_surf = prm->eps_surf;
        e->const_inf = (e->eps_surf < 1) ? 0.0 : 1.0;
        e->spherical = prm->spherical_sum;
        e->update_frequency = prm->update_frequency;
        e->alpha = prm->alpha;
        e->alpha2 = e->alpha * e->alpha;
        e->rc2 = prm->rc * prm->rc;
        e->kc = prm->kc;
        e->kappa = 1.0 / prm->debyelength;
        e->kc2 = e->kc * e->kc;
        e->kcc = (int)ceil(e->kc);
        e->eff = calloc((size_t)T, sizeof(double));
        e->qeqe = calloc((size_t)T * T, sizeof(double));
        /* effective charges, ewald.h:535-548 */
        for (int a = 0; a < T; ++a) {
            double qi = prm->tcharge[a], ri = prm->tradius[a];
            double qis = qi;
            if (fabs(e->kappa * ri) > 1e-9) qis *= sinh(ri * e->kappa) / ri / e->kappa;
            e->eff[a] = qis / qi;
            for (int b = 0; b < T; ++b) {
                double qj = prm->tcharge[b], rj = prm->tradius[b];
                double qjs = qj;
                if (fabs(e->kappa * rj) > 1e-9) qjs *= sinh(rj * e->kappa) / rj / e->kappa;
                e->qeqe[a * T + b] = qis * qjs / qi / qj;
            }
        }
        or_ewald_set_space(s);
    }
    or_mt_seed(&s->slump, 5489u);
    or_mt_seed(&s->internal, 5489u);
    return s;
}

void or_destroy(or_sys *s)
{
    if (!s) return;
    if (s->ew) {
        or_ewald *e = s->ew;
        free(e->kv); free(e->aks); free(e->Qre); free(e->Qim); free(e->Qtre); free(e->Qtim);
        free(e->eff); free(e->qeqe); free(e);
    }
    free(s->p); free(s->trial); free(s->g); free(s->s2); free(s->eps4); free(s->tdp);
    free(s);
}

void or_get_xyz(or_sys *s, int which, double *xyz)
{
    or_particle *p = vecp(s, which);
    for (int i = 0; i < s->n; ++i)
        for (int k = 0; k < 3; ++k) xyz[3 * i + k] = p[i].r[k];
}

void or_set_pos(or_sys *s, int which, int i, const double *xyz)
{
    or_particle *p = vecp(s, which);
    for (int k = 0; k < 3; ++k) p[i].r[k] = xyz[k];
}

void or_set_cm_trial(or_sys *s, int g, const double *cm_trial)
{
    memcpy(s->g[g].cm_trial, cm_trial, 3 * sizeof(double));
}

void or_get_cm(or_sys *s, int g, double *cm, double *cm_trial)
{
    if (cm) memcpy(cm, s->g[g].cm, 3 * sizeof(double));
    if (cm_trial) memcpy(cm_trial, s->g[g].cm_trial, 3 * sizeof(double));
}

/* ===================================================================== */
/* Moves                                                                   */
/* ===================================================================== */

int or_add_move(or_sys *s, int type, int molid, double prob, int peratom, int permol,
                const double *dir, const double *dir2, double dp, double dprot,
                int center_rotation)
{
    if (s->nmoves >= OR_MAX_MOVES) return -1;
    or_move *m = &s->moves[s->nmoves++];
    memset(m, 0, sizeof *m);
    m->type = type; m->molid = molid; m->prob = prob;
    m->peratom = peratom; m->permol = permol;
    for (int k = 0; k < 3; ++k) {
        m->dir[k] = dir ? dir[k] : 1.0;
        m->dir2[k] = dir2 ? dir2[k] : 1.0;
    }
    m->dp1 = dp;
    m->dp2 = dprot;
    if (m->dp2 > 4 * acos(-1.)) m->dp2 = 4 * acos(-1.);  /* move.h:1598-1599 */
    m->center_rotation = center_rotation;
    m->iparticle = -1;
    m->igroup = -1;
    return s->nmoves - 1;
}

void or_rng_seed(or_sys *s, uint32_t seed) { or_mt_seed(&s->slump, seed); }
int or_rng_set_text(or_sys *s, const char *txt) { return or_mt_from_text(&s->slump, txt); }

void or_propagator_init(or_sys *s)
{   /* move.h:7781: base::_slump().eng = slump.eng */
    s->internal = s->slump;
    s->propagator_started = 0;
}

int or_propagator_pick(or_sys *s, int nmoves)
{   /* Propagator::move's pick (move.h:7887): an element of its nmoves moves */
    return or_range(&s->internal, 0, nmoves - 1);
}

int or_propagator_peek(const or_sys *s, int nmoves)
{   /* the next pick, the engine left as it is */
    or_mt e = s->internal;
    return or_range(&e, 0, nmoves - 1);
}

/* groups with molId == molid in Space::molTrack order (= group order) */
static int find_molecules(or_sys *s, int molid, int *out)
{
    int c = 0;
    for (int k = 0; k < s->ng; ++k)
        if (s->g[k].molid == molid) out[c++] = k;
    return c;
}

static void record_trial(or_sys *s, int which)
{
    if (s->last_which && s->last_n < s->last_cap) s->last_which[s->last_n] = which;
    s->last_n++;
}

/* PointBase::ranunit, point.h:74-89 */
static void ranunit(or_mt *r, const double *dir, double *v)
{
    double r2;
    do {
        for (int i = 0; i < 3; ++i) v[i] = (2 * or_canonical(r) - 1) * dir[i];
        r2 = v[0] * v[0] + (v[1] * v[1] + v[2] * v[2]);
    } while (r2 > 1);
    double n = sqrt(r2);
    for (int i = 0; i < 3; ++i) v[i] = v[i] / n;
}

/* Geometry::QuaternionRotate (geometry.h:95-147, src/geometry.cpp:753-766)
 * with Eigen 3.2.9's AngleAxis->Quaternion and _transformVector */
typedef struct { double origin[3], w, v[3]; } or_qrot;

static void qrot_set(const or_sys *s, or_qrot *q, const double *beg, const double *end, double angle)
{
    double u[3] = {end[0] - beg[0], end[1] - beg[1], end[2] - beg[2]};
    memcpy(q->origin, beg, sizeof q->origin);
    box_boundary(&s->box, u);
    double n = sqrt(u[0] * u[0] + (u[1] * u[1] + u[2] * u[2]));
    for (int k = 0; k < 3; ++k) u[k] /= n;
    double ha = 0.5 * angle;
    q->w = cos(ha);
    double sh = sin(ha);
    for (int k = 0; k < 3; ++k) q->v[k] = sh * u[k];
}

static inline void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static void qrot_apply(const or_sys *s, const or_qrot *q, double *a)
{
    double t[3] = {a[0] - q->origin[0], a[1] - q->origin[1], a[2] - q->origin[2]};
    box_boundary(&s->box, t);
    double uv[3], c[3];
    cross3(q->v, t, uv);
    for (int k = 0; k < 3; ++k) uv[k] += uv[k];
    cross3(q->v, uv, c);
    for (int k = 0; k < 3; ++k) a[k] = ((t[k] + q->w * uv[k]) + c[k]) + q->origin[k];
    box_boundary(&s->box, a);
}

/* AtomicTranslation::_trialMove, move.h:818-864 */
static void at_trial(or_sys *s, or_move *m)
{
    int gvec[1024];
    int ngv = find_molecules(s, m->molid, gvec);
    m->igroup = gvec[or_range(&s->slump, 0, ngv - 1)];
    const or_group *g = &s->g[m->igroup];
    m->iparticle = g->front + or_range(&s->slump, 0, g->back - g->front);
    int i = m->iparticle;
    double dp = s->tdp[s->p[i].id];
    double t[3] = {m->dir[0] * dp, m->dir[1] * dp, m->dir[2] * dp};
    t[0] *= or_canonical(&s->slump) - 0.5;
    t[1] *= or_canonical(&s->slump) - 0.5;
    t[2] *= or_canonical(&s->slump) - 0.5;
    for (int k = 0; k < 3; ++k) s->trial[i].r[k] += t[k];
    box_boundary(&s->box, s->trial[i].r);
    or_group *gi = &s->g[find_group(s, i)];
    if (g_molecular(gi)) mass_center(s, s->trial, gi, gi->cm_trial);
    s->ch_n = 1;
    s->ch_g[0] = m->igroup;
    s->ch_off[0] = 0; s->ch_off[1] = 1;
    s->ch_idx[0] = i;
    record_trial(s, i);
}

/* TranslateRotate::_trialMove, move.h:1620-1686 */
static void tr_trial(or_sys *s, or_move *m)
{
    int gvec[1024];
    int ngv = find_molecules(s, m->molid, gvec);
    m->igroup = gvec[or_range(&s->slump, 0, ngv - 1)];
    or_group *g = &s->g[m->igroup];
    double p[3];
    if (m->dp2 > 1e-6) {
        ranunit(&s->slump, m->dir2, p);
        double center[3];
        memcpy(center, g->cm, sizeof center);
        if (!m->center_rotation) {
            if (or_canonical(&s->slump) < 0.5) memcpy(center, s->trial[g->front].r, sizeof center);
            else memcpy(center, s->trial[g->back].r, sizeof center);
        }
        for (int k = 0; k < 3; ++k) p[k] = center[k] + p[k];
        double angle = m->dp2 * (or_canonical(&s->slump) - 0.5);
        or_qrot q;
        qrot_set(s, &q, center, p, angle);          /* Group::newrotate, group.h:260-281 */
        qrot_apply(s, &q, g->cm_trial);
        for (int i = g->front; i <= g->back; ++i) qrot_apply(s, &q, s->trial[i].r);
    }
    if (m->dp1 > 1e-6) {
        p[0] = m->dir[0] * m->dp1 * (or_canonical(&s->slump) - 0.5);
        p[1] = m->dir[1] * m->dp1 * (or_canonical(&s->slump) - 0.5);
        p[2] = m->dir[2] * m->dp1 * (or_canonical(&s->slump) - 0.5);
        /* Group::translate, group.h:187-195 */
        for (int k = 0; k < 3; ++k) g->cm_trial[k] += p[k];
        box_boundary(&s->box, g->cm_trial);
        for (int i = g->front; i <= g->back; ++i) {
            for (int k = 0; k < 3; ++k) s->trial[i].r[k] += p[k];
            box_boundary(&s->box, s->trial[i].r);
        }
    }
    s->ch_n = 1;
    s->ch_g[0] = m->igroup;
    s->ch_off[0] = 0; s->ch_off[1] = 0;
    record_trial(s, m->igroup);
}

/* Movebase::move, move.h:515-555, plus the per-move overrides */
static double movebase_move(or_sys *s, or_move *m, double *du_out, int *acc_out, long cap,
                            long *ntr)
{
    double utot = 0;
    /* randomMolId (move.h:429-446): element over the mollist map (one entry) */
    (void)or_range(&s->internal, 0, 0);
    int gvec[1024];
    int nmol = find_molecules(s, m->molid, gvec);
    m->repeat = 1;
    if (m->permol) m->repeat *= nmol;
    if (m->peratom) m->repeat *= nmol > 0 ? g_size(&s->g[gvec[0]]) : 0;
    int n = m->repeat;
    /* AtomicTranslation::run (move.h:806-816) checks findMolecules first;
     * Movebase::run draws the runfraction canonical (move.h:575-581). */
    if (m->type == OR_MOVE_AT && nmol == 0) return 0;
    if (!(or_canonical(&s->internal) < m->prob)) return 0;
    while (n-- > 0) {
        if (m->cnt == 0)  /* Movebase::trialMove, move.h:466-475 */
            for (int k = 0; k < s->ng; ++k) {
                mass_center(s, s->p, &s->g[k], s->g[k].cm);
                memcpy(s->g[k].cm_trial, s->g[k].cm, sizeof s->g[k].cm);
            }
        m->cnt++;
        if (m->type == OR_MOVE_AT) at_trial(s, m); else tr_trial(s, m);
        or_ewald_update_change(s, s->ch_n, s->ch_g, s->ch_off, s->ch_idx);
        double du;
        if (m->type == OR_MOVE_TR && m->dp2 < 1e-6 && m->dp1 < 1e-6) du = 0;
        else du = or_energy_change(s, s->ch_n, s->ch_g, s->ch_off, s->ch_idx);
        /* Movebase::metropolis, move.h:567-573 */
        int accept = !(or_canonical(&s->slump) > exp(-du));
        if (m->type == OR_MOVE_AT) {
            int i = m->iparticle;
            or_group *gi = &s->g[find_group(s, i)];
            if (accept) {   /* move.h:866-877 */
                s->p[i] = s->trial[i];
                if (g_molecular(gi)) memcpy(gi->cm, gi->cm_trial, sizeof gi->cm);
            } else {        /* move.h:879-889 */
                s->trial[i] = s->p[i];
                if (g_molecular(gi)) memcpy(gi->cm_trial, gi->cm, sizeof gi->cm);
            }
        } else {
            or_commit(s, s->ch_n, s->ch_g, s->ch_off, s->ch_idx, accept);
        }
        if (accept) utot += du;
        utot += or_ewald_update(s, accept);
        if (du_out && *ntr < cap) { du_out[*ntr] = du; acc_out[*ntr] = accept; }
        (*ntr)++;
    }
    return utot;
}

/* Propagator::move, move.h:7877-7891 */
long or_propagate(or_sys *s, int ncalls, double *du_out, int *acc_out, long cap,
                  double *utot, double *uinit)
{
    long ntr = 0;
    double ut = 0;
    for (int c = 0; c < ncalls; ++c) {
        if (!s->propagator_started) {
            s->uinit = or_system_energy(s, 0);
            s->propagator_started = 1;
        }
        int k = or_range(&s->internal, 0, s->nmoves - 1);
        double du = movebase_move(s, &s->moves[k], du_out, acc_out, cap, &ntr);
        s->dusum += du;
        ut += du;
    }
    if (utot) *utot = ut;
    if (uinit) *uinit = s->uinit;
    return ntr;
}

long or_last_trials(or_sys *s, int *which_out, long cap)
{
    s->last_which = which_out;
    s->last_cap = cap;
    long n = s->last_n;
    s->last_n = 0;
    return n;
}

/* ===================================================================== */
/* ParallelTempering (move.h:6520-6772) restated for R ranks in one process */
/* ===================================================================== */

/* The decision of one exchange attempt, shared by both forms below.
 * findPartner (move.h:6621-6633): rank r pairs with r + dr (even r) or r - dr
 * (odd r); a partner outside [0, R) is no partner and _energyChange returns
 * +inf (move.h:6703-6704).  _energyChange returns (unew - uold) +
 * du_partner (move.h:6701-6723), i.e. du[r] = du_self[r] + du_self[partner].
 * metropolis (move.h:567-573) rejects iff u > exp(-du); each rank draws its
 * own u (uni[r]).  The reference takes the decision on both ranks
 * separately and relies on their global engines being in lock-step (the
 * precondition of move.h:557-566), so that both see the same u; here the
 * pair takes the lower rank's u, which is the same decision under that
 * precondition.  part[r] gets the partner (-1: none). */
#define OR_PT_MAXR 256

static void pt_decide(int R, int dr, const double *du_self, const double *uni, int *part,
                      int *acc, double *du)
{
    for (int r = 0; r < R; ++r) {
        int p = (r % 2 == 0) ? r + dr : r - dr;
        part[r] = (p >= 0 && p < R && p != r) ? p : -1;
    }
    for (int r = 0; r < R; ++r) {
        acc[r] = 0;
        if (part[r] < 0) {
            du[r] = INFINITY;
            continue;
        }
        du[r] = du_self[r] + du_self[part[r]];
    }
    for (int r = 0; r < R; ++r) {
        int p = part[r];
        if (p < 0 || p < r) continue;
        int a = !(uni[r] > exp(-du[r]));
        acc[r] = acc[p] = a;
    }
}

/* Energy-matrix form: u[r*R + c] is rank r's Hamiltonian on configuration
 * c, slot[r] the configuration rank r holds; du_self[r] = U_r(x_partner) -
 * U_r(x_own).  Every rank's global engine is the same (lock-step), so one
 * draw of `glob` is every rank's Metropolis uniform.  Returns the accepted
 * swaps; acc_out/du_out per rank (du = 0 for a rank without partner). */
int or_pt_exchange(int R, const double *u, int *slot, or_mt *ctl, or_mt *glob, int *acc_out,
                   double *du_out)
{
    int part[OR_PT_MAXR];
    double du_self[OR_PT_MAXR], uni[OR_PT_MAXR];
    if (R < 1 || R > OR_PT_MAXR) return -2;
    /* MPIController::random (mpi.h:48): default-seeded on every rank, so
     * the direction is shared */
    int dr = (or_canonical(ctl) > 0.5) ? 1 : -1;
    double g = or_canonical(glob);
    for (int r = 0; r < R; ++r) {
        int p = (r % 2 == 0) ? r + dr : r - dr;
        du_self[r] = (p >= 0 && p < R) ? u[r * R + slot[p]] - u[r * R + slot[r]] : 0.0;
        uni[r] = g;
    }
    pt_decide(R, dr, du_self, uni, part, acc_out, du_out);
    int nacc = 0;
    for (int r = 0; r < R; ++r) {
        if (part[r] < 0) { du_out[r] = 0.0; continue; }
        if (part[r] > r && acc_out[r]) {
            int t = slot[r]; slot[r] = slot[part[r]]; slot[part[r]] = t;
            nacc++;
        }
    }
    return nacc;
}

void or_pt_exchange_seeded(int R, const double *u, int *slot, uint32_t ctl_seed,
                           uint32_t glob_seed, int nsteps, int *acc_out, double *du_out)
{
    or_mt ctl, glob;
    or_mt_seed(&ctl, ctl_seed);
    or_mt_seed(&glob, glob_seed);
    for (int t = 0; t < nsteps; ++t)
        or_pt_exchange(R, u, slot, &ctl, &glob, acc_out + (long)t * R, du_out + (long)t * R);
}

or_mt *or_mt_new(uint32_t seed)
{
    or_mt *m = (or_mt *)malloc(sizeof(or_mt));
    if (m) or_mt_seed(m, seed);
    return m;
}

void or_mt_free(or_mt *m) { free(m); }

/* One ParallelTempering move on each of R ranks, as Movebase::move
 * (move.h:515-555) runs it when every rank calls it together.  Rank r is
 * the system rk[r]: its Hamiltonian (lB), its global and internal engines
 * and its Ewald state stay with it; configurations move between ranks
 * (ParticleTransmitter, mpi.h:238-331, format XYZQI).
 *   run(): one canonical of the rank's internal engine against prob
 *     (move.h:575-581; the move has no molecule list, so no randomMolId);
 *   _trialMove: findPartner from `ctl`; the partner's positions arrive in
 *     the trial vector and every group's cm_trial is its mass centre there
 *     (move.h:6648-6669);
 *   _energyChange: uold = systemEnergy(p), unew = systemEnergy(trial),
 *     both through Energy::systemEnergy, whose setSpace recomputes the
 *     Ewald S(k) from the rank's own p -- so unew's reciprocal term uses the
 *     rank's own S(k) (ewald.h:716-729, 731-757);
 *   metropolis on the rank's global engine; _acceptMove copies trial to p
 *     and cm_trial to cm; _rejectMove calls setSpace (S(k) recomputed from p)
 *     and copies p to trial; then pot->update(accepted) (ewald.h:613-648).
 * Volumes must agree (no Isobaric with tempering here) and so must the
 * particle compositions.  ret_out[r] = what Movebase::move returns for rank r
 * (unew - uold when accepted, move.h:6719-6722, plus the Ewald update drift).
 * Returns the number of accepted swaps, -1 if the ranks' run() draws
 * disagree (the reference would block), -2 for mismatched systems. */
int or_pt_step(int R, or_sys **rk, or_mt *ctl, double prob, int *acc_out, double *du_out,
               double *ret_out)
{
    int run0 = -1, part[OR_PT_MAXR];
    double du_self[OR_PT_MAXR], uni[OR_PT_MAXR];
    if (R < 1 || R > OR_PT_MAXR) return -2;
    for (int r = 0; r < R; ++r) {
        acc_out[r] = 0; du_out[r] = 0.0; ret_out[r] = 0.0;
        if (rk[r]->n != rk[0]->n) return -2;
        for (int k = 0; k < 3; ++k)
            if (rk[r]->box.len[k] != rk[0]->box.len[k]) return -2;
        for (int i = 0; i < rk[r]->n; ++i)
            if (rk[r]->p[i].charge != rk[0]->p[i].charge || rk[r]->p[i].id != rk[0]->p[i].id)
                return -2;
    }
    for (int r = 0; r < R; ++r) {
        int run = or_canonical(&rk[r]->internal) < prob;
        if (run0 < 0) run0 = run;
        else if (run != run0) return -1;
    }
    if (!run0) return 0;
    int dr = (or_canonical(ctl) > 0.5) ? 1 : -1;
    for (int r = 0; r < R; ++r) {
        int p = (r % 2 == 0) ? r + dr : r - dr;
        part[r] = (p >= 0 && p < R) ? p : -1;
    }
    /* every rank sends its accepted positions; each receives its partner's */
    for (int r = 0; r < R; ++r) {
        if (part[r] < 0) continue;
        const or_sys *src = rk[part[r]];
        for (int i = 0; i < rk[r]->n; ++i)
            memcpy(rk[r]->trial[i].r, src->p[i].r, sizeof(double) * 3);
        for (int g = 0; g < rk[r]->ng; ++g)
            mass_center(rk[r], rk[r]->trial, &rk[r]->g[g], rk[r]->g[g].cm_trial);
    }
    for (int r = 0; r < R; ++r) {
        du_self[r] = 0.0;
        if (part[r] < 0) continue;
        double uold = or_system_energy(rk[r], 0);
        double unew = or_system_energy(rk[r], 1);
        du_self[r] = unew - uold;
    }
    for (int r = 0; r < R; ++r) uni[r] = or_canonical(&rk[r]->slump);
    pt_decide(R, dr, du_self, uni, part, acc_out, du_out);
    int nacc = 0;
    for (int r = 0; r < R; ++r) {
        or_sys *s = rk[r];
        if (part[r] < 0) {
            ret_out[r] = or_ewald_update(s, 0);
            continue;
        }
        if (acc_out[r]) {
            for (int i = 0; i < s->n; ++i) memcpy(s->p[i].r, s->trial[i].r, sizeof(double) * 3);
            for (int g = 0; g < s->ng; ++g) memcpy(s->g[g].cm, s->g[g].cm_trial, sizeof(double) * 3);
            ret_out[r] = du_self[r];
            if (part[r] > r) nacc++;
        } else {
            if (s->ew) or_ewald_set_space(s);
            for (int i = 0; i < s->n; ++i) memcpy(s->trial[i].r, s->p[i].r, sizeof(double) * 3);
            for (int g = 0; g < s->ng; ++g) memcpy(s->g[g].cm_trial, s->g[g].cm, sizeof(double) * 3);
        }
        ret_out[r] += or_ewald_update(s, acc_out[r]);
    }
    return nacc;
}

/* ===================================================================== */
/* Move::Isobaric (move.h:5424-5640)                                       */
/* ===================================================================== */

void or_get_box(or_sys *s, double *len)
{
    memcpy(len, s->box.len, sizeof(double) * 3);
}

/* Group::scale (group.h:318-355) for a Cuboid scaled isotropically by xyz
 * (Cuboid::scale, src/geometry.cpp:326-332), after setMassCenter
 * (group.h:174-179); trial == p on entry. */
static int iso_scale_group(or_sys *s, or_group *g, double xyz, const double oldlen[3],
                           const double newlen[3])
{
    mass_center(s, s->p, g, g->cm);
    memcpy(g->cm_trial, g->cm, sizeof g->cm);
    if (g_empty(g)) return 0;
    if (g_atomic(g)) {
        for (int k = 0; k < 3; ++k) g->cm_trial[k] = g->cm[k] * 1.0 * xyz;
        for (int i = g->front; i <= g->back; ++i)
            for (int k = 0; k < 3; ++k) s->trial[i].r[k] = s->trial[i].r[k] * 1.0 * xyz;
        return 0;
    }
    if (g_molecular(g)) {
        double newcm[3], mcm[3];
        for (int k = 0; k < 3; ++k) {
            newcm[k] = g->cm[k] * 1.0 * xyz;
            mcm[k] = -g->cm[k];
        }
        or_box ob, nb;
        box_set(&ob, oldlen);
        box_set(&nb, newlen);
        /* translate(spc, -cm): Point::translate = += then boundary (point.h:95-100) */
        for (int k = 0; k < 3; ++k) g->cm_trial[k] += mcm[k];
        box_boundary(&ob, g->cm_trial);
        for (int i = g->front; i <= g->back; ++i) {
            for (int k = 0; k < 3; ++k) s->trial[i].r[k] += mcm[k];
            box_boundary(&ob, s->trial[i].r);
        }
        for (int i = g->front; i <= g->back; ++i) {
            for (int k = 0; k < 3; ++k) s->trial[i].r[k] += newcm[k];
            box_boundary(&nb, s->trial[i].r);
        }
        memcpy(g->cm_trial, newcm, sizeof newcm);
        return 0;
    }
    return -1;   /* several molecules in one group (isRange) */
}

/* One Movebase::move (move.h:515-555) of an Isobaric move with the given
 * run fraction, after the Propagator's pick: returns the call's utot
 * (accepted dU plus pot->update), records the trial at *nt, and sets *bad when
 * a group cannot be scaled.  mvg/off/idx are scratch of ng, ng+1, n ints.
 *
 * With NonbondedEwald every setSpace call (energyChange with the trial
 * geometry and back, _acceptMove, _rejectMove) regenerates the k-vectors and
 * A_k for the Space's current box and recomputes S(k) from the accepted
 * coordinates (ewald.h:731-757), then copies them to the trial set (undo).
 * updateChange's own geometry branch (ewald.h:653-658) is overwritten by the
 * first of these calls and leaves nothing behind. */
static double iso_move(or_sys *s, double dp, double runfraction, int *mvg, int *off, int *idx,
                       double *du_out, int *acc_out, long cap, long *nt, int *bad_out)
{
    const int ng = s->ng;
    double utot = 0.0;
    *bad_out = 0;
    if (!(or_canonical(&s->internal) < runfraction)) return 0.0;   /* Movebase::run */
    /* Isobaric::_trialMove (move.h:5529-5567) */
    s->iso_cnt++;
    double oldlen[3], newlen[3];
    memcpy(oldlen, s->box.len, sizeof oldlen);
    const double oldval = box_volume(&s->box);
    const double newval = exp(log(oldval) + (or_canonical(&s->slump) - 0.5) * dp);
    const double xyz = cbrt(newval / oldval);
    for (int k = 0; k < 3; ++k) newlen[k] = oldlen[k] * 1.0 * xyz;
    int nmv = 0, ni = 0, bad = 0;
    off[0] = 0;
    for (int g = 0; g < ng; ++g) {
        bad |= iso_scale_group(s, &s->g[g], xyz, oldlen, newlen);
        mvg[nmv] = g;
        if (g_atomic(&s->g[g]) && !g_empty(&s->g[g]))
            for (int i = s->g[g].front; i <= s->g[g].back; ++i) idx[ni++] = i;
        off[++nmv] = ni;
    }
    if (bad) { *bad_out = 1; return 0.0; }
    /* Energy::energyChange with a geometry change (energy.h:3577-3621): the
     * trial geometry and setSpace, the collision test (an early return that
     * leaves the trial geometry in place until _rejectMove), the trial
     * configuration, then the old geometry and setSpace */
    double du;
    box_set(&s->box, newlen);
    if (s->ew) or_ewald_set_space(s);
    int coll = 0;
    for (int i = 0; i < s->n && !coll; ++i) coll = box_collision(&s->box, s->trial[i].r);
    if (coll) {
        box_set(&s->box, oldlen);
        du = INFINITY;
    } else {
        const double dnew = change_config(s, 1, nmv, mvg, off, idx);
        box_set(&s->box, oldlen);
        if (s->ew) or_ewald_set_space(s);
        const double dold = change_config(s, 0, nmv, mvg, off, idx);
        du = dnew - dold;
    }
    /* Movebase::metropolis (move.h:567-573) */
    const int acc = !(or_canonical(&s->slump) > exp(-du));
    if (acc) {   /* Isobaric::_acceptMove (move.h:5570-5580): setSpace with the new
                  * box before the groups copy their trial coordinates */
        box_set(&s->box, newlen);
        if (s->ew) or_ewald_set_space(s);
        for (int i = 0; i < s->n; ++i) s->p[i] = s->trial[i];
        for (int g = 0; g < ng; ++g) memcpy(s->g[g].cm, s->g[g].cm_trial, sizeof s->g[g].cm);
        utot += du;
    } else {     /* Isobaric::_rejectMove (move.h:5582-5592) */
        if (s->ew) or_ewald_set_space(s);
        for (int i = 0; i < s->n; ++i) s->trial[i] = s->p[i];
        for (int g = 0; g < ng; ++g) memcpy(s->g[g].cm_trial, s->g[g].cm, sizeof s->g[g].cm);
    }
    /* pot->update(accepted): the Ewald accept counter and refresh (ewald.h:613-648) */
    utot += or_ewald_update(s, acc);
    record_trial(s, -1);
    if (*nt < cap) {
        if (du_out) du_out[*nt] = du;
        if (acc_out) acc_out[*nt] = acc;
    }
    (*nt)++;
    return utot;
}

/* Propagator::move calls (move.h:7877-7891) of a Propagator whose moves are
 * the registered ones with one Isobaric move inserted at position npt_index
 * (the input's key order); npt_index < 0 means the Isobaric move alone. */
static long propagate_with_isobaric(or_sys *s, double dp, double prob, int npt_index,
                                    int ncalls, double *du_out, int *acc_out, long cap,
                                    double *utot)
{
    const int ng = s->ng;
    int *mvg = malloc(sizeof(int) * (ng > 0 ? ng : 1));
    int *off = malloc(sizeof(int) * (ng + 1));
    int *idx = malloc(sizeof(int) * (s->n > 0 ? s->n : 1));
    long nt = 0;
    double ut = 0.0;
    const int nm = npt_index < 0 ? 1 : s->nmoves + 1;
    if (npt_index < 0) npt_index = 0;
    /* Isobaric constructor: runfraction = prob, 0 when dp < 1e-6 (move.h:5462-5464) */
    const double runfraction = dp < 1e-6 ? 0.0 : prob;
    for (int c = 0; c < ncalls; ++c) {
        if (!s->propagator_started) {   /* Propagator::move: uinit (move.h:7884-7885) */
            s->uinit = or_system_energy(s, 0);
            s->propagator_started = 1;
        }
        const int k = or_range(&s->internal, 0, nm - 1);   /* move.h:7887 */
        double du;
        if (k == npt_index) {
            int bad;
            du = iso_move(s, dp, runfraction, mvg, off, idx, du_out, acc_out, cap, &nt, &bad);
            if (bad) { nt = -1; break; }
        } else {
            du = movebase_move(s, &s->moves[k < npt_index ? k : k - 1], du_out, acc_out, cap,
                               &nt);
        }
        s->dusum += du;
        ut += du;
    }
    free(mvg); free(off); free(idx);
    if (utot) *utot = ut;
    return nt;
}

long or_isobaric_propagate(or_sys *s, double dp, double prob, int ncalls, double *du_out,
                           int *acc_out, long cap, double *utot)
{
    return propagate_with_isobaric(s, dp, prob, -1, ncalls, du_out, acc_out, cap, utot);
}

long or_propagate_npt(or_sys *s, double dp, double prob, int npt_index, int ncalls,
                      double *du_out, int *acc_out, long cap, double *utot)
{
    if (npt_index < 0 || npt_index > s->nmoves) return -1;
    return propagate_with_isobaric(s, dp, prob, npt_index, ncalls, du_out, acc_out, cap, utot);
}
