/*
 * oracle.h -- CPU restatement of the Faunus Metropolis Delta-U path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * engine is compared against; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product (libfgx.so and the
 * beofaunus_amd package) never links, loads or calls anything here.
 *
 * Plain C99, FP64 throughout, compiled with -ffp-contract=off so that every
 * add/mul rounds exactly as the reference's x86-64 build does.  Each
 * function names the reference file:line it restates (paths relative to the
 * reference tree).
 *
 * Pinning status (see tests/test_oracle_*.py):
 *   - RNG (mt19937, generate_canonical, uniform_int): pinned by the C++
 *     standard's known answer (10000th draw of seed 5489) and by golden
 *     draws of libstdc++ written by oracle/_ref.
 *   - Cuboid geometry, pair potentials, Nonbonded sums, energyChange,
 *     AtomicTranslation / TranslateRotate trajectories, NonbondedEwald:
 *     pinned by golden vectors of the compiled reference (oracle/_ref),
 *     including runs at the benchmark configurations' own sizes, by closed
 *     forms (lB known answers; the minimum image against the brute-force
 *     minimum over 27 images, tests/test_oracle_known_answers.py) and by
 *     the Ewald ion pair of the reference's unit test
 *     (src/examples/unittests.cpp:131-175, doi:10.1063/1.481216) at
 *     kappa -> 0.
 *   - ParallelTempering exchange: no reference run (the reference needs
 *     MPI, which is absent).  Pinned by hand values of _energyChange's sign
 *     and partner terms, by the stationary occupancy of a two-state toy
 *     (which a negated dU fails), and by swaps of identical replicas.
 */
#ifndef FGX_ORACLE_H
#define FGX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- RNG: libstdc++ mt19937 + distributions ---------------- */
typedef struct { uint32_t x[624]; int p; } or_mt;

void     or_mt_seed(or_mt *m, uint32_t s);
uint32_t or_mt_u32(or_mt *m);
double   or_canonical(or_mt *m);
int      or_range(or_mt *m, int a, int b);
/* parse / print the libstdc++ text state ("624 words then position") */
int      or_mt_from_text(or_mt *m, const char *txt);
int      or_mt_to_text(const or_mt *m, char *buf, int cap);
/* convenience for ctypes: fill out[] with draws of the given kind */
void     or_rng_draws(uint32_t seed, int kind, int a, int b, int n, double *out);

/* ---------------- system description ---------------- */
enum { OR_ELEC_NONE = 0, OR_ELEC_COULOMB = 1, OR_ELEC_DH = 2, OR_ELEC_EWALD = 3 };
enum { OR_SR_NONE = 0, OR_SR_HS = 1, OR_SR_LJ = 2 };
enum { OR_NB_PLAIN = 0, OR_NB_VECTOR = 1, OR_NB_CUTG2G = 2 };

typedef struct {
    int elec, sr, nbkind;
    double lB;            /* Bjerrum length (A) */
    double kappa;         /* DH inverse Debye length (1/A) */
    double g2g_cut;       /* NonbondedCutg2g mass-centre cutoff (A); <=0 = infinity */
    int ntypes;
    const double *sigma;  /* per type (A) */
    const double *eps;    /* per type (kT) */
    const double *tcharge;/* per type charge (Ewald effective charges) */
    const double *tradius;/* per type radius */
    const double *tdp;    /* per type AtomicTranslation dp (A) */
    /* Ewald */
    double alpha, rc, kc, debyelength, eps_surf;
    int spherical_sum, update_frequency;
    /* Potential::CutShift<LennardJonesLB> cutoff (A); <= 0 = uncut LJ */
    double lj_cut;
    /* Energy::ExternalPressure in the Hamiltonian (energy.h:1817-1894):
     * pressure in 1/A^3 (moves.isobaric.pressure in mM times 1.0_mM) */
    int has_pressure;
    double pressure;
} or_params;

typedef struct or_sys or_sys;

or_sys *or_create(const or_params *prm, const double box[3], int n,
                  const double *xyz, const double *charge, const double *radius,
                  const double *mw, const int *id,
                  int ng, const int *front, const int *back,
                  const int *molsize, const int *molid);
void    or_destroy(or_sys *s);

/* which: 0 = accepted vector p, 1 = trial vector */
void   or_get_xyz(or_sys *s, int which, double *xyz);
void   or_set_pos(or_sys *s, int which, int i, const double *xyz);
void   or_get_cm(or_sys *s, int g, double *cm, double *cm_trial);
void   or_set_cm_trial(or_sys *s, int g, const double *cm_trial);

/* ---------------- energies (Energy::Nonbonded family) ---------------- */
double or_pair(or_sys *s, int which, int i, int j);
double or_i2i(or_sys *s, int which, int i, int j);
double or_i2g(or_sys *s, int which, int g, int j);
double or_i2all(or_sys *s, int which, int i);
double or_g2g(or_sys *s, int which, int g1, int g2);
double or_g_internal(or_sys *s, int which, int g);
double or_external(or_sys *s, int which);
double or_g_external(or_sys *s, int which, int g);
double or_system_energy(or_sys *s, int which);

/* Energy::energyChange for a change given as moved groups; group k of the
 * change owns idx[off[k] .. off[k+1]) (empty = whole group moved). */
double or_energy_change(or_sys *s, int nmv, const int *mvg, const int *off, const int *idx);
/* accept / reject a change (copies trial->p or p->trial, cm handling) */
void   or_commit(or_sys *s, int nmv, const int *mvg, const int *off, const int *idx, int accepted);

/* ---------------- Ewald (Energy::NonbondedEwald) ---------------- */
int    or_ewald_nk(or_sys *s);
void   or_ewald_get(or_sys *s, int which, double *kvec, double *aks, double *Qre, double *Qim);
double or_ewald_update_change(or_sys *s, int nmv, const int *mvg, const int *off, const int *idx);
double or_ewald_update(or_sys *s, int accepted);
double or_ewald_self(or_sys *s, int which);
double or_ewald_surface(or_sys *s, int which);
double or_ewald_reciprocal(or_sys *s, int which);
void   or_ewald_set_space(or_sys *s);

/* ---------------- Moves / Propagator ---------------- */
enum { OR_MOVE_AT = 1, OR_MOVE_TR = 2 };
/* register one move (Propagator key order) acting on molecule molid */
int  or_add_move(or_sys *s, int type, int molid, double prob, int peratom, int permol,
                 const double *dir, const double *dir2, double dp, double dprot,
                 int center_rotation);
void or_rng_seed(or_sys *s, uint32_t seed);          /* global slump seed */
int  or_rng_set_text(or_sys *s, const char *txt);     /* global slump state */
/* Propagator construction copies the global engine into the move engine */
void or_propagator_init(or_sys *s);
/* the Propagator's move pick among nmoves moves from the system's move
 * engine (move.h:7887); returns the index */
int or_propagator_pick(or_sys *s, int nmoves);
int or_propagator_peek(const or_sys *s, int nmoves);
/* run `ncalls` Propagator::move() calls; records per-trial du and accept
 * flag (up to cap) and returns the number of trials run; *utot gets the sum
 * of returned energies and *uinit the initial system energy. */
long or_propagate(or_sys *s, int ncalls, double *du_out, int *acc_out, long cap,
                  double *utot, double *uinit);
/* per-trial extra record: moved index (AT) or group (TR) */
long or_last_trials(or_sys *s, int *which_out, long cap);

/* A Propagator holding one Move::Isobaric (move.h:5424-5640) with the
 * ExternalPressure term in the Hamiltonian: `ncalls` Propagator::move()
 * calls, each one volume trial.  Records dU and accept flags (up to cap);
 * returns the number of trials, or -1 if the system is not supported
 * (Ewald; groups holding several molecules). */
long or_isobaric_propagate(or_sys *s, double dp, double prob, int ncalls, double *du_out,
                           int *acc_out, long cap, double *utot);
/* Propagator::move calls of a Propagator holding the registered moves and one
 * Isobaric move at position npt_index of its move list (move.h:7877-7891) */
long or_propagate_npt(or_sys *s, double dp, double prob, int npt_index, int ncalls,
                      double *du_out, int *acc_out, long cap, double *utot);
void or_get_box(or_sys *s, double *len);

/* ParallelTempering (move.h:6520-6772) restated for R ranks in one process.
 * Pinned by hand values of the exchange energy and the stationary occupancy
 * of a two-state toy (tests/test_oracle_temper.py); no reference run (the
 * reference's ParallelTempering needs MPI, absent here).
 *
 * Energy-matrix form: u[r][c] is rank r's Hamiltonian on configuration c;
 * slot[r] is the configuration rank r holds.  One exchange attempt draws the
 * partner direction from `ctl` (MPIController::random) and one Metropolis
 * uniform from `glob`, standing for every rank's global engine in lock-step.
 * Returns the accepted swaps; acc_out[r] = 1 if r swapped, du_out[r] the
 * pair's energy change (0 without partner). */
int  or_pt_exchange(int R, const double *u, int *slot, or_mt *ctl, or_mt *glob, int *acc_out,
                    double *du_out);
void or_pt_exchange_seeded(int R, const double *u, int *slot, uint32_t ctl_seed,
                           uint32_t glob_seed, int nsteps, int *acc_out, double *du_out);
or_mt *or_mt_new(uint32_t seed);
void   or_mt_free(or_mt *m);
/* System form: one ParallelTempering move on every rank, rank r being rk[r]
 * with its own engines, Hamiltonian and Ewald state; configurations move
 * between ranks.  du_out[r] = +inf without partner.  ret_out[r] = the energy
 * Movebase::move returns on rank r.  Returns the accepted swaps, -1 if the
 * ranks' run() draws disagree, -2 for systems of different size, box or
 * composition. */
int  or_pt_step(int R, or_sys **rk, or_mt *ctl, double prob, int *acc_out, double *du_out,
                double *ret_out);

#ifdef __cplusplus
}
#endif
#endif
