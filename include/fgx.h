/*
 * fgx.h -- C ABI of the MI355X (gfx950) Metropolis Delta-U engine.
 *
 * One `fgx_system` is one Faunus `Space<Cuboid,PointParticle>` plus its
 * non-bonded Hamiltonian and registered moves, resident in HBM.  The entry
 * points below are the operations the reference exposes through C++ virtual
 * dispatch on `Energy::Energybase<Tspace>` and `Move::Movebase<Tspace>`; each
 * declaration names the reference interface it replaces (paths relative to
 * the reference tree).  Plain pointers and sizes only; no exceptions cross
 * this boundary.  Every call returns FGX_OK (0) or a negative FGX_E* code,
 * with a message retrievable through fgx_last_error().
 *
 * Units follow the reference: lengths in Angstrom, energies in kT, charges
 * in e.  `which` selects the particle vector: 0 = accepted (Space::p),
 * 1 = trial (Space::trial).
 *
 * Threading: a handle is used by one host thread at a time; calls on one
 * handle are ordered on that handle's HIP stream.
 */
#ifndef FGX_H
#define FGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FGX_OK 0
#define FGX_EARG (-1)        /* invalid argument */
#define FGX_EHIP (-2)        /* HIP runtime error */
#define FGX_ESTATE (-3)      /* call not valid in the current state */
#define FGX_ENOMEM (-4)      /* device allocation or HBM budget exceeded */
#define FGX_EUNSUPPORTED (-5)/* combination not implemented on the device */
#define FGX_EDEVICE (-6)     /* device-side check failed (e.g. bounded wait) */

/* electrostatic term: Potential::Coulomb (potentials.h:2515-2568),
 * Potential::DebyeHuckel (potentials.h:3303-3375),
 * Potential::EwaldReal + Energy::NonbondedEwald (ewald.h:35-130, 229-760) */
enum { FGX_ELEC_NONE = 0, FGX_ELEC_COULOMB = 1, FGX_ELEC_DH = 2, FGX_ELEC_EWALD = 3 };
/* short-range term: Potential::HardSphere (potentials.h:455-475),
 * Potential::LennardJonesLB (potentials.h:2003-2116, 1816-1831, 3891) */
enum { FGX_SR_NONE = 0, FGX_SR_HS = 1, FGX_SR_LJ = 2 };
/* energy class: Energy::Nonbonded (energy.h:781-951, sqdist),
 * Energy::NonbondedVector (energy.h:1062-1260, vdist),
 * Energy::NonbondedCutg2g (energy.h:1328-1444, mass-centre cutoff) */
enum { FGX_NB_PLAIN = 0, FGX_NB_VECTOR = 1, FGX_NB_CUTG2G = 2 };
/* pair arithmetic: FP64 throughout, or FP32 pair geometry/potentials with
 * FP64 accumulation, FP64 S(k) and FP64 Metropolis */
enum { FGX_FP64 = 0, FGX_MIXED = 1 };
/* moves: Move::AtomicTranslation (move.h:730-967),
 * Move::TranslateRotate (move.h:1505-1801) */
enum { FGX_MOVE_AT = 1, FGX_MOVE_TR = 2 };

typedef struct fgx_system fgx_system;

/* Hamiltonian parameters, as the reference derives them from its JSON
 * input (energy.nonbonded.*, atomlist.*); per-type arrays include the dummy
 * type 0 (species.h:494-504). */
typedef struct {
    int elec, sr, nbkind, precision;
    double lB;        /* Bjerrum length, pc::lB(epsr) (physconst.h:91-94) */
    double kappa;     /* DH inverse Debye length (src/potentials.cpp:770-782) */
    double g2g_cut;   /* NonbondedCutg2g cutoff_g2g (A); <= 0 means infinity */
    int ntypes;
    const double *type_sigma, *type_eps, *type_charge, *type_radius, *type_dp;
    /* NonbondedEwald (ewald.h:507-527) */
    double ewald_alpha, ewald_rc, ewald_kc, ewald_debyelength, ewald_eps_surf;
    int ewald_spherical_sum, ewald_update_frequency;
    /* LJ as Potential::CutShift<LennardJonesLB> (potentials.h:1780-1812):
     * zero beyond lj_cut (A), shifted by its value at lj_cut inside; <= 0
     * keeps the uncut LennardJonesLB.  With FGX_ELEC_EWALD and lj_cut equal
     * to ewald_rc every pair term vanishes beyond rc, and the mixed
     * AtomicTranslation sweep uses a cell list (fgx_set_cell_list). */
    double lj_cut;
    /* Energy::ExternalPressure (energy.h:1817-1894) in the Hamiltonian:
     * external() += P V - ln V, g_external() += -N ln V (N molecules of the
     * group); pressure in 1/A^3 (the input's mM times 1.0_mM) */
    int has_pressure;
    double pressure;
} fgx_hamiltonian;

typedef struct {
    long trials, accepted;
    double utot;        /* sum of accepted dU plus Ewald drift corrections */
    double dusum;       /* Propagator dusum */
    long rejected_inf;  /* trials rejected with dU = +inf (overlap) */
    long ewald_refresh; /* full S(k) recomputes */
    double seconds;     /* device time of the call */
} fgx_stats;

const char *fgx_last_error(void);
int fgx_device_info(char *name, int cap, int *cu_count, size_t *hbm_bytes);

/* Space construction + Space::load (space.h:807-880): particles as SoA,
 * groups as [front, back] ranges with molsize (1 = atomic) and molId. */
int fgx_create(const fgx_hamiltonian *h, const double box[3], int n,
               const double *xyz, const double *charge, const double *radius,
               const double *mw, const int *type,
               int ng, const int *front, const int *back,
               const int *molsize, const int *molid, fgx_system **out);
int fgx_destroy(fgx_system *s);
size_t fgx_device_bytes(const fgx_system *s);

/* explicit transfers (the handle never retains host pointers) */
int fgx_set_positions(fgx_system *s, int which, int first, int count, const double *xyz);
int fgx_get_positions(fgx_system *s, int which, double *xyz);
int fgx_get_group_cm(fgx_system *s, int g, double cm[3], double cm_trial[3]);
/* Group::cm / Group::cm_trial (group.h:55-56), e.g. after a host-side
 * Group::translate / newrotate of the trial vector.  A null pointer keeps
 * that centre at its current value (as the device sweeps left it). */
int fgx_set_group_cm(fgx_system *s, int g, const double cm[3], const double cm_trial[3]);

/* Energy::Energybase entry points (energy.h:54-224) */
int fgx_i2i(fgx_system *s, int which, int i, int j, double *u);        /* energy.h:842 */
int fgx_i2g(fgx_system *s, int which, int g, int j, double *u);        /* energy.h:847 */
int fgx_i2all(fgx_system *s, int which, int i, double *u);             /* energy.h:867 */
int fgx_g2g(fgx_system *s, int which, int g1, int g2, double *u);      /* energy.h:879 */
int fgx_g_internal(fgx_system *s, int which, int g, double *u);        /* energy.h:930 */
int fgx_g_external(fgx_system *s, int which, int g, double *u);        /* ewald.h:712 */
int fgx_i_external(fgx_system *s, int which, int i, double *u);        /* ewald.h:707 */
int fgx_external(fgx_system *s, int which, double *u);                 /* ewald.h:716 */
/* Energy::systemEnergy (energy.h:3486-3491 -> 169-180) */
int fgx_system_energy(fgx_system *s, int which, double *u);
/* fgx_system_energy without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops).  No
 * counterpart in the reference.  Its kernels' scratch (partial sums and the
 * NonbondedCutg2g group-pair mask) is one static device array per GPU of
 * 2 MiB + 64 KiB: systems needing more than 2^18 partial sums (about 65,000
 * particles) or more than 256 groups under NonbondedCutg2g get FGX_ENOMEM.
 * Calls on one GPU share the array, so calls that overlap on one GPU race. */
int fgx_system_energy_noalloc(fgx_system *s, int which, double *u);

/* Analysis::AtomRDF sampling (analysis.h:2117-2148): pair counts of types
 * {type_a, type_b} over i < j at bin int(r/dr + 0.5), r the minimum-image
 * distance; nbins must cover half the box diagonal.  Normalise as
 * PairFunctionBase::normalize (src/analysis.cpp:190-206). */
int fgx_rdf(fgx_system *s, int which, int type_a, int type_b, double dr, int nbins,
            unsigned long long *counts);

/* A Space::Change (space.h:287-318) restricted to one moved group:
 * `count` moved indices of group g, or count = 0 meaning the whole group
 * (TranslateRotate).  Trial coordinates must already be uploaded with
 * fgx_set_positions(which = 1). */
typedef struct {
    int group;
    int count;
    const int *index;
} fgx_change;

/* Energybase::updateChange (ewald.h:651-705): incremental trial S(k) */
int fgx_update_change(fgx_system *s, const fgx_change *c);
/* Energy::energyChange (energy.h:3577-3621) */
int fgx_energy_change(fgx_system *s, const fgx_change *c, double *du);
/* accept/reject bookkeeping of the moves (move.h:866-889, group.h:371-389)
 * followed by Energybase::update(accepted) (ewald.h:613-648); *drift gets
 * the value update() returns (Ewald refresh correction, else 0). */
int fgx_commit(fgx_system *s, const fgx_change *c, int accepted, double *drift);

/* ------------------------------------------------------------------------
 * Device-resident Propagator (move.h:7703-7926) and Movebase::move
 * (move.h:515-555).  Moves are registered in Propagator key order.  The
 * RNG is libstdc++'s mt19937 with generate_canonical<double,53> and
 * uniform_int_distribution (slump.h:24-126), replayed on the device so the
 * accept/reject sequence is the reference's.
 * ------------------------------------------------------------------------ */
int fgx_add_move(fgx_system *s, int type, int molid, double prob, int peratom, int permol,
                 const double dir[3], const double dir2[3], double dp, double dprot,
                 int center_rotation);
/* global slump engine (slump.h:126): seed, or full state (624 words + pos) */
int fgx_rng_seed(fgx_system *s, uint32_t seed);
int fgx_rng_set_state(fgx_system *s, const uint32_t words[624], int pos);
int fgx_rng_get_state(fgx_system *s, uint32_t words[624], int *pos);
/* the move-internal engine (Movebase::_slump, move.h:356-360): full state */
int fgx_rng_set_internal(fgx_system *s, const uint32_t words[624], int pos);
int fgx_rng_get_internal(fgx_system *s, uint32_t words[624], int *pos);
/* Propagator constructor: copies the global engine into the move engine
 * (move.h:7781) */
int fgx_propagator_init(fgx_system *s);
/* run `ncalls` Propagator::move() calls on the device.  Optional per-trial
 * record (dU and accept flag) up to `cap` trials. */
int fgx_propagate(fgx_system *s, int ncalls, fgx_stats *st, double *du_out, int *acc_out,
                  long cap);

/* Propagator::move() calls of a Propagator holding one Move::Isobaric
 * (move.h:5424-5640): each call draws the move, its run fraction (prob; 0
 * when dp < 1e-6), a volume trial V' = exp(ln V + dp (u - 1/2)) with every
 * group scaled isotropically (Group::scale, group.h:318-355), the energy
 * change at the trial geometry (energy.h:3577-3621) and the Metropolis
 * decision, from the system's RNG engines.  Needs has_pressure.  With Ewald
 * every setSpace the reference makes around a volume trial is followed
 * (ewald.h:731-757): k-vectors, A_k and V of the trial box, S(k) recomputed
 * from the accepted coordinates; an accepted trial leaves the S(k) of the
 * unscaled coordinates under the new k-vectors, and the Ewald accept counter
 * and refresh (ewald.h:613-648) advance.  The first call of a Propagator
 * (here or in fgx_propagate) recomputes S(k), as its uinit evaluation does.
 * Optional per-trial record (dU, accept) up to `cap`. */
int fgx_isobaric_propagate(fgx_system *s, double dp, double prob, int ncalls, fgx_stats *st,
                           double *du_out, int *acc_out, long cap);
/* Propagator::move() calls (move.h:7877-7891) of a Propagator holding the
 * system's one registered move (fgx_add_move) and one Move::Isobaric at
 * position npt_index (0 or 1) of its move list, i.e. the input's move keys in
 * sorted order ("atomtranslate" < "isobaric" < "moltransrot").  Each call
 * picks one of the two with the internal engine; volume trials run as in
 * fgx_isobaric_propagate, the other move's calls on the device sweep.
 * Records every trial (dU, accept) up to `cap`. */
int fgx_propagate_npt(fgx_system *s, double dp, double prob, int npt_index, int ncalls,
                      fgx_stats *st, double *du_out, int *acc_out, long cap);

/* Energybase::setSpace (energy.h:86-90) with the Space's box `len`
 * (Cuboid::setlen): sets the box, and for NonbondedEwald::setSpace
 * (ewald.h:731-757) the k-vectors, A_k and V of that box, with S(k)
 * recomputed from the accepted vector (trial S(k) = current).  The host
 * binding calls it wherever the reference calls setSpace. */
int fgx_set_space(fgx_system *s, const double len[3]);
/* energyChangeConfiguration (energy.h:3521-3572) of `which` for a change that
 * moves every group (Isobaric, move.h:5548-5565), at the current box */
int fgx_config_energy(fgx_system *s, int which, double *u);
/* accept/reject of a change that moved every group (Isobaric::_acceptMove /
 * _rejectMove, group.h:371-389: every group copied) and NonbondedEwald::update
 * (ewald.h:613-648); *drift = the refresh's energy correction */
int fgx_commit_all(fgx_system *s, int accepted, double *drift);
/* Cuboid side lengths (Cuboid::setlen after an accepted volume move) */
int fgx_get_box(const fgx_system *s, double len[3]);
/* NonbondedEwald's k-vectors in use (kVectorChange, ewald.h:330-378; the
 * "Wavefunctions" count of its info(), ewald.h:466); 0 without Ewald */
int fgx_ewald_k(const fgx_system *s, int *K);

/* Launch schedule of the device sweeps: workgroup size (256 or 512 threads
 * per chain) and partner-loop batch depth (4 or 8) of the mixed path; 0
 * picks the defaults.  No counterpart in the reference (a GPU schedule). */
int fgx_set_launch(fgx_system *s, int threads, int unroll);
int fgx_get_launch(fgx_system *s, int *threads, int *unroll);
/* Group-pair energy cache of the mixed TranslateRotate sweep, after
 * Energy::EnergyMatrix (energy.h:225-641): the accepted state's g2g energies
 * are kept per group pair (ng <= 256) so only the trial state is evaluated.
 * On by default; any change of the accepted state from the host rebuilds it. */
int fgx_set_energy_matrix(fgx_system *s, int on);
/* Cell list of the mixed AtomicTranslation sweep, used when every pair term
 * is cut at the Ewald rc (lj_cut == ewald_rc) in a cubic box with at least 3
 * cells of side rc + the largest displacement per axis: each trial visits the
 * 27 cells around the moved particle instead of all N partners.  The same
 * Hamiltonian as the all-pairs sweep; the reference's NonbondedEwald has no
 * cell list.  mode 0 = off, 1 = automatic capacity (default), 2..4096 =
 * records per cell.  A chain whose cell fills up finishes its launch on the
 * all-pairs screen and rebuilds the list (automatic mode: twice as large). */
int fgx_set_cell_list(fgx_system *s, int mode);
/* cells per axis and records per cell of the last sweep (0, 0: no list) */
int fgx_get_cell_list(const fgx_system *s, int *cells_per_axis, int *capacity);
/* Mixed AtomicTranslation sweep without Ewald: a chain of up to 2,048
 * particles keeps its partners (packed record, FP64 position, charge,
 * radius, type) in LDS for the launch.  The same pair arithmetic as the
 * device-memory loop; partners are visited in other batches, so dU agrees
 * to FP64 rounding of the sum.  on = 1 (default) or 0 (partners read from
 * device memory).  A schedule knob with no reference counterpart. */
int fgx_set_lds_partners(fgx_system *s, int on);
/* Ewald AtomicTranslation sweep: keep S(k), the trial's dS(k), A_k and the
 * k-vector triples in LDS for the launch when they fit (hbm = 0, default) or
 * in device memory (hbm = 1; always so beyond the LDS capacity), which frees
 * the LDS for a second workgroup per CU.  The same arithmetic either way.
 * A schedule knob with no reference counterpart. */
int fgx_set_ewald_state(fgx_system *s, int hbm);
/* Block AtomicTranslation sweep: B consecutive trials of a Propagator call
 * (their draws do not depend on earlier decisions) are evaluated against one
 * snapshot in a single pass over the partners and decided in order with
 * exact corrections for the block's own moves; the accept sequence is the
 * one-trial sweep's, energy changes agree to FP64 rounding.  mode 1 uses it
 * where an instance exists (Debye-Hueckel / Coulomb with or without hard
 * spheres, Lennard-Jones in FP64, Ewald + Lennard-Jones; no cell list), 0
 * keeps the one-trial sweep, 2 (default) is 1 except for batches of mixed
 * Ewald chains that all keep one workgroup with partners in device memory,
 * which run one trial at a time.  nsplit > 1 spreads one chain
 * over that many workgroups (at most 32), which exchange partial sums once
 * per block; 0 chooses automatically.  A launch whose chains cannot all be
 * co-resident is refused (FGX_EUNSUPPORTED); a part that waits too long for
 * the others stops the launch with FGX_EDEVICE.  A split launch should not
 * run beside other kernels that hold compute units for long (the job server
 * runs its split jobs' launch alone).  A schedule knob with no reference
 * counterpart. */
int fgx_set_block_sweep(fgx_system *s, int mode, int nsplit);
/* workgroups per chain of the last AtomicTranslation device sweep if it was a
 * block sweep, 0 if it was the one-trial sweep */
int fgx_get_block_parts(const fgx_system *s, int *parts);
/* Autotuner: times `ncalls` Propagator::move() calls of a probe chain under
 * every schedule and keeps the fastest on the probe (whose state advances). */
int fgx_autotune(fgx_system *probe, int ncalls, int *threads, int *unroll, double *best_seconds);
/* the same over a batch of chains advanced together (fgx_propagate_batch);
 * the winning schedule is set on every chain */
int fgx_autotune_batch(fgx_system **systems, int nsys, int ncalls, int *threads, int *unroll,
                       double *best_seconds);

/* ------------------------------------------------------------------------
 * Batched chains: many independent systems advanced by one persistent
 * launch (one workgroup per chain).  Systems keep their own state.
 * ------------------------------------------------------------------------ */
int fgx_propagate_batch(fgx_system **systems, int nsys, int ncalls, fgx_stats *st);
/* the same with fgx_propagate's per-trial record for every chain: chain c
 * writes its first min(cap, trials) dU values and accept flags to
 * du_out[c cap ...], acc_out[c cap ...] and that count to counts[c]. */
int fgx_propagate_batch_record(fgx_system **systems, int nsys, int ncalls, fgx_stats *st,
                               double *du_out, int *acc_out, long cap, long *counts);
/* wall time this chain spent inside the last launch that advanced it
 * (device clock; chains of a batch finish at different times).  No
 * counterpart in the reference (per-chain sweep latency). */
int fgx_chain_seconds(const fgx_system *s, double *seconds);

/* ------------------------------------------------------------------------
 * Job server: independent simulations of different size and Hamiltonian
 * served on one device.  Jobs are admitted in arrival order while their
 * device footprint fits the HBM budget, advance together in batched
 * launches (one per kernel kind, concurrent streams), and release their
 * memory when done.  A job whose chain runs the block AtomicTranslation
 * sweep takes one workgroup per 512 particles, at most 24 (fixed at
 * admission, so that its result does not depend on the jobs beside it); the
 * jobs of more than one workgroup of a tick are packed onto XCDs so that the
 * workgroups of each are resident together, and one that does not fit waits
 * for a later tick.  No counterpart in the reference (one simulation per
 * process there).
 * ------------------------------------------------------------------------ */
typedef struct fgx_server fgx_server;
typedef struct {
    const fgx_hamiltonian *h;              /* copied at submission */
    double box[3];
    int n;
    const double *xyz, *charge, *radius, *mw;
    const int *type;
    int ng;
    const int *front, *back, *molsize, *molid;
    /* the job's one move, as fgx_add_move */
    int move_type, move_molid, peratom, permol, center_rotation;
    double prob, dp, dprot;
    const double *dir, *dir2;              /* NULL: (1,1,1) */
    uint32_t seed;                         /* global slump seed (fgx_rng_seed) */
    int sweeps;                            /* Propagator::move() calls to run */
    int keep_positions;                    /* keep final positions for fgx_server_result */
} fgx_job;
typedef struct {
    int status;                            /* FGX_OK or the error that ended the job */
    int sweeps;
    long trials, accepted;
    double utot;                           /* sum of accepted dU + Ewald drift terms */
    double queue_seconds;                  /* submission to admission */
    double latency_seconds;                /* submission to completion */
    double mean_sweep_seconds, max_sweep_seconds;   /* device time per sweep */
    size_t device_bytes;
    int workgroups;                        /* per launch of its block sweeps (0: one-trial sweep) */
} fgx_job_result;
int fgx_server_create(size_t hbm_budget, int max_active, double tick_seconds, fgx_server **out);
int fgx_server_submit(fgx_server *sv, const fgx_job *job, int *id);
/* admit what fits, advance every resident job by one tick, retire finished.
 * A job the device sweeps cannot run is retired at admission with its error
 * (fgx_server_result status); the call itself fails only on runtime errors. */
int fgx_server_step(fgx_server *sv, int *n_admitted, int *n_finished);
int fgx_server_poll(fgx_server *sv, int *ids, int cap, int *count);
int fgx_server_result(fgx_server *sv, int id, fgx_job_result *r, double *xyz);
int fgx_server_info(const fgx_server *sv, size_t *resident, size_t *peak, int *active,
                    int *queued);
int fgx_server_destroy(fgx_server *sv);

/* ------------------------------------------------------------------------
 * Move::ParallelTempering (move.h:6520-6772) for R ranks on one device.  Rank
 * r is replica r's Hamiltonian (the replicas differ in Bjerrum length only:
 * temperature enters through lB, physconst.h:91-94), and what belongs to a
 * rank in the reference -- its lB, its global and internal engines
 * (slump.h:126, move.h:356-360) and its Ewald accept counter -- travels with
 * it; configurations stay in their handles.  An exchange attempt is one
 * ParallelTempering move on every rank (Movebase::move, move.h:515-555): the
 * run() draw of each rank's internal engine, the partner direction from a
 * controller engine shared by all ranks (MPIController::random, mpi.h:48),
 * du = (U_r(x_p) - U_r(x_r)) + (U_p(x_r) - U_p(x_p)) (_energyChange,
 * move.h:6701-6723) and the Metropolis draw of each rank's global engine; a
 * pair decides with its lower rank's draw, the reference's own decision when
 * the ranks' engines are in lock-step (move.h:557-566).  With Ewald the
 * partner energy's reciprocal term uses the rank's own S(k), as the
 * reference's systemEnergy does after setSpace (ewald.h:716-757).
 * ------------------------------------------------------------------------ */
typedef struct fgx_temper fgx_temper;
/* ParallelTempering constructor (move.h:6592-6613); prob = its runfraction
 * ("prob" key).  Replicas must agree in everything but lB (FGX_EUNSUPPORTED
 * otherwise; ExternalPressure and NonbondedCutg2g are not supported). */
int fgx_temper_create(fgx_system **replicas, int R, uint32_t ctl_seed, double prob,
                      fgx_temper **out);
/* one exchange attempt (one ParallelTempering::move on every rank);
 * acc[r] = 1 if rank r swapped; du[r] = the pair's energy change, +inf for a
 * rank without partner.  FGX_ESTATE if the ranks' run() draws disagree. */
int fgx_temper_exchange(fgx_temper *t, int *acc, double *du);
/* `rounds` times: `calls` Propagator::move() calls of every configuration's
 * registered move (one batched sweep of all tempers' configurations), then
 * one exchange attempt in each temper; the round is captured once as a HIP
 * graph and replayed.  Not for Ewald replicas.  *seconds = device time. */
int fgx_temper_run(fgx_temper **tempers, int nt, int rounds, int calls, fgx_stats *st,
                   double *seconds);
/* records of the last exchange / run call: *rounds, then per round and rank
 * acc, du and slot (after the round), and per round and configuration its
 * sweep time in seconds (run only).  Any output may be null. */
int fgx_temper_records(fgx_temper *t, int *rounds, int *acc, double *du, int *slot,
                       double *sweep_seconds);
/* slot[r] = index of the configuration (replica handle) rank r holds */
int fgx_temper_slots(fgx_temper *t, int *slot);
/* what each rank's ParallelTempering::move returned in the last exchange
 * (Movebase::move, move.h:515-555): after an accepted swap the rank's own
 * unew - uold (useAlternativeReturnEnergy, move.h:6604) plus the Ewald
 * update's energy change, else 0 */
int fgx_temper_returns(fgx_temper *t, double *ret);
int fgx_temper_destroy(fgx_temper *t);

#ifdef __cplusplus
}
#endif
#endif
