// fgx_internal.hpp -- host-side definition of an fgx_system handle and the
// helpers shared by the translation units of libfgx.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fgx.h"
#include "fgx_types.hpp"

namespace fgx {

void set_error(const std::string &msg);

#define FGX_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::fgx::set_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + \
                             __FILE__ + ":" + std::to_string(__LINE__) + " in " + #call); \
            return FGX_EHIP;                                                            \
        }                                                                               \
    } while (0)

#define FGX_CHECK(cond, code, msg)          \
    do {                                    \
        if (!(cond)) {                      \
            ::fgx::set_error(msg);          \
            return (code);                  \
        }                                   \
    } while (0)

// Device view of a system's reciprocal-space state (NonbondedEwald,
// ewald.h:229-760).  k-vectors are stored both as FP64 vectors (the
// reference's kVectors) and as integer triples for the factorised
// exp(i k.r) = prod_axis exp(i k_a r_a) used by the device sweeps.
struct EwaldDev {
    int K, kcc;
    double *kv;     // 3K
    double *aks;    // K
    int *kn;        // 3K integer triples
    double *Q;      // 2K (re, im) current
    double *Qt;     // 2K trial
    double *eff;    // per type
    double V;
    double const_inf, eps_surf, alpha, kappa;
    int update_frequency, cnt_accepted;
    double surf, surf_trial, reci, reci_trial, update_drift;
    int *cnt_dev = nullptr;   // device copy of cnt_accepted for the sweeps
    // Chebyshev series of erfcx(x) = exp(x^2) erfc(x) on [0, cheb_xmax] for
    // the real-space pairs of the mixed path's FP64 pass (plain Ewald only)
    double *cheb = nullptr;
    int ncheb = 0;
    double cheb_xmax = 0;
    // device sweeps with more k-vectors than LDS holds keep the trial's
    // dS(k) (2K doubles) and the packed n-triples (K ints) here
    double *sk_dS = nullptr;
    int *sk_kid = nullptr;
};

static constexpr int EW_CHEB_MAX = 48;
// records past the end of the partner arrays (packed records, types, FP32
// radii), zero-filled, so batched partner loads need no clamping: the
// double-buffered loop reads up to three batches (3 x unroll x block) beyond
// the last partner
static constexpr int PARTNER_PAD = 3 * 8 * 512;

// device pointers of one particle vector (FP64 SoA)
struct Vec64 {
    double *x, *y, *z;
};

}  // namespace fgx

struct fgx_system {
    int device = 0;
    hipStream_t stream = nullptr;
    fgx::Box box{};
    fgx::Pot pot{};
    int n = 0, ng = 0, ntypes = 0;
    // host mirrors
    std::vector<fgx::GroupDesc> groups;
    std::vector<double> cm, cm_trial;      // 3 * ng
    std::vector<double> h_charge, h_radius, h_mw;
    std::vector<int> h_type, h_gid;
    std::vector<double> type_dp, type_charge, type_radius;
    std::vector<double> tab;               // TAB_COUNT * ntypes^2
    // device particle data
    fgx::Vec64 v[2]{};                     // 0 = p, 1 = trial
    double *d_q = nullptr, *d_r = nullptr, *d_mw = nullptr;
    int *d_type = nullptr, *d_gid = nullptr;
    int4 *d_pk = nullptr;                  // packed accepted vector: fixed-point x, y, z + FP32 charge
    float *d_qf = nullptr, *d_rf = nullptr;
    fgx::GroupDesc *d_groups = nullptr;
    double *d_cm = nullptr;                // 6 * ng: cm, cm_trial
    double *d_tab = nullptr;
    float *d_tabf = nullptr;
    double *d_type_dp = nullptr;
    unsigned char *d_mask = nullptr;       // 2 * ng keep masks (new, old)
    double *d_partial = nullptr;           // reduction scratch
    double *d_stage = nullptr;             // 3n: coordinates as xyz triples (host transfers)
    int partial_cap = 0;
    double *d_out = nullptr;               // 8 result slots
    double *h_out = nullptr;               // pinned mirror
    int *d_idx = nullptr;                  // moved-index scratch
    int idx_cap = 0;
    // Ewald
    bool has_ewald = false;
    fgx::EwaldDev ew{};
    std::vector<double> h_eff;
    std::vector<int> h_kn;                 // k-vector index triples (3K)
    // moves and RNG
    std::vector<fgx::MoveDesc> moves;
    std::vector<long> move_cnt;
    uint32_t rng_global[625]{};            // 624 words + position
    uint32_t rng_internal[625]{};
    uint32_t *d_rng = nullptr;             // 2 x 625 words (global, internal)
    bool propagator_started = false;
    double uinit = 0, dusum = 0;
    long *d_stats_i = nullptr;
    double *d_stats_d = nullptr;
    size_t device_bytes = 0;
    int launch_threads = 0;                // 0 = default; set by fgx_set_launch / fgx_autotune
    int launch_unroll = 0;
    // The device copies of the RNG states and the Ewald accept counter are the
    // current ones after a device sweep (state_on_dev); host-side users pull
    // them first (pull_device_state)
    bool state_on_dev = false;
    double chain_seconds = 0;              // this chain's wall time in the last device launch
    int wall_khz = 0;
    int *d_gvec = nullptr;                 // move's group list (device sweeps)
    int *d_resume = nullptr;               // unfinished Propagator::move() call (4 ints)
    int gvec_n = 0;
    // per-batch launch buffers owned by the batch's first system
    void *d_batch = nullptr;               // ChainDev array
    int batch_cap = 0;
    long *d_bstat_i = nullptr;             // 8 per chain
    double *d_bstat_d = nullptr;
    // group-pair energy cache of the mixed TranslateRotate sweep (ng x ng)
    double *d_emat = nullptr;
    bool emat_on = true;                   // fgx_set_energy_matrix
    bool emat_valid = false;               // cleared whenever the accepted state or lB changes
    // cell list of the mixed AtomicTranslation sweep when every pair term is
    // cut at the Ewald rc (fgx_set_cell_list): mode 0 off, 1 automatic
    // capacity, >= 2 that many records per cell
    int cell_mode = 1;
    // mixed AtomicTranslation sweeps without Ewald stage the partners in LDS
    // when they fit (fgx_set_lds_partners): 1 on, 0 read them from device memory
    int lds_partners = 1;
    // where the AtomicTranslation sweep keeps the Ewald state (S(k), dS(k),
    // A_k, n) during a launch (fgx_set_ewald_state): 0 LDS when it fits, 1
    // device memory
    int ewald_state_hbm = 0;
    int cell_nc = 0, cell_cap = 0;         // cells per axis, records per cell (allocated)
    int cell_cap_min = 0;                  // automatic capacity after an overflow
    int *d_cell_cnt = nullptr;             // nc^3 counts + overflow word
    double4 *d_cell_rec = nullptr;         // nc^3 * cap records (FP64 x, y, z, charge)
    int *d_cell_idx = nullptr;             // nc^3 * cap particle indices
    int *d_cell_slot = nullptr;            // n: cell * cap + slot
    size_t cell_bytes = 0;
    bool cell_valid = false;               // cleared whenever the accepted state changes
    // block AtomicTranslation sweep (fgx_set_block_sweep): 0 off, 1 where an
    // instance exists, 2 automatic (1, except batches of unsplit mixed Ewald
    // chains with partners in device memory); workgroups per chain (0:
    // automatic)
    int blk_mode = 2;
    int blk_nsplit = 0;
    // XCD that holds all of the chain's workgroups (-1: chosen per launch);
    // set by callers that run several block-sweep launches at once and keep
    // every XCD's split workgroups within its compute units
    int blk_xcd = -1;
    int blk_parts = 0;                     // workgroups of the last block sweep (0: one-trial sweep)
    double *d_xbuf = nullptr;              // exchange slots of the chain's parts
    int *d_xcnt = nullptr;                 // arrival / exit counters (zero between launches)
    int *d_xerr = nullptr;                 // launch error word (batch owner)
    int2 *d_blkmap = nullptr;              // workgroup -> (chain, part) of the batch (owner)
    int blkmap_cap = 0;
    void *d_mirror = nullptr;              // trial-mirror copy jobs of the batch (owner)
    int mirror_cap = 0;
    // Energy::ExternalPressure term (1/A^3)
    bool press_on = false;
    double press = 0.0;
};

namespace fgx {
// implemented in fgx_system.hip, used by fgx_sweep.hip
int ensure_partials(fgx_system *s, int nblocks);
int ewald_full_recompute(fgx_system *s, int which, double *Qdst);
int ewald_set_space(fgx_system *s);
// NonbondedEwald::update(true) (ewald.h:613-648) on the host-side counter:
// advances cnt_accepted, refreshes S(k) from the trial vector when due
// (*drift = the reciprocal-energy correction), and makes Q = Q_trial
int ewald_accepted(fgx_system *s, double *drift);
int ewald_reciprocal(fgx_system *s, const double *Q, double V, double *out_host);
void mt_seed_host(uint32_t *state625, uint32_t seed);
int pull_device_state(fgx_system *s);
// asynchronous batched sweeps (fgx_sweep.hip), used by the job server:
// chain c runs up to ncalls[c] Propagator::move() calls and at most
// max_trials[c] trials (> 0); a call cut short resumes in the next launch.
// Statistics slot 5 holds the calls completed.  Block sweeps take each
// chain's blk_nsplit workgroups (on XCD blk_xcd when set).
int batch_launch(std::vector<fgx_system *> &sys, const std::vector<int> &ncalls,
                 const std::vector<long> &max_trials, void **handle);
// Workgroups of a job-server chain's block sweeps: fixed by the chain alone
// (one per 512 partner rows, at most cap); 0 when it cannot run them.
int blk_job_parts(const fgx_system *s, int cap);
int batch_finish(void *handle, double *seconds, std::vector<long> *stats_i,
                 std::vector<double> *stats_d);
// the device sweeps' checks of a system and its registered move (the ones a
// launch makes), without a launch: FGX_OK or the error the launch would give
int check_chain(const fgx_system *s);
}  // namespace fgx
// folds one chain's launch statistics (8 longs, 8 doubles) into the system
// and into *st (may be null)
void collect_stats(fgx_system *s, const long *si, const double *sd, fgx_stats *st,
                   double seconds);
namespace fgx {
}  // namespace fgx
