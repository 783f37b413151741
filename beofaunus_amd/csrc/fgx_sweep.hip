// fgx_sweep.hip -- host side of the device-resident Propagator::move
// (move.h:7877-7891) and Movebase::move (move.h:515-555): chain preparation,
// batched launches, RNG state, schedules and the exported entry points.  The
// kernels live in fgx_sweep_kernels.hpp and are instantiated per Hamiltonian
// family in fgx_sweep_at_*.hip and fgx_sweep_tr.hip.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "fgx_sweep_blk.hpp"

namespace fgx {

// Builds a chain's cell list from its accepted vector (one block per chain;
// cells from the fixed-point coordinates, records in FP64).  Slots follow particle order within each cell, so the list (and
// the order of every later sum over it) is deterministic.  A cell that would
// exceed `cell_cap` sets the overflow word cell_cnt[nc^3].
__global__ __launch_bounds__(256) void k_cell_build(const ChainDev *__restrict__ chains)
{
    const ChainDev &C = chains[blockIdx.x];
    if (C.cell_nc <= 0) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int nc = C.cell_nc, ncell = nc * nc * nc, cap = C.cell_cap;
    extern __shared__ int bcnt[];
    for (int c = tid; c < ncell; c += blockDim.x) bcnt[c] = 0;
    __syncthreads();
    if (tid < 64) {
        int over = 0;
        for (int j0 = 0; j0 < C.n; j0 += 64) {
            const int j = j0 + lane;
            int4 p = make_int4(0, 0, 0, 0);
            int c = -1;
            if (j < C.n) {
                p = C.pk[j];
                c = cell_of(p.x, p.y, p.z, nc);
            }
            // rank among lower lanes of this chunk in the same cell
            int below = 0, above = 0;
            for (int l = 0; l < 64; ++l) {
                const int o = __shfl(c, l, 64);
                below += (l < lane && o == c) ? 1 : 0;
                above += (l > lane && o == c) ? 1 : 0;
            }
            int slot = 0;
            if (c >= 0) slot = bcnt[c] + below;
            wave_sync_lds();
            if (c >= 0) {
                if (slot < cap) {
                    C.cell_rec[c * cap + slot] = make_double4(C.x[j], C.y[j], C.z[j], C.q[j]);
                    C.cell_idx[c * cap + slot] = j;
                    C.cell_slot[j] = c * cap + slot;
                } else {
                    over = 1;
                }
                if (above == 0) bcnt[c] = slot + 1;   // last of its cell in this chunk
            }
            wave_sync_lds();
        }
        over = __ballot(over) != 0ull;
        if (lane == 0) C.cell_cnt[ncell] = over;
    }
    __syncthreads();
    for (int c = tid; c < ncell; c += blockDim.x) C.cell_cnt[c] = min(bcnt[c], cap);
}


// AtomicTranslation sweep launch for the chains' Hamiltonian
template <bool MIX>
static int launch_at(const Pot &P, hipStream_t st, const ChainDev *d_chains, int nchains, int nthr,
                     size_t dyn, int ncalls, int unroll, bool fx, bool cell, bool skg)
{
    int rc;
    if (P.elec == ELEC_EWALD)
        rc = MIX ? launch_at_ewald_mixed(P, st, d_chains, nchains, nthr, dyn, ncalls, unroll, fx, cell,
                                         skg)
                 : launch_at_ewald_fp64(P, st, d_chains, nchains, nthr, dyn, ncalls, skg);
    else if (P.elec == ELEC_DH)
        rc = launch_at_dh(MIX, P, st, d_chains, nchains, nthr, dyn, ncalls, unroll, fx, cell);
    else
        rc = launch_at_plain(MIX, P, st, d_chains, nchains, nthr, dyn, ncalls, unroll, fx, cell);
    if (rc == FGX_EUNSUPPORTED) {
        set_error("device sweep: no kernel for this Hamiltonian");
        return rc;
    }
    FGX_HIP(hipGetLastError());
    return rc;
}

// ------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------

static int chain_block(const fgx_system *s)
{
    if (s->n >= 2048) return AT_THREADS;
    return 256;
}

struct Prepared {
    std::vector<int> gvec;
    int *d_gvec = nullptr;
};

static constexpr size_t AT_DYN_LDS_MAX = 110 * 1024;
// dynamic LDS for the partner slice of a mixed AtomicTranslation chain
// without Ewald (the sweep's static LDS is under 29 KiB of the CU's 160 KiB)
static constexpr size_t AT_LDS_PART_MAX = 128 * 1024;
// dynamic LDS for the Ewald state of one chain (the static LDS of the
// mixed sweep leaves this much of the CU's 160 KiB); more k-vectors than
// this holds keep S(k), dS(k), A_k and n in device memory instead
static constexpr size_t AT_DYN_LDS_EWALD_MAX = 110 * 1024;
static bool ewald_in_hbm(int K) { return at_dyn_lds(K) > AT_DYN_LDS_EWALD_MAX; }
// LDS bytes of the Ewald state of an AtomicTranslation chain (none when it
// lives in device memory)
static size_t at_dyn_lds_ewald(int K) { return ewald_in_hbm(K) ? 0 : at_dyn_lds(K); }
// ... and with a cell list of nc^3 cells
static size_t at_dyn_lds_cells(int K, int nc)
{
    return at_dyn_lds_ewald(K) + 4 * (size_t)nc * nc * nc;
}

// Cell list geometry and storage of a chain (cubic box, every pair term cut
// at rc).  Cells of side >= rc + the largest single-trial displacement per
// axis, at least 3 per axis; otherwise the chain keeps the all-pairs screen.
static int setup_cells(fgx_system *s, ChainDev &C, const MoveDesc &mv)
{
    double dpmax = 0;
    std::vector<char> present(s->pot.ntypes, 0);
    for (int t : s->h_type) present[t] = 1;
    for (int t = 0; t < s->pot.ntypes; ++t)
        if (present[t]) dpmax = std::max(dpmax, std::fabs(s->type_dp[t]));
    const double dmax = std::max({std::fabs(mv.dir[0]), std::fabs(mv.dir[1]), std::fabs(mv.dir[2])});
    const double reach = (std::sqrt(s->pot.ew_rc2) + 0.5 * dpmax * dmax) * (1.0 + 1e-6);
    const int nc = (int)std::floor(s->box.len[0] / reach);
    const int K = s->has_ewald ? s->ew.K : 0;
    if (nc < 3 || at_dyn_lds_cells(K, nc) > AT_DYN_LDS_MAX) return FGX_OK;
    const int ncell = nc * nc * nc;
    int cap = s->cell_mode >= 2 ? s->cell_mode : 0;
    if (!cap) {
        // twice the mean occupancy plus a margin, grown on overflow
        const double mean = (double)s->n / ncell;
        cap = std::max(32, ((int)std::ceil(2.0 * mean + 32.0) + 31) / 32 * 32);
        cap = std::max(cap, s->cell_cap_min);
    }
    if (s->cell_nc != nc || s->cell_cap != cap || !s->d_cell_cnt) {
        if (s->d_cell_cnt) (void)hipFree(s->d_cell_cnt);
        if (s->d_cell_rec) (void)hipFree(s->d_cell_rec);
        if (s->d_cell_idx) (void)hipFree(s->d_cell_idx);
        if (s->d_cell_slot) (void)hipFree(s->d_cell_slot);
        s->d_cell_cnt = nullptr; s->d_cell_rec = nullptr; s->d_cell_idx = nullptr;
        s->d_cell_slot = nullptr;
        s->device_bytes -= s->cell_bytes;
        s->cell_bytes = 0;
        s->cell_nc = s->cell_cap = 0;
        FGX_HIP(hipMalloc(&s->d_cell_cnt, sizeof(int) * (ncell + 1)));
        FGX_HIP(hipMalloc(&s->d_cell_rec, sizeof(double4) * (size_t)ncell * cap));
        FGX_HIP(hipMalloc(&s->d_cell_idx, sizeof(int) * (size_t)ncell * cap));
        FGX_HIP(hipMalloc(&s->d_cell_slot, sizeof(int) * (size_t)s->n));
        s->cell_bytes = sizeof(int) * (ncell + 1) + (sizeof(double4) + sizeof(int)) * (size_t)ncell * cap +
                        sizeof(int) * (size_t)s->n;
        s->device_bytes += s->cell_bytes;
        s->cell_nc = nc;
        s->cell_cap = cap;
        s->cell_valid = false;
    }
    C.cell_nc = nc;
    C.cell_cap = cap;
    C.cell_cnt = s->d_cell_cnt;
    C.cell_rec = (double4 *)s->d_cell_rec;
    C.cell_idx = s->d_cell_idx;
    C.cell_slot = s->d_cell_slot;
    // partner types are needed unless every present type pair shares the
    // LJ and Ewald effective-charge tables
    {
        const int nt = s->pot.ntypes, nt2 = s->pot.ntypes2;
        bool uni = true, first = true;
        double ref[TAB_COUNT] = {0};
        for (int a = 0; a < nt && uni; ++a)
            for (int b2 = 0; b2 < nt && uni; ++b2) {
                if (!present[a] || !present[b2]) continue;
                for (int t = 0; t < TAB_COUNT; ++t) {
                    const double v = s->tab[t * nt2 + a * nt + b2];
                    if (first) ref[t] = v;
                    else if (v != ref[t]) uni = false;
                }
                first = false;
            }
        C.cell_type = uni ? 0 : 1;
    }
    return FGX_OK;
}

// What the device sweeps support for a system's registered move, checked
// without touching the device: the move's groups and, per move, the limits of
// the kernels' LDS stages.  `gvec` gets the move's groups.
static int chain_limits(const fgx_system *s, std::vector<int> &gvec)
{
    FGX_CHECK(s->moves.size() == 1, FGX_EUNSUPPORTED,
              "device Propagator: exactly one registered move is supported");
    const MoveDesc &mv = s->moves[0];
    gvec.clear();
    for (int g = 0; g < s->ng; ++g)
        if (s->groups[g].molid == mv.molid) gvec.push_back(g);
    FGX_CHECK(!gvec.empty(), FGX_ESTATE, "device Propagator: no group carries the move's molid");
    if (mv.type == MOVE_AT) {
        for (int g : gvec)
            FGX_CHECK(s->groups[g].molsize == 1, FGX_EUNSUPPORTED,
                      "device AtomicTranslation: moved groups must be atomic");
        FGX_CHECK(!s->has_ewald || s->ew.kcc < MAXK_AXIS, FGX_EUNSUPPORTED,
                  "device Ewald: cutoffK too large");
    } else {
        FGX_CHECK(!s->has_ewald, FGX_EUNSUPPORTED, "device TranslateRotate with Ewald");
        FGX_CHECK(s->ng <= MAX_GROUPS_LDS, FGX_EUNSUPPORTED, "device TranslateRotate: too many groups");
        for (int g : gvec)
            FGX_CHECK(s->groups[g].back - s->groups[g].front + 1 <= MAX_GROUP_SITES,
                      FGX_EUNSUPPORTED, "device TranslateRotate: group larger than LDS stage");
    }
    return FGX_OK;
}

int check_chain(const fgx_system *s)
{
    std::vector<int> gvec;
    return chain_limits(s, gvec);
}

// CutShift LJ: the fixed-unit Ewald form takes it only at the Ewald rc,
// where every pair term vanishes beyond rc (the loop then only screens)
static bool pot_lj_at_rc(const fgx_system *s)
{
    return s->pot.lj_rc2 > 0.0 && s->pot.lj_rc2 == s->pot.ew_rc2;
}

static bool pot_ew_lj(const fgx_system *s)
{
    return s->pot.elec == ELEC_EWALD && s->pot.sr == SR_LJ && s->pot.ntypes2 <= LDS_TYPES2 &&
           (s->pot.lj_rc2 == 0.0 || pot_lj_at_rc(s));
}

// fixed-unit pair form (one fixed-point unit L / 2^32 on all axes): cubic
// boxes with Debye-Hueckel / Coulomb / plain hard spheres, or Ewald + LJ
static bool fixed_unit_form(const fgx_system *s)
{
    const bool cubic = s->box.len[0] == s->box.len[1] && s->box.len[0] == s->box.len[2];
    const bool dh_like = (s->pot.sr == SR_NONE || s->pot.sr == SR_HS) &&
                         (s->pot.elec == ELEC_NONE || s->pot.elec == ELEC_COULOMB ||
                          s->pot.elec == ELEC_DH);
    return cubic && (dh_like || pot_ew_lj(s));
}

static int prepare_chain(fgx_system *s, ChainDev &C, hipStream_t st, double *du_rec, int *acc_rec,
                         long cap)
{
    std::vector<int> gvec;
    if (int rc = chain_limits(s, gvec)) return rc;
    const MoveDesc &mv = s->moves[0];
    int repeat = 1;
    if (mv.permol) repeat *= (int)gvec.size();
    if (mv.peratom) repeat *= s->groups[gvec[0]].back - s->groups[gvec[0]].front + 1;
    if (!s->d_gvec || s->gvec_n != (int)gvec.size()) {
        if (s->d_gvec) (void)hipFree(s->d_gvec);
        s->d_gvec = nullptr;
        FGX_HIP(hipMalloc(&s->d_gvec, sizeof(int) * gvec.size()));
        FGX_HIP(hipMemcpyAsync(s->d_gvec, gvec.data(), sizeof(int) * gvec.size(),
                               hipMemcpyHostToDevice, st));
        s->gvec_n = (int)gvec.size();
    }
    C.box = s->box;
    C.pot = s->pot;
    C.n = s->n;
    C.ng = s->ng;
    C.x = s->v[0].x; C.y = s->v[0].y; C.z = s->v[0].z;
    C.q = s->d_q; C.r = s->d_r; C.t = s->d_type; C.gid = s->d_gid;
    C.pk = s->d_pk;
    C.pk_w = s->d_pk;
    C.qf = s->d_qf; C.rf = s->d_rf;
    C.tab = s->d_tab; C.tabf = s->d_tabf;
    C.groups = s->d_groups;
    C.cm = s->d_cm;
    C.type_dp = s->d_type_dp;
    C.mv = mv;
    C.nmoves = (int)s->moves.size();
    C.repeat = repeat;
    C.gvec = s->d_gvec;
    C.ngvec = (int)gvec.size();
    C.rng = s->d_rng;
    if (s->has_ewald) {
        const double pi = std::acos(-1.);
        C.K = s->ew.K;
        C.kcc = s->ew.kcc;
        C.kn = s->ew.kn;
        C.aks = s->ew.aks;
        C.Q = s->ew.Q;
        C.eff = s->ew.eff;
        C.V = s->ew.V;
        C.recip_pref = (2 * pi / s->ew.V) * s->pot.lB;
        C.surf_pref = s->ew.const_inf >= 0.5
                          ? s->ew.const_inf * (2 * pi / ((2 * s->ew.eps_surf + 1) * s->ew.V)) * s->pot.lB
                          : 0.0;
        C.update_frequency = s->ew.update_frequency;
        C.ew_cheb = s->ew.cheb;
        C.ew_ncheb = s->ew.ncheb;
        C.ew_cheb_t = s->ew.ncheb ? 2.0 / s->ew.cheb_xmax : 0.0;
        C.lj_rc2inv = s->pot.lj_rc2 > 0.0 ? 1.0 / s->pot.lj_rc2 : 0.0;
    } else {
        C.lj_rc2inv = 0.0;
        C.ew_cheb = nullptr;
        C.ew_ncheb = 0;
        C.ew_cheb_t = 0.0;
        C.K = 0;
        C.kcc = 0;
    }
    C.stats_i = s->d_stats_i;
    C.stats_d = s->d_stats_d;
    C.du_rec = du_rec;
    C.acc_rec = acc_rec;
    C.cap = cap;
    C.emat = nullptr;
    if (mv.type == MOVE_TR && s->pot.precision == FGX_MIXED && s->emat_on && s->ng <= MAT_MAXG &&
        s->n < (1 << 22)) {
        if (!s->d_emat) {
            FGX_HIP(hipMalloc(&s->d_emat, sizeof(double) * s->ng * s->ng));
            s->device_bytes += sizeof(double) * s->ng * s->ng;
            s->emat_valid = false;
        }
        C.emat = s->d_emat;
    }
    // fixed-unit pair form: one fixed-point unit s = L / 2^32 on all axes
    C.fx = 0;
    C.fx_lbs = C.fx_rs = C.fx_c1 = C.fx_rc2 = C.fx_s2scale = 0.f;
    C.lj_uniform = 0;
    C.lj_check = 1;
    C.lj_s2f = C.lj_eps4f = 0.f;
    const bool lj_at_rc = pot_lj_at_rc(s);
    const bool ew_lj = pot_ew_lj(s);
    C.lj_screen = 0;
    C.cell_nc = C.cell_cap = 0;
    C.cell_cnt = nullptr;
    C.cell_rec = nullptr;
    C.cell_idx = nullptr;
    C.cell_slot = nullptr;
    C.sk_dS = nullptr;
    C.sk_kid = nullptr;
    C.ldsp = 0;
    C.sp_j0 = 0;
    C.sp_j1 = s->n;
    C.cell_type = 1;
    if (fixed_unit_form(s)) {
        const double unit = s->box.len[0] / 4294967296.0;
        C.fx = 1;
        C.fx_lbs = (float)(s->pot.lB / unit);
        C.fx_rs = (float)(std::sqrt(1.000001) / unit);
        C.fx_c1 = (float)(-s->pot.kappa * unit / std::log(2.0));
        if (ew_lj) {
            C.fx_rc2 = (float)(s->pot.ew_rc2 / (unit * unit) * 1.00001);   // margin: see pair loop
            C.fx_s2scale = (float)(1.0 / (unit * unit));
            // LJ parameters over the type pairs present; and whether any LJ
            // pair beyond rc can be strong (|x^6 - x^3| <= y + y^2 with
            // y = (sigma^2 / rc^2)^3), which would need the |u| test
            const int nt = s->pot.ntypes, nt2 = s->pot.ntypes2;
            std::vector<char> present(nt, 0);
            for (int t : s->h_type) present[t] = 1;
            bool uni = true, first = true;
            double s2u = 0, e4u = 0, bound = 0;
            for (int a = 0; a < nt; ++a)
                for (int b = 0; b < nt; ++b) {
                    if (!present[a] || !present[b]) continue;
                    const double s2 = s->tab[TAB_S2 * nt2 + a * nt + b];
                    const double e4 = s->tab[TAB_EPS4 * nt2 + a * nt + b];
                    if (first) { s2u = s2; e4u = e4; first = false; }
                    else if (s2 != s2u || e4 != e4u) uni = false;
                    const double y = std::pow(s2 / s->pot.ew_rc2, 3.0);
                    bound = std::max(bound, std::fabs(e4) * (y + y * y));
                }
            C.lj_uniform = uni ? 1 : 0;
            C.lj_s2f = (float)(s2u / (unit * unit));
            C.lj_eps4f = (float)e4u;
            C.lj_check = bound < 0.5 * s->pot.exact_f ? 0 : 1;
            C.lj_screen = lj_at_rc ? 1 : 0;
        }
    }
    if (C.lj_screen && mv.type == MOVE_AT && s->pot.precision == FGX_MIXED && s->cell_mode != 0) {
        int rc = setup_cells(s, C, mv);
        if (rc) return rc;
    }
    return FGX_OK;
}

// Movebase::trialMove at cnt == 0 (move.h:466-475): all mass centres are
// recomputed from the accepted vector before the first trial of a move.
static int init_mass_centres(fgx_system *s)
{
    std::vector<double> xyz(3 * (size_t)s->n);
    int rc = fgx_get_positions(s, 0, xyz.data());
    if (rc) return rc;
    for (int g = 0; g < s->ng; ++g) {
        double cm[3];
        const GroupDesc &G = s->groups[g];
        cm[0] = cm[1] = cm[2] = 0;
        if (G.back >= G.front) {
            int io = (int)(G.front + (G.back - G.front) * 0.5);
            double o[3] = {xyz[3 * io], xyz[3 * io + 1], xyz[3 * io + 2]};
            double sum = 0;
            for (int i = G.front; i <= G.back; ++i) {
                double t[3] = {xyz[3 * i] - o[0], xyz[3 * i + 1] - o[1], xyz[3 * i + 2] - o[2]};
                for (int k = 0; k < 3; ++k)
                    if (std::fabs(t[k]) > s->box.half[k]) {
                        double a = t[k] * s->box.inv[k];
                        t[k] -= s->box.len[k] * (int)(a > 0. ? a + .5 : a - .5);
                    }
                for (int k = 0; k < 3; ++k) cm[k] += t[k] * s->h_mw[i];
                sum += s->h_mw[i];
            }
            if (std::fabs(sum) < 1e-6) sum = 1;
            for (int k = 0; k < 3; ++k) {
                cm[k] = cm[k] / sum + o[k];
                if (std::fabs(cm[k]) > s->box.half[k]) {
                    double a = cm[k] * s->box.inv[k];
                    cm[k] -= s->box.len[k] * (int)(a > 0. ? a + .5 : a - .5);
                }
            }
        }
        for (int k = 0; k < 3; ++k) { s->cm[3 * g + k] = cm[k]; s->cm_trial[3 * g + k] = cm[k]; }
    }
    std::vector<double> cm6(6 * (size_t)s->ng);
    for (int g = 0; g < s->ng; ++g)
        for (int k = 0; k < 3; ++k) {
            cm6[6 * g + k] = s->cm[3 * g + k];
            cm6[6 * g + 3 + k] = s->cm_trial[3 * g + k];
        }
    FGX_HIP(hipMemcpyAsync(s->d_cm, cm6.data(), sizeof(double) * cm6.size(), hipMemcpyHostToDevice,
                           s->stream));
    FGX_HIP(hipStreamSynchronize(s->stream));
    return FGX_OK;
}

// the device RNG states and Ewald accept counter become the host's again
int pull_device_state(fgx_system *s)
{
    if (!s->state_on_dev) return FGX_OK;
    FGX_HIP(hipMemcpyAsync(s->rng_global, s->d_rng, sizeof(uint32_t) * 625, hipMemcpyDeviceToHost,
                           s->stream));
    FGX_HIP(hipMemcpyAsync(s->rng_internal, s->d_rng + 625, sizeof(uint32_t) * 625,
                           hipMemcpyDeviceToHost, s->stream));
    if (s->has_ewald && s->ew.cnt_dev)
        FGX_HIP(hipMemcpyAsync(&s->ew.cnt_accepted, s->ew.cnt_dev, sizeof(int),
                               hipMemcpyDeviceToHost, s->stream));
    FGX_HIP(hipStreamSynchronize(s->stream));
    s->state_on_dev = false;
    return FGX_OK;
}

// One batched launch in flight (launch_chains .. finish_chains).
struct Launch {
    std::vector<fgx_system *> sys;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = FGX_OK;
    int *d_xerr = nullptr;   // block sweep: the launch's error word
};

static int blk_instance(bool mix, const Pot &P, size_t *static_lds, int *block)
{
    return blk_instance_pair(mix, P, static_lds, block) ||
           blk_instance_ewald(mix, P, static_lds, block);
}

// Automatic workgroups per chain for block sweeps: with at least half as many
// chains as compute units every chain keeps one workgroup; otherwise the
// units go, one at a time, to the chain whose sweep takes longest, while that
// shortens it, a part keeping at least 512 partner rows.  A sweep of n trials
// in P workgroups takes about n (F + (n + 2K) / P) partner-row times: the
// partner pass (and the k-vectors, about two rows each) shrinks with P, the
// per-trial share of the block's draws, reduction, exchange and decisions, F,
// does not.  On MI355X, fitted to Debye-Hueckel chains of 1,252-14,980
// particles, F is about 2,700 rows on the mixed path and 400 on the FP64
// path (whose rows cost four times as much).  The parts of a chain share one
// XCD, so the result is fitted to the XCDs' compute units at the end.
static constexpr double BLK_FIXED_ROWS_MIXED = 2700.0, BLK_FIXED_ROWS_FP64 = 400.0;

static std::vector<int> blk_auto_parts(const std::vector<fgx_system *> &sys, int ncu)
{
    const int nc = (int)sys.size();
    std::vector<int> parts(nc, 1);
    if (2 * nc >= ncu) return parts;
    auto sweep_time = [&](int c, int P) {
        const fgx_system *s = sys[c];
        const double rows = (double)s->n + 2.0 * (s->has_ewald ? s->ew.K : 0);
        const double F = s->pot.precision == FGX_MIXED ? BLK_FIXED_ROWS_MIXED : BLK_FIXED_ROWS_FP64;
        return (double)s->n * (F + rows / P);
    };
    int used = nc;
    while (used < ncu) {
        int worst = -1;
        double tw = -1.0;
        for (int c = 0; c < nc; ++c) {
            if (sys[c]->blk_nsplit > 0) continue;
            const double t = sweep_time(c, parts[c]);
            if (t > tw) { tw = t; worst = c; }
        }
        if (worst < 0) break;
        const int cap = std::min(BLK_NSPLIT_MAX, std::max(1, sys[worst]->n / 512));
        if (parts[worst] >= cap) break;   // the slowest chain cannot get faster
        ++parts[worst];
        ++used;
    }
    // fit the XCD packing of plan_blk (largest first onto the least-loaded
    // XCD): while an XCD holds more parts than it has compute units, the
    // largest split chain there gives one back
    const int nx = 8, cux = std::max(1, ncu / nx);
    for (int guard = 0; guard < 64 * nc; ++guard) {
        std::vector<int> order(nc), xcd(nc), load(nx, 0);
        for (int c = 0; c < nc; ++c) order[c] = c;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return (double)parts[a] * sys[a]->n > (double)parts[b] * sys[b]->n;
        });
        for (int c : order) {
            int x = 0;
            for (int y = 1; y < nx; ++y)
                if (load[y] < load[x]) x = y;
            xcd[c] = x;
            load[x] += sys[c]->blk_nsplit > 0 ? sys[c]->blk_nsplit : parts[c];
        }
        int over = -1;
        for (int x = 0; x < nx; ++x)
            if (load[x] > cux && (over < 0 || load[x] > load[over])) over = x;
        if (over < 0) break;
        int big = -1;
        for (int c = 0; c < nc; ++c)
            if (xcd[c] == over && sys[c]->blk_nsplit <= 0 && parts[c] > 1 &&
                (big < 0 || parts[c] > parts[big]))
                big = c;
        if (big < 0) break;   // explicit splits alone overfill it: plan_blk refuses
        --parts[big];
    }
    return parts;
}

// Workgroups of a job-server chain's block sweeps, fixed by the chain alone
// so that its result does not depend on the jobs beside it: one per 512
// partner rows, at most cap; 0 when the chain cannot run the block sweep (no
// instance, a cell list, no fixed-unit form on the mixed path, or LDS).
int blk_job_parts(const fgx_system *s, int cap)
{
    const bool mix = s->pot.precision == FGX_MIXED;
    size_t st_lds = 0;
    int bsz = 0;
    if (!s->blk_mode || s->moves.empty() || s->moves[0].type != MOVE_AT) return 0;
    if (!blk_instance(mix, s->pot, &st_lds, &bsz) || st_lds == 0) return 0;
    if (s->pot.ntypes2 > LDS_TYPES2 || s->n < 1 || s->n >= (1 << 24)) return 0;
    if (mix && !fixed_unit_form(s)) return 0;
    if (mix && pot_lj_at_rc(s) && pot_ew_lj(s) && s->cell_mode != 0) return 0;
    const int P = std::max(1, std::min({cap, BLK_NSPLIT_MAX, s->n / 512}));
    const int K = s->has_ewald ? s->ew.K : 0;
    const BlkLds base = blk_lds((K + P - 1) / P, bsz, s->has_ewald ? s->ew.kcc + 1 : 1,
                                s->has_ewald ? 1 : 0, (s->n + P - 1) / P, 0);
    return st_lds + base.total <= 160 * 1024 ? P : 0;
}

// Block-sweep plan of a batch: every chain must have an instance and fit,
// else the batch keeps the one-trial sweep (an explicit split that cannot be
// honoured is an error).  Sets the chains' parts, LDS layout and exchange
// buffers, and the workgroup map: a chain's parts on one XCD (workgroups are
// dealt to the 8 XCDs round-robin by index), the XCDs balanced by workgroups.
static int plan_blk(std::vector<fgx_system *> &sys, std::vector<ChainDev> &chains_io, bool skg,
                    BatchPlan &B)
{
    const int nc = (int)sys.size();
    // changes go to a copy, kept only when the batch takes the block sweep
    std::vector<ChainDev> chains = chains_io;
    fgx_system *s0 = sys[0];
    B.blk = false;
    const bool mix = s0->pot.precision == FGX_MIXED;
    size_t st_lds = 0;
    int bsz = 0;
    (void)skg;   // the block sweep lays out its own (per-part) Ewald state
    if (!s0->blk_mode || s0->moves[0].type != MOVE_AT) return FGX_OK;
    if (!blk_instance(mix, s0->pot, &st_lds, &bsz) || st_lds == 0) return FGX_OK;
    for (int c = 0; c < nc; ++c) {
        const fgx_system *s = sys[c];
        if (!s->blk_mode || chains[c].cell_nc > 0 || (mix && !chains[c].fx) ||
            s->pot.ntypes2 > LDS_TYPES2 || s->n < 1 || s->n >= (1 << 24))
            return FGX_OK;
    }
    int dev = 0, ncu = 0;
    FGX_HIP(hipGetDevice(&dev));
    FGX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const size_t lds_cap = 160 * 1024;
    std::vector<int> parts = blk_auto_parts(sys, ncu);
    size_t dyn = 0;
    bool split = false;
    for (int c = 0; c < nc; ++c) {
        fgx_system *s = sys[c];
        int P = s->blk_nsplit > 0 ? s->blk_nsplit : parts[c];
        FGX_CHECK(P <= BLK_NSPLIT_MAX && P <= s->n, FGX_EARG,
                  "block sweep: more workgroups per chain than partners or than 32");
        const int K = s->has_ewald ? s->ew.K : 0;
        const int ks = (K + P - 1) / P;
        const int kcc1 = s->has_ewald ? s->ew.kcc + 1 : 1;
        const int ns = (s->n + P - 1) / P;
        const BlkLds base = blk_lds(ks, bsz, kcc1, s->has_ewald ? 1 : 0, ns, 0);
        const BlkLds full = blk_lds(ks, bsz, kcc1, s->has_ewald ? 1 : 0, ns, ns);
        if (st_lds + base.total > lds_cap) {
            FGX_CHECK(s->blk_nsplit == 0, FGX_EUNSUPPORTED,
                      "block sweep: the Ewald k-vectors of one workgroup exceed its LDS");
            return FGX_OK;
        }
        const bool ldsp = s->lds_partners && ns <= BLK_LDS_ROWS && st_lds + full.total <= lds_cap;
        chains[c].ldsp = ldsp ? 1 : 0;
        chains[c].sp_j0 = 0;
        chains[c].sp_j1 = s->n;
        dyn = std::max(dyn, ldsp ? full.total : base.total);
        parts[c] = P;
        split = split || P > 1;
        if (!s->d_xbuf) {
            FGX_HIP(hipMalloc(&s->d_xbuf, sizeof(double) * 2 * BLK_NSPLIT_MAX * BLK_NV_MAX));
            FGX_HIP(hipMalloc(&s->d_xcnt, sizeof(int) * 2));
            FGX_HIP(hipMemset(s->d_xcnt, 0, sizeof(int) * 2));
            s->device_bytes += sizeof(double) * 2 * BLK_NSPLIT_MAX * BLK_NV_MAX + sizeof(int) * 2;
        }
        chains[c].nsplit = P;
        chains[c].xbuf = s->d_xbuf;
        chains[c].xcnt = s->d_xcnt;
    }
    if (!s0->d_xerr) {
        FGX_HIP(hipMalloc(&s0->d_xerr, sizeof(int)));
        FGX_HIP(hipMemset(s0->d_xerr, 0, sizeof(int)));
    }
    for (int c = 0; c < nc; ++c) chains[c].xerr = s0->d_xerr;
    const int nx = 8;
    std::vector<std::vector<int2>> per(nx);
    std::vector<int> order(nc);
    for (int c = 0; c < nc; ++c) order[c] = c;
    // chains with a given XCD first, so that their parts take the lowest rows
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const bool fa = sys[a]->blk_xcd >= 0, fb = sys[b]->blk_xcd >= 0;
        if (fa != fb) return fa;
        return (double)parts[a] * sys[a]->n > (double)parts[b] * sys[b]->n;
    });
    for (int c : order) {
        int x = 0;
        if (sys[c]->blk_xcd >= 0)
            x = sys[c]->blk_xcd % nx;
        else
            for (int y = 1; y < nx; ++y)
                if (per[y].size() < per[x].size()) x = y;
        for (int p = 0; p < parts[c]; ++p) per[x].push_back(make_int2(c, p));
    }
    size_t rows = 0;
    for (int x = 0; x < nx; ++x) rows = std::max(rows, per[x].size());
    const int ngrid = (int)rows * nx;
    // parts of a chain wait for each other: all workgroups must be resident
    // together (one 512-thread workgroup per compute unit)
    FGX_CHECK(!split || ngrid <= ncu, FGX_EUNSUPPORTED,
              "block sweep: chains split over more workgroups than the device holds at once");
    // automatic mode: unsplit mixed Ewald chains whose partners stay in
    // device memory run faster one trial at a time (the partner loop there
    // is the whole cost, and its one-trial form is the better tuned)
    bool keep = false;
    for (int c = 0; c < nc; ++c)
        keep = keep || sys[c]->blk_mode == 1 || parts[c] > 1 || chains[c].ldsp;
    if (!keep && s0->pot.elec == ELEC_EWALD && mix) return FGX_OK;
    std::vector<int2> map((size_t)ngrid, make_int2(-1, -1));
    for (int x = 0; x < nx; ++x)
        for (size_t r = 0; r < per[x].size(); ++r) map[r * nx + x] = per[x][r];
    if (s0->blkmap_cap < ngrid) {
        if (s0->d_blkmap) (void)hipFree(s0->d_blkmap);
        s0->d_blkmap = nullptr;
        s0->blkmap_cap = 0;
        FGX_HIP(hipMalloc(&s0->d_blkmap, sizeof(int2) * ngrid));
        s0->blkmap_cap = ngrid;
    }
    FGX_HIP(hipMemcpyAsync(s0->d_blkmap, map.data(), sizeof(int2) * ngrid, hipMemcpyHostToDevice,
                           s0->stream));
    for (int c = 0; c < nc; ++c) sys[c]->blk_parts = parts[c];
    chains_io = chains;
    B.blk = true;
    B.d_map = s0->d_blkmap;
    B.ngrid = ngrid;
    B.nthr = BLK_THREADS;
    B.dyn = dyn;
    B.d_xerr = s0->d_xerr;
    return FGX_OK;
}

// the launch error word of a block sweep: FGX_EDEVICE when a part gave up
// waiting for the others; clears it and the chains' counters
int blk_check(const std::vector<fgx_system *> &sys, int *d_xerr)
{
    if (!d_xerr) return FGX_OK;
    int err = 0;
    FGX_HIP(hipMemcpy(&err, d_xerr, sizeof(int), hipMemcpyDeviceToHost));
    if (!err) return FGX_OK;
    FGX_HIP(hipMemset(d_xerr, 0, sizeof(int)));
    for (fgx_system *s : sys)
        if (s->d_xcnt) FGX_HIP(hipMemset(s->d_xcnt, 0, sizeof(int) * 2));
    set_error("block sweep: a workgroup of a split chain waited too long for the others");
    return FGX_EDEVICE;
}

// Everything of a batched launch up to the kernel: per-chain state on the
// device, cell lists and energy caches rebuilt where stale, the ChainDev
// array copied to the first system's batch buffer.
int prepare_batch(std::vector<fgx_system *> &sys, const std::vector<int> &ncalls,
                  const std::vector<long> *max_trials, double *du_rec, int *acc_rec, long cap,
                  BatchPlan &B)
{
    fgx_system *s0 = sys[0];
    const int nc = (int)sys.size();
    hipStream_t st = s0->stream;
    std::vector<ChainDev> chains(nc);
    int rc = FGX_OK;
    int nthr = 0;
    size_t dyn = 0;
    bool fx = true;   // fixed-unit form only when every chain has it
    bool cell = false;   // every pair term cut at rc: the screening kernel
    bool cells = false;  // some chain sweeps with a cell list
    const int unroll = s0->launch_unroll > 0 ? s0->launch_unroll : 4;
    // batch buffers, kept by the first system
    if (s0->batch_cap < nc) {
        if (s0->d_batch) (void)hipFree(s0->d_batch);
        if (s0->d_bstat_i) (void)hipFree(s0->d_bstat_i);
        if (s0->d_bstat_d) (void)hipFree(s0->d_bstat_d);
        s0->d_batch = nullptr; s0->d_bstat_i = nullptr; s0->d_bstat_d = nullptr;
        s0->batch_cap = 0;
        FGX_HIP(hipMalloc(&s0->d_batch, sizeof(ChainDev) * nc));
        FGX_HIP(hipMalloc(&s0->d_bstat_i, sizeof(long) * 8 * nc));
        FGX_HIP(hipMalloc(&s0->d_bstat_d, sizeof(double) * 8 * nc));
        s0->batch_cap = nc;
    }
    for (int c = 0; c < nc && !rc; ++c) {
        fgx_system *s = sys[c];
        FGX_CHECK(s->moves.size() == s0->moves.size() && !s->moves.empty() &&
                      s->moves[0].type == s0->moves[0].type &&
                      s->pot.precision == s0->pot.precision && s->pot.elec == s0->pot.elec &&
                      s->pot.sr == s0->pot.sr,
                  FGX_EUNSUPPORTED, "batched chains must share move type and precision");
        // Propagator::move evaluates uinit = systemEnergy before its first move
        // (move.h:7884-7885).  The energy itself only serves drift reporting
        // and is left to callers that ask for it (fgx_system_energy), but its
        // setSpace recomputes S(k) from the accepted vector, which the first
        // call does here too (S(k) may differ from that sum, e.g. after an
        // accepted volume move).
        if (!s->propagator_started && s->has_ewald) {
            if ((rc = ewald_set_space(s))) break;
            if (hipStreamSynchronize(s->stream) != hipSuccess) {
                set_error("hipStreamSynchronize failed");
                rc = FGX_EHIP;
                break;
            }
        }
        s->propagator_started = true;
        if (s->move_cnt.empty()) s->move_cnt.assign(s->moves.size(), 0);
        if (s->move_cnt[0] == 0 && (rc = init_mass_centres(s))) break;
        if (s->has_ewald && !s->ew.cnt_dev) FGX_HIP(hipMalloc(&s->ew.cnt_dev, sizeof(int)));
        // work queued on the system's own stream precedes the sweep
        if (s != s0) FGX_HIP(hipStreamSynchronize(s->stream));
        if (!s->state_on_dev) {
            FGX_HIP(hipMemcpyAsync(s->d_rng, s->rng_global, sizeof(uint32_t) * 625,
                                   hipMemcpyHostToDevice, st));
            FGX_HIP(hipMemcpyAsync(s->d_rng + 625, s->rng_internal, sizeof(uint32_t) * 625,
                                   hipMemcpyHostToDevice, st));
            if (s->has_ewald)
                FGX_HIP(hipMemcpyAsync(s->ew.cnt_dev, &s->ew.cnt_accepted, sizeof(int),
                                       hipMemcpyHostToDevice, st));
            FGX_HIP(hipStreamSynchronize(st));   // host buffers are read at once
            s->state_on_dev = true;
        }
        // chain c records into entries [c cap, (c + 1) cap) of the buffers
        rc = prepare_chain(s, chains[c], st, du_rec ? du_rec + (size_t)c * cap : nullptr,
                           acc_rec ? acc_rec + (size_t)c * cap : nullptr, du_rec ? cap : 0);
        if (rc) break;
        if (s->has_ewald) chains[c].ew_cnt = s->ew.cnt_dev;
        chains[c].ncalls = ncalls[c];
        chains[c].max_trials = max_trials ? (*max_trials)[c] : 0;
        chains[c].resume = s->d_resume;
        chains[c].stats_i = s0->d_bstat_i + 8 * c;
        chains[c].stats_d = s0->d_bstat_d + 8 * c;
        nthr = std::max(nthr, s->launch_threads > 0 ? s->launch_threads : chain_block(s));
        // the partner loop's prefetch must stay inside the zero padding
        FGX_CHECK(3 * 8 * nthr <= PARTNER_PAD, FGX_EUNSUPPORTED,
                  "device sweep: block size exceeds the partner-array padding");
        if (s->has_ewald)
            dyn = std::max(dyn, chains[c].cell_nc > 0 ? at_dyn_lds_cells(s->ew.K, chains[c].cell_nc)
                                                      : at_dyn_lds_ewald(s->ew.K));
        if (!chains[c].fx) fx = false;
        if (chains[c].cell_nc > 0) cells = true;
        FGX_CHECK(chains[c].lj_screen == chains[0].lj_screen, FGX_EUNSUPPORTED,
                  "batched chains must agree on a Lennard-Jones cut at the Ewald rc");
        // mixed AtomicTranslation without Ewald: a chain whose partners fit
        // keeps them in LDS for the launch
        const size_t part = (size_t)s->n * LDS_PART_BYTES;
        if (s->lds_partners && !s->has_ewald && s->pot.precision == FGX_MIXED &&
            chains[c].mv.type == MOVE_AT && s->pot.ntypes2 <= LDS_TYPES2 && part <= AT_LDS_PART_MAX) {
            chains[c].ldsp = 1;
            dyn = std::max(dyn, part);
        }
    }
    // pair terms cut at rc: the screening kernel (cell lists where present)
    if (!rc) cell = chains[0].lj_screen != 0;
    // more k-vectors than the LDS-resident Ewald state holds: the batch's
    // Ewald chains keep S(k), dS(k), A_k and n in device memory
    bool skg = false;
    for (int c = 0; c < nc && !rc; ++c)
        skg = skg || (sys[c]->has_ewald && (ewald_in_hbm(sys[c]->ew.K) || sys[c]->ewald_state_hbm));
    if (!rc && skg) {
        // dynamic LDS then holds only the cell counts
        dyn = 0;
        for (int c = 0; c < nc; ++c)
            if (chains[c].cell_nc > 0)
                dyn = std::max(dyn, 4 * (size_t)chains[c].cell_nc * chains[c].cell_nc * chains[c].cell_nc);
        for (int c = 0; c < nc; ++c) {
            fgx_system *s = sys[c];
            if (!s->has_ewald) continue;
            if (!s->ew.sk_dS) {
                FGX_HIP(hipMalloc(&s->ew.sk_dS, sizeof(double) * 2 * s->ew.K));
                FGX_HIP(hipMalloc(&s->ew.sk_kid, sizeof(int) * s->ew.K));
                std::vector<int> kid(s->ew.K);
                for (int k = 0; k < s->ew.K; ++k)
                    kid[k] = kpack(s->h_kn[3 * k], s->h_kn[3 * k + 1], s->h_kn[3 * k + 2]);
                FGX_HIP(hipMemcpy(s->ew.sk_kid, kid.data(), sizeof(int) * s->ew.K,
                                  hipMemcpyHostToDevice));
                s->device_bytes += (16 + 4) * (size_t)s->ew.K;
            }
            chains[c].sk_dS = (double2 *)s->ew.sk_dS;
            chains[c].sk_kid = s->ew.sk_kid;
        }
    }
    // the cell scan spreads 27 cells over waves 1..7 of a 512-thread block
    if (cells) nthr = AT_THREADS;
    for (int c = 0; c < nc; ++c) {
        chains[c].nsplit = 1;
        chains[c].xbuf = nullptr;
        chains[c].xcnt = nullptr;
        chains[c].xerr = nullptr;
        if (!rc) sys[c]->blk_parts = 0;
    }
    if (!rc) {
        // cell lists that no longer match the accepted state
        std::vector<ChainDev> stale;
        int maxcell = 0;
        for (int c = 0; c < nc; ++c)
            if (chains[c].cell_nc > 0 && !sys[c]->cell_valid) {
                stale.push_back(chains[c]);
                maxcell = std::max(maxcell, chains[c].cell_nc * chains[c].cell_nc * chains[c].cell_nc);
            }
        if (!stale.empty()) {
            ChainDev *d_stale = nullptr;
            FGX_HIP(hipMalloc(&d_stale, sizeof(ChainDev) * stale.size()));
            FGX_HIP(hipMemcpyAsync(d_stale, stale.data(), sizeof(ChainDev) * stale.size(),
                                   hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_cell_build, dim3((unsigned)stale.size()), dim3(256),
                               sizeof(int) * (size_t)maxcell, st, d_stale);
            hipError_t err = hipGetLastError();
            if (err == hipSuccess) err = hipStreamSynchronize(st);
            (void)hipFree(d_stale);
            if (err != hipSuccess) {
                set_error(std::string("cell list kernel failed: ") + hipGetErrorString(err));
                rc = FGX_EHIP;
            }
            if (!rc)
                for (int c = 0; c < nc; ++c)
                    if (chains[c].cell_nc > 0) sys[c]->cell_valid = true;
        }
    }
    if (!rc) {
        // group-pair energy caches that no longer match the accepted state
        std::vector<ChainDev> stale;
        int maxng = 0;
        for (int c = 0; c < nc; ++c)
            if (chains[c].emat && !sys[c]->emat_valid) {
                stale.push_back(chains[c]);
                maxng = std::max(maxng, chains[c].ng);
            }
        if (!stale.empty()) {
            ChainDev *d_stale = nullptr;
            FGX_HIP(hipMalloc(&d_stale, sizeof(ChainDev) * stale.size()));
            FGX_HIP(hipMemcpyAsync(d_stale, stale.data(), sizeof(ChainDev) * stale.size(),
                                   hipMemcpyHostToDevice, st));
            rc = launch_emat_init(s0->pot, st, d_stale, (int)stale.size(), maxng);
            hipError_t err = hipStreamSynchronize(st);
            (void)hipFree(d_stale);
            if (!rc && err != hipSuccess) {
                set_error(std::string("energy cache kernel failed: ") + hipGetErrorString(err));
                rc = FGX_EHIP;
            }
            if (!rc)
                for (int c = 0; c < nc; ++c)
                    if (chains[c].emat) sys[c]->emat_valid = true;
        }
    }
    BatchPlan BB;
    if (!rc) rc = plan_blk(sys, chains, skg, BB);
    if (!rc) {
        ChainDev *d_chains = (ChainDev *)s0->d_batch;
        FGX_HIP(hipMemcpyAsync(d_chains, chains.data(), sizeof(ChainDev) * nc, hipMemcpyHostToDevice,
                               st));
        B.d_chains = d_chains;
        B.nchains = nc;
        B.nthr = nthr;
        B.dyn = dyn;
        B.ncalls = ncalls.empty() ? 0 : ncalls[0];
        B.unroll = unroll;
        B.fx = fx;
        B.cell = cell;
        B.skg = skg;
        B.pot = s0->pot;
        B.move_type = s0->moves[0].type;
        B.blk = BB.blk;
        if (BB.blk) {
            B.d_map = BB.d_map;
            B.ngrid = BB.ngrid;
            B.nthr = BB.nthr;
            B.dyn = BB.dyn;
            B.d_xerr = BB.d_xerr;
        }
    }
    return rc;
}

// The sweep kernel of a prepared batch on stream st (nothing else: graph
// capture records exactly this launch).
int enqueue_batch(const BatchPlan &B, hipStream_t st)
{
    if (B.blk) {
        const bool mix = B.pot.precision == FGX_MIXED;
        int rc = B.pot.elec == ELEC_EWALD
                     ? launch_blk_ewald(mix, B.pot, st, B.d_chains, B.d_map, B.ngrid, B.nthr, B.dyn)
                     : launch_blk_pair(mix, B.pot, st, B.d_chains, B.d_map, B.ngrid, B.nthr, B.dyn);
        if (rc) {
            set_error("block sweep: no kernel for this Hamiltonian");
            return rc;
        }
        FGX_HIP(hipGetLastError());
        return FGX_OK;
    }
    if (B.move_type == MOVE_AT)
        return B.pot.precision == FGX_MIXED
                   ? launch_at<true>(B.pot, st, B.d_chains, B.nchains, B.nthr, B.dyn, B.ncalls,
                                     B.unroll, B.fx, B.cell, B.skg)
                   : launch_at<false>(B.pot, st, B.d_chains, B.nchains, B.nthr, B.dyn, B.ncalls,
                                      B.unroll, B.fx, false, B.skg);
    return launch_tr(B.pot.precision == FGX_MIXED, B.pot, st, B.d_chains, B.nchains, B.nthr,
                     B.ncalls);
}

// Trial vector and trial S(k) mirror the accepted state after device sweeps
// (undo/accept leave p == trial between moves): one copy launch for the
// whole batch, each system a row of the grid.
struct MirrorJob {
    const double *sx, *sy, *sz, *sq;
    double *dx, *dy, *dz, *dq;
    int n, nq;
};

__global__ __launch_bounds__(256) void k_mirror(const MirrorJob *__restrict__ jobs)
{
    const MirrorJob J = jobs[blockIdx.y];
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < J.n; i += stride) {
        J.dx[i] = J.sx[i];
        J.dy[i] = J.sy[i];
        J.dz[i] = J.sz[i];
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < J.nq; i += stride) J.dq[i] = J.sq[i];
}

int mirror_trial(std::vector<fgx_system *> &sys, hipStream_t st)
{
    const int nc = (int)sys.size();
    if (nc == 0) return FGX_OK;
    fgx_system *s0 = sys[0];
    std::vector<MirrorJob> jobs(nc);
    int maxn = 1;
    for (int c = 0; c < nc; ++c) {
        fgx_system *s = sys[c];
        MirrorJob &J = jobs[c];
        J.sx = s->v[0].x; J.sy = s->v[0].y; J.sz = s->v[0].z;
        J.dx = s->v[1].x; J.dy = s->v[1].y; J.dz = s->v[1].z;
        J.n = s->n;
        J.sq = s->has_ewald ? s->ew.Q : nullptr;
        J.dq = s->has_ewald ? s->ew.Qt : nullptr;
        J.nq = s->has_ewald ? 2 * s->ew.K : 0;
        maxn = std::max(maxn, std::max(J.n, J.nq));
    }
    if (s0->mirror_cap < nc) {
        if (s0->d_mirror) (void)hipFree(s0->d_mirror);
        s0->d_mirror = nullptr;
        s0->mirror_cap = 0;
        FGX_HIP(hipMalloc(&s0->d_mirror, sizeof(MirrorJob) * nc));
        s0->mirror_cap = nc;
    }
    MirrorJob *d_jobs = (MirrorJob *)s0->d_mirror;
    FGX_HIP(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(MirrorJob) * nc, hipMemcpyHostToDevice, st));
    const int gx = std::min(64, (maxn + 255) / 256);
    for (int c0 = 0; c0 < nc; c0 += 65535) {
        const int rows = std::min(65535, nc - c0);
        hipLaunchKernelGGL(k_mirror, dim3(gx, rows), dim3(256), 0, st, d_jobs + c0);
    }
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

// Enqueues one launch advancing chain c by ncalls[c] Propagator::move()
// calls.  All transfers go on the first system's stream; RNG states and
// counters stay on the device between launches.  Nothing waits for the
// sweep itself, so launches of different Hamiltonians on different first
// systems overlap on the device.
static int launch_chains(std::vector<fgx_system *> &sys, const std::vector<int> &ncalls,
                         const std::vector<long> *max_trials, double *du_rec, int *acc_rec,
                         long cap, Launch &L)
{
    hipStream_t st = sys[0]->stream;
    L.sys = sys;
    L.st = st;
    BatchPlan B;
    int rc = prepare_batch(sys, ncalls, max_trials, du_rec, acc_rec, cap, B);
    if (!rc) {
        FGX_HIP(hipEventCreate(&L.e0));
        FGX_HIP(hipEventCreate(&L.e1));
        FGX_HIP(hipEventRecord(L.e0, st));
        rc = enqueue_batch(B, st);
        L.d_xerr = B.blk ? B.d_xerr : nullptr;
        FGX_HIP(hipEventRecord(L.e1, st));
        if (!rc) rc = mirror_trial(sys, st);
    }
    L.rc = rc;
    return rc;
}

// Waits for a launch; its device time, and per-chain statistics (8 longs +
// 8 doubles per chain) when asked for.
static int finish_chains(Launch &L, double *seconds, std::vector<long> *stats_i,
                         std::vector<double> *stats_d)
{
    int rc = L.rc;
    if (L.sys.empty()) return rc;
    fgx_system *s0 = L.sys[0];
    const int nc = (int)L.sys.size();
    hipError_t err = hipStreamSynchronize(L.st);
    if (err != hipSuccess && !rc) {
        set_error(std::string("sweep kernel failed: ") + hipGetErrorString(err));
        rc = FGX_EHIP;
    }
    if (!rc) rc = blk_check(L.sys, L.d_xerr);
    float ms = 0;
    if (!rc && L.e0 && L.e1) FGX_HIP(hipEventElapsedTime(&ms, L.e0, L.e1));
    if (seconds) *seconds = ms * 1e-3;
    if (L.e0) (void)hipEventDestroy(L.e0);
    if (L.e1) (void)hipEventDestroy(L.e1);
    L.e0 = L.e1 = nullptr;
    if (!rc && stats_i) {
        stats_i->assign(8 * (size_t)nc, 0);
        stats_d->assign(8 * (size_t)nc, 0.0);
        FGX_HIP(hipMemcpy(stats_i->data(), s0->d_bstat_i, sizeof(long) * 8 * nc,
                          hipMemcpyDeviceToHost));
        FGX_HIP(hipMemcpy(stats_d->data(), s0->d_bstat_d, sizeof(double) * 8 * nc,
                          hipMemcpyDeviceToHost));
    }
    return rc;
}

// asynchronous batch launches for the job server (fgx_server.hip)
int batch_launch(std::vector<fgx_system *> &sys, const std::vector<int> &ncalls,
                 const std::vector<long> &max_trials, void **handle)
{
    Launch *L = new Launch;
    *handle = L;
    return launch_chains(sys, ncalls, &max_trials, nullptr, nullptr, 0, *L);
}

int batch_finish(void *handle, double *seconds, std::vector<long> *stats_i,
                 std::vector<double> *stats_d)
{
    Launch *L = (Launch *)handle;
    int rc = finish_chains(*L, seconds, stats_i, stats_d);
    delete L;
    return rc;
}

static int run_chains(std::vector<fgx_system *> &sys, int ncalls, double *du_rec, int *acc_rec,
                      long cap, double *seconds, std::vector<long> *stats_i,
                      std::vector<double> *stats_d)
{
    Launch L;
    std::vector<int> nv(sys.size(), ncalls);
    int rc = launch_chains(sys, nv, nullptr, du_rec, acc_rec, cap, L);
    int rc2 = finish_chains(L, seconds, rc ? nullptr : stats_i, stats_d);
    return rc ? rc : rc2;
}

}  // namespace fgx

using namespace fgx;

extern "C" int fgx_add_move(fgx_system *s, int type, int molid, double prob, int peratom,
                            int permol, const double dir[3], const double dir2[3], double dp,
                            double dprot, int center_rotation)
{
    FGX_CHECK(s && (type == MOVE_AT || type == MOVE_TR), FGX_EARG, "fgx_add_move: bad move type");
    MoveDesc m{};
    m.type = type;
    m.molid = molid;
    m.prob = prob;
    m.peratom = peratom;
    m.permol = permol;
    for (int k = 0; k < 3; ++k) {
        m.dir[k] = dir ? dir[k] : 1.0;
        m.dir2[k] = dir2 ? dir2[k] : 1.0;
    }
    m.dp = dp;
    m.dprot = dprot;
    if (m.dprot > 4 * std::acos(-1.)) m.dprot = 4 * std::acos(-1.);   // move.h:1598-1599
    m.center_rotation = center_rotation;
    s->moves.push_back(m);
    s->move_cnt.push_back(0);
    return FGX_OK;
}

extern "C" int fgx_rng_seed(fgx_system *s, uint32_t seed)
{
    FGX_CHECK(s, FGX_EARG, "fgx_rng_seed: null system");
    if (int rc = pull_device_state(s)) return rc;
    mt_seed_host(s->rng_global, seed);
    return FGX_OK;
}

extern "C" int fgx_rng_set_state(fgx_system *s, const uint32_t words[624], int pos)
{
    FGX_CHECK(s && words && pos >= 0 && pos <= 624, FGX_EARG, "fgx_rng_set_state: bad arguments");
    if (int rc = pull_device_state(s)) return rc;
    std::memcpy(s->rng_global, words, sizeof(uint32_t) * 624);
    s->rng_global[624] = (uint32_t)pos;
    return FGX_OK;
}

extern "C" int fgx_rng_get_state(fgx_system *s, uint32_t words[624], int *pos)
{
    FGX_CHECK(s && words && pos, FGX_EARG, "fgx_rng_get_state: bad arguments");
    if (int rc = pull_device_state(s)) return rc;
    std::memcpy(words, s->rng_global, sizeof(uint32_t) * 624);
    *pos = (int)s->rng_global[624];
    return FGX_OK;
}

// the move-internal engine (Movebase::_slump, move.h:356-360)
extern "C" int fgx_rng_set_internal(fgx_system *s, const uint32_t words[624], int pos)
{
    FGX_CHECK(s && words && pos >= 0 && pos <= 624, FGX_EARG, "fgx_rng_set_internal: bad arguments");
    if (int rc = pull_device_state(s)) return rc;
    std::memcpy(s->rng_internal, words, sizeof(uint32_t) * 624);
    s->rng_internal[624] = (uint32_t)pos;
    return FGX_OK;
}

extern "C" int fgx_rng_get_internal(fgx_system *s, uint32_t words[624], int *pos)
{
    FGX_CHECK(s && words && pos, FGX_EARG, "fgx_rng_get_internal: bad arguments");
    if (int rc = pull_device_state(s)) return rc;
    std::memcpy(words, s->rng_internal, sizeof(uint32_t) * 624);
    *pos = (int)s->rng_internal[624];
    return FGX_OK;
}

extern "C" int fgx_propagator_init(fgx_system *s)
{
    FGX_CHECK(s, FGX_EARG, "fgx_propagator_init: null system");
    if (int rc = pull_device_state(s)) return rc;
    std::memcpy(s->rng_internal, s->rng_global, sizeof(s->rng_global));   // move.h:7781
    s->propagator_started = false;
    s->dusum = 0;
    return FGX_OK;
}

// folds one chain's launch statistics into the system and the caller's sum
void collect_stats(fgx_system *s, const long *si, const double *sd, fgx_stats *st,
                          double seconds)
{
    if (si[0] > 0) s->move_cnt[0] += si[0];
    s->dusum += sd[1];
    if (s->wall_khz <= 0) {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s->device) != hipSuccess)
            khz = 0;
        s->wall_khz = khz;
    }
    s->chain_seconds = s->wall_khz > 0 ? (double)si[4] / (1e3 * s->wall_khz) : 0.0;
    if (s->cell_nc > 0 && si[6]) {
        // a cell filled up: the chain finished the launch on the all-pairs
        // screen; the next launch rebuilds the list, with twice the capacity
        // unless a capacity was set
        s->cell_valid = false;
        if (s->cell_mode == 1) s->cell_cap_min = 2 * s->cell_cap;
    }
    if (st) {
        st->trials += si[0];
        st->accepted += si[1];
        st->rejected_inf += si[2];
        st->ewald_refresh += si[3];
        st->utot += sd[0];
        st->dusum += sd[1];
        st->seconds = seconds;
    }
}

extern "C" int fgx_propagate(fgx_system *s, int ncalls, fgx_stats *st, double *du_out,
                             int *acc_out, long cap)
{
    FGX_CHECK(s && ncalls >= 0, FGX_EARG, "fgx_propagate: bad arguments");
    if (st) std::memset(st, 0, sizeof(*st));
    double *d_du = nullptr;
    int *d_acc = nullptr;
    if (cap > 0 && du_out && acc_out) {
        FGX_HIP(hipMalloc(&d_du, sizeof(double) * cap));
        FGX_HIP(hipMalloc(&d_acc, sizeof(int) * cap));
    } else {
        cap = 0;
    }
    std::vector<fgx_system *> sys{s};
    double secs = 0;
    std::vector<long> si;
    std::vector<double> sd;
    int rc = run_chains(sys, ncalls, d_du, d_acc, cap, &secs, &si, &sd);
    if (!rc) collect_stats(s, si.data(), sd.data(), st, secs);
    if (!rc && cap > 0) {
        const long ncopy = std::min(cap, si[0]);
        if (ncopy > 0) {
            FGX_HIP(hipMemcpy(du_out, d_du, sizeof(double) * ncopy, hipMemcpyDeviceToHost));
            FGX_HIP(hipMemcpy(acc_out, d_acc, sizeof(int) * ncopy, hipMemcpyDeviceToHost));
        }
    }
    if (d_du) (void)hipFree(d_du);
    if (d_acc) (void)hipFree(d_acc);
    return rc;
}

extern "C" int fgx_chain_seconds(const fgx_system *s, double *seconds)
{
    FGX_CHECK(s && seconds, FGX_EARG, "fgx_chain_seconds: bad arguments");
    *seconds = s->chain_seconds;
    return FGX_OK;
}

extern "C" int fgx_propagate_batch(fgx_system **systems, int nsys, int ncalls, fgx_stats *st)
{
    FGX_CHECK(systems && nsys > 0 && ncalls >= 0, FGX_EARG, "fgx_propagate_batch: bad arguments");
    std::vector<fgx_system *> sys(systems, systems + nsys);
    for (auto *s : sys) FGX_CHECK(s, FGX_EARG, "fgx_propagate_batch: null system");
    for (int a = 0; a < nsys; ++a)
        for (int b = a + 1; b < nsys; ++b)
            FGX_CHECK(sys[a] != sys[b], FGX_EARG, "fgx_propagate_batch: a system appears twice");
    if (st) std::memset(st, 0, sizeof(*st));
    double secs = 0;
    std::vector<long> si;
    std::vector<double> sd;
    int rc = run_chains(sys, ncalls, nullptr, nullptr, 0, &secs, &si, &sd);
    for (int c = 0; c < nsys && !rc; ++c)
        collect_stats(sys[c], &si[8 * (size_t)c], &sd[8 * (size_t)c], st, secs);
    return rc;
}

extern "C" int fgx_propagate_batch_record(fgx_system **systems, int nsys, int ncalls,
                                          fgx_stats *st, double *du_out, int *acc_out, long cap,
                                          long *counts)
{
    FGX_CHECK(systems && nsys > 0 && ncalls >= 0 && du_out && acc_out && cap > 0 && counts,
              FGX_EARG, "fgx_propagate_batch_record: bad arguments");
    std::vector<fgx_system *> sys(systems, systems + nsys);
    for (auto *s : sys) FGX_CHECK(s, FGX_EARG, "fgx_propagate_batch_record: null system");
    for (int a = 0; a < nsys; ++a)
        for (int b = a + 1; b < nsys; ++b)
            FGX_CHECK(sys[a] != sys[b], FGX_EARG,
                      "fgx_propagate_batch_record: a system appears twice");
    if (st) std::memset(st, 0, sizeof(*st));
    const size_t tot = (size_t)nsys * (size_t)cap;
    double *d_du = nullptr;
    int *d_acc = nullptr;
    FGX_HIP(hipMalloc(&d_du, sizeof(double) * tot));
    if (hipMalloc(&d_acc, sizeof(int) * tot) != hipSuccess) {
        (void)hipFree(d_du);
        set_error("fgx_propagate_batch_record: out of device memory");
        return FGX_EHIP;
    }
    double secs = 0;
    std::vector<long> si;
    std::vector<double> sd;
    int rc = run_chains(sys, ncalls, d_du, d_acc, cap, &secs, &si, &sd);
    for (int c = 0; c < nsys && !rc; ++c)
        collect_stats(sys[c], &si[8 * (size_t)c], &sd[8 * (size_t)c], st, secs);
    if (!rc) {
        for (int c = 0; c < nsys; ++c) counts[c] = std::min(cap, si[8 * (size_t)c]);
        if (hipMemcpy(du_out, d_du, sizeof(double) * tot, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(acc_out, d_acc, sizeof(int) * tot, hipMemcpyDeviceToHost) != hipSuccess) {
            set_error("fgx_propagate_batch_record: copy of the records failed");
            rc = FGX_EHIP;
        }
    }
    (void)hipFree(d_du);
    (void)hipFree(d_acc);
    return rc;
}

// ------------------------------------------------------------------------
// launch configuration and autotuner
// ------------------------------------------------------------------------

extern "C" int fgx_set_launch(fgx_system *s, int threads, int unroll)
{
    FGX_CHECK(s && (threads == 0 || threads == 256 || threads == 512) &&
                  (unroll == 0 || unroll == 4 || unroll == 8),
              FGX_EARG, "fgx_set_launch: threads in {0,256,512}, unroll in {0,4,8}");
    s->launch_threads = threads;
    s->launch_unroll = unroll;
    return FGX_OK;
}

extern "C" int fgx_set_energy_matrix(fgx_system *s, int on)
{
    FGX_CHECK(s, FGX_EARG, "fgx_set_energy_matrix: null system");
    s->emat_on = on != 0;
    s->emat_valid = false;
    return FGX_OK;
}

extern "C" int fgx_set_cell_list(fgx_system *s, int mode)
{
    FGX_CHECK(s && mode >= 0 && (mode <= 1 || (mode >= 2 && mode <= 4096)), FGX_EARG,
              "fgx_set_cell_list: mode is 0 (off), 1 (automatic) or a capacity 2..4096");
    s->cell_mode = mode;
    s->cell_cap_min = 0;
    s->cell_valid = false;
    return FGX_OK;
}

extern "C" int fgx_set_ewald_state(fgx_system *s, int hbm)
{
    FGX_CHECK(s && (hbm == 0 || hbm == 1), FGX_EARG, "fgx_set_ewald_state: hbm is 0 or 1");
    s->ewald_state_hbm = hbm;
    return FGX_OK;
}

extern "C" int fgx_set_lds_partners(fgx_system *s, int on)
{
    FGX_CHECK(s && (on == 0 || on == 1), FGX_EARG, "fgx_set_lds_partners: on is 0 or 1");
    s->lds_partners = on;
    return FGX_OK;
}

extern "C" int fgx_set_block_sweep(fgx_system *s, int mode, int nsplit)
{
    FGX_CHECK(s && mode >= 0 && mode <= 2 && nsplit >= 0 && nsplit <= BLK_NSPLIT_MAX, FGX_EARG,
              "fgx_set_block_sweep: mode is 0, 1 or 2, nsplit in 0..32");
    s->blk_mode = mode;
    s->blk_nsplit = nsplit;
    return FGX_OK;
}

extern "C" int fgx_get_block_parts(const fgx_system *s, int *parts)
{
    FGX_CHECK(s && parts, FGX_EARG, "fgx_get_block_parts: bad arguments");
    *parts = s->blk_parts;
    return FGX_OK;
}

extern "C" int fgx_get_cell_list(const fgx_system *s, int *cells_per_axis, int *capacity)
{
    FGX_CHECK(s && cells_per_axis && capacity, FGX_EARG, "fgx_get_cell_list: bad arguments");
    *cells_per_axis = s->cell_mode ? s->cell_nc : 0;
    *capacity = s->cell_mode ? s->cell_cap : 0;
    return FGX_OK;
}

extern "C" int fgx_get_launch(fgx_system *s, int *threads, int *unroll)
{
    FGX_CHECK(s && threads && unroll, FGX_EARG, "fgx_get_launch: bad arguments");
    *threads = s->launch_threads > 0 ? s->launch_threads : chain_block(s);
    *unroll = s->launch_unroll > 0 ? s->launch_unroll : 4;
    return FGX_OK;
}

// Time `ncalls` Propagator::move() calls of a batch of chains under each
// schedule (block size x partner-loop unroll) and keep the fastest on all of
// them.  The chains advance (use it as the warm-up of a production batch).
extern "C" int fgx_autotune_batch(fgx_system **systems, int nsys, int ncalls, int *threads,
                                  int *unroll, double *best_seconds)
{
    FGX_CHECK(systems && nsys > 0 && ncalls > 0, FGX_EARG, "fgx_autotune_batch: bad arguments");
    const int cand_t[2] = {256, 512};
    const int cand_u[2] = {4, 8};
    double best = 1e300;
    int bt = 0, bu = 0;
    for (int t : cand_t)
        for (int u : cand_u) {
            if (systems[0]->pot.precision != FGX_MIXED && u != 4) continue;
            for (int c = 0; c < nsys; ++c) {
                systems[c]->launch_threads = t;
                systems[c]->launch_unroll = u;
            }
            fgx_stats st;
            int rc = fgx_propagate_batch(systems, nsys, ncalls, &st);
            if (rc) return rc;
            if (st.seconds < best) { best = st.seconds; bt = t; bu = u; }
        }
    for (int c = 0; c < nsys; ++c) {
        systems[c]->launch_threads = bt;
        systems[c]->launch_unroll = bu;
    }
    if (threads) *threads = bt;
    if (unroll) *unroll = bu;
    if (best_seconds) *best_seconds = best;
    return FGX_OK;
}

extern "C" int fgx_autotune(fgx_system *probe, int ncalls, int *threads, int *unroll,
                            double *best_seconds)
{
    return fgx_autotune_batch(&probe, 1, ncalls, threads, unroll, best_seconds);
}
