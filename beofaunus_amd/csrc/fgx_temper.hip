// fgx_temper.hip -- Move::ParallelTempering (move.h:6520-6772) for R ranks
// on one device.
//
// The reference runs one rank per MPI process: a rank keeps its Hamiltonian
// (here its Bjerrum length: temperature enters through lB, physconst.h:91-94),
// its global and internal random engines (slump.h:126, move.h:356-360) and its
// Ewald state, and configurations travel between ranks (ParticleTransmitter,
// mpi.h:238-331).  Here every configuration is an fgx_system resident on the
// same GPU and stays where it is; what travels on an accepted swap is the
// rank's label: lB, both engines and the Ewald accept counter (and, with
// Ewald, the structure factor the reference leaves behind).  Swapping labels
// instead of coordinates keeps every handle's device buffers, cell lists and
// prepared launches valid.
//
// One exchange attempt = one ParallelTempering move on every rank
// (Movebase::move, move.h:515-555, with an empty molecule list):
//   run():      one canonical of the rank's internal engine against prob;
//   _trialMove: the partner direction from the controller engine
//               (MPIController::random, mpi.h:48), shared by all ranks;
//   _energyChange: du = (U_r(x_p) - U_r(x_r)) + (U_p(x_r) - U_p(x_p))
//               (move.h:6701-6723), each U a full Energy::systemEnergy;
//   metropolis: one canonical of each rank's global engine; the pair takes
//               the lower rank's (the reference decides on both ranks and
//               relies on their engines being in lock-step, move.h:557-566).
// Energies come from one split per configuration, U(lB) = lB A + B: A the
// electrostatic part per unit lB, B the rest, one O(N^2) pass for all
// replicas instead of four systemEnergy calls per pair.  With Ewald the
// reference's systemEnergy(trial) after setSpace takes the reciprocal term
// from the rank's own S(k) (ewald.h:716-757), so at equal volume the
// reciprocal terms cancel in each rank's unew - uold and A carries the
// real-space and surface terms only.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "fgx_sweep_kernels.hpp"

namespace fgx {

static constexpr int PT_MAXR = 256;      // ranks per temper (LDS tables)
static constexpr int PT_RB = 256;        // energy kernel block
static constexpr int PT_TILE = 64;       // staged rows per energy block
static constexpr int PT_GX = 8;          // column blocks per row tile

// per-configuration view for the exchange kernels
struct PTSys {
    Box box;
    Pot pot;
    const double *x, *y, *z, *q, *r;
    const int *t;
    const double *tab, *eff;
    int n;
    double surf1;        // Ewald surface prefactor per unit lB (0: none)
    uint32_t *rng;       // 2 x 625 words: global, internal engine
    int *ew_cnt;         // Ewald accept counter on the device, or null
    ChainDev *chain;     // this configuration's entry of a prepared batch, or null
    double fx_unit;      // L / 2^32 of a cubic box (ChainDev::fx_lbs = lB / unit)
    double recip_base;   // 2 pi / V
    double surf_base;    // const_inf 2 pi / ((2 eps + 1) V), or 0
    double *partial;     // energy partial sums, 2 per block
    int bstat;           // index of its chain in the batch statistics (graph mode)
};

// state of one temper on the device
struct PTDev {
    int R;
    const PTSys *sys;    // per configuration (handle)
    int *slot;           // slot[r] = configuration held by rank r
    const double *rank_lB;
    const double *AB;    // per configuration {A, B}
    uint32_t *ctl;       // controller engine, 625 words
    double prob;
    int *err;            // set when the ranks' run() draws disagree
    // records: round index (or null: record 0), per round and rank
    int *round;
    int cap_rounds;
    int *rec_acc, *rec_part;
    double *rec_du, *rec_ret;
    int *rec_slot;
    long *rec_cycles;    // per round and configuration: the sweep's wall cycles
    // graph mode: per-configuration launch statistics folded into totals
    const long *bstat_i;
    const double *bstat_d;
    long *tot_i;
    double *tot_d;
};

// ------------------------------------------------------------------------
// libstdc++ mt19937 on one thread, state in device memory (624 words + pos)
// ------------------------------------------------------------------------
__device__ uint32_t mtg_next(uint32_t *st)
{
    int p = (int)st[624];
    if (p >= 624) {
        for (int k = 0; k < 624; ++k) {
            const uint32_t y = (st[k] & 0x80000000u) | (st[(k + 1) % 624] & 0x7fffffffu);
            st[k] = st[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        p = 0;
    }
    uint32_t z = st[p];
    st[624] = (uint32_t)(p + 1);
    z ^= (z >> 11);
    z ^= (z << 7) & 0x9d2c5680u;
    z ^= (z << 15) & 0xefc60000u;
    z ^= (z >> 18);
    return z;
}

#pragma clang fp contract(off)
__device__ double mtg_canonical(uint32_t *st)
{
    double sum = (double)mtg_next(st);
    sum += (double)mtg_next(st) * 4294967296.0;
    double r = sum / 18446744073709551616.0;
    if (r >= 1.0) r = 0x1.fffffffffffffp-1;
    return r;
}
#pragma clang fp contract(on)

// ------------------------------------------------------------------------
// Energy split of every configuration: sum over i < j of the electrostatic
// pair term at lB = 1 and of the short-range term, as Energybase::systemEnergy
// (energy.h:169-180) sums every pair of the space.  Grid: (PT_GX, row tiles,
// configurations).
// ------------------------------------------------------------------------
template <int ELEC, int SR>
__global__ __launch_bounds__(PT_RB) void k_pt_energy(const PTSys *__restrict__ sys)
{
    const PTSys &S = sys[blockIdx.z];
    __shared__ double sx[PT_TILE], sy[PT_TILE], sz[PT_TILE], sq[PT_TILE], sr[PT_TILE];
    __shared__ int st[PT_TILE];
    __shared__ double red[PT_RB / 64];
    const int i0 = blockIdx.y * PT_TILE;
    const int ni = max(0, min(PT_TILE, S.n - i0));
    const int tid = threadIdx.x;
    if (tid < ni) {
        const int i = i0 + tid;
        sx[tid] = S.x[i]; sy[tid] = S.y[i]; sz[tid] = S.z[i];
        sq[tid] = S.q[i]; sr[tid] = S.r[i]; st[tid] = S.t[i];
    }
    __syncthreads();
    Pot P1 = S.pot;
    P1.lB = 1.0;
    double ea = 0.0, eb = 0.0;
    if (ni > 0) {
        for (int j = i0 + 1 + blockIdx.x * PT_RB + tid; j < S.n; j += gridDim.x * PT_RB) {
            const double bx = S.x[j], by = S.y[j], bz = S.z[j], bq = S.q[j], br = S.r[j];
            const int bt = S.t[j];
            for (int k = 0; k < ni; ++k) {
                if (i0 + k >= j) break;
                double r2;
                if (ELEC == ELEC_EWALD || S.pot.nbkind == NB_VECTOR) {
                    double rx, ry, rz;
                    vdist_d(S.box, sx[k], sy[k], sz[k], bx, by, bz, rx, ry, rz);
                    r2 = rx * rx + (ry * ry + rz * rz);
                } else {
                    r2 = sqdist_d(S.box, sx[k], sy[k], sz[k], bx, by, bz);
                }
                if constexpr (ELEC != ELEC_NONE)
                    ea += pair_u64_t<ELEC, SR_NONE>(P1, S.tab, r2, sq[k], bq, sr[k], br, st[k], bt);
                if constexpr (SR != SR_NONE)
                    eb += pair_u64_t<ELEC_NONE, SR>(S.pot, S.tab, r2, sq[k], bq, sr[k], br, st[k], bt);
            }
        }
    }
    ea = block_sum(ea, red);
    eb = block_sum(eb, red);
    if (tid == 0) {
        const int b = blockIdx.y * gridDim.x + blockIdx.x;
        S.partial[2 * b] = ea;
        S.partial[2 * b + 1] = eb;
    }
}

// fixed-order sum of a configuration's partials, plus the Ewald surface term
// per unit lB (getSurfaceEnergy, ewald.h:278-292) from the dipole sum
__global__ __launch_bounds__(PT_RB) void k_pt_reduce(const PTSys *__restrict__ sys, int nblk,
                                                      double *AB)
{
    const PTSys &S = sys[blockIdx.x];
    __shared__ double red[PT_RB / 64];
    const int tid = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (int k = tid; k < nblk; k += PT_RB) {
        a += S.partial[2 * k];
        b += S.partial[2 * k + 1];
    }
    a = block_sum(a, red);
    b = block_sum(b, red);
    double mx = 0.0, my = 0.0, mz = 0.0;
    if (S.surf1 != 0.0) {
        for (int j = tid; j < S.n; j += PT_RB) {
            const double w = S.q[j] * S.eff[S.t[j]];
            mx += w * S.x[j];
            my += w * S.y[j];
            mz += w * S.z[j];
        }
    }
    mx = block_sum(mx, red);
    my = block_sum(my, red);
    mz = block_sum(mz, red);
    if (tid == 0) {
        if (S.surf1 != 0.0) a += S.surf1 * (mx * mx + (my * my + mz * mz));
        AB[2 * blockIdx.x] = a;
        AB[2 * blockIdx.x + 1] = b;
    }
}

// A configuration takes the label of rank lB: the scalars a prepared
// ChainDev derives from lB (prepare_chain) follow it.
__device__ void pt_patch_chain(const PTSys &S, double lB)
{
    ChainDev *C = S.chain;
    if (!C) return;
    C->pot.lB = lB;
    C->pot.lBf = (float)lB;
    if (C->fx) C->fx_lbs = (float)(lB / S.fx_unit);
    C->recip_pref = S.recip_base * lB;
    C->surf_pref = S.surf_base * lB;
}

// One exchange attempt of every rank (one block of 64 threads).
__global__ __launch_bounds__(64) void k_pt_decide(PTDev D)
{
    __shared__ double du_self[PT_MAXR], du[PT_MAXR], uni[PT_MAXR], ret[PT_MAXR];
    __shared__ int part[PT_MAXR], acc[PT_MAXR], old_slot[PT_MAXR];
    __shared__ int swaps[PT_MAXR][2];
    __shared__ int nswap, go, rd;
    const int tid = threadIdx.x, R = D.R;
    if (tid == 0) {
        rd = D.round ? *D.round : 0;
        if (D.round) *D.round = rd + 1;
        // graph mode: fold each configuration's launch statistics
        if (D.bstat_i)
            for (int c = 0; c < R; ++c) {
                const int b = D.sys[c].bstat;
                for (int k = 0; k < 8; ++k) {
                    if (k != 4 && k != 6) D.tot_i[8 * c + k] += D.bstat_i[8 * b + k];
                    D.tot_d[8 * c + k] += D.bstat_d[8 * b + k];
                }
                D.tot_i[8 * c + 4] = D.bstat_i[8 * b + 4];
                D.tot_i[8 * c + 6] |= D.bstat_i[8 * b + 6];
                if (D.rec_cycles && rd < D.cap_rounds)
                    D.rec_cycles[(size_t)rd * R + c] = D.bstat_i[8 * b + 4];
            }
        for (int r = 0; r < R; ++r) {
            old_slot[r] = D.slot[r];
            acc[r] = 0;
            du[r] = 0.0;
            ret[r] = 0.0;
            part[r] = -1;
        }
        nswap = 0;
        // Movebase::run (move.h:575-581) on every rank's internal engine
        int run0 = -1;
        bool same = true;
        for (int r = 0; r < R; ++r) {
            const int run = mtg_canonical(D.sys[D.slot[r]].rng + 625) < D.prob ? 1 : 0;
            if (run0 < 0) run0 = run;
            else if (run != run0) same = false;
        }
        if (!same) *D.err = 1;
        go = (same && run0 == 1) ? 1 : 0;
        if (go) {
            // findPartner (move.h:6621-6633)
            const int dr = (mtg_canonical(D.ctl) > 0.5) ? 1 : -1;
            for (int r = 0; r < R; ++r) {
                const int p = (r % 2 == 0) ? r + dr : r - dr;
                part[r] = (p >= 0 && p < R) ? p : -1;
            }
            // _energyChange (move.h:6701-6723): unew - uold on each rank
            for (int r = 0; r < R; ++r) {
                du_self[r] = 0.0;
                if (part[r] < 0) continue;
                const int a = D.slot[r], b = D.slot[part[r]];
                const double lB = D.rank_lB[r];
                const double unew = lB * D.AB[2 * b] + D.AB[2 * b + 1];
                const double uold = lB * D.AB[2 * a] + D.AB[2 * a + 1];
                du_self[r] = unew - uold;
            }
            for (int r = 0; r < R; ++r) du[r] = part[r] < 0 ? __builtin_inf() : du_self[r] + du_self[part[r]];
            // metropolis (move.h:567-573) on every rank's global engine
            for (int r = 0; r < R; ++r) uni[r] = mtg_canonical(D.sys[D.slot[r]].rng);
            for (int r = 0; r < R; ++r) {
                const int p = part[r];
                if (p < 0 || p < r) continue;
                const int a = !(uni[r] > exp(-du[r]));
                acc[r] = acc[p] = a;
                if (a) {
                    const int ha = D.slot[r], hb = D.slot[p];
                    swaps[nswap][0] = ha;
                    swaps[nswap][1] = hb;
                    ++nswap;
                    D.slot[r] = hb;
                    D.slot[p] = ha;
                    ret[r] = du_self[r];
                    ret[p] = du_self[p];
                    pt_patch_chain(D.sys[hb], D.rank_lB[r]);
                    pt_patch_chain(D.sys[ha], D.rank_lB[p]);
                    // the ranks' Ewald accept counters travel with them
                    if (D.sys[ha].ew_cnt && D.sys[hb].ew_cnt) {
                        const int t = *D.sys[ha].ew_cnt;
                        *D.sys[ha].ew_cnt = *D.sys[hb].ew_cnt;
                        *D.sys[hb].ew_cnt = t;
                    }
                }
            }
        }
        if (rd < D.cap_rounds)
            for (int r = 0; r < R; ++r) {
                D.rec_acc[(size_t)rd * R + r] = acc[r];
                D.rec_part[(size_t)rd * R + r] = part[r];
                D.rec_du[(size_t)rd * R + r] = du[r];
                D.rec_ret[(size_t)rd * R + r] = ret[r];
                D.rec_slot[(size_t)rd * R + r] = D.slot[r];
            }
    }
    __syncthreads();
    // the ranks' engines travel with them: swap both engines of each pair
    for (int w = 0; w < nswap; ++w) {
        uint32_t *ra = D.sys[swaps[w][0]].rng, *rb = D.sys[swaps[w][1]].rng;
        for (int k = tid; k < 2 * 625; k += 64) {
            const uint32_t t = ra[k];
            ra[k] = rb[k];
            rb[k] = t;
        }
    }
}

// ------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------
static int launch_pt_energy(const Pot &P, hipStream_t st, const PTSys *d_sys, int R, int tiles)
{
    const dim3 grid(PT_GX, tiles, R);
#define FGX_PE(E, S) \
    if (P.elec == E && P.sr == S) { hipLaunchKernelGGL((k_pt_energy<E, S>), grid, dim3(PT_RB), 0, st, d_sys); return FGX_OK; }
    FGX_PE(ELEC_NONE, SR_HS) FGX_PE(ELEC_NONE, SR_LJ)
    FGX_PE(ELEC_COULOMB, SR_NONE) FGX_PE(ELEC_COULOMB, SR_HS) FGX_PE(ELEC_COULOMB, SR_LJ)
    FGX_PE(ELEC_DH, SR_NONE) FGX_PE(ELEC_DH, SR_HS) FGX_PE(ELEC_DH, SR_LJ)
    FGX_PE(ELEC_EWALD, SR_NONE) FGX_PE(ELEC_EWALD, SR_HS) FGX_PE(ELEC_EWALD, SR_LJ)
#undef FGX_PE
    set_error("ParallelTempering: no energy kernel for this Hamiltonian");
    return FGX_EUNSUPPORTED;
}

}  // namespace fgx

using namespace fgx;

struct fgx_temper {
    int R = 0;
    std::vector<fgx_system *> cfg;   // configurations (handles), fixed order
    std::vector<double> lB;          // lB of rank r
    std::vector<int> slot;           // host mirror: slot[r] = configuration at rank r
    double prob = 1.0;
    int tiles = 0;
    hipStream_t st = nullptr;        // the first configuration's stream
    // device state
    PTSys *d_sys = nullptr;
    std::vector<PTSys> h_sys;
    int *d_slot = nullptr, *d_err = nullptr, *d_round = nullptr;
    double *d_rank_lB = nullptr, *d_AB = nullptr, *d_partial = nullptr;
    uint32_t *d_ctl = nullptr;
    int cap_rounds = 0;
    int *d_rec_acc = nullptr, *d_rec_slot = nullptr, *d_rec_part = nullptr;
    double *d_rec_du = nullptr, *d_rec_ret = nullptr;
    long *d_rec_cycles = nullptr;
    long *d_tot_i = nullptr;
    double *d_tot_d = nullptr;
    // records of the last call (rounds x R)
    int last_rounds = 0;
    std::vector<int> rec_acc, rec_slot, rec_part;
    std::vector<double> rec_du, rec_ret;
    std::vector<long> rec_cycles;
    // each rank's ParallelTempering::move return value in the last exchange
    std::vector<double> last_ret;
};

static void temper_free(fgx_temper *t)
{
    void *ptrs[] = {t->d_sys, t->d_slot, t->d_err, t->d_round, t->d_rank_lB, t->d_AB,
                    t->d_partial, t->d_ctl, t->d_rec_acc, t->d_rec_slot, t->d_rec_part, t->d_rec_du,
                    t->d_rec_ret, t->d_rec_cycles, t->d_tot_i, t->d_tot_d};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
}

// record buffers for `rounds` exchange rounds
static int ensure_records(fgx_temper *t, int rounds)
{
    if (t->cap_rounds >= rounds) return FGX_OK;
    void *ptrs[] = {t->d_rec_acc, t->d_rec_slot, t->d_rec_part, t->d_rec_du, t->d_rec_ret,
                    t->d_rec_cycles};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    t->d_rec_acc = t->d_rec_slot = t->d_rec_part = nullptr;
    t->d_rec_du = t->d_rec_ret = nullptr;
    t->d_rec_cycles = nullptr;
    t->cap_rounds = 0;
    const size_t n = (size_t)rounds * t->R;
    FGX_HIP(hipMalloc(&t->d_rec_acc, sizeof(int) * n));
    FGX_HIP(hipMalloc(&t->d_rec_slot, sizeof(int) * n));
    FGX_HIP(hipMalloc(&t->d_rec_part, sizeof(int) * n));
    FGX_HIP(hipMalloc(&t->d_rec_du, sizeof(double) * n));
    FGX_HIP(hipMalloc(&t->d_rec_ret, sizeof(double) * n));
    FGX_HIP(hipMalloc(&t->d_rec_cycles, sizeof(long) * n));
    t->cap_rounds = rounds;
    return FGX_OK;
}

// bitwise equality: tables may hold NaN (the effective-charge ratio 0/0 of a
// neutral type, ewald.h:535-548), which == never matches
static bool same_bits(const std::vector<double> &a, const std::vector<double> &b)
{
    return a.size() == b.size() &&
           (a.empty() || std::memcmp(a.data(), b.data(), sizeof(double) * a.size()) == 0);
}

extern "C" int fgx_temper_create(fgx_system **replicas, int R, uint32_t ctl_seed, double prob,
                                 fgx_temper **out)
{
    FGX_CHECK(replicas && R >= 2 && R <= PT_MAXR && out, FGX_EARG,
              "fgx_temper_create: need 2..256 replicas");
    FGX_CHECK(prob > 0.0 && prob <= 1.0, FGX_EARG, "fgx_temper_create: prob must be in (0, 1]");
    const fgx_system *s0 = replicas[0];
    for (int r = 0; r < R; ++r) {
        const fgx_system *s = replicas[r];
        FGX_CHECK(s, FGX_EARG, "fgx_temper_create: null replica");
        for (int q = 0; q < r; ++q)
            FGX_CHECK(replicas[q] != s, FGX_EARG, "fgx_temper_create: a replica is listed twice");
        // the reference aborts when partners differ in size (move.h:6686-6687)
        FGX_CHECK(s->n == s0->n, FGX_EARG,
                  "fgx_temper_create: replicas must have the same particle count");
        // ranks may differ in lB only: an exchange evaluates each configuration
        // with the partner's Bjerrum length and nothing else (e.g. LJ eps given
        // in kJ/mol converts at each rank's own temperature, species.h:293)
        const Pot &a = s->pot, &b = s0->pot;
        bool same = a.elec == b.elec && a.sr == b.sr && a.nbkind == b.nbkind &&
                    a.ntypes == b.ntypes && a.kappa == b.kappa && a.g2g_rc2 == b.g2g_rc2 &&
                    a.ew_alpha == b.ew_alpha && a.ew_kappa == b.ew_kappa && a.ew_rc2 == b.ew_rc2 &&
                    a.lj_rc2 == b.lj_rc2 && s->has_ewald == s0->has_ewald && same_bits(s->tab, s0->tab) &&
                    s->h_charge == s0->h_charge && s->h_radius == s0->h_radius &&
                    s->h_type == s0->h_type && s->ng == s0->ng;
        for (int k = 0; k < 3 && same; ++k) same = s->box.len[k] == s0->box.len[k];
        if (same && s->has_ewald)
            same = s->ew.K == s0->ew.K && s->ew.kcc == s0->ew.kcc &&
                   s->ew.eps_surf == s0->ew.eps_surf && s->ew.const_inf == s0->ew.const_inf &&
                   s->ew.update_frequency == s0->ew.update_frequency && same_bits(s->h_eff, s0->h_eff);
        FGX_CHECK(same, FGX_EUNSUPPORTED,
                  "fgx_temper_create: replicas differ in more than lB (Hamiltonian, box, "
                  "composition or Ewald settings)");
        // ExternalPressure and tempering together would swap volumes as well
        // (move.h:6651, 6710-6711), which this label exchange does not do
        FGX_CHECK(!s->press_on, FGX_EUNSUPPORTED,
                  "fgx_temper_create: replicas with ExternalPressure are not supported");
        FGX_CHECK(s->pot.nbkind != NB_CUTG2G, FGX_EUNSUPPORTED,
                  "fgx_temper_create: NonbondedCutg2g replicas are not supported");
    }
    fgx_temper *t = new fgx_temper();
    t->R = R;
    t->prob = prob;
    t->cfg.assign(replicas, replicas + R);
    t->st = replicas[0]->stream;
    for (int r = 0; r < R; ++r) {
        t->lB.push_back(replicas[r]->pot.lB);
        t->slot.push_back(r);
    }
    const int n = s0->n;
    t->tiles = std::max(1, (n + PT_TILE - 1) / PT_TILE);
    const int nblk = t->tiles * PT_GX;
    int rc = FGX_OK;
    auto fail = [&](int code) {
        temper_free(t);
        delete t;
        return code;
    };
#define PT_HIP(call)                                                     \
    do {                                                                 \
        hipError_t e_ = (call);                                          \
        if (e_ != hipSuccess) {                                          \
            set_error(std::string("HIP error ") + hipGetErrorString(e_) + \
                      " in fgx_temper_create");                          \
            return fail(FGX_EHIP);                                       \
        }                                                                \
    } while (0)
    PT_HIP(hipMalloc(&t->d_sys, sizeof(PTSys) * R));
    PT_HIP(hipMalloc(&t->d_slot, sizeof(int) * R));
    PT_HIP(hipMalloc(&t->d_err, sizeof(int)));
    PT_HIP(hipMalloc(&t->d_round, sizeof(int)));
    PT_HIP(hipMalloc(&t->d_rank_lB, sizeof(double) * R));
    PT_HIP(hipMalloc(&t->d_AB, sizeof(double) * 2 * R));
    PT_HIP(hipMalloc(&t->d_partial, sizeof(double) * 2 * (size_t)nblk * R));
    PT_HIP(hipMalloc(&t->d_ctl, sizeof(uint32_t) * 625));
    PT_HIP(hipMalloc(&t->d_tot_i, sizeof(long) * 8 * R));
    PT_HIP(hipMalloc(&t->d_tot_d, sizeof(double) * 8 * R));
    const double pi = std::acos(-1.);
    t->h_sys.resize(R);
    for (int c = 0; c < R; ++c) {
        fgx_system *s = replicas[c];
        PTSys &P = t->h_sys[c];
        P.box = s->box;
        P.pot = s->pot;
        P.x = s->v[0].x; P.y = s->v[0].y; P.z = s->v[0].z;
        P.q = s->d_q; P.r = s->d_r; P.t = s->d_type;
        P.tab = s->d_tab;
        P.eff = s->has_ewald ? s->ew.eff : nullptr;
        P.n = s->n;
        P.surf1 = (s->has_ewald && s->ew.const_inf >= 0.5)
                      ? s->ew.const_inf * (2 * pi / ((2 * s->ew.eps_surf + 1) * s->ew.V))
                      : 0.0;
        P.rng = s->d_rng;
        P.ew_cnt = nullptr;
        P.chain = nullptr;
        P.fx_unit = s->box.len[0] / 4294967296.0;
        P.recip_base = s->has_ewald ? 2 * pi / s->ew.V : 0.0;
        P.surf_base = P.surf1;
        P.partial = t->d_partial + 2 * (size_t)nblk * c;
        P.bstat = c;
        if (s->has_ewald && !s->ew.cnt_dev) PT_HIP(hipMalloc(&s->ew.cnt_dev, sizeof(int)));
        P.ew_cnt = s->has_ewald ? s->ew.cnt_dev : nullptr;
    }
    std::vector<uint32_t> ctl(625);
    mt_seed_host(ctl.data(), ctl_seed);
    PT_HIP(hipMemcpyAsync(t->d_sys, t->h_sys.data(), sizeof(PTSys) * R, hipMemcpyHostToDevice, t->st));
    PT_HIP(hipMemcpyAsync(t->d_slot, t->slot.data(), sizeof(int) * R, hipMemcpyHostToDevice, t->st));
    PT_HIP(hipMemcpyAsync(t->d_rank_lB, t->lB.data(), sizeof(double) * R, hipMemcpyHostToDevice, t->st));
    PT_HIP(hipMemcpyAsync(t->d_ctl, ctl.data(), sizeof(uint32_t) * 625, hipMemcpyHostToDevice, t->st));
    PT_HIP(hipStreamSynchronize(t->st));
#undef PT_HIP
    (void)rc;
    *out = t;
    return FGX_OK;
}

// the configurations' RNG states and Ewald counters become the device's
static int push_state(fgx_temper *t)
{
    for (fgx_system *s : t->cfg) {
        if (s->state_on_dev) continue;
        FGX_HIP(hipMemcpyAsync(s->d_rng, s->rng_global, sizeof(uint32_t) * 625,
                               hipMemcpyHostToDevice, t->st));
        FGX_HIP(hipMemcpyAsync(s->d_rng + 625, s->rng_internal, sizeof(uint32_t) * 625,
                               hipMemcpyHostToDevice, t->st));
        if (s->has_ewald)
            FGX_HIP(hipMemcpyAsync(s->ew.cnt_dev, &s->ew.cnt_accepted, sizeof(int),
                                   hipMemcpyHostToDevice, t->st));
        FGX_HIP(hipStreamSynchronize(t->st));
        s->state_on_dev = true;
    }
    return FGX_OK;
}

// energy split + decision, enqueued on t->st (graph capture takes them as they are)
static int enqueue_exchange(fgx_temper *t, const long *bstat_i, const double *bstat_d)
{
    const int R = t->R;
    int rc = launch_pt_energy(t->cfg[0]->pot, t->st, t->d_sys, R, t->tiles);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pt_reduce, dim3(R), dim3(PT_RB), 0, t->st, t->d_sys, t->tiles * PT_GX,
                       t->d_AB);
    PTDev D{};
    D.R = R;
    D.sys = t->d_sys;
    D.slot = t->d_slot;
    D.rank_lB = t->d_rank_lB;
    D.AB = t->d_AB;
    D.ctl = t->d_ctl;
    D.prob = t->prob;
    D.err = t->d_err;
    D.round = t->d_round;
    D.cap_rounds = t->cap_rounds;
    D.rec_acc = t->d_rec_acc;
    D.rec_part = t->d_rec_part;
    D.rec_du = t->d_rec_du;
    D.rec_ret = t->d_rec_ret;
    D.rec_slot = t->d_rec_slot;
    D.rec_cycles = bstat_i ? t->d_rec_cycles : nullptr;
    D.bstat_i = bstat_i;
    D.bstat_d = bstat_d;
    D.tot_i = t->d_tot_i;
    D.tot_d = t->d_tot_d;
    hipLaunchKernelGGL(k_pt_decide, dim3(1), dim3(64), 0, t->st, D);
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

// clears the round counter, the error flag and the statistics totals
__global__ void k_pt_reset(int *round, int *err, long *tot_i, double *tot_d, int n8)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) { *round = 0; *err = 0; }
    if (k < n8) { tot_i[k] = 0; tot_d[k] = 0.0; }
}

static int reset_counters(fgx_temper *t)
{
    const int n8 = 8 * t->R;
    hipLaunchKernelGGL(k_pt_reset, dim3((n8 + 255) / 256), dim3(256), 0, t->st, t->d_round,
                       t->d_err, t->d_tot_i, t->d_tot_d, n8);
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

// copies the records of `rounds` rounds back and brings the host mirrors of
// every configuration (lB, slots) up to date
static int pull_records(fgx_temper *t, int rounds)
{
    const int R = t->R;
    const size_t n = (size_t)rounds * R;
    int err = 0;
    t->rec_acc.resize(n);
    t->rec_slot.resize(n);
    t->rec_part.resize(n);
    t->rec_du.resize(n);
    t->rec_ret.resize(n);
    t->rec_cycles.assign(n, 0);
    FGX_HIP(hipMemcpyAsync(&err, t->d_err, sizeof(int), hipMemcpyDeviceToHost, t->st));
    FGX_HIP(hipMemcpyAsync(t->rec_acc.data(), t->d_rec_acc, sizeof(int) * n, hipMemcpyDeviceToHost, t->st));
    FGX_HIP(hipMemcpyAsync(t->rec_slot.data(), t->d_rec_slot, sizeof(int) * n, hipMemcpyDeviceToHost, t->st));
    FGX_HIP(hipMemcpyAsync(t->rec_part.data(), t->d_rec_part, sizeof(int) * n, hipMemcpyDeviceToHost, t->st));
    FGX_HIP(hipMemcpyAsync(t->rec_du.data(), t->d_rec_du, sizeof(double) * n, hipMemcpyDeviceToHost, t->st));
    FGX_HIP(hipMemcpyAsync(t->rec_ret.data(), t->d_rec_ret, sizeof(double) * n, hipMemcpyDeviceToHost, t->st));
    FGX_HIP(hipMemcpyAsync(t->slot.data(), t->d_slot, sizeof(int) * R, hipMemcpyDeviceToHost, t->st));
    FGX_HIP(hipStreamSynchronize(t->st));
    t->last_rounds = rounds;
    for (int r = 0; r < R; ++r) {
        fgx_system *s = t->cfg[t->slot[r]];
        if (s->pot.lB != t->lB[r]) {
            s->pot.lB = t->lB[r];
            s->pot.lBf = (float)t->lB[r];
            s->emat_valid = false;   // cached pair energies scale with lB
        }
    }
    FGX_CHECK(!err, FGX_ESTATE,
              "ParallelTempering: the ranks' run() draws disagree (their internal engines are "
              "not in lock-step; the reference would block here)");
    return FGX_OK;
}

// With Ewald, the structure factors the reference leaves after one exchange
// (ewald.h:613-648, 731-757): a rank with a partner recomputes S(k) from its
// own configuration (setSpace in systemEnergy); after an accepted swap it
// keeps that S(k) -- of the configuration it gave away -- and advances its
// accept counter (update(true): a refresh from the new configuration when
// due); after a rejection S(k) is recomputed again from its own
// configuration (_rejectMove's setSpace), i.e. it stays fresh.
static int ewald_after_exchange(fgx_temper *t, const std::vector<int> &before, double *drift)
{
    const int R = t->R;
    for (fgx_system *s : t->cfg)
        if (int rc = pull_device_state(s)) return rc;
    const int *part = t->rec_part.data();
    const int *acc = t->rec_acc.data();
    for (int r = 0; r < R; ++r)
        if (part[r] >= 0)
            if (int rc = ewald_set_space(t->cfg[before[r]])) return rc;
    // the recomputed structure factors change hands below
    for (fgx_system *s : t->cfg) FGX_HIP(hipStreamSynchronize(s->stream));
    for (int r = 0; r < R; ++r) {
        const int p = part[r];
        if (p < r || !acc[r]) continue;
        // configurations a (was at r) and b (was at p) swapped ranks: each now
        // carries the S(k) its new rank computed for the other
        fgx_system *a = t->cfg[before[r]], *b = t->cfg[before[p]];
        std::swap(a->ew.Q, b->ew.Q);
        std::swap(a->ew.Qt, b->ew.Qt);
    }
    for (int r = 0; r < R; ++r) {
        drift[r] = 0.0;
        if (!acc[r]) continue;
        fgx_system *s = t->cfg[t->slot[r]];
        if (int rc = ewald_accepted(s, &drift[r])) return rc;
    }
    for (fgx_system *s : t->cfg) FGX_HIP(hipStreamSynchronize(s->stream));
    return FGX_OK;
}

// one exchange attempt; acc[r] = 1 if rank r swapped; du[r] = the pair's
// energy change (+inf for a rank without partner)
extern "C" int fgx_temper_exchange(fgx_temper *t, int *acc, double *du)
{
    FGX_CHECK(t, FGX_EARG, "fgx_temper_exchange: null handle");
    const int R = t->R;
    for (fgx_system *s : t->cfg) FGX_HIP(hipStreamSynchronize(s->stream));
    int rc;
    if ((rc = push_state(t))) return rc;
    if ((rc = ensure_records(t, 1))) return rc;
    if ((rc = reset_counters(t))) return rc;
    const std::vector<int> before = t->slot;
    if ((rc = enqueue_exchange(t, nullptr, nullptr))) return rc;
    if ((rc = pull_records(t, 1))) return rc;
    std::vector<double> drift(R, 0.0);
    if (t->cfg[0]->has_ewald && (rc = ewald_after_exchange(t, before, drift.data()))) return rc;
    t->last_ret.assign(R, 0.0);
    for (int r = 0; r < R; ++r) {
        if (acc) acc[r] = t->rec_acc[r];
        if (du) du[r] = t->rec_du[r];
        // Movebase::move: the accepted move's return value, then
        // pot->update(acceptance) (move.h:538-546)
        if (t->rec_acc[r]) t->last_ret[r] = t->rec_ret[r] + drift[r];
    }
    return FGX_OK;
}

extern "C" int fgx_temper_returns(fgx_temper *t, double *ret)
{
    FGX_CHECK(t && ret, FGX_EARG, "fgx_temper_returns: bad arguments");
    FGX_CHECK((int)t->last_ret.size() == t->R, FGX_ESTATE, "fgx_temper_returns: no exchange yet");
    for (int r = 0; r < t->R; ++r) ret[r] = t->last_ret[r];
    return FGX_OK;
}

// `rounds` times: every configuration runs `calls` Propagator::move() calls
// of its registered move (one batched sweep launch), then one exchange
// attempt.  The round is captured once as a hipGraph (sweep kernel, energy
// split, decision) and replayed; host work happens before and after only.
extern "C" int fgx_temper_run(fgx_temper **tempers, int nt, int rounds, int calls,
                              fgx_stats *st, double *seconds)
{
    FGX_CHECK(tempers && nt >= 1 && rounds >= 1 && calls >= 0, FGX_EARG,
              "fgx_temper_run: bad arguments");
    std::vector<fgx_system *> sys;
    for (int k = 0; k < nt; ++k) {
        FGX_CHECK(tempers[k], FGX_EARG, "fgx_temper_run: null temper");
        FGX_CHECK(!tempers[k]->cfg[0]->has_ewald, FGX_EUNSUPPORTED,
                  "fgx_temper_run: Ewald replicas exchange through fgx_temper_exchange");
        for (fgx_system *s : tempers[k]->cfg)
            FGX_CHECK(!s->moves.empty() && !(s->moves[0].type == MOVE_TR &&
                                             s->pot.precision == FGX_MIXED && s->emat_on),
                      FGX_EUNSUPPORTED,
                      "fgx_temper_run: the group-pair energy cache scales with lB; switch it off "
                      "(fgx_set_energy_matrix) for TranslateRotate replicas");
        for (fgx_system *s : tempers[k]->cfg) {
            for (fgx_system *o : sys)
                FGX_CHECK(o != s, FGX_EARG, "fgx_temper_run: a configuration is in two tempers");
            sys.push_back(s);
        }
    }
    fgx_system *s0 = sys[0];
    hipStream_t st0 = s0->stream;
    for (fgx_system *s : sys) FGX_HIP(hipStreamSynchronize(s->stream));
    int rc;
    for (int k = 0; k < nt; ++k) {
        fgx_temper *t = tempers[k];
        if ((rc = push_state(t))) return rc;
        if ((rc = ensure_records(t, rounds))) return rc;
    }
    BatchPlan B;
    std::vector<int> nv(sys.size(), calls);
    if ((rc = prepare_batch(sys, nv, nullptr, nullptr, nullptr, 0, B))) return rc;
    // each temper's view of the prepared chains and their statistics
    size_t off = 0;
    for (int k = 0; k < nt; ++k) {
        fgx_temper *t = tempers[k];
        for (int c = 0; c < t->R; ++c) {
            t->h_sys[c].chain = B.d_chains + off + c;
            t->h_sys[c].bstat = (int)(off + c);
        }
        FGX_HIP(hipMemcpyAsync(t->d_sys, t->h_sys.data(), sizeof(PTSys) * t->R,
                               hipMemcpyHostToDevice, st0));
        t->st = st0;
        if ((rc = reset_counters(t))) return rc;
        off += t->R;
    }
    FGX_HIP(hipStreamSynchronize(st0));
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    FGX_HIP(hipStreamBeginCapture(st0, hipStreamCaptureModeThreadLocal));
    rc = enqueue_batch(B, st0);
    for (int k = 0; k < nt && !rc; ++k)
        rc = enqueue_exchange(tempers[k], s0->d_bstat_i, s0->d_bstat_d);
    hipError_t ce = hipStreamEndCapture(st0, &graph);
    if (!rc && ce != hipSuccess) {
        set_error(std::string("fgx_temper_run: graph capture failed: ") + hipGetErrorString(ce));
        rc = FGX_EHIP;
    }
    if (!rc && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        set_error("fgx_temper_run: hipGraphInstantiate failed");
        rc = FGX_EHIP;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc) {
        FGX_HIP(hipEventCreate(&e0));
        FGX_HIP(hipEventCreate(&e1));
        FGX_HIP(hipEventRecord(e0, st0));
        for (int i = 0; i < rounds && !rc; ++i)
            if (hipGraphLaunch(exec, st0) != hipSuccess) {
                set_error("fgx_temper_run: hipGraphLaunch failed");
                rc = FGX_EHIP;
            }
        FGX_HIP(hipEventRecord(e1, st0));
        hipError_t se = hipStreamSynchronize(st0);
        if (!rc && se != hipSuccess) {
            set_error(std::string("fgx_temper_run: ") + hipGetErrorString(se));
            rc = FGX_EHIP;
        }
        if (!rc && B.blk) rc = blk_check(sys, B.d_xerr);
    }
    float ms = 0.f;
    if (!rc) (void)hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (seconds) *seconds = ms * 1e-3;
    // the chains no longer belong to a prepared launch
    for (int k = 0; k < nt; ++k) {
        fgx_temper *t = tempers[k];
        for (int c = 0; c < t->R; ++c) {
            t->h_sys[c].chain = nullptr;
            t->h_sys[c].bstat = c;
        }
        FGX_HIP(hipMemcpyAsync(t->d_sys, t->h_sys.data(), sizeof(PTSys) * t->R,
                               hipMemcpyHostToDevice, st0));
    }
    if (rc) return rc;
    if ((rc = mirror_trial(sys, st0))) return rc;
    if (st) *st = fgx_stats{};
    for (int k = 0; k < nt && !rc; ++k) {
        fgx_temper *t = tempers[k];
        if ((rc = pull_records(t, rounds))) break;
        // the last round's exchange (no Ewald replicas here: no update term)
        t->last_ret.assign(t->R, 0.0);
        if (rounds > 0)
            for (int r = 0; r < t->R; ++r)
                if (t->rec_acc[(size_t)(rounds - 1) * t->R + r])
                    t->last_ret[r] = t->rec_ret[(size_t)(rounds - 1) * t->R + r];
        std::vector<long> ti(8 * t->R);
        std::vector<double> td(8 * t->R);
        FGX_HIP(hipMemcpy(ti.data(), t->d_tot_i, sizeof(long) * ti.size(), hipMemcpyDeviceToHost));
        FGX_HIP(hipMemcpy(td.data(), t->d_tot_d, sizeof(double) * td.size(), hipMemcpyDeviceToHost));
        std::vector<long> cyc((size_t)rounds * t->R);
        FGX_HIP(hipMemcpy(cyc.data(), t->d_rec_cycles, sizeof(long) * cyc.size(), hipMemcpyDeviceToHost));
        t->rec_cycles = cyc;
        for (int c = 0; c < t->R; ++c) collect_stats(t->cfg[c], &ti[8 * c], &td[8 * c], st, ms * 1e-3);
        t->st = t->cfg[0]->stream;
    }
    return rc;
}

// records of the last fgx_temper_exchange / fgx_temper_run call, rounds x R:
// acc and du per rank, slot[r] after each round, and each configuration's
// sweep time per round (seconds, fgx_temper_run only)
extern "C" int fgx_temper_records(fgx_temper *t, int *rounds, int *acc, double *du, int *slot,
                                  double *sweep_seconds)
{
    FGX_CHECK(t && rounds, FGX_EARG, "fgx_temper_records: bad arguments");
    *rounds = t->last_rounds;
    const size_t n = (size_t)t->last_rounds * t->R;
    for (size_t k = 0; k < n; ++k) {
        if (acc) acc[k] = t->rec_acc[k];
        if (du) du[k] = t->rec_du[k];
        if (slot) slot[k] = t->rec_slot[k];
        if (sweep_seconds) {
            const fgx_system *s = t->cfg[k % t->R];
            sweep_seconds[k] = (s->wall_khz > 0 && k < t->rec_cycles.size())
                                   ? (double)t->rec_cycles[k] / (1e3 * s->wall_khz)
                                   : 0.0;
        }
    }
    return FGX_OK;
}

extern "C" int fgx_temper_slots(fgx_temper *t, int *slot)
{
    FGX_CHECK(t && slot, FGX_EARG, "fgx_temper_slots: bad arguments");
    for (int r = 0; r < t->R; ++r) slot[r] = t->slot[r];
    return FGX_OK;
}

extern "C" int fgx_temper_destroy(fgx_temper *t)
{
    if (!t) return FGX_OK;
    temper_free(t);
    delete t;
    return FGX_OK;
}
