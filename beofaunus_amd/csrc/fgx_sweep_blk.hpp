// fgx_sweep_blk.hpp -- block AtomicTranslation sweep: the device-resident
// Propagator::move (move.h:7877-7891) over AtomicTranslation trials
// (move.h:818-901) evaluated B at a time.
//
// The random numbers of a trial (group, particle, three displacement
// canonicals, the Metropolis uniform; move.h:825, group.h:132, move.h:843-845,
// move.h:570) do not depend on earlier decisions, so B consecutive trials of a
// Propagator call are drawn ahead and their energy changes evaluated against
// one snapshot of the accepted configuration in a single pass over the
// partners: every partner record is loaded once for B trials.  The trials are
// then decided in the reference's order with exact corrections for what the
// earlier trials of the block changed:
//
//   dU_t = P_t + sum_{j in block, j != i_t} [u(n_t, x_j) - u(o_t, x_j)]
//               + 2 (2 pi / V) lB sum_{s < t accepted} sum_k A_k Re(dS_s* dS_t)
//               + surface term with the current dipole,
//
// where P_t sums the partners outside the block at their snapshot positions
// and x_j is the block particle's position just before trial t (its new
// position if an earlier trial of the block moved it and was accepted).  A
// trial whose particle an earlier accepted trial of the block moved, and the
// trials after an accepted move that triggers the Ewald S(k) refresh
// (ewald.h:626-643), are left for the next block, with their draws.  The
// decisions are the sequential ones; energy changes differ from a one-trial
// sweep in rounding only.
//
// A chain may span several workgroups ("parts"): part p owns the partner rows
// [p n / P, (p + 1) n / P) and the k-vectors [p K / P, (p + 1) K / P).  Every
// part replays the same RNG streams and takes the same decisions from the
// same totals; per block the parts exchange their partial sums through
// device memory (one counter barrier per block, bounded spin: a part that
// waits too long sets the launch's error word and the chain stops).  All
// parts write every accepted move, so each part reads its own writes.
#pragma once

#include "fgx_sweep_kernels.hpp"

namespace fgx {

static constexpr int BLK_NSPLIT_MAX = 32;    // workgroups per chain
static constexpr int BLK_NV_MAX = 40;        // partial sums exchanged per block
static constexpr int BLK_THREADS = 512;
// partners in device memory (mixed path): trials per pass over the partners
// and records per thread and step (the next step's records in flight)
static constexpr int BLK_CB_PAIR = 2, BLK_U_PAIR = 2;
static constexpr int BLK_CB_EWALD = 2, BLK_U_EWALD = 2;
static constexpr int BLK_LDS_ROWS = 2048;    // partner rows of an LDS-resident slice, at most
static constexpr long BLK_SPIN_MAX = 1L << 22;   // s_sleep rounds of a part waiting for the others

// one trial of a block: its draws, the moved particle's data, the proposal
struct BlkTrial {
    int i, ti, collide, first;
    int fetched;                            // particle data read since the draw
    int32_t fnx, fny, fnz, fox, foy, foz;   // fixed-point new / old position
    int pkw;                                // FP32 charge bits of the packed record
    float rf;                               // FP32 radius
    double q64, r64, wq, u, lu;   // lu = -ln u
    double t3[3];
    double ox, oy, oz, nx, ny, nz;
};

// byte offsets of a block-sweep workgroup's dynamic LDS: Ewald S(k) slice
// (double2), A_k (double), packed n (int); per-trial phase tables
// [B][old, new][3 kcc1] (double2); a bit per partner row of the slice (the
// block's own particles, excluded from the pass); the LDS-resident partner
// slice (LdsPart, ns rows; 0 when the partners stay in device memory)
struct BlkLds {
    size_t s, a, kid, etab, excl, part, total;
};
__host__ __device__ inline int blk_excl_words(int rows) { return (rows + 31) / 32 + 2; }
__host__ __device__ inline BlkLds blk_lds(int ks, int B, int kcc1, int ew, int rows, int ns)
{
    BlkLds o;
    o.s = 0;
    o.a = (size_t)16 * ks;
    o.kid = (size_t)24 * ks;
    o.etab = ((size_t)28 * ks + 15) / 16 * 16;
    o.excl = o.etab + (ew ? (size_t)B * 2 * 3 * kcc1 * 16 : 0);
    o.part = (o.excl + 4 * (size_t)blk_excl_words(rows) + 15) / 16 * 16;
    o.total = o.part + (size_t)ns * LDS_PART_BYTES;
    return o;
}

// trial data a pass over the partners needs, per trial of the block
struct BlkFx {
    int32_t nx, ny, nz, ox, oy, oz;
    float aq, ar;
    int ti;
};

// lane l's double, in every lane
__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                            (unsigned long long)(unsigned)lo));
}

// sum over the four lanes of a quad in a fixed order; every lane of the
// quad receives the same bits (IEEE addition commutes)
__device__ __forceinline__ double quad_sum(double v)
{
    v += dpp_f64<0xb1>(v);          // quad_perm 1,0,3,2
    return v + dpp_f64<0x4e>(v);    // quad_perm 2,3,0,1
}

__device__ __forceinline__ int quad_isum(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0xb1, 0xf, 0xf, false);
    return v + __builtin_amdgcn_update_dpp(0, v, 0x4e, 0xf, 0xf, false);
}

// bit 4t of a wave mask -> bit t (t < 16)
__device__ __forceinline__ unsigned quad_bits(uint64_t x)
{
    x &= 0x1111111111111111ull;
    x = (x | (x >> 3)) & 0x0303030303030303ull;
    x = (x | (x >> 6)) & 0x000f000f000f000full;
    x = (x | (x >> 12)) & 0x000000ff000000ffull;
    x = (x | (x >> 24)) & 0xffffull;
    return (unsigned)x;
}

// uniform_int_distribution<int>(0, erange - 1) from one word (Lemire's
// method, as mt_range); ok = false when the method needs another word
__device__ __forceinline__ uint32_t lemire_word(uint32_t word, uint32_t erange, bool &ok)
{
    const uint64_t prod = (uint64_t)word * (uint64_t)erange;
    const uint32_t low = (uint32_t)prod;
    if (low < erange) {
        const uint32_t thr = (uint32_t)(0u - erange) % erange;
        if (low < thr) ok = false;
    }
    return (uint32_t)(prod >> 32);
}

#pragma clang fp contract(off)
// generate_canonical<double,53> from its two words (as mt_canonical)
__device__ __forceinline__ double canonical_words(uint32_t w1, uint32_t w2)
{
    double sum = (double)w1;
    sum += (double)w2 * 4294967296.0;
    double r = sum / 18446744073709551616.0;
    if (r >= 1.0) r = 0x1.fffffffffffffp-1;
    return r;
}
#pragma clang fp contract(on)

// constants of the fixed-unit pair form, wave-uniform
struct BlkFxConst {
    float rc2, s2, e4, rs, c1, exact;
    int ntypes, ntypes2;
};

// Fixed-unit pair term of one trial against one partner record (mixed path):
// returns u(new) - u(old) in FP32; `bad` marks a pair whose value must come
// from FP64 (inside the Ewald rc, strong, non-finite or near hard-sphere
// contact).  Trial: fixed-point new / old position, type, charge lB q / s and
// radius r c / s; partner: packed record, FP32 radius, type.  MODE (Ewald):
// 0 type tables, 1 one LJ parameter pair, 2 nothing beyond rc.
template <int ELEC, int SR, int MODE>
__device__ __forceinline__ float blk_pair_fx(const BlkFxConst &K, const float *tabf, int32_t tnx,
                                             int32_t tny, int32_t tnz, int32_t tox, int32_t toy,
                                             int32_t toz, int tti, float taq, float tar, int4 P,
                                             float Rb, int Tb, bool &bad)
{
    float dx = (float)(int32_t)(tnx - P.x), dy = (float)(int32_t)(tny - P.y),
          dz = (float)(int32_t)(tnz - P.z);
    const float r2n = dx * dx + dy * dy + dz * dz;
    dx = (float)(int32_t)(tox - P.x);
    dy = (float)(int32_t)(toy - P.y);
    dz = (float)(int32_t)(toz - P.z);
    const float r2o = dx * dx + dy * dy + dz * dz;
    if constexpr (ELEC == ELEC_EWALD) {
        if constexpr (MODE == 2) {
            bad = fminf(r2n, r2o) <= K.rc2;
            return 0.f;
        } else {
            float s2 = K.s2, e4 = K.e4;
            if constexpr (MODE == 0) {
                const int k = tti * K.ntypes + Tb;
                s2 = tabf[TAB_S2 * K.ntypes2 + k];
                e4 = tabf[TAB_EPS4 * K.ntypes2 + k];
            }
            float xn = s2 * __builtin_amdgcn_rcpf(r2n);
            float xo = s2 * __builtin_amdgcn_rcpf(r2o);
            xn = xn * xn * xn;
            xo = xo * xo * xo;
            const float vn = xn * xn - xn, vo = xo * xo - xo;
            bad = fminf(r2n, r2o) <= K.rc2;
            if constexpr (MODE == 0)
                bad = bad || !(fabsf(e4 * vn) <= K.exact) || !(fabsf(e4 * vo) <= K.exact);
            return e4 * (vn - vo);
        }
    } else {
        // Coulomb / Debye-Hueckel (+ hard sphere)
        const float bq = __int_as_float(P.w);
        float un = 0.f, uo = 0.f;
        const float rin = __builtin_amdgcn_rsqf(r2n);
        const float rio = __builtin_amdgcn_rsqf(r2o);
        if constexpr (ELEC == ELEC_COULOMB) {
            un = taq * bq * rin;
            uo = taq * bq * rio;
        } else if constexpr (ELEC == ELEC_DH) {
            un = taq * bq * rin * __builtin_amdgcn_exp2f(K.c1 * (r2n * rin));
            uo = taq * bq * rio * __builtin_amdgcn_exp2f(K.c1 * (r2o * rio));
        }
        bad = !(fabsf(un) <= K.exact) || !(fabsf(uo) <= K.exact);
        if constexpr (SR == SR_HS) {
            const float m = tar + Rb * K.rs;
            bad = bad || (r2n < m * m) || (r2o < m * m);
        }
        return un - uo;
    }
}

template <int ELEC, int SR, bool MIX, int B>
__global__ __launch_bounds__(BLK_THREADS) void k_sweep_blk(const ChainDev *__restrict__ chains,
                                                           const int2 *__restrict__ wgmap)
{
    constexpr bool EW = (ELEC == ELEC_EWALD);
    constexpr int NG = EW ? B * (B - 1) / 2 : 0;
    constexpr int NV = B + NG;
    static_assert(NV <= BLK_NV_MAX, "block too large");
    static_assert(B <= 16, "the decisions hold up to four lanes per trial");
    constexpr bool NEED_T = (SR == SR_LJ || ELEC == ELEC_EWALD);
    // trials per pass over the partners (mixed path)
    constexpr int CB = EW ? BLK_CB_EWALD : BLK_CB_PAIR;
    const int2 wm = wgmap[blockIdx.x];
    if (wm.x < 0) return;
    const ChainDev &C = chains[wm.x];
    const int part = wm.y, nsplit = C.nsplit;
    const int tid = threadIdx.x, nthr = blockDim.x, lane = tid & 63, wv = tid >> 6;
    const int nw = nthr >> 6;
    const bool wave0 = tid < 64;

    __shared__ uint32_t mtg[624], mti[624];
    // the ring: the block's trials and up to B more drawn ahead (the draw
    // sequence does not depend on the decisions)
    __shared__ BlkTrial R[2 * B];
    __shared__ int sgp;   // position of the trial generator between the drawing waves
    __shared__ int bidx[B];
    __shared__ BlkFx F[B];
    __shared__ double red[BLK_THREADS / 64][NV];
    __shared__ double redp[BLK_THREADS / 64][B];   // partner-pass sums per wave and trial
    __shared__ double tot[BLK_NV_MAX];
    __shared__ double xc[B][B], yc[B][B];
    __shared__ int dlist[MIX ? DCAP : 1];
    __shared__ int nqs[BLK_THREADS / 64];
    __shared__ double ltab[TAB_COUNT * LDS_TYPES2];
    __shared__ float ltabf[MIX ? TAB_COUNT * LDS_TYPES2 : 1];
    __shared__ int lgvec[AT_LDS_GROUPS];
    __shared__ GroupDesc lgroups[AT_LDS_GROUPS];
    __shared__ double ltdp[AT_LDS_TYPES], leff[AT_LDS_TYPES];
    __shared__ double lcheb[EW && MIX ? EW_CHEB_MAX : 1];
    __shared__ double2 rtab[EW ? REFRESH_CHUNK * 3 * REFRESH_KAX : 1];
    __shared__ double rw[EW ? REFRESH_CHUNK : 1];
    // accepted moves of the last block (index, new position), for the fetch
    __shared__ int lmv_i[B];
    __shared__ double lmv_x[B], lmv_y[B], lmv_z[B];
    __shared__ int sgo, sabort, snlmv;
    __shared__ int snd, saccm, srefresh;   // wave 0's decision of the block
    __shared__ double sdu[B], swq[B];
    extern __shared__ double2 dyn[];

    // ---- this part's slices
    const int j0 = (int)((long)part * C.n / nsplit);
    const int j1 = (int)((long)(part + 1) * C.n / nsplit);
    const int K = EW ? C.K : 0;
    const int k0 = (int)((long)part * K / nsplit);
    const int k1 = (int)((long)(part + 1) * K / nsplit);
    const int ks = k1 - k0;
    const int kcc1 = C.kcc + 1;
    const bool ldsp = C.ldsp != 0;
    const BlkLds lo = blk_lds(ks, B, kcc1, EW ? 1 : 0, j1 - j0, ldsp ? j1 - j0 : 0);
    char *dbase = (char *)dyn;
    double2 *S = (double2 *)(dbase + lo.s);
    double *Ak = (double *)(dbase + lo.a);
    int *kid = (int *)(dbase + lo.kid);
    double2 *etab = (double2 *)(dbase + lo.etab);
    uint32_t *exclb = (uint32_t *)(dbase + lo.excl);   // block particles in the slice
    LdsPart L;
    L.j0 = j0;
    L.ns = ldsp ? j1 - j0 : 0;
    L.carve((FGX_LDS void *)(dbase + lo.part), L.ns);

    // ---- chain barrier across the parts.  Wave 0 alone writes what the
    // parts exchange, so one release fence of its thread 0 publishes it; the
    // spin reads the counter relaxed and one acquire fence follows.
    int epoch = 0;          // barriers passed (uniform)
    int xpar = 0;           // exchange buffer in use (alternates per exchange)
    auto chain_sync = [&]() {
        ++epoch;
        if (tid == 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_fetch_add(C.xcnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int target = epoch * nsplit;
            long spins = 0;
            while (__hip_atomic_load(C.xcnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > BLK_SPIN_MAX || __hip_atomic_load(C.xerr, __ATOMIC_RELAXED,
                                                                __HIP_MEMORY_SCOPE_AGENT)) {
                    __hip_atomic_store(C.xerr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    sabort = 1;
                    break;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
    };
    // tot[0..nv) of this part -> the chain's totals (parts summed in a fixed
    // order); tot is written by wave 0 (tid < nv) before the exchange
    auto exchange = [&](int nv) {
        if (nsplit == 1) return;
        double *xb = C.xbuf + (size_t)xpar * BLK_NSPLIT_MAX * BLK_NV_MAX;
        if (tid < nv)
            __hip_atomic_store(&xb[part * BLK_NV_MAX + tid], tot[tid], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        chain_sync();
        // value v summed by one wavefront, lane p loading part p's: every
        // load in flight at once, the parts added in a fixed (DPP) order
        for (int v = tid >> 6; v < nv; v += nthr >> 6) {
            const int ln = tid & 63;
            const double x = ln < nsplit ? __hip_atomic_load(&xb[ln * BLK_NV_MAX + v],
                                                             __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT)
                                         : 0.0;
            const double sv = wave_sum_dpp(x);
            if (ln == 0) tot[v] = sv;
        }
        __syncthreads();
        xpar ^= 1;
    };

    // ---- chain state into LDS
    if (tid == 0) {
        sabort = 0;
        snlmv = 0;
    }
    for (int k = tid; k < blk_excl_words(j1 - j0); k += nthr) exclb[k] = 0u;
    if (tid < B) bidx[tid] = -1;
    for (int k = tid; k < 624; k += nthr) {
        mtg[k] = C.rng[k];
        mti[k] = C.rng[625 + k];
    }
    WaveMT G{mtg, (int)C.rng[624]}, I{mti, (int)C.rng[625 + 624]};
    if (tid == 0) sgp = (int)C.rng[624];
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
    // type tables (launched only with ntypes^2 <= LDS_TYPES2)
    for (int k = tid; k < TAB_COUNT * C.pot.ntypes2; k += nthr) {
        const int t = k / C.pot.ntypes2;
        ltab[k] = C.tab[k];
        if (MIX) ltabf[k] = C.tabf[k] * ((EW && t == TAB_S2) ? C.fx_s2scale : 1.f);
    }
    const double *tab = ltab;
    const float *tabf = ltabf;
    if (ldsp) {
        for (int k = tid; k < L.ns; k += nthr) {
            const int j = j0 + k;
            L.pk[k] = int4_to_v4i(gload(C.pk, j));
            L.x[k] = C.x[j]; L.y[k] = C.y[j]; L.z[k] = C.z[j];
            L.q[k] = gload(C.q, j);
            L.r[k] = gload(C.r, j);
            L.rf[k] = gload(C.rf, j);
            L.t[k] = gload(C.t, j);
        }
    }
    if constexpr (EW) {
        for (int k = tid; k < ks; k += nthr) {
            const int kg = k0 + k;
            S[k] = make_double2(C.Q[2 * kg], C.Q[2 * kg + 1]);
            Ak[k] = C.aks[kg];
            kid[k] = kpack(C.kn[3 * kg], C.kn[3 * kg + 1], C.kn[3 * kg + 2]);
        }
    }
    // surface dipole M = sum q eff r over all particles (getSurfaceEnergy,
    // ewald.h:278-292); every part sums it alike
    double Mx = 0.0, My = 0.0, Mz = 0.0;
    __shared__ double redm[BLK_THREADS / 64];
    if (EW && C.surf_pref != 0.0) {
        for (int j = tid; j < C.n; j += nthr) {
            const double w = C.q[j] * C.eff[C.t[j]];
            Mx += w * C.x[j]; My += w * C.y[j]; Mz += w * C.z[j];
        }
        Mx = block_sum(Mx, redm);
        My = block_sum(My, redm);
        Mz = block_sum(Mz, redm);
    }
    int ew_cnt = EW ? *C.ew_cnt : 0;
    double utot = 0.0, dusum = 0.0;
    long n_trials = 0, n_acc = 0, n_inf = 0, n_ref = 0;
    const long long t_start = wall_clock64();
    long rec = 0;
    __syncthreads();
    // every part has read the chain state before any part writes it back
    if (nsplit > 1) chain_sync();

    const double dirx = C.mv.dir[0], diry = C.mv.dir[1], dirz = C.mv.dir[2];
    const long budget = C.max_trials > 0 ? C.max_trials : LONG_MAX;
    long used = 0;
    int calls_done = 0;
    bool stop = false;
    int resume_rep = (C.resume && C.resume[0]) ? C.resume[1] : -1;
    int nr = 0;   // trials in the ring (uniform)
    const int qcap = DCAP / nw;
    // fixed-unit pair constants, kept in registers across the partner loop
    const BlkFxConst KF{C.fx_rc2, C.lj_s2f, C.lj_eps4f, C.fx_rs, C.fx_c1, C.pot.exact_f,
                        C.pot.ntypes, C.pot.ntypes2};
    const float fx_rs = C.fx_rs;
    // waves per trial of the LDS-resident partner pass
    const int wpt = nw >= B ? nw / B : 1;
    const volatile BlkTrial *Rv = R;

    // draws of `cnt` trials into ring slots at.. (one wave: wave 0 at a
    // block's start, wave 1 during the decisions): a trial takes ten words
    // (two uniform_int, four canonicals) unless a uniform_int rejects; lane
    // l draws trial l from its ten words of the current state, up to the
    // first rejection, and the rest are drawn one by one
    auto draw = [&](int at, int cnt) {
        G.p = sgp;
        G.wb = -1;
        while (cnt > 0) {
            // trials whose ten words lie in the current state: in parallel
            const int k = min(cnt, (624 - G.p) / 10);
            int done = 0;
            if (k > 0) {
                bool okl = true;
                int pi = 0;
                double a = 0.0, b = 0.0, c = 0.0, u = 0.0;
                if (lane < k) {
                    const int p = G.p + 10 * lane;
                    uint32_t w[10];
#pragma unroll
                    for (int q = 0; q < 10; ++q) w[q] = mt_temper(mtg[p + q]);
                    const int gk = (int)lemire_word(w[0], (uint32_t)C.ngvec, okl);
                    const int g = okl ? (lds_grp ? lgvec[gk] : C.gvec[gk]) : 0;
                    const GroupDesc gd = lds_grp ? lgroups[g] : C.groups[g];
                    pi = gd.front + (int)lemire_word(w[1], (uint32_t)(gd.back - gd.front) + 1u, okl);
                    a = canonical_words(w[2], w[3]);
                    b = canonical_words(w[4], w[5]);
                    c = canonical_words(w[6], w[7]);
                    u = canonical_words(w[8], w[9]);
                }
                const uint64_t bad = __ballot(lane < k && !okl);
                done = bad ? __ffsll((unsigned long long)bad) - 1 : k;
                if (lane < done) {
                    BlkTrial &t = R[at + lane];
                    t.i = pi;
                    t.t3[0] = a; t.t3[1] = b; t.t3[2] = c;
                    t.u = u;
                    t.fetched = 0;
                }
                G.p += 10 * done;
                G.wb = -1;
                at += done;
                cnt -= done;
            }
            if (cnt > 0) {
                // one trial word by word: the state twists within its words,
                // or one of its uniform_int draws needs a second word
                const int gk = mt_range(G, 0, C.ngvec - 1);
                const int g = lds_grp ? lgvec[gk] : C.gvec[gk];
                const GroupDesc gd = lds_grp ? lgroups[g] : C.groups[g];
                const int i = gd.front + mt_range(G, 0, gd.back - gd.front);
                const double a = mt_canonical(G), b = mt_canonical(G), c = mt_canonical(G);
                const double u = mt_canonical(G);
                if (lane == 0) {
                    BlkTrial &t = R[at];
                    t.i = i;
                    t.t3[0] = a; t.t3[1] = b; t.t3[2] = c;
                    t.u = u;
                    t.fetched = 0;
                }
                ++at;
                --cnt;
            }
        }
        wave_sync_lds();
        if (lane == 0) sgp = G.p;
    };
    // the moved particle's data from the accepted vector (one lane)
    auto fetch_particle = [&](BlkTrial &t) {
        const int i = t.i;
        t.ox = gload(C.x, i); t.oy = gload(C.y, i); t.oz = gload(C.z, i);
        const int ti = gload(C.t, i);
        t.ti = ti;
        t.q64 = gload(C.q, i);
        t.r64 = gload(C.r, i);
        t.rf = gload(C.rf, i);
        t.pkw = gload(C.pk, i).w;
        t.wq = EW ? t.q64 * (lds_grp ? leff[ti] : C.eff[ti]) : 0.0;
        t.fetched = 1;
    };

    for (int call = 0; call < C.ncalls && !stop && !sabort; ++call) {
        int rep0 = 0;
        bool go;
        if (resume_rep >= 0) {
            go = true;
            rep0 = resume_rep;
            resume_rep = -1;
        } else {
            if (wave0) {
                // ---- Propagator::move (move.h:7887) + Movebase::move prologue
                (void)mt_range(I, 0, C.nmoves - 1);
                (void)mt_range(I, 0, 0);
                const bool run = mt_canonical(I) < C.mv.prob;
                if (tid == 0) sgo = run ? 1 : 0;
            }
            __syncthreads();
            go = sgo != 0;
            __syncthreads();
        }
        if (!go) { ++calls_done; continue; }
        int drawn = rep0;
        nr = 0;
        for (;;) {
            // ---- fill the ring with the next trials of this call (wave 0;
            // usually drawn ahead already)
            const int nfill = nr >= B ? 0 : (int)min((long)min(B - nr, C.repeat - drawn), budget - used);
            if (nr == 0 && nfill == 0) {
                if (drawn < C.repeat) {   // trial budget of the launch spent
                    stop = true;
                    if (tid == 0 && part == 0 && C.resume) { C.resume[0] = 1; C.resume[1] = drawn; }
                }
                break;
            }
            if (wave0 && nfill > 0) draw(nr, nfill);
            nr += nfill;
            used += nfill;
            drawn += nfill;
            const int nb = min(nr, B);
            if (wave0) {
                wave_sync_lds();
                // ---- the moved particles' data (read once per draw),
                // corrected for the moves the last block accepted
                {
                    const int nl = snlmv;
                    const int li = lane < nl ? lmv_i[lane] : -1;
                    if (lane < nr) {
                        BlkTrial &t = R[lane];
                        if (!t.fetched) fetch_particle(t);
                        int hit = -1;
#pragma unroll
                        for (int m = 0; m < B; ++m)
                            if (m < nl && __builtin_amdgcn_readlane(li, m) == t.i) hit = m;
                        if (hit >= 0) { t.ox = lmv_x[hit]; t.oy = lmv_y[hit]; t.oz = lmv_z[hit]; }
                    }
                }
                wave_sync_lds();
                // the last block's particles leave the slice's exclusion bits
                if (lane < B) {
                    const int io = bidx[lane];
                    if (io >= j0 && io < j1) exclb[(io - j0) >> 5] = 0u;
                }
                wave_sync_lds();
                // ---- proposals (one lane per trial of the block)
                if (lane < B) {
                    if (lane < nb) {
                        BlkTrial &t = R[lane];
                        const int i = t.i, ti = t.ti;
                        const double ox = t.ox, oy = t.oy, oz = t.oz;
                        // Metropolis (move.h:567-573) rejects when u > exp(-dU),
                        // i.e. when dU > -ln u
                        t.lu = -log(t.u);
                        const double dp = lds_grp ? ltdp[ti] : C.type_dp[ti];
                        double nx, ny, nz;
                        {
#pragma clang fp contract(off)
                            double tx = dirx * dp, ty = diry * dp, tz = dirz * dp;
                            tx *= t.t3[0] - 0.5;
                            ty *= t.t3[1] - 0.5;
                            tz *= t.t3[2] - 0.5;
                            nx = ox + tx; ny = oy + ty; nz = oz + tz;
                            boundary_d(C.box, nx, ny, nz);
                        }
                        t.nx = nx; t.ny = ny; t.nz = nz;
                        t.collide = collision_d(C.box, nx, ny, nz) ? 1 : 0;
                        t.fox = to_fixed(ox, C.box.inv[0]);
                        t.foy = to_fixed(oy, C.box.inv[1]);
                        t.foz = to_fixed(oz, C.box.inv[2]);
                        t.fnx = to_fixed(nx, C.box.inv[0]);
                        t.fny = to_fixed(ny, C.box.inv[1]);
                        t.fnz = to_fixed(nz, C.box.inv[2]);
                        int first = lane;
#pragma unroll
                        for (int s2 = B - 1; s2 >= 0; --s2)
                            if (s2 < lane && __builtin_amdgcn_readlane(i, s2) == i) first = s2;
                        t.first = first;
                        bidx[lane] = i;
                        if (i >= j0 && i < j1)
                            atomicOr(&exclb[(i - j0) >> 5], 1u << ((i - j0) & 31));
                        BlkFx f;
                        f.nx = t.fnx; f.ny = t.fny; f.nz = t.fnz;
                        f.ox = t.fox; f.oy = t.foy; f.oz = t.foz;
                        const float aq = __int_as_float(t.pkw);
                        f.aq = C.fx ? aq * C.fx_lbs : aq;
                        f.ar = C.fx ? t.rf * fx_rs : t.rf;
                        f.ti = ti;
                        F[lane] = f;
                    } else {
                        bidx[lane] = -1;
                    }
                }
            }
            __syncthreads();
            // ---- Ewald phase tables exp(i 2 pi n x / L), old and new, per trial
            if constexpr (EW) {
                const int per = 3 * kcc1, ntab = nb * 2 * per;
                for (int e = tid; e < ntab; e += nthr) {
                    const int b = e / (2 * per), rem = e % (2 * per);
                    const int which = rem / per, r2 = rem % per;
                    const int axis = r2 / kcc1, nn = r2 % kcc1;
                    const BlkTrial &t = R[b];
                    const double coord = which ? (axis == 0 ? t.nx : axis == 1 ? t.ny : t.nz)
                                               : (axis == 0 ? t.ox : axis == 1 ? t.oy : t.oz);
                    double sn, cs;
                    sincospi(2.0 * (double)nn * coord * C.box.inv[axis], &sn, &cs);
                    etab[(b * 2 + which) * per + r2] = make_double2(cs, sn);
                }
            }
            // ---- pair part over this part's partners outside the block, CB
            // trials per pass over the partners (a trial's data is read from
            // LDS per partner, not held across the loop); each chunk's sums
            // go to redp at once
            double acc[B];
#pragma unroll
            for (int b = 0; b < B; ++b) acc[b] = 0.0;
            // a partner row of the slice that is one of the block's particles
            auto excluded = [&](int j) { return (exclb[(j - j0) >> 5] >> ((j - j0) & 31)) & 1u; };
            if constexpr (MIX) {
                const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
                int *dq = dlist + wv * qcap;
                int cnt = 0;
                // mode 1: one LJ parameter pair and no strong-pair test beyond
                // rc; 2: nothing beyond rc (CutShift LJ at rc)
                const int mode = (EW && C.lj_screen) ? 2
                                 : (EW && C.lj_uniform && !C.lj_check) ? 1 : 0;
                auto pass = [&](auto M, auto UU, int bc) __attribute__((always_inline)) {
                    constexpr int MODE = decltype(M)::value;
                    constexpr int U = decltype(UU)::value;   // records per thread per step
                    // this chunk's trial constants, wave-uniform (scalar registers)
                    int32_t tnx[CB], tny[CB], tnz[CB], tox[CB], toy[CB], toz[CB], tti[CB];
                    float taq[CB], tar[CB];
#pragma unroll
                    for (int c = 0; c < CB; ++c) {
                        const int b = min(bc + c, B - 1);
                        tnx[c] = __builtin_amdgcn_readfirstlane(F[b].nx);
                        tny[c] = __builtin_amdgcn_readfirstlane(F[b].ny);
                        tnz[c] = __builtin_amdgcn_readfirstlane(F[b].nz);
                        tox[c] = __builtin_amdgcn_readfirstlane(F[b].ox);
                        toy[c] = __builtin_amdgcn_readfirstlane(F[b].oy);
                        toz[c] = __builtin_amdgcn_readfirstlane(F[b].oz);
                        tti[c] = __builtin_amdgcn_readfirstlane(F[b].ti);
                        taq[c] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(F[b].aq)));
                        tar[c] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(F[b].ar)));
                    }
                    double ac[CB];
#pragma unroll
                    for (int c = 0; c < CB; ++c) ac[c] = 0.0;
                    struct Rec {
                        int4 P[U];
                        float R[U];
                        int T[U];
                    };
                    auto fetch = [&](Rec &X, int base) __attribute__((always_inline)) {
#pragma unroll
                        for (int u = 0; u < U; ++u) {
                            const int j = base + u * nthr + tid;
                            // rows past n read the zero padding
                            X.P[u] = gload(C.pk, j);
                            X.R[u] = SR == SR_HS ? gload(C.rf, j) : 0.f;
                            X.T[u] = (NEED_T && MODE == 0) ? gload(C.t, j) : 0;
                        }
                    };
                    // one LJ parameter pair beyond rc (|u| < 1e-2 kT per pair): a
                    // step's terms are summed in FP32 and widened once
                    constexpr bool BATCH32 = EW && MODE == 1;
                    float bs[CB];
#pragma unroll
                    for (int c = 0; c < CB; ++c) bs[c] = 0.f;
                    auto eval = [&](const Rec &X, int base) __attribute__((always_inline)) {
                        unsigned dm = 0;
#pragma unroll
                        for (int u = 0; u < U; ++u) {
                            const int j = base + u * nthr + tid;
                            const bool ok = j < j1 && !excluded(j);
                            const int4 P = X.P[u];
#pragma unroll
                            for (int c = 0; c < CB; ++c) {
                                if (bc + c >= nb) break;
                                bool bad;
                                const float d = blk_pair_fx<ELEC, SR, MODE>(
                                    KF, tabf, tnx[c], tny[c], tnz[c], tox[c], toy[c], toz[c], tti[c],
                                    taq[c], tar[c], P, X.R[u], X.T[u], bad);
                                dm |= ((ok && bad) ? 1u : 0u) << (u * CB + c);
                                const float v = __int_as_float(__float_as_int(d) & ((ok && !bad) ? -1 : 0));
                                if constexpr (BATCH32) bs[c] += v;
                                else ac[c] += (double)v;
                            }
                        }
                        if constexpr (BATCH32) {
#pragma unroll
                            for (int c = 0; c < CB; ++c) { ac[c] += (double)bs[c]; bs[c] = 0.f; }
                        }
                        if (__ballot(dm != 0)) {
#pragma unroll
                            for (int u = 0; u < U; ++u)
#pragma unroll
                                for (int c = 0; c < CB; ++c) {
                                    const bool dd = (dm >> (u * CB + c)) & 1u;
                                    const uint64_t m = __ballot(dd);
                                    if (m) {
                                        if (dd) {
                                            const int j = base + u * nthr + tid;
                                            const int k = cnt + __popcll(m & below);
                                            if (k < qcap) {
                                                dq[k] = ((bc + c) << 24) | j;
                                            } else {
                                                const BlkTrial &t = R[bc + c];
                                                ac[c] += du_pair64_mixed<ELEC, SR>(C, tab, lcheb, t.nx, t.ny,
                                                                                   t.nz, t.ox, t.oy, t.oz,
                                                                                   t.q64, t.r64, t.ti, j);
                                            }
                                        }
                                        cnt += __popcll(m);
                                    }
                                }
                        }
                    };
                    const int step = U * nthr;
                    Rec A, Bq;
                    fetch(A, j0);
                    for (int base = j0; base < j1; base += 2 * step) {
                        fetch(Bq, base + step);   // device memory: within the zero padding
                        eval(A, base);
                        fetch(A, base + 2 * step);
                        if (base + step < j1) eval(Bq, base + step);
                    }
#pragma unroll
                    for (int c = 0; c < CB; ++c) {
                        const double sv = wave_sum_dpp(ac[c]);
                        if (lane == 0 && bc + c < B) redp[wv][bc + c] = sv;
                    }
                };
                // partners in LDS: wave w takes trial w / wpt over rows
                // (w % wpt) 64 + lane (mod 64 wpt) of the slice, block particles
                // excluded by their bit
                auto lds_pass = [&](auto M) __attribute__((always_inline)) {
                    constexpr int MODE = decltype(M)::value;
                    const int sub = wv % wpt;
                    for (int b = wv / wpt; b < nb; b += nw / wpt) {
                    const int32_t tnx = __builtin_amdgcn_readfirstlane(F[b].nx);
                    const int32_t tny = __builtin_amdgcn_readfirstlane(F[b].ny);
                    const int32_t tnz = __builtin_amdgcn_readfirstlane(F[b].nz);
                    const int32_t tox = __builtin_amdgcn_readfirstlane(F[b].ox);
                    const int32_t toy = __builtin_amdgcn_readfirstlane(F[b].oy);
                    const int32_t toz = __builtin_amdgcn_readfirstlane(F[b].oz);
                    const int tti = __builtin_amdgcn_readfirstlane(F[b].ti);
                    const float taq = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(F[b].aq)));
                    const float tar = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(F[b].ar)));
                    const int ns = j1 - j0;
                    double ac = 0.0;
                    for (int r0 = sub * 64; r0 < ns; r0 += 64 * wpt) {
                        const int r = r0 + lane;
                        const bool in = r < ns;
                        const int4 P = v4i_to_int4(in ? L.pk[r] : lds_v4i{0, 0, 0, 0});
                        const float Rb = (SR == SR_HS && in) ? L.rf[r] : 0.f;
                        const int Tb = (NEED_T && MODE == 0 && in) ? L.t[r] : 0;
                        const bool ok = in && !((exclb[r >> 5] >> (r & 31)) & 1u);
                        bool bad;
                        const float d = blk_pair_fx<ELEC, SR, MODE>(KF, tabf, tnx, tny, tnz, tox, toy, toz,
                                                                    tti, taq, tar, P, Rb, Tb, bad);
                        ac += (double)__int_as_float(__float_as_int(d) & ((ok && !bad) ? -1 : 0));
                        const bool dd = ok && bad;
                        const uint64_t m = __ballot(dd);
                        if (m) {
                            if (dd) {
                                const int k = cnt + __popcll(m & below);
                                if (k < qcap) {
                                    dq[k] = (b << 24) | (j0 + r);
                                } else {
                                    const BlkTrial &t = R[b];
                                    ac += du_pair64_v<ELEC, SR>(C, tab, t.nx, t.ny, t.nz, t.ox, t.oy, t.oz,
                                                                t.q64, t.r64, t.ti, L.x[r], L.y[r], L.z[r],
                                                                L.q[r], L.r[r], L.t[r]);
                                }
                            }
                            cnt += __popcll(m);
                        }
                    }
                    const double sv = wave_sum_dpp(ac);
                    if (lane == 0) redp[sub][b] = sv;
                    }
                };
                if (ldsp) {
                    if constexpr (EW) {
                        if (mode == 2) lds_pass(std::integral_constant<int, 2>{});
                        else if (mode == 1) lds_pass(std::integral_constant<int, 1>{});
                        else lds_pass(std::integral_constant<int, 0>{});
                    } else {
                        lds_pass(std::integral_constant<int, 0>{});
                    }
                } else {
                    // partners in device memory: CB trials per pass, UM records
                    // per thread and step, the next step's in flight
                    constexpr int UM = EW ? BLK_U_EWALD : BLK_U_PAIR;
                    for (int bc = 0; bc < nb; bc += CB) {
                        if constexpr (EW) {
                            if (mode == 2) pass(std::integral_constant<int, 2>{}, std::integral_constant<int, UM>{}, bc);
                            else if (mode == 1) pass(std::integral_constant<int, 1>{}, std::integral_constant<int, UM>{}, bc);
                            else pass(std::integral_constant<int, 0>{}, std::integral_constant<int, UM>{}, bc);
                        } else {
                            pass(std::integral_constant<int, 0>{}, std::integral_constant<int, UM>{}, bc);
                        }
                    }
                }
                if (lane == 0) nqs[wv] = min(cnt, qcap);
            } else {
                if (ldsp) {
                    // partners in LDS: wave w takes trials w / wpt, as above
                    const int sub = wv % wpt;
                    for (int b = wv / wpt; b < nb; b += nw / wpt) {
                        const double tnx = R[b].nx, tny = R[b].ny, tnz = R[b].nz;
                        const double tox = R[b].ox, toy = R[b].oy, toz = R[b].oz;
                        const double tq = R[b].q64, tr = R[b].r64;
                        const int tt = R[b].ti;
                        const int ns = j1 - j0;
                        double ac = 0.0;
                        for (int r = sub * 64 + lane; r < ns; r += 64 * wpt) {
                            if ((exclb[r >> 5] >> (r & 31)) & 1u) continue;
                            ac += du_pair64_v<ELEC, SR>(C, tab, tnx, tny, tnz, tox, toy, toz, tq, tr, tt,
                                                        L.x[r], L.y[r], L.z[r], L.q[r], L.r[r], L.t[r]);
                        }
                        const double sv = wave_sum_dpp(ac);
                        if (lane == 0) redp[sub][b] = sv;
                    }
                } else {
                    // partners in device memory: one trial per pass
                    for (int b = 0; b < nb; ++b) {
                        const double tnx = Rv[b].nx, tny = Rv[b].ny, tnz = Rv[b].nz;
                        const double tox = Rv[b].ox, toy = Rv[b].oy, toz = Rv[b].oz;
                        const double tq = Rv[b].q64, tr = Rv[b].r64;
                        const int tt = Rv[b].ti;
                        double ac = 0.0;
                        for (int j = j0 + tid; j < j1; j += nthr) {
                            if (excluded(j)) continue;
                            ac += du_pair64_v<ELEC, SR>(C, tab, tnx, tny, tnz, tox, toy, toz, tq, tr, tt,
                                                        C.x[j], C.y[j], C.z[j], C.q[j], C.r[j], C.t[j]);
                        }
                        const double sv = wave_sum_dpp(ac);
                        if (lane == 0) redp[wv][b] = sv;
                    }
                }
            }
            // ---- corrections: the block's own particles at their position
            // before (X) and after (Y) an accepted move, over all threads
            for (int e = tid; e < 2 * nb * nb; e += nthr) {
                const int e2 = e % (nb * nb);
                const int s2 = e2 / nb, t = e2 % nb;
                const bool yv = e >= nb * nb;
                const BlkTrial &a = R[t], &o = R[s2];
                double v = 0.0;
                if (s2 != t && a.i != o.i && (!yv || s2 < t)) {
                    // mixed path with Coulomb or Debye-Hueckel (+ hard spheres):
                    // the partner pass's fixed-unit FP32 form, FP64 for the
                    // pairs it marks
                    bool bad = true;
                    if constexpr (MIX && !EW && SR != SR_LJ) {
                        const BlkFx &ft = F[t], &fs = F[s2];
                        const int4 P = make_int4(yv ? fs.nx : fs.ox, yv ? fs.ny : fs.oy,
                                                 yv ? fs.nz : fs.oz, o.pkw);
                        const float d = blk_pair_fx<ELEC, SR, 0>(KF, tabf, ft.nx, ft.ny, ft.nz, ft.ox,
                                                                 ft.oy, ft.oz, ft.ti, ft.aq, ft.ar, P,
                                                                 o.rf, o.ti, bad);
                        v = (double)d;
                    }
                    if (bad)
                        v = du_pair64_v<ELEC, SR>(C, tab, a.nx, a.ny, a.nz, a.ox, a.oy, a.oz, a.q64,
                                                  a.r64, a.ti, yv ? o.nx : o.ox, yv ? o.ny : o.oy,
                                                  yv ? o.nz : o.oz, o.q64, o.r64, o.ti);
                }
                if (yv) yc[s2][t] = v;
                else xc[s2][t] = v;
            }
            __syncthreads();   // FP64 queues, corrections and phase tables complete
            // ---- FP64 pass of the mixed path over the queued pairs
            if constexpr (MIX) {
                int Qn = 0;
                for (int w = 0; w < nw; ++w) Qn += nqs[w];
                for (int e = tid; e < Qn; e += nthr) {
                    int w = 0, k = e;
                    while (k >= nqs[w]) { k -= nqs[w]; ++w; }
                    const int ent = dlist[w * qcap + k];
                    const int b = ent >> 24, j = ent & 0xffffff;
                    const BlkTrial &t = R[b];
                    const double v =
                        ldsp ? du_pair64_v<ELEC, SR>(C, tab, t.nx, t.ny, t.nz, t.ox, t.oy, t.oz, t.q64,
                                                     t.r64, t.ti, L.x[j - j0], L.y[j - j0], L.z[j - j0],
                                                     L.q[j - j0], L.r[j - j0], L.t[j - j0])
                             : du_pair64_mixed<ELEC, SR>(C, tab, lcheb, t.nx, t.ny, t.nz, t.ox, t.oy,
                                                         t.oz, t.q64, t.r64, t.ti, j);
#pragma unroll
                    for (int bb = 0; bb < B; ++bb)
                        if (bb == b) acc[bb] += v;
                }
            }
            // ---- reciprocal part over this part's k-vectors: per trial
            // A_k (2 Re(S* dS_t) + |dS_t|^2), and A_k Re(dS_s* dS_t) between
            // trials (the accept update recomputes dS of the accepted trials)
            double gacc[NG > 0 ? NG : 1];
            if constexpr (EW) {
                double rec_e[B];
#pragma unroll
                for (int b = 0; b < B; ++b) rec_e[b] = 0.0;
#pragma unroll
                for (int g = 0; g < NG; ++g) gacc[g] = 0.0;
                const int per = 3 * kcc1;
                for (int kl = tid; kl < ks; kl += nthr) {
                    double dsr[B], dsi[B];
#pragma unroll
                    for (int b = 0; b < B; ++b) { dsr[b] = 0.0; dsi[b] = 0.0; }
                    {
                        int nx, ny, nz;
                        kunpack(kid[kl], nx, ny, nz);
                        const double2 sk = S[kl];
                        const double a = Ak[kl];
#pragma unroll
                        for (int b = 0; b < B; ++b) {
                            if (b < nb) {
                                double er, ei, orr, oi;
                                ew_phase(etab + (b * 2 + 1) * per, kcc1, nx, ny, nz, er, ei);
                                ew_phase(etab + (b * 2) * per, kcc1, nx, ny, nz, orr, oi);
                                const double wq = R[b].wq;
                                const double dr = wq * (er - orr), di = wq * (ei - oi);
                                dsr[b] = dr;
                                dsi[b] = di;
                                rec_e[b] += a * (2.0 * (sk.x * dr + sk.y * di) + (dr * dr + di * di));
                            }
                        }
                        int g = 0;
#pragma unroll
                        for (int s2 = 0; s2 < B; ++s2)
#pragma unroll
                            for (int t = s2 + 1; t < B; ++t, ++g)
                                gacc[g] += a * (dsr[s2] * dsr[t] + dsi[s2] * dsi[t]);
                    }
                }
#pragma unroll
                for (int b = 0; b < B; ++b) acc[b] += C.recip_pref * rec_e[b];
            }
            // ---- block reduction of the partial sums (fixed order); on the
            // FP64 path a wavefront holding nothing for a trial (partners in
            // LDS are taken trial-per-wave, so most hold nothing) skips its
            // sum (measured: 5 % on the 500-ion salt; the mixed path's extra
            // branch costs more than it saves)
#pragma unroll
            for (int v = 0; v < B; ++v) {
                double sv;
                if constexpr (MIX)
                    sv = wave_sum_dpp(acc[v]);
                else
                    sv = __ballot(acc[v] != 0.0) ? wave_sum_dpp(acc[v]) : 0.0;
                if (lane == 0) red[wv][v] = sv;
            }
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                const double sv = wave_sum_dpp(gacc[g]);
                if (lane == 0) red[wv][B + g] = sv;
            }
            __syncthreads();
            if (tid < NV) {
                double sv = 0.0;
                for (int w = 0; w < nw; ++w) sv += red[w][tid];
                if (tid < nb)
                    for (int w = 0; w < (ldsp ? wpt : nw); ++w) sv += redp[w][tid];
                tot[tid] = sv;
            }
            exchange(NV);
            if (sabort) break;
            // trials drawn ahead by wave 1 while wave 0 decides: the ring's
            // free slots, within this call and the launch's trial budget
            const int kpre = nw >= 2 ? (int)min((long)min(2 * B - nr, C.repeat - drawn), budget - used) : 0;
            if constexpr (B >= 8) {
                // ---- Metropolis decisions in order (wave 0).  Blocks of 8 or more
                // trials: the fast form puts trial t on lanes 4t..4t+3, lane 4t+g
                // holding the earlier trials s = g, g+4, ... of it; the partial
                // sums meet by quad DPP.
                // A trial's energy change there is a finite sum and a count of
                // +inf terms.  A block with a NaN or -inf term, or with the
                // surface term, takes the sequential form (one lane per trial,
                // counts of +inf, -inf and NaN terms, which add up as IEEE sums do).
                if (wave0) {
                    wave_sync_lds();
                    constexpr int LPT = 4;                      // lanes per trial
                    constexpr int SPL = (B + LPT - 1) / LPT;    // earlier trials per lane
                    const int qt = lane / LPT, qg = lane % LPT;
                    const bool qmine = qt < nb;
                    const int qs = qmine ? qt : 0;
                    const int myi = qmine ? R[qs].i : -1;
                    const int myf = qmine ? R[qs].first : 0;
                    const double lu = qmine ? R[qs].lu : 0.0;
                    const int mycol = qmine ? R[qs].collide : 0;
                    const double base = qmine ? tot[qs] : 0.0;
                    // bit s: earlier trial s moves this trial's particle (eqm), or
                    // is the first trial of its particle in the block (ffm)
                    unsigned eqm = 0u, ffm = 0u;
    #pragma unroll
                    for (int s2 = 0; s2 < B; ++s2) {
                        const int is = __builtin_amdgcn_readlane(myi, LPT * s2);
                        const int fs = __builtin_amdgcn_readlane(myf, LPT * s2);
                        eqm |= is == myi ? 1u << s2 : 0u;
                        ffm |= fs == s2 ? 1u << s2 : 0u;
                    }
                    const unsigned usem = qmine ? (nb >= 32 ? ~0u : (1u << nb) - 1u) : 0u;
                    const unsigned ltm = (1u << qt) - 1u;      // qt < 16
                    const unsigned same = eqm & usem;          // trials of this trial's particle
                    const unsigned app = ~eqm & usem & ltm;     // earlier trials of other particles
                    const unsigned firsts = ~eqm & usem & ffm;  // ... each particle's first one
                    // this lane's earlier trials s: the first-trial term, X (the
                    // snapshot term of s's particle) and Y (after s's move); bit k
                    // of myfi / myap: s = qg + LPT k is in firsts / app
                    double xt[SPL], Xs[SPL], Ys[SPL];
                    unsigned myfi = 0u, myap = 0u;
    #pragma unroll
                    for (int k = 0; k < SPL; ++k) {
                        const int s2 = qg + LPT * k;
                        const bool in = s2 < B;
                        const int sc = in ? s2 : 0;
                        const int fs = s2 < nb ? R[sc].first : 0;
                        xt[k] = xc[sc][qs];
                        Xs[k] = xc[fs][qs];
                        Ys[k] = yc[sc][qs];
                        if constexpr (EW) {
                            const bool gg = qmine && in && s2 < qt;
                            const int gi = gg ? B + s2 * (2 * B - s2 - 1) / 2 + (qt - s2 - 1) : 0;
                            const double gv = tot[gi];
                            Ys[k] += gg ? 2.0 * C.recip_pref * gv : 0.0;
                        }
                        myfi |= (in && ((firsts >> sc) & 1u)) ? 1u << k : 0u;
                        myap |= (in && ((app >> sc) & 1u)) ? 1u << k : 0u;
                    }
                    // a NaN or -inf among a trial's terms makes their IEEE sum NaN
                    // or -inf (+inf + -inf is NaN); an overflow to -inf only sends
                    // the block down the sequential path
                    double pp = 0.0;
    #pragma unroll
                    for (int k = 0; k < SPL; ++k) {
                        const double x1 = ((myfi >> k) & 1u) ? xt[k] : 0.0;
                        const bool ap = (myap >> k) & 1u;
                        pp += x1 + ((ap ? Xs[k] : 0.0) + (ap ? Ys[k] : 0.0));
                    }
                    const double probe = base + quad_sum(pp);
                    const bool odd = probe != probe || probe == -__builtin_inf();
                    int nd = nb;
                    unsigned accm = 0;
                    bool refresh = false;
                    if (!__ballot(odd) && !(EW && C.surf_pref != 0.0)) {
                        // a collision counts as one +inf term (energy.h:3597-3607)
                        auto fpart = [](double v) { return v == __builtin_inf() ? 0.0 : v; };
                        auto ipart = [](double v) { return v == __builtin_inf() ? 1 : 0; };
                        double pf = 0.0;
                        int pi = 0;
                        double dfin[SPL];
                        int dinf[SPL];
    #pragma unroll
                        for (int k = 0; k < SPL; ++k) {
                            const bool fi = (myfi >> k) & 1u;
                            pf += fi ? fpart(xt[k]) : 0.0;
                            pi += fi ? ipart(xt[k]) : 0;
                            const bool ap = (myap >> k) & 1u;
                            dfin[k] = ap ? fpart(Ys[k]) - fpart(Xs[k]) : 0.0;
                            dinf[k] = ap ? ipart(Ys[k]) - ipart(Xs[k]) : 0;
                        }
                        const double fin0 = fpart(base) + quad_sum(pf);
                        const int npi0 = (mycol ? 1 : 0) + ipart(base) + quad_isum(pi);
                        // The decisions as a fixed point: from a guess of the
                        // accepted set, every trial forms its energy change and
                        // Metropolis decision (move.h:567-573) at once; trial t
                        // depends on trials s < t only, so after k rounds the first
                        // k decisions are the sequential ones, and a set that
                        // reproduces itself is the sequential answer.
                        unsigned am = 0;
                        double mydu = 0.0;
                        for (int it = 0; it <= B; ++it) {
                            double f = 0.0;
                            int ni = 0;
    #pragma unroll
                            for (int k = 0; k < SPL; ++k) {
                                const bool on = (am >> (qg + LPT * k)) & 1u;
                                f += on ? dfin[k] : 0.0;
                                ni += on ? dinf[k] : 0;
                            }
                            f = quad_sum(f);
                            ni = npi0 + quad_isum(ni);
                            mydu = ni > 0 ? __builtin_inf() : fin0 + f;
                            const bool acc = qmine && (mydu <= 0.0 || !(mydu > lu));
                            const unsigned nm = quad_bits(__ballot(acc && qg == 0));
                            if (nm == am) break;
                            am = nm;
                        }
                        // a trial whose particle an accepted trial of this block
                        // moved waits for the next block
                        const unsigned cut = quad_bits(__ballot(qg == 0 && qmine && (same & am & ltm) != 0u));
                        if (cut) nd = __ffs(cut) - 1;
                        accm = am & ((nd >= 32 ? 0u : (1u << nd)) - 1u);
                        if constexpr (EW) {
                            // NonbondedEwald::update: the accepted move that takes
                            // the counter past update_frequency - 1 ends the block
                            const int need = C.update_frequency - ew_cnt;   // accepts until the refresh
                            if (need >= 1 && __popc(accm) >= need) {
                                unsigned m2 = accm;
                                for (int k = 1; k < need; ++k) m2 &= m2 - 1u;
                                const int tr = __ffs(m2) - 1;
                                nd = tr + 1;
                                accm &= (2u << tr) - 1u;
                                refresh = true;
                            }
                            ew_cnt += __popc(accm);
                        }
    #pragma unroll
                        for (int t = 0; t < B; ++t)
                            if ((accm >> t) & 1u) {
                                const double d = readlane_f64(mydu, LPT * t);
                                utot += d;
                                dusum += d;
                            }
                        n_inf += __popcll(__ballot(qg == 0 && qt < nd && !((accm >> qt) & 1u) &&
                                                   mydu == __builtin_inf()));
                        if (qg == 0 && qt < B) sdu[qt] = mydu;
                    } else {
                        // one lane per trial; the masks again, and earlier trial
                        // s's terms read from LDS when used
                        const bool mine = lane < nb;
                        const int sl = mine ? lane : 0;
                        const int si = mine ? R[sl].i : -1;
                        const int sf = mine ? R[sl].first : 0;
                        const double slu = mine ? R[sl].lu : 0.0;
                        const int scol = mine ? R[sl].collide : 0;
                        unsigned seq = 0u, sff = 0u;
    #pragma unroll
                        for (int s2 = 0; s2 < B; ++s2) {
                            const int is = __builtin_amdgcn_readlane(si, s2);
                            const int fs = __builtin_amdgcn_readlane(sf, s2);
                            seq |= is == si ? 1u << s2 : 0u;
                            sff |= fs == s2 ? 1u << s2 : 0u;
                        }
                        const unsigned suse = mine ? (nb >= 32 ? ~0u : (1u << nb) - 1u) : 0u;
                        const unsigned slt = lane < 32 ? (1u << lane) - 1u : ~0u;
                        const unsigned ssame = seq & suse;
                        const unsigned sapp = ~seq & suse & slt;
                        const unsigned sfirsts = ~seq & suse & sff;
                        auto Xof = [&](int s2) { return xc[__builtin_amdgcn_readlane(sf, s2)][sl]; };
                        auto Yof = [&](int s2) {
                            double y = yc[s2][sl];
                            if constexpr (EW) {
                                const bool gg = mine && ((slt >> s2) & 1u);
                                const int gi = gg ? B + s2 * (2 * B - s2 - 1) / 2 + (lane - s2 - 1) : 0;
                                const double gv = tot[gi];
                                y += gg ? 2.0 * C.recip_pref * gv : 0.0;
                            }
                            return y;
                        };
                        double mydu = 0.0;
                        int moved = 0;   // this lane's particle was moved by an accepted trial
                        double fin = 0.0;
                        int npi = 0, nmi = 0, nna = 0;
                        auto add = [&](double v, int sgn) {
                            const bool isn = v != v, isp = v == __builtin_inf(),
                                       ism = v == -__builtin_inf();
                            nna += isn ? sgn : 0;
                            npi += isp ? sgn : 0;
                            nmi += ism ? sgn : 0;
                            fin += (isn || isp || ism) ? 0.0 : (sgn > 0 ? v : -v);
                        };
                        add(mine ? tot[sl] : 0.0, 1);
    #pragma unroll
                        for (int s2 = 0; s2 < B; ++s2)
                            if ((sfirsts >> s2) & 1u) add(xc[s2][sl], 1);
    #pragma unroll
                        for (int t = 0; t < B; ++t) {
                            if (t >= nb) break;
                            if (__builtin_amdgcn_readlane(moved, t)) { nd = t; break; }
                            const double v = (nna > 0 || (npi > 0 && nmi > 0)) ? __builtin_nan("")
                                             : npi > 0 ? __builtin_inf()
                                             : nmi > 0 ? -__builtin_inf() : fin;
                            double du = readlane_f64(v, t);
                            double dMx = 0.0, dMy = 0.0, dMz = 0.0;
                            if (EW && C.surf_pref != 0.0) {
                                const BlkTrial &T = R[t];
                                dMx = T.wq * (T.nx - T.ox);
                                dMy = T.wq * (T.ny - T.oy);
                                dMz = T.wq * (T.nz - T.oz);
                                const double Nx = Mx + dMx, Ny = My + dMy, Nz = Mz + dMz;
                                du += C.surf_pref *
                                      ((Nx * Nx + (Ny * Ny + Nz * Nz)) - (Mx * Mx + (My * My + Mz * Mz)));
                            }
                            if (__builtin_amdgcn_readlane(scol, t)) du = __builtin_inf();   // energy.h:3597-3607
                            const bool accept = du <= 0.0 || !(du > readlane_f64(slu, t));
                            if (lane == t) mydu = du;
                            if (accept) {
                                accm |= 1u << t;
                                utot += du;
                                dusum += du;
                                if (EW) { Mx += dMx; My += dMy; Mz += dMz; }
                                if ((sapp >> t) & 1u) {
                                    add(Xof(t), -1);
                                    add(Yof(t), 1);
                                }
                                moved |= (int)((ssame >> t) & 1u);
                                if (EW && ++ew_cnt > C.update_frequency - 1) {
                                    refresh = true;
                                    nd = t + 1;
                                    break;
                                }
                            } else if (du == __builtin_inf()) {
                                n_inf++;
                            }
                        }
                        if (lane < B) sdu[lane] = mydu;
                    }
                    n_trials += nd;
                    n_acc += __popc(accm);
                    if (lane == 0) { snd = nd; saccm = (int)accm; srefresh = refresh ? 1 : 0; }
                    if (lane < B) swq[lane] = lane < nb ? R[lane].wq : 0.0;
                }
            } else {
                // ---- Metropolis decisions in order (wave 0), smaller blocks: lane
                // t holds trial t's energy change as a finite sum and counts of
                // +inf, -inf and NaN terms, which add up as IEEE sums do (the
                // four-lane form spills registers in the Ewald instance and runs
                // slower there)
                if (wave0) {
                    wave_sync_lds();
                    const bool mine = lane < nb;
                    // every load first (lane-parallel), then the sequential part on
                    // per-lane registers (lane t: trial t) and a few scalars
                    const int myi = mine ? R[lane].i : -1;
                    const int myf = mine ? R[lane].first : 0;
                    const double lu = mine ? R[lane].lu : 0.0;
                    const int mycol = mine ? R[lane].collide : 0;
                    const double base = mine ? tot[lane] : 0.0;
                    // per earlier trial s: X (its particle's snapshot term), Y (after
                    // its accepted move), and whether it moves this lane's particle
                    // (same) or changes this lane's energy (app)
                    // the masks as integer bit sets (bit s: earlier trial s), the
                    // terms loaded by every lane (lanes past the block read column
                    // 0) and used only under the masks, so nothing here branches
                    unsigned eqm = 0u, ffm = 0u;
                    double xt[B], Xs[B], Ys[B];
                    const int sl = mine ? lane : 0;
                    const unsigned ltm = lane < 32 ? (1u << lane) - 1u : ~0u;
    #pragma unroll
                    for (int s2 = 0; s2 < B; ++s2) {
                        const int is = __builtin_amdgcn_readlane(myi, s2);
                        const int fs = __builtin_amdgcn_readlane(myf, s2);
                        eqm |= is == myi ? 1u << s2 : 0u;
                        ffm |= fs == s2 ? 1u << s2 : 0u;
                        xt[s2] = xc[s2][sl];
                        Xs[s2] = xc[fs][sl];
                        Ys[s2] = yc[s2][sl];
                        if constexpr (EW) {
                            const bool g = mine && ((ltm >> s2) & 1u);
                            const int gi = g ? B + s2 * (2 * B - s2 - 1) / 2 + (lane - s2 - 1) : 0;
                            const double gv = tot[gi];
                            Ys[s2] += g ? 2.0 * C.recip_pref * gv : 0.0;
                        }
                    }
                    const unsigned usem = mine ? (nb >= 32 ? ~0u : (1u << nb) - 1u) : 0u;
                    const unsigned same = eqm & usem;          // trials of this lane's particle
                    const unsigned app = ~eqm & usem & ltm;     // earlier trials of other particles
                    const unsigned firsts = ~eqm & usem & ffm;  // ... each particle's first one
                    // a NaN or -inf among this lane's terms makes their IEEE sum NaN
                    // or -inf (+inf + -inf is NaN); an overflow to -inf only sends
                    // the block down the sequential path
                    double probe = base;
    #pragma unroll
                    for (int s2 = 0; s2 < B; ++s2) {
                        const double x1 = ((firsts >> s2) & 1u) ? xt[s2] : 0.0;
                        const bool ap = (app >> s2) & 1u;
                        probe += x1 + ((ap ? Xs[s2] : 0.0) + (ap ? Ys[s2] : 0.0));
                    }
                    const bool odd = probe != probe || probe == -__builtin_inf();
                    int nd = nb;
                    unsigned accm = 0;
                    bool refresh = false;
                    double mydu = 0.0;
                    int moved = 0;   // this lane's particle was moved by an accepted trial
                    if (!__ballot(odd) && !(EW && C.surf_pref != 0.0)) {
                        // lane t's energy change: a finite sum and a count of +inf
                        // terms (a collision counts as one: energy.h:3597-3607),
                        // and what an accepted earlier trial s changes in them
                        auto fpart = [](double v) { return v == __builtin_inf() ? 0.0 : v; };
                        auto ipart = [](double v) { return v == __builtin_inf() ? 1 : 0; };
                        double fin0 = fpart(base);
                        int npi0 = (mycol ? 1 : 0) + ipart(base);
                        double dfin[B];
                        int dinf[B];
    #pragma unroll
                        for (int s2 = 0; s2 < B; ++s2) {
                            const bool fi = (firsts >> s2) & 1u;
                            fin0 += fi ? fpart(xt[s2]) : 0.0;
                            npi0 += fi ? ipart(xt[s2]) : 0;
                            const bool ap = (app >> s2) & 1u;
                            dfin[s2] = ap ? fpart(Ys[s2]) - fpart(Xs[s2]) : 0.0;
                            dinf[s2] = ap ? ipart(Ys[s2]) - ipart(Xs[s2]) : 0;
                        }
                        // The decisions as a fixed point: from a guess of the
                        // accepted set, every lane forms its energy change and
                        // Metropolis decision (move.h:567-573) at once; trial t
                        // depends on trials s < t only, so after k rounds the first
                        // k decisions are the sequential ones, and a set that
                        // reproduces itself is the sequential answer.
                        unsigned am = 0;
                        for (int it = 0; it <= B; ++it) {
                            double f = fin0;
                            int ni = npi0;
    #pragma unroll
                            for (int s2 = 0; s2 < B; ++s2)
                                if ((am >> s2) & 1u) { f += dfin[s2]; ni += dinf[s2]; }
                            mydu = ni > 0 ? __builtin_inf() : f;
                            const bool acc = mine && (mydu <= 0.0 || !(mydu > lu));
                            const unsigned nm = (unsigned)__ballot(acc);
                            if (nm == am) break;
                            am = nm;
                        }
                        // a trial whose particle an accepted trial of this block
                        // moved waits for the next block
                        const unsigned below_me = lane < 32 ? (1u << lane) - 1u : ~0u;
                        const unsigned cut = (unsigned)__ballot(mine && (same & am & below_me) != 0u);
                        if (cut) nd = __ffs(cut) - 1;
                        accm = am & ((nd >= 32 ? 0u : (1u << nd)) - 1u);
                        if constexpr (EW) {
                            // NonbondedEwald::update: the accepted move that takes
                            // the counter past update_frequency - 1 ends the block
                            const int need = C.update_frequency - ew_cnt;   // accepts until the refresh
                            if (need >= 1 && __popc(accm) >= need) {
                                unsigned m2 = accm;
                                for (int k = 1; k < need; ++k) m2 &= m2 - 1u;
                                const int tr = __ffs(m2) - 1;
                                nd = tr + 1;
                                accm &= (2u << tr) - 1u;
                                refresh = true;
                            }
                            ew_cnt += __popc(accm);
                        }
    #pragma unroll
                        for (int t = 0; t < B; ++t)
                            if ((accm >> t) & 1u) {
                                const double d = readlane_f64(mydu, t);
                                utot += d;
                                dusum += d;
                            }
                        n_inf += __popcll(__ballot(lane < nd && !((accm >> lane) & 1u) &&
                                                   mydu == __builtin_inf()));
                    } else {
                        // the same with counts of +inf, -inf and NaN terms, which
                        // add up as IEEE sums do
                        double fin = 0.0;
                        int npi = 0, nmi = 0, nna = 0;
                        auto add = [&](double v, int sgn) {
                            const bool isn = v != v, isp = v == __builtin_inf(),
                                       ism = v == -__builtin_inf();
                            nna += isn ? sgn : 0;
                            npi += isp ? sgn : 0;
                            nmi += ism ? sgn : 0;
                            fin += (isn || isp || ism) ? 0.0 : (sgn > 0 ? v : -v);
                        };
                        add(base, 1);
    #pragma unroll
                        for (int s2 = 0; s2 < B; ++s2)
                            if ((firsts >> s2) & 1u) add(xt[s2], 1);
    #pragma unroll
                        for (int t = 0; t < B; ++t) {
                            if (t >= nb) break;
                            if (__builtin_amdgcn_readlane(moved, t)) { nd = t; break; }
                            const double v = (nna > 0 || (npi > 0 && nmi > 0)) ? __builtin_nan("")
                                             : npi > 0 ? __builtin_inf()
                                             : nmi > 0 ? -__builtin_inf() : fin;
                            double du = readlane_f64(v, t);
                            double dMx = 0.0, dMy = 0.0, dMz = 0.0;
                            if (EW && C.surf_pref != 0.0) {
                                const BlkTrial &T = R[t];
                                dMx = T.wq * (T.nx - T.ox);
                                dMy = T.wq * (T.ny - T.oy);
                                dMz = T.wq * (T.nz - T.oz);
                                const double Nx = Mx + dMx, Ny = My + dMy, Nz = Mz + dMz;
                                du += C.surf_pref *
                                      ((Nx * Nx + (Ny * Ny + Nz * Nz)) - (Mx * Mx + (My * My + Mz * Mz)));
                            }
                            if (__builtin_amdgcn_readlane(mycol, t)) du = __builtin_inf();   // energy.h:3597-3607
                            const bool accept = du <= 0.0 || !(du > readlane_f64(lu, t));
                            if (lane == t) mydu = du;
                            if (accept) {
                                accm |= 1u << t;
                                utot += du;
                                dusum += du;
                                if (EW) { Mx += dMx; My += dMy; Mz += dMz; }
                                if ((app >> t) & 1u) {
                                    add(Xs[t], -1);
                                    add(Ys[t], 1);
                                }
                                moved |= (int)((same >> t) & 1u);
                                if (EW && ++ew_cnt > C.update_frequency - 1) {
                                    refresh = true;
                                    nd = t + 1;
                                    break;
                                }
                            } else if (du == __builtin_inf()) {
                                n_inf++;
                            }
                        }
                    }
                    n_trials += nd;
                    n_acc += __popc(accm);
                    if (lane == 0) { snd = nd; saccm = (int)accm; srefresh = refresh ? 1 : 0; }
                    if (lane < B) {
                        sdu[lane] = mydu;
                        swq[lane] = lane < nb ? R[lane].wq : 0.0;
                    }
                }
            }
            if (wv == 1 && kpre > 0) {
                // with their particles' data: the accepted vector holds every
                // move before this block's (written blocks ago, barriers
                // between); this block's moves are corrected for at the next
                // block's start, as for every trial in the ring
                draw(nr, kpre);
                wave_sync_lds();
                if (lane < kpre) fetch_particle(R[nr + lane]);
            }
            __syncthreads();
            nr += kpre;
            used += kpre;
            drawn += kpre;
            const int nd = snd;
            const unsigned accm = (unsigned)saccm;
            const bool refresh = srefresh != 0;
            // ---- accepted moves into the chain state
            if (tid < nd) {
                const BlkTrial &T = R[tid];
                const bool a = (accm >> tid) & 1u;
                if (a) {   // AtomicTranslation::_acceptMove (move.h:866-877)
                    const int i = T.i;
                    C.x[i] = T.nx; C.y[i] = T.ny; C.z[i] = T.nz;
                    C.pk_w[i] = make_int4(T.fnx, T.fny, T.fnz, T.pkw);
                    if (ldsp && i >= j0 && i < j1) {
                        const int r = i - j0;
                        L.pk[r] = int4_to_v4i(make_int4(T.fnx, T.fny, T.fnz, T.pkw));
                        L.x[r] = T.nx; L.y[r] = T.ny; L.z[r] = T.nz;
                    }
                }
                if (part == 0 && C.du_rec && rec + tid < C.cap) {
                    C.du_rec[rec + tid] = sdu[tid];
                    C.acc_rec[rec + tid] = a ? 1 : 0;
                }
            }
            rec += nd;
            if constexpr (EW) {
                // S(k) += dS(k) of the accepted trials, in order (dS as the
                // reciprocal pass formed it, from the same phase tables)
                if (accm) {
                    const int per = 3 * kcc1;
                    for (int kl = tid; kl < ks; kl += nthr) {
                        int nx, ny, nz;
                        kunpack(kid[kl], nx, ny, nz);
                        double2 sk = S[kl];
#pragma unroll
                        for (int b = 0; b < B; ++b)
                            if (b < nd && ((accm >> b) & 1u)) {
                                double er, ei, orr, oi;
                                ew_phase(etab + (b * 2 + 1) * per, kcc1, nx, ny, nz, er, ei);
                                ew_phase(etab + (b * 2) * per, kcc1, nx, ny, nz, orr, oi);
                                const double wq = swq[b];
                                sk.x = sk.x + wq * (er - orr);
                                sk.y = sk.y + wq * (ei - oi);
                            }
                        S[kl] = sk;
                    }
                }
            }
            if (wave0) {
                wave_sync_lds();   // the wave's reads of the ring are done
                // the accepted moves, in order, one lane each (accm holds
                // trials below nd only)
                if (lane < nd && ((accm >> lane) & 1u)) {
                    const int m = __popc(accm & ((1u << lane) - 1u));
                    lmv_i[m] = R[lane].i;
                    lmv_x[m] = R[lane].nx; lmv_y[m] = R[lane].ny; lmv_z[m] = R[lane].nz;
                }
                if (lane == 0) snlmv = __popc(accm);
                // undecided trials keep their draws for the next block: read
                // by every lane first, written after the wave's reads are done
                constexpr int NW = (int)(sizeof(BlkTrial) / 4);
                static_assert(sizeof(BlkTrial) % 4 == 0, "ring records copy as words");
                const bool keep = lane >= nd && lane < nr;
                uint32_t kt[NW];
                const uint32_t *src = (const uint32_t *)&R[keep ? lane : 0];
#pragma unroll
                for (int w = 0; w < NW; ++w) kt[w] = src[w];
                wave_sync_lds();
                if (keep) {
                    uint32_t *dst = (uint32_t *)&R[lane - nd];
#pragma unroll
                    for (int w = 0; w < NW; ++w) dst[w] = kt[w];
                }
                wave_sync_lds();
            }
            nr -= nd;
            if constexpr (EW) {
                if (refresh) {
                    // NonbondedEwald::update (ewald.h:626-643): S(k) from all
                    // positions; the reciprocal-energy change enters utot
                    __syncthreads();
                    double eb = 0.0;
                    for (int k = tid; k < ks; k += nthr)
                        eb += Ak[k] * (S[k].x * S[k].x + S[k].y * S[k].y);
                    for (int k = tid; k < ks; k += nthr) S[k] = make_double2(0.0, 0.0);
                    const bool tabled = kcc1 <= REFRESH_KAX;
                    const int per = 3 * kcc1;
                    for (int jc = 0; jc < C.n; jc += REFRESH_CHUNK) {
                        const int nj = min(REFRESH_CHUNK, C.n - jc);
                        __syncthreads();
                        if (tabled) {
                            for (int e = tid; e < nj * per; e += nthr) {
                                const int jj = e / per, rem = e % per;
                                const int axis = rem / kcc1, nn = rem % kcc1;
                                const int j = jc + jj;
                                const double coord = axis == 0 ? C.x[j] : axis == 1 ? C.y[j] : C.z[j];
                                double sn, cs;
                                sincospi(2.0 * (double)nn * coord * C.box.inv[axis], &sn, &cs);
                                rtab[jj * per + axis * kcc1 + nn] = make_double2(cs, sn);
                            }
                        }
                        for (int jj = tid; jj < nj; jj += nthr)
                            rw[jj] = C.q[jc + jj] * C.eff[C.t[jc + jj]];
                        __syncthreads();
                        for (int k = tid; k < ks; k += nthr) {
                            int nx, ny, nz;
                            kunpack(kid[k], nx, ny, nz);
                            double sr = 0.0, si = 0.0;
                            for (int jj = 0; jj < nj; ++jj) {
                                double er, ei;
                                if (tabled) {
                                    ew_phase(&rtab[jj * per], kcc1, nx, ny, nz, er, ei);
                                } else {
                                    const int j = jc + jj;
                                    const double ph = 2.0 * ((double)nx * C.x[j] * C.box.inv[0] +
                                                             (double)ny * C.y[j] * C.box.inv[1] +
                                                             (double)nz * C.z[j] * C.box.inv[2]);
                                    sincospi(ph, &ei, &er);
                                }
                                sr += rw[jj] * er;
                                si += rw[jj] * ei;
                            }
                            const double2 sk = S[k];
                            S[k] = make_double2(sk.x + sr, sk.y + si);
                        }
                    }
                    double ea = 0.0;
                    for (int k = tid; k < ks; k += nthr)
                        ea += Ak[k] * (S[k].x * S[k].x + S[k].y * S[k].y);
                    eb = block_sum(eb, redm);
                    ea = block_sum(ea, redm);
                    if (tid == 0) { tot[0] = eb; tot[1] = ea; }
                    __syncthreads();
                    exchange(2);
                    if (sabort) break;
                    utot += C.recip_pref * (tot[1] - tot[0]);
                    if (C.surf_pref != 0.0) {
                        Mx = My = Mz = 0.0;
                        for (int j = tid; j < C.n; j += nthr) {
                            const double w = C.q[j] * C.eff[C.t[j]];
                            Mx += w * C.x[j]; My += w * C.y[j]; Mz += w * C.z[j];
                        }
                        Mx = block_sum(Mx, redm);
                        My = block_sum(My, redm);
                        Mz = block_sum(Mz, redm);
                    }
                    ew_cnt = 0;
                    n_ref++;
                    __syncthreads();
                }
            }
        }
        if (!stop && !sabort) ++calls_done;
    }
    // ---- write back the chain state
    if (nsplit > 1 && !sabort) chain_sync();
    __syncthreads();
    if (part == 0) {
        if (!stop && tid == 0 && C.resume) C.resume[0] = 0;
        for (int k = tid; k < 624; k += nthr) {
            C.rng[k] = mtg[k];
            C.rng[625 + k] = mti[k];
        }
        if (tid == 0) {
            C.rng[624] = (uint32_t)sgp;
            C.rng[625 + 624] = (uint32_t)I.p;
            C.stats_i[0] = n_trials;
            C.stats_i[1] = n_acc;
            C.stats_i[2] = n_inf;
            C.stats_i[3] = n_ref;
            C.stats_i[4] = wall_clock64() - t_start;
            C.stats_i[5] = calls_done;
            C.stats_i[6] = 0;
            C.stats_d[0] = utot;
            C.stats_d[1] = dusum;
            if (EW) *C.ew_cnt = ew_cnt;
        }
    }
    if constexpr (EW) {
        for (int k = tid; k < ks; k += nthr) {
            C.Q[2 * (k0 + k)] = S[k].x;
            C.Q[2 * (k0 + k) + 1] = S[k].y;
        }
    }
    // the last part out resets the chain's counters for the next launch
    if (nsplit > 1 && tid == 0) {
        __threadfence();
        const int d = __hip_atomic_fetch_add(C.xcnt + 1, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (d == nsplit - 1) {
            __hip_atomic_store(C.xcnt, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(C.xcnt + 1, 0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

template <int ELEC, int SR, bool MIX, int B>
static void launch_blk_f(hipStream_t st, const ChainDev *d_chains, const int2 *d_map, int ngrid,
                         int nthr, size_t dyn)
{
    if (dyn > 0)
        (void)hipFuncSetAttribute((const void *)k_sweep_blk<ELEC, SR, MIX, B>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
    hipLaunchKernelGGL((k_sweep_blk<ELEC, SR, MIX, B>), dim3(ngrid), dim3(nthr), dyn, st, d_chains,
                       d_map);
}

// static LDS of a block-sweep instance (for the host's dynamic-LDS budget)
template <int ELEC, int SR, bool MIX, int B>
static size_t blk_static_lds()
{
    hipFuncAttributes a{};
    if (hipFuncGetAttributes(&a, (const void *)k_sweep_blk<ELEC, SR, MIX, B>) != hipSuccess) return 0;
    return a.sharedSizeBytes;
}

// Block sweep launchers (fgx_sweep_blk_*.hip): FGX_EUNSUPPORTED without an
// error message for a Hamiltonian they have no instance of.
int launch_blk_pair(bool mix, const Pot &P, hipStream_t st, const ChainDev *d_chains,
                    const int2 *d_map, int ngrid, int nthr, size_t dyn);
int launch_blk_ewald(bool mix, const Pot &P, hipStream_t st, const ChainDev *d_chains,
                     const int2 *d_map, int ngrid, int nthr, size_t dyn);
// whether an instance exists for this Hamiltonian (1) and its static LDS and
// trials per block
int blk_instance_pair(bool mix, const Pot &P, size_t *static_lds, int *block);
int blk_instance_ewald(bool mix, const Pot &P, size_t *static_lds, int *block);

}  // namespace fgx
