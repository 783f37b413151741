// fgx_system.hip -- fgx_system handle, transfers, and the per-call energy
// entry points that replace Energy::Energybase's virtual interface
// (energy.h:54-224), Energy::energyChange (energy.h:3577-3621) and the
// NonbondedEwald state machine (ewald.h:229-760).
//
// Every sum here is an FP64 block reduction with a fixed combination order,
// so results are deterministic run to run.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>

#include "fgx_device.hpp"
#include "fgx_internal.hpp"

namespace fgx {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

static constexpr int RB = 256;      // reduction block size
static constexpr int TILE = 64;     // LDS tile of "row" particles

// ------------------------------------------------------------------------
// kernel-side view of one particle vector
// ------------------------------------------------------------------------
struct PV {
    const double *x, *y, *z, *q, *r;
    const int *t, *gid;
};

static PV pv(const fgx_system *s, int which)
{
    return PV{s->v[which].x, s->v[which].y, s->v[which].z, s->d_q, s->d_r, s->d_type, s->d_gid};
}

// squared distance as the energy class computes it: Nonbonded uses
// Cuboid::sqdist (energy.h:842-845), NonbondedVector and the Ewald real
// space use Cuboid::vdist (energy.h:1119-1122, ewald.h:104)
__device__ __forceinline__ double r2_of(const Box &b, const Pot &P, double ax, double ay, double az,
                                        double bx, double by, double bz)
{
    if (P.nbkind == NB_VECTOR || P.elec == ELEC_EWALD) {
        double rx, ry, rz;
        vdist_d(b, ax, ay, az, bx, by, bz, rx, ry, rz);
        return rx * rx + (ry * ry + rz * rz);
    }
    return sqdist_d(b, ax, ay, az, bx, by, bz);
}

__device__ __forceinline__ double pair_ab(const Box &b, const Pot &P, const double *tab, double ax,
                                          double ay, double az, double aq, double ar, int at,
                                          double bx, double by, double bz, double bq, double br,
                                          int bt)
{
    double r2 = r2_of(b, P, ax, ay, az, bx, by, bz);
    return pair_u64(P, tab, r2, aq, bq, ar, br, at, bt);
}

// write a block's partial sum
__device__ __forceinline__ void block_store(double acc, double *partial, int slot)
{
    __shared__ double red[RB / 64];
    double t = block_sum(acc, red);
    if (threadIdx.x == 0) partial[slot] = t;
}

// sum_{j in [j0,j1), j != skip} u(A, j)  (Nonbonded::i2g / i2all rows)
__global__ __launch_bounds__(RB) void k_row(Box b, Pot P, const double *tab, PV v, int ia, int j0,
                                            int j1, int skip, double *partial)
{
    double ax = v.x[ia], ay = v.y[ia], az = v.z[ia], aq = v.q[ia], ar = v.r[ia];
    int at = v.t[ia];
    double acc = 0.0;
    for (int j = j0 + blockIdx.x * RB + threadIdx.x; j < j1; j += gridDim.x * RB) {
        if (j == skip) continue;
        acc += pair_ab(b, P, tab, ax, ay, az, aq, ar, at, v.x[j], v.y[j], v.z[j], v.q[j], v.r[j],
                       v.t[j]);
    }
    block_store(acc, partial, blockIdx.x);
}

// sum_{i in [a0,a1)} sum_{j in [b0,b1)} u(i,j)   (Nonbonded::g2g)
// tri != 0: only pairs with i < j                 (Nonbonded::g_internal)
// mask != nullptr: pairs of different groups gi, gj count only if
//                  mask[gi*ng+gj] (NonbondedCutg2g over the whole system)
__global__ __launch_bounds__(RB) void k_block(Box b, Pot P, const double *tab, PV v, int a0, int a1,
                                              int b0, int b1, int tri, const unsigned char *mask,
                                              int ng, double *partial)
{
    __shared__ double sx[TILE], sy[TILE], sz[TILE], sq[TILE], sr[TILE];
    __shared__ int st[TILE], sg[TILE];
    const int i0 = a0 + blockIdx.y * TILE;
    const int ni = min(TILE, a1 - i0);
    if (threadIdx.x < ni) {
        int i = i0 + threadIdx.x;
        sx[threadIdx.x] = v.x[i]; sy[threadIdx.x] = v.y[i]; sz[threadIdx.x] = v.z[i];
        sq[threadIdx.x] = v.q[i]; sr[threadIdx.x] = v.r[i]; st[threadIdx.x] = v.t[i];
        sg[threadIdx.x] = v.gid[i];
    }
    __syncthreads();
    double acc = 0.0;
    int j = b0 + blockIdx.x * RB + threadIdx.x;
    if (j < b1 && ni > 0) {
        double bx = v.x[j], by = v.y[j], bz = v.z[j], bq = v.q[j], br = v.r[j];
        int bt = v.t[j], bg = v.gid[j];
        for (int k = 0; k < ni; ++k) {
            int i = i0 + k;
            if (tri && i >= j) break;
            if (mask && sg[k] != bg && !mask[sg[k] * ng + bg]) continue;
            acc += pair_ab(b, P, tab, sx[k], sy[k], sz[k], sq[k], sr[k], st[k], bx, by, bz, bq, br,
                           bt);
        }
    }
    block_store(acc, partial, blockIdx.y * gridDim.x + blockIdx.x);
}

// Pair part of Energy::energyChange for a change confined to one group:
// the moved sites (rows, trial and accepted coordinates) against every
// particle that did not move.  Per other group h, keep flags decide whether
// the trial / accepted interaction counts (NonbondedCutg2g; all ones
// otherwise).  Moved-vs-moved pairs are added by block (0,0) when `list`
// (explicit indices in an atomic group, energy.h:3545-3552).
__global__ __launch_bounds__(RB) void k_delta(Box b, Pot P, const double *tab, PV pp, PV pt,
                                              const int *idx, int nm, int range_front,
                                              int range_back, int list, int n,
                                              const unsigned char *keep_new,
                                              const unsigned char *keep_old, double *partial)
{
    __shared__ double nx[TILE], ny[TILE], nz[TILE], ox[TILE], oy[TILE], oz[TILE];
    __shared__ double sq[TILE], sr[TILE];
    __shared__ int st[TILE], sidx[TILE];
    const int k0 = blockIdx.y * TILE;
    const int nk = min(TILE, nm - k0);
    if (threadIdx.x < nk) {
        int m = idx ? idx[k0 + threadIdx.x] : range_front + k0 + threadIdx.x;
        sidx[threadIdx.x] = m;
        nx[threadIdx.x] = pt.x[m]; ny[threadIdx.x] = pt.y[m]; nz[threadIdx.x] = pt.z[m];
        ox[threadIdx.x] = pp.x[m]; oy[threadIdx.x] = pp.y[m]; oz[threadIdx.x] = pp.z[m];
        sq[threadIdx.x] = pp.q[m]; sr[threadIdx.x] = pp.r[m]; st[threadIdx.x] = pp.t[m];
    }
    __syncthreads();
    double acc = 0.0;
    int j = blockIdx.x * RB + threadIdx.x;
    if (j < n && nk > 0) {
        bool moved;
        if (list) {
            moved = false;
            for (int k = 0; k < nm; ++k) moved |= (idx[k] == j);
        } else {
            moved = (j >= range_front && j <= range_back);
        }
        if (!moved) {
            double bx = pp.x[j], by = pp.y[j], bz = pp.z[j], bq = pp.q[j], br = pp.r[j];
            int bt = pp.t[j], h = pp.gid[j];
            bool kn = keep_new[h], ko = keep_old[h];
            for (int k = 0; k < nk; ++k) {
                double un = kn ? pair_ab(b, P, tab, nx[k], ny[k], nz[k], sq[k], sr[k], st[k], bx,
                                         by, bz, bq, br, bt)
                               : 0.0;
                double uo = ko ? pair_ab(b, P, tab, ox[k], oy[k], oz[k], sq[k], sr[k], st[k], bx,
                                         by, bz, bq, br, bt)
                               : 0.0;
                acc += un - uo;
            }
        }
    }
    if (list && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        for (int a = 0; a < nm; ++a)
            for (int c = 0; c < nm; ++c) {
                int ja = idx[a], jc = idx[c];
                if (ja <= jc) continue;
                double un = pair_ab(b, P, tab, pt.x[ja], pt.y[ja], pt.z[ja], pt.q[ja], pt.r[ja],
                                    pt.t[ja], pt.x[jc], pt.y[jc], pt.z[jc], pt.q[jc], pt.r[jc],
                                    pt.t[jc]);
                double uo = pair_ab(b, P, tab, pp.x[ja], pp.y[ja], pp.z[ja], pp.q[ja], pp.r[ja],
                                    pp.t[ja], pp.x[jc], pp.y[jc], pp.z[jc], pp.q[jc], pp.r[jc],
                                    pp.t[jc]);
                acc += un - uo;
            }
    }
    block_store(acc, partial, blockIdx.y * gridDim.x + blockIdx.x);
}

// fixed-order final reduction of partials -> out[slot]
__global__ __launch_bounds__(RB) void k_reduce(const double *partial, int n, double *out, int slot)
{
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += RB) acc += partial[i];
    __shared__ double red[RB / 64];
    double t = block_sum(acc, red);
    if (threadIdx.x == 0) out[slot] = t;
}

// copy moved particles between vectors (accept: trial -> p, reject: p -> trial)
__device__ __forceinline__ int4 pack_particle(double x, double y, double z, float q, const Box &b)
{
    return make_int4(to_fixed(x, b.inv[0]), to_fixed(y, b.inv[1]), to_fixed(z, b.inv[2]),
                     __float_as_int(q));
}

__global__ void k_copy_moved(Vec64 dst, Vec64 src, const int *idx, int nm, int front, int4 *pk,
                             const float *qf, Box b, int update_fixed)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nm) return;
    int m = idx ? idx[k] : front + k;
    dst.x[m] = src.x[m]; dst.y[m] = src.y[m]; dst.z[m] = src.z[m];
    if (update_fixed) pk[m] = pack_particle(src.x[m], src.y[m], src.z[m], qf[m], b);
}

// coordinates between the host's xyz triples and the device's per-axis arrays
__global__ void k_to_triples(Vec64 v, int n, double *out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[3 * i] = v.x[i]; out[3 * i + 1] = v.y[i]; out[3 * i + 2] = v.z[i];
}

__global__ void k_from_triples(Vec64 v, int first, int count, const double *in)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    v.x[first + i] = in[3 * i]; v.y[first + i] = in[3 * i + 1]; v.z[first + i] = in[3 * i + 2];
}

__global__ void k_fixed_all(const double *x, const double *y, const double *z, const float *qf,
                            int n, Box b, int4 *pk)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pk[i] = pack_particle(x[i], y[i], z[i], qf[i], b);
}

// ------------------------------------------------------------------------
// Ewald reciprocal space
// ------------------------------------------------------------------------

// updateAllComplexNumbers (ewald.h:388-415): one block per k-vector
__global__ __launch_bounds__(RB) void k_ewald_full(const double *kv, int K, PV v, int n,
                                                   const double *eff, double *Qdst)
{
    int k = blockIdx.x;
    if (k >= K) return;
    double k0 = kv[3 * k], k1 = kv[3 * k + 1], k2 = kv[3 * k + 2];
    double re = 0.0, im = 0.0;
    for (int i = threadIdx.x; i < n; i += RB) {
        double dot = k0 * v.x[i] + (k1 * v.y[i] + k2 * v.z[i]);
        double s, c;
        sincos(dot, &s, &c);
        double w = v.q[i] * eff[v.t[i]];
        re += c * w;
        im += s * w;
    }
    __shared__ double red[RB / 64];
    re = block_sum(re, red);
    im = block_sum(im, red);
    if (threadIdx.x == 0) { Qdst[2 * k] = re; Qdst[2 * k + 1] = im; }
}

// updateChange (ewald.h:651-705): Qt[k] = Q[k] + sum_moved q eff (e^{ik.r'} - e^{ik.r})
__global__ void k_ewald_change(const double *kv, int K, PV pp, PV pt, const int *idx, int nm,
                               int front, const double *eff, const double *Q, double *Qt)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    double k0 = kv[3 * k], k1 = kv[3 * k + 1], k2 = kv[3 * k + 2];
    double re = Q[2 * k], im = Q[2 * k + 1];
    for (int t = 0; t < nm; ++t) {
        int i = idx ? idx[t] : front + t;
        double dT = (k0 * pt.x[i] + k1 * pt.y[i]) + k2 * pt.z[i];
        double dO = (k0 * pp.x[i] + k1 * pp.y[i]) + k2 * pp.z[i];
        double sT, cT, sO, cO;
        sincos(dT, &sT, &cT);
        sincos(dO, &sO, &cO);
        double w = pp.q[i] * eff[pp.t[i]];
        re += cT * w;
        im += sT * w;
        re -= cO * w;
        im -= sO * w;
    }
    Qt[2 * k] = re;
    Qt[2 * k + 1] = im;
}

// sum_k A_k |Q_k|^2 (getReciprocalEnergy, ewald.h:298-312) and, when Qt is
// given, sum_k A_k (|Qt_k|^2 - |Q_k|^2) evaluated per k (no cancellation
// between two large sums)
__global__ __launch_bounds__(RB) void k_ewald_energy(const double *aks, int K, const double *Q,
                                                     const double *Qt, double *out, int slot)
{
    double e = 0.0;
    for (int k = threadIdx.x; k < K; k += RB) {
        double a = Q[2 * k], c = Q[2 * k + 1];
        if (Qt) {
            double at = Qt[2 * k], ct = Qt[2 * k + 1];
            e += aks[k] * ((at - a) * (at + a) + (ct - c) * (ct + c));
        } else {
            e += aks[k] * (a * a + c * c);
        }
    }
    __shared__ double red[RB / 64];
    e = block_sum(e, red);
    if (threadIdx.x == 0) out[slot] = e;
}

// M = sum_i q_i eff_i r_i (getSurfaceEnergy, ewald.h:278-292) and
// self = sum_i q_i^2 eff_i^2 over [f, b] (getSelfEnergy, ewald.h:260-272)
__global__ __launch_bounds__(RB) void k_ewald_moments(PV v, int f, int b, const double *eff,
                                                      double *out, int slot)
{
    double mx = 0, my = 0, mz = 0, sq = 0;
    for (int i = f + threadIdx.x; i <= b; i += RB) {
        double e = eff[v.t[i]];
        double w = v.q[i] * e;
        mx += v.q[i] * v.x[i] * e;
        my += v.q[i] * v.y[i] * e;
        mz += v.q[i] * v.z[i] * e;
        sq += v.q[i] * v.q[i] * e * e;
        (void)w;
    }
    __shared__ double red[RB / 64];
    mx = block_sum(mx, red);
    my = block_sum(my, red);
    mz = block_sum(mz, red);
    sq = block_sum(sq, red);
    if (threadIdx.x == 0) {
        out[slot] = mx; out[slot + 1] = my; out[slot + 2] = mz; out[slot + 3] = sq;
    }
}

// ------------------------------------------------------------------------
// host helpers
// ------------------------------------------------------------------------

int ensure_partials(fgx_system *s, int nblocks)
{
    if (nblocks <= s->partial_cap) return FGX_OK;
    if (s->d_partial) (void)hipFree(s->d_partial);
    s->d_partial = nullptr;
    int cap = std::max(nblocks, 1024);
    FGX_HIP(hipMalloc(&s->d_partial, sizeof(double) * cap));
    s->partial_cap = cap;
    return FGX_OK;
}

static int ensure_idx(fgx_system *s, int n)
{
    if (n <= s->idx_cap) return FGX_OK;
    if (s->d_idx) (void)hipFree(s->d_idx);
    s->d_idx = nullptr;
    int cap = std::max(n, 256);
    FGX_HIP(hipMalloc(&s->d_idx, sizeof(int) * cap));
    s->idx_cap = cap;
    return FGX_OK;
}

// fixed-order sum of `nblocks` partials into out slot
static int finish_sum(fgx_system *s, int nblocks, int slot, const double *partial = nullptr)
{
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(RB), 0, s->stream,
                       partial ? partial : s->d_partial, nblocks, s->d_out, slot);
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

// Where a pair kernel writes its per-block partial sums: the system's own
// buffer, grown on demand, or a fixed buffer of `cap` doubles that is never
// grown (the allocation-free entry point's scratch).
struct FixedPartials {
    double *buf;
    long cap;
};

static int partials_for(fgx_system *s, int nblocks, const FixedPartials *fixed, double **out)
{
    if (fixed) {
        FGX_CHECK(nblocks <= fixed->cap, FGX_ENOMEM,
                  "allocation-free call: the system needs more partial sums than the fixed "
                  "scratch holds (particle count too large)");
        *out = fixed->buf;
        return FGX_OK;
    }
    int rc = ensure_partials(s, nblocks);
    if (rc) return rc;
    *out = s->d_partial;
    return FGX_OK;
}

static int fetch(fgx_system *s, int nslots)
{
    FGX_HIP(hipMemcpyAsync(s->h_out, s->d_out, sizeof(double) * nslots, hipMemcpyDeviceToHost,
                           s->stream));
    FGX_HIP(hipStreamSynchronize(s->stream));
    return FGX_OK;
}

static int grid_for(int n) { return std::max(1, std::min((n + RB - 1) / RB, 4096)); }

// sum over a row: A = particle ia of `which`, j in [j0, j1) except skip
static int row_sum(fgx_system *s, int which, int ia, int j0, int j1, int skip, int slot)
{
    int nb = grid_for(j1 - j0);
    int rc = ensure_partials(s, nb);
    if (rc) return rc;
    if (j1 <= j0) {
        FGX_HIP(hipMemsetAsync(s->d_partial, 0, sizeof(double), s->stream));
        nb = 1;
    } else {
        hipLaunchKernelGGL(k_row, dim3(nb), dim3(RB), 0, s->stream, s->box, s->pot, s->d_tab,
                           pv(s, which), ia, j0, j1, skip, s->d_partial);
        FGX_HIP(hipGetLastError());
    }
    return finish_sum(s, nb, slot);
}

static int block_sum_launch(fgx_system *s, int which, int a0, int a1, int b0, int b1, int tri,
                            const unsigned char *mask, int slot,
                            const FixedPartials *fixed = nullptr)
{
    if (a1 <= a0 || b1 <= b0) {
        FGX_HIP(hipMemsetAsync(s->d_out + slot, 0, sizeof(double), s->stream));
        return FGX_OK;
    }
    dim3 g((b1 - b0 + RB - 1) / RB, (a1 - a0 + TILE - 1) / TILE);
    int nb = g.x * g.y;
    double *partial = nullptr;
    int rc = partials_for(s, nb, fixed, &partial);
    if (rc) return rc;
    hipLaunchKernelGGL(k_block, g, dim3(RB), 0, s->stream, s->box, s->pot, s->d_tab,
                       pv(s, which), a0, a1, b0, b1, tri, mask, s->ng, partial);
    FGX_HIP(hipGetLastError());
    return finish_sum(s, nb, slot, partial);
}

int ewald_full_recompute(fgx_system *s, int which, double *Qdst)
{
    if (!s->has_ewald) return FGX_OK;
    hipLaunchKernelGGL(k_ewald_full, dim3(s->ew.K), dim3(RB), 0, s->stream, s->ew.kv, s->ew.K,
                       pv(s, which), s->n, s->ew.eff, Qdst);
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

// getReciprocalEnergy for Q (device) -> *out_host
int ewald_reciprocal(fgx_system *s, const double *Q, double V, double *out_host)
{
    hipLaunchKernelGGL(k_ewald_energy, dim3(1), dim3(RB), 0, s->stream, s->ew.aks, s->ew.K, Q,
                       (const double *)nullptr, s->d_out, 0);
    FGX_HIP(hipGetLastError());
    int rc = fetch(s, 1);
    if (rc) return rc;
    const double pi = std::acos(-1.);
    *out_host = (2 * pi / V) * s->h_out[0] * s->pot.lB;
    return FGX_OK;
}

static int ewald_surface(fgx_system *s, int which, double V, double *out)
{
    if (s->ew.const_inf < 0.5) { *out = 0.0; return FGX_OK; }
    hipLaunchKernelGGL(k_ewald_moments, dim3(1), dim3(RB), 0, s->stream, pv(s, which), 0,
                       s->n - 1, s->ew.eff, s->d_out, 0);
    FGX_HIP(hipGetLastError());
    int rc = fetch(s, 4);
    if (rc) return rc;
    double *m = s->h_out;
    double qq = m[0] * m[0] + (m[1] * m[1] + m[2] * m[2]);
    const double pi = std::acos(-1.);
    *out = s->ew.const_inf * (2 * pi / ((2 * s->ew.eps_surf + 1) * V)) * qq * s->pot.lB;
    return FGX_OK;
}

static int ewald_self(fgx_system *s, int which, int f, int b, double *out)
{
    hipLaunchKernelGGL(k_ewald_moments, dim3(1), dim3(RB), 0, s->stream, pv(s, which), f, b,
                       s->ew.eff, s->d_out, 0);
    FGX_HIP(hipGetLastError());
    int rc = fetch(s, 4);
    if (rc) return rc;
    const double pi = std::acos(-1.);
    *out = (-s->ew.alpha * s->h_out[3] / std::sqrt(pi)) * s->pot.lB;
    return FGX_OK;
}

void mt_seed_host(uint32_t *st, uint32_t seed)
{
    st[0] = seed;
    for (int i = 1; i < 624; ++i) st[i] = 1812433253u * (st[i - 1] ^ (st[i - 1] >> 30)) + (uint32_t)i;
    st[624] = 624;
}

static void mass_center_host(const fgx_system *s, const std::vector<double> &xyz, int g,
                             double *cm)
{   // Geometry::massCenter / anyCenter, geometry.h:713-740
    const GroupDesc &G = s->groups[g];
    cm[0] = cm[1] = cm[2] = 0;
    if (G.back < G.front) return;
    int io = (int)(G.front + (G.back - G.front) * 0.5);
    double o[3] = {xyz[3 * io], xyz[3 * io + 1], xyz[3 * io + 2]};
    double sum = 0;
    for (int i = G.front; i <= G.back; ++i) {
        double t[3] = {xyz[3 * i] - o[0], xyz[3 * i + 1] - o[1], xyz[3 * i + 2] - o[2]};
        for (int k = 0; k < 3; ++k)
            if (std::fabs(t[k]) > s->box.half[k])
                t[k] -= s->box.len[k] * (int)(t[k] * s->box.inv[k] > 0. ? t[k] * s->box.inv[k] + .5
                                                                       : t[k] * s->box.inv[k] - .5);
        double w = s->h_mw[i];
        for (int k = 0; k < 3; ++k) cm[k] += t[k] * w;
        sum += w;
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

static double sqdist_host(const Box &b, const double *a, const double *c)
{
    double d[3];
    for (int k = 0; k < 3; ++k) {
        d[k] = std::fabs(a[k] - c[k]);
        if (d[k] > b.half[k]) d[k] -= 2 * b.half[k];
    }
    return d[0] * d[0] + (d[1] * d[1] + d[2] * d[2]);
}

static bool molecular(const GroupDesc &g) { return g.molsize == g.back - g.front + 1; }
static bool atomic(const GroupDesc &g) { return g.molsize == 1; }

// NonbondedCutg2g::cut (energy.h:1343-1358)
static bool cut_host(const fgx_system *s, int g1, bool trial1, int g2, bool trial2)
{
    if (s->pot.nbkind != NB_CUTG2G) return false;
    if (!molecular(s->groups[g1]) || !molecular(s->groups[g2])) return false;
    const double *a = trial1 ? &s->cm_trial[3 * g1] : &s->cm[3 * g1];
    const double *b = trial2 ? &s->cm_trial[3 * g2] : &s->cm[3 * g2];
    return sqdist_host(s->box, a, b) > s->pot.g2g_rc2;
}

}  // namespace fgx

using namespace fgx;

// ==========================================================================
// C ABI
// ==========================================================================

extern "C" const char *fgx_last_error(void) { return fgx::g_err.c_str(); }

extern "C" int fgx_device_info(char *name, int cap, int *cu_count, size_t *hbm_bytes)
{
    int dev = 0;
    FGX_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    FGX_HIP(hipGetDeviceProperties(&p, dev));
    if (name && cap > 0) {
        std::strncpy(name, p.gcnArchName, (size_t)cap - 1);
        name[cap - 1] = 0;
    }
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = p.totalGlobalMem;
    return FGX_OK;
}

template <class T>
static int dalloc(fgx_system *s, T **p, size_t count)
{
    FGX_HIP(hipMalloc((void **)p, sizeof(T) * std::max<size_t>(count, 1)));
    s->device_bytes += sizeof(T) * std::max<size_t>(count, 1);
    return FGX_OK;
}

template <class T>
static int upload(fgx_system *s, T *dst, const T *src, size_t count)
{
    if (count == 0) return FGX_OK;
    FGX_HIP(hipMemcpyAsync(dst, src, sizeof(T) * count, hipMemcpyHostToDevice, s->stream));
    return FGX_OK;
}

// k-vectors 2 pi n / L and A_k = f exp(-(k^2 + kappa^2) / 4 alpha^2) / (k^2 +
// kappa^2) of the index triples h_kn for the box `L` (kVectorChange,
// ewald.h:330-378).  Which triples enter depends on kc only: every nonzero
// triple has k^2 above check_k2_zero = 0.1 (2 pi / max L)^2 for any box, so a
// box change keeps the set and rescales the vectors and A_k.
static void ewald_tables(const fgx_system *s, const double *L, std::vector<double> &kv,
                         std::vector<double> &aks)
{
    const EwaldDev &e = s->ew;
    const double pi = std::acos(-1.), twopi = 2 * pi;
    const double alpha2 = e.alpha * e.alpha;
    kv.resize(3 * (size_t)e.K);
    aks.resize(e.K);
    for (int k = 0; k < e.K; ++k) {
        const int kx = s->h_kn[3 * k], ky = s->h_kn[3 * k + 1], kz = s->h_kn[3 * k + 2];
        double factor = (kx > 0) ? 2.0 : 1.0;
        double v[3] = {twopi * (kx / L[0]), twopi * (ky / L[1]), twopi * (kz / L[2])};
        double k2 = v[0] * v[0] + (v[1] * v[1] + v[2] * v[2]);
        for (int a = 0; a < 3; ++a) kv[3 * k + a] = v[a];
        aks[k] = factor * std::exp(-(k2 + e.kappa * e.kappa) / (4.0 * alpha2)) /
                 (k2 + e.kappa * e.kappa);
    }
}

static int build_ewald(fgx_system *s, const fgx_hamiltonian *h)
{   // NonbondedEwald ctor + kVectorChange (ewald.h:507-550, 330-378)
    EwaldDev &e = s->ew;
    e.alpha = h->ewald_alpha;
    e.kappa = 1.0 / h->ewald_debyelength;
    e.eps_surf = h->ewald_eps_surf;
    e.const_inf = (e.eps_surf < 1) ? 0.0 : 1.0;
    e.update_frequency = h->ewald_update_frequency;
    e.cnt_accepted = 0;
    const double kc = h->ewald_kc, kc2 = kc * kc;
    e.kcc = (int)std::ceil(kc);
    const double pi = std::acos(-1.);
    const double *L = s->box.len;
    double maxL = std::max(L[0], std::max(L[1], L[2]));
    double check_k2_zero = 0.1 * (4 * pi * pi) / (maxL * maxL);
    std::vector<int> kn;
    for (int kx = 0; kx <= e.kcc; kx++) {
        double dkx2 = (double)(kx * kx);
        for (int ky = -e.kcc; ky <= e.kcc; ky++) {
            double dky2 = (double)(ky * ky);
            for (int kz = -e.kcc; kz <= e.kcc; kz++) {
                double dkz2 = (double)(kz * kz);
                double twopi = 2 * pi;
                double v[3] = {twopi * (kx / L[0]), twopi * (ky / L[1]), twopi * (kz / L[2])};
                double k2 = v[0] * v[0] + (v[1] * v[1] + v[2] * v[2]);
                if (k2 < check_k2_zero) continue;
                if (h->ewald_spherical_sum)
                    if ((dkx2 / kc2) + (dky2 / kc2) + (dkz2 / kc2) > 1.0) continue;
                kn.push_back(kx); kn.push_back(ky); kn.push_back(kz);
            }
        }
    }
    FGX_CHECK(!kn.empty(), FGX_EUNSUPPORTED, "Ewald: no k-vectors (cutoffK too small)");
    e.K = (int)kn.size() / 3;
    s->h_kn = kn;
    std::vector<double> kv, aks;
    ewald_tables(s, L, kv, aks);
    int rc;
    if ((rc = dalloc(s, &e.kv, kv.size()))) return rc;
    if ((rc = dalloc(s, &e.aks, aks.size()))) return rc;
    if ((rc = dalloc(s, &e.kn, kn.size()))) return rc;
    if ((rc = dalloc(s, &e.Q, 2 * (size_t)e.K))) return rc;
    if ((rc = dalloc(s, &e.Qt, 2 * (size_t)e.K))) return rc;
    if ((rc = dalloc(s, &e.eff, s->h_eff.size()))) return rc;
    if ((rc = upload(s, e.kv, kv.data(), kv.size()))) return rc;
    if ((rc = upload(s, e.aks, aks.data(), aks.size()))) return rc;
    if ((rc = upload(s, e.kn, kn.data(), kn.size()))) return rc;
    if ((rc = upload(s, e.eff, s->h_eff.data(), s->h_eff.size()))) return rc;
    e.V = L[0] * L[1] * L[2];
    return FGX_OK;
}

// NonbondedEwald::setSpace (ewald.h:750-757): recompute S(k) from p, trial = current
int fgx::ewald_set_space(fgx_system *s)
{
    if (!s->has_ewald) return FGX_OK;
    int rc = ewald_full_recompute(s, 0, s->ew.Q);
    if (rc) return rc;
    FGX_HIP(hipMemcpyAsync(s->ew.Qt, s->ew.Q, sizeof(double) * 2 * s->ew.K,
                           hipMemcpyDeviceToDevice, s->stream));
    return FGX_OK;
}

// NonbondedEwald::setGeometry (ewald.h:731-746) for a box `len`: k-vectors,
// A_k and V of that box.  S(k) are left as they are (setSpace recomputes them).
static int ewald_rebox(fgx_system *s, const double len[3])
{
    if (!s->has_ewald) return FGX_OK;
    std::vector<double> kv, aks;
    ewald_tables(s, len, kv, aks);
    int rc;
    if ((rc = upload(s, s->ew.kv, kv.data(), kv.size()))) return rc;
    if ((rc = upload(s, s->ew.aks, aks.data(), aks.size()))) return rc;
    FGX_HIP(hipStreamSynchronize(s->stream));
    s->ew.V = len[0] * len[1] * len[2];
    return FGX_OK;
}

// NonbondedEwald::update(true) (ewald.h:626-648): the trial S(k) become the
// current ones; every update_frequency accepted moves they are recomputed
// from the (now accepted) trial vector and the change of the reciprocal
// energy is returned in *drift
int fgx::ewald_accepted(fgx_system *s, double *drift)
{
    EwaldDev &e = s->ew;
    *drift = 0.0;
    if (++e.cnt_accepted > e.update_frequency - 1) {
        double duB, duA;
        int rc;
        if ((rc = ewald_reciprocal(s, e.Qt, e.V, &duB))) return rc;
        if ((rc = ewald_full_recompute(s, 1, e.Qt))) return rc;
        if ((rc = ewald_reciprocal(s, e.Qt, e.V, &duA))) return rc;
        FGX_HIP(hipMemcpyAsync(e.Q, e.Qt, sizeof(double) * 2 * e.K, hipMemcpyDeviceToDevice,
                               s->stream));
        e.cnt_accepted = 0;
        e.update_drift += std::fabs(duA - duB);
        *drift = duA - duB;
    } else {
        FGX_HIP(hipMemcpyAsync(e.Q, e.Qt, sizeof(double) * 2 * e.K, hipMemcpyDeviceToDevice,
                               s->stream));
    }
    return FGX_OK;
}

// Chebyshev coefficients of erfcx(x) = exp(x^2) erfc(x) on [0, xmax], xmax =
// alpha rc (the largest real-space argument), from long-double samples at the
// Chebyshev nodes; the series is cut where the remaining coefficients are
// below 1e-17 (erfcx is entire, so they fall off faster than geometrically).
// Used by the mixed path's FP64 pass: erfc(x) = exp(-x^2) * series(x).
static int build_erfc_series(fgx_system *s)
{
    const int M = 96;
    const long double xmax = (long double)s->pot.ew_alpha * std::sqrt((long double)s->pot.ew_rc2) *
                             (1.0L + 1e-9L);
    const long double pi = 3.14159265358979323846264338327950288L;
    std::vector<long double> f(M), c(M);
    for (int k = 0; k < M; ++k) {
        const long double t = std::cos(pi * (k + 0.5L) / M);
        const long double x = 0.5L * (t + 1.0L) * xmax;
        f[k] = std::exp(x * x) * std::erfc(x);
    }
    for (int j = 0; j < M; ++j) {
        long double a = 0;
        for (int k = 0; k < M; ++k) a += f[k] * std::cos(pi * j * (k + 0.5L) / M);
        c[j] = 2.0L * a / M;
    }
    c[0] *= 0.5L;
    int n = 0;
    for (int j = M - 1; j >= 0; --j)
        if (std::fabs(c[j]) > 1e-17L) { n = j + 1; break; }
    if (n > EW_CHEB_MAX || n == 0) return FGX_OK;   // series not used: generic erfc stays
    std::vector<double> cd(n);
    for (int j = 0; j < n; ++j) cd[j] = (double)c[j];
    int rc;
    if ((rc = dalloc(s, &s->ew.cheb, n)) || (rc = upload(s, s->ew.cheb, cd.data(), n))) return rc;
    s->ew.ncheb = n;
    s->ew.cheb_xmax = (double)xmax;
    return FGX_OK;
}

extern "C" int fgx_create(const fgx_hamiltonian *h, const double box[3], int n, const double *xyz,
                          const double *charge, const double *radius, const double *mw,
                          const int *type, int ng, const int *front, const int *back,
                          const int *molsize, const int *molid, fgx_system **out)
{
    FGX_CHECK(h && box && out && n >= 0 && ng >= 0, FGX_EARG, "fgx_create: bad arguments");
    FGX_CHECK(n == 0 || (xyz && charge && radius && mw && type), FGX_EARG,
              "fgx_create: particle arrays missing");
    FGX_CHECK(h->ntypes > 0 && h->ntypes <= 256, FGX_EARG, "fgx_create: ntypes must be 1..256");
    FGX_CHECK(h->elec >= 0 && h->elec <= 3 && h->sr >= 0 && h->sr <= 2 && h->nbkind >= 0 &&
                  h->nbkind <= 2,
              FGX_EARG, "fgx_create: unknown potential kind");
    for (int k = 0; k < 3; ++k)
        FGX_CHECK(box[k] > 0, FGX_EARG, "fgx_create: box lengths must be positive");
    for (int i = 0; i < n; ++i)
        FGX_CHECK(type[i] >= 0 && type[i] < h->ntypes, FGX_EARG,
                  "fgx_create: particle type out of range");
    for (int g = 0; g < ng; ++g)
        FGX_CHECK(front[g] >= 0 && back[g] < n && back[g] >= front[g] - 1 && molsize[g] != 0,
                  FGX_EARG, "fgx_create: bad group range");

    fgx_system *s = new fgx_system();
    *out = nullptr;
    auto fail = [&](int rc) {
        fgx_destroy(s);
        return rc;
    };
    if (hipGetDevice(&s->device) != hipSuccess) return fail(FGX_EHIP);
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed");
        return fail(FGX_EHIP);
    }
    s->n = n;
    s->ng = ng;
    s->ntypes = h->ntypes;
    // Cuboid::setlen
    for (int k = 0; k < 3; ++k) {
        s->box.len[k] = box[k];
        s->box.half[k] = box[k] * 0.5;
        s->box.inv[k] = 1 / box[k];
        s->box.lenf[k] = (float)box[k];
        s->box.fscale[k] = (float)(box[k] / 4294967296.0);
    }
    Pot &P = s->pot;
    P.elec = h->elec; P.sr = h->sr; P.nbkind = h->nbkind; P.precision = h->precision;
    P.ntypes = h->ntypes; P.ntypes2 = h->ntypes * h->ntypes;
    P.lB = h->lB; P.kappa = h->kappa;
    P.g2g_rc2 = h->g2g_cut > 0 ? h->g2g_cut * h->g2g_cut : HUGE_VAL;
    P.lBf = (float)h->lB; P.kappaf = (float)h->kappa;
    P.exact_f = 0.5f;
    // CutShift: rc2 = pow(cutoff, 2) (potentials.h:1786)
    FGX_CHECK(h->lj_cut <= 0 || h->sr == SR_LJ, FGX_EARG,
              "fgx_create: lj_cut needs the Lennard-Jones short-range term");
    P.lj_rc2 = h->lj_cut > 0 ? std::pow(h->lj_cut, 2) : 0.0;
    P.lj_rc2f = (float)P.lj_rc2;
    FGX_CHECK(!h->has_pressure || h->pressure >= 0, FGX_EARG,
              "fgx_create: negative pressure (energy.h:1871-1872)");
    s->press_on = h->has_pressure != 0;
    s->press = h->pressure;
    // tables: LennardJonesMixed<LorentzBerthelot>::init (potentials.h:2010-2019)
    const int T = h->ntypes;
    s->tab.assign((size_t)TAB_COUNT * T * T, 0.0);
    s->type_dp.assign(T, 0.0);
    s->type_charge.assign(T, 0.0);
    s->type_radius.assign(T, 0.0);
    for (int a = 0; a < T; ++a) {
        if (h->type_dp) s->type_dp[a] = h->type_dp[a];
        if (h->type_charge) s->type_charge[a] = h->type_charge[a];
        if (h->type_radius) s->type_radius[a] = h->type_radius[a];
        for (int b = 0; b < T; ++b) {
            double sa = h->type_sigma ? h->type_sigma[a] : 0.0;
            double sb = h->type_sigma ? h->type_sigma[b] : 0.0;
            double ea = h->type_eps ? h->type_eps[a] : 0.0;
            double eb = h->type_eps ? h->type_eps[b] : 0.0;
            double sigma = 0.5 * (sa + sb), epsilon = std::sqrt(ea * eb);
            s->tab[TAB_S2 * T * T + a * T + b] = sigma * sigma;
            s->tab[TAB_EPS4 * T * T + a * T + b] = 4 * epsilon;
        }
    }
    if (h->elec == ELEC_EWALD) {
        FGX_CHECK(h->type_charge && h->type_radius, FGX_EARG,
                  "fgx_create: Ewald needs per-type charge and radius");
        FGX_CHECK(h->ewald_debyelength > 0 && h->ewald_alpha > 0 && h->ewald_rc > 0, FGX_EARG,
                  "fgx_create: Ewald needs alpha, cutoff and debyelength (ewald.h:521)");
        double kappa = 1.0 / h->ewald_debyelength;
        s->h_eff.assign(T, 0.0);
        // effective charges (ewald.h:535-548)
        for (int a = 0; a < T; ++a) {
            double qi = h->type_charge[a], ri = h->type_radius[a];
            double qis = qi;
            if (std::fabs(kappa * ri) > 1e-9) qis *= std::sinh(ri * kappa) / ri / kappa;
            s->h_eff[a] = qis / qi;
            for (int b = 0; b < T; ++b) {
                double qj = h->type_charge[b], rj = h->type_radius[b];
                double qjs = qj;
                if (std::fabs(kappa * rj) > 1e-9) qjs *= std::sinh(rj * kappa) / rj / kappa;
                s->tab[TAB_QEQE * T * T + a * T + b] = qis * qjs / qi / qj;
            }
        }
        P.ew_alpha = h->ewald_alpha;
        P.ew_kappa = kappa;
        P.ew_rc2 = h->ewald_rc * h->ewald_rc;
        P.ew_alphaf = (float)P.ew_alpha;
        P.ew_kappaf = (float)kappa;
        P.ew_rc2f = (float)P.ew_rc2;
        P.ew_sf = (float)(kappa / 2.0 / P.ew_alpha);
        P.ew_plain = (kappa * h->ewald_rc < 1e-12) ? 1 : 0;
        s->has_ewald = true;
    }
    // host mirrors
    s->groups.resize(ng);
    for (int g = 0; g < ng; ++g) s->groups[g] = GroupDesc{front[g], back[g], molsize[g], molid[g]};
    s->h_charge.assign(charge, charge + n);
    s->h_radius.assign(radius, radius + n);
    s->h_mw.assign(mw, mw + n);
    s->h_type.assign(type, type + n);
    s->h_gid.assign(n, -1);
    for (int g = 0; g < ng; ++g)
        for (int i = front[g]; i <= back[g]; ++i) s->h_gid[i] = g;
    for (int i = 0; i < n; ++i)
        if (s->h_gid[i] < 0) {
            set_error("fgx_create: every particle must belong to a group");
            return fail(FGX_EARG);
        }
    // device buffers
    int rc;
    for (int w = 0; w < 2; ++w) {
        if ((rc = dalloc(s, &s->v[w].x, n)) || (rc = dalloc(s, &s->v[w].y, n)) ||
            (rc = dalloc(s, &s->v[w].z, n)))
            return fail(rc);
    }
    if ((rc = dalloc(s, &s->d_q, n)) || (rc = dalloc(s, &s->d_r, n)) ||
        (rc = dalloc(s, &s->d_mw, n)) || (rc = dalloc(s, &s->d_type, n + PARTNER_PAD)) ||
        (rc = dalloc(s, &s->d_gid, n)) || (rc = dalloc(s, &s->d_pk, n + PARTNER_PAD)) ||
        (rc = dalloc(s, &s->d_qf, n)) || (rc = dalloc(s, &s->d_rf, n + PARTNER_PAD)) ||
        (rc = dalloc(s, &s->d_groups, ng)) || (rc = dalloc(s, &s->d_cm, 6 * (size_t)ng)) ||
        (rc = dalloc(s, &s->d_tab, s->tab.size())) || (rc = dalloc(s, &s->d_tabf, s->tab.size())) ||
        (rc = dalloc(s, &s->d_type_dp, T)) || (rc = dalloc(s, &s->d_mask, 2 * (size_t)ng + 2)) ||
        (rc = dalloc(s, &s->d_out, 16)) || (rc = dalloc(s, &s->d_rng, 2 * 625)) ||
        (rc = dalloc(s, &s->d_stats_i, 16)) || (rc = dalloc(s, &s->d_stats_d, 16)) ||
        (rc = dalloc(s, &s->d_resume, 4)))
        return fail(rc);
    if (hipHostMalloc((void **)&s->h_out, sizeof(double) * 16, hipHostMallocDefault) !=
        hipSuccess) {
        set_error("hipHostMalloc failed");
        return fail(FGX_EHIP);
    }
    if ((rc = ensure_partials(s, 1024))) return fail(rc);
    if (hipMemsetAsync(s->d_resume, 0, sizeof(int) * 4, s->stream) != hipSuccess ||
        hipMemsetAsync(s->d_type + n, 0, sizeof(int) * PARTNER_PAD, s->stream) != hipSuccess ||
        hipMemsetAsync(s->d_pk + n, 0, sizeof(int4) * PARTNER_PAD, s->stream) != hipSuccess ||
        hipMemsetAsync(s->d_rf + n, 0, sizeof(float) * PARTNER_PAD, s->stream) != hipSuccess) {
        set_error("fgx_create: hipMemsetAsync failed");
        return fail(FGX_EHIP);
    }
    std::vector<double> tmp(n);
    for (int k = 0; k < 3; ++k) {
        for (int i = 0; i < n; ++i) tmp[i] = xyz[3 * i + k];
        double *dp = k == 0 ? s->v[0].x : k == 1 ? s->v[0].y : s->v[0].z;
        double *dt = k == 0 ? s->v[1].x : k == 1 ? s->v[1].y : s->v[1].z;
        if ((rc = upload(s, dp, tmp.data(), n)) || (rc = upload(s, dt, tmp.data(), n)))
            return fail(rc);
        if (hipStreamSynchronize(s->stream) != hipSuccess) return fail(FGX_EHIP);
    }
    std::vector<float> qf(n), rf(n), tabf(s->tab.size());
    for (int i = 0; i < n; ++i) { qf[i] = (float)charge[i]; rf[i] = (float)radius[i]; }
    for (size_t k = 0; k < s->tab.size(); ++k) tabf[k] = (float)s->tab[k];
    if ((rc = upload(s, s->d_q, charge, n)) || (rc = upload(s, s->d_r, radius, n)) ||
        (rc = upload(s, s->d_mw, mw, n)) || (rc = upload(s, s->d_type, type, n)) ||
        (rc = upload(s, s->d_gid, s->h_gid.data(), n)) || (rc = upload(s, s->d_qf, qf.data(), n)) ||
        (rc = upload(s, s->d_rf, rf.data(), n)) ||
        (rc = upload(s, s->d_groups, s->groups.data(), ng)) ||
        (rc = upload(s, s->d_tab, s->tab.data(), s->tab.size())) ||
        (rc = upload(s, s->d_tabf, tabf.data(), tabf.size())) ||
        (rc = upload(s, s->d_type_dp, s->type_dp.data(), T)))
        return fail(rc);
    if (n > 0) {
        hipLaunchKernelGGL(k_fixed_all, dim3((n + 255) / 256), dim3(256), 0, s->stream,
                           s->v[0].x, s->v[0].y, s->v[0].z, s->d_qf, n, s->box, s->d_pk);
    }
    // mass centres (Space::load -> Group::setMassCenter, space.h:850)
    std::vector<double> hx(xyz, xyz + 3 * (size_t)n);
    s->cm.assign(3 * (size_t)ng, 0.0);
    for (int g = 0; g < ng; ++g) mass_center_host(s, hx, g, &s->cm[3 * g]);
    s->cm_trial = s->cm;
    std::vector<double> cm6(6 * (size_t)ng);
    for (int g = 0; g < ng; ++g)
        for (int k = 0; k < 3; ++k) { cm6[6 * g + k] = s->cm[3 * g + k]; cm6[6 * g + 3 + k] = s->cm[3 * g + k]; }
    if ((rc = upload(s, s->d_cm, cm6.data(), cm6.size()))) return fail(rc);
    if (s->has_ewald) {
        if ((rc = build_ewald(s, h))) return fail(rc);
        if ((rc = ewald_set_space(s))) return fail(rc);
        if (P.ew_plain && (rc = build_erfc_series(s))) return fail(rc);
    }
    mt_seed_host(s->rng_global, 5489u);
    mt_seed_host(s->rng_internal, 5489u);
    if (hipStreamSynchronize(s->stream) != hipSuccess) {
        set_error("fgx_create: device error during upload");
        return fail(FGX_EHIP);
    }
    *out = s;
    return FGX_OK;
}

extern "C" int fgx_destroy(fgx_system *s)
{
    if (!s) return FGX_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    void *ptrs[] = {s->v[0].x, s->v[0].y, s->v[0].z, s->v[1].x, s->v[1].y, s->v[1].z, s->d_q,
                    s->d_r, s->d_mw, s->d_type, s->d_gid, s->d_pk, s->d_qf,
                    s->d_rf, s->d_groups, s->d_cm, s->d_tab, s->d_tabf, s->d_type_dp, s->d_mask,
                    s->d_partial, s->d_out, s->d_idx, s->d_rng, s->d_stats_i, s->d_stats_d,
                    s->ew.kv, s->ew.aks, s->ew.kn, s->ew.Q, s->ew.Qt, s->ew.eff,
                    s->ew.cnt_dev, s->d_emat, s->ew.cheb, s->ew.sk_dS, s->ew.sk_kid,
                    s->d_gvec, s->d_batch,
                    s->d_bstat_i, s->d_bstat_d, s->d_resume, s->d_cell_cnt, s->d_cell_rec,
                    s->d_cell_idx, s->d_cell_slot, s->d_stage, s->d_xbuf, s->d_xcnt,
                    s->d_xerr, s->d_blkmap, s->d_mirror};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (s->h_out) (void)hipHostFree(s->h_out);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return FGX_OK;
}

extern "C" size_t fgx_device_bytes(const fgx_system *s) { return s ? s->device_bytes : 0; }

// staging for host transfers of coordinates, 3n doubles, kept until destroy
static int ensure_stage(fgx_system *s)
{
    if (s->d_stage || s->n == 0) return FGX_OK;
    FGX_HIP(hipMalloc(&s->d_stage, sizeof(double) * 3 * (size_t)s->n));
    s->device_bytes += sizeof(double) * 3 * (size_t)s->n;
    return FGX_OK;
}

extern "C" int fgx_set_positions(fgx_system *s, int which, int first, int count, const double *xyz)
{
    FGX_CHECK(s && xyz && which >= 0 && which <= 2 && first >= 0 && count >= 0 &&
                  first + count <= s->n,
              FGX_EARG, "fgx_set_positions: bad arguments");
    if (which != 1) s->emat_valid = s->cell_valid = false;
    int rc;
    if (count > 0 && (rc = ensure_stage(s))) return rc;
    if (count > 0) {   // one copy of the triples, then a scatter per vector
        FGX_HIP(hipMemcpyAsync(s->d_stage, xyz, sizeof(double) * 3 * (size_t)count,
                               hipMemcpyHostToDevice, s->stream));
        for (int w = 0; w < 2; ++w) {
            if (which != 2 && which != w) continue;
            hipLaunchKernelGGL(k_from_triples, dim3((count + 255) / 256), dim3(256), 0, s->stream,
                               s->v[w], first, count, (const double *)s->d_stage);
            FGX_HIP(hipGetLastError());
        }
    }
    if (which != 1 && count > 0) {
        hipLaunchKernelGGL(k_fixed_all, dim3((s->n + 255) / 256), dim3(256), 0, s->stream,
                           s->v[0].x, s->v[0].y, s->v[0].z, s->d_qf, s->n, s->box, s->d_pk);
        FGX_HIP(hipGetLastError());
    }
    FGX_HIP(hipStreamSynchronize(s->stream));
    return FGX_OK;
}

extern "C" int fgx_get_positions(fgx_system *s, int which, double *xyz)
{
    FGX_CHECK(s && xyz && (which == 0 || which == 1), FGX_EARG, "fgx_get_positions: bad arguments");
    if (s->n == 0) return FGX_OK;
    int rc = ensure_stage(s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_to_triples, dim3((s->n + 255) / 256), dim3(256), 0, s->stream,
                       s->v[which], s->n, s->d_stage);
    FGX_HIP(hipGetLastError());
    FGX_HIP(hipMemcpyAsync(xyz, s->d_stage, sizeof(double) * 3 * (size_t)s->n,
                           hipMemcpyDeviceToHost, s->stream));
    FGX_HIP(hipStreamSynchronize(s->stream));
    return FGX_OK;
}

static int sync_cm_from_device(fgx_system *s)
{
    std::vector<double> cm6(6 * (size_t)s->ng);
    if (s->ng == 0) return FGX_OK;
    FGX_HIP(hipMemcpyAsync(cm6.data(), s->d_cm, sizeof(double) * cm6.size(), hipMemcpyDeviceToHost,
                           s->stream));
    FGX_HIP(hipStreamSynchronize(s->stream));
    for (int g = 0; g < s->ng; ++g)
        for (int k = 0; k < 3; ++k) {
            s->cm[3 * g + k] = cm6[6 * g + k];
            s->cm_trial[3 * g + k] = cm6[6 * g + 3 + k];
        }
    return FGX_OK;
}

static int push_cm(fgx_system *s)
{
    std::vector<double> cm6(6 * (size_t)s->ng);
    for (int g = 0; g < s->ng; ++g)
        for (int k = 0; k < 3; ++k) {
            cm6[6 * g + k] = s->cm[3 * g + k];
            cm6[6 * g + 3 + k] = s->cm_trial[3 * g + k];
        }
    int rc = upload(s, s->d_cm, cm6.data(), cm6.size());
    if (rc) return rc;
    FGX_HIP(hipStreamSynchronize(s->stream));
    return FGX_OK;
}

extern "C" int fgx_get_group_cm(fgx_system *s, int g, double cm[3], double cm_trial[3])
{
    FGX_CHECK(s && g >= 0 && g < s->ng, FGX_EARG, "fgx_get_group_cm: bad group");
    int rc = sync_cm_from_device(s);
    if (rc) return rc;
    for (int k = 0; k < 3; ++k) {
        if (cm) cm[k] = s->cm[3 * g + k];
        if (cm_trial) cm_trial[k] = s->cm_trial[3 * g + k];
    }
    return FGX_OK;
}

extern "C" int fgx_set_group_cm(fgx_system *s, int g, const double cm[3], const double cm_trial[3])
{
    FGX_CHECK(s && g >= 0 && g < s->ng, FGX_EARG, "fgx_set_group_cm: bad group");
    if (cm) s->emat_valid = false;
    // the device sweeps move mass centres on the device: start from those, so
    // the value not given here is kept as it is
    if (int rc = sync_cm_from_device(s)) return rc;
    for (int k = 0; k < 3; ++k) {
        if (cm) s->cm[3 * g + k] = cm[k];
        if (cm_trial) s->cm_trial[3 * g + k] = cm_trial[k];
    }
    return push_cm(s);
}

// ---------------------------------------------------------------- energies

static int check_which(fgx_system *s, int which) { return s && (which == 0 || which == 1); }

// ------------------------------------------------------------- analysis

static constexpr int RDF_TILE = 256;
static constexpr int RDF_MAXBINS = 8192;

// Analysis::AtomRDF::update (analysis.h:2119-2134): every pair i < j whose
// types are {ta, tb} adds one count at bin int(r / dr + 0.5) (Table2D::round,
// auxiliary.h:831), r the minimum-image distance (Cuboid::dist = sqrt of
// sqdist).  Block (bx, by) takes partners j of tile bx against rows i of
// tile by; counts go to an LDS histogram first, then to `hist`.
__global__ __launch_bounds__(RDF_TILE) void k_rdf(Box b, PV v, int n, int ta, int tb, double dr,
                                                  int nbins, unsigned long long *hist)
{
    __shared__ unsigned int lh[RDF_MAXBINS];
    __shared__ double sx[RDF_TILE], sy[RDF_TILE], sz[RDF_TILE];
    __shared__ int st[RDF_TILE];
    for (int k = threadIdx.x; k < nbins; k += blockDim.x) lh[k] = 0;
    const int i0 = blockIdx.y * RDF_TILE;
    const int ni = min(RDF_TILE, n - i0);
    if ((int)threadIdx.x < ni) {
        const int i = i0 + threadIdx.x;
        sx[threadIdx.x] = v.x[i]; sy[threadIdx.x] = v.y[i]; sz[threadIdx.x] = v.z[i];
        st[threadIdx.x] = v.t[i];
    }
    __syncthreads();
    const int j = blockIdx.x * RDF_TILE + threadIdx.x;
    if (j < n && blockIdx.x >= blockIdx.y) {
        const double bx = v.x[j], by = v.y[j], bz = v.z[j];
        const int bt = v.t[j];
        for (int k = 0; k < ni; ++k) {
            const int i = i0 + k;
            if (i >= j) break;
            const int at = st[k];
            if (!((at == ta && bt == tb) || (at == tb && bt == ta))) continue;
            const double r = sqrt(sqdist_d(b, sx[k], sy[k], sz[k], bx, by, bz));
            const int bin = (int)(r / dr + 0.5);
            if (bin < nbins) atomicAdd(&lh[bin], 1u);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nbins; k += blockDim.x)
        if (lh[k]) atomicAdd(&hist[k], (unsigned long long)lh[k]);
}

extern "C" int fgx_rdf(fgx_system *s, int which, int type_a, int type_b, double dr, int nbins,
                       unsigned long long *counts)
{
    FGX_CHECK(check_which(s, which) && counts && dr > 0 && nbins > 0 && nbins <= RDF_MAXBINS,
              FGX_EARG, "fgx_rdf: bad arguments");
    // every minimum-image distance must land in a bin
    const double rmax = 0.5 * std::sqrt(s->box.len[0] * s->box.len[0] +
                                        s->box.len[1] * s->box.len[1] +
                                        s->box.len[2] * s->box.len[2]);
    FGX_CHECK((int)(rmax / dr + 0.5) < nbins, FGX_EARG,
              "fgx_rdf: nbins too small for the box diagonal");
    unsigned long long *d = nullptr;
    FGX_HIP(hipMalloc(&d, sizeof(unsigned long long) * nbins));
    int rc = FGX_OK;
    if (hipMemsetAsync(d, 0, sizeof(unsigned long long) * nbins, s->stream) != hipSuccess) {
        set_error("fgx_rdf: hipMemsetAsync failed");
        rc = FGX_EHIP;
    }
    const int nt = (s->n + RDF_TILE - 1) / RDF_TILE;
    if (!rc && s->n > 1) {
        hipLaunchKernelGGL(k_rdf, dim3(nt, nt), dim3(RDF_TILE), 0, s->stream, s->box, pv(s, which),
                           s->n, type_a, type_b, dr, nbins, d);
        if (hipGetLastError() != hipSuccess) {
            set_error("fgx_rdf: launch failed");
            rc = FGX_EHIP;
        }
    }
    if (!rc && (hipMemcpyAsync(counts, d, sizeof(unsigned long long) * nbins, hipMemcpyDeviceToHost,
                               s->stream) != hipSuccess ||
                hipStreamSynchronize(s->stream) != hipSuccess)) {
        set_error("fgx_rdf: device error");
        rc = FGX_EHIP;
    }
    (void)hipFree(d);
    return rc;
}

extern "C" int fgx_i2i(fgx_system *s, int which, int i, int j, double *u)
{
    FGX_CHECK(check_which(s, which) && u && i >= 0 && j >= 0 && i < s->n && j < s->n, FGX_EARG,
              "fgx_i2i: bad arguments");
    int rc = row_sum(s, which, i, j, j + 1, -1, 0);
    if (rc) return rc;
    if ((rc = fetch(s, 1))) return rc;
    *u = s->h_out[0];
    return FGX_OK;
}

// Nonbonded::i2g (energy.h:847-865), with the NonbondedCutg2g cut
// (energy.h:1378-1382)
extern "C" int fgx_i2g(fgx_system *s, int which, int g, int j, double *u)
{
    FGX_CHECK(check_which(s, which) && u && g >= 0 && g < s->ng && j >= 0 && j < s->n, FGX_EARG,
              "fgx_i2g: bad arguments");
    const GroupDesc &G = s->groups[g];
    if (s->pot.nbkind == NB_CUTG2G) {
        int rc = sync_cm_from_device(s);
        if (rc) return rc;
        if (cut_host(s, g, which, s->h_gid[j], which)) { *u = 0.0; return FGX_OK; }
    }
    int rc = row_sum(s, which, j, G.front, G.back + 1, j, 0);
    if (rc) return rc;
    if ((rc = fetch(s, 1))) return rc;
    *u = s->h_out[0];
    return FGX_OK;
}

extern "C" int fgx_i2all(fgx_system *s, int which, int i, double *u)
{
    FGX_CHECK(check_which(s, which) && u && i >= 0 && i < s->n, FGX_EARG, "fgx_i2all: bad arguments");
    if (s->pot.nbkind == NB_CUTG2G) {
        // NonbondedCutg2g::i2all with noPairPotentialCutoff = false
        // (energy.h:1390-1413): sum of i2g over the groups that are not cut
        int rc = sync_cm_from_device(s);
        if (rc) return rc;
        double t = 0;
        int gi = s->h_gid[i];
        for (int g = 0; g < s->ng; ++g) {
            if (cut_host(s, gi, which, g, which)) continue;
            const GroupDesc &G = s->groups[g];
            if ((rc = row_sum(s, which, i, G.front, G.back + 1, i, 0))) return rc;
            if ((rc = fetch(s, 1))) return rc;
            t += s->h_out[0];
        }
        *u = t;
        return FGX_OK;
    }
    int rc = row_sum(s, which, i, 0, s->n, i, 0);
    if (rc) return rc;
    if ((rc = fetch(s, 1))) return rc;
    *u = s->h_out[0];
    return FGX_OK;
}

// Nonbonded::g2g (energy.h:879-919) incl. the subgroup cases
static int g2g_raw(fgx_system *s, int which, int g1, int g2, double *u)
{
    const GroupDesc &A = s->groups[g1], &B = s->groups[g2];
    *u = 0.0;
    if (A.back < A.front || B.back < B.front) return FGX_OK;
    auto find = [](const GroupDesc &g, int i) { return i >= g.front && i <= g.back; };
    int rc;
    if (find(A, B.front) && find(A, B.back)) {  // B subgroup of A
        if ((rc = block_sum_launch(s, which, A.front, B.front, B.front, B.back + 1, 0, nullptr, 0)))
            return rc;
        if ((rc = block_sum_launch(s, which, B.back + 1, A.back + 1, B.front, B.back + 1, 0, nullptr,
                                   1)))
            return rc;
        if ((rc = fetch(s, 2))) return rc;
        *u = s->h_out[0] + s->h_out[1];
        return FGX_OK;
    }
    if (find(B, A.front) && find(B, A.back)) {  // A subgroup of B
        if ((rc = block_sum_launch(s, which, B.front, A.front, A.front, A.back + 1, 0, nullptr, 0)))
            return rc;
        if ((rc = block_sum_launch(s, which, A.back + 1, B.back + 1, A.front, A.back + 1, 0, nullptr,
                                   1)))
            return rc;
        if ((rc = fetch(s, 2))) return rc;
        *u = s->h_out[0] + s->h_out[1];
        return FGX_OK;
    }
    if ((rc = block_sum_launch(s, which, A.front, A.back + 1, B.front, B.back + 1, 0, nullptr, 0)))
        return rc;
    if ((rc = fetch(s, 1))) return rc;
    *u = s->h_out[0];
    return FGX_OK;
}

extern "C" int fgx_g2g(fgx_system *s, int which, int g1, int g2, double *u)
{
    FGX_CHECK(check_which(s, which) && u && g1 >= 0 && g2 >= 0 && g1 < s->ng && g2 < s->ng,
              FGX_EARG, "fgx_g2g: bad arguments");
    if (s->pot.nbkind == NB_CUTG2G) {
        int rc = sync_cm_from_device(s);
        if (rc) return rc;
        if (cut_host(s, g1, which, g2, which)) { *u = 0.0; return FGX_OK; }
    }
    return g2g_raw(s, which, g1, g2, u);
}

extern "C" int fgx_g_internal(fgx_system *s, int which, int g, double *u)
{
    FGX_CHECK(check_which(s, which) && u && g >= 0 && g < s->ng, FGX_EARG,
              "fgx_g_internal: bad arguments");
    const GroupDesc &G = s->groups[g];
    int rc = block_sum_launch(s, which, G.front, G.back + 1, G.front, G.back + 1, 1, nullptr, 0);
    if (rc) return rc;
    if ((rc = fetch(s, 1))) return rc;
    *u = s->h_out[0];
    return FGX_OK;
}

// Cuboid::_getVolume (src/geometry.cpp:293-300)
static double box_volume(const Box &b) { return b.len[0] * b.len[1] * b.len[2]; }

// NonbondedEwald::i_external (ewald.h:707-710): the self energy of particle i
// (Energybase's default 0 without Ewald, energy.h:128-129; ExternalPressure
// has no per-particle term)
extern "C" int fgx_i_external(fgx_system *s, int which, int i, double *u)
{
    FGX_CHECK(check_which(s, which) && u && i >= 0 && i < s->n, FGX_EARG,
              "fgx_i_external: bad arguments");
    double e = 0.0;
    if (s->has_ewald) {
        int rc = ewald_self(s, which, i, i, &e);
        if (rc) return rc;
    }
    *u = e;
    return FGX_OK;
}

// NonbondedEwald::g_external (ewald.h:712-714), then ExternalPressure::
// g_external = -N ln V for the group's N molecules (energy.h:1885-1893)
extern "C" int fgx_g_external(fgx_system *s, int which, int g, double *u)
{
    FGX_CHECK(check_which(s, which) && u && g >= 0 && g < s->ng, FGX_EARG,
              "fgx_g_external: bad arguments");
    double e = 0.0;
    if (s->has_ewald) {
        int rc = ewald_self(s, which, s->groups[g].front, s->groups[g].back, &e);
        if (rc) return rc;
    }
    if (s->press_on) {
        const GroupDesc &G = s->groups[g];
        const int N = G.back < G.front ? 0 : (G.back - G.front + 1) / G.molsize;
        e += -N * std::log(box_volume(s->box));
    }
    *u = e;
    return FGX_OK;
}

// NonbondedEwald::external (ewald.h:716-729), then ExternalPressure::external
// = P V - ln V (energy.h:1878-1883), as CombinedEnergy adds them
extern "C" int fgx_external(fgx_system *s, int which, double *u)
{
    FGX_CHECK(check_which(s, which) && u, FGX_EARG, "fgx_external: bad arguments");
    double e = 0.0;
    if (s->has_ewald) {
        double surf, reci;
        int rc = ewald_surface(s, which, s->ew.V, &surf);
        if (rc) return rc;
        if ((rc = ewald_reciprocal(s, which ? s->ew.Qt : s->ew.Q, s->ew.V, &reci))) return rc;
        if (which) { s->ew.surf_trial = surf; s->ew.reci_trial = reci; }
        else { s->ew.surf = surf; s->ew.reci = reci; }
        e = surf + reci;
    }
    if (s->press_on) {
        const double V = box_volume(s->box);
        e += s->press * V - std::log(V);
    }
    *u = e;
    return FGX_OK;
}

// Energy::systemEnergy (energy.h:3486-3491): setSpace, then Energybase::
// systemEnergy (energy.h:169-180) = external + sum_g (g_external + g_internal)
// + sum_{g<h} g2g.  Pair sums run as one masked i<j sweep.  `fixed` and
// `mask_buf` (mask_cap bytes): the allocation-free form's scratch; null
// pointers take the system's own buffers.
static int system_energy_impl(fgx_system *s, int which, double *u, const FixedPartials *fixed,
                              unsigned char *mask_buf, size_t mask_cap)
{
    int rc;
    if (s->has_ewald && (rc = ewald_set_space(s))) return rc;
    double ext = 0.0;
    if ((rc = fgx_external(s, which, &ext))) return rc;
    double gext = 0.0;
    for (int g = 0; g < s->ng && (s->has_ewald || s->press_on); ++g) {
        if (s->groups[g].back < s->groups[g].front) continue;
        double t;
        if ((rc = fgx_g_external(s, which, g, &t))) return rc;
        gext += t;
    }
    const unsigned char *mask = nullptr;
    std::vector<unsigned char> hm;
    unsigned char *dm = nullptr;
    if (s->pot.nbkind == NB_CUTG2G && s->ng > 1) {
        if ((rc = sync_cm_from_device(s))) return rc;
        hm.assign((size_t)s->ng * s->ng, 1);
        for (int a = 0; a < s->ng; ++a)
            for (int b = 0; b < s->ng; ++b)
                if (a != b && cut_host(s, a, which, b, which)) hm[(size_t)a * s->ng + b] = 0;
        if (fixed) {
            FGX_CHECK(hm.size() <= mask_cap, FGX_ENOMEM,
                      "allocation-free call: too many groups for the fixed group-pair mask");
            FGX_HIP(hipMemcpyAsync(mask_buf, hm.data(), hm.size(), hipMemcpyHostToDevice,
                                   s->stream));
            mask = mask_buf;
        } else {
            FGX_HIP(hipMalloc(&dm, hm.size()));
            FGX_HIP(hipMemcpyAsync(dm, hm.data(), hm.size(), hipMemcpyHostToDevice, s->stream));
            mask = dm;
        }
    }
    rc = block_sum_launch(s, which, 0, s->n, 0, s->n, 1, mask, 0, fixed);
    if (!rc) rc = fetch(s, 1);
    if (dm) (void)hipFree(dm);
    if (rc) return rc;
    *u = ext + gext + s->h_out[0];
    return FGX_OK;
}

extern "C" int fgx_system_energy(fgx_system *s, int which, double *u)
{
    FGX_CHECK(check_which(s, which) && u, FGX_EARG, "fgx_system_energy: bad arguments");
    return system_energy_impl(s, which, u, nullptr, nullptr, 0);
}

// ------------------------------------------------------------------------
// Allocation-free Energy::systemEnergy.  Its pair kernels write their
// partial sums, and NonbondedCutg2g its group-pair mask, into this file-scope
// device array instead of buffers the library allocates.  Every GPU holds its
// own copy for the life of the process; its address on each device is looked
// up once (hipGetSymbolAddress with that device current) and only that
// address is handed to kernels and copies.  Calls on one GPU share the array:
// calls that overlap on one GPU race.  Everything a call reads from it is
// written by that call first.
// ------------------------------------------------------------------------
static constexpr long NOALLOC_PARTIALS = 1L << 18;   // FP64 partial sums (2 MiB)
static constexpr size_t NOALLOC_MASK = 1u << 16;     // group-pair mask bytes (ng <= 256)
static __device__ __align__(256) unsigned char scratch[NOALLOC_PARTIALS * 8 + NOALLOC_MASK];

static constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// the array's device address on `device` (looked up once per device)
static int scratch_on(int device, unsigned char **out)
{
    FGX_CHECK(device >= 0 && device < NOALLOC_MAX_DEVICES, FGX_EARG,
              "allocation-free call: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        FGX_HIP(hipGetDevice(&cur));
        if (cur != device) FGX_HIP(hipSetDevice(device));
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != device) FGX_HIP(hipSetDevice(cur));
        if (e != hipSuccess || !ptr) {
            set_error(std::string("allocation-free call: hipGetSymbolAddress failed: ") +
                      hipGetErrorString(e));
            return FGX_EHIP;
        }
        g_scratch_ptr[device] = ptr;
    }
    *out = (unsigned char *)g_scratch_ptr[device];
    return FGX_OK;
}

extern "C" int fgx_system_energy_noalloc(fgx_system *s, int which, double *u)
{
    FGX_CHECK(check_which(s, which) && u, FGX_EARG, "fgx_system_energy_noalloc: bad arguments");
    unsigned char *base = nullptr;
    int rc = scratch_on(s->device, &base);
    if (rc) return rc;
    const FixedPartials fixed{(double *)base, NOALLOC_PARTIALS};
    return system_energy_impl(s, which, u, &fixed, base + NOALLOC_PARTIALS * 8, NOALLOC_MASK);
}

// ------------------------------------------------------------- Delta U

static int stage_change(fgx_system *s, const fgx_change *c, int *nm, bool *list)
{
    FGX_CHECK(c && c->group >= 0 && c->group < s->ng && c->count >= 0, FGX_EARG,
              "fgx_change: bad group or count");
    const GroupDesc &G = s->groups[c->group];
    if (c->count == 0) {
        *nm = G.back - G.front + 1;
        *list = false;
        return FGX_OK;
    }
    FGX_CHECK(c->index, FGX_EARG, "fgx_change: index array missing");
    for (int k = 0; k < c->count; ++k)
        FGX_CHECK(c->index[k] >= G.front && c->index[k] <= G.back, FGX_EARG,
                  "fgx_change: moved index outside its group");
    int rc = ensure_idx(s, c->count);
    if (rc) return rc;
    rc = upload(s, s->d_idx, c->index, c->count);
    if (rc) return rc;
    *nm = c->count;
    *list = true;
    return FGX_OK;
}

extern "C" int fgx_update_change(fgx_system *s, const fgx_change *c)
{
    FGX_CHECK(s, FGX_EARG, "fgx_update_change: null system");
    if (!s->has_ewald) return FGX_OK;
    int nm;
    bool list;
    int rc = stage_change(s, c, &nm, &list);
    if (rc) return rc;
    if (!list) {
        // the reference only updates S(k) for explicitly listed indices
        // (ewald.h:668-670): a whole-group change leaves Q_trial = Q
        FGX_HIP(hipMemcpyAsync(s->ew.Qt, s->ew.Q, sizeof(double) * 2 * s->ew.K,
                               hipMemcpyDeviceToDevice, s->stream));
        return FGX_OK;
    }
    hipLaunchKernelGGL(k_ewald_change, dim3((s->ew.K + 255) / 256), dim3(256), 0, s->stream,
                       s->ew.kv, s->ew.K, pv(s, 0), pv(s, 1), s->d_idx, nm, 0, s->ew.eff, s->ew.Q,
                       s->ew.Qt);
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

// Energy::energyChange (energy.h:3577-3621) for a one-group change.
extern "C" int fgx_energy_change(fgx_system *s, const fgx_change *c, double *du)
{
    FGX_CHECK(s && du, FGX_EARG, "fgx_energy_change: bad arguments");
    int nm;
    bool list;
    int rc = stage_change(s, c, &nm, &list);
    if (rc) return rc;
    const GroupDesc &G = s->groups[c->group];
    // container overlap of the moved particles (energy.h:3597-3607)
    {
        std::vector<int> ids;
        if (list) ids.assign(c->index, c->index + c->count);
        else for (int i = G.front; i <= G.back; ++i) ids.push_back(i);
        std::vector<double> xyz(3 * (size_t)s->n);
        if ((rc = fgx_get_positions(s, 1, xyz.data()))) return rc;
        for (int i : ids)
            for (int k = 0; k < 3; ++k)
                if (std::fabs(xyz[3 * i + k]) > s->box.half[k]) { *du = HUGE_VAL; return FGX_OK; }
    }
    // keep flags per other group (NonbondedCutg2g::cut with cm_trial for the
    // trial vector, energy.h:1336-1358); the moved group's own partners
    // always count (g_internal / atomic i2i, energy.h:3541-3557)
    if (s->pot.nbkind == NB_CUTG2G && (rc = sync_cm_from_device(s))) return rc;
    std::vector<unsigned char> km(2 * (size_t)s->ng, 1);
    bool flip = false;
    for (int h = 0; h < s->ng; ++h) {
        if (h == c->group) continue;
        km[h] = !cut_host(s, c->group, true, h, true);
        km[s->ng + h] = !cut_host(s, c->group, false, h, false);
        flip |= (km[h] != km[s->ng + h]);
    }
    // a whole molecular group moved: no internal energy (energy.h:3553-3557);
    // an atomic group: i2i to unmoved members; a molecular group with indices:
    // g_internal(trial) - g_internal(p) = moved sites vs the rest of the group
    if ((rc = upload(s, s->d_mask, km.data(), km.size()))) return rc;
    dim3 g((s->n + RB - 1) / RB, (nm + TILE - 1) / TILE);
    int nb = g.x * g.y;
    if ((rc = ensure_partials(s, nb))) return rc;
    hipLaunchKernelGGL(k_delta, g, dim3(RB), 0, s->stream, s->box, s->pot, s->d_tab, pv(s, 0),
                       pv(s, 1), list ? s->d_idx : (const int *)nullptr, nm, G.front, G.back,
                       list ? 1 : 0, s->n, s->d_mask, s->d_mask + s->ng, s->d_partial);
    FGX_HIP(hipGetLastError());
    if ((rc = finish_sum(s, nb, 0))) return rc;
    if ((rc = fetch(s, 1))) return rc;
    double dpair = s->h_out[0];
    // g2g terms of the unmoved members of a moved molecular group whose cut
    // flag flips between the trial and accepted mass centres
    if (flip && list) {
        for (int h = 0; h < s->ng; ++h) {
            if (h == c->group || km[h] == km[s->ng + h]) continue;
            // unmoved members of G vs group h, same in p and trial
            std::vector<char> mv(G.back - G.front + 1, 0);
            for (int k = 0; k < c->count; ++k) mv[c->index[k] - G.front] = 1;
            double sgn = km[h] ? 1.0 : -1.0;
            const GroupDesc &H = s->groups[h];
            int a = G.front;
            while (a <= G.back) {
                if (mv[a - G.front]) { ++a; continue; }
                int b = a;
                while (b + 1 <= G.back && !mv[b + 1 - G.front]) ++b;
                if ((rc = block_sum_launch(s, 0, a, b + 1, H.front, H.back + 1, 0, nullptr, 0)))
                    return rc;
                if ((rc = fetch(s, 1))) return rc;
                dpair += sgn * s->h_out[0];
                a = b + 1;
            }
        }
    }
    double dext = 0.0;
    if (s->has_ewald) {
        // external(trial) - external(p): reciprocal per-k difference plus the
        // surface term; the self energy (g_external) is identical for both
        // vectors and cancels
        hipLaunchKernelGGL(k_ewald_energy, dim3(1), dim3(RB), 0, s->stream, s->ew.aks, s->ew.K,
                           s->ew.Q, s->ew.Qt, s->d_out, 0);
        FGX_HIP(hipGetLastError());
        if ((rc = fetch(s, 1))) return rc;
        const double pi = std::acos(-1.);
        dext = (2 * pi / s->ew.V) * s->h_out[0] * s->pot.lB;
        if (s->ew.const_inf >= 0.5) {
            double st, sp;
            if ((rc = ewald_surface(s, 1, s->ew.V, &st))) return rc;
            if ((rc = ewald_surface(s, 0, s->ew.V, &sp))) return rc;
            dext += st - sp;
        }
    }
    *du = dpair + dext;
    return FGX_OK;
}

// accept / reject: particle copies and mass centres (move.h:866-889,
// group.h:371-389), then NonbondedEwald::update (ewald.h:613-648)
extern "C" int fgx_commit(fgx_system *s, const fgx_change *c, int accepted, double *drift)
{
    FGX_CHECK(s, FGX_EARG, "fgx_commit: null system");
    if (accepted) s->emat_valid = s->cell_valid = false;
    if (int prc = pull_device_state(s)) return prc;
    s->state_on_dev = false;
    int nm;
    bool list;
    int rc = stage_change(s, c, &nm, &list);
    if (rc) return rc;
    const GroupDesc &G = s->groups[c->group];
    if (nm > 0) {
        Vec64 dst = accepted ? s->v[0] : s->v[1];
        Vec64 src = accepted ? s->v[1] : s->v[0];
        hipLaunchKernelGGL(k_copy_moved, dim3((nm + 255) / 256), dim3(256), 0, s->stream, dst, src,
                           list ? s->d_idx : (const int *)nullptr, nm, G.front, s->d_pk, s->d_qf,
                           s->box, accepted ? 1 : 0);
        FGX_HIP(hipGetLastError());
    }
    // mass centres: whole-group moves and molecular groups carry cm/cm_trial
    if (!list || molecular(G)) {
        if ((rc = sync_cm_from_device(s))) return rc;
        for (int k = 0; k < 3; ++k) {
            if (accepted) s->cm[3 * c->group + k] = s->cm_trial[3 * c->group + k];
            else s->cm_trial[3 * c->group + k] = s->cm[3 * c->group + k];
        }
        if ((rc = push_cm(s))) return rc;
    }
    double d = 0.0;
    if (s->has_ewald) {
        EwaldDev &e = s->ew;
        if (!accepted) {
            FGX_HIP(hipMemcpyAsync(e.Qt, e.Q, sizeof(double) * 2 * e.K, hipMemcpyDeviceToDevice,
                                   s->stream));
        } else if ((rc = ewald_accepted(s, &d))) {
            return rc;
        }
    }
    FGX_HIP(hipStreamSynchronize(s->stream));
    if (drift) *drift = d;
    return FGX_OK;
}

// ------------------------------------------------------------------------
// Move::Isobaric (move.h:5424-5640) with Energy::ExternalPressure
// ------------------------------------------------------------------------
namespace {

#pragma clang fp contract(off)

// libstdc++ std::mt19937 over a host state (624 words, then the position),
// and the draws the moves take from it (slump.h:24-126)
struct HostMT {
    uint32_t *x;
    void twist()
    {
        for (int i = 0; i < 624; ++i) {
            const uint32_t y = (x[i] & 0x80000000u) | (x[(i + 1) % 624] & 0x7fffffffu);
            x[i] = x[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
    }
    uint32_t next()
    {
        if (x[624] >= 624) {
            twist();
            x[624] = 0;
        }
        uint32_t z = x[x[624]++];
        z ^= (z >> 11);
        z ^= (z << 7) & 0x9d2c5680u;
        z ^= (z << 15) & 0xefc60000u;
        z ^= (z >> 18);
        return z;
    }
    // generate_canonical<double,53>: (w1 + w2 2^32) / 2^64, clamped below 1
    double canonical()
    {
        double sum = (double)next();
        sum += (double)next() * 4294967296.0;
        double r = sum / 18446744073709551616.0;
        if (r >= 1.0) r = 0x1.fffffffffffffp-1;
        return r;
    }
    // uniform_int_distribution<int>(a, b), Lemire's method
    int range(int a, int b)
    {
        const uint32_t erange = (uint32_t)((uint32_t)b - (uint32_t)a) + 1u;
        uint64_t prod = (uint64_t)next() * (uint64_t)erange;
        uint32_t low = (uint32_t)prod;
        if (low < erange) {
            const uint32_t thr = (uint32_t)(0u - erange) % erange;
            while (low < thr) {
                prod = (uint64_t)next() * (uint64_t)erange;
                low = (uint32_t)prod;
            }
        }
        return (int)((uint32_t)(prod >> 32) + (uint32_t)a);
    }
};

int anint_h(double x) { return (int)(x > 0. ? x + .5 : x - .5); }

// Cuboid::boundary (geometry.h:539-547)
void boundary_h(const double len[3], double *a)
{
    for (int k = 0; k < 3; ++k)
        if (std::fabs(a[k]) > len[k] * 0.5) a[k] -= len[k] * anint_h(a[k] * (1 / len[k]));
}

void set_box(Box &b, const double len[3])
{   // Cuboid::setlen (src/geometry.cpp:265-274)
    for (int k = 0; k < 3; ++k) {
        b.len[k] = len[k];
        b.half[k] = len[k] * 0.5;
        b.inv[k] = 1 / len[k];
        b.lenf[k] = (float)len[k];
        b.fscale[k] = (float)(len[k] / 4294967296.0);
    }
}

#pragma clang fp contract(on)

}  // namespace

// energyChangeConfiguration (energy.h:3521-3572) for a change that moves every
// group (atomic groups with all their indices, molecular groups whole):
// g2g over all group pairs, external, each group's g_external, and the
// internal pairs of the atomic groups.
static int all_groups_config(fgx_system *s, int which, double *u)
{
    double t = 0.0, v;
    int rc;
    for (int a = 0; a < s->ng; ++a)
        for (int b = a + 1; b < s->ng; ++b) {
            if ((rc = fgx_g2g(s, which, a, b, &v))) return rc;
            t += v;
        }
    if ((rc = fgx_external(s, which, &v))) return rc;
    t += v;
    for (int g = 0; g < s->ng; ++g) {
        const GroupDesc &G = s->groups[g];
        if (G.back < G.front) continue;
        if ((rc = fgx_g_external(s, which, g, &v))) return rc;
        t += v;
        if (G.molsize == 1) {
            if ((rc = fgx_g_internal(s, which, g, &v))) return rc;
            t += v;
        }
    }
    *u = t;
    return FGX_OK;
}

// Energybase::setSpace (energy.h:86-90) with the Space's box (Cuboid::setlen,
// src/geometry.cpp:265-274): the box, and for NonbondedEwald::setSpace
// (ewald.h:731-757) the k-vectors, A_k and V of that box and S(k) recomputed
// from the accepted vector (trial S(k) = current)
extern "C" int fgx_set_space(fgx_system *s, const double len[3])
{
    FGX_CHECK(s && len && len[0] > 0 && len[1] > 0 && len[2] > 0, FGX_EARG,
              "fgx_set_space: bad arguments");
    int rc;
    if (len[0] != s->box.len[0] || len[1] != s->box.len[1] || len[2] != s->box.len[2]) {
        set_box(s->box, len);
        if ((rc = ewald_rebox(s, len))) return rc;
        // the fixed-point copies, cell lists and group-pair cache depend on the box
        hipLaunchKernelGGL(k_fixed_all, dim3((s->n + 255) / 256), dim3(256), 0, s->stream,
                           s->v[0].x, s->v[0].y, s->v[0].z, s->d_qf, s->n, s->box, s->d_pk);
        FGX_HIP(hipGetLastError());
        s->emat_valid = s->cell_valid = false;
    }
    if ((rc = ewald_set_space(s))) return rc;
    FGX_HIP(hipStreamSynchronize(s->stream));
    return FGX_OK;
}

// energyChangeConfiguration (energy.h:3521-3572) of a change that moves every
// group (Isobaric::_trialMove, move.h:5548-5565), for `which` at the current
// box: g2g over all group pairs, external, each group's g_external and the
// internal pairs of the atomic groups
extern "C" int fgx_config_energy(fgx_system *s, int which, double *u)
{
    FGX_CHECK(check_which(s, which) && u, FGX_EARG, "fgx_config_energy: bad arguments");
    return all_groups_config(s, which, u);
}

// accept / reject of a change that moved every group: Isobaric::_acceptMove /
// _rejectMove copy every group (group.h:371-389), then NonbondedEwald::update
// (ewald.h:613-648) with its accept counter and refresh
extern "C" int fgx_commit_all(fgx_system *s, int accepted, double *drift)
{
    FGX_CHECK(s, FGX_EARG, "fgx_commit_all: null system");
    if (int prc = pull_device_state(s)) return prc;
    s->state_on_dev = false;
    s->emat_valid = s->cell_valid = false;
    Vec64 dst = accepted ? s->v[0] : s->v[1];
    Vec64 src = accepted ? s->v[1] : s->v[0];
    if (s->n > 0) {
        hipLaunchKernelGGL(k_copy_moved, dim3((s->n + 255) / 256), dim3(256), 0, s->stream, dst,
                           src, (const int *)nullptr, s->n, 0, s->d_pk, s->d_qf, s->box,
                           accepted ? 1 : 0);
        FGX_HIP(hipGetLastError());
    }
    int rc;
    if ((rc = sync_cm_from_device(s))) return rc;
    for (size_t k = 0; k < s->cm.size(); ++k) {
        if (accepted) s->cm[k] = s->cm_trial[k];
        else s->cm_trial[k] = s->cm[k];
    }
    if ((rc = push_cm(s))) return rc;
    double d = 0.0;
    if (s->has_ewald) {
        if (!accepted) {
            FGX_HIP(hipMemcpyAsync(s->ew.Qt, s->ew.Q, sizeof(double) * 2 * s->ew.K,
                                   hipMemcpyDeviceToDevice, s->stream));
        } else if ((rc = ewald_accepted(s, &d))) {
            return rc;
        }
    }
    FGX_HIP(hipStreamSynchronize(s->stream));
    if (drift) *drift = d;
    return FGX_OK;
}

// NonbondedEwald's k-vectors in use for the current box (kVectorChange,
// ewald.h:330-378; the count its info() reports as "Wavefunctions",
// ewald.h:466), 0 without Ewald
extern "C" int fgx_ewald_k(const fgx_system *s, int *K)
{
    FGX_CHECK(s && K, FGX_EARG, "fgx_ewald_k: bad arguments");
    *K = s->has_ewald ? s->ew.K : 0;
    return FGX_OK;
}

extern "C" int fgx_get_box(const fgx_system *s, double len[3])
{
    FGX_CHECK(s && len, FGX_EARG, "fgx_get_box: bad arguments");
    for (int k = 0; k < 3; ++k) len[k] = s->box.len[k];
    return FGX_OK;
}

// Propagator::move() calls (move.h:7877-7891) of a Propagator holding one
// Isobaric move (npt_index < 0), or the system's one registered move and an
// Isobaric move at position npt_index (0 or 1) of its move list.  The pick is
// drawn on the host from a copy of the internal engine; a pick of the
// registered move leaves that draw to the device sweep, whose own pick over
// its one move (range(0, 0)) takes exactly one engine output as range(0, 1)
// does, and runs that one call there.
static int propagate_with_isobaric(fgx_system *s, double dp, double prob, int npt_index,
                                   int ncalls, fgx_stats *st, double *du_out, int *acc_out,
                                   long cap)
{
    FGX_CHECK(s && ncalls >= 0 && (cap <= 0 || (du_out && acc_out)), FGX_EARG,
              "fgx_isobaric_propagate: bad arguments");
    FGX_CHECK(s->press_on, FGX_ESTATE,
              "Isobaric needs the ExternalPressure term in the Hamiltonian (move.h:5466-5472)");
    FGX_CHECK(npt_index < 0 || ((npt_index == 0 || npt_index == 1) && s->moves.size() == 1),
              FGX_EUNSUPPORTED,
              "Propagator with Isobaric: exactly one registered move besides it, npt_index 0 or 1");
    for (int g = 0; g < s->ng; ++g) {
        const GroupDesc &G = s->groups[g];
        const int sz = G.back - G.front + 1;
        FGX_CHECK(G.back < G.front || G.molsize == 1 || G.molsize == sz, FGX_EUNSUPPORTED,
                  "Isobaric: a group holding several molecules (Group::isRange)");
    }
    if (st) std::memset(st, 0, sizeof(*st));
    if (cap < 0) cap = 0;
    int rc = FGX_OK;
    const int n = s->n, ng = s->ng;
    std::vector<double> p(3 * (size_t)n), t(3 * (size_t)n);
    std::vector<double> cm(3 * (size_t)ng), cmt(3 * (size_t)ng);
    bool have_p = false;   // p holds the accepted positions
    // Isobaric constructor: runfraction = prob, 0 when dp < 1e-6 (move.h:5462-5464)
    const double runfraction = dp < 1e-6 ? 0.0 : prob;
    long trials = 0, accepted = 0, inf = 0, refresh = 0;
    double utot = 0.0;   // Movebase::move: accepted dU plus update() drift terms
    // one Movebase::move of the Isobaric move after the pick; adds the call's
    // utot (accepted dU, then pot->update) to `ucall`
    auto iso_call = [&](HostMT &G, HostMT &I, double &ucall) -> int {
        if (!(I.canonical() < runfraction)) return FGX_OK;   // Movebase::run (move.h:575-581)
        // ---- Isobaric::_trialMove (move.h:5529-5567)
        double oldlen[3], newlen[3];
        for (int k = 0; k < 3; ++k) oldlen[k] = s->box.len[k];
        double newval, xyz;
        {
#pragma clang fp contract(off)
            const double oldval = box_volume(s->box);
            newval = std::exp(std::log(oldval) + (G.canonical() - 0.5) * dp);
            xyz = std::cbrt(newval / oldval);
            for (int k = 0; k < 3; ++k) newlen[k] = oldlen[k] * 1.0 * xyz;
            t = p;
            for (int g = 0; g < ng; ++g) {
                const GroupDesc &Gd = s->groups[g];
                double c[3];
                mass_center_host(s, p, g, c);      // Group::setMassCenter (group.h:174-179)
                for (int k = 0; k < 3; ++k) cm[3 * g + k] = cmt[3 * g + k] = c[k];
                if (Gd.back < Gd.front) continue;
                if (Gd.molsize == 1) {             // Group::scale, atomic (group.h:323-330)
                    for (int k = 0; k < 3; ++k) cmt[3 * g + k] = c[k] * 1.0 * xyz;
                    for (int i = Gd.front; i <= Gd.back; ++i)
                        for (int k = 0; k < 3; ++k) t[3 * i + k] = t[3 * i + k] * 1.0 * xyz;
                } else {                           // molecular (group.h:332-353)
                    double ncm[3];
                    for (int k = 0; k < 3; ++k) ncm[k] = c[k] * 1.0 * xyz;
                    for (int i = Gd.front; i <= Gd.back; ++i) {
                        for (int k = 0; k < 3; ++k) t[3 * i + k] += -c[k];
                        boundary_h(oldlen, &t[3 * i]);
                    }
                    for (int i = Gd.front; i <= Gd.back; ++i) {
                        for (int k = 0; k < 3; ++k) t[3 * i + k] += ncm[k];
                        boundary_h(newlen, &t[3 * i]);
                    }
                    for (int k = 0; k < 3; ++k) cmt[3 * g + k] = ncm[k];
                }
            }
        }
        // ---- Energy::energyChange with the trial geometry (energy.h:3577-3621).
        // Each setSpace there (and in _acceptMove / _rejectMove) makes
        // NonbondedEwald take the k-vectors, A_k and V of the Space's box and
        // recompute S(k) from the accepted coordinates (ewald.h:731-757); the
        // trial configuration's reciprocal energy therefore uses the S(k) of
        // the unscaled coordinates under the trial box's k-vectors.  (On a
        // collision the trial geometry stays set until _rejectMove, which
        // restores it; nothing is computed with it in between.)
        if ((rc = fgx_set_positions(s, 1, 0, n, t.data()))) return rc;
        // Group::setMassCenter set cm to the recomputed mass centre, scale set cm_trial
        for (int g = 0; g < ng && !rc; ++g) rc = fgx_set_group_cm(s, g, &cm[3 * g], &cmt[3 * g]);
        if (rc) return rc;
        double du;
        bool coll = false;
        for (int i = 0; i < n && !coll; ++i)
            for (int k = 0; k < 3; ++k)
                if (std::fabs(t[3 * i + k]) > newlen[k] * 0.5) coll = true;
        if (coll) {
            du = HUGE_VAL;
        } else {
            set_box(s->box, newlen);
            double unew = 0.0, uold = 0.0;
            if (!(rc = ewald_rebox(s, newlen)) && !(rc = ewald_set_space(s)))
                rc = all_groups_config(s, 1, &unew);
            set_box(s->box, oldlen);
            if (!rc && !(rc = ewald_rebox(s, oldlen)) && !(rc = ewald_set_space(s)))
                rc = all_groups_config(s, 0, &uold);
            if (rc) return rc;
            du = unew - uold;
        }
        // ---- Movebase::metropolis (move.h:567-573)
        const bool acc = !(G.canonical() > std::exp(-du));
        double drift = 0.0;
        if (acc) {   // Isobaric::_acceptMove (move.h:5570-5580): setSpace with the
                     // new box while p still holds the unscaled coordinates, then
                     // the groups take their trial coordinates
            set_box(s->box, newlen);
            if (!(rc = ewald_rebox(s, newlen))) rc = ewald_set_space(s);
            p = t;
            if (!rc) rc = fgx_set_positions(s, 2, 0, n, p.data());
            for (int g = 0; g < ng && !rc; ++g) rc = fgx_set_group_cm(s, g, &cmt[3 * g], &cmt[3 * g]);
            // Movebase::move adds pot->update(true) (move.h:547-548)
            if (!rc && s->has_ewald) rc = ewald_accepted(s, &drift);
            ucall += du;
            ucall += drift;
            accepted++;
        } else {     // Isobaric::_rejectMove (move.h:5582-5592), then update(false);
                     // _trialMove's setMassCenter left cm = cm_trial = the mass
                     // centre of the accepted coordinates (group.h:174-179)
            if (s->has_ewald && !(rc = ewald_rebox(s, oldlen))) rc = ewald_set_space(s);
            if (!rc) rc = fgx_set_positions(s, 1, 0, n, p.data());
            for (int g = 0; g < ng && !rc; ++g) rc = fgx_set_group_cm(s, g, &cm[3 * g], &cm[3 * g]);
            if (du == HUGE_VAL) inf++;
        }
        if (trials < cap) {
            du_out[trials] = du;
            acc_out[trials] = acc ? 1 : 0;
        }
        trials++;
        return rc;
    };
    hipEvent_t e0, e1;
    FGX_HIP(hipEventCreate(&e0));
    FGX_HIP(hipEventCreate(&e1));
    FGX_HIP(hipEventRecord(e0, s->stream));
    for (int call = 0; call < ncalls && !rc; ++call) {
        if ((rc = pull_device_state(s))) break;
        // Propagator::move evaluates uinit = systemEnergy before its first
        // move (move.h:7884-7885), whose setSpace recomputes S(k)
        if (!s->propagator_started) {
            if ((rc = ewald_set_space(s))) break;
            s->propagator_started = true;
        }
        int k = 0;
        if (npt_index >= 0) {
            uint32_t peek[625];
            std::memcpy(peek, s->rng_internal, sizeof(peek));
            HostMT P{peek};
            k = P.range(0, 1);
        }
        if (npt_index < 0 || k == npt_index) {
            HostMT G{s->rng_global}, I{s->rng_internal};
            (void)I.range(0, npt_index < 0 ? 0 : 1);   // Propagator::move (move.h:7887)
            if (!have_p && (rc = fgx_get_positions(s, 0, p.data()))) break;
            have_p = true;
            double ucall = 0.0;
            rc = iso_call(G, I, ucall);
            utot += ucall;
        } else {
            fgx_stats s1;
            const long room = cap > trials ? cap - trials : 0;
            rc = fgx_propagate(s, 1, &s1, room ? du_out + trials : nullptr,
                               room ? acc_out + trials : nullptr, room);
            if (rc) break;
            trials += s1.trials;
            accepted += s1.accepted;
            inf += s1.rejected_inf;
            refresh += s1.ewald_refresh;
            utot += s1.utot;
            have_p = false;
        }
    }
    FGX_HIP(hipEventRecord(e1, s->stream));
    FGX_HIP(hipEventSynchronize(e1));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    s->dusum += utot;   // Propagator::move adds what Movebase::move returns
    if (st) {
        st->trials = trials;
        st->accepted = accepted;
        st->rejected_inf = inf;
        st->ewald_refresh = refresh;
        st->utot = utot;
        st->dusum = utot;
        st->seconds = ms * 1e-3;
    }
    return FGX_OK;
}

extern "C" int fgx_isobaric_propagate(fgx_system *s, double dp, double prob, int ncalls,
                                      fgx_stats *st, double *du_out, int *acc_out, long cap)
{
    return propagate_with_isobaric(s, dp, prob, -1, ncalls, st, du_out, acc_out, cap);
}

extern "C" int fgx_propagate_npt(fgx_system *s, double dp, double prob, int npt_index,
                                 int ncalls, fgx_stats *st, double *du_out, int *acc_out,
                                 long cap)
{
    FGX_CHECK(npt_index == 0 || npt_index == 1, FGX_EARG, "fgx_propagate_npt: npt_index 0 or 1");
    return propagate_with_isobaric(s, dp, prob, npt_index, ncalls, st, du_out, acc_out, cap);
}
