// fgx_device.hpp -- device-side building blocks of the Delta-U engine:
// Cuboid geometry, pair potentials (FP64 and mixed FP32/FP64), the libstdc++
// mt19937 stream replayed by one wavefront, and reductions.
//
// Written for gfx950 (CDNA4): wave64, 160 KB LDS per CU.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fgx_types.hpp"

namespace fgx {

// -------------------------------------------------------------------------
// Cuboid geometry (geometry.h:468-555).  FP64 versions follow the reference
// expression order and are compiled without FMA contraction so trial
// coordinates are bit-identical to the reference's.
// -------------------------------------------------------------------------
#pragma clang fp contract(off)

__device__ __forceinline__ int anint_d(double x) { return (int)(x > 0. ? x + .5 : x - .5); }

// Cuboid::boundary, geometry.h:539-547
__device__ __forceinline__ void boundary_d(const Box &b, double &x, double &y, double &z)
{
    if (fabs(x) > b.half[0]) x -= b.len[0] * anint_d(x * b.inv[0]);
    if (fabs(y) > b.half[1]) y -= b.len[1] * anint_d(y * b.inv[1]);
    if (fabs(z) > b.half[2]) z -= b.len[2] * anint_d(z * b.inv[2]);
}

// Cuboid::collision, geometry.h:492-501
__device__ __forceinline__ bool collision_d(const Box &b, double x, double y, double z)
{
    return fabs(x) > b.half[0] || fabs(y) > b.half[1] || fabs(z) > b.half[2];
}

// Cuboid::sqdist, geometry.h:507-518
__device__ __forceinline__ double sqdist_d(const Box &b, double ax, double ay, double az,
                                           double bx, double by, double bz)
{
    double dx = fabs(ax - bx), dy = fabs(ay - by), dz = fabs(az - bz);
    if (dx > b.half[0]) dx -= 2 * b.half[0];
    if (dy > b.half[1]) dy -= 2 * b.half[1];
    if (dz > b.half[2]) dz -= 2 * b.half[2];
    return dx * dx + (dy * dy + dz * dz);
}

// Cuboid::vdist, geometry.h:521-537
__device__ __forceinline__ void vdist_d(const Box &b, double ax, double ay, double az, double bx,
                                        double by, double bz, double &rx, double &ry, double &rz)
{
    rx = ax - bx; ry = ay - by; rz = az - bz;
    if (rx > b.half[0]) rx -= b.len[0]; else if (rx < -b.half[0]) rx += b.len[0];
    if (ry > b.half[1]) ry -= b.len[1]; else if (ry < -b.half[1]) ry += b.len[1];
    if (rz > b.half[2]) rz -= b.len[2]; else if (rz < -b.half[2]) rz += b.len[2];
}

#pragma clang fp contract(on)

// -------------------------------------------------------------------------
// Pair energies.  `Pot` carries the Hamiltonian; the type tables (LJ
// sigma^2 / 4 eps, Ewald effective-charge products) are read through `tab`.
// -------------------------------------------------------------------------

// LennardJonesMixed<LorentzBerthelot> (potentials.h:2035-2040), or
// CutShift<LennardJonesLB> (potentials.h:1805-1811) when Pot::lj_rc2 > 0:
// zero beyond the cutoff, minus the energy at the cutoff inside it.
__device__ __forceinline__ double lj64(const Pot &P, const double *__restrict__ tab, double r2,
                                       int k)
{
    const double s2 = tab[TAB_S2 * P.ntypes2 + k], e4 = tab[TAB_EPS4 * P.ntypes2 + k];
    double x = s2 / r2;
    x = x * x * x;
    double u = e4 * (x * x - x);
    if (P.lj_rc2 > 0.0) {
        double xc = s2 / P.lj_rc2;
        xc = xc * xc * xc;
        u = (r2 > P.lj_rc2) ? 0.0 : u - e4 * (xc * xc - xc);
    }
    return u;
}

__device__ __forceinline__ float lj32(const Pot &P, const float *__restrict__ tabf, float r2, int k)
{
    const float s2 = tabf[TAB_S2 * P.ntypes2 + k], e4 = tabf[TAB_EPS4 * P.ntypes2 + k];
    float x = s2 * __builtin_amdgcn_rcpf(r2);
    x = x * x * x;
    float u = e4 * (x * x - x);
    if (P.lj_rc2f > 0.f) {
        float xc = s2 * __builtin_amdgcn_rcpf(P.lj_rc2f);
        xc = xc * xc * xc;
        u = (r2 > P.lj_rc2f) ? 0.f : u - e4 * (xc * xc - xc);
    }
    return u;
}

// FP64 pair energy from squared distance (and the vector for Ewald, which
// the reference evaluates with vdist).  Same formulas as the reference's
// CombinedPairPotential (potentials.h:3768-3771): elec + short-range.
__device__ __forceinline__ double pair_u64(const Pot &P, const double *__restrict__ tab, double r2,
                                           double qa, double qb, double ra, double rb, int ta,
                                           int tb)
{
    double u1 = 0.0, u2 = 0.0;
    if (P.elec == ELEC_COULOMB) {
        u1 = P.lB * qa * qb / sqrt(r2);
    } else if (P.elec == ELEC_DH) {
        double r = sqrt(r2);
        u1 = P.lB * qa * qb / r * exp(-P.kappa * r);
    } else if (P.elec == ELEC_EWALD) {
        if (r2 <= P.ew_rc2) {
            double r1 = sqrt(r2);
            double s = P.ew_kappa / 2.0 / P.ew_alpha;
            u1 = qa * qb *
                 (erfc(P.ew_alpha * r1 + s) * exp(P.ew_kappa * r1) +
                  erfc(P.ew_alpha * r1 - s) * exp(-P.ew_kappa * r1)) /
                 (2.0 * r1) * tab[TAB_QEQE * P.ntypes2 + ta * P.ntypes + tb] * P.lB;
        }
    }
    if (P.sr == SR_HS) {
        double m = ra + rb;
        u2 = (r2 < m * m) ? __builtin_inf() : 0.0;
    } else if (P.sr == SR_LJ) {
        u2 = lj64(P, tab, r2, ta * P.ntypes + tb);
    }
    return u1 + u2;
}

// FP32 pair energy for the mixed path; the caller accumulates in FP64.
__device__ __forceinline__ float pair_u32(const Pot &P, const float *__restrict__ tabf, float r2,
                                          float qa, float qb, float ra, float rb, int ta, int tb)
{
    float u1 = 0.f, u2 = 0.f;
    if (P.elec == ELEC_COULOMB) {
        u1 = P.lBf * qa * qb * __frsqrt_rn(r2);
    } else if (P.elec == ELEC_DH) {
        float rinv = __frsqrt_rn(r2);
        float r = r2 * rinv;
        u1 = P.lBf * qa * qb * rinv * __expf(-P.kappaf * r);
    } else if (P.elec == ELEC_EWALD) {
        if (r2 <= P.ew_rc2f) {
            float rinv = __frsqrt_rn(r2);
            float r1 = r2 * rinv;
            // kappa -> 0 (plain Ewald) reduces to erfc(alpha r) / r
            float e;
            if (P.ew_plain)
                e = erfcf(P.ew_alphaf * r1) * rinv;
            else
                e = (erfcf(P.ew_alphaf * r1 + P.ew_sf) * __expf(P.ew_kappaf * r1) +
                     erfcf(P.ew_alphaf * r1 - P.ew_sf) * __expf(-P.ew_kappaf * r1)) *
                    (0.5f * rinv);
            u1 = qa * qb * e * tabf[TAB_QEQE * P.ntypes2 + ta * P.ntypes + tb] * P.lBf;
        }
    }
    if (P.sr == SR_HS) {
        float m = ra + rb;
        u2 = (r2 < m * m) ? __builtin_inff() : 0.f;
    } else if (P.sr == SR_LJ) {
        u2 = lj32(P, tabf, r2, ta * P.ntypes + tb);
    }
    return u1 + u2;
}


// Compile-time specialisations for the sweep kernels: only the terms of
// the chosen Hamiltonian are generated, which keeps register pressure down.
template <int ELEC, int SR>
__device__ __forceinline__ double pair_u64_t(const Pot &P, const double *__restrict__ tab, double r2,
                                             double qa, double qb, double ra, double rb, int ta,
                                             int tb)
{
    double u1 = 0.0, u2 = 0.0;
    if constexpr (ELEC == ELEC_COULOMB) {
        u1 = P.lB * qa * qb / sqrt(r2);
    } else if constexpr (ELEC == ELEC_DH) {
        double r = sqrt(r2);
        u1 = P.lB * qa * qb / r * exp(-P.kappa * r);
    } else if constexpr (ELEC == ELEC_EWALD) {
        if (r2 <= P.ew_rc2) {
            double r1 = sqrt(r2);
            double s = P.ew_kappa / 2.0 / P.ew_alpha;
            u1 = qa * qb *
                 (erfc(P.ew_alpha * r1 + s) * exp(P.ew_kappa * r1) +
                  erfc(P.ew_alpha * r1 - s) * exp(-P.ew_kappa * r1)) /
                 (2.0 * r1) * tab[TAB_QEQE * P.ntypes2 + ta * P.ntypes + tb] * P.lB;
        }
    }
    if constexpr (SR == SR_HS) {
        double m = ra + rb;
        u2 = (r2 < m * m) ? __builtin_inf() : 0.0;
    } else if constexpr (SR == SR_LJ) {
        u2 = lj64(P, tab, r2, ta * P.ntypes + tb);
    }
    return u1 + u2;
}

template <int ELEC, int SR>
__device__ __forceinline__ float pair_u32_t(const Pot &P, const float *__restrict__ tabf, float r2,
                                            float qa, float qb, float ra, float rb, int ta, int tb)
{
    float u1 = 0.f, u2 = 0.f;
    if constexpr (ELEC == ELEC_COULOMB) {
        u1 = P.lBf * qa * qb * __builtin_amdgcn_rsqf(r2);
    } else if constexpr (ELEC == ELEC_DH) {
        float rinv = __builtin_amdgcn_rsqf(r2);
        float r = r2 * rinv;
        u1 = P.lBf * qa * qb * rinv * __builtin_amdgcn_exp2f(-P.kappaf * 1.44269504088896341f * r);
    } else if constexpr (ELEC == ELEC_EWALD) {
        if (r2 <= P.ew_rc2f) {
            float rinv = __builtin_amdgcn_rsqf(r2);
            float r1 = r2 * rinv;
            float e;
            if (P.ew_plain)
                e = erfcf(P.ew_alphaf * r1) * rinv;
            else
                e = (erfcf(P.ew_alphaf * r1 + P.ew_sf) * __expf(P.ew_kappaf * r1) +
                     erfcf(P.ew_alphaf * r1 - P.ew_sf) * __expf(-P.ew_kappaf * r1)) *
                    (0.5f * rinv);
            u1 = qa * qb * e * tabf[TAB_QEQE * P.ntypes2 + ta * P.ntypes + tb] * P.lBf;
        }
    }
    if constexpr (SR == SR_HS) {
        // conservative contact test: anything within FP32 rounding of contact
        // reports +inf so the caller settles it in FP64
        float m = ra + rb;
        u2 = (r2 < m * m * 1.000001f) ? __builtin_inff() : 0.f;
    } else if constexpr (SR == SR_LJ) {
        u2 = lj32(P, tabf, r2, ta * P.ntypes + tb);
    }
    if constexpr (ELEC == ELEC_NONE) return u2;
    else if constexpr (SR == SR_NONE) return u1;
    else return u1 + u2;
}

// fixed-point coordinates: x -> round(x / L * 2^32) as a wrapping int32, so
// the difference of two coordinates is the minimum image exactly.
__host__ __device__ __forceinline__ int32_t to_fixed(double x, double inv_len)
{
    double s = x * inv_len * 4294967296.0;
    long long v = (long long)__builtin_rint(s);
    return (int32_t)(uint32_t)(unsigned long long)v;
}

// -------------------------------------------------------------------------
// mt19937 replay (libstdc++ std::mt19937; slump.h:24-126).  The 624-word
// state lives in LDS; all 64 lanes of one wavefront execute the draw
// functions uniformly (each lane obtains the same value), so the twist can
// be done wave-parallel.
// -------------------------------------------------------------------------
struct WaveMT {
    uint32_t *x;  // LDS, 624 words
    int p;        // wave-uniform position
    // window: lane l holds the tempered word x[wb + l] of the current state,
    // for draws wb <= p < wb + 64 (wb < 0: empty), so a draw is a lane read
    // instead of a dependent LDS load
    int wb = -1;
    uint32_t win = 0;
};

__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Standard twist, done in ten chunks of 64 words.  Within a chunk every
// lane reads x[i], x[i+1] and x[(i+397) % 624] before any lane writes;
// across chunks the sequential data dependencies of the reference loop are
// respected (words >= i+1 are still old, words i-227 are already new).
__device__ __forceinline__ void mt_twist_wave(uint32_t *x)
{
    const int lane = threadIdx.x & 63;
    for (int c = 0; c < 624; c += 64) {
        int i = c + lane;
        uint32_t nv = 0;
        if (i < 624) {
            uint32_t xi = x[i];
            uint32_t xn = x[(i + 1) % 624];
            uint32_t xm = x[(i + 397) % 624];
            uint32_t y = (xi & 0x80000000u) | (xn & 0x7fffffffu);
            nv = xm ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        wave_sync_lds();
        if (i < 624) x[i] = nv;
        wave_sync_lds();
    }
}

__device__ __forceinline__ uint32_t mt_temper(uint32_t z)
{
    z ^= (z >> 11);
    z ^= (z << 7) & 0x9d2c5680u;
    z ^= (z << 15) & 0xefc60000u;
    z ^= (z >> 18);
    return z;
}

__device__ __forceinline__ uint32_t mt_next(WaveMT &m)
{
    if (m.p >= 624) {
        mt_twist_wave(m.x);
        m.p = 0;
        m.wb = -1;
    }
    if (m.wb < 0 || m.p >= m.wb + 64) {
        // refill: the next (up to) 64 words of this state, tempered per lane
        const int k = m.p + (int)(threadIdx.x & 63);
        m.win = k < 624 ? mt_temper(m.x[k]) : 0u;
        m.wb = m.p;
    }
    const uint32_t z = (uint32_t)__builtin_amdgcn_readlane((int)m.win, m.p - m.wb);
    m.p++;
    return z;
}

#pragma clang fp contract(off)
// generate_canonical<double,53>: (w1 + w2 * 2^32) / 2^64, clamped below 1
__device__ __forceinline__ double mt_canonical(WaveMT &m)
{
    double sum = (double)mt_next(m);
    sum += (double)mt_next(m) * 4294967296.0;
    double r = sum / 18446744073709551616.0;
    if (r >= 1.0) r = 0x1.fffffffffffffp-1;
    return r;
}
#pragma clang fp contract(on)

// uniform_int_distribution<int>(a,b): Lemire's nearly-divisionless method
__device__ __forceinline__ int mt_range(WaveMT &m, int a, int b)
{
    uint32_t erange = (uint32_t)((uint32_t)b - (uint32_t)a) + 1u;
    uint64_t prod = (uint64_t)mt_next(m) * (uint64_t)erange;
    uint32_t low = (uint32_t)prod;
    if (low < erange) {
        uint32_t thr = (uint32_t)(0u - erange) % erange;
        while (low < thr) {
            prod = (uint64_t)mt_next(m) * (uint64_t)erange;
            low = (uint32_t)prod;
        }
    }
    return (int)((uint32_t)(prod >> 32) + (uint32_t)a);
}

// -------------------------------------------------------------------------
// reductions
// -------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ int wave_or(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
    return v;
}

// block-wide FP64 sum; `red` is LDS scratch of >= blockDim/64 doubles.
// Result valid in every thread.  Contains two __syncthreads.
__device__ __forceinline__ double block_sum(double v, double *red)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wave_sum(v);
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < nw; ++k) t += red[k];  // fixed order: deterministic
    __syncthreads();
    return t;
}

// workgroup barrier that orders LDS only: waits for this wave's LDS
// operations, not for its global loads and stores in flight (for loops whose
// waves exchange data through LDS alone between two such barriers)
__device__ __forceinline__ void lds_sync()
{
    __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// FP64 wave sum through DPP row operations (quad permutes, row shifts, row
// broadcasts; rows of 16 lanes), which avoid the LDS crossbar of a shuffle;
// fixed order, every lane receives lane 63's total
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                            (unsigned long long)(unsigned)lo));
}

__device__ __forceinline__ double wave_sum_dpp(double v)
{
    v += dpp_f64<0xb1>(v);    // quad_perm 1,0,3,2
    v += dpp_f64<0x4e>(v);    // quad_perm 2,3,0,1
    v += dpp_f64<0x114>(v);   // row_shr:4
    v += dpp_f64<0x118>(v);   // row_shr:8   (lane 15 of each row: the row's sum)
    v += dpp_f64<0x142>(v);   // row_bcast:15
    v += dpp_f64<0x143>(v);   // row_bcast:31 (lane 63: the wave's sum)
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, 63);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                            (unsigned long long)(unsigned)lo));
}

// block_sum with one barrier: `red` must not be written again before every
// wave has passed another barrier (callers alternate two buffers between
// calls separated by a barrier).  LDS-only barrier when `lds_only`.
__device__ __forceinline__ double block_sum1(double v, double *red, bool lds_only)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wave_sum_dpp(v);
    if (lane == 0) red[w] = v;
    if (lds_only)
        lds_sync();
    else
        __syncthreads();
    double t = 0.0;
    for (int k = 0; k < nw; ++k) t += red[k];  // fixed order: deterministic
    return t;
}

}  // namespace fgx
