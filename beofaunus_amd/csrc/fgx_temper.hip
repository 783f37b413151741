// fgx_temper.hip -- ParallelTempering (move.h:6520-6772) for R replicas on
// one device.
//
// The reference runs one replica per MPI rank and swaps coordinates over
// MPI (mpi.h:238-331).  Here every replica is an fgx_system resident on the
// same GPU; replicas differ only in their Bjerrum length (temperature enters
// through lB, physconst.h:91-94), so a swap exchanges the lB labels of two
// handles instead of copying 5N+1 doubles.  Energies U_r(x_c) -- replica r's
// Hamiltonian on configuration c -- are full Energy::systemEnergy sums
// evaluated on the device with the label of r.

#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "fgx_internal.hpp"

namespace fgx {
static void mt_twist(uint32_t *x)
{
    for (int k = 0; k < 624; ++k) {
        uint32_t y = (x[k] & 0x80000000u) | (x[(k + 1) % 624] & 0x7fffffffu);
        x[k] = x[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
}
static uint32_t mt_u32(uint32_t *st)
{
    if (st[624] >= 624) { mt_twist(st); st[624] = 0; }
    uint32_t z = st[st[624]++];
    z ^= (z >> 11);
    z ^= (z << 7) & 0x9d2c5680u;
    z ^= (z << 15) & 0xefc60000u;
    z ^= (z >> 18);
    return z;
}
static double mt_canon(uint32_t *st)
{
    double sum = (double)mt_u32(st);
    sum += (double)mt_u32(st) * 4294967296.0;
    double r = sum / 18446744073709551616.0;
    return r >= 1.0 ? std::nextafter(1.0, 0.0) : r;
}
}  // namespace fgx

struct fgx_temper {
    std::vector<fgx_system *> cfg;   // configurations (handles)
    std::vector<double> lB;          // lB of replica (temperature) r
    std::vector<int> slot;           // slot[r] = handle holding temperature r
    uint32_t ctl[625], glob[625];
};

using namespace fgx;

// U_r(x_c): system energy of handle c evaluated with replica r's lB
static int energy_as(fgx_system *c, double lB, double *u)
{
    double keep = c->pot.lB;
    float keepf = c->pot.lBf;
    c->pot.lB = lB;
    c->pot.lBf = (float)lB;
    int rc = fgx_system_energy(c, 0, u);
    c->pot.lB = keep;
    c->pot.lBf = keepf;
    return rc;
}

extern "C" int fgx_temper_create(fgx_system **replicas, int R, uint32_t ctl_seed,
                                 uint32_t glob_seed, fgx_temper **out)
{
    FGX_CHECK(replicas && R >= 2 && out, FGX_EARG, "fgx_temper_create: need >= 2 replicas");
    for (int r = 0; r < R; ++r) {
        FGX_CHECK(replicas[r], FGX_EARG, "fgx_temper_create: null replica");
        // the reference aborts when partners differ in size (move.h:6686-6687)
        FGX_CHECK(replicas[r]->n == replicas[0]->n, FGX_EARG,
                  "fgx_temper_create: replicas must have the same particle count");
        // an exchange evaluates each configuration with the partner's Bjerrum
        // length only; replicas whose Hamiltonians differ otherwise (e.g. LJ
        // eps in kT converted at their own temperatures, species.h:293) are
        // outside that
        const Pot &a = replicas[r]->pot, &b = replicas[0]->pot;
        FGX_CHECK(a.elec == b.elec && a.sr == b.sr && a.nbkind == b.nbkind &&
                      a.ntypes == b.ntypes && a.kappa == b.kappa &&
                      replicas[r]->has_ewald == replicas[0]->has_ewald,
                  FGX_EUNSUPPORTED, "fgx_temper_create: replicas differ in more than lB");
        if (a.sr == SR_LJ) {
            const size_t T2 = (size_t)a.ntypes * a.ntypes;
            bool same = true;
            for (size_t k = 0; k < 2 * T2; ++k)   // sigma^2 and 4 eps
                same = same && replicas[r]->tab[k] == replicas[0]->tab[k];
            FGX_CHECK(same, FGX_EUNSUPPORTED,
                      "fgx_temper_create: replicas with different LJ tables (only lB is swapped)");
        }
    }
    fgx_temper *t = new fgx_temper();
    t->cfg.assign(replicas, replicas + R);
    for (int r = 0; r < R; ++r) {
        t->lB.push_back(replicas[r]->pot.lB);
        t->slot.push_back(r);
    }
    mt_seed_host(t->ctl, ctl_seed);
    mt_seed_host(t->glob, glob_seed);
    *out = t;
    return FGX_OK;
}

// one exchange attempt: findPartner (move.h:6621-6633), _energyChange
// (move.h:6701-6723), metropolis on both partners with the same uniform
extern "C" int fgx_temper_exchange(fgx_temper *t, int *acc, double *du)
{
    FGX_CHECK(t, FGX_EARG, "fgx_temper_exchange: null handle");
    const int R = (int)t->cfg.size();
    int dr = (mt_canon(t->ctl) > 0.5) ? 1 : -1;
    double uni = mt_canon(t->glob);
    for (int r = 0; r < R; ++r) {
        if (acc) acc[r] = 0;
        if (du) du[r] = 0.0;
    }
    for (int r = 0; r < R; ++r) {
        int p = (r % 2 == 0) ? r + dr : r - dr;
        if (p < 0 || p >= R || p < r) continue;
        fgx_system *a = t->cfg[t->slot[r]], *b = t->cfg[t->slot[p]];
        double Ura, Urb, Upa, Upb;
        int rc;
        if ((rc = energy_as(a, t->lB[r], &Ura))) return rc;
        if ((rc = energy_as(b, t->lB[r], &Urb))) return rc;
        if ((rc = energy_as(a, t->lB[p], &Upa))) return rc;
        if ((rc = energy_as(b, t->lB[p], &Upb))) return rc;
        double d = (Urb - Ura) + (Upa - Upb);
        bool accept = !(uni > std::exp(-d));
        if (du) { du[r] = d; du[p] = d; }
        if (accept) {
            std::swap(t->slot[r], t->slot[p]);
            // the handle now at temperature r takes lB_r, and vice versa
            t->cfg[t->slot[r]]->pot.lB = t->lB[r];
            t->cfg[t->slot[r]]->pot.lBf = (float)t->lB[r];
            t->cfg[t->slot[p]]->pot.lB = t->lB[p];
            t->cfg[t->slot[p]]->pot.lBf = (float)t->lB[p];
            t->cfg[t->slot[r]]->emat_valid = false;   // cached pair energies scale with lB
            t->cfg[t->slot[p]]->emat_valid = false;
            if (acc) { acc[r] = 1; acc[p] = 1; }
        }
    }
    return FGX_OK;
}

extern "C" int fgx_temper_slots(fgx_temper *t, int *slot)
{
    FGX_CHECK(t && slot, FGX_EARG, "fgx_temper_slots: bad arguments");
    for (size_t r = 0; r < t->slot.size(); ++r) slot[r] = t->slot[r];
    return FGX_OK;
}

extern "C" int fgx_temper_destroy(fgx_temper *t)
{
    delete t;
    return FGX_OK;
}
