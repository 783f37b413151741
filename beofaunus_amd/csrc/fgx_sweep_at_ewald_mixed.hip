// fgx_sweep_at_ewald_mixed.hip -- AtomicTranslation sweep kernels of the
// mixed path for Energy::NonbondedEwald (ewald.h:229-760) with its
// EwaldReal pair term (ewald.h:97-110), alone or with HardSphere or
// LennardJonesLB (uncut, or cut and shifted at rc with the cell list).
#include "fgx_sweep_kernels.hpp"

namespace fgx {

int launch_at_ewald_mixed(const Pot &P, hipStream_t st, const ChainDev *d_chains, int nchains,
                          int nthr, size_t dyn, int ncalls, int unroll, bool fx, bool cell,
                          bool skg)
{
    if (P.elec != ELEC_EWALD) return FGX_EUNSUPPORTED;
    if (skg) {
        if (P.sr == SR_NONE) launch_at_skg<SR_NONE, true>(st, d_chains, nchains, nthr, dyn, ncalls, fx, cell);
        else if (P.sr == SR_HS) launch_at_skg<SR_HS, true>(st, d_chains, nchains, nthr, dyn, ncalls, fx, cell);
        else if (P.sr == SR_LJ) launch_at_skg<SR_LJ, true>(st, d_chains, nchains, nthr, dyn, ncalls, fx, cell);
        else return FGX_EUNSUPPORTED;
        return FGX_OK;
    }
#define FGX_AT(S) \
    if (P.sr == S) { launch_at_k<ELEC_EWALD, S, true>(st, d_chains, nchains, nthr, dyn, ncalls, unroll, fx, cell); return FGX_OK; }
    FGX_AT(SR_NONE)
    FGX_AT(SR_HS)
    FGX_AT(SR_LJ)
#undef FGX_AT
    return FGX_EUNSUPPORTED;
}

}  // namespace fgx
