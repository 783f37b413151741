// fgx_sweep_at_ewald64.hip -- AtomicTranslation sweep kernels of the FP64
// path for Energy::NonbondedEwald (ewald.h:229-760) with its EwaldReal pair
// term (ewald.h:97-110), alone or with HardSphere or LennardJonesLB.
#include "fgx_sweep_kernels.hpp"

namespace fgx {

int launch_at_ewald_fp64(const Pot &P, hipStream_t st, const ChainDev *d_chains, int nchains,
                         int nthr, size_t dyn, int ncalls, bool skg)
{
    if (P.elec != ELEC_EWALD) return FGX_EUNSUPPORTED;
    if (skg) {
        if (P.sr == SR_NONE) launch_at_skg<SR_NONE, false>(st, d_chains, nchains, nthr, dyn, ncalls, false, false);
        else if (P.sr == SR_HS) launch_at_skg<SR_HS, false>(st, d_chains, nchains, nthr, dyn, ncalls, false, false);
        else if (P.sr == SR_LJ) launch_at_skg<SR_LJ, false>(st, d_chains, nchains, nthr, dyn, ncalls, false, false);
        else return FGX_EUNSUPPORTED;
        return FGX_OK;
    }
#define FGX_AT(S) \
    if (P.sr == S) { launch_at_k<ELEC_EWALD, S, false>(st, d_chains, nchains, nthr, dyn, ncalls, 4, false, false); return FGX_OK; }
    FGX_AT(SR_NONE)
    FGX_AT(SR_HS)
    FGX_AT(SR_LJ)
#undef FGX_AT
    return FGX_EUNSUPPORTED;
}

}  // namespace fgx
