// fgx_sweep_at_dh.hip -- AtomicTranslation sweep kernels for
// Potential::DebyeHuckel (potentials.h:3303-3375), alone or with HardSphere or
// LennardJonesLB.
#include "fgx_sweep_kernels.hpp"

namespace fgx {

template <bool MIX>
static int at_dh(const Pot &P, hipStream_t st, const ChainDev *d, int nc, int nthr, size_t dyn,
                 int ncalls, int unroll, bool fx, bool cell)
{
#define FGX_AT(S) \
    if (P.sr == S) { launch_at_k<ELEC_DH, S, MIX>(st, d, nc, nthr, dyn, ncalls, unroll, fx, cell); return FGX_OK; }
    FGX_AT(SR_NONE)
    FGX_AT(SR_HS)
    FGX_AT(SR_LJ)
#undef FGX_AT
    return FGX_EUNSUPPORTED;
}

int launch_at_dh(bool mix, const Pot &P, hipStream_t st, const ChainDev *d_chains, int nchains,
                 int nthr, size_t dyn, int ncalls, int unroll, bool fx, bool cell)
{
    if (P.elec != ELEC_DH) return FGX_EUNSUPPORTED;
    return mix ? at_dh<true>(P, st, d_chains, nchains, nthr, dyn, ncalls, unroll, fx, cell)
               : at_dh<false>(P, st, d_chains, nchains, nthr, dyn, ncalls, unroll, fx, cell);
}

}  // namespace fgx
