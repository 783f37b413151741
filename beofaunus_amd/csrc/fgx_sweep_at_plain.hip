// fgx_sweep_at_plain.hip -- AtomicTranslation sweep kernels for Hamiltonians
// without a screened electrostatic term: Potential::Coulomb or none, with
// HardSphere or LennardJonesLB (potentials.h:455-475, 2003-2116, 2515-2568).
#include "fgx_sweep_kernels.hpp"

namespace fgx {

template <bool MIX>
static int at_plain(const Pot &P, hipStream_t st, const ChainDev *d, int nc, int nthr, size_t dyn,
                    int ncalls, int unroll, bool fx, bool cell)
{
#define FGX_AT(E, S) \
    if (P.elec == E && P.sr == S) { launch_at_k<E, S, MIX>(st, d, nc, nthr, dyn, ncalls, unroll, fx, cell); return FGX_OK; }
    FGX_AT(ELEC_NONE, SR_HS)
    FGX_AT(ELEC_NONE, SR_LJ)
    FGX_AT(ELEC_COULOMB, SR_NONE)
    FGX_AT(ELEC_COULOMB, SR_HS)
    FGX_AT(ELEC_COULOMB, SR_LJ)
#undef FGX_AT
    return FGX_EUNSUPPORTED;
}

int launch_at_plain(bool mix, const Pot &P, hipStream_t st, const ChainDev *d_chains, int nchains,
                    int nthr, size_t dyn, int ncalls, int unroll, bool fx, bool cell)
{
    return mix ? at_plain<true>(P, st, d_chains, nchains, nthr, dyn, ncalls, unroll, fx, cell)
               : at_plain<false>(P, st, d_chains, nchains, nthr, dyn, ncalls, unroll, fx, cell);
}

}  // namespace fgx
