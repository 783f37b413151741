// fgx_sweep_blk_ewald.hip -- block AtomicTranslation sweep instances for
// NonbondedEwald<LennardJonesLB> (ewald.h:229-760, potentials.h:2003-2116):
// four trials per block.
#include "fgx_sweep_blk.hpp"

namespace fgx {

static constexpr int BLK_B_EWALD = 4;

int launch_blk_ewald(bool mix, const Pot &P, hipStream_t st, const ChainDev *d_chains,
                     const int2 *d_map, int ngrid, int nthr, size_t dyn)
{
    if (P.elec != ELEC_EWALD || P.sr != SR_LJ) return FGX_EUNSUPPORTED;
    if (mix)
        launch_blk_f<ELEC_EWALD, SR_LJ, true, BLK_B_EWALD>(st, d_chains, d_map, ngrid, nthr, dyn);
    else
        launch_blk_f<ELEC_EWALD, SR_LJ, false, BLK_B_EWALD>(st, d_chains, d_map, ngrid, nthr, dyn);
    return FGX_OK;
}

int blk_instance_ewald(bool mix, const Pot &P, size_t *static_lds, int *block)
{
    if (P.elec != ELEC_EWALD || P.sr != SR_LJ) return 0;
    *static_lds = mix ? blk_static_lds<ELEC_EWALD, SR_LJ, true, BLK_B_EWALD>()
                      : blk_static_lds<ELEC_EWALD, SR_LJ, false, BLK_B_EWALD>();
    *block = BLK_B_EWALD;
    return 1;
}

}  // namespace fgx
