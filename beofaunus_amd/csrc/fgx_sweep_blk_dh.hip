// fgx_sweep_blk_dh.hip -- block AtomicTranslation sweep instances for
// Debye-Hueckel (potentials.h:3303-3375) and Coulomb (potentials.h:2515-2568)
// with or without HardSphere (potentials.h:455-475), and LennardJonesLB
// alone in FP64: sixteen trials per block.
#include "fgx_sweep_blk.hpp"

namespace fgx {

static constexpr int BLK_B_PAIR = 16;

#define FGX_BLK_CASES(X)                                  \
    X(ELEC_DH, SR_NONE, true) X(ELEC_DH, SR_NONE, false)  \
    X(ELEC_DH, SR_HS, true) X(ELEC_DH, SR_HS, false)      \
    X(ELEC_COULOMB, SR_HS, true) X(ELEC_COULOMB, SR_HS, false) \
    X(ELEC_NONE, SR_LJ, false)

int launch_blk_pair(bool mix, const Pot &P, hipStream_t st, const ChainDev *d_chains,
                    const int2 *d_map, int ngrid, int nthr, size_t dyn)
{
#define FGX_X(E, S, M)                                                                     \
    if (P.elec == E && P.sr == S && mix == M) {                                            \
        launch_blk_f<E, S, M, BLK_B_PAIR>(st, d_chains, d_map, ngrid, nthr, dyn);         \
        return FGX_OK;                                                                     \
    }
    FGX_BLK_CASES(FGX_X)
#undef FGX_X
    return FGX_EUNSUPPORTED;
}

int blk_instance_pair(bool mix, const Pot &P, size_t *static_lds, int *block)
{
#define FGX_X(E, S, M)                                       \
    if (P.elec == E && P.sr == S && mix == M) {              \
        *static_lds = blk_static_lds<E, S, M, BLK_B_PAIR>(); \
        *block = BLK_B_PAIR;                                 \
        return 1;                                            \
    }
    FGX_BLK_CASES(FGX_X)
#undef FGX_X
    return 0;
}

}  // namespace fgx
