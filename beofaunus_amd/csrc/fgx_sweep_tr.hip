// fgx_sweep_tr.hip -- TranslateRotate sweep kernels (move.h:1620-1713) and
// the group-pair energy cache fill (EnergyMatrix, energy.h:225-641), for the
// pair terms without Ewald.
#include "fgx_sweep_kernels.hpp"

namespace fgx {

template <bool MIX>
static int tr_launch(const Pot &P, hipStream_t st, const ChainDev *d, int nc, int nthr, int ncalls)
{
#define FGX_TR(E, S) \
    if (P.elec == E && P.sr == S) { hipLaunchKernelGGL((k_sweep_tr<E, S, MIX>), dim3(nc), dim3(nthr), 0, st, d, ncalls); return FGX_OK; }
    FGX_TR(ELEC_NONE, SR_HS)
    FGX_TR(ELEC_NONE, SR_LJ)
    FGX_TR(ELEC_COULOMB, SR_NONE)
    FGX_TR(ELEC_COULOMB, SR_HS)
    FGX_TR(ELEC_COULOMB, SR_LJ)
    FGX_TR(ELEC_DH, SR_NONE)
    FGX_TR(ELEC_DH, SR_HS)
    FGX_TR(ELEC_DH, SR_LJ)
#undef FGX_TR
    return FGX_EUNSUPPORTED;
}

int launch_tr(bool mix, const Pot &P, hipStream_t st, const ChainDev *d_chains, int nchains,
              int nthr, int ncalls)
{
    int rc = mix ? tr_launch<true>(P, st, d_chains, nchains, nthr, ncalls)
                 : tr_launch<false>(P, st, d_chains, nchains, nthr, ncalls);
    if (rc) {
        set_error("device TranslateRotate: no kernel for this Hamiltonian");
        return rc;
    }
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

int launch_emat_init(const Pot &P, hipStream_t st, const ChainDev *d_chains, int nchains,
                     int maxng)
{
    bool done = false;
#define FGX_EI(E, S) \
    if (!done && P.elec == E && P.sr == S) { hipLaunchKernelGGL((k_emat_init<E, S>), dim3(maxng, nchains), dim3(AT_THREADS), 0, st, d_chains); done = true; }
    FGX_EI(ELEC_NONE, SR_HS)
    FGX_EI(ELEC_NONE, SR_LJ)
    FGX_EI(ELEC_COULOMB, SR_NONE)
    FGX_EI(ELEC_COULOMB, SR_HS)
    FGX_EI(ELEC_COULOMB, SR_LJ)
    FGX_EI(ELEC_DH, SR_NONE)
    FGX_EI(ELEC_DH, SR_HS)
    FGX_EI(ELEC_DH, SR_LJ)
#undef FGX_EI
    if (!done) {
        set_error("group-pair energy cache: no kernel for this Hamiltonian");
        return FGX_EUNSUPPORTED;
    }
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

}  // namespace fgx
