// FP64 MFMA contraction instances (gh_mfma.h), bra momentum 5 and 6.
#include "gh_mfma.h"

namespace gh {

hipError_t launch_mfma_b(int lab, int lcd, int var, const WorkItem* items, int n, const DevShell* shells, const double* exps,
                         const double* coefs, const double* norms, const double* kappa, const zc* seeds, zc* out,
                         hipStream_t s) {
#define GH_MF(A, B) \
    if (lab == A && lcd == B) return mf::launch_instance<A, B>(var, items, n, shells, exps, coefs, norms, kappa, seeds, out, s);
    GH_MF(5, 2)
    GH_MF(5, 3)
    GH_MF(5, 4)
    GH_MF(5, 5)
    GH_MF(6, 2)
    GH_MF(6, 3)
    GH_MF(6, 4)
    GH_MF(6, 5)
    GH_MF(6, 6)
#undef GH_MF
    return hipErrorInvalidValue;
}

}  // namespace gh
