// FP64 MFMA contraction instances (gh_mfma.h), bra momentum 7 and 8.
#include "gh_mfma.h"

namespace gh {

hipError_t launch_mfma_c(int lab, int lcd, int var, const WorkItem* items, int n, const DevShell* shells, const double* exps,
                         const double* coefs, const double* norms, const double* kappa, const zc* seeds, zc* out,
                         hipStream_t s) {
#define GH_MF(A, B) \
    if (lab == A && lcd == B) return mf::launch_instance<A, B>(var, items, n, shells, exps, coefs, norms, kappa, seeds, out, s);
    GH_MF(7, 7)
    GH_MF(8, 8)
#undef GH_MF
    return hipErrorInvalidValue;
}

}  // namespace gh
