// FP64 MFMA contraction instances (gh_mfma.h), bra momentum 2..4, and the
// dispatch over every instance.
#include "gh_mfma.h"

namespace gh {

hipError_t launch_mfma_b(int lab, int lcd, int var, const WorkItem* items, int n, const DevShell* shells, const double* exps,
                         const double* coefs, const double* norms, const double* kappa, const zc* seeds, zc* out,
                         hipStream_t s);
hipError_t launch_mfma_c(int lab, int lcd, int var, const WorkItem* items, int n, const DevShell* shells, const double* exps,
                         const double* coefs, const double* norms, const double* kappa, const zc* seeds, zc* out,
                         hipStream_t s);

hipError_t launch_mfma_a(int lab, int lcd, int var, const WorkItem* items, int n, const DevShell* shells, const double* exps,
                         const double* coefs, const double* norms, const double* kappa, const zc* seeds, zc* out,
                         hipStream_t s) {
#define GH_MF(A, B) \
    if (lab == A && lcd == B) return mf::launch_instance<A, B>(var, items, n, shells, exps, coefs, norms, kappa, seeds, out, s);
    GH_MF(2, 2)
    GH_MF(3, 2)
    GH_MF(3, 3)
    GH_MF(4, 2)
    GH_MF(4, 3)
    GH_MF(4, 4)
#undef GH_MF
    return hipErrorInvalidValue;
}

}  // namespace gh

namespace gh {

bool mfma_instance(int lab, int lcd) {
    if (lab >= 2 && lab <= 6) return lcd >= 2 && lcd <= lab;
    return (lab == 7 || lab == 8) && lcd == lab;
}

int mfma_cd_tile(int lab) { return 16 * mf::n_tiles(lab); }

hipError_t launch_mfma(int lab, int lcd, int var, const WorkItem* items, int n, const DevShell* shells, const double* exps,
                       const double* coefs, const double* norms, const double* kappa, const zc* seeds, zc* out,
                       hipStream_t s) {
    if (!mfma_instance(lab, lcd)) return hipErrorInvalidValue;
    if (lab <= 4) return launch_mfma_a(lab, lcd, var, items, n, shells, exps, coefs, norms, kappa, seeds, out, s);
    if (lab <= 6) return launch_mfma_b(lab, lcd, var, items, n, shells, exps, coefs, norms, kappa, seeds, out, s);
    return launch_mfma_c(lab, lcd, var, items, n, shells, exps, coefs, norms, kappa, seeds, out, s);
}

}  // namespace gh
