// Split-FP32 matrix-core contraction instances (gh_mfma.h, SPLIT = true) for
// the precision study of the high angular momentum classes: (gg|gg) as
// (LAB, LCD) = (8, 8) and (ff|dd) as (6, 4).
#include "gh_mfma.h"

namespace gh {

bool mfma_split_instance(int lab, int lcd) { return (lab == 8 && lcd == 8) || (lab == 6 && lcd == 4); }

hipError_t launch_mfma_split(int lab, int lcd, const WorkItem* items, int n, const DevShell* shells,
                             const double* exps, const double* coefs, const double* norms, const double* kappa,
                             const zc* seeds, zc* out, hipStream_t s) {
    if (lab == 8 && lcd == 8)
        return mf::launch_instance<8, 8, true>(0, items, n, shells, exps, coefs, norms, kappa, seeds, out, s);
    if (lab == 6 && lcd == 4)
        return mf::launch_instance<6, 4, true>(0, items, n, shells, exps, coefs, norms, kappa, seeds, out, s);
    return hipErrorInvalidValue;
}

}  // namespace gh
