// Batched transform kernels of the classes (33|20), (33|21), (33|30).
#include "gh_fast.h"

namespace gh {
FastClassInfo fast_info_g14(int la, int lb, int lc, int ld) {
#define GH_FAST(a, b, c, d) \
    if (la == a && lb == b && lc == c && ld == d) return fast::info<a, b, c, d>();
    GH_FAST(3, 3, 2, 0)
    GH_FAST(3, 3, 2, 1)
    GH_FAST(3, 3, 3, 0)
#undef GH_FAST
    return FastClassInfo{false, 0, 0, 0, 0, nullptr, nullptr};
}
}  // namespace gh
