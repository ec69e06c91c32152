// Batched transform kernels of the classes (21|20), (31|01), (21|12), (31|03), (32|02).
#include "gh_fast.h"

namespace gh {
FastClassInfo fast_info_g4(int la, int lb, int lc, int ld) {
#define GH_FAST(a, b, c, d) \
    if (la == a && lb == b && lc == c && ld == d) return fast::info<a, b, c, d>();
    GH_FAST(2, 1, 2, 0)
    GH_FAST(3, 1, 0, 1)
    GH_FAST(2, 1, 1, 2)
    GH_FAST(3, 1, 0, 3)
    GH_FAST(3, 2, 0, 2)
#undef GH_FAST
    return FastClassInfo{false, 0, 0, 0, 0, nullptr, nullptr};
}
}  // namespace gh
