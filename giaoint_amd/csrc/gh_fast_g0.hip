// Batched transform kernels of the classes (20|10), (21|02), (21|11), (32|00), (32|20), (31|13).
#include "gh_fast.h"

namespace gh {
FastClassInfo fast_info_g0(int la, int lb, int lc, int ld) {
#define GH_FAST(a, b, c, d) \
    if (la == a && lb == b && lc == c && ld == d) return fast::info<a, b, c, d>();
    GH_FAST(2, 0, 1, 0)
    GH_FAST(2, 1, 0, 2)
    GH_FAST(2, 1, 1, 1)
    GH_FAST(3, 2, 0, 0)
    GH_FAST(3, 2, 2, 0)
    GH_FAST(3, 1, 1, 3)
#undef GH_FAST
    return FastClassInfo{false, 0, 0, 0, 0, nullptr, nullptr};
}
}  // namespace gh
