// Batched transform kernels of the classes (21|00), (30|00), (30|11), (30|21), (32|03).
#include "gh_fast.h"

namespace gh {
FastClassInfo fast_info_g2(int la, int lb, int lc, int ld) {
#define GH_FAST(a, b, c, d) \
    if (la == a && lb == b && lc == c && ld == d) return fast::info<a, b, c, d>();
    GH_FAST(2, 1, 0, 0)
    GH_FAST(3, 0, 0, 0)
    GH_FAST(3, 0, 1, 1)
    GH_FAST(3, 0, 2, 1)
    GH_FAST(3, 2, 0, 3)
#undef GH_FAST
    return FastClassInfo{false, 0, 0, 0, 0, nullptr, nullptr};
}
}  // namespace gh
