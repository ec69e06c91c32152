// Batched transform kernels of the classes (31|00), (30|02), (30|12), (31|31).
#include "gh_fast.h"

namespace gh {
FastClassInfo fast_info_g1(int la, int lb, int lc, int ld) {
#define GH_FAST(a, b, c, d) \
    if (la == a && lb == b && lc == c && ld == d) return fast::info<a, b, c, d>();
    GH_FAST(3, 1, 0, 0)
    GH_FAST(3, 0, 0, 2)
    GH_FAST(3, 0, 1, 2)
    GH_FAST(3, 1, 3, 1)
#undef GH_FAST
    return FastClassInfo{false, 0, 0, 0, 0, nullptr, nullptr};
}
}  // namespace gh
