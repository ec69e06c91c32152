// Batched transform kernels of the classes (21|10), (30|10), (31|20), (31|22), (32|21).
#include "gh_fast.h"

namespace gh {
FastClassInfo fast_info_g8(int la, int lb, int lc, int ld) {
#define GH_FAST(a, b, c, d) \
    if (la == a && lb == b && lc == c && ld == d) return fast::info<a, b, c, d>();
    GH_FAST(2, 1, 1, 0)
    GH_FAST(3, 0, 1, 0)
    GH_FAST(3, 1, 2, 0)
    GH_FAST(3, 1, 2, 2)
    GH_FAST(3, 2, 2, 1)
#undef GH_FAST
    return FastClassInfo{false, 0, 0, 0, 0, nullptr, nullptr};
}
}  // namespace gh
