// Batched transform kernels of the classes (20|01), (20|20), (30|01), (31|11), (22|30), (32|12).
#include "gh_fast.h"

namespace gh {
FastClassInfo fast_info_g9(int la, int lb, int lc, int ld) {
#define GH_FAST(a, b, c, d) \
    if (la == a && lb == b && lc == c && ld == d) return fast::info<a, b, c, d>();
    GH_FAST(2, 0, 0, 1)
    GH_FAST(2, 0, 2, 0)
    GH_FAST(3, 0, 0, 1)
    GH_FAST(3, 1, 1, 1)
    GH_FAST(2, 2, 3, 0)
    GH_FAST(3, 2, 1, 2)
#undef GH_FAST
    return FastClassInfo{false, 0, 0, 0, 0, nullptr, nullptr};
}
}  // namespace gh
