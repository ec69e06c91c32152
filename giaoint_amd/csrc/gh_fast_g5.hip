// Batched transform kernels of the classes (22|10), (31|10), (21|21), (31|30), (32|11).
#include "gh_fast.h"

namespace gh {
FastClassInfo fast_info_g5(int la, int lb, int lc, int ld) {
#define GH_FAST(a, b, c, d) \
    if (la == a && lb == b && lc == c && ld == d) return fast::info<a, b, c, d>();
    GH_FAST(2, 2, 1, 0)
    GH_FAST(3, 1, 1, 0)
    GH_FAST(2, 1, 2, 1)
    GH_FAST(3, 1, 3, 0)
    GH_FAST(3, 2, 1, 1)
#undef GH_FAST
    return FastClassInfo{false, 0, 0, 0, 0, nullptr, nullptr};
}
}  // namespace gh
