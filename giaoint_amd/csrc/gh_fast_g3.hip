// Batched transform kernels of the classes (21|01), (30|20), (31|02), (22|21), (32|30).
#include "gh_fast.h"

namespace gh {
FastClassInfo fast_info_g3(int la, int lb, int lc, int ld) {
#define GH_FAST(a, b, c, d) \
    if (la == a && lb == b && lc == c && ld == d) return fast::info<a, b, c, d>();
    GH_FAST(2, 1, 0, 1)
    GH_FAST(3, 0, 2, 0)
    GH_FAST(3, 1, 0, 2)
    GH_FAST(2, 2, 2, 1)
    GH_FAST(3, 2, 3, 0)
#undef GH_FAST
    return FastClassInfo{false, 0, 0, 0, 0, nullptr, nullptr};
}
}  // namespace gh
