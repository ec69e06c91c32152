// Batched transform kernels of the class (32|32).
#include "gh_fast.h"

namespace gh {
FastClassInfo fast_info_g12(int la, int lb, int lc, int ld) {
    if (la == 3 && lb == 2 && lc == 3 && ld == 2) return fast::info<3, 2, 3, 2>();
    return FastClassInfo{false, 0, 0, 0, 0, nullptr, nullptr};
}
}  // namespace gh
