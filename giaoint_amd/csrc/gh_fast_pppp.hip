// Batched transform kernels of the (11|11) class.
#include "gh_fast.h"

namespace gh {
FastClassInfo fast_info_1111() { return fast::info<1, 1, 1, 1>(); }
}  // namespace gh
