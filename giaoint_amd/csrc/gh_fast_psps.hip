// Batched transform kernels of the (10|10) class.
#include "gh_fast.h"

namespace gh {
FastClassInfo fast_info_1010() { return fast::info<1, 0, 1, 0>(); }
}  // namespace gh
