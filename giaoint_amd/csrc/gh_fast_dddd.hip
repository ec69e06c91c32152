// Batched transform kernels of the (22|22) class.
#include "gh_fast.h"

namespace gh {
FastClassInfo fast_info_2222() { return fast::info<2, 2, 2, 2>(); }
}  // namespace gh
