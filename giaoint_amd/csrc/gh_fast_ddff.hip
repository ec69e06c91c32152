// Batched transform kernels of the (22|33) class.
#include "gh_fast.h"

namespace gh {
FastClassInfo fast_info_2233() { return fast::info<2, 2, 3, 3>(); }
}  // namespace gh
