// Batched transform kernels of the (00|00) class.
#include "gh_fast.h"

namespace gh {
FastClassInfo fast_info_0000() { return fast::info<0, 0, 0, 0>(); }
}  // namespace gh
