// Class dispatch of the batched transform. The kernels are instantiated in
// gh_fast_*.hip, one translation unit per group of classes.
#include "gh_internal.h"

namespace gh {

FastClassInfo fast_info_2222();
FastClassInfo fast_info_2233();
FastClassInfo fast_info_1111();
FastClassInfo fast_info_1010();
FastClassInfo fast_info_0000();

int fast_waves(int nhab, int kab, int cap) {
    const int maxw = cap >= kFastBraGroups && cap <= kFastMaxWaves ? cap : kFastMaxWaves;
    const int fit = 64 * maxw / nhab;
    const int ppp = kab < fit ? kab : fit;
    const int w = (ppp * nhab + 63) / 64;
    return w < kFastBraGroups ? kFastBraGroups : w;
}

// Classes with a specialized kernel. The orientation chosen by the host puts
// the pair with fewer components on the bra side.
FastClassInfo fast_class_info(int la, int lb, int lc, int ld) {
#define GH_FAST(a, b, c, d) \
    if (la == a && lb == b && lc == c && ld == d) return fast_info_##a##b##c##d();
    GH_FAST(2, 2, 2, 2)
    GH_FAST(2, 2, 3, 3)
    GH_FAST(1, 1, 1, 1)
    GH_FAST(1, 0, 1, 0)
    GH_FAST(0, 0, 0, 0)
#undef GH_FAST
    return FastClassInfo{false, 0, 0, 0, 0, nullptr, nullptr};
}

}  // namespace gh
