// Class dispatch of the batched transform. The kernels are instantiated in
// gh_fast_*.hip, one translation unit per group of classes.
#include "gh_internal.h"

namespace gh {

FastClassInfo fast_info_2222();
FastClassInfo fast_info_2233();
FastClassInfo fast_info_1111();
FastClassInfo fast_info_1010();
FastClassInfo fast_info_0000();
// groups of classes with la + lb + lc + ld = 3..8 (gh_fast_g*.hip)
FastClassInfo fast_info_g0(int la, int lb, int lc, int ld);
FastClassInfo fast_info_g1(int la, int lb, int lc, int ld);
FastClassInfo fast_info_g2(int la, int lb, int lc, int ld);
FastClassInfo fast_info_g3(int la, int lb, int lc, int ld);
FastClassInfo fast_info_g4(int la, int lb, int lc, int ld);
FastClassInfo fast_info_g5(int la, int lb, int lc, int ld);
FastClassInfo fast_info_g6(int la, int lb, int lc, int ld);
FastClassInfo fast_info_g7(int la, int lb, int lc, int ld);
FastClassInfo fast_info_g8(int la, int lb, int lc, int ld);
FastClassInfo fast_info_g9(int la, int lb, int lc, int ld);
// and the (fd) classes with la + lb + lc + ld = 9, 10
FastClassInfo fast_info_g10(int la, int lb, int lc, int ld);
FastClassInfo fast_info_g11(int la, int lb, int lc, int ld);
FastClassInfo fast_info_g12(int la, int lb, int lc, int ld);
// and the classes with an (ff) pair
FastClassInfo fast_info_g13(int la, int lb, int lc, int ld);
FastClassInfo fast_info_g14(int la, int lb, int lc, int ld);
FastClassInfo fast_info_g15(int la, int lb, int lc, int ld);
FastClassInfo fast_info_g16(int la, int lb, int lc, int ld);

int fast_waves(int nhab, int kab, int cap) {
    const int maxw = cap >= kFastBraGroups && cap <= kFastMaxWaves ? cap : kFastMaxWaves;
    const int fit = 64 * maxw / nhab;
    const int ppp = kab < fit ? kab : fit;
    const int w = (ppp * nhab + 63) / 64;
    return w < kFastBraGroups ? kFastBraGroups : w;
}

// Classes with a specialized kernel: every class with l <= 3, except those
// with la + lb + lc + ld <= 2 (the lane-per-primitive-quartet kernel), through
// one of its images (ab|cd), (cd|ab), conj (ba|dc), conj (dc|ba); the host
// takes the first image listed here. The groups hold the image whose bra pair
// has the more Hermite indices (the transform's rows); (dd|ff) keeps dd on the
// bra side.
FastClassInfo fast_class_info(int la, int lb, int lc, int ld) {
#define GH_FAST(a, b, c, d) \
    if (la == a && lb == b && lc == c && ld == d) return fast_info_##a##b##c##d();
    GH_FAST(2, 2, 2, 2)
    GH_FAST(2, 2, 3, 3)
    GH_FAST(1, 1, 1, 1)
    GH_FAST(1, 0, 1, 0)
    GH_FAST(0, 0, 0, 0)
#undef GH_FAST
    for (FastClassInfo (*group)(int, int, int, int) :
         {fast_info_g0, fast_info_g1, fast_info_g2, fast_info_g3, fast_info_g4, fast_info_g5, fast_info_g6,
          fast_info_g7, fast_info_g8, fast_info_g9, fast_info_g10, fast_info_g11, fast_info_g12,
          fast_info_g13, fast_info_g14, fast_info_g15, fast_info_g16}) {
        const FastClassInfo f = group(la, lb, lc, ld);
        if (f.available) return f;
    }
    return FastClassInfo{false, 0, 0, 0, 0, nullptr, nullptr};
}

}  // namespace gh
