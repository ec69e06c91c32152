// Compile-time plans of the class-specialized ERI transform (gh_eri_fast.hip).
//
// For a primitive pair of angular momenta (L1, L2) the complex Hermite
// product of Cartesian component c = (a, b) at Hermite index h = (t, u, v)
// is nonzero only inside the box t <= ax+bx, u <= ay+by, v <= az+bz. A plan
// lists these (h, c) entries sorted by h, so that a kernel walking h in order
// reads the products sequentially (wave-uniform scalar loads) and, for each
// h, touches only the components whose box contains h.
#pragma once

#include <array>
#include <cstdint>

namespace gh::plan {

constexpr int nc(int l) { return (l + 1) * (l + 2) / 2; }
constexpr int nh(int l) { return (l + 1) * (l + 2) * (l + 3) / 6; }

struct P3 {
    int x, y, z;
};

// x-major Cartesian component c of shell l (types.cpp:221-237).
constexpr P3 cart(int l, int c) {
    int k = 0;
    for (int x = l; x >= 0; --x)
        for (int y = l - x; y >= 0; --y) {
            if (k == c) return P3{x, y, l - x - y};
            ++k;
        }
    return P3{0, 0, 0};
}

// Hermite index h of degree <= D in lexicographic (t, u, v) order.
constexpr P3 herm(int D, int h) {
    int k = 0;
    for (int t = 0; t <= D; ++t)
        for (int u = 0; u + t <= D; ++u)
            for (int v = 0; v + u + t <= D; ++v) {
                if (k == h) return P3{t, u, v};
                ++k;
            }
    return P3{0, 0, 0};
}

// lexicographic offset of (t, u, 0) among Hermite indices of degree <= D
constexpr int lexoff(int D, int t, int u) { return nh(D) - nh(D - t) + u * (D - t + 1) - u * (u - 1) / 2; }
constexpr int lexidx(int D, int t, int u, int v) { return lexoff(D, t, u) + v; }

// Component c (index into the pair's nc(L1)*nc(L2) list, bra-major) box test.
constexpr bool in_box(int L1, int L2, int c, P3 h) {
    const P3 a = cart(L1, c / nc(L2)), b = cart(L2, c % nc(L2));
    return h.x <= a.x + b.x && h.y <= a.y + b.y && h.z <= a.z + b.z;
}

// 1/sqrt((2x-1)!! (2y-1)!! (2z-1)!!): the component part of the primitive
// normalization (types.cpp:207-219 factorized into radial x angular parts).
constexpr double angular_norm(P3 p) {
    auto dfact = [](int n) {
        double r = 1.0;
        for (int k = 2 * n - 1; k > 1; k -= 2) r *= (double)k;
        return r;
    };
    const double d = dfact(p.x) * dfact(p.y) * dfact(p.z);
    double s = d > 1.0 ? d : 1.0;  // Newton for sqrt(d) from above
    for (int i = 0; i < 200; ++i) s = 0.5 * (s + d / s);
    return 1.0 / s;
}

// Components of a pair split into NG groups: MODE 0 contiguous chunks of
// ceil(NC/NG), MODE 1 round robin (c % NG). Entries are sorted by group, then
// Hermite index, then component; start[g][h] indexes the flattened arrays.
template <int L1, int L2, int NG, int MODE>
struct Side {
    static constexpr int D = L1 + L2;
    static constexpr int NH = nh(D);
    static constexpr int NC = nc(L1) * nc(L2);
    static constexpr int CH = (NC + NG - 1) / NG;
    static constexpr bool member(int g, int c) { return MODE == 0 ? (c / CH == g) : (c % NG == g); }
    struct Boxes {
        int bx[NC], by[NC], bz[NC];
    };
    static constexpr Boxes boxes() {
        Boxes b{};
        for (int c = 0; c < NC; ++c) {
            const P3 x = cart(L1, c / nc(L2)), y = cart(L2, c % nc(L2));
            b.bx[c] = x.x + y.x;
            b.by[c] = x.y + y.y;
            b.bz[c] = x.z + y.z;
        }
        return b;
    }
    template <class F>
    static constexpr void walk(F&& f) {  // (g, h, c) in plan order
        const Boxes b = boxes();
        for (int g = 0; g < NG; ++g) {
            int h = 0;
            for (int t = 0; t <= D; ++t)
                for (int u = 0; u + t <= D; ++u)
                    for (int v = 0; v + u + t <= D; ++v, ++h) {
                        f(g, h, -1);
                        for (int c = 0; c < NC; ++c)
                            if (member(g, c) && t <= b.bx[c] && u <= b.by[c] && v <= b.bz[c]) f(g, h, c);
                    }
        }
    }
    static constexpr int count() {
        int n = 0;
        walk([&](int, int, int c) {
            if (c >= 0) ++n;
        });
        return n;
    }
    static constexpr int NE = count();
    // Stored tables pad every group to a multiple of kBatch entries, so that a
    // wavefront loads one batch with one 16-byte load per lane.
    static constexpr int kBatch = 16;
    static constexpr int padded_count() {
        int n = 0;
        for (int g = 0; g < NG; ++g) {
            int cnt = 0;
            walk([&](int gg, int, int c) {
                if (gg == g && c >= 0) ++cnt;
            });
            n += (cnt + kBatch - 1) / kBatch * kBatch;
        }
        return n;
    }
    static constexpr int NEP = padded_count();
    struct Tables {
        std::array<uint16_t, NE> comp{};
        std::array<uint16_t, NE> hidx{};
        std::array<std::array<uint16_t, NH + 1>, NG> start{};
        std::array<uint16_t, NG> pbase{};  // padded start of each group
        std::array<uint16_t, NEP> ent{};   // entry stored in a padded slot, 0xffff for padding
    };
    static constexpr Tables make() {
        Tables t{};
        int n = 0;
        walk([&](int g, int h, int c) {
            if (c < 0) {
                t.start[g][h] = (uint16_t)n;
            } else {
                t.comp[n] = (uint16_t)c;
                t.hidx[n] = (uint16_t)h;
                ++n;
            }
        });
        // close each group's start table
        int acc = 0;
        for (int g = 0; g < NG; ++g) {
            int cnt = 0;
            for (int e = 0; e < NE; ++e) {
                const int c = t.comp[e];
                if (member(g, c)) ++cnt;
            }
            acc += cnt;
            t.start[g][NH] = (uint16_t)acc;
        }
        for (int p = 0; p < NEP; ++p) t.ent[p] = 0xffff;
        int pb = 0;
        for (int g = 0; g < NG; ++g) {
            t.pbase[g] = (uint16_t)pb;
            const int b = t.start[g][0], e = t.start[g][NH];
            for (int i = b; i < e; ++i) t.ent[pb + (i - b)] = (uint16_t)i;
            pb += (e - b + kBatch - 1) / kBatch * kBatch;
        }
        return t;
    }
    static constexpr Tables T = make();
};

}  // namespace gh::plan
