// C entry points over the compiled reference (GIAOINT src/*.cpp built from
// /root/reference by oracle/Makefile into oracle/_ref/). TEST INFRASTRUCTURE
// ONLY: used to pin the C oracle and as the "reference" CPU baseline. The
// signatures equal those of giao_oracle.h with the prefix ref_ instead of or_.
#include "giao_integrals/boys.hpp"
#include "giao_integrals/derivatives.hpp"
#include "giao_integrals/eri.hpp"
#include "giao_integrals/nuclear.hpp"
#include "cholesky.hpp"

#include <array>
#include <cstdio>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <vector>

namespace {

thread_local char g_msg[512];

int report(int code, const char* what) {
    std::snprintf(g_msg, sizeof g_msg, "%s", what);
    return code;
}

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const giao::BoysNumericalError& e) {
        return report(4, e.what());
    } catch (const std::invalid_argument& e) {
        return report(1, e.what());
    } catch (const std::length_error& e) {
        return report(2, e.what());
    } catch (const std::out_of_range& e) {
        return report(3, e.what());
    } catch (const std::exception& e) {
        return report(5, e.what());
    }
}

struct RefBasis {
    int nshell;
    const double* centers;
    const int* L;
    const int* nprim;
    const int* ncontr;
    const double* exps;
    const double* coefs;
};

giao::Basis make_basis(const RefBasis* b) {
    std::vector<giao::Shell> shells;
    std::size_t po = 0, co = 0;
    for (int s = 0; s < b->nshell; ++s) {
        const auto np = static_cast<std::size_t>(b->nprim[s]);
        const auto nc = static_cast<std::size_t>(b->ncontr[s]);
        std::vector<double> e(b->exps + po, b->exps + po + np);
        std::vector<double> c(b->coefs + co, b->coefs + co + np * nc);
        shells.emplace_back(giao::Vec3{b->centers[3 * s], b->centers[3 * s + 1], b->centers[3 * s + 2]},
                            static_cast<std::uint16_t>(b->L[s]), std::move(e), std::move(c), nc,
                            giao::ContractionNormalization::as_provided);
        po += np;
        co += np * nc;
    }
    return giao::Basis(std::move(shells));
}

giao::MagneticField make_field(const double f[6]) {
    return giao::MagneticField({f[0], f[1], f[2]}, {f[3], f[4], f[5]});
}

struct PackData {
    const std::uint64_t* offsets;
    double* out;
    std::size_t index;
};

void pack_block(const giao::ShellQuartetBlockView& block, void* user) {
    auto& d = *static_cast<PackData*>(user);
    auto* dst = reinterpret_cast<giao::Complex*>(d.out) + d.offsets[d.index];
    std::copy(block.values.begin(), block.values.end(), dst);
    ++d.index;
}

}  // namespace

extern "C" {

const char* ref_last_error() { return g_msg; }

int ref_normalize_shell(int L, int nprim, int ncontr, const double* exps, const double* coefs_in,
                        double* coefs_out) {
    return guarded([&] {
        giao::Shell s({0, 0, 0}, static_cast<std::uint16_t>(L),
                      std::vector<double>(exps, exps + nprim),
                      std::vector<double>(coefs_in, coefs_in + nprim * ncontr),
                      static_cast<std::size_t>(ncontr), giao::ContractionNormalization::normalize);
        std::copy(s.coefficients().begin(), s.coefficients().end(), coefs_out);
    });
}

int ref_boys(double zr, double zi, int nmax, int scaled, double* out, int* region, double* err, int* segments) {
    return guarded([&] {
        std::vector<giao::Complex> v(static_cast<std::size_t>(nmax) + 1U);
        const auto d = giao::compute_boys({zr, zi}, v, scaled ? giao::BoysScaling::exp_z : giao::BoysScaling::unscaled);
        std::memcpy(out, v.data(), v.size() * sizeof(giao::Complex));
        if (region) *region = static_cast<int>(d.region);
        if (err) *err = d.estimated_absolute_error;
        if (segments) *segments = static_cast<int>(d.quadrature_segments);
    });
}

int ref_primitive_eri(const double prim[4][6], const int ang[4][3], const double field[6], double out[2]) {
    return guarded([&] {
        std::vector<giao::PrimitiveGaussian> p;
        for (int i = 0; i < 4; ++i) {
            p.emplace_back(prim[i][0], giao::Vec3{prim[i][1], prim[i][2], prim[i][3]},
                           giao::CartesianExponent{static_cast<std::uint16_t>(ang[i][0]),
                                                   static_cast<std::uint16_t>(ang[i][1]),
                                                   static_cast<std::uint16_t>(ang[i][2])},
                           prim[i][4], prim[i][5] != 0.0);
        }
        const auto v = giao::primitive_eri(p[0], p[1], p[2], p[3], make_field(field));
        out[0] = v.real();
        out[1] = v.imag();
    });
}

int ref_eri_batch(const RefBasis* b, const std::uint32_t* quartets, std::size_t n, const double field[6],
                  const std::uint64_t* offsets, double* out, int threads) {
    return guarded([&] {
        const auto basis = make_basis(b);
        std::vector<giao::ShellQuartetIndex> q(n);
        for (std::size_t i = 0; i < n; ++i)
            q[i] = {quartets[4 * i], quartets[4 * i + 1], quartets[4 * i + 2], quartets[4 * i + 3]};
        PackData d{offsets, out, 0};
        giao::EriEvaluationOptions opt;
        opt.thread_count = static_cast<std::size_t>(threads);
        [[maybe_unused]] const auto st =
            giao::evaluate_eri_shell_quartets(basis, q, make_field(field), opt, nullptr, pack_block, &d);
    });
}

int ref_schwarz_bounds(const RefBasis* b, const double field[6], double* out) {
    return guarded([&] {
        const auto basis = make_basis(b);
        const giao::EriSchwarzBounds bounds(basis, make_field(field));
        const auto v = bounds.values();
        std::copy(v.begin(), v.end(), out);
    });
}

int ref_eri_tensor(const RefBasis* b, const double field[6], double* out, int threads) {
    return guarded([&] {
        const auto basis = make_basis(b);
        const auto n = basis.ao_count();
        giao::EriEvaluationOptions opt;
        opt.thread_count = static_cast<std::size_t>(threads);
        [[maybe_unused]] const auto st = giao::compute_eri_tensor(
            basis, make_field(field), opt, nullptr,
            std::span<giao::Complex>(reinterpret_cast<giao::Complex*>(out), n * n * n * n));
    });
}

int ref_cholesky_diagonal(const RefBasis* b, const std::uint64_t* pairs, std::size_t npairs, const double field[6],
                          double* out, int threads) {
    return guarded([&] {
        const auto basis = make_basis(b);
        std::vector<std::array<std::size_t, 2>> p(npairs);
        for (std::size_t i = 0; i < npairs; ++i) p[i] = {pairs[2 * i], pairs[2 * i + 1]};
        const auto n = basis.ao_count();
        giao::detail::cholesky_diagonal(basis, p, make_field(field), std::span<double>(out, n * n),
                                        static_cast<std::size_t>(threads));
    });
}

int ref_cholesky_columns(const RefBasis* b, const std::uint64_t* pivots, std::size_t npiv, const double field[6],
                         double* out, int real_output, int threads) {
    return guarded([&] {
        const auto basis = make_basis(b);
        std::vector<std::array<std::size_t, 2>> p(npiv);
        for (std::size_t i = 0; i < npiv; ++i) p[i] = {pivots[2 * i], pivots[2 * i + 1]};
        const auto n = basis.ao_count();
        if (real_output) {
            giao::detail::cholesky_columns(basis, p, make_field(field), std::span<double>(out, npiv * n * n),
                                           static_cast<std::size_t>(threads));
        } else {
            giao::detail::cholesky_columns(
                basis, p, make_field(field),
                std::span<giao::Complex>(reinterpret_cast<giao::Complex*>(out), npiv * n * n),
                static_cast<std::size_t>(threads));
        }
    });
}

int ref_nuclear_matrix(const RefBasis* b, int nnuc, const double* charges, const double* centers,
                       const double field[6], double* out) {
    return guarded([&] {
        const auto basis = make_basis(b);
        const auto n = basis.ao_count();
        std::vector<giao::Nucleus> nuclei;
        for (int c = 0; c < nnuc; ++c)
            nuclei.emplace_back(charges[c], giao::Vec3{centers[3 * c], centers[3 * c + 1], centers[3 * c + 2]});
        giao::compute_nuclear_attraction_matrix(basis, nuclei, make_field(field),
                                                std::span<giao::Complex>(reinterpret_cast<giao::Complex*>(out), n * n));
    });
}

int ref_eri_derivative_blocks(const RefBasis* b, const std::uint32_t* quartets, std::size_t n, const double field[6],
                              const std::uint64_t* offsets, double* out, std::size_t out_len, int threads) {
    return guarded([&] {
        const auto basis = make_basis(b);
        std::vector<giao::ShellQuartetIndex> q(n);
        for (std::size_t i = 0; i < n; ++i)
            q[i] = {quartets[4 * i], quartets[4 * i + 1], quartets[4 * i + 2], quartets[4 * i + 3]};
        const std::vector<std::size_t> offs(offsets, offsets + n + 1);
        giao::compute_eri_derivative_blocks(basis, q, offs, make_field(field),
                                            {reinterpret_cast<giao::Complex*>(out), out_len},
                                            static_cast<std::size_t>(threads));
    });
}

int ref_eri_derivative_pair_bounds(const RefBasis* b, const double field[6], double* out) {
    return guarded([&] {
        const auto basis = make_basis(b);
        const giao::EriDerivativeSchwarzBounds bounds(basis, make_field(field));
        const auto v = bounds.values();
        std::copy(v.begin(), v.end(), out);
    });
}

int ref_eri_derivative_quartet_bounds(const RefBasis* b, const double field[6], const std::uint32_t* quartets,
                                      std::size_t n, double* out) {
    return guarded([&] {
        const auto basis = make_basis(b);
        const giao::EriDerivativeSchwarzBounds bounds(basis, make_field(field));
        for (std::size_t i = 0; i < n; ++i)
            out[i] = bounds.quartet_bound({quartets[4 * i], quartets[4 * i + 1], quartets[4 * i + 2], quartets[4 * i + 3]});
    });
}

}  // extern "C"
