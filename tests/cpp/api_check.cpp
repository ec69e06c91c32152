// Driver for the C++ host layer tests (include/giao_hip.hpp).
//   api_check host <dir>   checks that need no GPU
//   api_check gpu <dir>    evaluates the integrals of <dir>/shells.txt and
//                          writes raw complex128/float64 outputs to <dir>
// shells.txt: one shell per line: cx cy cz l nprim ncontr exps... coefs...
// Prints one "name value" line per result; exceptions print their type.
#include "giao_hip.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <typeinfo>

using namespace giao;

template <class F>
static void expect(const char* name, F&& f) {
    try {
        f();
        std::printf("%s ok\n", name);
    } catch (const BoysNumericalError& e) {
        std::printf("%s BoysNumericalError %s\n", name, e.what());
    } catch (const std::invalid_argument& e) {
        std::printf("%s invalid_argument %s\n", name, e.what());
    } catch (const std::length_error& e) {
        std::printf("%s length_error %s\n", name, e.what());
    } catch (const std::out_of_range& e) {
        std::printf("%s out_of_range %s\n", name, e.what());
    } catch (const std::runtime_error& e) {
        std::printf("%s runtime_error %s\n", name, e.what());
    }
}

static std::vector<Shell> read_shells(const std::string& path) {
    std::ifstream in(path);
    std::vector<Shell> out;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream s(line);
        double cx, cy, cz;
        int l, np, nc;
        if (!(s >> cx >> cy >> cz >> l >> np >> nc)) continue;
        std::vector<double> e(np), c(np * nc);
        for (auto& x : e) s >> x;
        for (auto& x : c) s >> x;
        out.emplace_back(Vec3{cx, cy, cz}, static_cast<std::uint16_t>(l), e, c, static_cast<std::size_t>(nc));
    }
    return out;
}

template <class T>
static void dump(const std::string& path, const std::vector<T>& v) {
    std::ofstream(path, std::ios::binary).write(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(T));
}

static void consume(const ShellQuartetBlockView& v, void* user) {
    auto* out = static_cast<std::vector<Complex>*>(user);
    out->insert(out->end(), v.values.begin(), v.values.end());
}

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const std::string mode = argv[1], dir = argv[2];
    if (mode == "host") {
        expect("shell_count", [] { Shell(Vec3{}, 1, {1.0, 2.0}, {1.0}); });
        expect("shell_exponent", [] { Shell(Vec3{}, 0, {-1.0}, {1.0}); });
        expect("primitive_exponent", [] { PrimitiveGaussian(0.0, Vec3{}); });
        const auto k = canonicalize_shell_quartet({3, 1, 0, 2});
        std::printf("canonical %u %u %u %u %d\n", k.shells.a, k.shells.b, k.shells.c, k.shells.d, int(k.conjugate));
        const Shell d(Vec3{}, 2, {1.0, 0.5}, {0.3, 0.7, 1.0, -1.0}, 2);
        std::printf("shell %zu %zu %zu %zu\n", d.primitive_count(), d.contraction_count(), d.cartesian_count(),
                    d.ao_count());
        const Vec3 kappa = MagneticField({0.1, -0.2, 0.3}, {1.0, 0.0, -1.0}).london_wave_vector({0.5, 2.0, -3.0});
        std::printf("kappa %.17g %.17g %.17g\n", kappa.x, kappa.y, kappa.z);
        expect("boys_without_gpu", [] { (void)boys_values({1.0, 0.5}, 4); });
        expect("nucleus_zero_charge", [] { Nucleus(0.0, Vec3{}); });
        expect("nucleus_negative_charge", [] { Nucleus(-1.0, Vec3{}); });
        expect("nucleus_nan_center", [] { Nucleus(8.0, Vec3{0.0, std::nan(""), 0.0}); });
        expect("nucleus_ok", [] { Nucleus(1.0, Vec3{0.0, 0.5, 0.0}); });
        std::printf("primitive_normalization %.17g\n", primitive_normalization(0.7, CartesianExponent{2, 1, 0}));
        expect("normalization_exponent", [] { (void)primitive_normalization(-1.0, CartesianExponent{}); });
        {
            const auto comps = cartesian_components(3);
            std::printf("components %zu", comps.size());
            for (const auto& c : comps) std::printf(" %u%u%u", c.x, c.y, c.z);
            std::printf("\n");
        }
        std::printf("max_threads %zu\n", openmp_max_threads());
        return 0;
    }
    const Basis basis(read_shells(dir + "/shells.txt"));
    const MagneticField field({0.1, -0.3, 0.25}, {0.2, 0.1, -0.1});
    const auto sh = basis.shells();
    // one block through compute_eri
    std::vector<Complex> block(shell_quartet_size(sh[0], sh[1], sh[2], sh[3]));
    EriWorkspace ws;
    compute_eri(sh[0], sh[1], sh[2], sh[3], field, block, ws);
    dump(dir + "/block.bin", block);
    // every quartet through evaluate_eri_shell_quartets, with and without screening
    const std::uint32_t ns = static_cast<std::uint32_t>(sh.size());
    std::vector<ShellQuartetIndex> all;
    for (std::uint32_t a = 0; a < ns; ++a)
        for (std::uint32_t b = 0; b < ns; ++b)
            for (std::uint32_t c = 0; c < ns; ++c)
                for (std::uint32_t d = 0; d < ns; ++d) all.push_back({a, b, c, d});
    std::vector<Complex> packed;
    auto st = evaluate_eri_shell_quartets(basis, all, field, {0.0, 4}, nullptr, consume, &packed);
    std::printf("stats %zu %zu %zu\n", st.requested_quartets, st.computed_quartets, st.screened_quartets);
    dump(dir + "/packed.bin", packed);
    const EriSchwarzBounds bounds(basis, field);
    dump(dir + "/bounds.bin", std::vector<double>(bounds.values().begin(), bounds.values().end()));
    double thr = 0.0;
    {
        std::vector<double> prods;
        for (const auto& q : all) prods.push_back(bounds(q.a, q.b) * bounds(q.c, q.d));
        std::nth_element(prods.begin(), prods.begin() + prods.size() / 2, prods.end());
        thr = prods[prods.size() / 2];
    }
    std::vector<Complex> screened;
    st = evaluate_eri_shell_quartets(basis, all, field, {thr, 1}, &bounds, consume, &screened);
    std::printf("screened %.17g %zu %zu %zu %zu\n", thr, st.requested_quartets, st.computed_quartets,
                st.screened_quartets, screened.size());
    // tensors
    const std::size_t n = basis.ao_count();
    std::vector<Complex> tensor(n * n * n * n), tensor2(n * n * n * n);
    compute_eri_tensor(basis, field, tensor);
    dump(dir + "/tensor.bin", tensor);
    st = compute_eri_tensor(basis, field, {0.0, 1}, nullptr, tensor2);
    dump(dir + "/tensor2.bin", tensor2);
    std::printf("tensor_stats %zu %zu\n", st.requested_quartets, st.computed_quartets);
    // Boys and one primitive integral
    std::vector<Complex> f(9);
    const auto diag = compute_boys({20.0, -35.0}, f);
    dump(dir + "/boys.bin", f);
    std::printf("boys_region %d\n", int(diag.region));
    const Complex p = primitive_eri(PrimitiveGaussian(0.8, {0.0, -0.2, 0.0}, {0, 0, 1}),
                                    PrimitiveGaussian(0.9, {0.1, -0.2, 0.3}, {1, 0, 1}),
                                    PrimitiveGaussian(1.0, {0.2, -0.2, 0.6}, {0, 0, 1}),
                                    PrimitiveGaussian(1.1, {0.3, -0.2, 0.9}, {1, 0, 1}), field);
    std::printf("primitive %.17g %.17g\n", p.real(), p.imag());
    // selected Cholesky integrals
    std::vector<std::array<std::size_t, 2>> pairs;
    for (std::size_t a = 0; a < ns; ++a)
        for (std::size_t b = a; b < ns; ++b) pairs.push_back({a, b});
    std::vector<double> d(n * n);
    detail::cholesky_diagonal(basis, pairs, field, d, 4);
    dump(dir + "/diag.bin", d);
    const std::array<std::size_t, 2> piv{1, 0};
    std::vector<Complex> col(n * n);
    detail::cholesky_columns(basis, std::span(&piv, 1), field, col, 4);
    dump(dir + "/column.bin", col);
    // warm compute_eri: 1000 calls over changing shell quartets reuse the
    // workspace's device basis, so no call allocates device memory
    {
        std::vector<Complex> warm(64 * 64 * 64);
        std::size_t max_size = 0;
        for (std::size_t a = 0; a < sh.size(); ++a) max_size = std::max(max_size, sh[a].ao_count());
        warm.resize(max_size * max_size * max_size * max_size);
        // first pass over every quartet of the basis grows the tables once
        for (std::size_t i = 0; i < 256; ++i) {
            const Shell &x = sh[i % 4], &y = sh[(i / 4) % 4], &z = sh[(i / 16) % 4], &w = sh[(i / 64) % 4];
            compute_eri(x, y, z, w, field, std::span(warm.data(), shell_quartet_size(x, y, z, w)), ws);
        }
        const unsigned long long before = giaohip_device_allocations();
        const auto t0 = std::chrono::steady_clock::now();
        for (std::size_t i = 0; i < 1000; ++i) {
            const Shell &x = sh[(i * 7) % 4], &y = sh[(i * 3 + 1) % 4], &z = sh[(i / 3) % 4], &w = sh[(i / 5) % 4];
            compute_eri(x, y, z, w, field, std::span(warm.data(), shell_quartet_size(x, y, z, w)), ws);
        }
        const auto t1 = std::chrono::steady_clock::now();
        const unsigned long long after = giaohip_device_allocations();
        // the last call's block, for the oracle comparison
        const std::size_t i = 999;
        std::printf("warm_last %zu %zu %zu %zu\n", (i * 7) % 4, (i * 3 + 1) % 4, (i / 3) % 4, (i / 5) % 4);
        const Shell &x = sh[(i * 7) % 4], &y = sh[(i * 3 + 1) % 4], &z = sh[(i / 3) % 4], &w = sh[(i / 5) % 4];
        dump(dir + "/warm_last.bin",
             std::vector<Complex>(warm.begin(), warm.begin() + (long)shell_quartet_size(x, y, z, w)));
        std::printf("warm_allocations %llu\n", after - before);
        std::printf("warm_us_per_call %.3f\n", std::chrono::duration<double, std::micro>(t1 - t0).count() / 1000.0);
    }
    // centre derivatives: one block, a packed batch, and the derivative bounds
    {
        std::vector<Complex> dblock(12 * shell_quartet_size(sh[0], sh[1], sh[2], sh[3]));
        compute_eri_center_derivatives(sh[0], sh[1], sh[2], sh[3], field, dblock, ws);
        dump(dir + "/deriv_block.bin", dblock);
        const std::vector<ShellQuartetIndex> dq{{0, 1, 2, 3}, {3, 3, 1, 0}, {2, 2, 2, 2}};
        std::vector<std::size_t> offs{0};
        for (const auto& q : dq)
            offs.push_back(offs.back() + 12 * shell_quartet_size(sh[q.a], sh[q.b], sh[q.c], sh[q.d]));
        std::vector<Complex> dpacked(offs.back());
        compute_eri_derivative_blocks(basis, dq, offs, field, dpacked);
        dump(dir + "/deriv_packed.bin", dpacked);
        const EriDerivativeSchwarzBounds dbounds(basis, field);
        dump(dir + "/deriv_bounds.bin", std::vector<double>(dbounds.values().begin(), dbounds.values().end()));
        std::printf("deriv_quartet_bound %.17g\n", dbounds.quartet_bound({0, 1, 2, 3}));
        expect("deriv_bad_offsets", [&] {
            std::vector<std::size_t> bad(offs);
            bad[1] += 1;
            compute_eri_derivative_blocks(basis, dq, bad, field, dpacked);
        });
        expect("deriv_bad_start", [&] {
            std::vector<std::size_t> bad(offs);
            bad[0] = 1;
            compute_eri_derivative_blocks(basis, dq, bad, field, dpacked);
        });
        expect("deriv_bad_index", [&] { (void)dbounds.quartet_bound({0, 0, 0, ns}); });
    }
    // error mapping
    expect("bad_output_size", [&] {
        std::vector<Complex> small(3);
        compute_eri(sh[0], sh[1], sh[2], sh[3], field, small, ws);
    });
    expect("bad_index", [&] {
        const ShellQuartetIndex q{0, 0, 0, ns};
        (void)evaluate_eri_shell_quartets(basis, std::span(&q, 1), field, {}, nullptr, consume, &packed);
    });
    expect("null_consumer", [&] { (void)evaluate_eri_shell_quartets(basis, all, field, {}, nullptr, nullptr, nullptr); });
    expect("screening_without_bounds",
           [&] { (void)evaluate_eri_shell_quartets(basis, all, field, {1e-8, 1}, nullptr, consume, &packed); });
    expect("bounds_other_field", [&] {
        (void)evaluate_eri_shell_quartets(basis, all, MagneticField{}, {1e-8, 1}, &bounds, consume, &packed);
    });
    expect("zero_threads", [&] { (void)evaluate_eri_shell_quartets(basis, all, field, {0.0, 0}, nullptr, consume, &packed); });
    expect("boys_outside_domain", [] { (void)boys_values({-400.0, 10.0}, 4); });
    expect("boys_order", [] { (void)boys_values({1.0, 0.0}, 33); });
    expect("real_columns_in_field", [&] {
        std::vector<double> r(n * n);
        detail::cholesky_columns(basis, std::span(&piv, 1), field, r, 1);
    });
    return 0;
}
