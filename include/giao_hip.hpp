// C++20 host layer over the giaohip C ABI with the names, signatures and
// exception types of GIAOINT's ERI and Cholesky entry points
// (include/giao_integrals/types.hpp:12-140, boys.hpp:12-42, eri.hpp:19-146,
// src/cholesky.hpp:9-17), so that code written against those headers
// compiles unchanged and runs its integrals on the GPU.
//
// Header only: include it and link libgiaohip.so. All calls go to one
// process-wide context on device 0, created on first use. A Basis uploads
// its shell tables on first use and keeps them while any copy of it lives.
// EriWorkspace exists for signature compatibility and holds nothing (the
// device work buffers live in the context); thread counts are validated as
// in the reference and otherwise ignored (results never depend on them).
#pragma once

#include "giaohip.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>
#include <new>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

namespace giao {

using Complex = std::complex<double>;

struct Vec3 {
    double x{};
    double y{};
    double z{};
    [[nodiscard]] constexpr double operator[](std::size_t axis) const { return axis == 0 ? x : (axis == 1 ? y : z); }
};

struct CartesianExponent {
    std::uint16_t x{};
    std::uint16_t y{};
    std::uint16_t z{};
    [[nodiscard]] constexpr std::uint32_t total() const noexcept { return std::uint32_t{x} + y + z; }
    [[nodiscard]] constexpr std::uint16_t operator[](std::size_t axis) const {
        return axis == 0 ? x : (axis == 1 ? y : z);
    }
};

class BoysNumericalError : public std::runtime_error {
  public:
    using std::runtime_error::runtime_error;
};

namespace hip_detail {

// Status of a C ABI call -> the reference's exception type, with the message
// the library recorded.
inline void check(giaohip_ctx* ctx, int status) {
    if (status == GIAOHIP_OK) return;
    char buf[1024] = {0};
    if (ctx) giaohip_last_error(ctx, buf, sizeof buf);
    const std::string msg = buf[0] ? buf : "giaohip call failed";
    switch (status) {
        case GIAOHIP_INVALID_ARGUMENT: throw std::invalid_argument(msg);
        case GIAOHIP_LENGTH_ERROR: throw std::length_error(msg);
        case GIAOHIP_OUT_OF_RANGE: throw std::out_of_range(msg);
        case GIAOHIP_BOYS_ERROR: throw BoysNumericalError(msg);
        case GIAOHIP_OUT_OF_MEMORY: throw std::bad_alloc();
        default: throw std::runtime_error(msg);
    }
}

inline giaohip_ctx* context() {
    static std::once_flag once;
    static giaohip_ctx* ctx = nullptr;
    static int status = GIAOHIP_OK;
    std::call_once(once, [] { status = giaohip_create(0, &ctx); });
    if (status != GIAOHIP_OK || !ctx) throw std::runtime_error("no usable GPU for the giaohip library (device 0)");
    return ctx;
}

inline std::array<double, 6> field6(const Vec3& B, const Vec3& O) { return {B.x, B.y, B.z, O.x, O.y, O.z}; }

inline std::vector<CartesianExponent> cartesian(std::uint16_t l) {
    std::vector<CartesianExponent> out;
    for (int x = l; x >= 0; --x)
        for (int y = l - x; y >= 0; --y)
            out.push_back({static_cast<std::uint16_t>(x), static_cast<std::uint16_t>(y),
                           static_cast<std::uint16_t>(l - x - y)});
    return out;
}

}  // namespace hip_detail

// primitive_normalization (types.hpp:141-142, types.cpp:60-72): the norm of
// a Cartesian primitive x^ax y^ay z^az exp(-alpha r^2).
[[nodiscard]] inline double primitive_normalization(double exponent, CartesianExponent angular) {
    if (!std::isfinite(exponent) || exponent <= 0.0)
        throw std::invalid_argument("Gaussian exponent must be finite and positive");
    return giaohip_primitive_normalization(exponent, angular.x, angular.y, angular.z);
}

// cartesian_components (types.hpp:144-145, types.cpp:74-90): x-major order,
// x descending, then y descending.
[[nodiscard]] inline std::vector<CartesianExponent> cartesian_components(std::uint16_t total_angular_momentum) {
    return hip_detail::cartesian(total_angular_momentum);
}

struct PrimitiveGaussian {
    double exponent{};
    double coefficient{1.0};
    Vec3 center{};
    CartesianExponent angular{};
    bool normalized{true};

    PrimitiveGaussian(double exponent_in, Vec3 center_in, CartesianExponent angular_in = {},
                      double coefficient_in = 1.0, bool normalized_in = true)
        : exponent(exponent_in), coefficient(coefficient_in), center(center_in), angular(angular_in),
          normalized(normalized_in) {
        if (!std::isfinite(exponent) || exponent <= 0.0)
            throw std::invalid_argument("Gaussian exponent must be finite and positive");
        if (!std::isfinite(coefficient)) throw std::invalid_argument("Gaussian coefficient must be finite");
    }
};

enum class ContractionNormalization { normalize, as_provided };

class Shell {
  public:
    Shell(Vec3 center, std::uint16_t angular_momentum, std::vector<double> exponents,
          std::vector<double> coefficients, std::size_t contraction_count = 1,
          ContractionNormalization normalization = ContractionNormalization::normalize)
        : center_(center), angular_momentum_(angular_momentum), exponents_(std::move(exponents)),
          coefficients_(std::move(coefficients)), contraction_count_(contraction_count),
          normalization_(normalization), components_(hip_detail::cartesian(angular_momentum)) {
        if (contraction_count_ == 0 || coefficients_.size() != contraction_count_ * exponents_.size())
            throw std::invalid_argument("shell coefficients must have contraction_count * primitive_count values");
        const int L = angular_momentum_, np = static_cast<int>(exponents_.size()),
                  nc = static_cast<int>(contraction_count_);
        std::vector<double> normalized(coefficients_.size());
        char msg[512] = {0};
        // validation (and the contraction normalization) by the library's host code
        if (giaohip_normalize_shells(1, &L, &np, &nc, exponents_.data(), coefficients_.data(), normalized.data(), msg,
                                     sizeof msg) != GIAOHIP_OK)
            throw std::invalid_argument(msg);
        if (normalization_ == ContractionNormalization::normalize) coefficients_ = std::move(normalized);
    }

    [[nodiscard]] Vec3 center() const noexcept { return center_; }
    [[nodiscard]] std::uint16_t angular_momentum() const noexcept { return angular_momentum_; }
    [[nodiscard]] std::size_t primitive_count() const noexcept { return exponents_.size(); }
    [[nodiscard]] std::size_t contraction_count() const noexcept { return contraction_count_; }
    [[nodiscard]] std::size_t cartesian_count() const noexcept { return components_.size(); }
    [[nodiscard]] std::size_t ao_count() const noexcept { return contraction_count_ * cartesian_count(); }
    [[nodiscard]] ContractionNormalization normalization_policy() const noexcept { return normalization_; }
    [[nodiscard]] std::span<const double> exponents() const noexcept { return exponents_; }
    [[nodiscard]] std::span<const double> coefficients() const noexcept { return coefficients_; }
    [[nodiscard]] std::span<const CartesianExponent> components() const noexcept { return components_; }
    [[nodiscard]] double coefficient(std::size_t contraction, std::size_t primitive) const noexcept {
        return coefficients_[contraction * primitive_count() + primitive];
    }

  private:
    Vec3 center_{};
    std::uint16_t angular_momentum_{};
    std::vector<double> exponents_;
    std::vector<double> coefficients_;
    std::size_t contraction_count_{};
    ContractionNormalization normalization_{};
    std::vector<CartesianExponent> components_;
};

class Basis {
  public:
    explicit Basis(std::vector<Shell> shells) : shells_(std::move(shells)) {
        std::size_t off = 0;
        for (const Shell& s : shells_) {
            ao_offsets_.push_back(off);
            off += s.ao_count();
        }
        ao_count_ = off;
        device_ = std::make_shared<Device>();
    }
    [[nodiscard]] std::span<const Shell> shells() const noexcept { return shells_; }
    [[nodiscard]] std::span<const std::size_t> ao_offsets() const noexcept { return ao_offsets_; }
    [[nodiscard]] std::size_t ao_count() const noexcept { return ao_count_; }

    // The uploaded shell tables (giaohip_basis), created on first use.
    [[nodiscard]] giaohip_basis* device() const {
        std::lock_guard<std::mutex> lock(device_->mu);
        if (!device_->handle) {
            giaohip_ctx* ctx = hip_detail::context();
            std::vector<double> centers, exps, coefs;
            std::vector<int> L, np, nc;
            for (const Shell& s : shells_) {
                centers.insert(centers.end(), {s.center().x, s.center().y, s.center().z});
                L.push_back(s.angular_momentum());
                np.push_back(static_cast<int>(s.primitive_count()));
                nc.push_back(static_cast<int>(s.contraction_count()));
                exps.insert(exps.end(), s.exponents().begin(), s.exponents().end());
                coefs.insert(coefs.end(), s.coefficients().begin(), s.coefficients().end());
            }
            giaohip_basis* b = nullptr;
            hip_detail::check(ctx, giaohip_basis_create(ctx, static_cast<int>(shells_.size()), centers.data(),
                                                        L.data(), np.data(), nc.data(), exps.data(), coefs.data(),
                                                        &b));
            device_->handle = b;
        }
        return device_->handle;
    }

  private:
    struct Device {
        std::mutex mu;
        giaohip_basis* handle = nullptr;
        ~Device() {
            if (handle) giaohip_basis_destroy(hip_detail::context(), handle);
        }
    };
    std::vector<Shell> shells_;
    std::vector<std::size_t> ao_offsets_;
    std::size_t ao_count_{};
    std::shared_ptr<Device> device_;
};

struct MagneticField {
    Vec3 B{};
    Vec3 gauge_origin{};
    MagneticField(Vec3 field = {}, Vec3 origin = {}) : B(field), gauge_origin(origin) {}
    // kappa = 1/2 B x (C - O) (types.cpp:179-186)
    [[nodiscard]] Vec3 london_wave_vector(Vec3 c) const noexcept {
        const Vec3 d{c.x - gauge_origin.x, c.y - gauge_origin.y, c.z - gauge_origin.z};
        return {0.5 * (B.y * d.z - B.z * d.y), 0.5 * (B.z * d.x - B.x * d.z), 0.5 * (B.x * d.y - B.y * d.x)};
    }
};

// --- Boys function (boys.hpp:12-42) -----------------------------------------
enum class BoysScaling { unscaled, exp_z };
enum class BoysRegion { power_series, adaptive_quadrature, scaled_quadrature, positive_asymptotic };
struct BoysDiagnostics {
    BoysRegion region{};
    BoysScaling scaling{};
    double estimated_absolute_error{};
    std::size_t quadrature_segments{};
};

[[nodiscard]] inline BoysDiagnostics compute_boys(Complex argument, std::span<Complex> values,
                                                  BoysScaling scaling = BoysScaling::unscaled) {
    if (values.empty()) throw std::invalid_argument("Boys output must contain at least F_0");
    giaohip_ctx* ctx = hip_detail::context();
    const double z[2] = {argument.real(), argument.imag()};
    int region = 0, segments = 0;
    double err = 0.0;
    hip_detail::check(ctx, giaohip_boys(ctx, z, 1, static_cast<int>(values.size() - 1),
                                        scaling == BoysScaling::exp_z ? 1 : 0,
                                        reinterpret_cast<double*>(values.data()), &region, &err, &segments));
    return {static_cast<BoysRegion>(region), scaling, err, static_cast<std::size_t>(segments)};
}

[[nodiscard]] inline std::vector<Complex> boys_values(Complex argument, std::size_t maximum_order,
                                                      BoysScaling scaling = BoysScaling::unscaled) {
    std::vector<Complex> v(maximum_order + 1);
    [[maybe_unused]] const auto d = compute_boys(argument, v, scaling);
    return v;
}

// --- ERIs (eri.hpp:19-146) ---------------------------------------------------
// Per-caller scratch of compute_eri (eri.hpp:19-46): a four-shell device
// basis that each call refills in place (giaohip_basis_update). Its device
// tables only grow, so once warm a call allocates no device memory, as the
// reference's workspace allocates nothing after warm-up
// (tests/cpp/test_overlap.cpp:387-401). Not shared between threads.
class EriWorkspace {
  public:
    EriWorkspace() = default;
    EriWorkspace(const EriWorkspace&) = delete;
    EriWorkspace& operator=(const EriWorkspace&) = delete;
    ~EriWorkspace() {
        if (basis_) giaohip_basis_destroy(hip_detail::context(), basis_);
    }
    void prepare(std::array<std::size_t, 7>, std::size_t, std::array<std::size_t, 4> = {}) {}

    // The scratch basis holding the four given shells.
    giaohip_basis* load(giaohip_ctx* ctx, const std::array<const Shell*, 4>& shells);

  private:
    giaohip_basis* basis_ = nullptr;
    std::vector<double> centers_, exps_, coefs_;
    std::vector<int> L_, np_, nc_;
};

struct ShellQuartetIndex {
    std::uint32_t a{};
    std::uint32_t b{};
    std::uint32_t c{};
    std::uint32_t d{};
    [[nodiscard]] constexpr auto as_array() const noexcept { return std::array<std::uint32_t, 4>{a, b, c, d}; }
};

struct CanonicalShellQuartet {
    ShellQuartetIndex shells{};
    bool conjugate{};
};

struct ShellQuartetBlockView {
    ShellQuartetIndex shells{};
    std::array<std::size_t, 4> shape{};
    std::span<const Complex> values{};
};

using QuartetConsumer = void (*)(const ShellQuartetBlockView&, void* user_data);

struct EriEvaluationOptions {
    double screening_threshold{};
    std::size_t thread_count{1U};
};

struct EriStatistics {
    std::size_t requested_quartets{};
    std::size_t computed_quartets{};
    std::size_t screened_quartets{};
};

class EriSchwarzBounds {
  public:
    explicit EriSchwarzBounds(const Basis& basis, const MagneticField& field = {})
        : basis_identity_(&basis), field_(field), shell_count_(basis.shells().size()),
          values_(shell_count_ * shell_count_) {
        giaohip_ctx* ctx = hip_detail::context();
        const auto f = hip_detail::field6(field.B, field.gauge_origin);
        hip_detail::check(ctx, giaohip_schwarz_bounds(ctx, basis.device(), f.data(), values_.data()));
    }
    [[nodiscard]] std::size_t shell_count() const noexcept { return shell_count_; }
    [[nodiscard]] double operator()(std::size_t a, std::size_t b) const {
        if (a >= shell_count_ || b >= shell_count_)
            throw std::out_of_range("ERI Schwarz shell-pair index is out of range");
        return values_[a * shell_count_ + b];
    }
    [[nodiscard]] std::span<const double> values() const noexcept { return values_; }
    [[nodiscard]] bool matches(const Basis& basis, const MagneticField& f) const noexcept {
        return basis_identity_ == &basis && field_.B.x == f.B.x && field_.B.y == f.B.y && field_.B.z == f.B.z &&
               field_.gauge_origin.x == f.gauge_origin.x && field_.gauge_origin.y == f.gauge_origin.y &&
               field_.gauge_origin.z == f.gauge_origin.z;
    }

  private:
    const Basis* basis_identity_{};
    MagneticField field_{};
    std::size_t shell_count_{};
    std::vector<double> values_;
};

// Lexicographic minimum of {abcd, cdab, badc*, dcba*} (eri.cpp:414-429).
[[nodiscard]] inline CanonicalShellQuartet canonicalize_shell_quartet(ShellQuartetIndex q) noexcept {
    const std::array<std::array<std::uint32_t, 4>, 4> images{{{q.a, q.b, q.c, q.d},
                                                              {q.c, q.d, q.a, q.b},
                                                              {q.b, q.a, q.d, q.c},
                                                              {q.d, q.c, q.b, q.a}}};
    std::size_t best = 0;
    for (std::size_t i = 1; i < 4; ++i)
        if (images[i] < images[best]) best = i;
    const auto& m = images[best];
    return {{m[0], m[1], m[2], m[3]}, best >= 2};
}

[[nodiscard]] inline std::size_t shell_quartet_size(const Shell& a, const Shell& b, const Shell& c, const Shell& d) {
    return a.ao_count() * b.ao_count() * c.ao_count() * d.ao_count();
}

[[nodiscard]] inline Complex primitive_eri(const PrimitiveGaussian& a, const PrimitiveGaussian& b,
                                           const PrimitiveGaussian& c, const PrimitiveGaussian& d,
                                           const MagneticField& field = {}) {
    giaohip_ctx* ctx = hip_detail::context();
    double prim[4][6];
    int ang[4][3];
    const PrimitiveGaussian* p[4] = {&a, &b, &c, &d};
    for (int i = 0; i < 4; ++i) {
        prim[i][0] = p[i]->exponent;
        prim[i][1] = p[i]->center.x;
        prim[i][2] = p[i]->center.y;
        prim[i][3] = p[i]->center.z;
        prim[i][4] = p[i]->coefficient;
        prim[i][5] = p[i]->normalized ? 1.0 : 0.0;
        ang[i][0] = p[i]->angular.x;
        ang[i][1] = p[i]->angular.y;
        ang[i][2] = p[i]->angular.z;
    }
    const auto f = hip_detail::field6(field.B, field.gauge_origin);
    double out[2];
    hip_detail::check(ctx, giaohip_primitive_eri(ctx, prim, ang, f.data(), out));
    return {out[0], out[1]};
}

[[nodiscard]] inline Complex primitive_eri(const PrimitiveGaussian& a, const PrimitiveGaussian& b,
                                           const PrimitiveGaussian& c, const PrimitiveGaussian& d,
                                           const MagneticField& field, EriWorkspace&) {
    return primitive_eri(a, b, c, d, field);
}

inline giaohip_basis* EriWorkspace::load(giaohip_ctx* ctx, const std::array<const Shell*, 4>& shells) {
    centers_.clear();
    exps_.clear();
    coefs_.clear();
    L_.clear();
    np_.clear();
    nc_.clear();
    for (const Shell* s : shells) {
        centers_.insert(centers_.end(), {s->center().x, s->center().y, s->center().z});
        L_.push_back(s->angular_momentum());
        np_.push_back(static_cast<int>(s->primitive_count()));
        nc_.push_back(static_cast<int>(s->contraction_count()));
        exps_.insert(exps_.end(), s->exponents().begin(), s->exponents().end());
        coefs_.insert(coefs_.end(), s->coefficients().begin(), s->coefficients().end());
    }
    if (!basis_) {
        hip_detail::check(ctx, giaohip_basis_create(ctx, 4, centers_.data(), L_.data(), np_.data(), nc_.data(),
                                                    exps_.data(), coefs_.data(), &basis_));
    } else {
        hip_detail::check(ctx, giaohip_basis_update(ctx, basis_, 4, centers_.data(), L_.data(), np_.data(),
                                                    nc_.data(), exps_.data(), coefs_.data()));
    }
    return basis_;
}

// compute_eri (eri.cpp:454-585): the contracted block (ab|cd) into output,
// through the workspace's scratch basis.
inline void compute_eri(const Shell& a, const Shell& b, const Shell& c, const Shell& d, const MagneticField& field,
                        std::span<Complex> output, EriWorkspace& workspace) {
    if (output.size() != shell_quartet_size(a, b, c, d)) throw std::length_error("ERI output has an incorrect size");
    giaohip_ctx* ctx = hip_detail::context();
    giaohip_basis* four = workspace.load(ctx, {&a, &b, &c, &d});
    const std::uint32_t q[4] = {0, 1, 2, 3};
    const auto f = hip_detail::field6(field.B, field.gauge_origin);
    hip_detail::check(ctx, giaohip_eri_batch(ctx, four, q, 1, f.data(), nullptr,
                                             reinterpret_cast<double*>(output.data()), 0));
}

[[nodiscard]] inline EriStatistics evaluate_eri_shell_quartets(const Basis& basis,
                                                               std::span<const ShellQuartetIndex> quartets,
                                                               const MagneticField& field,
                                                               const EriEvaluationOptions& options,
                                                               const EriSchwarzBounds* bounds,
                                                               QuartetConsumer consumer, void* user_data) {
    // eri.cpp:644-767: the checks in the reference's order, screening, then
    // the blocks handed to the consumer in input order.
    if (consumer == nullptr) throw std::invalid_argument("ERI quartet consumer must not be null");
    if (!std::isfinite(options.screening_threshold) || options.screening_threshold < 0.0)
        throw std::invalid_argument("ERI screening threshold must be finite and non-negative");
    if (options.thread_count == 0U) throw std::invalid_argument("ERI thread count must be positive");
    if (options.screening_threshold > 0.0 && bounds == nullptr)
        throw std::invalid_argument("positive ERI screening requires Schwarz bounds");
    if (bounds != nullptr && !bounds->matches(basis, field))
        throw std::invalid_argument("ERI Schwarz bounds do not match the basis");
    giaohip_ctx* ctx = hip_detail::context();
    giaohip_basis* dev = basis.device();
    std::vector<std::uint32_t> q(4 * quartets.size()), kept(4 * quartets.size());
    for (std::size_t i = 0; i < quartets.size(); ++i) {
        const auto x = quartets[i].as_array();
        std::copy(x.begin(), x.end(), q.begin() + 4 * i);
    }
    std::size_t nkept = 0;
    hip_detail::check(ctx, giaohip_screen_quartets(ctx, dev, q.data(), quartets.size(),
                                                   bounds ? bounds->values().data() : nullptr,
                                                   options.screening_threshold, kept.data(), &nkept));
    EriStatistics st{quartets.size(), nkept, quartets.size() - nkept};
    if (nkept == 0) return st;
    std::uint64_t total = 0;
    hip_detail::check(ctx, giaohip_eri_batch_size(dev, kept.data(), nkept, &total));
    std::vector<Complex> values(total);
    const auto f = hip_detail::field6(field.B, field.gauge_origin);
    hip_detail::check(ctx, giaohip_eri_batch(ctx, dev, kept.data(), nkept, f.data(), nullptr,
                                             reinterpret_cast<double*>(values.data()), 0));
    const auto shells = basis.shells();
    std::size_t off = 0;
    for (std::size_t i = 0; i < nkept; ++i) {
        const std::uint32_t* x = kept.data() + 4 * i;
        const std::array<std::size_t, 4> shape{shells[x[0]].ao_count(), shells[x[1]].ao_count(),
                                               shells[x[2]].ao_count(), shells[x[3]].ao_count()};
        const std::size_t size = shape[0] * shape[1] * shape[2] * shape[3];
        consumer({{x[0], x[1], x[2], x[3]}, shape, std::span<const Complex>(values.data() + off, size)}, user_data);
        off += size;
    }
    return st;
}

inline void for_each_eri_shell_quartet(const Basis& basis, std::span<const ShellQuartetIndex> quartets,
                                       const MagneticField& field, QuartetConsumer consumer, void* user_data) {
    [[maybe_unused]] const auto st =
        evaluate_eri_shell_quartets(basis, quartets, field, {}, nullptr, consumer, user_data);
}

[[nodiscard]] inline EriStatistics compute_eri_tensor(const Basis& basis, const MagneticField& field,
                                                      const EriEvaluationOptions& options,
                                                      const EriSchwarzBounds* bounds, std::span<Complex> output) {
    // eri.cpp:777-847: canonical quartets, four symmetry images per block
    const std::size_t n = basis.ao_count(), ns = basis.shells().size();
    if (output.size() != n * n * n * n) throw std::length_error("basis ERI output has an incorrect size");
    std::fill(output.begin(), output.end(), Complex{});
    std::vector<ShellQuartetIndex> canon;
    for (std::uint32_t a = 0; a < ns; ++a)
        for (std::uint32_t b = 0; b < ns; ++b)
            for (std::uint32_t c = 0; c < ns; ++c)
                for (std::uint32_t d = 0; d < ns; ++d) {
                    const ShellQuartetIndex q{a, b, c, d};
                    const auto k = canonicalize_shell_quartet(q);
                    if (k.shells.as_array() == q.as_array()) canon.push_back(q);
                }
    struct Sink {
        const Basis* basis;
        std::span<Complex> out;
        std::size_t n;
    } sink{&basis, output, n};
    auto put = [](const ShellQuartetBlockView& v, void* user) {
        auto& s = *static_cast<Sink*>(user);
        const auto off = s.basis->ao_offsets();
        const auto idx = v.shells.as_array();
        const std::size_t N = s.n;
        for (std::size_t i = 0; i < v.shape[0]; ++i)
            for (std::size_t j = 0; j < v.shape[1]; ++j)
                for (std::size_t k = 0; k < v.shape[2]; ++k)
                    for (std::size_t l = 0; l < v.shape[3]; ++l) {
                        const Complex x = v.values[((i * v.shape[1] + j) * v.shape[2] + k) * v.shape[3] + l];
                        const std::size_t p = off[idx[0]] + i, q = off[idx[1]] + j, r = off[idx[2]] + k,
                                          t = off[idx[3]] + l;
                        s.out[((p * N + q) * N + r) * N + t] = x;
                        s.out[((r * N + t) * N + p) * N + q] = x;
                        s.out[((q * N + p) * N + t) * N + r] = std::conj(x);
                        s.out[((t * N + r) * N + q) * N + p] = std::conj(x);
                    }
    };
    return evaluate_eri_shell_quartets(basis, canon, field, options, bounds, put, &sink);
}

inline void compute_eri_tensor(const Basis& basis, const MagneticField& field, std::span<Complex> output) {
    const std::size_t n = basis.ao_count();
    if (output.size() != n * n * n * n) throw std::length_error("basis ERI output has an incorrect size");
    giaohip_ctx* ctx = hip_detail::context();
    const auto f = hip_detail::field6(field.B, field.gauge_origin);
    hip_detail::check(ctx, giaohip_eri_tensor(ctx, basis.device(), f.data(), reinterpret_cast<double*>(output.data()),
                                              0));
}

// --- ERI centre derivatives (derivatives.hpp:226-261) ------------------------
// compute_eri_center_derivatives (derivatives.cpp:952-959): the (4, 3, na,
// nb, nc, nd) block of (ab|cd), London phase response included, through the
// workspace's scratch basis.
inline void compute_eri_center_derivatives(const Shell& a, const Shell& b, const Shell& c, const Shell& d,
                                           const MagneticField& field, std::span<Complex> output,
                                           EriWorkspace& workspace) {
    if (output.size() != 12U * shell_quartet_size(a, b, c, d))
        throw std::length_error("center-derivative shell-quartet output has an incorrect size");
    giaohip_ctx* ctx = hip_detail::context();
    giaohip_basis* four = workspace.load(ctx, {&a, &b, &c, &d});
    const std::uint32_t q[4] = {0, 1, 2, 3};
    const auto f = hip_detail::field6(field.B, field.gauge_origin);
    hip_detail::check(ctx, giaohip_eri_derivative_batch(ctx, four, q, 1, f.data(), nullptr,
                                                        reinterpret_cast<double*>(output.data()), output.size(), 0));
}
inline void compute_eri_center_derivatives(const Shell& a, const Shell& b, const Shell& c, const Shell& d,
                                           const MagneticField& field, std::span<Complex> output) {
    EriWorkspace workspace;
    compute_eri_center_derivatives(a, b, c, d, field, output, workspace);
}

// EriDerivativeSchwarzBounds (derivatives.hpp:244-255, derivatives.cpp:
// 1307-1387): both tables evaluated on the device.
class EriDerivativeSchwarzBounds {
  public:
    explicit EriDerivativeSchwarzBounds(const Basis& basis, const MagneticField& field = {}, std::size_t = 1U)
        : ordinary_(basis, field), values_(basis.shells().size() * basis.shells().size()) {
        giaohip_ctx* ctx = hip_detail::context();
        const auto f = hip_detail::field6(field.B, field.gauge_origin);
        hip_detail::check(ctx, giaohip_eri_derivative_pair_bounds(ctx, basis.device(), f.data(), values_.data()));
    }
    [[nodiscard]] double quartet_bound(ShellQuartetIndex quartet) const {
        const auto n = shell_count();
        for (const auto index : quartet.as_array())
            if (index >= n) throw std::out_of_range("derivative shell quartet index is out of range");
        const double bound = std::max(values_[quartet.a * n + quartet.b] * ordinary_(quartet.c, quartet.d),
                                      ordinary_(quartet.a, quartet.b) * values_[quartet.c * n + quartet.d]);
        return std::nextafter(bound * (1.0 + 128.0 * std::numeric_limits<double>::epsilon()),
                              std::numeric_limits<double>::infinity());
    }
    [[nodiscard]] bool matches(const Basis& basis, const MagneticField& field) const noexcept {
        return ordinary_.matches(basis, field);
    }
    [[nodiscard]] std::span<const double> values() const noexcept { return values_; }
    [[nodiscard]] std::size_t shell_count() const noexcept { return ordinary_.shell_count(); }

  private:
    EriSchwarzBounds ordinary_;
    std::vector<double> values_;
};

// compute_eri_derivative_blocks (derivatives.cpp:1389-1415): concatenated
// (4, 3, na, nb, nc, nd) blocks at explicit cumulative offsets (n + 1 of
// them); the checks and messages are the reference's, made by the library.
inline void compute_eri_derivative_blocks(const Basis& basis, std::span<const ShellQuartetIndex> quartets,
                                          std::span<const std::size_t> offsets, const MagneticField& field,
                                          std::span<Complex> output, std::size_t = 1U) {
    if (offsets.size() != quartets.size() + 1U || offsets[0] != 0U)
        throw std::length_error("derivative batch offsets have an incorrect size or start");
    giaohip_ctx* ctx = hip_detail::context();
    std::vector<std::uint32_t> q(4 * quartets.size());
    for (std::size_t i = 0; i < quartets.size(); ++i) {
        const auto x = quartets[i].as_array();
        std::copy(x.begin(), x.end(), q.begin() + 4 * i);
    }
    const std::vector<std::uint64_t> offs(offsets.begin(), offsets.end());
    const auto f = hip_detail::field6(field.B, field.gauge_origin);
    hip_detail::check(ctx, giaohip_eri_derivative_batch(ctx, basis.device(), q.data(), quartets.size(), f.data(),
                                                        offs.data(), reinterpret_cast<double*>(output.data()),
                                                        output.size(), 0));
}

// Point charge (types.hpp:134-139), validated as types.cpp:188-194 does.
struct Nucleus {
    double charge{};
    Vec3 center{};
    Nucleus(double charge_in, Vec3 center_in) : charge(charge_in), center(center_in) {
        if (!std::isfinite(charge) || charge <= 0.0)
            throw std::invalid_argument("nuclear charge must be finite and positive");
        if (!std::isfinite(center.x) || !std::isfinite(center.y) || !std::isfinite(center.z))
            throw std::invalid_argument("nuclear center values must be finite");
    }
};

// compute_nuclear_attraction_matrix (nuclear.cpp:282-312, nuclear.hpp:59-61):
// the (n, n) matrix of sum_C -Z_C <mu| 1/|r - C| |nu>, row-major.
inline void compute_nuclear_attraction_matrix(const Basis& basis, std::span<const Nucleus> nuclei,
                                              const MagneticField& field, std::span<Complex> output) {
    const std::size_t n = basis.ao_count();
    if (output.size() != n * n) throw std::length_error("basis nuclear-attraction output has an incorrect size");
    std::vector<double> charges, centers;
    for (const Nucleus& nu : nuclei) {
        charges.push_back(nu.charge);
        centers.insert(centers.end(), {nu.center.x, nu.center.y, nu.center.z});
    }
    giaohip_ctx* ctx = hip_detail::context();
    const auto f = hip_detail::field6(field.B, field.gauge_origin);
    hip_detail::check(ctx, giaohip_nuclear_attraction(ctx, basis.device(), static_cast<int>(charges.size()),
                                                      charges.data(), centers.data(), f.data(),
                                                      reinterpret_cast<double*>(output.data())));
}

// Parallel evaluation is available (on the GPU), so thread counts above one
// are accepted; the host threads the library plans with (eri.cpp:855-869).
[[nodiscard]] inline bool openmp_enabled() noexcept { return true; }
[[nodiscard]] inline std::size_t openmp_max_threads() noexcept {
    return static_cast<std::size_t>(giaohip_host_threads());
}

// --- selected Cholesky integrals (src/cholesky.hpp:9-17) --------------------
namespace detail {

inline void cholesky_diagonal(const Basis& basis, std::span<const std::array<std::size_t, 2>> pairs,
                              const MagneticField& field, std::span<double> output, std::size_t threads) {
    if (threads == 0U) throw std::invalid_argument("ERI thread count must be positive");
    giaohip_ctx* ctx = hip_detail::context();
    std::vector<std::uint64_t> p;
    for (const auto& x : pairs) p.insert(p.end(), {x[0], x[1]});
    const std::size_t n = basis.ao_count();
    if (output.size() != n * n) throw std::length_error("Cholesky diagonal output has incorrect size");
    const auto f = hip_detail::field6(field.B, field.gauge_origin);
    hip_detail::check(ctx, giaohip_cholesky_diagonal(ctx, basis.device(), p.data(), pairs.size(), f.data(),
                                                     output.data()));
}

inline void cholesky_columns_impl(const Basis& basis, std::span<const std::array<std::size_t, 2>> pivots,
                                  const MagneticField& field, double* out, std::size_t count, bool real,
                                  std::size_t threads) {
    if (threads == 0U) throw std::invalid_argument("ERI thread count must be positive");
    const std::size_t n = basis.ao_count();
    if (count != pivots.size() * n * n) throw std::length_error("selected Cholesky column output has incorrect size");
    giaohip_ctx* ctx = hip_detail::context();
    std::vector<std::uint64_t> p;
    for (const auto& x : pivots) p.insert(p.end(), {x[0], x[1]});
    const auto f = hip_detail::field6(field.B, field.gauge_origin);
    hip_detail::check(ctx, giaohip_cholesky_columns(ctx, basis.device(), p.data(), pivots.size(), f.data(), out,
                                                    real ? 1 : 0));
}

inline void cholesky_columns(const Basis& basis, std::span<const std::array<std::size_t, 2>> pivots,
                             const MagneticField& field, std::span<double> output, std::size_t threads) {
    cholesky_columns_impl(basis, pivots, field, output.data(), output.size(), true, threads);
}

inline void cholesky_columns(const Basis& basis, std::span<const std::array<std::size_t, 2>> pivots,
                             const MagneticField& field, std::span<Complex> output, std::size_t threads) {
    cholesky_columns_impl(basis, pivots, field, reinterpret_cast<double*>(output.data()), output.size(), false,
                          threads);
}

}  // namespace detail

}  // namespace giao
