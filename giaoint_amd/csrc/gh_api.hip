// C ABI of the gfx950 London-orbital ERI library (declared in include/giaohip.h).
//
// Host responsibilities: validation with the reference's error semantics,
// per-shell tables (normalizations, London wave vectors), planning of device
// work (seed chunks bounded in memory, work items grouped by angular-momentum
// class, general contractions split into contraction rows), and copies.
// Every integral is computed by the kernels in gh_eri.hip.
#include "giaohip.h"
#include "gh_internal.h"

#include <rocblas/rocblas.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <limits>
#include <list>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <set>
#include <sstream>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <span>
#include <vector>

using gh::DerivItem;
using gh::DevError;
using gh::DevShell;
using gh::SeedQuartet;
using gh::WorkItem;
using gh::zc;

// Device memory of giaohip_eri_batch_noalloc: one array on every device the
// library runs on, held from the load of the code object until the process
// ends. On the host the name is only a stand-in for the symbol: host code
// uses the device address that hipGetSymbolAddress gives per device
// (noalloc_scratch_init), never the name itself.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

// Device allocations the library has made in this process (giaohip_device_allocations).
std::atomic<unsigned long long> g_device_allocations{0};

hipError_t counted_malloc(void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) g_device_allocations.fetch_add(1, std::memory_order_relaxed);
    return e;
}

// Per-device address of `scratch`, looked up once per device.
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_addr;

// Looks up `scratch` on `device` (made current here) the first time; later
// calls return the recorded address.
hipError_t noalloc_scratch_init(int device, void** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_addr.find(device);
    if (it != g_scratch_addr.end()) {
        *out = it->second;
        return hipSuccess;
    }
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return e;
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (ptr == nullptr) return hipErrorInvalidSymbol;
    g_scratch_addr[device] = ptr;
    *out = ptr;
    return hipSuccess;
}

// Layout of the allocation-free scratch (bytes): wave vectors, the device
// error record, the work-queue counters, seed-quartet descriptors, work
// items, then the seeds and the Boys quadrature lists.
constexpr size_t kNaKappa = 0, kNaKappaBytes = size_t(64) << 10;
constexpr size_t kNaErr = kNaKappa + kNaKappaBytes;
constexpr size_t kNaQueue = kNaErr + 128;
constexpr size_t kNaSq = kNaErr + 4096, kNaSqBytes = size_t(1) << 20;
constexpr size_t kNaItems = kNaSq + kNaSqBytes, kNaItemsBytes = size_t(4) << 20;
constexpr size_t kNaSeeds = kNaItems + kNaItemsBytes;
static_assert(kNaSeeds % 256 == 0 && kNaSeeds < kNoAllocScratchBytes, "scratch layout");

struct DevBuffer {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && p) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            p = nullptr;
            cap = 0;
            if (e != hipSuccess) return e;
        }
        size_t want = std::max<size_t>(bytes, 256);
        hipError_t e = counted_malloc(&p, want);
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// A DevBuffer released when it goes out of scope (temporaries of one call).
struct ScopedBuffer : DevBuffer {
    ScopedBuffer() = default;
    ScopedBuffer(const ScopedBuffer&) = delete;
    ScopedBuffer& operator=(const ScopedBuffer&) = delete;
    ~ScopedBuffer() { release(); }
};

// Page-locked host staging: copies from it are asynchronous with respect to
// the host, so the host can plan the next chunk while the device works.
struct PinnedBuffer {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && p) return hipSuccess;
        if (p) {
            hipError_t e = hipHostFree(p);
            p = nullptr;
            cap = 0;
            if (e != hipSuccess) return e;
        }
        size_t want = std::max<size_t>(bytes, 4096);
        hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

struct Status {
    int code = GIAOHIP_OK;
    std::string msg;
};

constexpr double kPi = 3.141592653589793;
constexpr size_t kMaxAuxEntries = 4000000u;  // eri.cpp:24

}  // namespace

struct giaohip_basis {
    int ns = 0;
    long long nao = 0;
    // Per shell-class (la, lb, lc, ld) tables, index ((la * lm + lb) * lm + lc) * lm + ld:
    // admission status (see class_status) and the batched kernel's orientation.
    int lm = 1;
    std::vector<signed char> cls_status, cls_orient;
    std::vector<gh::FastClassInfo> cls_info;
    std::vector<DevShell> shells;
    std::vector<long long> ao_off;
    std::vector<double> exps, coefs, norms, radial;
    std::vector<int> owner;
    DevBuffer d_shells, d_exps, d_coefs, d_norms, d_aooff, d_owner, d_radial;
    // shifted-momentum companion basis of the centre derivatives, built on
    // first use and dropped when the shells change
    giaohip_basis* deriv = nullptr;
};

void release_basis(giaohip_basis* b) {
    if (!b) return;
    release_basis(b->deriv);
    b->deriv = nullptr;
    for (DevBuffer* d : {&b->d_shells, &b->d_exps, &b->d_coefs, &b->d_norms, &b->d_aooff, &b->d_owner, &b->d_radial})
        d->release();
    delete b;
}

struct giaohip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t aux = nullptr;  // producer stream of the batched path (default priority, like `stream`)
    std::mutex mu;
    std::string err;
    giaohip_stats stats{};
    DevBuffer seeds, squartets, items, kappa, errrec, staging, blocks, misc_a, misc_b, misc_c, fpq, fitems;
    int eri_variant = GIAOHIP_ERI_BATCHED;
    hipEvent_t ev[5]{};
    std::vector<hipEvent_t> chunk_ev;  // per-chunk timing events of the batched path
    PinnedBuffer h_sq, h_pq, h_items;  // host staging of the batched path's descriptors
    // host outputs: a second device staging buffer (one group is computed
    // while the previous one is copied out), page-locked bounce buffers and
    // the copy stream with its events
    DevBuffer staging2;
    PinnedBuffer bounce[2];
    hipStream_t copy_stream = nullptr;
    hipEvent_t copy_ev[2]{};
    DevBuffer work[2];                 // seeds, R0 tables and pair tables of two chunks in flight
    std::set<void*> user_allocs;
    size_t seed_budget_bytes = size_t(16384) << 20;
    int chol_block = 32;  // candidates per Cholesky pivot block
    int chol_pivot_batch = 32;  // cholesky_eri's pivot_batch_size
    size_t staging_budget_bytes = size_t(1) << 30;
    int eri_queue = GIAOHIP_QUEUE_NONE;  // generic-kernel launch form (GIAOHIP_OPTION_ERI_QUEUE)
    int fast_waves_cap = 4;                  // class-specialized workgroup cap (0: kFastMaxWaves)
    int mfma_mode = GIAOHIP_MFMA_DEFAULT;    // FP64-MFMA contraction use (GIAOHIP_OPTION_MFMA)
    int pipeline = GIAOHIP_PIPELINE_OVERLAP; // producer stream of the batched path (GIAOHIP_OPTION_PIPELINE)
    int wave_max_l = 7;                      // one-wavefront generic items up to this L (GIAOHIP_OPTION_WAVE_MAX_L)
    size_t fast_pending = 0;                 // batched-path chunks queued, statistics not yet collected
    int mfma_min_l = 9;                      // default-mode matrix-core classes from this L (GIAOHIP_OPTION_MFMA_MIN_L)
    int chol_graphs = 0;                     // Cholesky column launches as hipGraphs (GIAOHIP_OPTION_CHOL_GRAPHS)
    int mfma_variant = 0;                    // compile-time variant of the matrix-core kernels (GIAOHIP_OPTION_MFMA_VARIANT)
    int small_max_l = gh::kSmallMaxL;        // lane-per-primitive-quartet classes up to this L (GIAOHIP_OPTION_SMALL_MAX_L)
    DevBuffer queue;                         // work-queue counters: [0] context stream, [1] aux stream
    // pipelined chunks of the plain path (run_plain): per slot the seeds,
    // descriptors and work items of one chunk in flight, their page-locked
    // host staging, and events; the
    // seed kernels run on plain_stream (default priority)
    hipStream_t plain_stream = nullptr;
    DevBuffer pl_seeds[2], pl_sq[2], pl_items[2];
    PinnedBuffer pl_hsq[2], pl_hitems[2];
    hipEvent_t pl_ev[2][4]{};  // seeds start, seeds end, transforms done on the context and on the aux stream
    hipEvent_t pl_init = nullptr;
    void* noalloc = nullptr;                 // device address of `scratch` on this context's device
    rocblas_handle blas = nullptr;           // library GEMMs (exchange contraction), created on first use
};

// A device-resident Cholesky: factors, residual and every buffer a pivot step
// needs, so that a run can be continued (giaohip_cholesky_resume).
struct giaohip_cholesky {
    long long n = 0;
    int rows_per_chunk = 0;
    std::vector<zc*> chunks;  // device rows [chunk][row][n*n]
    std::vector<int64_t> pivots;
    uint64_t rank = 0;
    double max_residual = 0.0, tolerance = 0.0;
    int converged = 0;
    double column_ms = 0.0, update_ms = 0.0, diagonal_ms = 0.0;
    // state carried between steps
    const giaohip_basis* basis = nullptr;
    int failed = 0;         // status that ended the run early (the factors stop there)
    double psd_tol = 0.0;
    uint64_t limit = 0;     // rank limit (max_rank, or n*n)
    gh::CholState state{};  // current maximum residual and its pivot
    std::vector<std::array<int, 4>> groups;  // bra (la, lb, first unit, count)
    int nunits = 0, max_lab = 0;
    long long p1 = 0;
    DevBuffer d_units, d_sq, d_wi, d_res, d_col, d_coef, d_pmax, d_pidx, d_state, d_flags, d_chunkptr, d_seeds,
        d_kappa, d_err, d_proj, d_cand;
    int* d_scratch = nullptr;
    // pivot block: candidates (host copy), slots, slots not yet taken
    int block = 1, slots = 1;
    int ncand = 0;
    long long cand[gh::kCholMaxBlock]{};
    uint64_t complete[gh::kCholMaxBlock]{};  // rank up to which each slot's projection is summed
    // Boys seeds of recent ket shell pairs (a pivot's seeds depend only on
    // its two shells): seed_slots buffers of seed_slot_zc complex each,
    // least recently used replaced first
    long long seed_slot_zc = 0;
    int seed_slots = 1;
    std::vector<long long> seed_key;     // ket shell pair per buffer (-1: empty)
    std::vector<uint64_t> seed_used;     // last use (pivot count) per buffer
    uint64_t seed_hits = 0, seed_misses = 0;
    // Pivot policy of complex runs on bases with general contractions
    // (cholesky.py:193-262): the pivot may come from the latest candidate
    // block; the candidate sets and the membership of the raw-column cache
    // decide when. Only the membership matters here (every column is
    // computed when its pivot is taken), so the cache holds indices.
    bool policy = false;
    int pivot_batch = 32;
    std::vector<long long> pol_block, pol_cand;
    std::list<long long> pol_lru;
    std::unordered_map<long long, std::list<long long>::iterator> pol_pos;
    DevBuffer d_gidx, d_gval;
    unsigned open = 0;
    uint64_t blocks = 0;
    uint64_t rows_read = 0;  // factor rows read by the block projections
    uint64_t projected = 0;  // candidates projected (the others carried over)
    uint64_t cand_rows = 0;  // sum of candidates projected x rank
    int real = 0;            // real arithmetic: zero field or one centre (symmetrized columns)
    double field[6]{};       // the field the integrals are taken at (zero on one centre)
    double block_ms = 0.0;   // candidate selection + projection
    hipEvent_t ev[10]{};
    hipStream_t s3 = nullptr;  // second stream of the column launches
    void* host = nullptr;  // page-locked copy of the state and the PSD flag
    // A pivot's column launches as hipGraphs, one executable graph per launch
    // topology (ket momenta, whether the seeds are computed, launch stream,
    // kernel and stream of each bra group): the first pivot of a topology
    // launches directly, later ones capture their launches and update the
    // graph.
    struct ColumnGraph {
        hipGraphExec_t exec = nullptr;
        uint64_t uses = 0;
    };
    std::map<std::vector<int>, ColumnGraph> graphs;
    uint64_t graph_launches = 0, graph_instantiations = 0, graph_updates = 0;
    void release() {
        for (auto& [key, g] : graphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        graphs.clear();
        for (zc* c : chunks) (void)hipFree(c);
        chunks.clear();
        for (DevBuffer* x : {&d_units, &d_sq, &d_wi, &d_res, &d_col, &d_coef, &d_pmax, &d_pidx, &d_state, &d_flags,
                             &d_chunkptr, &d_seeds, &d_kappa, &d_err, &d_proj, &d_cand, &d_gidx, &d_gval})
            x->release();
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (s3) (void)hipStreamDestroy(s3);
        s3 = nullptr;
        if (host) (void)hipHostFree(host);
        host = nullptr;
    }
};

namespace {

int set_error(giaohip_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int hip_fail(giaohip_ctx* ctx, hipError_t e, const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return set_error(ctx, GIAOHIP_DEVICE_ERROR, m);
}

#define HIP_TRY(expr, what)                                  \
    do {                                                     \
        hipError_t _e = (expr);                              \
        if (_e != hipSuccess) return hip_fail(ctx, _e, what); \
    } while (0)

// types.cpp:168-175,207-219
double log_odd_double_factorial(int l) {
    if (l == 0) return 0.0;
    const double n = (double)l;
    return n * std::log(2.0) + std::lgamma(n + 0.5) - 0.5 * std::log(kPi);
}

double primitive_normalization(double exponent, int ax, int ay, int az) {
    const double total = (double)(ax + ay + az);
    const double log_den = log_odd_double_factorial(ax) + log_odd_double_factorial(ay) + log_odd_double_factorial(az);
    const double log_norm = 0.75 * std::log(2.0 * exponent / kPi) + 0.5 * (total * std::log(4.0 * exponent) - log_den);
    return std::exp(log_norm);
}

std::vector<std::array<int, 3>> cartesian_components(int l) {
    std::vector<std::array<int, 3>> out;
    for (int x = l; x >= 0; --x)
        for (int y = l - x; y >= 0; --y) out.push_back({x, y, l - x - y});
    return out;
}

// MagneticField::london_wave_vector (types.cpp:326-333)
std::vector<double> wave_vectors(const giaohip_basis* b, const double f[6]) {
    std::vector<double> k(3 * (size_t)b->ns);
    for (int s = 0; s < b->ns; ++s) {
        const double* c = b->shells[s].center;
        const double dx = c[0] - f[3], dy = c[1] - f[4], dz = c[2] - f[5];
        k[3 * s + 0] = 0.5 * (f[1] * dz - f[2] * dy);
        k[3 * s + 1] = 0.5 * (f[2] * dx - f[0] * dz);
        k[3 * s + 2] = 0.5 * (f[0] * dy - f[1] * dx);
    }
    return k;
}

int nmax_bucket(int L) {
    for (int b : {4, 8, 12, 16, 24}) if (L <= b) return b;
    return 32;
}

// Maximum over component quartets of (L+1) * prod over the six axes of
// (power sum + 1), the reference's auxiliary admission size (eri.cpp:85-98).
size_t max_aux_entries(int la, int lb, int lc, int ld) {
    auto pair_max = [](int l1, int l2) {
        size_t best = 0;
        for (auto& a : cartesian_components(l1))
            for (auto& b : cartesian_components(l2)) {
                size_t v = 1;
                for (int i = 0; i < 3; ++i) v *= (size_t)(a[i] + b[i] + 1);
                best = std::max(best, v);
            }
        return best;
    };
    return (size_t)(la + lb + lc + ld + 1) * pair_max(la, lb) * pair_max(lc, ld);
}

// A request: one shell quartet, optionally restricted to fixed ket
// contraction rows and a ket component range, written with AO strides.
struct Request {
    int sh[4];
    int xfix[4];  // -1 = every contraction row
    int cd0, cd1; // ket component range; cd1 < 0 = all
    long long out_base;
    long long stride[4];
};

// Per-class admission, memoized (the reference checks every component quartet).
// Admission of a shell class, in the reference's check order.
enum ClassStatus : signed char { kClsOk = 0, kClsOrder = 1, kClsAux = 2, kClsShellL = 3, kClsTotalL = 4 };

ClassStatus class_admission(int la, int lb, int lc, int ld) {
    const int L = la + lb + lc + ld;
    if (L > 32) return kClsOrder;
    if (max_aux_entries(la, lb, lc, ld) > kMaxAuxEntries) return kClsAux;
    if (la > gh::kMaxShellL || lb > gh::kMaxShellL || lc > gh::kMaxShellL || ld > gh::kMaxShellL) return kClsShellL;
    if (L > gh::kMaxTotalL) return kClsTotalL;
    return kClsOk;
}

int class_error(giaohip_ctx* ctx, int status) {
    switch (status) {
        case kClsOrder:
            return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "electron repulsion supports combined angular order through 32");
        case kClsAux: return set_error(ctx, GIAOHIP_LENGTH_ERROR, "ERI auxiliary exceeds the workspace limit");
        case kClsShellL:
            return set_error(ctx, GIAOHIP_UNSUPPORTED, "GPU ERI kernels support shell angular momentum through 8");
        case kClsTotalL:
            return set_error(ctx, GIAOHIP_UNSUPPORTED, "GPU ERI kernels support combined angular order through 24");
        default: return GIAOHIP_OK;
    }
}

inline size_t class_index(const giaohip_basis* b, const int sh[4]) {
    const int lm = b->lm;
    return (((size_t)b->shells[sh[0]].l * lm + b->shells[sh[1]].l) * lm + b->shells[sh[2]].l) * lm +
           b->shells[sh[3]].l;
}

int check_quartet_supported(giaohip_ctx* ctx, const giaohip_basis* b, const int sh[4]) {
    return class_error(ctx, b->cls_status[class_index(b, sh)]);
}

// Host threads of the library's planning work: the machine's hardware
// threads, at most 16.
size_t host_threads() { return std::min<size_t>(16, std::max(1u, std::thread::hardware_concurrency())); }

// Runs f(begin, end) over [0, n) split across host_threads() threads.
template <class F>
void parallel_for(size_t n, F&& f) {
    constexpr size_t kGrain = 8192;
    size_t T = host_threads();
    T = std::min(T, (n + kGrain - 1) / kGrain);
    if (T <= 1) {
        f((size_t)0, n);
        return;
    }
    const size_t per = (n + T - 1) / T;
    std::vector<std::thread> threads;
    threads.reserve(T - 1);
    for (size_t t = 1; t < T; ++t)
        threads.emplace_back([&, t] {
            const size_t lo = t * per, hi = std::min(n, lo + per);
            if (lo < hi) f(lo, hi);
        });
    f((size_t)0, std::min(n, per));
    for (auto& th : threads) th.join();
}

// Allocator whose value-less construction leaves trivial elements
// uninitialized: large request and plan arrays are filled by parallel loops
// right after allocation, so zeroing them first (single-threaded) is waste.
template <class T>
struct NoInit : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInit<U>;
    };
    NoInit() = default;
    template <class U>
    NoInit(const NoInit<U>&) noexcept {}
    template <class U>
    void construct(U* p) noexcept {
        ::new (static_cast<void*>(p)) U;
    }
    template <class U, class... Args>
    void construct(U* p, Args&&... args) {
        ::new (static_cast<void*>(p)) U(std::forward<Args>(args)...);
    }
};
template <class T>
using BigVec = std::vector<T, NoInit<T>>;

int boys_error(giaohip_ctx* ctx, const DevError& e) {
    switch (e.code) {
        case gh::kErrBoysDomain: {
            std::ostringstream m;
            m << "complex Boys argument " << std::complex<double>(e.zr, e.zi) << " at order " << e.order
              << " is outside the verified direct domain |z| <= " << 160.0
              << " and the conservative large-positive asymptotic sector";
            return set_error(ctx, GIAOHIP_BOYS_ERROR, m.str());
        }
        case gh::kErrBoysSeries: return set_error(ctx, GIAOHIP_BOYS_ERROR, "complex Boys power series did not converge");
        case gh::kErrBoysQuadrature:
            return set_error(ctx, GIAOHIP_BOYS_ERROR, "adaptive complex Boys quadrature did not reach tolerance");
        case gh::kErrBoysScaling:
            return set_error(ctx, GIAOHIP_BOYS_ERROR, "requested Boys scaling is not representable for this argument");
        case gh::kErrBoysUnscaled:
            return set_error(ctx, GIAOHIP_BOYS_ERROR, "unscaled Boys value is not representable for this argument");
        default: return set_error(ctx, GIAOHIP_BOYS_ERROR, "complex Boys evaluation failed");
    }
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0f;
    return ms;
}

// Ket-component tile, LDS bytes and number of work items of one request on
// the plain transform kernel.
// Work items of a request on the generic transform: the class, the ket
// components per item (tile), LDS bytes, item count and workgroup size nt.
// Classes with la + lb + lc + ld <= ctx->wave_max_l run on one wavefront
// (nt = 64) when a tile of at least one component fits its registers and the
// whole ket range takes no more passes than on four wavefronts.
int plain_plan(giaohip_ctx* ctx, const giaohip_basis* b, const Request& r, gh::ClassArgs& ca, int& lds,
               long long& nitems, int& nt) {
    const DevShell* S[4] = {&b->shells[r.sh[0]], &b->shells[r.sh[1]], &b->shells[r.sh[2]], &b->shells[r.sh[3]]};
    const int ncd_all = S[2]->ncart * S[3]->ncart;
    const int cd0 = r.cd1 < 0 ? 0 : r.cd0, cd1 = r.cd1 < 0 ? ncd_all : r.cd1;
    const int nab = S[0]->ncart * S[1]->ncart;
    const int NHab = gh::nherm(S[0]->l + S[1]->l);
    const int L = S[0]->l + S[1]->l + S[2]->l + S[3]->l;
    auto tile_for = [&](int threads) {
        return std::min({cd1 - cd0, gh::transform_max_w(threads) / NHab, gh::transform_max_o(threads) / nab});
    };
    nt = 256;
    // (the single persistent launch form keeps every item on four wavefronts)
    if (L <= ctx->wave_max_l && ctx->eri_queue != GIAOHIP_QUEUE_SINGLE) {
        const int t64 = tile_for(64), t256 = tile_for(256);
        if (t64 >= 1 && t256 >= 1 && (cd1 - cd0 + t64 - 1) / t64 == (cd1 - cd0 + t256 - 1) / t256) nt = 64;
    }
    int tile = tile_for(nt);
    if (tile < 1) return set_error(ctx, GIAOHIP_UNSUPPORTED, "ERI class too large for the GPU transform kernel");
    ca = gh::ClassArgs{{S[0]->l, S[1]->l, S[2]->l, S[3]->l}, tile};
    while (tile > 1 && gh::transform_lds_bytes(ca) > 160 * 1024) ca.tile = --tile;
    lds = gh::transform_lds_bytes(ca);
    if (lds > 160 * 1024)
        return set_error(ctx, GIAOHIP_UNSUPPORTED, "ERI class exceeds the LDS capacity of the transform kernel");
    long long rows = 1;
    for (int k = 0; k < 4; ++k) rows *= r.xfix[k] < 0 ? S[k]->ncontr : 1;
    nitems = rows * ((cd1 - cd0 + ca.tile - 1) / ca.tile);
    return GIAOHIP_OK;
}

// FP64-MFMA instance of a request, with the request turned into its (cd|ab)
// image when that image's bra side is the larger one (the kernels are
// instantiated for lab >= lcd): returns lab * 16 + lcd of the instance, or -1.
// A request restricted to part of its ket components is not turned.
// Under GIAOHIP_MFMA_DEFAULT only classes of total momentum >= min_l are used:
// below that the FP64 VALU kernels are faster (the dense products carry the
// structural zeros of the Hermite tables).
// Instance keys of the split-FP32 form carry this flag.
constexpr int kMfmaSplitKey = 256;

int mfma_orient(const giaohip_basis* b, Request& r, int mode, int min_l) {
    const DevShell *S[4] = {&b->shells[r.sh[0]], &b->shells[r.sh[1]], &b->shells[r.sh[2]], &b->shells[r.sh[3]]};
    const int lab = S[0]->l + S[1]->l, lcd = S[2]->l + S[3]->l;
    if (mode == GIAOHIP_MFMA_OFF) return -1;
    if (mode == GIAOHIP_MFMA_DEFAULT && lab + lcd < min_l) return -1;
    const bool split = mode == GIAOHIP_MFMA_SPLIT32 &&
                       (gh::mfma_split_instance(lab, lcd) || (r.cd1 < 0 && gh::mfma_split_instance(lcd, lab)));
    const bool keep = lab >= lcd && (split ? gh::mfma_split_instance(lab, lcd) : gh::mfma_instance(lab, lcd));
    const bool turn = r.cd1 < 0 && lcd >= lab && (split ? gh::mfma_split_instance(lcd, lab) : gh::mfma_instance(lcd, lab));
    if (!keep && !turn) return -1;
    const int flag = split ? kMfmaSplitKey : 0;
    bool do_turn = turn && !keep;
    if (keep && turn) {
        // equal momenta: the ket with fewer components and primitive pairs costs less
        const double kab = (double)S[0]->nprim * S[1]->nprim, kcd = (double)S[2]->nprim * S[3]->nprim;
        const double nab = (double)S[0]->ncart * S[1]->ncart, ncd = (double)S[2]->ncart * S[3]->ncart;
        const double nh = (double)gh::nherm(lab);
        const double as_is = kab * kcd * nh * nh * ncd + kab * nab * nh * ncd;
        const double turned = kab * kcd * nh * nh * nab + kcd * ncd * nh * nab;
        do_turn = turned < as_is;
    }
    if (do_turn) {
        Request t = r;
        const int map[4] = {2, 3, 0, 1};
        for (int k = 0; k < 4; ++k) {
            t.sh[k] = r.sh[map[k]];
            t.xfix[k] = r.xfix[map[k]];
            t.stride[k] = r.stride[map[k]];
        }
        r = t;
        return (lcd * 16 + lab) | flag;
    }
    return (lab * 16 + lcd) | flag;
}

// Fixed device memory of the plain path: slices of the library's static
// scratch array, used by giaohip_eri_batch_noalloc instead of the context's
// buffers (which grow on demand). Capacities are in elements.
struct FixedPlainMem {
    double* kappa;
    size_t kappa_cap;
    DevError* err;
    int* queue;
    SeedQuartet* sq;
    size_t sq_cap;
    WorkItem* items;
    size_t items_cap;
    zc* seeds;  // seeds, then the Boys quadrature lists
    size_t seeds_cap;
};

size_t seed_scratch_zc(long long pq_total) {
    return (gh::seed_scratch_ints(pq_total) * sizeof(int) + sizeof(zc) - 1) / sizeof(zc);
}

// Runs requests with the plain per-class transform kernel (with `mfma`, the
// classes the context's MFMA mode selects on the matrix cores; with `small`,
// the smallest classes on the lane-per-primitive-quartet kernel). With `fixed`, all
// device memory comes from its slices (no allocation); chunks are then also
// bounded by the slices' capacities, and run one after the other. Otherwise
// chunks are pipelined over two slots: while the transforms of chunk c run
// on the context and aux streams, the host plans chunk c + 1 and its seed
// kernel runs on the plain-path stream; a slot is reused once the chunk two
// back has finished.
// Chunk timing of a finished slot of the pipelined generic path into the statistics.
hipError_t plain_slot_stats(giaohip_ctx* ctx, int k) {
    hipError_t e = hipEventSynchronize(ctx->pl_ev[k][2]);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->pl_ev[k][3]);
    if (e != hipSuccess) return e;
    const float t = std::max(elapsed(ctx->pl_ev[k][1], ctx->pl_ev[k][2]), elapsed(ctx->pl_ev[k][1], ctx->pl_ev[k][3]));
    ctx->stats.seed_ms += elapsed(ctx->pl_ev[k][0], ctx->pl_ev[k][1]);
    ctx->stats.transform_ms += t;
    ctx->stats.contract_ms += t;
    return hipSuccess;
}

// With `prepared`, the caller has already queued the wave vectors and the
// zeroed error record on the context stream (work queued before this call
// may have recorded an error there, which this call's check then reports).
int run_plain(giaohip_ctx* ctx, const giaohip_basis* b, const double field[6], std::span<const Request> reqs_in,
              zc* d_out, const FixedPlainMem* fixed = nullptr, bool mfma = false, bool small = false,
              bool prepared = false) {
    if (reqs_in.empty()) return GIAOHIP_OK;
    // requests with an FP64-MFMA instance, oriented for it (mfk: instance key or -1)
    BigVec<Request> oriented;
    std::vector<int> mfk;
    if (mfma && !fixed) {
        mfk.assign(reqs_in.size(), -1);
        std::atomic<bool> any{false};
        parallel_for(reqs_in.size(), [&](size_t lo, size_t hi) {
            for (size_t i = lo; i < hi; ++i) {
                Request r = reqs_in[i];
                mfk[i] = mfma_orient(b, r, ctx->mfma_mode, ctx->mfma_min_l);
                if (mfk[i] >= 0) any.store(true, std::memory_order_relaxed);
            }
        });
        if (any) {
            oriented.resize(reqs_in.size());
            parallel_for(reqs_in.size(), [&](size_t lo, size_t hi) {
                for (size_t i = lo; i < hi; ++i) {
                    oriented[i] = reqs_in[i];
                    if (mfk[i] >= 0) (void)mfma_orient(b, oriented[i], ctx->mfma_mode, ctx->mfma_min_l);
                }
            });
        } else {
            mfk.clear();
        }
    }
    const std::span<const Request> reqs = mfk.empty() ? reqs_in : std::span<const Request>(oriented);
    hipStream_t s = ctx->stream;
    const std::vector<double> kap = wave_vectors(b, field);
    double* d_kappa;
    DevError* d_err;
    if (fixed) {
        if (kap.size() > fixed->kappa_cap)
            return set_error(ctx, GIAOHIP_INVALID_ARGUMENT,
                             "basis has more shells than the allocation-free ERI scratch holds");
        d_kappa = fixed->kappa;
        d_err = fixed->err;
        HIP_TRY(gh::launch_zero_ints(reinterpret_cast<int*>(d_err), (int)(sizeof(DevError) / sizeof(int)), s),
                "fill kernel launch");
    } else {
        HIP_TRY(ctx->kappa.ensure(kap.size() * sizeof(double)), "hipMalloc");
        HIP_TRY(ctx->errrec.ensure(sizeof(DevError)), "hipMalloc");
        d_kappa = ctx->kappa.as<double>();
        d_err = ctx->errrec.as<DevError>();
        if (!prepared) HIP_TRY(hipMemsetAsync(d_err, 0, sizeof(DevError), s), "hipMemsetAsync");
    }
    if (!prepared || fixed)
        HIP_TRY(hipMemcpyAsync(d_kappa, kap.data(), kap.size() * sizeof(double), hipMemcpyHostToDevice, s),
                "hipMemcpyAsync");
    const bool piped = !fixed;
    hipStream_t ps = piped ? ctx->plain_stream : s;  // the seed kernels' stream
    if (piped) {
        // the producer stream sees the zeroed error record and the wave vectors
        HIP_TRY(hipEventRecord(ctx->pl_init, s), "hipEventRecord");
        HIP_TRY(hipStreamWaitEvent(ps, ctx->pl_init, 0), "hipStreamWaitEvent");
    }
    auto slot_stats = [&](int k) -> hipError_t { return plain_slot_stats(ctx, k); };
    // a chunk's seeds stay within two chunk budgets (per slot when pipelined:
    // the longest items' tails are paid once per chunk, so chunks are large)
    const size_t seed_cap = 2 * ctx->seed_budget_bytes;

    size_t begin = 0;
    int chunk = 0;
    while (begin < reqs.size()) {
        // Chunk so that the seed buffer stays within budget.
        size_t end = begin;
        size_t seed_zc = 0;
        long long pq_total = 0, item_total = 0;
        int maxL = 0;
        std::vector<SeedQuartet> sq;
        while (end < reqs.size()) {
            const Request& r = reqs[end];
            const DevShell &A = b->shells[r.sh[0]], &B = b->shells[r.sh[1]], &C = b->shells[r.sh[2]],
                           &D = b->shells[r.sh[3]];
            const int L = A.l + B.l + C.l + D.l;
            const long long kab = (long long)A.nprim * B.nprim, kcd = (long long)C.nprim * D.nprim;
            const size_t need = (size_t)(kab * kcd) * (size_t)(L + 4);
            // seeds are this path's only large buffer
            if (end > begin && (seed_zc + need) * sizeof(zc) > seed_cap) break;
            if (fixed) {
                gh::ClassArgs ca;
                int lds, nt;
                long long nit;
                const int st = plain_plan(ctx, b, r, ca, lds, nit, nt);
                if (st) return st;
                const bool fits = sq.size() + 1 <= fixed->sq_cap && (size_t)(item_total + nit) <= fixed->items_cap &&
                                  seed_zc + need + seed_scratch_zc(pq_total + kab * kcd) <= fixed->seeds_cap;
                if (!fits) {
                    if (end == begin)
                        return set_error(ctx, GIAOHIP_INVALID_ARGUMENT,
                                         "shell quartet exceeds the allocation-free ERI scratch");
                    break;
                }
                item_total += nit;
            }
            SeedQuartet q{};
            for (int i = 0; i < 4; ++i) q.sh[i] = r.sh[i];
            q.L = L;
            q.kab = (int)kab;
            q.kcd = (int)kcd;
            q.pq_begin = pq_total;
            q.seed_off = (long long)seed_zc;
            sq.push_back(q);
            seed_zc += need;
            pq_total += kab * kcd;
            maxL = std::max(maxL, L);
            ++end;
        }
        const size_t scr_zc = seed_scratch_zc(pq_total);

        // Work items of every class, one per pass of `tile` ket components,
        // each carrying its class; launched in a few LDS-size buckets, the
        // most expensive items first (a mixed-class batch would otherwise
        // leave the GPU idle behind launches of a handful of workgroups).
        struct Planned {
            WorkItem w;
            int lds;
            double cost;
            int nt = 256;  // workgroup size of the generic transform
        };
        std::vector<Planned> planned;
        std::map<int, std::vector<Planned>> mf_planned;  // FP64-MFMA items per instance
        std::map<int, std::vector<Planned>> sm_planned;  // lane-per-primitive-quartet items per class
        for (size_t i = begin; i < end; ++i) {
            const Request& r = reqs[i];
            const DevShell* S[4] = {&b->shells[r.sh[0]], &b->shells[r.sh[1]], &b->shells[r.sh[2]], &b->shells[r.sh[3]]};
            const int ncd_all = S[2]->ncart * S[3]->ncart;
            const int cd0 = r.cd1 < 0 ? 0 : r.cd0, cd1 = r.cd1 < 0 ? ncd_all : r.cd1;
            const int NHab = gh::nherm(S[0]->l + S[1]->l);
            if (!mfk.empty() && mfk[i] >= 0) {
                const int key = mfk[i], tile_w = gh::mfma_cd_tile((key >> 4) & 15);
                const double kq = (double)S[0]->nprim * S[1]->nprim * S[2]->nprim * S[3]->nprim;
                int lo[4], hi[4];
                for (int k = 0; k < 4; ++k) {
                    lo[k] = r.xfix[k] < 0 ? 0 : r.xfix[k];
                    hi[k] = r.xfix[k] < 0 ? S[k]->ncontr : r.xfix[k] + 1;
                }
                for (int xa = lo[0]; xa < hi[0]; ++xa)
                    for (int xb = lo[1]; xb < hi[1]; ++xb)
                        for (int xc = lo[2]; xc < hi[2]; ++xc)
                            for (int xd = lo[3]; xd < hi[3]; ++xd)
                                for (int t0 = cd0; t0 < cd1; t0 += tile_w) {
                                    Planned pl{};
                                    WorkItem& w = pl.w;
                                    for (int k = 0; k < 4; ++k) w.sh[k] = r.sh[k];
                                    w.x[0] = xa;
                                    w.x[1] = xb;
                                    w.x[2] = xc;
                                    w.x[3] = xd;
                                    w.cd0 = t0;
                                    w.cd1 = std::min(t0 + tile_w, cd1);
                                    w.seed_off = sq[i - begin].seed_off;
                                    w.out_base = r.out_base;
                                    for (int k = 0; k < 4; ++k) w.stride[k] = r.stride[k];
                                    w.cls = 0;
                                    // matrix-core work is about a tenth of the generic kernel's per MAC
                                    pl.cost = 0.1 * kq * NHab * gh::nherm(S[2]->l + S[3]->l) * tile_w;
                                    mf_planned[key].push_back(pl);
                                }
                continue;
            }
            const int L = S[0]->l + S[1]->l + S[2]->l + S[3]->l;
            if (small && !fixed && L <= ctx->small_max_l) {
                // one lane-per-primitive-quartet item per combination of contraction rows
                const int key = S[0]->l | (S[1]->l << 4) | (S[2]->l << 8) | (S[3]->l << 12);
                const double kq = (double)S[0]->nprim * S[1]->nprim * S[2]->nprim * S[3]->nprim;
                int lo[4], hi[4];
                for (int k = 0; k < 4; ++k) {
                    lo[k] = r.xfix[k] < 0 ? 0 : r.xfix[k];
                    hi[k] = r.xfix[k] < 0 ? S[k]->ncontr : r.xfix[k] + 1;
                }
                for (int xa = lo[0]; xa < hi[0]; ++xa)
                    for (int xb = lo[1]; xb < hi[1]; ++xb)
                        for (int xc = lo[2]; xc < hi[2]; ++xc)
                            for (int xd = lo[3]; xd < hi[3]; ++xd) {
                                Planned pl{};
                                WorkItem& w = pl.w;
                                for (int k = 0; k < 4; ++k) w.sh[k] = r.sh[k];
                                w.x[0] = xa;
                                w.x[1] = xb;
                                w.x[2] = xc;
                                w.x[3] = xd;
                                w.cd0 = cd0;
                                w.cd1 = cd1;
                                w.seed_off = sq[i - begin].seed_off;
                                w.out_base = r.out_base;
                                for (int k = 0; k < 4; ++k) w.stride[k] = r.stride[k];
                                w.cls = 0;
                                pl.cost = 0.05 * kq;
                                sm_planned[key].push_back(pl);
                            }
                continue;
            }
            gh::ClassArgs ca;
            int lds, nt;
            long long nit;
            const int st = plain_plan(ctx, b, r, ca, lds, nit, nt);
            if (st) return st;
            const double kq = (double)S[0]->nprim * S[1]->nprim * S[2]->nprim * S[3]->nprim;
            const int cls = S[0]->l | (S[1]->l << 4) | (S[2]->l << 8) | (S[3]->l << 12) | (ca.tile << 16);
            int lo[4], hi[4];
            for (int k = 0; k < 4; ++k) {
                lo[k] = r.xfix[k] < 0 ? 0 : r.xfix[k];
                hi[k] = r.xfix[k] < 0 ? S[k]->ncontr : r.xfix[k] + 1;
            }
            for (int xa = lo[0]; xa < hi[0]; ++xa)
                for (int xb = lo[1]; xb < hi[1]; ++xb)
                    for (int xc = lo[2]; xc < hi[2]; ++xc)
                        for (int xd = lo[3]; xd < hi[3]; ++xd)
                            for (int t0 = cd0; t0 < cd1; t0 += ca.tile) {
                                Planned pl{};
                                WorkItem& w = pl.w;
                                for (int k = 0; k < 4; ++k) w.sh[k] = r.sh[k];
                                w.x[0] = xa;
                                w.x[1] = xb;
                                w.x[2] = xc;
                                w.x[3] = xd;
                                w.cd0 = t0;
                                w.cd1 = std::min(t0 + ca.tile, cd1);
                                w.seed_off = sq[i - begin].seed_off;
                                w.out_base = r.out_base;
                                for (int k = 0; k < 4; ++k) w.stride[k] = r.stride[k];
                                w.cls = cls;
                                pl.lds = lds;
                                pl.nt = nt;
                                pl.cost = kq * (gh::nherm(L) + 0.5 * NHab * gh::nherm(S[2]->l + S[3]->l) *
                                                                   (w.cd1 - w.cd0));
                                planned.push_back(pl);
                            }
        }
        constexpr int kBuckets[] = {16 * 1024, 32 * 1024, 64 * 1024, 96 * 1024, 160 * 1024};
        std::vector<WorkItem> all;
        all.reserve(planned.size());
        struct MixedLaunch {
            int lds;
            size_t first, count;
            double cost;
            int mf = -1;   // FP64-MFMA instance key, or -1 for the generic kernel
            int nt = 256;  // workgroup size of the generic kernel
            int sm = -1;   // class of the lane-per-primitive-quartet kernel, or -1
        };
        std::vector<MixedLaunch> launches;
        for (const int nt : {64, 256})
            for (size_t bk = 0; bk < std::size(kBuckets); ++bk) {
                const int lo_b = bk ? kBuckets[bk - 1] : 0, hi_b = kBuckets[bk];
                std::vector<const Planned*> sel;
                int lds_max = 0;
                for (const Planned& pl : planned)
                    if (pl.nt == nt && pl.lds > lo_b && pl.lds <= hi_b) {
                        sel.push_back(&pl);
                        lds_max = std::max(lds_max, pl.lds);
                    }
                if (sel.empty()) continue;
                std::stable_sort(sel.begin(), sel.end(),
                                 [](const Planned* x, const Planned* y) { return x->cost > y->cost; });
                double cost = 0.0;
                for (const Planned* pl : sel) cost += pl->cost;
                launches.push_back({lds_max, all.size(), sel.size(), cost, -1, nt});
                for (const Planned* pl : sel) all.push_back(pl->w);
            }
        if (ctx->eri_queue == GIAOHIP_QUEUE_SINGLE && launches.size() > 1) {
            // one persistent launch over every item, longest first, with the
            // largest LDS footprint of any of them
            std::vector<const Planned*> sel;
            sel.reserve(planned.size());
            for (const Planned& pl : planned) sel.push_back(&pl);
            std::stable_sort(sel.begin(), sel.end(), [](const Planned* x, const Planned* y) { return x->cost > y->cost; });
            MixedLaunch one{0, 0, sel.size(), 0.0};
            all.clear();
            for (const Planned* pl : sel) {
                one.lds = std::max(one.lds, pl->lds);
                one.cost += pl->cost;
                all.push_back(pl->w);
            }
            launches.assign(1, one);
        }
        for (auto& [key, list] : mf_planned) {
            std::stable_sort(list.begin(), list.end(),
                             [](const Planned& x, const Planned& y) { return x.cost > y.cost; });
            double cost = 0.0;
            for (const Planned& pl : list) cost += pl.cost;
            launches.push_back({0, all.size(), list.size(), cost, key});
            for (const Planned& pl : list) all.push_back(pl.w);
        }
        for (auto& [key, list] : sm_planned) {
            std::stable_sort(list.begin(), list.end(),
                             [](const Planned& x, const Planned& y) { return x.cost > y.cost; });
            double cost = 0.0;
            for (const Planned& pl : list) cost += pl.cost;
            launches.push_back({0, all.size(), list.size(), cost, -1, 64, key});
            for (const Planned& pl : list) all.push_back(pl.w);
        }
        int* d_queue = nullptr;
        if (ctx->eri_queue != GIAOHIP_QUEUE_NONE) {
            if (fixed) {
                d_queue = fixed->queue;
            } else {
                HIP_TRY(ctx->queue.ensure(2 * sizeof(int)), "hipMalloc");
                d_queue = ctx->queue.as<int>();
            }
        }
        const int k = chunk & 1;  // slot of a pipelined chunk
        zc* d_seeds;
        SeedQuartet* d_sq;
        WorkItem* d_items;
        hipEvent_t ev_s0 = ctx->ev[0], ev_s1 = ctx->ev[1];
        if (fixed) {
            d_seeds = fixed->seeds;
            d_sq = fixed->sq;
            d_items = fixed->items;
            HIP_TRY(hipMemcpyAsync(d_sq, sq.data(), sq.size() * sizeof(SeedQuartet), hipMemcpyHostToDevice, s),
                    "hipMemcpyAsync");
            HIP_TRY(hipMemcpyAsync(d_items, all.data(), all.size() * sizeof(WorkItem), hipMemcpyHostToDevice, s),
                    "hipMemcpyAsync");
        } else {
            // the slot is free once the chunk two back has finished
            if (chunk >= 2) HIP_TRY(slot_stats(k), "kernel execution");
            HIP_TRY(ctx->pl_seeds[k].ensure((seed_zc + scr_zc) * sizeof(zc)), "hipMalloc(seeds)");
            HIP_TRY(ctx->pl_sq[k].ensure(sq.size() * sizeof(SeedQuartet)), "hipMalloc");
            HIP_TRY(ctx->pl_items[k].ensure(all.size() * sizeof(WorkItem)), "hipMalloc");
            HIP_TRY(ctx->pl_hsq[k].ensure(sq.size() * sizeof(SeedQuartet)), "hipHostMalloc");
            HIP_TRY(ctx->pl_hitems[k].ensure(all.size() * sizeof(WorkItem)), "hipHostMalloc");
            d_seeds = ctx->pl_seeds[k].as<zc>();
            d_sq = ctx->pl_sq[k].as<SeedQuartet>();
            d_items = ctx->pl_items[k].as<WorkItem>();
            std::memcpy(ctx->pl_hsq[k].p, sq.data(), sq.size() * sizeof(SeedQuartet));
            std::memcpy(ctx->pl_hitems[k].p, all.data(), all.size() * sizeof(WorkItem));
            HIP_TRY(hipMemcpyAsync(d_sq, ctx->pl_hsq[k].p, sq.size() * sizeof(SeedQuartet), hipMemcpyHostToDevice, ps),
                    "hipMemcpyAsync");
            HIP_TRY(hipMemcpyAsync(d_items, ctx->pl_hitems[k].p, all.size() * sizeof(WorkItem),
                                   hipMemcpyHostToDevice, ps),
                    "hipMemcpyAsync");
            ev_s0 = ctx->pl_ev[k][0];
            ev_s1 = ctx->pl_ev[k][1];
        }
        int* d_scratch = reinterpret_cast<int*>(d_seeds + seed_zc);

        HIP_TRY(hipEventRecord(ev_s0, ps), "hipEventRecord");
        HIP_TRY(gh::launch_seed(nmax_bucket(maxL), d_sq, (int)sq.size(), pq_total, b->d_shells.as<DevShell>(),
                                b->d_exps.as<double>(), d_kappa, d_seeds, d_err, d_scratch, ps),
                "seed kernel launch");
        HIP_TRY(hipEventRecord(ev_s1, ps), "hipEventRecord");
        // class launches alternate between the two streams (disjoint outputs),
        // so that a launch with few workgroups does not leave the GPU idle
        if (piped) HIP_TRY(hipStreamWaitEvent(s, ev_s1, 0), "hipStreamWaitEvent");
        HIP_TRY(hipStreamWaitEvent(ctx->aux, ev_s1, 0), "hipStreamWaitEvent");
        // greedy: the next launch goes to the stream with less work queued
        double queued[2] = {0.0, 0.0};
        std::vector<size_t> order(launches.size());
        std::iota(order.begin(), order.end(), size_t{0});
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t x, size_t y) { return launches[x].cost > launches[y].cost; });
        for (size_t li : order) {
            const MixedLaunch& L = launches[li];
            const int side = queued[1] < queued[0] ? 1 : 0;
            queued[side] += L.cost;
            hipStream_t ls = side ? ctx->aux : s;
            if (L.sm >= 0) {
                HIP_TRY(gh::launch_small(L.sm & 15, (L.sm >> 4) & 15, (L.sm >> 8) & 15, (L.sm >> 12) & 15,
                                         d_items + L.first, (int)L.count, b->d_shells.as<DevShell>(),
                                         b->d_exps.as<double>(), b->d_coefs.as<double>(), b->d_norms.as<double>(),
                                         d_kappa, d_seeds, d_out, ls),
                        "small-class kernel launch");
                ctx->stats.launches += 1;
                continue;
            }
            if (L.mf >= 0) {
                const int lab = (L.mf >> 4) & 15, lcd = L.mf & 15;
                if (L.mf & kMfmaSplitKey)
                    HIP_TRY(gh::launch_mfma_split(lab, lcd, d_items + L.first, (int)L.count, b->d_shells.as<DevShell>(),
                                                  b->d_exps.as<double>(), b->d_coefs.as<double>(),
                                                  b->d_norms.as<double>(), d_kappa, d_seeds, d_out, ls),
                            "MFMA transform kernel launch");
                else
                    HIP_TRY(gh::launch_mfma(lab, lcd, ctx->mfma_variant, d_items + L.first, (int)L.count,
                                            b->d_shells.as<DevShell>(), b->d_exps.as<double>(),
                                            b->d_coefs.as<double>(), b->d_norms.as<double>(), d_kappa, d_seeds,
                                            d_out, ls),
                            "MFMA transform kernel launch");
                ctx->stats.launches += 1;
                continue;
            }
            HIP_TRY(gh::launch_transform_mixed(L.nt, L.lds, d_items + L.first, (int)L.count, b->d_shells.as<DevShell>(),
                                               b->d_exps.as<double>(), b->d_coefs.as<double>(),
                                               b->d_norms.as<double>(), d_kappa, d_seeds, d_out,
                                               d_queue ? d_queue + side : nullptr, ls),
                    "transform kernel launch");
            ctx->stats.launches += 1;
        }
        ctx->stats.contract_launches += launches.size();
        ctx->stats.launches += 1;
        ctx->stats.primitive_quartets += (uint64_t)pq_total;
        if (piped) {
            // no join: each stream runs on into the next chunk; the slot is
            // free once both have passed this chunk
            HIP_TRY(hipEventRecord(ctx->pl_ev[k][2], s), "hipEventRecord");
            HIP_TRY(hipEventRecord(ctx->pl_ev[k][3], ctx->aux), "hipEventRecord");
        } else {
            HIP_TRY(hipEventRecord(ctx->ev[4], ctx->aux), "hipEventRecord");
            HIP_TRY(hipStreamWaitEvent(s, ctx->ev[4], 0), "hipStreamWaitEvent");
            HIP_TRY(hipEventRecord(ctx->ev[2], s), "hipEventRecord");
            HIP_TRY(hipEventSynchronize(ctx->ev[2]), "kernel execution");
            ctx->stats.seed_ms += elapsed(ctx->ev[0], ctx->ev[1]);
            ctx->stats.transform_ms += elapsed(ctx->ev[1], ctx->ev[2]);
            ctx->stats.contract_ms += elapsed(ctx->ev[1], ctx->ev[2]);
            DevError de{};
            HIP_TRY(hipMemcpy(&de, d_err, sizeof de, hipMemcpyDeviceToHost), "hipMemcpy");
            if (de.code != 0) return boys_error(ctx, de);
        }
        begin = end;
        ++chunk;
    }
    if (piped) {
        for (int c = std::max(0, chunk - 2); c < chunk; ++c) HIP_TRY(slot_stats(c & 1), "kernel execution");
        // later work on the context stream follows the aux stream's transforms
        HIP_TRY(hipEventRecord(ctx->ev[4], ctx->aux), "hipEventRecord");
        HIP_TRY(hipStreamWaitEvent(s, ctx->ev[4], 0), "hipStreamWaitEvent");
        DevError de{};
        HIP_TRY(hipMemcpy(&de, d_err, sizeof de, hipMemcpyDeviceToHost), "hipMemcpy");
        if (de.code != 0) return boys_error(ctx, de);
    }
    return GIAOHIP_OK;
}


// One request evaluated through a symmetry image of its quartet by the
// class-specialized kernels: (ab|cd) = (cd|ab) and (ab|cd) = conj (ba|dc).
struct FastReq {
    int sh[4];    // oriented shells
    int perm[4];  // oriented slot k holds original slot perm[k]
    int conj;
    const Request* r;
    gh::FastClassInfo info;
};

const int kPerms[4][4] = {{0, 1, 2, 3}, {2, 3, 0, 1}, {1, 0, 3, 2}, {3, 2, 1, 0}};

// The batched kernel's orientation of a class: the first of the four images
// (ab|cd), (cd|ab), conj (ba|dc), conj (dc|ba) with a specialized kernel, or -1.
signed char class_orientation(int la, int lb, int lc, int ld, gh::FastClassInfo* info) {
    const int l[4] = {la, lb, lc, ld};
    for (int o = 0; o < 4; ++o) {
        const gh::FastClassInfo f = gh::fast_class_info(l[kPerms[o][0]], l[kPerms[o][1]], l[kPerms[o][2]], l[kPerms[o][3]]);
        if (f.available) {
            *info = f;
            return (signed char)o;
        }
    }
    *info = gh::FastClassInfo{};
    return -1;
}

// Admission status and batched-kernel orientation of every class of the basis.
void build_class_tables(giaohip_basis* b) {
    int lmax = 0;
    for (const DevShell& s : b->shells) lmax = std::max(lmax, s.l);
    const int lm = lmax + 1;
    b->lm = lm;
    const size_t n = (size_t)lm * lm * lm * lm;
    b->cls_status.assign(n, kClsOk);
    b->cls_orient.assign(n, -1);
    b->cls_info.assign(n, gh::FastClassInfo{});
    for (int la = 0; la < lm; ++la)
        for (int lb = 0; lb < lm; ++lb)
            for (int lc = 0; lc < lm; ++lc)
                for (int ld = 0; ld < lm; ++ld) {
                    const size_t i = (((size_t)la * lm + lb) * lm + lc) * lm + ld;
                    b->cls_status[i] = class_admission(la, lb, lc, ld);
                    if (b->cls_status[i] == kClsOk) b->cls_orient[i] = class_orientation(la, lb, lc, ld, &b->cls_info[i]);
                }
}

bool orient(const giaohip_basis* b, const Request& r, FastReq& f) {
    const size_t ci = class_index(b, r.sh);
    int o = b->cls_orient[ci];
    if (o < 0) return false;
    // Where the image with the pairs exchanged has the same class (bra and
    // ket pairs of equal momenta), it uses the same kernel: the pair with
    // more primitive pairs goes to the bra side, whose primitive pairs are
    // the kernel's rows.
    {
        const int o2 = o ^ 1;
        bool same = true;
        for (int k = 0; k < 4; ++k)
            same = same && b->shells[r.sh[kPerms[o2][k]]].l == b->shells[r.sh[kPerms[o][k]]].l;
        auto kbra = [&](int img) {
            return b->shells[r.sh[kPerms[img][0]]].nprim * b->shells[r.sh[kPerms[img][1]]].nprim;
        };
        if (same && kbra(o2) > kbra(o)) o = o2;
    }
    for (int k = 0; k < 4; ++k) {
        f.sh[k] = r.sh[kPerms[o][k]];
        f.perm[k] = kPerms[o][k];
    }
    f.conj = o >= 2 ? 1 : 0;
    f.r = &r;
    f.info = b->cls_info[ci];
    return true;
}

// The batched path. Requests are cut into chunks whose seeds, R0 tables, pair
// tables and quadrature lists fit the chunk budget. Two streams pipeline the
// chunks: the producer stream (`aux`) runs a chunk's seed, pair-table and R0
// kernels while the consumer stream (`stream`) runs the previous chunk's
// transform, so the latency-bound producer kernels share the CUs with the
// FMA-bound transform. Chunks alternate between two work buffers; events order
// a chunk's transform after its producer and a buffer's reuse after the
// transform that read it. Descriptors are planned on the host into
// page-locked staging while the device works, and the host waits once.
int fast_finish(giaohip_ctx* ctx);

int run_fast(giaohip_ctx* ctx, const giaohip_basis* b, std::span<const FastReq> reqs, zc* d_out,
             bool finish = true) {
    if (reqs.empty()) return GIAOHIP_OK;
    // the producers run on the second stream (overlapping the previous
    // chunk's contraction) or, serially, on the context stream
    hipStream_t sc = ctx->stream, sp = ctx->pipeline == GIAOHIP_PIPELINE_SERIAL ? ctx->stream : ctx->aux;
    constexpr size_t kScratchHead = (gh::kSeedScratchHead * sizeof(int) + sizeof(zc) - 1) / sizeof(zc);  // counters
    struct Need {
        size_t seed, r0, tab, scr;
        long long kab, kcd;
        int L;
        size_t items;
    };
    // pass 1: sizes
    BigVec<Need> need(reqs.size());
    parallel_for(reqs.size(), [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            const FastReq& f = reqs[i];
            const DevShell* S[4] = {&b->shells[f.sh[0]], &b->shells[f.sh[1]], &b->shells[f.sh[2]], &b->shells[f.sh[3]]};
            Need& n = need[i];
            n.L = S[0]->l + S[1]->l + S[2]->l + S[3]->l;
            n.kab = (long long)S[0]->nprim * S[1]->nprim;
            n.kcd = (long long)S[2]->nprim * S[3]->nprim;
            n.seed = (size_t)(n.kab * n.kcd) * (size_t)(n.L + 4);
            n.r0 = (size_t)(n.kab * n.kcd) * (size_t)f.info.nhl;
            n.tab = (size_t)(n.kab * f.info.bra_ne + n.kcd * f.info.ket_ne);
            // quadrature lists (index and quartet per entry) and the block map of the seed kernel
            const size_t units = (size_t)(n.kab * n.kcd);
            n.scr = ((2 * (size_t)gh::kBoysBins * units + units / 256 + 2) * sizeof(int) + sizeof(zc) - 1) / sizeof(zc);
            n.items = 1;
            for (int k = 0; k < 4; ++k)
                if (f.r->xfix[f.perm[k]] < 0) n.items *= (size_t)S[k]->ncontr;
        }
    });
    size_t total_items = 0, max_one = 0;
    for (const Need& n : need) {
        total_items += n.items;
        max_one = std::max(max_one, n.seed + n.r0 + n.tab + n.scr + kScratchHead);
    }
    const size_t work_zc = std::max(max_one, ctx->seed_budget_bytes / sizeof(zc));
    for (DevBuffer& w : ctx->work) HIP_TRY(w.ensure(work_zc * sizeof(zc)), "hipMalloc(batched work buffer)");
    HIP_TRY(ctx->squartets.ensure(reqs.size() * sizeof(SeedQuartet)), "hipMalloc");
    HIP_TRY(ctx->fpq.ensure(reqs.size() * sizeof(gh::FastPairQuartet)), "hipMalloc");
    HIP_TRY(ctx->fitems.ensure(total_items * sizeof(gh::FastItem)), "hipMalloc");
    HIP_TRY(ctx->h_sq.ensure(reqs.size() * sizeof(SeedQuartet)), "hipHostMalloc");
    HIP_TRY(ctx->h_pq.ensure(reqs.size() * sizeof(gh::FastPairQuartet)), "hipHostMalloc");
    HIP_TRY(ctx->h_items.ensure(total_items * sizeof(gh::FastItem)), "hipHostMalloc");
    SeedQuartet* hsq = ctx->h_sq.as<SeedQuartet>();
    gh::FastPairQuartet* hpq = ctx->h_pq.as<gh::FastPairQuartet>();
    gh::FastItem* hit = ctx->h_items.as<gh::FastItem>();
    // events per chunk: 0 producer start, 1 seeds done, 2 producer done, 3 transform start, 4 transform done
    constexpr int kEv = 5;
    auto event = [&](size_t c, int k) -> hipEvent_t& { return ctx->chunk_ev[kEv * c + k]; };
    // the producer starts after the work already queued on the main stream (field, error record)
    hipEvent_t start_ev = ctx->ev[2];
    HIP_TRY(hipEventRecord(start_ev, sc), "hipEventRecord");
    HIP_TRY(hipStreamWaitEvent(sp, start_ev, 0), "hipStreamWaitEvent");

    struct Group {
        std::pair<gh::FastLaunch, gh::FastLaunch> fn;
        int waves;
        std::vector<size_t> idx;
    };
    std::vector<Group> groups;
    size_t nsq = 0, npq = 0, nit = 0, nchunks = 0;
    size_t begin = 0;
    while (begin < reqs.size()) {
        // pass 2: one chunk within the budget
        size_t end = begin, seed_zc = 0, r0_zc = 0, tab_zc = 0, scr_zc = kScratchHead;
        long long pq_total = 0;
        int maxL = 0;
        const size_t sq0 = nsq;
        while (end < reqs.size()) {
            const Need& n = need[end];
            if (end > begin && seed_zc + r0_zc + tab_zc + scr_zc + n.seed + n.r0 + n.tab + n.scr > work_zc) break;
            SeedQuartet& q = hsq[nsq++];
            q = SeedQuartet{};
            for (int k = 0; k < 4; ++k) q.sh[k] = reqs[end].sh[k];
            q.L = n.L;
            q.kab = (int)n.kab;
            q.kcd = (int)n.kcd;
            q.pq_begin = pq_total;
            q.seed_off = (long long)seed_zc;
            seed_zc += n.seed;
            r0_zc += n.r0;
            tab_zc += n.tab;
            scr_zc += n.scr;
            pq_total += n.kab * n.kcd;
            maxL = std::max(maxL, n.L);
            ++end;
        }
        zc* work = ctx->work[nchunks & 1].as<zc>();
        zc* d_seeds = work;
        zc* d_r0 = d_seeds + seed_zc;
        zc* d_tab = d_r0 + r0_zc;
        int* d_scratch = reinterpret_cast<int*>(d_tab + tab_zc);
        // group by class kernels and transform workgroup size
        for (Group& g : groups) g.idx.clear();
        for (size_t i = begin; i < end; ++i) {
            const FastReq& f = reqs[i];
            const int waves = gh::fast_waves(f.info.nhab, (int)need[i].kab, ctx->fast_waves_cap);
            const std::pair<gh::FastLaunch, gh::FastLaunch> fn{f.info.produce, f.info.transform};
            Group* g = nullptr;
            for (Group& x : groups)
                if (x.fn == fn && x.waves == waves) g = &x;
            if (!g) {
                groups.push_back(Group{fn, waves, {}});
                g = &groups.back();
            }
            g->idx.push_back(i);
        }
        struct Launch {
            std::pair<gh::FastLaunch, gh::FastLaunch> fn;
            int waves;
            size_t pq0, npq, it0, nit;
        };
        std::vector<Launch> launches;
        const size_t pq_chunk0 = npq, it_chunk0 = nit;
        long long toff = 0, roff = 0;
        for (const Group& g : groups) {
            if (g.idx.empty()) continue;
            Launch L{g.fn, g.waves, npq, g.idx.size(), nit, 0};
            for (size_t i : g.idx) {
                const FastReq& f = reqs[i];
                const Need& n = need[i];
                gh::FastPairQuartet& pq = hpq[npq++];
                pq = gh::FastPairQuartet{};
                for (int k = 0; k < 4; ++k) pq.sh[k] = f.sh[k];
                pq.table_off[0] = toff;
                toff += n.kab * f.info.bra_ne;
                pq.table_off[1] = toff;
                toff += n.kcd * f.info.ket_ne;
                pq.seed_off = hsq[sq0 + (i - begin)].seed_off;
                pq.r0_off = roff;
                pq.kab = (int)n.kab;
                pq.kcd = (int)n.kcd;
                roff += n.kab * n.kcd * f.info.nhl;
                int lo[4], hi[4];
                for (int k = 0; k < 4; ++k) {
                    const int orig = f.perm[k];
                    lo[k] = f.r->xfix[orig] < 0 ? 0 : f.r->xfix[orig];
                    hi[k] = f.r->xfix[orig] < 0 ? b->shells[f.sh[k]].ncontr : f.r->xfix[orig] + 1;
                }
                for (int xa = lo[0]; xa < hi[0]; ++xa)
                    for (int xb = lo[1]; xb < hi[1]; ++xb)
                        for (int xc = lo[2]; xc < hi[2]; ++xc)
                            for (int xd = lo[3]; xd < hi[3]; ++xd) {
                                gh::FastItem& w = hit[nit++];
                                w = gh::FastItem{};
                                for (int k = 0; k < 4; ++k) {
                                    w.sh[k] = f.sh[k];
                                    w.stride[k] = f.r->stride[f.perm[k]];
                                }
                                w.x[0] = xa;
                                w.x[1] = xb;
                                w.x[2] = xc;
                                w.x[3] = xd;
                                w.conj = f.conj;
                                w.r0_off = pq.r0_off;
                                w.ebra_off = pq.table_off[0];
                                w.eket_off = pq.table_off[1];
                                w.out_base = f.r->out_base;
                            }
            }
            L.nit = nit - L.it0;
            launches.push_back(L);
        }
        while (ctx->chunk_ev.size() < kEv * (nchunks + 1)) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e), "hipEventCreate");
            ctx->chunk_ev.push_back(e);
        }
        // producer: wait until the transform that last read this work buffer is done
        if (nchunks >= 2) HIP_TRY(hipStreamWaitEvent(sp, event(nchunks - 2, 4), 0), "hipStreamWaitEvent");
        HIP_TRY(hipMemcpyAsync(ctx->squartets.as<SeedQuartet>() + sq0, hsq + sq0, (nsq - sq0) * sizeof(SeedQuartet),
                               hipMemcpyHostToDevice, sp),
                "hipMemcpyAsync");
        HIP_TRY(hipMemcpyAsync(ctx->fpq.as<gh::FastPairQuartet>() + pq_chunk0, hpq + pq_chunk0,
                               (npq - pq_chunk0) * sizeof(gh::FastPairQuartet), hipMemcpyHostToDevice, sp),
                "hipMemcpyAsync");
        HIP_TRY(hipMemcpyAsync(ctx->fitems.as<gh::FastItem>() + it_chunk0, hit + it_chunk0,
                               (nit - it_chunk0) * sizeof(gh::FastItem), hipMemcpyHostToDevice, sp),
                "hipMemcpyAsync");
        HIP_TRY(hipEventRecord(event(nchunks, 0), sp), "hipEventRecord");
        HIP_TRY(gh::launch_seed(nmax_bucket(maxL), ctx->squartets.as<SeedQuartet>() + sq0, (int)(nsq - sq0),
                                pq_total, b->d_shells.as<DevShell>(), b->d_exps.as<double>(),
                                ctx->kappa.as<double>(), d_seeds, ctx->errrec.as<DevError>(), d_scratch, sp,
                                b->d_radial.as<double>()),
                "seed kernel launch");
        HIP_TRY(hipEventRecord(event(nchunks, 1), sp), "hipEventRecord");
        auto args = [&](const Launch& L) {
            gh::FastLaunchArgs a{};
            a.pq = ctx->fpq.as<gh::FastPairQuartet>() + L.pq0;
            a.npq = (int)L.npq;
            a.items = ctx->fitems.as<gh::FastItem>() + L.it0;
            a.nitems = (int)L.nit;
            a.waves = L.waves;
            a.shells = b->d_shells.as<DevShell>();
            a.exps = b->d_exps.as<double>();
            a.coefs = b->d_coefs.as<double>();
            a.kappa = ctx->kappa.as<double>();
            a.seeds = d_seeds;
            a.r0 = d_r0;
            a.etab = d_tab;
            a.out = d_out;
            return a;
        };
        for (const Launch& L : launches) {
            HIP_TRY(L.fn.first(args(L), sp), "batched pair/R0 launch");
            ctx->stats.launches += 3;
        }
        HIP_TRY(hipEventRecord(event(nchunks, 2), sp), "hipEventRecord");
        // consumer: the transform after this chunk's producer
        HIP_TRY(hipStreamWaitEvent(sc, event(nchunks, 2), 0), "hipStreamWaitEvent");
        HIP_TRY(hipEventRecord(event(nchunks, 3), sc), "hipEventRecord");
        for (const Launch& L : launches) {
            HIP_TRY(L.fn.second(args(L), sc), "batched transform launch");
            ctx->stats.launches += 1;
            ctx->stats.contract_launches += 1;
        }
        HIP_TRY(hipEventRecord(event(nchunks, 4), sc), "hipEventRecord");
        ctx->stats.launches += 2;
        ctx->stats.primitive_quartets += (uint64_t)pq_total;
        ++nchunks;
        begin = end;
    }
    ctx->fast_pending = nchunks;
    return finish ? fast_finish(ctx) : GIAOHIP_OK;
}

// Waits for the queued batched-path chunks, collects their timings and checks
// the error record.
int fast_finish(giaohip_ctx* ctx) {
    const size_t nchunks = ctx->fast_pending;
    ctx->fast_pending = 0;
    if (nchunks == 0) return GIAOHIP_OK;
    constexpr int kEv = 5;
    auto event = [&](size_t c, int k) -> hipEvent_t& { return ctx->chunk_ev[kEv * c + k]; };
    HIP_TRY(hipEventSynchronize(event(nchunks - 1, 4)), "kernel execution");
    for (size_t c = 0; c < nchunks; ++c) {
        ctx->stats.seed_ms += elapsed(event(c, 0), event(c, 1));
        const double produce = elapsed(event(c, 1), event(c, 2)), contract = elapsed(event(c, 3), event(c, 4));
        ctx->stats.transform_ms += produce + contract;
        ctx->stats.produce_ms += produce;
        ctx->stats.contract_ms += contract;
    }
    DevError de{};
    HIP_TRY(hipMemcpy(&de, ctx->errrec.p, sizeof de, hipMemcpyDeviceToHost), "hipMemcpy");
    if (de.code != 0) return boys_error(ctx, de);
    return GIAOHIP_OK;
}

// Runs a list of requests; outputs go to d_out (device). Full-block requests
// of classes with a specialized kernel take the batched path (unless the
// plain variant is selected); the rest the plain per-class kernel.
int run_requests(giaohip_ctx* ctx, const giaohip_basis* b, const double field[6], std::span<const Request> reqs,
                 zc* d_out) {
    if (reqs.empty()) return GIAOHIP_OK;
    // Split into plain and batched requests, the batched ones grouped by class
    // kernel (input order kept within a class, classes in class-table order):
    // a chunk then holds few classes and each launch many items. Two parallel
    // passes over contiguous thread ranges: the class key of every request and
    // per-range counts, then a stable scatter.
    const size_t n = reqs.size();
    const bool batched = ctx->eri_variant == GIAOHIP_ERI_BATCHED;
    const bool mfma = batched && ctx->mfma_mode != GIAOHIP_MFMA_OFF;
    const size_t nkey = b->cls_orient.size() + 1;  // key 0: plain, 1 + class index: batched
    auto key_of = [&](const Request& r, FastReq& f) -> size_t {
        if (!batched || r.cd1 >= 0 || !orient(b, r, f)) return 0;
        if (ctx->mfma_mode >= GIAOHIP_MFMA_ALL) {
            Request m = r;
            if (mfma_orient(b, m, ctx->mfma_mode, ctx->mfma_min_l) >= 0) return 0;
        }
        return 1 + class_index(b, f.sh);
    };
    const size_t T = std::max<size_t>(1, std::min(host_threads(), (n + 8191) / 8192));
    const size_t per = (n + T - 1) / T;
    std::vector<size_t> count(T * nkey, 0);
    auto in_ranges = [&](auto&& body) {
        std::vector<std::thread> th;
        for (size_t t = 1; t < T; ++t) th.emplace_back([&, t] { body(t); });
        body(0);
        for (auto& x : th) x.join();
    };
    in_ranges([&](size_t t) {
        size_t* c = count.data() + t * nkey;
        FastReq f;
        for (size_t i = t * per, e = std::min(n, (t + 1) * per); i < e; ++i) ++c[key_of(reqs[i], f)];
    });
    // exclusive offsets: plain list for key 0, batched list in key order
    size_t nplain = 0, nfast = 0;
    for (size_t k = 0; k < nkey; ++k)
        for (size_t t = 0; t < T; ++t) {
            size_t& c = count[t * nkey + k];
            const size_t v = c;
            c = k == 0 ? nplain : nfast;
            (k == 0 ? nplain : nfast) += v;
        }
    BigVec<Request> plain(nplain);
    BigVec<FastReq> fast(nfast);
    in_ranges([&](size_t t) {
        size_t* c = count.data() + t * nkey;
        FastReq f;
        for (size_t i = t * per, e = std::min(n, (t + 1) * per); i < e; ++i) {
            const size_t k = key_of(reqs[i], f);
            if (k == 0) plain[c[0]++] = reqs[i];
            else fast[c[k]++] = f;
        }
    });
    if (fast.empty()) return run_plain(ctx, b, field, plain, d_out, nullptr, mfma, batched);
    // The batched part first, queued without waiting: the plain part's host
    // planning then overlaps its execution, and the plain part's final wait
    // and error check cover both (shared context stream and error record).
    hipStream_t s = ctx->stream;
    const std::vector<double> kap = wave_vectors(b, field);
    HIP_TRY(ctx->kappa.ensure(kap.size() * sizeof(double)), "hipMalloc");
    HIP_TRY(hipMemcpyAsync(ctx->kappa.p, kap.data(), kap.size() * sizeof(double), hipMemcpyHostToDevice, s),
            "hipMemcpyAsync");
    HIP_TRY(ctx->errrec.ensure(sizeof(DevError)), "hipMalloc");
    HIP_TRY(hipMemsetAsync(ctx->errrec.p, 0, sizeof(DevError), s), "hipMemsetAsync");
    int st = run_fast(ctx, b, fast, d_out, plain.empty());
    if (st || plain.empty()) return st;
    st = run_plain(ctx, b, field, plain, d_out, nullptr, mfma, batched, true);
    const int fst = fast_finish(ctx);
    return st ? st : fst;
}

void reset_stats(giaohip_ctx* ctx) { ctx->stats = giaohip_stats{}; }

long long ao_count(const giaohip_basis* b, int s) { return b->shells[s].ao_count; }

// Cholesky diagonal D[nu,mu] = D[mu,nu] = Re (mu nu|nu mu) into device memory
// (cholesky.cpp:285-314): (ab|ba) blocks, then the diagonal extraction.
int diagonal_to_device(giaohip_ctx* ctx, const giaohip_basis* b, const uint64_t* pairs, size_t npairs,
                       const double field[6], double* d_diag) {
    const int ns = b->ns;
    const long long nao = b->nao;
    std::vector<char> seen((size_t)ns * ns + 1, 0);
    std::vector<int> pr;
    std::vector<Request> reqs;
    std::vector<long long> boff;
    long long total = 0;
    for (size_t i = 0; i < npairs; ++i) {
        const uint64_t a = pairs[2 * i], bb = pairs[2 * i + 1];
        if (a >= (uint64_t)ns || bb >= (uint64_t)ns || a > bb)
            return set_error(ctx, GIAOHIP_OUT_OF_RANGE, "Cholesky diagonal shell pair must be valid and ordered");
        if (seen[a * ns + bb]) return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "duplicate Cholesky diagonal shell pair");
        seen[a * ns + bb] = 1;
    }
    for (size_t i = 0; i < npairs; ++i) {
        const int a = (int)pairs[2 * i], bb = (int)pairs[2 * i + 1];
        const int q[4] = {a, bb, bb, a};
        const int st = check_quartet_supported(ctx, b, q);
        if (st) return st;
        Request r{};
        for (int k = 0; k < 4; ++k) {
            r.sh[k] = q[k];
            r.xfix[k] = -1;
        }
        r.cd0 = 0;
        r.cd1 = -1;
        const long long na = ao_count(b, a), nb = ao_count(b, bb);
        r.out_base = total;
        r.stride[0] = nb * nb * na;
        r.stride[1] = nb * na;
        r.stride[2] = na;
        r.stride[3] = 1;
        reqs.push_back(r);
        boff.push_back(total);
        pr.push_back(a);
        pr.push_back(bb);
        total += na * nb * nb * na;
    }
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    HIP_TRY(ctx->blocks.ensure(std::max<long long>(total, 1) * sizeof(zc)), "hipMalloc(blocks)");
    HIP_TRY(ctx->misc_a.ensure(boff.size() * sizeof(long long) + pr.size() * sizeof(int) + 64), "hipMalloc");
    int st = run_requests(ctx, b, field, reqs, ctx->blocks.as<zc>());
    if (st) return st;
    long long* d_boff = ctx->misc_a.as<long long>();
    int* d_pr = reinterpret_cast<int*>(d_boff + boff.size());
    if (!boff.empty()) {
        HIP_TRY(hipMemcpyAsync(d_boff, boff.data(), boff.size() * sizeof(long long), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
        HIP_TRY(hipMemcpyAsync(d_pr, pr.data(), pr.size() * sizeof(int), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    }
    HIP_TRY(hipMemsetAsync(d_diag, 0, (size_t)nao * nao * sizeof(double), s), "hipMemsetAsync");
    HIP_TRY(gh::launch_diag_extract(d_boff, d_pr, (int)npairs, b->d_shells.as<DevShell>(), b->d_aooff.as<long long>(), nao,
                                    ctx->blocks.as<zc>(), d_diag, s),
            "diagonal launch");
    HIP_TRY(hipStreamSynchronize(s), "diagonal execution");
    return GIAOHIP_OK;
}

}  // namespace

extern "C" {

int giaohip_host_threads(void) { return (int)host_threads(); }

unsigned long long giaohip_device_allocations(void) { return g_device_allocations.load(); }

double giaohip_primitive_normalization(double exponent, int ax, int ay, int az) {
    return primitive_normalization(exponent, ax, ay, az);
}

// types.cpp:177-188 (same-centre overlap of two normalized primitives)
static double same_center_overlap(double alpha, double beta, int l) {
    const double norm = primitive_normalization(alpha, l, 0, 0) * primitive_normalization(beta, l, 0, 0);
    const double p = alpha + beta;
    const double lm = std::lgamma((double)l + 0.5) - ((double)l + 0.5) * std::log(p);
    return norm * std::exp(lm) * kPi / p;
}

int giaohip_normalize_shells(int n, const int* L, const int* nprim, const int* ncontr, const double* exps,
                             const double* coefs_in, double* coefs_out, char* message, size_t message_size) {
    auto fail = [&](const char* m) {
        if (message && message_size) {
            std::snprintf(message, message_size, "%s", m);
        }
        return GIAOHIP_INVALID_ARGUMENT;
    };
    size_t po = 0, co = 0;
    for (int s = 0; s < n; ++s) {
        const int np = nprim[s], nc = ncontr[s];
        if (np < 1) return fail("a shell requires at least one primitive");
        if (nc < 1) return fail("a shell requires at least one contraction");
        const double* e = exps + po;
        for (int i = 0; i < np; ++i)
            if (!std::isfinite(e[i]) || e[i] <= 0.0) return fail("shell exponents must be finite and positive");
        for (int i = 0; i < np * nc; ++i)
            if (!std::isfinite(coefs_in[co + i])) return fail("shell coefficients must be finite");
        if (coefs_out != coefs_in) std::memcpy(coefs_out + co, coefs_in + co, sizeof(double) * (size_t)(np * nc));
        for (int x = 0; x < nc; ++x) {
            double* c = coefs_out + co + (size_t)x * np;
            bool any = false;
            for (int i = 0; i < np; ++i) any = any || c[i] != 0.0;
            if (!any) return fail("a shell contraction cannot have all-zero coefficients");
            double norm2 = 0.0;
            for (int i = 0; i < np; ++i)
                for (int j = 0; j < np; ++j) norm2 += c[i] * c[j] * same_center_overlap(e[i], e[j], L[s]);
            if (!std::isfinite(norm2) || norm2 <= 0.0) return fail("contraction has a non-positive or non-finite norm");
            const double scale = 1.0 / std::sqrt(norm2);
            for (int i = 0; i < np; ++i) c[i] *= scale;
        }
        po += (size_t)np;
        co += (size_t)np * nc;
    }
    return GIAOHIP_OK;
}

int giaohip_create(int device, giaohip_ctx** out) {
    if (!out) return GIAOHIP_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return GIAOHIP_DEVICE_ERROR;
    if (device < 0 || device >= count) return GIAOHIP_INVALID_ARGUMENT;
    auto* ctx = new giaohip_ctx;
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->plain_stream, hipStreamNonBlocking);
    for (int i = 0; i < 5 && e == hipSuccess; ++i) e = hipEventCreate(&ctx->ev[i]);
    for (int i = 0; i < 8 && e == hipSuccess; ++i) e = hipEventCreate(&ctx->pl_ev[i / 4][i % 4]);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->pl_init, hipEventDisableTiming);
    for (int i = 0; i < 2 && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&ctx->copy_ev[i], hipEventDisableTiming);
    if (e == hipSuccess) e = gh::upload_constants();
    if (e == hipSuccess) e = noalloc_scratch_init(device, &ctx->noalloc);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        giaohip_destroy(ctx);
        return GIAOHIP_DEVICE_ERROR;
    }
    *out = ctx;
    return GIAOHIP_OK;
}

int giaohip_destroy(giaohip_ctx* ctx) {
    if (!ctx) return GIAOHIP_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->aux) (void)hipStreamSynchronize(ctx->aux);
    if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
    if (ctx->plain_stream) (void)hipStreamSynchronize(ctx->plain_stream);
    for (int k = 0; k < 2; ++k) {
        for (DevBuffer* b : {&ctx->pl_seeds[k], &ctx->pl_sq[k], &ctx->pl_items[k]}) b->release();
        for (PinnedBuffer* b : {&ctx->pl_hsq[k], &ctx->pl_hitems[k]}) b->release();
        for (auto& ev : ctx->pl_ev[k])
            if (ev) (void)hipEventDestroy(ev);
    }
    if (ctx->pl_init) (void)hipEventDestroy(ctx->pl_init);
    ctx->staging2.release();
    for (PinnedBuffer& b : ctx->bounce) b.release();
    for (auto& ev : ctx->copy_ev)
        if (ev) (void)hipEventDestroy(ev);
    for (DevBuffer* b : {&ctx->seeds, &ctx->squartets, &ctx->items, &ctx->kappa, &ctx->errrec, &ctx->staging,
                         &ctx->blocks, &ctx->misc_a, &ctx->misc_b, &ctx->misc_c, &ctx->fpq, &ctx->fitems,
                         &ctx->queue})
        b->release();
    for (auto& ev : ctx->chunk_ev)
        if (ev) (void)hipEventDestroy(ev);
    for (DevBuffer& w : ctx->work) w.release();
    for (PinnedBuffer* b : {&ctx->h_sq, &ctx->h_pq, &ctx->h_items}) b->release();
    for (void* p : ctx->user_allocs) (void)hipFree(p);
    for (auto& ev : ctx->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->blas) (void)rocblas_destroy_handle(ctx->blas);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->aux) (void)hipStreamDestroy(ctx->aux);
    if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
    if (ctx->plain_stream) (void)hipStreamDestroy(ctx->plain_stream);
    delete ctx;
    return GIAOHIP_OK;
}

size_t giaohip_last_error(const giaohip_ctx* ctx, char* buffer, size_t size) {
    if (!ctx) return 0;
    if (buffer && size) {
        const size_t n = std::min(size - 1, ctx->err.size());
        std::memcpy(buffer, ctx->err.data(), n);
        buffer[n] = '\0';
    }
    return ctx->err.size();
}

void* giaohip_stream(giaohip_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int giaohip_set_option(giaohip_ctx* ctx, int option, int value) {
    if (!ctx) return GIAOHIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (option == GIAOHIP_OPTION_ERI_VARIANT && (value == GIAOHIP_ERI_PLAIN || value == GIAOHIP_ERI_BATCHED)) {
        ctx->eri_variant = value;
        return GIAOHIP_OK;
    }
    if (option == GIAOHIP_OPTION_CHUNK_MIB && value >= 16 && value <= 65536) {
        ctx->seed_budget_bytes = (size_t)value << 20;
        return GIAOHIP_OK;
    }
    if (option == GIAOHIP_OPTION_CHOL_BLOCK && value >= 1 && value <= gh::kCholMaxBlock) {
        ctx->chol_block = value;
        return GIAOHIP_OK;
    }
    if (option == GIAOHIP_OPTION_CHOL_PIVOT_BATCH && value >= 1) {
        ctx->chol_pivot_batch = value;
        return GIAOHIP_OK;
    }
    if (option == GIAOHIP_OPTION_STAGING_MIB && value >= 1 && value <= 65536) {
        ctx->staging_budget_bytes = (size_t)value << 20;
        return GIAOHIP_OK;
    }
    if (option == GIAOHIP_OPTION_ERI_QUEUE && value >= GIAOHIP_QUEUE_NONE && value <= GIAOHIP_QUEUE_SINGLE) {
        ctx->eri_queue = value;
        return GIAOHIP_OK;
    }
    if (option == GIAOHIP_OPTION_FAST_WAVES &&
        (value == 0 || (value >= gh::kFastBraGroups && value <= gh::kFastMaxWaves))) {
        ctx->fast_waves_cap = value;
        return GIAOHIP_OK;
    }
    if (option == GIAOHIP_OPTION_MFMA && value >= GIAOHIP_MFMA_OFF && value <= GIAOHIP_MFMA_SPLIT32) {
        ctx->mfma_mode = value;
        return GIAOHIP_OK;
    }
    if (option == GIAOHIP_OPTION_SMALL_MAX_L && value >= -1 && value <= gh::kSmallMaxL) {
        ctx->small_max_l = value;
        return GIAOHIP_OK;
    }
    if (option == GIAOHIP_OPTION_MFMA_VARIANT && value >= 0 && value < gh::kMfmaVariantCount) {
        ctx->mfma_variant = value;
        return GIAOHIP_OK;
    }
    if (option == GIAOHIP_OPTION_CHOL_GRAPHS && value >= 0 && value <= 2) {
        ctx->chol_graphs = value;
        return GIAOHIP_OK;
    }
    if (option == GIAOHIP_OPTION_MFMA_MIN_L && value >= 0 && value <= 33) {
        ctx->mfma_min_l = value;
        return GIAOHIP_OK;
    }
    if (option == GIAOHIP_OPTION_WAVE_MAX_L && value >= -1 && value <= gh::kMaxTotalL) {
        ctx->wave_max_l = value;
        return GIAOHIP_OK;
    }
    if (option == GIAOHIP_OPTION_PIPELINE && (value == GIAOHIP_PIPELINE_OVERLAP || value == GIAOHIP_PIPELINE_SERIAL)) {
        ctx->pipeline = value;
        return GIAOHIP_OK;
    }
    return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "unknown option or value");
}

int giaohip_last_stats(const giaohip_ctx* ctx, giaohip_stats* out) {
    if (!ctx || !out) return GIAOHIP_INVALID_ARGUMENT;
    *out = ctx->stats;
    return GIAOHIP_OK;
}

int giaohip_device_alloc(giaohip_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return GIAOHIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    (void)hipSetDevice(ctx->device);
    void* p = nullptr;
    HIP_TRY(counted_malloc(&p, std::max<size_t>(bytes, 16)), "hipMalloc");
    ctx->user_allocs.insert(p);
    *out = p;
    return GIAOHIP_OK;
}

int giaohip_device_free(giaohip_ctx* ctx, void* ptr) {
    if (!ctx) return GIAOHIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!ctx->user_allocs.erase(ptr)) return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "pointer not allocated by this context");
    HIP_TRY(hipFree(ptr), "hipFree");
    return GIAOHIP_OK;
}

int giaohip_memcpy_to_host(giaohip_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return GIAOHIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return GIAOHIP_OK;
}

int giaohip_synchronize(giaohip_ctx* ctx) {
    if (!ctx) return GIAOHIP_INVALID_ARGUMENT;
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return GIAOHIP_OK;
}

}  // extern "C"

namespace {

// Fills b's host tables from the C-ABI shell layout and uploads them. Device
// buffers only grow (DevBuffer::ensure), so refilling a basis with shells that
// fit its earlier tables allocates nothing. On failure b is left empty.
int load_basis(giaohip_ctx* ctx, giaohip_basis* b, int nshell, const double* centers, const int* L,
               const int* nprim, const int* ncontr, const double* exps, const double* coefs) {
    auto clear = [b] {
        b->ns = 0;
        b->nao = 0;
        b->shells.clear();
        b->ao_off.clear();
        b->exps.clear();
        b->coefs.clear();
        b->norms.clear();
        b->radial.clear();
        b->owner.clear();
        release_basis(b->deriv);
        b->deriv = nullptr;
    };
    clear();
    if (nshell < 0) return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "shell count must be non-negative");
    if (nshell > 0 && (!centers || !L || !nprim || !ncontr || !exps || !coefs))
        return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "shell description arrays must not be null");
    int po = 0, co = 0, no = 0;
    long long ao = 0;
    for (int s = 0; s < nshell; ++s) {
        if (L[s] < 0 || nprim[s] < 1 || ncontr[s] < 1) {
            clear();
            return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "invalid shell description");
        }
        DevShell d{};
        for (int i = 0; i < 3; ++i) d.center[i] = centers[3 * s + i];
        d.l = L[s];
        d.nprim = nprim[s];
        d.ncontr = ncontr[s];
        d.ncart = gh::ncart(L[s]);
        d.prim_off = po;
        d.coef_off = co;
        d.norm_off = no;
        d.ao_count = d.ncart * d.ncontr;
        for (int p = 0; p < d.nprim; ++p) {
            const double e = exps[po + p];
            if (!std::isfinite(e) || e <= 0.0) {
                clear();
                return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "shell exponents must be finite and positive");
            }
            b->exps.push_back(e);
            // radial part of primitive_normalization (types.cpp:207-219)
            b->radial.push_back(std::exp(0.75 * std::log(2.0 * e / kPi) + 0.5 * (double)d.l * std::log(4.0 * e)));
        }
        for (int k = 0; k < d.nprim * d.ncontr; ++k) b->coefs.push_back(coefs[co + k]);
        const auto comps = cartesian_components(d.l);
        for (const auto& c : comps)
            for (int p = 0; p < d.nprim; ++p) b->norms.push_back(primitive_normalization(exps[po + p], c[0], c[1], c[2]));
        b->shells.push_back(d);
        b->ao_off.push_back(ao);
        for (int k = 0; k < d.ao_count; ++k) b->owner.push_back(s);
        ao += d.ao_count;
        po += d.nprim;
        co += d.nprim * d.ncontr;
        no += (int)comps.size() * d.nprim;
    }
    b->ns = nshell;
    b->nao = ao;
    build_class_tables(b);
    (void)hipSetDevice(ctx->device);
    auto up = [&](DevBuffer& buf, const void* src, size_t bytes) -> hipError_t {
        hipError_t e = buf.ensure(bytes);
        if (e != hipSuccess) return e;
        if (bytes) e = hipMemcpy(buf.p, src, bytes, hipMemcpyHostToDevice);
        return e;
    };
    hipError_t e = up(b->d_shells, b->shells.data(), b->shells.size() * sizeof(DevShell));
    if (e == hipSuccess) e = up(b->d_exps, b->exps.data(), b->exps.size() * sizeof(double));
    if (e == hipSuccess) e = up(b->d_coefs, b->coefs.data(), b->coefs.size() * sizeof(double));
    if (e == hipSuccess) e = up(b->d_norms, b->norms.data(), b->norms.size() * sizeof(double));
    if (e == hipSuccess) e = up(b->d_aooff, b->ao_off.data(), b->ao_off.size() * sizeof(long long));
    if (e == hipSuccess) e = up(b->d_owner, b->owner.data(), b->owner.size() * sizeof(int));
    if (e == hipSuccess) e = up(b->d_radial, b->radial.data(), b->radial.size() * sizeof(double));
    if (e != hipSuccess) {
        clear();
        return hip_fail(ctx, e, "basis upload");
    }
    return GIAOHIP_OK;
}

// The shifted-momentum companion of b for centre derivatives (built on first
// use, kept with b): shells [0, ns) are b's; s + ns has l + 1 and 2K rows,
// c sqrt(alpha) (the 2 alpha (a + 1| term with the radial normalization
// ratio R_l / R_{l+1} = 1 / (2 sqrt(alpha)) folded in) then
// c / (2 sqrt(alpha)) (the London phase's (a + 1| term); s + 2 ns has the
// first K of those rows only (zero field); s + 3 ns has l - 1 and
// c 2 sqrt(alpha) (R_l / R_{l-1}), or is a copy of s when l = 0 (never used).
int derivative_basis(giaohip_ctx* ctx, const giaohip_basis* b, giaohip_basis** out) {
    auto* mb = const_cast<giaohip_basis*>(b);
    if (mb->deriv) {
        *out = mb->deriv;
        return GIAOHIP_OK;
    }
    const int ns = b->ns;
    std::vector<double> centers, exps, coefs;
    std::vector<int> L, nprim, ncontr;
    for (int part = 0; part < 4; ++part)
        for (int s = 0; s < ns; ++s) {
            const DevShell& d = b->shells[s];
            const int K = d.ncontr, P = d.nprim;
            const double* e = b->exps.data() + d.prim_off;
            const double* c = b->coefs.data() + d.coef_off;
            int l = d.l, rows = K;
            if (part == 1) l = d.l + 1, rows = 2 * K;
            if (part == 2) l = d.l + 1;
            if (part == 3 && d.l > 0) l = d.l - 1;
            for (int i = 0; i < 3; ++i) centers.push_back(d.center[i]);
            L.push_back(l);
            nprim.push_back(P);
            ncontr.push_back(rows);
            exps.insert(exps.end(), e, e + P);
            for (int r = 0; r < rows; ++r)
                for (int p = 0; p < P; ++p) {
                    const double ra = std::sqrt(e[p]);
                    double f = 1.0;
                    if (part == 1 || part == 2) f = r < K ? ra : 1.0 / (2.0 * ra);
                    if (part == 3 && d.l > 0) f = 2.0 * ra;
                    coefs.push_back(c[(size_t)(r % K) * P + p] * f);
                }
        }
    auto* db = new giaohip_basis;
    const int st = load_basis(ctx, db, 4 * ns, centers.data(), L.data(), nprim.data(), ncontr.data(), exps.data(),
                              coefs.data());
    if (st != GIAOHIP_OK) {
        release_basis(db);
        return st;
    }
    mb->deriv = db;
    *out = db;
    return GIAOHIP_OK;
}

}  // namespace

extern "C" {

int giaohip_basis_create(giaohip_ctx* ctx, int nshell, const double* centers, const int* L, const int* nprim,
                         const int* ncontr, const double* exps, const double* coefs, giaohip_basis** out) {
    if (!ctx || !out) return GIAOHIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    *out = nullptr;
    auto* b = new giaohip_basis;
    const int st = load_basis(ctx, b, nshell, centers, L, nprim, ncontr, exps, coefs);
    if (st != GIAOHIP_OK) {
        release_basis(b);
        return st;
    }
    *out = b;
    return GIAOHIP_OK;
}

int giaohip_basis_update(giaohip_ctx* ctx, giaohip_basis* b, int nshell, const double* centers, const int* L,
                         const int* nprim, const int* ncontr, const double* exps, const double* coefs) {
    if (!ctx || !b) return GIAOHIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    (void)hipSetDevice(ctx->device);
    // the previous call's kernels may still read the tables being replaced
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    HIP_TRY(hipStreamSynchronize(ctx->aux), "hipStreamSynchronize");
    return load_basis(ctx, b, nshell, centers, L, nprim, ncontr, exps, coefs);
}

int giaohip_basis_destroy(giaohip_ctx* ctx, giaohip_basis* b) {
    if (!b) return GIAOHIP_OK;
    if (ctx) {
        std::lock_guard<std::mutex> lock(ctx->mu);
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
    }
    release_basis(b);
    return GIAOHIP_OK;
}

size_t giaohip_basis_ao_count(const giaohip_basis* b) { return b ? (size_t)b->nao : 0; }

int giaohip_eri_batch_size(const giaohip_basis* b, const uint32_t* quartets, size_t n, uint64_t* total) {
    if (!b || !total) return GIAOHIP_INVALID_ARGUMENT;
    uint64_t t = 0;
    for (size_t i = 0; i < n; ++i) {
        uint64_t sz = 1;
        for (int k = 0; k < 4; ++k) {
            if (quartets[4 * i + k] >= (uint32_t)b->ns) return GIAOHIP_OUT_OF_RANGE;
            sz *= (uint64_t)b->shells[quartets[4 * i + k]].ao_count;
        }
        t += sz;
    }
    *total = t;
    return GIAOHIP_OK;
}

int giaohip_boys(giaohip_ctx* ctx, const double* z, size_t n, int nmax, int scaled, double* out, int* region,
                 double* error_estimate, int* segments) {
    if (!ctx || (n > 0 && (!z || !out))) return GIAOHIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (nmax < 0) return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "Boys output must contain at least F_0");
    if (nmax > gh::kMaxBoysOrder) return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "complex Boys supports orders through 32");
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(z[2 * i]) || !std::isfinite(z[2 * i + 1]))
            return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "complex Boys argument must be finite");
    if (n == 0) return GIAOHIP_OK;
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    const size_t zb = n * sizeof(zc), ob = n * (size_t)(nmax + 1) * sizeof(zc);
    HIP_TRY(ctx->misc_a.ensure(zb + ob), "hipMalloc");
    HIP_TRY(ctx->misc_b.ensure(n * (2 * sizeof(int) + sizeof(double))), "hipMalloc");
    HIP_TRY(ctx->errrec.ensure(sizeof(DevError)), "hipMalloc");
    char* a = ctx->misc_a.as<char>();
    char* bb = ctx->misc_b.as<char>();
    int* d_region = reinterpret_cast<int*>(bb);
    int* d_seg = d_region + n;
    double* d_err = reinterpret_cast<double*>(bb + 2 * n * sizeof(int));
    HIP_TRY(hipMemsetAsync(ctx->errrec.p, 0, sizeof(DevError), s), "hipMemsetAsync");
    HIP_TRY(hipMemcpyAsync(a, z, zb, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    HIP_TRY(gh::launch_boys_probe(reinterpret_cast<zc*>(a), (long long)n, nmax, scaled, reinterpret_cast<zc*>(a + zb),
                                  d_region, d_err, d_seg, ctx->errrec.as<DevError>(), s),
            "Boys kernel launch");
    HIP_TRY(hipMemcpyAsync(out, a + zb, ob, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    std::vector<int> reg(n), seg(n);
    std::vector<double> er(n);
    HIP_TRY(hipMemcpyAsync(reg.data(), d_region, n * sizeof(int), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    HIP_TRY(hipMemcpyAsync(seg.data(), d_seg, n * sizeof(int), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    HIP_TRY(hipMemcpyAsync(er.data(), d_err, n * sizeof(double), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    DevError de{};
    HIP_TRY(hipMemcpyAsync(&de, ctx->errrec.p, sizeof de, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    HIP_TRY(hipStreamSynchronize(s), "Boys kernel execution");
    if (de.code != 0) return boys_error(ctx, de);
    if (region) std::copy(reg.begin(), reg.end(), region);
    if (segments) std::copy(seg.begin(), seg.end(), segments);
    if (error_estimate) std::copy(er.begin(), er.end(), error_estimate);
    return GIAOHIP_OK;
}

// The checks of evaluate_eri_shell_quartets (eri.cpp:644-688) on a quartet
// list, in parallel (the first failing quartet decides the error), and the
// size of each block in complex elements.
static int check_quartets(giaohip_ctx* ctx, const giaohip_basis* b, const uint32_t* quartets, size_t n,
                   const double field[6], std::vector<uint64_t>& sizes) {
    for (int k = 0; k < 6; ++k)
        if (!std::isfinite(field[k])) return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "field values must be finite");
    // validation and block sizes in parallel; the first failing quartet decides the error
    std::vector<size_t> first_bad(16, n);
    std::vector<int> bad_status(16, 0);
    std::atomic<size_t> slot{0};
    parallel_for(n, [&](size_t lo, size_t hi) {
        const size_t me = slot.fetch_add(1);
        for (size_t i = lo; i < hi; ++i) {
            int sh[4];
            bool in_range = true;
            for (int k = 0; k < 4; ++k) {
                sh[k] = (int)quartets[4 * i + k];
                in_range = in_range && quartets[4 * i + k] < (uint32_t)b->ns;
            }
            const int st = in_range ? b->cls_status[class_index(b, sh)] : -1;
            if (st != 0) {
                first_bad[me] = i;
                bad_status[me] = st;
                return;
            }
            sizes[i] = (uint64_t)b->shells[sh[0]].ao_count * b->shells[sh[1]].ao_count * b->shells[sh[2]].ao_count *
                       b->shells[sh[3]].ao_count;
        }
    });
    size_t bad = n;
    int bad_st = 0;
    for (size_t t = 0; t < first_bad.size(); ++t)
        if (first_bad[t] < bad) {
            bad = first_bad[t];
            bad_st = bad_status[t];
        }
    if (bad < n) {
        if (bad_st < 0) return set_error(ctx, GIAOHIP_OUT_OF_RANGE, "ERI shell-quartet index is out of range");
        return class_error(ctx, bad_st);
    }
    return GIAOHIP_OK;
}

// Request of quartet i of a list, written at complex offset `base` with the
// C-order strides of its (a, b, c, d) block.
static Request quartet_request(const giaohip_basis* b, const uint32_t* quartets, size_t i, long long base) {
    Request r{};
    for (int k = 0; k < 4; ++k) {
        r.sh[k] = (int)quartets[4 * i + k];
        r.xfix[k] = -1;
    }
    r.cd0 = 0;
    r.cd1 = -1;
    r.out_base = base;
    const long long nb = b->shells[r.sh[1]].ao_count, nc = b->shells[r.sh[2]].ao_count,
                    nd = b->shells[r.sh[3]].ao_count;
    r.stride[0] = nb * nc * nd;
    r.stride[1] = nc * nd;
    r.stride[2] = nd;
    r.stride[3] = 1;
    return r;
}

// Device -> pageable host copy through two page-locked bounce buffers: the
// DMA of piece k + 1 runs while piece k is copied out by host threads.
// Returns a HIP error (hipSuccess when done).
hipError_t copy_to_host(giaohip_ctx* ctx, const char* src, char* dst, size_t bytes) {
    constexpr size_t kPiece = size_t(64) << 20;
    for (PinnedBuffer& b : ctx->bounce)
        if (hipError_t e = b.ensure(kPiece); e != hipSuccess) return e;
    hipStream_t cs = ctx->copy_stream;
    const size_t pieces = (bytes + kPiece - 1) / kPiece;
    auto drain = [&](size_t k) -> hipError_t {
        if (hipError_t e = hipEventSynchronize(ctx->copy_ev[k % 2]); e != hipSuccess) return e;
        const size_t off = k * kPiece, len = std::min(kPiece, bytes - off);
        const char* from = ctx->bounce[k % 2].as<char>();
        parallel_for(len / 4096 + 1, [&](size_t lo, size_t hi) {
            const size_t a = std::min(len, lo * 4096), z = std::min(len, hi * 4096);
            if (z > a) std::memcpy(dst + off + a, from + a, z - a);
        });
        return hipSuccess;
    };
    for (size_t k = 0; k < pieces; ++k) {
        const size_t off = k * kPiece, len = std::min(kPiece, bytes - off);
        if (hipError_t e = hipMemcpyAsync(ctx->bounce[k % 2].p, src + off, len, hipMemcpyDeviceToHost, cs);
            e != hipSuccess)
            return e;
        if (hipError_t e = hipEventRecord(ctx->copy_ev[k % 2], cs); e != hipSuccess) return e;
        if (k > 0)
            if (hipError_t e = drain(k - 1); e != hipSuccess) return e;
    }
    return pieces ? drain(pieces - 1) : hipSuccess;
}

int giaohip_eri_batch(giaohip_ctx* ctx, const giaohip_basis* b, const uint32_t* quartets, size_t n,
                      const double field[6], const uint64_t* offsets, double* out, int out_on_device) {
    if (!ctx || !b || !field || (n > 0 && (!quartets || !out))) return GIAOHIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    reset_stats(ctx);
    ctx->stats.requested_quartets = n;
    const auto host_t0 = std::chrono::steady_clock::now();
    std::vector<uint64_t> offs(n), sizes(n);
    if (const int vst = check_quartets(ctx, b, quartets, n, field, sizes)) return vst;
    uint64_t run = 0;
    for (size_t i = 0; i < n; ++i) {
        offs[i] = offsets ? offsets[i] : run;
        run += sizes[i];
    }
    ctx->stats.computed_quartets = n;
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    HIP_TRY(hipEventRecord(ctx->ev[3], s), "hipEventRecord");
    hipEvent_t t0;
    HIP_TRY(hipEventCreate(&t0), "hipEventCreate");
    HIP_TRY(hipEventRecord(t0, s), "hipEventRecord");
    auto make_req = [&](size_t i, long long base) { return quartet_request(b, quartets, i, base); };
    int st = GIAOHIP_OK;
    if (out_on_device) {
        BigVec<Request> reqs(n);
        parallel_for(n, [&](size_t lo, size_t hi) {
            for (size_t i = lo; i < hi; ++i) reqs[i] = make_req(i, (long long)offs[i]);
        });
        st = run_requests(ctx, b, field, reqs, reinterpret_cast<zc*>(out));
    } else {
        // Stage groups of quartets through two device buffers: a group is
        // computed while a copier thread moves the previous one to the host.
        const size_t budget = ctx->staging_budget_bytes / sizeof(zc);
        DevBuffer* sbuf[2] = {&ctx->staging, &ctx->staging2};
        std::thread copier;
        hipError_t copy_err = hipSuccess;
        auto join_copier = [&] {
            if (copier.joinable()) copier.join();
            if (copy_err != hipSuccess && st == GIAOHIP_OK) st = hip_fail(ctx, copy_err, "copy to host");
        };
        size_t i = 0;
        for (int g = 0; i < n && st == GIAOHIP_OK; ++g) {
            size_t j = i;
            uint64_t tot = 0;
            while (j < n && (j == i || tot + sizes[j] <= budget)) tot += sizes[j++];
            DevBuffer& buf = *sbuf[g % 2];
            if (hipError_t e = buf.ensure(tot * sizeof(zc)); e != hipSuccess) {
                st = hip_fail(ctx, e, "hipMalloc(staging)");
                break;
            }
            std::vector<Request> reqs;
            uint64_t local = 0;
            for (size_t k = i; k < j; ++k) {
                reqs.push_back(make_req(k, (long long)local));
                local += sizes[k];
            }
            st = run_requests(ctx, b, field, reqs, buf.as<zc>());
            join_copier();  // the previous group is on the host; its buffer is free again
            if (st) break;
            bool packed = true;
            for (size_t k = i + 1; k < j; ++k) packed = packed && offs[k] == offs[k - 1] + sizes[k - 1];
            const char* src = static_cast<const char*>(buf.p);
            if (packed) {
                char* dst = reinterpret_cast<char*>(out + 2 * offs[i]);
                const size_t bytes = tot * sizeof(zc);
                copier = std::thread([&ctx, &copy_err, src, dst, bytes] {
                    (void)hipSetDevice(ctx->device);
                    copy_err = copy_to_host(ctx, src, dst, bytes);
                });
            } else {
                std::vector<zc> host_tmp(tot);
                if (hipError_t e = copy_to_host(ctx, src, reinterpret_cast<char*>(host_tmp.data()), tot * sizeof(zc));
                    e != hipSuccess) {
                    st = hip_fail(ctx, e, "copy to host");
                    break;
                }
                uint64_t lo = 0;
                for (size_t k = i; k < j; ++k) {
                    std::memcpy(out + 2 * offs[k], host_tmp.data() + lo, sizes[k] * sizeof(zc));
                    lo += sizes[k];
                }
            }
            i = j;
        }
        join_copier();
    }
    hipEvent_t t1;
    (void)hipEventCreate(&t1);
    (void)hipEventRecord(t1, s);
    (void)hipEventSynchronize(t1);
    ctx->stats.total_ms = elapsed(t0, t1);
    ctx->stats.host_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - host_t0).count();
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return st;
}

int giaohip_eri_batch_noalloc(giaohip_ctx* ctx, const giaohip_basis* b, const uint32_t* quartets, size_t n,
                              const double field[6], const uint64_t* offsets, double* out_device) {
    if (!ctx || !b || !field || (n > 0 && (!quartets || !out_device))) return GIAOHIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    reset_stats(ctx);
    ctx->stats.requested_quartets = n;
    if (!ctx->noalloc) return set_error(ctx, GIAOHIP_DEVICE_ERROR, "allocation-free ERI scratch is not available");
    std::vector<uint64_t> sizes(n);
    if (const int vst = check_quartets(ctx, b, quartets, n, field, sizes)) return vst;
    ctx->stats.computed_quartets = n;
    if (n == 0) return GIAOHIP_OK;
    (void)hipSetDevice(ctx->device);
    char* base = static_cast<char*>(ctx->noalloc);
    FixedPlainMem mem{};
    mem.kappa = reinterpret_cast<double*>(base + kNaKappa);
    mem.kappa_cap = kNaKappaBytes / sizeof(double);
    mem.err = reinterpret_cast<DevError*>(base + kNaErr);
    mem.queue = reinterpret_cast<int*>(base + kNaQueue);
    mem.sq = reinterpret_cast<SeedQuartet*>(base + kNaSq);
    mem.sq_cap = kNaSqBytes / sizeof(SeedQuartet);
    mem.items = reinterpret_cast<WorkItem*>(base + kNaItems);
    mem.items_cap = kNaItemsBytes / sizeof(WorkItem);
    mem.seeds = reinterpret_cast<zc*>(base + kNaSeeds);
    mem.seeds_cap = (kNoAllocScratchBytes - kNaSeeds) / sizeof(zc);
    std::vector<Request> reqs(n);
    uint64_t run = 0;
    for (size_t i = 0; i < n; ++i) {
        reqs[i] = quartet_request(b, quartets, i, (long long)(offsets ? offsets[i] : run));
        run += sizes[i];
    }
    return run_plain(ctx, b, field, reqs, reinterpret_cast<zc*>(out_device), &mem);
}

// compute_eri_derivative_blocks (derivatives.cpp:1389-1415) on the device:
// the shifted-momentum quartets of each derivative quartet through the ERI
// path on the companion basis, then the combination kernel (gh_deriv.hip).
int giaohip_eri_derivative_batch(giaohip_ctx* ctx, const giaohip_basis* b, const uint32_t* quartets, size_t n,
                                 const double field[6], const uint64_t* offsets, double* out, uint64_t out_len,
                                 int out_on_device) {
    if (!ctx || !b || !field || (n > 0 && (!quartets || !out))) return GIAOHIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    reset_stats(ctx);
    ctx->stats.requested_quartets = n;
    for (int k = 0; k < 6; ++k)
        if (!std::isfinite(field[k])) return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "field values must be finite");
    if (offsets && offsets[0] != 0)
        return set_error(ctx, GIAOHIP_LENGTH_ERROR, "derivative batch offsets have an incorrect size or start");
    std::vector<uint64_t> sizes(n), offs(n + 1, 0);
    for (size_t i = 0; i < n; ++i) {
        uint64_t size = 1;
        for (int k = 0; k < 4; ++k) {
            if (quartets[4 * i + k] >= (uint32_t)b->ns)
                return set_error(ctx, GIAOHIP_OUT_OF_RANGE, "derivative shell quartet index is out of range");
            size *= (uint64_t)b->shells[quartets[4 * i + k]].ao_count;
        }
        sizes[i] = size;
        if (offsets) {
            if (offsets[i] > out_len || offsets[i + 1] < offsets[i] || offsets[i + 1] - offsets[i] != 12 * size)
                return set_error(ctx, GIAOHIP_LENGTH_ERROR, "derivative batch offsets do not match block sizes");
            offs[i + 1] = offsets[i + 1];
        } else {
            offs[i + 1] = offs[i] + 12 * size;
        }
    }
    if (offs[n] != out_len) return set_error(ctx, GIAOHIP_LENGTH_ERROR, "derivative batch output has an incorrect size");
    if (n == 0) return GIAOHIP_OK;
    giaohip_basis* db = nullptr;
    if (const int st = derivative_basis(ctx, b, &db)) return st;
    const int ns = b->ns;
    const bool phase = field[0] != 0.0 || field[1] != 0.0 || field[2] != 0.0;
    const int plus_part = phase ? 1 : 2;
    // Every shifted class must be admitted (the reference evaluates them
    // all). The reference's own limits (combined order, auxiliary size) are
    // reported before the device kernels' limits.
    for (int pass = 0; pass < 2; ++pass)
        for (size_t i = 0; i < n; ++i) {
            const uint32_t* q = quartets + 4 * i;
            const int sh[4] = {(int)q[0], (int)q[1], (int)q[2], (int)q[3]};
            auto check = [&](const int* t) {
                const int cs = db->cls_status[class_index(db, t)];
                const bool reference_limit = cs == kClsOrder || cs == kClsAux;
                return (pass == 0) == reference_limit ? class_error(ctx, cs) : GIAOHIP_OK;
            };
            if (const int st = check(sh)) return st;
            for (int k = 0; k < 4; ++k) {
                int t[4] = {sh[0], sh[1], sh[2], sh[3]};
                t[k] = sh[k] + plus_part * ns;
                if (const int st = check(t)) return st;
                if (b->shells[sh[k]].l > 0) {
                    t[k] = sh[k] + 3 * ns;
                    if (const int st = check(t)) return st;
                }
            }
        }
    ctx->stats.computed_quartets = n;
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0), "hipEventCreate");
    if (hipError_t e = hipEventCreate(&t1); e != hipSuccess) {
        (void)hipEventDestroy(t0);
        return hip_fail(ctx, e, "hipEventCreate");
    }
    (void)hipEventRecord(t0, s);
    // work sizes of one derivative quartet: the shifted integrals and the output
    auto temp_size = [&](size_t i) {
        uint64_t t = sizes[i];
        for (int k = 0; k < 4; ++k) {
            const DevShell& d = b->shells[quartets[4 * i + k]];
            const uint64_t rest = sizes[i] / (uint64_t)d.ao_count;
            t += rest * (uint64_t)db->shells[quartets[4 * i + k] + plus_part * ns].ao_count;
            if (d.l > 0) t += rest * (uint64_t)db->shells[quartets[4 * i + k] + 3 * ns].ao_count;
        }
        return t;
    };
    const uint64_t budget = std::max<uint64_t>(ctx->staging_budget_bytes / sizeof(zc), 1);
    int st = GIAOHIP_OK;
    std::vector<Request> reqs;
    std::vector<DerivItem> items;
    std::vector<int> tiles;
    size_t i = 0;
#define DERIV_TRY(expr, what)                  \
    if (hipError_t _e = (expr); _e != hipSuccess) { \
        st = hip_fail(ctx, _e, what);          \
        break;                                 \
    }
    while (i < n && st == GIAOHIP_OK) {
        size_t j = i;
        uint64_t tt = 0, to = 0;
        while (j < n && (j == i || (tt + temp_size(j) <= budget && to + 12 * sizes[j] <= budget))) {
            tt += temp_size(j);
            to += 12 * sizes[j];
            ++j;
        }
        DERIV_TRY(ctx->blocks.ensure(tt * sizeof(zc)), "hipMalloc(blocks)");
        zc* temp = ctx->blocks.as<zc>();
        zc* dst;
        if (out_on_device) {
            dst = reinterpret_cast<zc*>(out) + offs[i];
        } else {
            DERIV_TRY(ctx->staging.ensure(to * sizeof(zc)), "hipMalloc(staging)");
            dst = ctx->staging.as<zc>();
        }
        reqs.clear();
        items.clear();
        tiles.clear();
        long long tpos = 0;
        auto add_request = [&](const int sh[4], long long base) {
            Request r{};
            for (int k = 0; k < 4; ++k) {
                r.sh[k] = sh[k];
                r.xfix[k] = -1;
            }
            r.cd0 = 0;
            r.cd1 = -1;
            r.out_base = base;
            const long long nb = db->shells[sh[1]].ao_count, nc = db->shells[sh[2]].ao_count,
                            nd = db->shells[sh[3]].ao_count;
            r.stride[0] = nb * nc * nd;
            r.stride[1] = nc * nd;
            r.stride[2] = nd;
            r.stride[3] = 1;
            reqs.push_back(r);
            return (long long)db->shells[sh[0]].ao_count * nb * nc * nd;
        };
        for (size_t k = i; k < j; ++k) {
            const uint32_t* q = quartets + 4 * k;
            const int sh[4] = {(int)q[0], (int)q[1], (int)q[2], (int)q[3]};
            DerivItem it{};
            it.out = (long long)(offs[k] - offs[i]);
            it.size = (long long)sizes[k];
            it.base = tpos;
            tpos += add_request(sh, tpos);
            for (int m = 0; m < 4; ++m) {
                const DevShell& d = b->shells[sh[m]];
                it.n[m] = d.ao_count;
                it.l[m] = d.l;
                it.krow[m] = d.ncontr;
                it.prow[m] = phase ? 2 * d.ncontr : d.ncontr;
                for (int x = 0; x < 3; ++x) it.center[m][x] = d.center[x];
                int t[4] = {sh[0], sh[1], sh[2], sh[3]};
                t[m] = sh[m] + plus_part * ns;
                it.plus[m] = tpos;
                tpos += add_request(t, tpos);
                it.minus[m] = tpos;
                if (d.l > 0) {
                    t[m] = sh[m] + 3 * ns;
                    tpos += add_request(t, tpos);
                }
            }
            it.first_tile = (int)tiles.size();
            const long long nt = (it.size + 255) / 256;
            for (long long x = 0; x < nt; ++x) tiles.push_back((int)items.size());
            items.push_back(it);
        }
        st = run_requests(ctx, db, field, reqs, temp);
        if (st) break;
        const size_t ib = items.size() * sizeof(DerivItem);
        DERIV_TRY(ctx->misc_b.ensure(ib + tiles.size() * sizeof(int) + 64), "hipMalloc");
        auto* d_items = ctx->misc_b.as<DerivItem>();
        int* d_tiles = reinterpret_cast<int*>(ctx->misc_b.as<char>() + ib);
        DERIV_TRY(hipMemcpyAsync(d_items, items.data(), ib, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
        DERIV_TRY(hipMemcpyAsync(d_tiles, tiles.data(), tiles.size() * sizeof(int), hipMemcpyHostToDevice, s),
                "hipMemcpyAsync");
        DERIV_TRY(launch_deriv_assemble(d_items, d_tiles, (int)tiles.size(), temp, dst, field, phase, s),
                "deriv_assemble_kernel");
        // the descriptors and the work buffers are reused by the next group
        DERIV_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
        if (!out_on_device)
            DERIV_TRY(copy_to_host(ctx, reinterpret_cast<const char*>(dst), reinterpret_cast<char*>(out + 2 * offs[i]),
                                 to * sizeof(zc)),
                    "copy to host");
        i = j;
    }
#undef DERIV_TRY
    (void)hipEventRecord(t1, s);
    (void)hipEventSynchronize(t1);
    ctx->stats.total_ms = elapsed(t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return st;
}

// EriDerivativeSchwarzBounds values (derivatives.cpp:1307-1372). The
// primitive Coulomb norms sqrt|(right left|left right)| of the shifted raw
// primitives are diagonal elements of ERI blocks of single-primitive shells
// (l, l + 1, l - 1 at each primitive), evaluated on the device; the device
// normalizes the shifted component, so each norm is rescaled by
// N(alpha, c) / N(alpha, c'), then combined as the reference does.
int giaohip_eri_derivative_pair_bounds(giaohip_ctx* ctx, const giaohip_basis* b, const double field[6], double* out) {
    if (!ctx || !b || !field || !out) return GIAOHIP_INVALID_ARGUMENT;
    for (int k = 0; k < 6; ++k)
        if (!std::isfinite(field[k])) return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "field values must be finite");
    const int ns = b->ns;
    if (ns == 0) return GIAOHIP_OK;
    std::vector<int> pstart(ns + 1, 0);
    for (int s = 0; s < ns; ++s) pstart[s + 1] = pstart[s] + b->shells[s].nprim;
    const int np = pstart[ns];
    std::vector<double> centers, exps, coefs(3 * (size_t)np, 1.0);
    std::vector<int> L, nprim(3 * (size_t)np, 1), ncontr(3 * (size_t)np, 1);
    for (int s = 0; s < ns; ++s) {
        const DevShell& d = b->shells[s];
        for (int p = 0; p < d.nprim; ++p)
            for (int v = 0; v < 3; ++v) {
                for (int i = 0; i < 3; ++i) centers.push_back(d.center[i]);
                exps.push_back(b->exps[d.prim_off + p]);
                L.push_back(v == 0 ? d.l : (v == 1 ? d.l + 1 : std::max(d.l - 1, 0)));
            }
    }
    giaohip_basis* pb = nullptr;
    if (const int st = giaohip_basis_create(ctx, 3 * np, centers.data(), L.data(), nprim.data(), ncontr.data(),
                                            exps.data(), coefs.data(), &pb))
        return st;
    struct Guard {
        giaohip_ctx* c;
        giaohip_basis* p;
        ~Guard() { giaohip_basis_destroy(c, p); }
    } guard{ctx, pb};
    // quartets of one shell pair a <= b and primitive pair: (B A|A B) with
    // A in {l, l+1, l-1} at fixed B, then B in {l+1, l-1} at fixed A
    auto quartets_of = [&](int a, int c, std::vector<uint32_t>& q) {
        const int la = b->shells[a].l, lc = b->shells[c].l;
        for (int pa = 0; pa < b->shells[a].nprim; ++pa)
            for (int pc = 0; pc < b->shells[c].nprim; ++pc) {
                const uint32_t A = 3 * (uint32_t)(pstart[a] + pa), C = 3 * (uint32_t)(pstart[c] + pc);
                q.insert(q.end(), {C, A, A, C, C, A + 1, A + 1, C});
                if (la > 0) q.insert(q.end(), {C, A + 2, A + 2, C});
                q.insert(q.end(), {C + 1, A, A, C + 1});
                if (lc > 0) q.insert(q.end(), {C + 2, A, A, C + 2});
            }
    };
    std::vector<std::pair<int, int>> pairs;
    for (int a = 0; a < ns; ++a)
        for (int c = a; c < ns; ++c) pairs.emplace_back(a, c);
    const uint64_t budget = std::max<uint64_t>(ctx->staging_budget_bytes / sizeof(zc), 1);
    size_t i = 0;
    std::vector<double> blocks;
    while (i < pairs.size()) {
        // a group of shell pairs within the staging budget
        size_t j = i;
        std::vector<uint32_t> q;
        std::vector<size_t> qfirst;
        uint64_t tot = 0;
        while (j < pairs.size()) {
            std::vector<uint32_t> mine;
            quartets_of(pairs[j].first, pairs[j].second, mine);
            uint64_t sz = 0;
            for (size_t k = 0; k < mine.size() / 4; ++k) {
                uint64_t e = 1;
                for (int m = 0; m < 4; ++m) e *= (uint64_t)pb->shells[mine[4 * k + m]].ao_count;
                sz += e;
            }
            if (j > i && tot + sz > budget) break;
            qfirst.push_back(q.size() / 4);
            q.insert(q.end(), mine.begin(), mine.end());
            tot += sz;
            ++j;
        }
        const size_t nq = q.size() / 4;
        std::vector<uint64_t> offs(nq + 1, 0);
        for (size_t k = 0; k < nq; ++k) {
            uint64_t e = 1;
            for (int m = 0; m < 4; ++m) e *= (uint64_t)pb->shells[q[4 * k + m]].ao_count;
            offs[k + 1] = offs[k] + e;
        }
        blocks.resize(2 * std::max<uint64_t>(offs[nq], 1));
        if (const int st = giaohip_eri_batch(ctx, pb, q.data(), nq, field, offs.data(), blocks.data(), 0)) return st;
        parallel_for(j - i, [&](size_t lo, size_t hi) {
            for (size_t t = lo; t < hi; ++t) {
                const int a = pairs[i + t].first, c = pairs[i + t].second;
                const DevShell &sa = b->shells[a], &sc = b->shells[c];
                const auto ca_list = cartesian_components(sa.l), cc_list = cartesian_components(sc.l);
                const int Ka = sa.ncontr, Kc = sc.ncontr, Pa = sa.nprim, Pc = sc.nprim;
                std::vector<double> contracted(6 * (size_t)Ka * Kc);
                double maximum = 0.0;
                for (size_t ia = 0; ia < ca_list.size(); ++ia)
                    for (size_t ic = 0; ic < cc_list.size(); ++ic) {
                        std::fill(contracted.begin(), contracted.end(), 0.0);
                        size_t k = qfirst[t];
                        for (int pa = 0; pa < Pa; ++pa)
                            for (int pc = 0; pc < Pc; ++pc) {
                                const double alpha = b->exps[sa.prim_off + pa], gamma = b->exps[sc.prim_off + pc];
                                // the primitive pair's blocks: k0 (B A|A B), k1 A+1, k2 A-1, k3 B+1, k4 B-1
                                const size_t k0 = k++, k1 = k++, k2 = sa.l > 0 ? k++ : k0, k3 = k++,
                                             k4 = sc.l > 0 ? k++ : k0;
                                // sqrt|(C x|x C)| of block kb at AO pair (cc_i, ca_i) with
                                // the unshifted normalization restored on the shifted side
                                auto norm = [&](size_t kb, int na, int nc, int ai, int ci, double ratio) {
                                    const uint64_t e = (((uint64_t)ci * na + ai) * na + ai) * nc + ci;
                                    const double* v = blocks.data() + 2 * (offs[kb] + e);
                                    return std::sqrt(std::hypot(v[0], v[1])) * ratio;
                                };
                                const auto& pa3 = ca_list[ia];
                                const auto& pc3 = cc_list[ic];
                                const int na0 = (int)ca_list.size(), nc0 = (int)cc_list.size();
                                const double same = norm(k0, na0, nc0, (int)ia, (int)ic, 1.0);
                                double pbnd[6];
                                for (int slot = 0; slot < 2; ++slot) {
                                    const auto& raw = slot == 0 ? pa3 : pc3;
                                    const double ex = slot == 0 ? alpha : gamma;
                                    const int l = slot == 0 ? sa.l : sc.l;
                                    const double* cen = slot == 0 ? sa.center : sc.center;
                                    const double n0 = primitive_normalization(ex, raw[0], raw[1], raw[2]);
                                    // norm with the slot's function at powers w (|w| = l +- 1)
                                    auto changed = [&](std::array<int, 3> w, bool up) {
                                        const int lw = l + (up ? 1 : -1);
                                        const int idx = (lw - w[0]) * (lw - w[0] + 1) / 2 + w[2];
                                        const int nw = (lw + 1) * (lw + 2) / 2;
                                        const double ratio = n0 / primitive_normalization(ex, w[0], w[1], w[2]);
                                        if (slot == 0) return norm(up ? k1 : k2, nw, nc0, idx, (int)ic, ratio);
                                        return norm(up ? k3 : k4, na0, nw, (int)ia, idx, ratio);
                                    };
                                    for (int axis = 0; axis < 3; ++axis) {
                                        std::array<int, 3> w = raw;
                                        w[axis] += 1;
                                        double bound = 2.0 * ex * changed(w, true);
                                        if (raw[axis] != 0) {
                                            w = raw;
                                            w[axis] -= 1;
                                            bound += (double)raw[axis] * changed(w, false);
                                        }
                                        const double e3[3] = {axis == 0 ? 1.0 : 0.0, axis == 1 ? 1.0 : 0.0,
                                                              axis == 2 ? 1.0 : 0.0};
                                        const double phase[3] = {field[1] * e3[2] - field[2] * e3[1],
                                                                 field[2] * e3[0] - field[0] * e3[2],
                                                                 field[0] * e3[1] - field[1] * e3[0]};
                                        for (int y = 0; y < 3; ++y) {
                                            const double weight = 0.5 * std::abs(phase[y]);
                                            if (weight == 0.0) continue;
                                            w = raw;
                                            w[y] += 1;
                                            bound += weight * (changed(w, true) + std::abs(cen[y]) * same);
                                        }
                                        pbnd[slot * 3 + axis] = bound;
                                    }
                                }
                                for (int xa = 0; xa < Ka; ++xa)
                                    for (int xc = 0; xc < Kc; ++xc) {
                                        const double weight = std::abs(b->coefs[sa.coef_off + (size_t)xa * Pa + pa] *
                                                                       b->coefs[sc.coef_off + (size_t)xc * Pc + pc]);
                                        for (int m = 0; m < 6; ++m)
                                            contracted[((size_t)xa * Kc + xc) * 6 + m] += weight * pbnd[m];
                                    }
                            }
                        for (double v : contracted) maximum = std::max(maximum, v);
                    }
                const double bound = std::nextafter(maximum * (1.0 + 128.0 * 2.220446049250313e-16),
                                                    std::numeric_limits<double>::infinity());
                out[(size_t)a * ns + c] = bound;
                out[(size_t)c * ns + a] = bound;
            }
        });
        i = j;
    }
    return GIAOHIP_OK;
}

int giaohip_schwarz_bounds(giaohip_ctx* ctx, const giaohip_basis* b, const double field[6], double* out) {
    if (!ctx || !b || !out || !field) return GIAOHIP_INVALID_ARGUMENT;
    const int ns = b->ns;
    std::vector<uint32_t> q;
    q.reserve((size_t)ns * (ns + 1) * 2);
    for (int a = 0; a < ns; ++a)
        for (int c = a; c < ns; ++c) q.insert(q.end(), {(uint32_t)a, (uint32_t)c, (uint32_t)c, (uint32_t)a});
    const size_t nq = q.size() / 4;
    std::vector<uint64_t> offs(nq + 1, 0);
    for (size_t i = 0; i < nq; ++i) {
        const uint64_t na = b->shells[q[4 * i]].ao_count, nb = b->shells[q[4 * i + 1]].ao_count;
        offs[i + 1] = offs[i] + na * nb * nb * na;
    }
    std::vector<double> blocks(2 * std::max<uint64_t>(offs[nq], 1));
    const int st = giaohip_eri_batch(ctx, b, q.data(), nq, field, offs.data(), blocks.data(), 0);
    if (st != GIAOHIP_OK) return st;
    constexpr double safety = 1.0 + 64.0 * 2.220446049250313e-16;
    parallel_for(nq, [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            const int a = (int)q[4 * i], c = (int)q[4 * i + 1];
            const uint64_t na = b->shells[a].ao_count, nb = b->shells[c].ao_count;
            const double* blk = blocks.data() + 2 * offs[i];
            double m = 0.0;
            for (uint64_t ia = 0; ia < na; ++ia)
                for (uint64_t ib = 0; ib < nb; ++ib) {
                    const uint64_t e = ((ia * nb + ib) * nb + ib) * na + ia;
                    m = std::max(m, std::hypot(blk[2 * e], blk[2 * e + 1]));
                }
            const double bound = std::nextafter(std::sqrt(m) * safety, std::numeric_limits<double>::infinity());
            out[(size_t)a * ns + c] = bound;
            out[(size_t)c * ns + a] = bound;
        }
    });
    return GIAOHIP_OK;
}

int giaohip_screen_quartets(giaohip_ctx* ctx, const giaohip_basis* b, const uint32_t* quartets, size_t n,
                            const double* bounds, double threshold, uint32_t* kept, size_t* nkept) {
    if (!ctx || !b || !nkept || (n && (!quartets || !kept))) return GIAOHIP_INVALID_ARGUMENT;
    if (!std::isfinite(threshold) || threshold < 0.0)
        return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "ERI screening threshold must be finite and non-negative");
    if (threshold > 0.0 && !bounds)
        return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "positive ERI screening requires Schwarz bounds");
    const uint32_t ns = (uint32_t)b->ns;
    for (size_t i = 0; i < 4 * n; ++i)
        if (quartets[i] >= ns) return set_error(ctx, GIAOHIP_OUT_OF_RANGE, "ERI shell-quartet index is out of range");
    size_t m = 0;
    for (size_t i = 0; i < n; ++i) {
        const uint32_t* x = quartets + 4 * i;
        if (threshold > 0.0 && bounds[(size_t)x[0] * ns + x[1]] * bounds[(size_t)x[2] * ns + x[3]] < threshold)
            continue;
        if (kept + 4 * m != x) std::copy(x, x + 4, kept + 4 * m);
        ++m;
    }
    *nkept = m;
    return GIAOHIP_OK;
}

int giaohip_eri_tensor(giaohip_ctx* ctx, const giaohip_basis* b, const double field[6], double* out, int out_on_device) {
    if (!ctx || !b || !field || !out) return GIAOHIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    reset_stats(ctx);
    const int ns = b->ns;
    const long long nao = b->nao;
    // canonical quartets (eri.cpp:414-429,825-843)
    std::vector<int> qs;
    std::vector<Request> reqs;
    std::vector<long long> boff;
    long long total = 0;
    for (int a = 0; a < ns; ++a)
        for (int bb = 0; bb < ns; ++bb)
            for (int c = 0; c < ns; ++c)
                for (int d = 0; d < ns; ++d) {
                    const std::array<int, 4> q{a, bb, c, d};
                    const std::array<std::array<int, 4>, 4> cand{q, std::array<int, 4>{c, d, a, bb},
                                                                 std::array<int, 4>{bb, a, d, c},
                                                                 std::array<int, 4>{d, c, bb, a}};
                    if (*std::min_element(cand.begin(), cand.end()) != q) continue;
                    const int st = check_quartet_supported(ctx, b, q.data());
                    if (st) return st;
                    Request r{};
                    for (int k = 0; k < 4; ++k) {
                        r.sh[k] = q[k];
                        r.xfix[k] = -1;
                        qs.push_back(q[k]);
                    }
                    r.cd0 = 0;
                    r.cd1 = -1;
                    const long long nb = ao_count(b, bb), nc = ao_count(b, c), nd = ao_count(b, d);
                    r.out_base = total;
                    r.stride[0] = nb * nc * nd;
                    r.stride[1] = nc * nd;
                    r.stride[2] = nd;
                    r.stride[3] = 1;
                    reqs.push_back(r);
                    boff.push_back(total);
                    total += ao_count(b, a) * nb * nc * nd;
                }
    ctx->stats.requested_quartets = ctx->stats.computed_quartets = reqs.size();
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    const size_t tensor_bytes = (size_t)nao * nao * nao * nao * sizeof(zc);
    HIP_TRY(ctx->blocks.ensure(std::max<size_t>(total, 1) * sizeof(zc)), "hipMalloc(blocks)");
    HIP_TRY(ctx->misc_a.ensure(boff.size() * sizeof(long long) + qs.size() * sizeof(int) + 64), "hipMalloc");
    zc* d_tensor = nullptr;
    if (out_on_device) d_tensor = reinterpret_cast<zc*>(out);
    else {
        HIP_TRY(ctx->staging.ensure(std::max<size_t>(tensor_bytes, 16)), "hipMalloc(tensor)");
        d_tensor = ctx->staging.as<zc>();
    }
    int st = run_requests(ctx, b, field, reqs, ctx->blocks.as<zc>());
    if (st) return st;
    long long* d_boff = ctx->misc_a.as<long long>();
    int* d_q = reinterpret_cast<int*>(d_boff + boff.size());
    HIP_TRY(hipMemcpyAsync(d_boff, boff.data(), boff.size() * sizeof(long long), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    HIP_TRY(hipMemcpyAsync(d_q, qs.data(), qs.size() * sizeof(int), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    HIP_TRY(hipMemsetAsync(d_tensor, 0, tensor_bytes, s), "hipMemsetAsync");
    HIP_TRY(gh::launch_tensor_scatter(d_boff, d_q, (int)boff.size(), total, b->d_shells.as<DevShell>(),
                                      b->d_aooff.as<long long>(), nao, ctx->blocks.as<zc>(), d_tensor, s),
            "scatter launch");
    HIP_TRY(hipStreamSynchronize(s), "tensor execution");
    if (!out_on_device)
        HIP_TRY(copy_to_host(ctx, reinterpret_cast<const char*>(d_tensor), reinterpret_cast<char*>(out), tensor_bytes),
                "copy");
    return GIAOHIP_OK;
}

int giaohip_nuclear_attraction(giaohip_ctx* ctx, const giaohip_basis* b, int nnuc, const double* charges,
                               const double* centers, const double field[6], double* out) {
    if (!ctx || !b || nnuc < 0 || (nnuc > 0 && (!charges || !centers)) || !field || !out)
        return GIAOHIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    reset_stats(ctx);
    // Nucleus::Nucleus (types.cpp:188-194): the charge first, then the centre
    for (int c = 0; c < nnuc; ++c) {
        if (!std::isfinite(charges[c]) || charges[c] <= 0.0)
            return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "nuclear charge must be finite and positive");
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(centers[3 * c + k]))
                return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "nuclear center values must be finite");
    }
    int lmax = 0;
    for (const DevShell& sh : b->shells) lmax = std::max(lmax, sh.l);
    if (2 * lmax > 32)
        return set_error(ctx, GIAOHIP_INVALID_ARGUMENT,
                         "nuclear attraction supports combined angular order through 32");
    if (2 * lmax > gh::kNucMaxL)
        return set_error(ctx, GIAOHIP_UNSUPPORTED, "nuclear attraction on the GPU supports shells through l = 4");
    const long long nao = b->nao;
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    const std::vector<double> kap = wave_vectors(b, field);
    std::vector<double> nuc(4 * (size_t)std::max(nnuc, 1), 0.0);
    for (int c = 0; c < nnuc; ++c) {
        nuc[4 * c] = charges[c];
        for (int k = 0; k < 3; ++k) nuc[4 * c + 1 + k] = centers[3 * c + k];
    }
    const size_t bytes = (size_t)nao * (size_t)nao * sizeof(zc);
    HIP_TRY(ctx->kappa.ensure(kap.size() * sizeof(double)), "hipMalloc");
    HIP_TRY(ctx->misc_a.ensure(nuc.size() * sizeof(double)), "hipMalloc");
    HIP_TRY(ctx->errrec.ensure(sizeof(DevError)), "hipMalloc");
    HIP_TRY(ctx->staging.ensure(std::max<size_t>(bytes, 16)), "hipMalloc");
    HIP_TRY(hipMemcpyAsync(ctx->kappa.p, kap.data(), kap.size() * sizeof(double), hipMemcpyHostToDevice, s),
            "hipMemcpyAsync");
    HIP_TRY(hipMemcpyAsync(ctx->misc_a.p, nuc.data(), nuc.size() * sizeof(double), hipMemcpyHostToDevice, s),
            "hipMemcpyAsync");
    HIP_TRY(hipMemsetAsync(ctx->errrec.p, 0, sizeof(DevError), s), "hipMemsetAsync");
    HIP_TRY(hipEventRecord(ctx->ev[0], s), "hipEventRecord");
    HIP_TRY(gh::launch_nuclear(b->ns, b->d_shells.as<DevShell>(), b->d_exps.as<double>(), b->d_coefs.as<double>(),
                               b->d_norms.as<double>(), ctx->kappa.as<double>(), ctx->misc_a.as<double>(), nnuc,
                               b->d_aooff.as<long long>(), nao, ctx->staging.as<zc>(), ctx->errrec.as<DevError>(), s),
            "nuclear attraction launch");
    HIP_TRY(hipEventRecord(ctx->ev[1], s), "hipEventRecord");
    HIP_TRY(hipStreamSynchronize(s), "nuclear attraction execution");
    ctx->stats.transform_ms = elapsed(ctx->ev[0], ctx->ev[1]);
    ctx->stats.launches = 1;
    DevError de{};
    HIP_TRY(hipMemcpy(&de, ctx->errrec.p, sizeof de, hipMemcpyDeviceToHost), "hipMemcpy");
    if (de.code != 0) return boys_error(ctx, de);
    if (bytes)
        HIP_TRY(copy_to_host(ctx, reinterpret_cast<const char*>(ctx->staging.p), reinterpret_cast<char*>(out), bytes),
                "copy");
    return GIAOHIP_OK;
}

int giaohip_primitive_eri(giaohip_ctx* ctx, const double prim[4][6], const int ang[4][3], const double field[6],
                          double out[2]) {
    if (!ctx || !prim || !ang || !field || !out) return GIAOHIP_INVALID_ARGUMENT;
    double centers[12], exps[4], coefs[4] = {1.0, 1.0, 1.0, 1.0};
    int L[4], np[4] = {1, 1, 1, 1}, nc[4] = {1, 1, 1, 1};
    for (int i = 0; i < 4; ++i) {
        if (!std::isfinite(prim[i][0]) || prim[i][0] <= 0.0)
            return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "Gaussian exponent must be finite and positive");
        exps[i] = prim[i][0];
        for (int k = 0; k < 3; ++k) centers[3 * i + k] = prim[i][1 + k];
        L[i] = ang[i][0] + ang[i][1] + ang[i][2];
    }
    giaohip_basis* b = nullptr;
    int st = giaohip_basis_create(ctx, 4, centers, L, np, nc, exps, coefs, &b);
    if (st) return st;
    const uint32_t q[4] = {0, 1, 2, 3};
    uint64_t total = 0;
    giaohip_eri_batch_size(b, q, 1, &total);
    std::vector<double> blk(2 * total);
    st = giaohip_eri_batch(ctx, b, q, 1, field, nullptr, blk.data(), 0);
    if (st == GIAOHIP_OK) {
        int idx[4];
        for (int i = 0; i < 4; ++i) {
            const auto comps = cartesian_components(L[i]);
            idx[i] = 0;
            for (size_t k = 0; k < comps.size(); ++k)
                if (comps[k][0] == ang[i][0] && comps[k][1] == ang[i][1] && comps[k][2] == ang[i][2]) idx[i] = (int)k;
        }
        const size_t k = (((size_t)idx[0] * gh::ncart(L[1]) + idx[1]) * gh::ncart(L[2]) + idx[2]) * gh::ncart(L[3]) + idx[3];
        double scale = 1.0;
        for (int i = 0; i < 4; ++i) {
            scale *= prim[i][4];
            if (prim[i][5] == 0.0) scale /= primitive_normalization(prim[i][0], ang[i][0], ang[i][1], ang[i][2]);
        }
        out[0] = scale * blk[2 * k];
        out[1] = scale * blk[2 * k + 1];
    }
    giaohip_basis_destroy(ctx, b);
    return st;
}

int giaohip_cholesky_diagonal(giaohip_ctx* ctx, const giaohip_basis* b, const uint64_t* pairs, size_t npairs,
                              const double field[6], double* out) {
    if (!ctx || !b || !field || !out || (npairs > 0 && !pairs)) return GIAOHIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    reset_stats(ctx);
    (void)hipSetDevice(ctx->device);
    const size_t dbytes = (size_t)b->nao * b->nao * sizeof(double);
    HIP_TRY(ctx->staging.ensure(std::max<size_t>(dbytes, 16)), "hipMalloc");
    int st = diagonal_to_device(ctx, b, pairs, npairs, field, ctx->staging.as<double>());
    if (st) return st;
    HIP_TRY(hipMemcpyAsync(out, ctx->staging.p, dbytes, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "diagonal execution");
    return GIAOHIP_OK;
}

int giaohip_cholesky_columns(giaohip_ctx* ctx, const giaohip_basis* b, const uint64_t* pivots, size_t npiv,
                             const double field[6], double* out, int real_output) {
    if (!ctx || !b || !field || (npiv > 0 && (!pivots || !out))) return GIAOHIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    reset_stats(ctx);
    const bool zero_field = field[0] == 0.0 && field[1] == 0.0 && field[2] == 0.0;
    if (real_output && !zero_field)
        return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "real Cholesky columns require zero phase field");
    if (npiv == 0) return GIAOHIP_OK;
    const long long n = b->nao;
    for (size_t i = 0; i < npiv; ++i)
        if (pivots[2 * i] >= (uint64_t)n || pivots[2 * i + 1] >= (uint64_t)n)
            return set_error(ctx, GIAOHIP_OUT_OF_RANGE, "Cholesky pivot AO index is out of range");
    const int sa = b->owner[pivots[0]], sb = b->owner[pivots[1]];
    for (size_t i = 0; i < npiv; ++i)
        if (b->owner[pivots[2 * i]] != sa || b->owner[pivots[2 * i + 1]] != sb)
            return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "Cholesky candidate pivots must share a shell pair");
    const DevShell& C = b->shells[sb];
    const DevShell& D = b->shells[sa];
    std::vector<Request> reqs;
    reqs.reserve(npiv * (size_t)b->ns * b->ns);
    for (size_t cand = 0; cand < npiv; ++cand) {
        const long long cl = (long long)pivots[2 * cand + 1] - b->ao_off[sb];
        const long long dl = (long long)pivots[2 * cand] - b->ao_off[sa];
        const int xc = (int)(cl / C.ncart), cc = (int)(cl % C.ncart);
        const int xd = (int)(dl / D.ncart), dc = (int)(dl % D.ncart);
        const int comp = cc * D.ncart + dc;
        for (int a = 0; a < b->ns; ++a)
            for (int bb = 0; bb < b->ns; ++bb) {
                const int q[4] = {a, bb, sb, sa};
                const int st = check_quartet_supported(ctx, b, q);
                if (st) return st;
                Request r{};
                for (int k = 0; k < 4; ++k) r.sh[k] = q[k];
                r.xfix[0] = r.xfix[1] = -1;
                r.xfix[2] = xc;
                r.xfix[3] = xd;
                r.cd0 = comp;
                r.cd1 = comp + 1;
                r.out_base = (long long)cand * n * n + b->ao_off[a] * n + b->ao_off[bb];
                r.stride[0] = n;
                r.stride[1] = 1;
                r.stride[2] = 0;
                r.stride[3] = 0;
                reqs.push_back(r);
            }
    }
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    const size_t ncol = npiv * (size_t)n * n;
    HIP_TRY(ctx->blocks.ensure(ncol * sizeof(zc)), "hipMalloc(columns)");
    int st = run_requests(ctx, b, field, reqs, ctx->blocks.as<zc>());
    if (st) return st;
    if (real_output) {
        HIP_TRY(ctx->staging.ensure(ncol * sizeof(double)), "hipMalloc");
        HIP_TRY(gh::launch_real_columns(ctx->blocks.as<zc>(), b->d_owner.as<int>(), n, (long long)npiv,
                                        ctx->staging.as<double>(), s),
                "real columns launch");
        HIP_TRY(hipMemcpyAsync(out, ctx->staging.p, ncol * sizeof(double), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    } else {
        HIP_TRY(hipMemcpyAsync(out, ctx->blocks.p, ncol * sizeof(zc), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    }
    HIP_TRY(hipStreamSynchronize(s), "columns execution");
    return GIAOHIP_OK;
}

}  // extern "C"

namespace {

struct CholHost {
    gh::CholState state;
    gh::CholState chosen;  // the policy's pivot, uploaded for the accept step
    int flags;
    int ntop;
    long long top[gh::kCholMaxBlock];  // a block's candidates, largest residual first
    // uploads of a block start: new candidates and their slots
    long long new_piv[gh::kCholMaxBlock];
    int new_slot[gh::kCholMaxBlock];
};
// Device layout of giaohip_cholesky::d_cand (bytes).
constexpr size_t kCandTop = 0, kCandNewPiv = 256, kCandNewSlot = 512, kCandCount = 640, kCandBytes = 768;

// Residual entries res[idx[i]] read back to the host.
int chol_gather(giaohip_ctx* ctx, giaohip_cholesky* c, const std::vector<long long>& idx, std::vector<double>& vals) {
    hipStream_t s = ctx->stream;
    vals.assign(idx.size(), 0.0);
    if (idx.empty()) return GIAOHIP_OK;
    HIP_TRY(c->d_gidx.ensure(idx.size() * sizeof(long long)), "hipMalloc");
    HIP_TRY(c->d_gval.ensure(idx.size() * sizeof(double)), "hipMalloc");
    HIP_TRY(hipMemcpyAsync(c->d_gidx.p, idx.data(), idx.size() * sizeof(long long), hipMemcpyHostToDevice, s),
            "hipMemcpyAsync");
    HIP_TRY(gh::launch_chol_gather(c->d_res.as<double>(), c->d_gidx.as<long long>(), (int)idx.size(),
                                   c->d_gval.as<double>(), s),
            "gather launch");
    HIP_TRY(hipMemcpyAsync(vals.data(), c->d_gval.p, idx.size() * sizeof(double), hipMemcpyDeviceToHost, s),
            "hipMemcpyAsync");
    HIP_TRY(hipStreamSynchronize(s), "Cholesky policy");
    return GIAOHIP_OK;
}

// The pivot of a complex run on a basis with general contractions
// (cholesky.py:193-262), from the global maximum (p, value): the largest
// residual of the latest candidate block replaces it when above the
// tolerance and at least 1% of the maximum. A pivot outside the current
// candidate set (and not in the raw-column cache) starts a new set: the
// contraction variants of its Cartesian component pair above the tolerance,
// largest residual first (stable), at most 32, the pivot forced in; the
// block is its first min(32, pivot_batch_size); the others enter the cache
// (least recently inserted leave first beyond min(1024, max(8, rank + 8))).
int policy_pivot(giaohip_ctx* ctx, giaohip_cholesky* c, long long& p, double& value) {
    const giaohip_basis* b = c->basis;
    const long long n = c->n;
    std::vector<double> vals;
    if (!c->pol_block.empty()) {
        if (int st = chol_gather(ctx, c, c->pol_block, vals)) return st;
        size_t best = 0;
        for (size_t i = 1; i < vals.size(); ++i)
            if (vals[i] > vals[best]) best = i;
        if (vals[best] > c->tolerance && vals[best] >= 0.01 * value) {
            p = c->pol_block[best];
            value = vals[best];
        }
    }
    auto cache_erase = [&](long long k) {
        auto it = c->pol_pos.find(k);
        if (it == c->pol_pos.end()) return false;
        c->pol_lru.erase(it->second);
        c->pol_pos.erase(it);
        return true;
    };
    if (cache_erase(p)) c->pol_cand.assign(1, p);
    if (std::find(c->pol_cand.begin(), c->pol_cand.end(), p) == c->pol_cand.end()) {
        const long long mu = p / n, nu = p % n;
        const int sa = b->owner[mu], sb = b->owner[nu];
        const DevShell &A = b->shells[sa], &B = b->shells[sb];
        const long long oa = b->ao_off[sa], ob = b->ao_off[sb];
        std::vector<long long> idx;
        for (int i = 0; i < A.ncontr; ++i)
            for (int j = 0; j < B.ncontr; ++j)
                idx.push_back((oa + (mu - oa) % A.ncart + (long long)A.ncart * i) * n +
                              (ob + (nu - ob) % B.ncart + (long long)B.ncart * j));
        if (int st = chol_gather(ctx, c, idx, vals)) return st;
        std::vector<size_t> keep;
        for (size_t i = 0; i < idx.size(); ++i)
            if (vals[i] > c->tolerance) keep.push_back(i);
        std::stable_sort(keep.begin(), keep.end(), [&](size_t x, size_t y) { return vals[x] > vals[y]; });
        if (keep.size() > 32) keep.resize(32);
        c->pol_cand.clear();
        for (size_t i : keep) c->pol_cand.push_back(idx[i]);
        if (std::find(c->pol_cand.begin(), c->pol_cand.end(), p) == c->pol_cand.end()) {
            if (c->pol_cand.empty()) c->pol_cand.push_back(p);
            else c->pol_cand.back() = p;
        }
        c->pol_block.assign(c->pol_cand.begin(),
                            c->pol_cand.begin() + std::min<size_t>(c->pol_cand.size(), std::min(32, c->pivot_batch)));
        for (long long k : c->pol_cand)
            if (k != p && !c->pol_pos.count(k)) {
                c->pol_lru.push_back(k);
                c->pol_pos[k] = std::prev(c->pol_lru.end());
            }
        const size_t cap = (size_t)std::min<uint64_t>(1024, std::max<uint64_t>(8, c->rank + 8));
        while (c->pol_lru.size() > cap) {
            c->pol_pos.erase(c->pol_lru.front());
            c->pol_lru.pop_front();
        }
    }
    cache_erase(p);
    return GIAOHIP_OK;
}

// Pivot steps of a run until the residual is below the tolerance or the rank
// reaches `target` (0: no target). Per step: when the pivot is not an open
// candidate of the current block, a new block (top-B candidates and their
// projections in one pass over the factors); then the pivot's column of ERIs,
// the accept step (the pivot's projection brought up to the current rank,
// factor row, residual update) and the next argmax. The host reads the pivot back to
// choose the column's ket class and to test block membership.
int chol_steps(giaohip_ctx* ctx, giaohip_cholesky* c, uint64_t target) {
    const giaohip_basis* b = c->basis;
    hipStream_t s = ctx->stream;
    const long long n = c->n, n2 = n * n;
#define CHOL_STEP_TRY(expr, what)                        \
    do {                                                 \
        hipError_t _e = (expr);                          \
        if (_e != hipSuccess) {                          \
            c->failed = hip_fail(ctx, _e, what);         \
            return c->failed;                            \
        }                                                \
    } while (0)
    CholHost* host = static_cast<CholHost*>(c->host);
    while (c->state.value > c->tolerance) {
        if (target && c->rank >= target) break;
        if (c->rank >= c->limit) {
            std::ostringstream m;
            m << "Cholesky ERI output requires at least " << (c->rank + 1) * (uint64_t)(n2 * 16 + 16)
              << " bytes, exceeding the requested rank limit " << c->limit;
            c->failed = set_error(ctx, GIAOHIP_OUT_OF_MEMORY, m.str());
            return c->failed;
        }
        long long p = c->state.pivot;
        double pvalue = c->state.value;
        if (c->policy) {
            if (int st = policy_pivot(ctx, c, p, pvalue)) {
                c->failed = st;
                return st;
            }
        }
        int slot = -1;
        for (int m = 0; m < c->ncand; ++m)
            if (((c->open >> m) & 1u) && c->cand[m] == p) slot = m;
        const bool new_block = slot < 0;
        bool block_pending = false;
        const long long mu = p / n, nu = p % n;
        const int sa = b->owner[mu], sb = b->owner[nu];
        const DevShell& C = b->shells[sb];
        const DevShell& D = b->shells[sa];
        gh::CholKet ket{};
        ket.sh_c = sb;
        ket.sh_d = sa;
        ket.lcd = C.l + D.l;
        ket.kcd = C.nprim * D.nprim;
        const long long cl = nu - b->ao_off[sb], dl = mu - b->ao_off[sa];
        ket.xc = (int)(cl / C.ncart);
        ket.xd = (int)(dl / D.ncart);
        ket.comp = (int)(cl % C.ncart) * D.ncart + (int)(dl % D.ncart);
        ket.n = n;
        // factor storage for row `rank`
        const int chunk = (int)(c->rank / c->rows_per_chunk);
        if (chunk >= (int)c->chunks.size()) {
            void* mem = nullptr;
            CHOL_STEP_TRY(counted_malloc(&mem, (size_t)c->rows_per_chunk * n2 * sizeof(zc)), "hipMalloc(factors)");
            c->chunks.push_back(static_cast<zc*>(mem));
            CHOL_STEP_TRY(c->d_chunkptr.ensure(c->chunks.size() * sizeof(zc*)), "hipMalloc");
            CHOL_STEP_TRY(hipMemcpy(c->d_chunkptr.p, c->chunks.data(), c->chunks.size() * sizeof(zc*),
                                    hipMemcpyHostToDevice),
                          "hipMemcpy");
        }
        zc* row = c->chunks[chunk] + (long long)(c->rank % c->rows_per_chunk) * n2;
        // The pivot's column ERIs; at a block start they run on the second
        // stream, beside the block's candidate selection and projection.
        hipStream_t sc = new_block ? ctx->aux : s;
        CHOL_STEP_TRY(hipEventRecord(c->ev[0], sc), "hipEventRecord");
        const int maxL = c->max_lab + ket.lcd;
        // the seeds of this ket shell pair, computed now unless still buffered
        const long long skey = (long long)sb * (long long)b->shells.size() + sa;
        int sslot = -1, lru = 0;
        for (int q = 0; q < c->seed_slots; ++q) {
            if (c->seed_key[q] == skey) sslot = q;
            if (c->seed_used[q] < c->seed_used[lru]) lru = q;
        }
        const bool seeds_ready = sslot >= 0;
        if (!seeds_ready) {
            sslot = lru;
            c->seed_key[sslot] = skey;
        }
        c->seed_used[sslot] = c->rank + 1;
        zc* seeds = c->d_seeds.as<zc>() + (size_t)sslot * c->seed_slot_zc;
        if (seeds_ready) c->seed_hits += 1;
        else c->seed_misses += 1;
        // the (la, lb) groups of bra pairs write disjoint rows of the column:
        // the larger ones first, spread over two streams; the launch topology
        // (the graph key) is the ket momenta, whether the seeds are computed,
        // and per group its kernel and stream
        std::vector<std::pair<double, size_t>> gorder;
        for (size_t gi = 0; gi < c->groups.size(); ++gi) {
            const auto& g = c->groups[gi];
            gorder.push_back({(double)g[3] * gh::nherm(g[0] + g[1] + C.l + D.l), gi});
        }
        std::sort(gorder.begin(), gorder.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
        // (an executable graph is only ever launched on one stream: updated in
        // place and launched on another stream than its previous launch, its
        // kernels were measured to run with stale arguments on this runtime)
        std::vector<int> topo = {C.l, D.l, seeds_ready ? 1 : 0, c->nunits > 0 ? 1 : 0, new_block ? 1 : 0};
        std::vector<int> gside(gorder.size()), gcol(gorder.size());
        {
            double gq[2] = {0.0, 0.0};
            for (size_t o = 0; o < gorder.size(); ++o) {
                const auto& g = c->groups[gorder[o].second];
                gside[o] = gq[1] < gq[0] ? 1 : 0;
                gq[gside[o]] += gorder[o].first;
                gcol[o] = g[0] + g[1] + C.l + D.l <= gh::kColumnMaxL &&
                          gh::chol_column_lds_bytes(g[0], g[1], C.l, D.l, ket.comp) <= 64 * 1024;
                topo.push_back((int)gorder[o].second * 4 + gside[o] * 2 + gcol[o]);
            }
        }
        auto issue_column = [&]() -> hipError_t {
            hipError_t e = gh::launch_chol_expand(c->d_units.as<gh::CholUnit>(), c->nunits, ket,
                                                  c->d_sq.as<SeedQuartet>(), c->d_wi.as<WorkItem>(), sc);
            if (e == hipSuccess && !seeds_ready)
                e = gh::launch_seed(nmax_bucket(maxL), c->d_sq.as<SeedQuartet>(), c->nunits, c->p1 * ket.kcd,
                                    b->d_shells.as<DevShell>(), b->d_exps.as<double>(), c->d_kappa.as<double>(),
                                    seeds, c->d_err.as<DevError>(), c->d_scratch, sc);
            if (e == hipSuccess) e = hipEventRecord(c->ev[8], sc);
            if (e == hipSuccess) e = hipStreamWaitEvent(c->s3, c->ev[8], 0);
            for (size_t o = 0; o < gorder.size() && e == hipSuccess; ++o) {
                const auto& g = c->groups[gorder[o].second];
                hipStream_t gs = gside[o] ? c->s3 : sc;
                gh::ClassArgs ca{{g[0], g[1], C.l, D.l}, 1};
                if (gcol[o])
                    e = gh::launch_chol_column(ca, ket.comp, c->d_wi.as<WorkItem>() + g[2], g[3],
                                               b->d_shells.as<DevShell>(), b->d_exps.as<double>(),
                                               b->d_coefs.as<double>(), b->d_norms.as<double>(),
                                               c->d_kappa.as<double>(), seeds, c->d_col.as<zc>(), gs);
                else
                    e = gh::launch_transform(ca, c->d_wi.as<WorkItem>() + g[2], g[3], b->d_shells.as<DevShell>(),
                                             b->d_exps.as<double>(), b->d_coefs.as<double>(), b->d_norms.as<double>(),
                                             c->d_kappa.as<double>(), seeds, c->d_col.as<zc>(), gs);
            }
            if (e == hipSuccess) e = hipEventRecord(c->ev[9], c->s3);
            if (e == hipSuccess) e = hipStreamWaitEvent(sc, c->ev[9], 0);
            return e;
        };
        // graphs for every pivot (1) or only inside a block (2: a block's
        // first column runs beside the candidate projection)
        const bool use_graph = ctx->chol_graphs == 1 || (ctx->chol_graphs == 2 && !new_block);
        giaohip_cholesky::ColumnGraph* cg = use_graph ? &c->graphs[topo] : nullptr;
        if (!cg || cg->uses++ == 0) {
            // direct launches (also the first pivot of a topology, so that
            // kernel attributes are set outside any capture)
            CHOL_STEP_TRY(issue_column(), "column launches");
        } else {
            // capture this pivot's launches, update the topology's graph with
            // them (or instantiate it), launch it
            CHOL_STEP_TRY(hipStreamBeginCapture(sc, hipStreamCaptureModeThreadLocal), "graph capture");
            const hipError_t e1 = issue_column();
            hipGraph_t graph = nullptr;
            const hipError_t e2 = hipStreamEndCapture(sc, &graph);
            if (e1 != hipSuccess || e2 != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                CHOL_STEP_TRY(e1 != hipSuccess ? e1 : e2, "graph capture");
            }
            if (cg->exec) {
                hipGraphNode_t bad = nullptr;
                hipGraphExecUpdateResult res = hipGraphExecUpdateError;
                if (hipGraphExecUpdate(cg->exec, graph, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
                    c->graph_updates += 1;
                } else {
                    (void)hipGetLastError();
                    (void)hipGraphExecDestroy(cg->exec);
                    cg->exec = nullptr;
                }
            }
            if (!cg->exec) {
                const hipError_t e3 = hipGraphInstantiate(&cg->exec, graph, nullptr, nullptr, 0);
                if (e3 != hipSuccess) {
                    cg->exec = nullptr;
                    (void)hipGraphDestroy(graph);
                    CHOL_STEP_TRY(e3, "graph instantiate");
                }
                c->graph_instantiations += 1;
            }
            (void)hipGraphDestroy(graph);
            CHOL_STEP_TRY(hipGraphLaunch(cg->exec, sc), "graph launch");
            c->graph_launches += 1;
        }
        CHOL_STEP_TRY(hipEventRecord(c->ev[1], sc), "hipEventRecord");
        if (new_block) {
            // New block: the largest residuals as candidates. Candidates still
            // open from the previous block keep their slots and projections
            // (complete up to this rank); only the others are projected, in
            // one pass over the factors.
            char* dc = static_cast<char*>(c->d_cand.p);
            CHOL_STEP_TRY(hipEventRecord(c->ev[3], s), "hipEventRecord");
            if (c->policy) {
                // the policy's pivot is the block's only candidate
                CHOL_STEP_TRY(hipEventRecord(c->ev[6], s), "hipEventRecord");
                host->top[0] = p;
                host->ntop = 1;
            } else {
                CHOL_STEP_TRY(gh::launch_chol_topk(c->d_res.as<double>(), n2, c->d_pmax.as<double>(),
                                                   c->d_pidx.as<long long>(), c->block,
                                                   reinterpret_cast<long long*>(dc + kCandTop),
                                                   reinterpret_cast<int*>(dc + kCandCount), s),
                              "top-k launch");
                CHOL_STEP_TRY(hipEventRecord(c->ev[6], s), "hipEventRecord");
                CHOL_STEP_TRY(hipMemcpyAsync(host->top, dc + kCandTop, sizeof host->top, hipMemcpyDeviceToHost, s),
                              "copy");
                CHOL_STEP_TRY(hipMemcpyAsync(&host->ntop, dc + kCandCount, sizeof(int), hipMemcpyDeviceToHost, s),
                              "copy");
                CHOL_STEP_TRY(hipStreamSynchronize(s), "Cholesky block");
            }
            const int ntop = host->ntop;
            if (ntop < 1 || host->top[0] != p) {
                (void)hipStreamSynchronize(sc);
                c->failed = set_error(ctx, GIAOHIP_RUNTIME_ERROR,
                                      "Cholesky pivot block does not start at the maximum residual");
                return c->failed;
            }
            // the block projects the factor rows below rank rounded down to a
            // multiple of 4 (the matrix-core row groups); the catch-up adds the rest
            const uint64_t rows4 = c->rank & ~uint64_t(3);
            unsigned keep = 0;
            int nnew = 0;
            for (int m = 0; m < ntop; ++m) {
                int found = -1;
                for (int q = 0; q < gh::kCholMaxBlock; ++q)
                    if (((c->open >> q) & 1u) && c->cand[q] == host->top[m]) found = q;
                if (found >= 0) {
                    keep |= 1u << found;
                } else {
                    host->new_piv[nnew] = host->top[m];
                    ++nnew;
                }
            }
            unsigned used = keep;
            for (int m = 0, q = 0; m < nnew; ++m) {
                while ((used >> q) & 1u) ++q;  // a free slot: at most ntop <= slots are in use
                host->new_slot[m] = q;
                c->cand[q] = host->new_piv[m];
                c->complete[q] = rows4;
                used |= 1u << q;
            }
            c->open = used;
            CHOL_STEP_TRY(hipMemcpyAsync(dc + kCandNewPiv, host->new_piv, sizeof host->new_piv, hipMemcpyHostToDevice,
                                         s),
                          "copy");
            CHOL_STEP_TRY(hipMemcpyAsync(dc + kCandNewSlot, host->new_slot, sizeof host->new_slot,
                                         hipMemcpyHostToDevice, s),
                          "copy");
            CHOL_STEP_TRY(hipEventRecord(c->ev[7], s), "hipEventRecord");
            if (nnew > 0) {
                const int nbt = nnew <= 16 ? 16 : 32;
                CHOL_STEP_TRY(c->d_coef.ensure((size_t)std::max<uint64_t>(rows4, 1) * nbt * sizeof(zc)), "hipMalloc");
                CHOL_STEP_TRY(gh::launch_chol_coef(c->d_chunkptr.as<const zc*>(), c->rows_per_chunk, n2, (int)rows4,
                                                   reinterpret_cast<const long long*>(dc + kCandNewPiv), nnew, nbt,
                                                   c->d_coef.as<zc>(), s),
                              "coefficient launch");
                CHOL_STEP_TRY(gh::launch_chol_project_mfma(c->d_chunkptr.as<const zc*>(), c->rows_per_chunk, n2,
                                                           (int)rows4, c->d_coef.as<zc>(), nnew, nbt,
                                                           reinterpret_cast<const int*>(dc + kCandNewSlot),
                                                           c->d_proj.as<zc>(), s),
                              "projection launch");
                c->rows_read += rows4;
            }
            CHOL_STEP_TRY(hipEventRecord(c->ev[4], s), "hipEventRecord");
            c->blocks += 1;
            c->projected += nnew;
            c->cand_rows += (uint64_t)nnew * rows4;
            c->ncand = gh::kCholMaxBlock;  // slot table is searched by the open mask
            for (int q = 0; q < gh::kCholMaxBlock; ++q)
                if (((c->open >> q) & 1u) && c->cand[q] == p) slot = q;
            CHOL_STEP_TRY(hipStreamWaitEvent(s, c->ev[1], 0), "hipStreamWaitEvent");
            block_pending = true;
        }
        CHOL_STEP_TRY(hipEventRecord(c->ev[5], s), "hipEventRecord");
        // the pivot's projection lacks the factors accepted since it was projected
        CHOL_STEP_TRY(gh::launch_chol_catchup_mfma(c->d_chunkptr.as<const zc*>(), c->rows_per_chunk, n2,
                                                   (int)c->complete[slot], (int)c->rank, p,
                                                   c->d_proj.as<zc>() + (long long)slot * n2, s),
                      "catch-up launch");
        c->complete[slot] = c->rank;
        if (c->policy) {
            host->chosen = gh::CholState{pvalue, p};
            CHOL_STEP_TRY(hipMemcpyAsync(c->d_state.p, &host->chosen, sizeof(gh::CholState), hipMemcpyHostToDevice, s),
                          "copy");
        }
        CHOL_STEP_TRY(gh::launch_chol_accept(c->d_col.as<zc>(), c->d_proj.as<zc>() + (long long)slot * n2, c->real,
                                             n, n2, c->d_state.as<gh::CholState>(), c->d_res.as<double>(), row,
                                             c->psd_tol, c->d_flags.as<int>(), c->d_pmax.as<double>(),
                                             c->d_pidx.as<long long>(), s),
                      "accept launch");
        CHOL_STEP_TRY(gh::launch_chol_argmax(c->d_res.as<double>(), n2, c->d_pmax.as<double>(),
                                             c->d_pidx.as<long long>(), c->d_state.as<gh::CholState>(), false, s),
                      "argmax launch");
        CHOL_STEP_TRY(hipEventRecord(c->ev[2], s), "hipEventRecord");
        CHOL_STEP_TRY(hipMemcpyAsync(&host->state, c->d_state.p, sizeof(gh::CholState), hipMemcpyDeviceToHost, s),
                      "copy");
        CHOL_STEP_TRY(hipMemcpyAsync(&host->flags, c->d_flags.p, sizeof(int), hipMemcpyDeviceToHost, s), "copy");
        CHOL_STEP_TRY(hipStreamSynchronize(s), "Cholesky step");
        c->column_ms += elapsed(c->ev[0], c->ev[1]);
        c->update_ms += elapsed(c->ev[5], c->ev[2]);
        if (block_pending) {
            const double bms = elapsed(c->ev[3], c->ev[6]) + elapsed(c->ev[7], c->ev[4]);
            c->update_ms += bms;
            c->block_ms += bms;
            block_pending = false;
        }
        c->pivots.push_back(mu);
        c->pivots.push_back(nu);
        c->rank += 1;
        c->open &= ~(1u << slot);
        if (host->flags & 1) {
            c->failed = set_error(ctx, GIAOHIP_RUNTIME_ERROR, "Coulomb Cholesky residual lost positive semidefiniteness");
            return c->failed;
        }
        c->state = host->state;
        DevError de{};
        if (c->rank % 64 == 0 || c->state.value <= c->tolerance) {
            CHOL_STEP_TRY(hipMemcpy(&de, c->d_err.p, sizeof de, hipMemcpyDeviceToHost), "hipMemcpy");
            if (de.code != 0) {
                c->failed = boys_error(ctx, de);
                return c->failed;
            }
        }
    }
#undef CHOL_STEP_TRY
    DevError de{};
    if (hipMemcpy(&de, c->d_err.p, sizeof de, hipMemcpyDeviceToHost) != hipSuccess)
        return c->failed = set_error(ctx, GIAOHIP_DEVICE_ERROR, "hipMemcpy");
    if (de.code != 0) return c->failed = boys_error(ctx, de);
    c->max_residual = std::max(c->state.value, 0.0);
    c->converged = c->state.value <= c->tolerance ? 1 : 0;
    return GIAOHIP_OK;
}

}  // namespace

extern "C" {

int giaohip_cholesky_run(giaohip_ctx* ctx, const giaohip_basis* b, const double field[6], double threshold,
                         uint64_t max_rank, uint64_t stop_after, giaohip_cholesky** out) {
    if (!ctx || !b || !out || !field) return GIAOHIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    reset_stats(ctx);
    *out = nullptr;
    if (!std::isfinite(threshold) || threshold < 0.0)
        return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "threshold must be finite and non-negative");
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    const long long n = b->nao, n2 = n * n;
    const int ns = b->ns;
    // cholesky.py:128-133,175-176: on one centre the London phase cancels and
    // the integrals are taken at zero field; one centre or zero field runs in
    // real arithmetic (symmetrized columns)
    bool single_center = true;
    for (const auto& sh : b->shells)
        for (int k = 0; k < 3; ++k) single_center = single_center && sh.center[k] == b->shells[0].center[k];
    double ifield_store[6];
    for (int k = 0; k < 6; ++k) ifield_store[k] = single_center ? 0.0 : field[k];
    const double* ifield = ifield_store;
    const bool real = single_center || (field[0] == 0.0 && field[1] == 0.0 && field[2] == 0.0);
    auto* c = new giaohip_cholesky;
    c->n = n;
    c->basis = b;
    c->real = real ? 1 : 0;
    for (int k = 0; k < 6; ++k) c->field[k] = ifield[k];
    // Factor rows live in chunks of about 2 GiB (at least one row).
    c->rows_per_chunk = (int)std::max<long long>(1, (2LL << 30) / (n2 * (long long)sizeof(zc)));
    auto cleanup = [&](int code) {
        c->release();
        delete c;
        return code;
    };

    // Static bra table: every ordered shell pair, grouped by (la, lb).
    std::vector<gh::CholUnit> units;
    for (int la = 0; la <= gh::kMaxShellL; ++la)
        for (int lb = 0; lb <= gh::kMaxShellL; ++lb) {
            const int first = (int)units.size();
            for (int a = 0; a < ns; ++a) {
                if (b->shells[a].l != la) continue;
                for (int bb = 0; bb < ns; ++bb) {
                    if (b->shells[bb].l != lb) continue;
                    for (int xa = 0; xa < b->shells[a].ncontr; ++xa)
                        for (int xb = 0; xb < b->shells[bb].ncontr; ++xb) {
                            gh::CholUnit u{};
                            u.sh_a = a;
                            u.sh_b = bb;
                            u.xa = xa;
                            u.xb = xb;
                            u.lab = la + lb;
                            u.kab = b->shells[a].nprim * b->shells[bb].nprim;
                            u.out_base = b->ao_off[a] * n + b->ao_off[bb];
                            units.push_back(u);
                        }
                }
            }
            if ((int)units.size() > first) {
                c->groups.push_back({la, lb, first, (int)units.size() - first});
                c->max_lab = std::max(c->max_lab, la + lb);
            }
        }
    long long p1 = 0, p2 = 0;
    for (auto& u : units) {
        u.p1 = p1;
        u.p2 = p2;
        p1 += u.kab;
        p2 += (long long)u.kab * u.lab;
    }
    c->p1 = p1;
    int max_l = 0, max_k = 1;
    for (const auto& sh : b->shells) {
        max_l = std::max(max_l, sh.l);
        max_k = std::max(max_k, sh.nprim);
    }
    if (2 * max_l + c->max_lab > gh::kMaxTotalL || max_l > gh::kMaxShellL)
        return cleanup(set_error(ctx, GIAOHIP_UNSUPPORTED, "GPU ERI kernels support combined angular order through 24"));
    c->nunits = (int)units.size();
    const long long seed_zc = (p2 + p1 * (long long)(2 * max_l + 4)) * (long long)max_k * max_k;
    const long long scr_zc =
        (long long)((gh::seed_scratch_ints(p1 * (long long)max_k * max_k) * sizeof(int) + sizeof(zc) - 1) / sizeof(zc));
    const int parts = gh::chol_partials(n2);
#define CHOL_TRY(expr, what)                             \
    do {                                                 \
        hipError_t _e = (expr);                          \
        if (_e != hipSuccess) return cleanup(hip_fail(ctx, _e, what)); \
    } while (0)
    CHOL_TRY(c->d_units.ensure(units.size() * sizeof(gh::CholUnit)), "hipMalloc");
    CHOL_TRY(hipMemcpy(c->d_units.p, units.data(), units.size() * sizeof(gh::CholUnit), hipMemcpyHostToDevice),
             "hipMemcpy");
    CHOL_TRY(c->d_sq.ensure(units.size() * sizeof(SeedQuartet)), "hipMalloc");
    CHOL_TRY(c->d_wi.ensure(units.size() * sizeof(WorkItem)), "hipMalloc");
    CHOL_TRY(c->d_res.ensure(n2 * sizeof(double)), "hipMalloc");
    CHOL_TRY(c->d_col.ensure(n2 * sizeof(zc)), "hipMalloc");
    CHOL_TRY(c->d_pmax.ensure(parts * sizeof(double)), "hipMalloc");
    CHOL_TRY(c->d_pidx.ensure(parts * sizeof(long long)), "hipMalloc");
    CHOL_TRY(c->d_state.ensure(sizeof(gh::CholState)), "hipMalloc");
    CHOL_TRY(c->d_flags.ensure(sizeof(int)), "hipMalloc");
    c->pivot_batch = ctx->chol_pivot_batch;
    {
        bool general = false;
        for (const auto& sh : b->shells) general = general || sh.ncontr > 1;
        c->policy = !real && general && c->pivot_batch > 1;
    }
    c->block = c->policy ? 1 : ctx->chol_block;
    c->slots = gh::chol_block_slots(c->block);
    CHOL_TRY(c->d_proj.ensure((size_t)c->slots * n2 * sizeof(zc)), "hipMalloc(projections)");
    CHOL_TRY(c->d_cand.ensure(kCandBytes), "hipMalloc");
    {
        // seed buffers: up to 1/16 of the free device memory, at most 8 GiB
        size_t free_b = 0, total_b = 0;
        CHOL_TRY(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        const size_t budget = std::min<size_t>(size_t(8) << 30, free_b / 16);
        const size_t per = (size_t)seed_zc * sizeof(zc);
        c->seed_slot_zc = seed_zc;
        c->seed_slots = (int)std::max<size_t>(1, std::min<size_t>(4096, per ? budget / per : 1));
        c->seed_key.assign(c->seed_slots, -1);
        c->seed_used.assign(c->seed_slots, 0);
    }
    CHOL_TRY(c->d_seeds.ensure(((size_t)c->seed_slots * seed_zc + scr_zc) * sizeof(zc)), "hipMalloc(seeds)");
    c->d_scratch = reinterpret_cast<int*>(c->d_seeds.as<zc>() + (size_t)c->seed_slots * seed_zc);
    CHOL_TRY(hipMemset(c->d_flags.p, 0, sizeof(int)), "hipMemset");
    for (auto& e : c->ev) CHOL_TRY(hipEventCreate(&e), "hipEventCreate");
    CHOL_TRY(hipStreamCreateWithFlags(&c->s3, hipStreamNonBlocking), "hipStreamCreate");
    CHOL_TRY(hipHostMalloc(&c->host, sizeof(CholHost), hipHostMallocDefault), "hipHostMalloc");

    // Diagonal over all ordered shell pairs a <= b (cholesky.py:143-169).
    std::vector<uint64_t> pairs;
    for (int a = 0; a < ns; ++a)
        for (int bb = a; bb < ns; ++bb) {
            pairs.push_back((uint64_t)a);
            pairs.push_back((uint64_t)bb);
        }
    CHOL_TRY(hipEventRecord(c->ev[0], s), "hipEventRecord");
    const int st = diagonal_to_device(ctx, b, pairs.data(), pairs.size() / 2, ifield, c->d_res.as<double>());
    if (st) return cleanup(st);
    CHOL_TRY(hipEventRecord(c->ev[1], s), "hipEventRecord");
    CHOL_TRY(gh::launch_chol_argmax(c->d_res.as<double>(), n2, c->d_pmax.as<double>(), c->d_pidx.as<long long>(),
                                    c->d_state.as<gh::CholState>(), true, s),
             "argmax launch");
    CHOL_TRY(hipMemcpyAsync(&c->state, c->d_state.p, sizeof c->state, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    CHOL_TRY(hipStreamSynchronize(s), "diagonal");
    c->diagonal_ms = elapsed(c->ev[0], c->ev[1]);
    const double floor = 64.0 * 2.220446049250313e-16 * std::max(c->state.value, 0.0);
    c->tolerance = std::max(threshold, floor);
    c->psd_tol = std::max(floor * 16.0, 1e-12);
    c->limit = max_rank ? max_rank : (uint64_t)n2;
    const std::vector<double> kap = wave_vectors(b, ifield);
    CHOL_TRY(c->d_kappa.ensure(kap.size() * sizeof(double)), "hipMalloc");
    CHOL_TRY(hipMemcpy(c->d_kappa.p, kap.data(), kap.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    CHOL_TRY(c->d_err.ensure(sizeof(DevError)), "hipMalloc");
    CHOL_TRY(hipMemset(c->d_err.p, 0, sizeof(DevError)), "hipMemset");
#undef CHOL_TRY
    const int status = chol_steps(ctx, c, stop_after);
    if (status != GIAOHIP_OK) return cleanup(status);
    *out = c;
    return GIAOHIP_OK;
}

int giaohip_cholesky_resume(giaohip_ctx* ctx, giaohip_cholesky* c, uint64_t stop_after) {
    if (!ctx || !c) return GIAOHIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    reset_stats(ctx);
    if (c->failed) return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "this Cholesky run ended with an error");
    if (c->converged) return GIAOHIP_OK;
    (void)hipSetDevice(ctx->device);
    return chol_steps(ctx, c, stop_after);
}

int giaohip_cholesky_remeasure(giaohip_ctx* ctx, giaohip_cholesky* c, double* max_residual) {
    if (!ctx || !c) return GIAOHIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    reset_stats(ctx);
    if (c->failed) return set_error(ctx, GIAOHIP_INVALID_ARGUMENT, "this Cholesky run ended with an error");
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    const giaohip_basis* b = c->basis;
    const long long n2 = c->n * c->n;
    const int ns = b->ns;
    std::vector<uint64_t> pairs;
    for (int a = 0; a < ns; ++a)
        for (int bb = a; bb < ns; ++bb) {
            pairs.push_back((uint64_t)a);
            pairs.push_back((uint64_t)bb);
        }
    HIP_TRY(ctx->staging.ensure(std::max<size_t>((size_t)n2 * sizeof(double), 16)), "hipMalloc");
    const int st = diagonal_to_device(ctx, b, pairs.data(), pairs.size() / 2, c->field, ctx->staging.as<double>());
    if (st) return st;
    const long long nb = (n2 + 255) / 256;
    HIP_TRY(ctx->misc_b.ensure(std::max<size_t>((size_t)nb * sizeof(double), 16)), "hipMalloc");
    HIP_TRY(gh::launch_chol_remeasure(c->d_chunkptr.as<const zc*>(), c->rows_per_chunk, n2, (int)c->rank,
                                      ctx->staging.as<double>(), ctx->misc_b.as<double>(), s),
            "remeasure launch");
    std::vector<double> part((size_t)nb);
    if (nb)
        HIP_TRY(hipMemcpyAsync(part.data(), ctx->misc_b.p, (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, s),
                "hipMemcpyAsync");
    HIP_TRY(hipStreamSynchronize(s), "remeasure execution");
    // np.max(residual, initial=0.0) (cholesky.py:297)
    double m = 0.0;
    for (double v : part) m = std::max(m, v);
    c->max_residual = m;
    if (max_residual) *max_residual = m;
    return GIAOHIP_OK;
}

int giaohip_cholesky_info(const giaohip_cholesky* c, uint64_t* rank, double* max_residual, double* tolerance,
                          int* converged, double* column_ms, double* update_ms, double* diagonal_ms) {
    if (!c) return GIAOHIP_INVALID_ARGUMENT;
    if (rank) *rank = c->rank;
    if (max_residual) *max_residual = c->max_residual;
    if (tolerance) *tolerance = c->tolerance;
    if (converged) *converged = c->converged;
    if (column_ms) *column_ms = c->column_ms;
    if (update_ms) *update_ms = c->update_ms;
    if (diagonal_ms) *diagonal_ms = c->diagonal_ms;
    return GIAOHIP_OK;
}

int giaohip_cholesky_graphs(const giaohip_cholesky* c, uint64_t* launches, uint64_t* instantiations,
                            uint64_t* updates) {
    if (!c) return GIAOHIP_INVALID_ARGUMENT;
    if (launches) *launches = c->graph_launches;
    if (instantiations) *instantiations = c->graph_instantiations;
    if (updates) *updates = c->graph_updates;
    return GIAOHIP_OK;
}

int giaohip_cholesky_blocks(const giaohip_cholesky* c, uint64_t* blocks, uint64_t* factor_rows_read,
                            double* block_ms, uint64_t* projected, uint64_t* candidate_rows) {
    if (!c) return GIAOHIP_INVALID_ARGUMENT;
    if (blocks) *blocks = c->blocks;
    if (projected) *projected = c->projected;
    if (candidate_rows) *candidate_rows = c->cand_rows;
    if (factor_rows_read) *factor_rows_read = c->rows_read;
    if (block_ms) *block_ms = c->block_ms;
    return GIAOHIP_OK;
}

int giaohip_cholesky_reconstruct(giaohip_ctx* ctx, const giaohip_cholesky* c, const int64_t ranges[8], double* out) {
    if (!ctx || !c || !ranges || !out) return GIAOHIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    long long r[8];
    long long total = 1;
    for (int k = 0; k < 4; ++k) {
        r[2 * k] = ranges[2 * k];
        r[2 * k + 1] = ranges[2 * k + 1];
        if (r[2 * k] < 0 || r[2 * k] > r[2 * k + 1] || r[2 * k + 1] > c->n)
            return set_error(ctx, GIAOHIP_OUT_OF_RANGE, "Cholesky reconstruction range is out of range");
        total *= r[2 * k + 1] - r[2 * k];
    }
    if (total == 0) return GIAOHIP_OK;
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    ScopedBuffer d_r, d_out;
    HIP_TRY(d_r.ensure(sizeof r), "hipMalloc");
    HIP_TRY(d_out.ensure((size_t)total * sizeof(zc)), "hipMalloc");
    HIP_TRY(hipMemcpyAsync(d_r.p, r, sizeof r, hipMemcpyHostToDevice, s), "copy");
    HIP_TRY(gh::launch_chol_reconstruct(c->d_chunkptr.as<const zc*>(), c->rows_per_chunk, c->n, (int)c->rank,
                                        d_r.as<long long>(), total, d_out.as<zc>(), s),
            "reconstruct launch");
    HIP_TRY(hipMemcpyAsync(out, d_out.p, (size_t)total * sizeof(zc), hipMemcpyDeviceToHost, s), "copy");
    HIP_TRY(hipStreamSynchronize(s), "reconstruct");
    return GIAOHIP_OK;
}

int giaohip_cholesky_coulomb(giaohip_ctx* ctx, const giaohip_cholesky* c, const double* density, double* out) {
    if (!ctx || !c || !density || !out) return GIAOHIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    const long long n = c->n, n2 = n * n;
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    ScopedBuffer d_dens, d_gamma, d_j, d_slot;
    HIP_TRY(d_dens.ensure((size_t)n2 * sizeof(zc)), "hipMalloc");
    HIP_TRY(d_gamma.ensure((size_t)std::max<uint64_t>(c->rank, 1) * sizeof(zc)), "hipMalloc");
    HIP_TRY(d_j.ensure((size_t)n2 * sizeof(zc)), "hipMalloc");
    HIP_TRY(d_slot.ensure(sizeof(int)), "hipMalloc");
    const int zero = 0;
    HIP_TRY(hipMemcpyAsync(d_slot.p, &zero, sizeof zero, hipMemcpyHostToDevice, s), "copy");
    HIP_TRY(hipMemcpyAsync(d_dens.p, density, (size_t)n2 * sizeof(zc), hipMemcpyHostToDevice, s), "copy");
    // gamma_P = sum conj(L_P[s,l]) D[l,s]; J = sum_P gamma_P L_P (one pass over the factors each)
    HIP_TRY(gh::launch_chol_gamma(c->d_chunkptr.as<const zc*>(), c->rows_per_chunk, n, (int)c->rank,
                                  d_dens.as<zc>(), d_gamma.as<zc>(), s),
            "gamma launch");
    HIP_TRY(gh::launch_chol_project(c->d_chunkptr.as<const zc*>(), c->rows_per_chunk, n2, (int)c->rank,
                                    d_gamma.as<zc>(), 1, 1, d_slot.as<int>(), d_j.as<zc>(), s),
            "coulomb launch");
    HIP_TRY(hipMemcpyAsync(out, d_j.p, (size_t)n2 * sizeof(zc), hipMemcpyDeviceToHost, s), "copy");
    HIP_TRY(hipStreamSynchronize(s), "coulomb");
    return GIAOHIP_OK;
}

int giaohip_cholesky_exchange(giaohip_ctx* ctx, const giaohip_cholesky* c, const double* density, double* out) {
    if (!ctx || !c || !density || !out) return GIAOHIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    reset_stats(ctx);
    const long long n = c->n, n2 = n * n;
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    if (!ctx->blas) {
        if (rocblas_create_handle(&ctx->blas) != rocblas_status_success) {
            ctx->blas = nullptr;
            return set_error(ctx, GIAOHIP_DEVICE_ERROR, "rocBLAS handle creation failed");
        }
    }
    if (rocblas_set_stream(ctx->blas, s) != rocblas_status_success)
        return set_error(ctx, GIAOHIP_DEVICE_ERROR, "rocBLAS stream setup failed");
    ScopedBuffer d_dens, d_t, d_k;
    HIP_TRY(d_dens.ensure((size_t)n2 * sizeof(zc)), "hipMalloc");
    HIP_TRY(d_t.ensure((size_t)n2 * sizeof(zc)), "hipMalloc");
    HIP_TRY(d_k.ensure((size_t)n2 * sizeof(zc)), "hipMalloc");
    HIP_TRY(hipMemcpyAsync(d_dens.p, density, (size_t)n2 * sizeof(zc), hipMemcpyHostToDevice, s), "copy");
    HIP_TRY(gh::launch_zero_ints(d_k.as<int>(), (int)(n2 * (long long)(sizeof(zc) / sizeof(int))), s), "fill");
    HIP_TRY(hipEventRecord(ctx->ev[0], s), "hipEventRecord");
    // K = sum_P L_P D L_P^H (row-major n x n). Column-major, the memory of a
    // row-major X is X^T, so per factor T^T = D^T L_P^T and
    // K^T += conj(L_P) T^T = (L_P^T)^H T^T.
    const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0};
    const rocblas_int m = (rocblas_int)n;
    auto* D = reinterpret_cast<const rocblas_double_complex*>(d_dens.p);
    auto* T = reinterpret_cast<rocblas_double_complex*>(d_t.p);
    auto* K = reinterpret_cast<rocblas_double_complex*>(d_k.p);
    for (uint64_t p = 0; p < c->rank; ++p) {
        auto* Lp = reinterpret_cast<const rocblas_double_complex*>(c->chunks[p / c->rows_per_chunk] +
                                                                   (long long)(p % c->rows_per_chunk) * n2);
        if (rocblas_zgemm(ctx->blas, rocblas_operation_none, rocblas_operation_none, m, m, m, &one, D, m, Lp, m,
                          &zero, T, m) != rocblas_status_success ||
            rocblas_zgemm(ctx->blas, rocblas_operation_conjugate_transpose, rocblas_operation_none, m, m, m, &one,
                          Lp, m, T, m, &one, K, m) != rocblas_status_success)
            return set_error(ctx, GIAOHIP_DEVICE_ERROR, "rocBLAS zgemm failed");
    }
    HIP_TRY(hipEventRecord(ctx->ev[1], s), "hipEventRecord");
    HIP_TRY(hipMemcpyAsync(out, d_k.p, (size_t)n2 * sizeof(zc), hipMemcpyDeviceToHost, s), "copy");
    HIP_TRY(hipStreamSynchronize(s), "exchange");
    ctx->stats.transform_ms = elapsed(ctx->ev[0], ctx->ev[1]);
    ctx->stats.launches = 2 * c->rank;
    return GIAOHIP_OK;
}

int giaohip_cholesky_pivots(const giaohip_cholesky* c, int64_t* out) {
    if (!c || !out) return GIAOHIP_INVALID_ARGUMENT;
    std::copy(c->pivots.begin(), c->pivots.end(), out);
    return GIAOHIP_OK;
}

int giaohip_cholesky_vectors(giaohip_ctx* ctx, const giaohip_cholesky* c, uint64_t first, uint64_t count, double* out) {
    if (!ctx || !c || !out) return GIAOHIP_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (first + count > c->rank) return set_error(ctx, GIAOHIP_OUT_OF_RANGE, "Cholesky vector range is out of range");
    const long long n2 = c->n * c->n;
    HIP_TRY(hipStreamSynchronize(ctx->stream), "copy");
    // rows of one factor chunk are contiguous: one copy per run of rows
    for (uint64_t k = 0; k < count;) {
        const uint64_t r = first + k;
        const uint64_t in_chunk = std::min<uint64_t>(count - k, c->rows_per_chunk - r % c->rows_per_chunk);
        const zc* row = c->chunks[r / c->rows_per_chunk] + (long long)(r % c->rows_per_chunk) * n2;
        HIP_TRY(copy_to_host(ctx, reinterpret_cast<const char*>(row), reinterpret_cast<char*>(out + 2 * k * n2),
                             in_chunk * n2 * sizeof(zc)),
                "copy");
        k += in_chunk;
    }
    return GIAOHIP_OK;
}

int giaohip_cholesky_destroy(giaohip_ctx* ctx, giaohip_cholesky* c) {
    if (!c) return GIAOHIP_OK;
    if (ctx) {
        std::lock_guard<std::mutex> lock(ctx->mu);
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamSynchronize(ctx->aux);
        if (c->s3) (void)hipStreamSynchronize(c->s3);
    }
    c->release();
    delete c;
    return GIAOHIP_OK;
}

}  // extern "C"
