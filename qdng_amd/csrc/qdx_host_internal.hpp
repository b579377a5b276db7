// Host-side internals of libqdx.so shared by the single-propagation runtime
// (qdx_host.cpp) and the batched-jobs runtime (qdx_batch.cpp): error
// handling, device buffers, grids and radix plans, the context, wave
// functions, Hamiltonians with QDng's composition and bound rules, and the
// Chebyshev parameters of OCheby::Init.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <tuple>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/qdx.h"
#include "qdx_impl.hpp"

using namespace qdx;

namespace qdx {

inline thread_local std::string g_err;

struct QdxError : std::runtime_error {
    int code;
    QdxError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string &msg) { throw QdxError(code, msg); }

#define HIPCHK(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) fail(QDX_ERR_DEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

template <class F>
int guarded(F &&f) {
    try {
        f();
        return QDX_OK;
    } catch (const QdxError &e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_err = "host allocation failed";
        return QDX_ERR_MEMORY;
    } catch (const std::exception &e) {
        g_err = e.what();
        return QDX_ERR_DEVICE;
    }
}

// ---------------------------------------------------------------------------
// device memory owned by RAII holders
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    void alloc(size_t n) {
        release();
        if (n == 0) return;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            fail(QDX_ERR_MEMORY, std::string("hipMalloc failed: ") + hipGetErrorString(e));
        }
        bytes = n;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

// ---------------------------------------------------------------------------
// geometry (GridSystem.cpp:207-211, Kspace.h:20-23)
struct Grid {
    int nd = 1;
    int N[4] = {1, 1, 1, 1};
    double xmin[4] = {0, 0, 0, 0}, xmax[4] = {0, 0, 0, 0};
    long size() const {
        long s = 1;
        for (int d = 0; d < nd; ++d) s *= N[d];
        return s;
    }
    double dx(int d) const { return (xmax[d] - xmin[d]) / double(N[d] - 1); }
    double dk(int d) const { return (2 * M_PI) / (xmax[d] - xmin[d] + dx(d)); }
    double weight() const {
        double w = 1;
        for (int d = 0; d < nd; ++d) w *= dx(d);
        return w;
    }
    bool operator==(const Grid &o) const {
        if (nd != o.nd) return false;
        for (int d = 0; d < nd; ++d) {  // GridSystem.cpp:248-261, GRID_EPS 1e-10
            if (N[d] != o.N[d]) return false;
            if (std::fabs(xmin[d] - o.xmin[d]) > 1e-10 || std::fabs(xmax[d] - o.xmax[d]) > 1e-10) return false;
        }
        return true;
    }
    GridDev dev() const {
        GridDev g;
        g.nd = nd;
        for (int d = 0; d < kMaxDims; ++d) g.N[d] = d < nd ? N[d] : 1;
        g.size = size();
        return g;
    }
};

// A line length L = 2^a 3^b p_1 p_2 ... with odd primes 5 <= p_i <= kMaxPrime
// (the device transforms run radix 8/4/2/3 stages and, for the other primes,
// a generic stage of r^2 operations per butterfly).  Returns false when L
// has a larger prime factor.
constexpr int kMaxPrime = 31;
inline bool factor_line(int L, int &a, int &b, std::vector<int> &big) {
    a = b = 0;
    big.clear();
    if (L < 1) return false;
    int m = L;
    while (m % 2 == 0) { ++a; m /= 2; }
    while (m % 3 == 0) { ++b; m /= 3; }
    for (int f = 5; m > 1 && f <= kMaxPrime; f += 2)
        while (m % f == 0) { big.push_back(f); m /= f; }
    return m == 1;
}
// only factors 2 and 3 (the batched kernels' in-place transforms)
inline bool is_smooth23(int n) {
    int a, b;
    std::vector<int> big;
    return factor_line(n, a, b, big) && big.empty();
}
inline bool is_small_prime(int x) {
    if (x < 5 || x > kMaxPrime || x % 2 == 0) return false;
    for (int f = 3; f * f <= x; f += 2)
        if (x % f == 0) return false;
    return true;
}

inline Grid make_grid(const qdx_grid *g) {
    if (!g) fail(QDX_ERR_PARAM, "grid is NULL");
    if (g->ndims < 1 || g->ndims > kMaxDims) fail(QDX_ERR_PARAM, "grids have 1..4 dimensions");
    Grid r;
    r.nd = g->ndims;
    for (int d = 0; d < r.nd; ++d) {
        if (g->N[d] < 2) fail(QDX_ERR_PARAM, "Zero points grid defined");
        if (g->xmax[d] - g->xmin[d] <= 0) fail(QDX_ERR_PARAM, "Zero length grid defined");
        int fa, fb;
        std::vector<int> fbig;
        if (!factor_line(g->N[d], fa, fb, fbig) || g->N[d] % 8 != 0)
            fail(QDX_ERR_PARAM, "grid sizes must be multiples of 8 without prime factors above 31");
        if (g->N[d] > kMaxLine) fail(QDX_ERR_PARAM, "grid dimension longer than 4096 points");
        r.N[d] = g->N[d];
        r.xmin[d] = g->xmin[d];
        r.xmax[d] = g->xmax[d];
    }
    return r;
}

// ---------------------------------------------------------------------------
// radix plans and twiddle tables exp(-2 pi i q / L).  L = 2^a 3^b p_1 ...
// with a >= 3; the first and the last stage take a radix 8, 4 or 2 (they run
// on registers in fft_line), radix-3 stages and the other primes sit in the
// middle.
inline void factor_or_fail(int L, int &a, int &b, std::vector<int> &big) {
    if (!factor_line(L, a, b, big) || a < 3)
        fail(QDX_ERR_PARAM, "line length must be a multiple of 8 without prime factors above 31");
}
inline std::vector<int> radix_plan(int L) {
    int a, b;
    std::vector<int> big;
    factor_or_fail(L, a, b, big);
    auto pw = [](int e) { return e >= 3 ? 8 : (e == 2 ? 4 : 2); };
    std::vector<int> r;
    if (b == 0 && big.empty()) {
        while (a >= 3) { r.push_back(8); a -= 3; }
        if (a > 0) r.push_back(pw(a));
    } else {
        const int first = a >= 4 ? 8 : 4;
        a -= first == 8 ? 3 : 2;
        const int last = pw(a);
        a -= last == 8 ? 3 : (last == 4 ? 2 : 1);
        r.push_back(first);
        while (a >= 3) { r.push_back(8); a -= 3; }
        if (a > 0) r.push_back(pw(a));
        for (int i = 0; i < b; ++i) r.push_back(3);
        for (int f : big) r.push_back(f);
        r.push_back(last);
    }
    if ((int)r.size() > kMaxStages) fail(QDX_ERR_PARAM, "too many radix stages");
    return r;
}

// Default plan of the 4-points-per-thread kernels: first and last radix 4
// (2 where L has too few factors 2), radix-8, -4, -2, -3 and other prime
// stages between.
inline std::vector<int> radix_plan4(int L) {
    int a, b;
    std::vector<int> big;
    factor_or_fail(L, a, b, big);
    const int first = a >= 5 ? 4 : 2;
    a -= first == 4 ? 2 : 1;
    const int last = a >= 2 ? 4 : 2;
    a -= last == 4 ? 2 : 1;
    std::vector<int> r{first};
    while (a >= 3) { r.push_back(8); a -= 3; }
    if (a == 2) r.push_back(4);
    if (a == 1) r.push_back(2);
    for (int i = 0; i < b; ++i) r.push_back(3);
    for (int f : big) r.push_back(f);
    r.push_back(last);
    if ((int)r.size() > kMaxStages) fail(QDX_ERR_PARAM, "too many radix stages");
    return r;
}

// A radix sequence fft_line can run for length L: product L, first and last
// radix 8, 4 or 2 (or, L = 8, the single radix 8), middle radices 8, 4, 3, 2
// or a prime 5 .. kMaxPrime.
inline bool radix_ok(int L, const std::vector<int> &r) {
    if (r.empty() || (int)r.size() > kMaxStages) return false;
    long prod = 1;
    for (size_t i = 0; i < r.size(); ++i) {
        const int x = r[i];
        const bool edge = i == 0 || i + 1 == r.size();
        if (!(x == 8 || x == 4 || x == 2 || (!edge && (x == 3 || is_small_prime(x))))) return false;
        prod *= x;
    }
    if (L == 8 && r.size() == 1) return r[0] == 8;   // one register stage (8 points per thread)
    return r.size() >= 2 && prod == L;
}

// Radix sequences for L with at most `extra` stages more than the fewest
// possible: every ordering of the middle radices, edges from {8, 4, 2}.
inline std::vector<std::vector<int>> radix_candidates(int L, int extra) {
    std::vector<std::vector<int>> out;
    if (L == 8) return {{8}};
    const int nmin = (int)radix_plan(L).size();
    int a, b;
    std::vector<int> radices{8, 4, 3, 2}, big;
    factor_line(L, a, b, big);
    for (int f : big)
        if (std::find(radices.begin(), radices.end(), f) == radices.end()) radices.push_back(f);
    std::vector<int> cur;
    // depth-first over sequences whose prefix product divides L
    std::function<void(long)> rec = [&](long prod) {
        const int n = (int)cur.size();
        if (prod == L) {
            if (radix_ok(L, cur)) out.push_back(cur);
            return;
        }
        if (n >= nmin + extra) return;
        for (int x : radices) {
            if ((L / prod) % x != 0) continue;
            cur.push_back(x);
            rec(prod * x);
            cur.pop_back();
        }
    };
    rec(1);
    return out;
}

}  // namespace

struct qdx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int refs = 0;          // live wave functions, Hamiltonians and propagators
    bool closing = false;  // qdx_ctx_destroy called while children were alive
    std::map<int, std::unique_ptr<DevBuf>> tw64, tw32;
    // radix plans set (or tuned) per line length and precision (fp32 flag)
    std::map<std::pair<int, bool>, std::vector<int>> radix;
    DevBuf partials;   // reduction scratch
    DevBuf scal;       // small device result area
    DevBuf wfx, ktab;  // scratch of the momentum expectation (a transformed copy, k table)
    double *pinned = nullptr;

    // per-launch HIP-event timing (qdx_ctx_profile): a fixed pool of events;
    // when it is used up the recorded spans are folded into the totals
    static constexpr size_t kEventPool = 8192;
    bool profiling = false;
    std::vector<hipEvent_t> events;
    size_t next_event = 0;
    std::vector<std::pair<size_t, size_t>> spans[QDX_NUM_KERNEL_CLASSES];
    double acc_ms[QDX_NUM_KERNEL_CLASSES] = {};
    int64_t acc_n[QDX_NUM_KERNEL_CLASSES] = {};
    void fold() {
        HIPCHK(hipDeviceSynchronize());   // timed launches may sit on side streams
        for (int c = 0; c < QDX_NUM_KERNEL_CLASSES; ++c) {
            for (auto &sp : spans[c]) {
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, events[sp.first], events[sp.second]));
                acc_ms[c] += ms;
                acc_n[c] += 1;
            }
            spans[c].clear();
        }
        next_event = 0;
    }
    // index of a fresh event; the caller records it right away
    size_t take_event() {
        if (events.empty()) {
            events.resize(kEventPool);
            for (auto &e : events) HIPCHK(hipEventCreate(&e));
        }
        return next_event++;
    }

    // radix sequence in use for length L: set / tuned, else the default
    std::vector<int> radix_of(int L, bool fp32) const {
        auto rt = radix.find({L, fp32});
        return rt != radix.end() ? rt->second : radix_plan(L);
    }
    LinePlan plan(int L, bool fp32) { return plan(L, fp32, radix_of(L, fp32)); }
    // plan of the kernels with ppt points per thread (8 or 4).  At 4 the
    // first and last radix must be 4 or 2: the plan set for L when it has
    // such edges, else radix_plan4.
    LinePlan plan_ppt(int L, bool fp32, int ppt) {
        if (ppt == 8) return plan(L, fp32);
        std::vector<int> r = radix_of(L, fp32);
        if (r.size() < 2 || r.front() > 4 || r.back() > 4) r = radix_plan4(L);
        LinePlan p = plan(L, fp32, r);
        p.ppt = 4;
        return p;
    }
    // points per thread of the single propagator's Chebyshev line kernels:
    // 0 chooses (4 on grids that give fewer than two waves per SIMD at 8)
    int line_ppt = 0;
    LinePlan plan(int L, bool fp32, const std::vector<int> &r) {
        auto &tab = fp32 ? tw32 : tw64;
        auto it = tab.find(L);
        if (it == tab.end()) {
            std::vector<double> t(2 * L);
            for (int q = 0; q < L; ++q) {
                long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)q / (long double)L;
                t[2 * q] = (double)cosl(a);
                t[2 * q + 1] = (double)sinl(a);
            }
            auto buf = std::make_unique<DevBuf>(fp32 ? 2 * L * sizeof(float) : 2 * L * sizeof(double));
            if (fp32) {
                std::vector<float> f(t.begin(), t.end());
                HIPCHK(hipMemcpy(buf->p, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
            } else {
                HIPCHK(hipMemcpy(buf->p, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
            }
            it = tab.emplace(L, std::move(buf)).first;
        }
        if (!radix_ok(L, r)) fail(QDX_ERR_PARAM, "radix plan does not fit the line length");
        LinePlan p;
        p.L = L;
        p.nst = (int)r.size();
        for (int i = 0; i < p.nst; ++i) p.rad[i] = r[i];
        p.tw = it->second->p;
        return p;
    }
};

namespace qdx {
// brackets one launch with events when the context is profiling
// Times exactly one kernel launch: the events ride on that launch's dispatch
// packet (hipExtLaunchKernelGGL), so the span is the kernel's own execution.
struct KTimer {
    qdx_ctx *c;
    int cls;
    size_t a = 0, b = 0;
    KTimer(qdx_ctx *ctx, int kind) : c(ctx), cls(kind) {
        if (!c->profiling) return;
        if (c->next_event + 2 > qdx_ctx::kEventPool) c->fold();
        a = c->take_event();
        b = c->take_event();
        g_launch_timing.start = c->events[a];
        g_launch_timing.stop = c->events[b];
    }
    ~KTimer() {
        if (!c->profiling) return;
        if (!g_launch_timing.start) c->spans[cls].push_back({a, b});   // consumed by the launch
        g_launch_timing = LaunchTiming();
    }
};

inline void ctx_free(qdx_ctx *ctx) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    ctx->pinned = nullptr;
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}
inline void ctx_acquire(qdx_ctx *ctx) {
    if (ctx) ctx->refs++;
}
inline void ctx_release(qdx_ctx *ctx) {
    if (!ctx) return;
    if (--ctx->refs == 0 && ctx->closing) ctx_free(ctx);
}
}  // namespace

struct qdx_wf {
    qdx_ctx *ctx = nullptr;
    Grid grid;
    int S = 1;
    DevBuf d;   // complex128 [S][N]
    double *ptr() const { return d.as<double>(); }
    long n() const { return grid.size(); }
};

namespace qdx {

// T_NIP: negative imaginary potential -i v(x) (OGobbler with nip)
enum TermKind { T_KIN = 0, T_POT = 1, T_DIP = 2, T_NIP = 3 };

struct Term {
    TermKind kind;
    int row, col;
    std::vector<double> mass;     // kinetic
    std::vector<double> v;        // potential / dipole mu / NIP
    std::vector<double> laser;    // dipole field
    double scale = 1.0;
    // bounds fixed by the operator instead of min / max of v (OGobbler:
    // Emax = gain, Emin = -i gain with nip, else 0; OGobbler.cpp:183-192)
    bool fixed = false;
    std::complex<double> fmax = 0, fmin = 0;
};

struct Elem {             // applied matrix element after QDng's mirroring
    std::vector<int> terms;
};

inline void sym_eigen(int n, std::vector<double> a, std::vector<double> &ev) {
    // cyclic Jacobi rotations; S <= 4 here
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += a[p * n + q] * a[p * n + q];
        if (off < 1e-300) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = a[p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                double theta = (a[q * n + q] - a[p * n + p]) / (2 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
                double c = 1 / std::sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; ++k) {
                    double akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = c * akp - s * akq;
                    a[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = c * apk - s * aqk;
                    a[q * n + k] = s * apk + c * aqk;
                }
            }
    }
    ev.resize(n);
    for (int i = 0; i < n; ++i) ev[i] = a[i * n + i];
}

}  // namespace

namespace qdx {
// k-space factors of -1/(2m) d^2/dx^2 along dimension d in FFT order
// (Kspace::Init1Dd2dx2, Kspace.cpp:22-50, with the factor of
// OGridNablaSq.cpp:51 and dk of Kspace.h:20-23)
inline std::vector<double> kspace_d2(const Grid &grid, int d, double mass) {
    const int Nx = grid.N[d];
    double dp = grid.dk(d);
    dp *= dp;
    dp *= -(-1. / 2. / mass);
    std::vector<double> k1(Nx);
    if (Nx % 2 == 0) {
        for (int q = 0; q < Nx / 2; ++q) {
            k1[q] = (double(q) * double(q)) * dp;
            k1[Nx - q - 1] = (double(q + 1) * double(q + 1)) * dp;
        }
    } else {
        k1[0] = 0;
        for (int q = 1; q < (Nx + 1) / 2; ++q) {
            k1[q] = (double(q) * double(q)) * dp;
            k1[Nx - q] = (double(q) * double(q)) * dp;
        }
    }
    return k1;
}

// k-space factors of the momentum -i d/dx along dimension d in FFT order
// (Kspace::Init1Dddx, Kspace.cpp:64-88; the Nyquist point of an even N is
// taken as -N/2 dk)
inline std::vector<double> kspace_d1(const Grid &grid, int d) {
    const int Nx = grid.N[d];
    const double dp = grid.dk(d);
    std::vector<double> k1(Nx);
    if (Nx % 2 == 0) {
        for (int q = 0; q < Nx / 2; ++q) {
            k1[q] = double(q) * dp;
            k1[Nx - q - 1] = -1 * double(q + 1) * dp;
        }
    } else {
        k1[0] = 0;
        for (int q = 1; q < (Nx + 1) / 2; ++q) {
            k1[q] = double(q) * dp;
            k1[Nx - q] = -1 * double(q) * dp;
        }
    }
    return k1;
}
}  // namespace qdx

struct qdx_ham {
    qdx_ctx *ctx = nullptr;
    Grid grid;
    int S = 1;
    bool multistate = false, unity = true;
    std::vector<Term> terms;
    // finalised state
    bool final_ = false;
    bool hermitian = true;
    std::vector<Elem> elem;        // S*S applied elements
    std::vector<bool> given;       // S*S elements added by the caller
    // device arrays (FP64 always, FP32 on demand)
    struct DevSet {
        std::vector<std::unique_ptr<DevBuf>> bufs;
        HamDev<double> h64;
        HamDev<float> h32;
        bool have32 = false;
    } dev;
    DevBuf tk;                     // kinetic tables, FP64
    DevBuf X64;                    // scratch for ham_apply
    DevBuf tmp;                    // complex128 scratch for expec

    void check_open() const {
        if (final_) fail(QDX_ERR_STATE, "Hamiltonian is already in use; add terms before the first use");
    }

    // QDng OMultistate::Add / Init mirroring (OMultistate.cpp:34-53, 99-142)
    void finalize() {
        if (final_) return;
        if (terms.empty()) fail(QDX_ERR_PARAM, multistate ? "Multistate operator is empty" : "Sum Operator is empty");
        given.assign(S * S, false);
        for (auto &t : terms) given[t.row * S + t.col] = true;
        hermitian = true;
        for (int i = 0; i < S; ++i)
            for (int j = 0; j < S; ++j)
                if (i != j && given[i * S + j] && given[j * S + i]) hermitian = false;
        if (has_nip() && multistate)
            fail(QDX_ERR_INCOMPAT, "negative imaginary potential: single-state (Sum) Hamiltonians only");
        elem.assign(S * S, Elem());
        for (int k = 0; k < (int)terms.size(); ++k) {
            const Term &t = terms[k];
            elem[t.row * S + t.col].terms.push_back(k);
            if (hermitian && t.row != t.col) elem[t.col * S + t.row].terms.push_back(k);
        }
        final_ = true;
        upload();
    }

    bool has_nip() const {
        for (auto &t : terms)
            if (t.kind == T_NIP) return true;
        return false;
    }
    // bounds of one matrix element: the sum of its terms (OSum.cpp:28-51)
    double term_emax(const Term &t) const {
        if (t.fixed) return t.fmax.real();
        switch (t.kind) {
        case T_KIN: {  // OGridNablaSq.cpp:89-101
            double T = 0;
            for (int d = 0; d < grid.nd; ++d)
                if (t.mass[d] > 0) T += 1 / (t.mass[d] * grid.dx(d) * grid.dx(d));
            return T * (M_PI * M_PI / 2);
        }
        case T_POT: return *std::max_element(t.v.begin(), t.v.end());
        case T_DIP:   // OGridDipole.cpp:36-39
            return *std::max_element(t.v.begin(), t.v.end()) * *std::max_element(t.laser.begin(), t.laser.end());
        case T_NIP: return 0;
        }
        return 0;
    }
    double term_emin(const Term &t) const {
        if (t.fixed) return t.fmin.real();
        if (t.kind == T_POT) return *std::min_element(t.v.begin(), t.v.end());
        return 0;
    }
    // Emin / Emax; emin_i / emax_i: imaginary parts (non-zero only for a
    // Sum with a negative imaginary potential; OMultistate takes |E_ij|)
    void bounds(double &emin, double &emax, double *emin_i = nullptr, double *emax_i = nullptr) {
        finalize();
        if (emin_i) *emin_i = 0;
        if (emax_i) *emax_i = 0;
        if (!multistate) {
            emin = emax = 0;
            for (auto &t : terms) {
                emax += term_emax(t);
                emin += term_emin(t);
                if (t.fixed && emin_i) *emin_i += t.fmin.imag();
                if (t.fixed && emax_i) *emax_i += t.fmax.imag();
            }
            return;
        }
        // OMultistate.cpp:163-204: eigenvalues of |Emax_ij| and |Emin_ij|
        std::vector<double> mx(S * S, 0.0), mn(S * S, 0.0), ev;
        for (int i = 0; i < S; ++i)
            for (int j = 0; j < S; ++j) {
                const Elem &e = elem[i * S + j];
                if (e.terms.empty()) continue;
                double a = 0, b = 0;
                for (int k : e.terms) {
                    a += term_emax(terms[k]);
                    b += term_emin(terms[k]);
                }
                mx[i * S + j] = std::fabs(a);
                mn[i * S + j] = std::fabs(b);
            }
        sym_eigen(S, mx, ev);
        emax = *std::max_element(ev.begin(), ev.end());
        sym_eigen(S, mn, ev);
        emin = *std::min_element(ev.begin(), ev.end());
    }

    // upload the applied matrix: W_ij = sum of potentials, mu_ij one dipole,
    // kinetic tables per state and dimension (Kspace.cpp:22-50)
    std::vector<std::vector<double>> hostW, hostMu, hostWi;
    void upload() {
        if (!ctx) return;
        const long n = grid.size();
        HamDev<double> &h = dev.h64;
        h.S = S;
        h.inv_n = 1.0 / double(n);
        for (int d = 0; d < kMaxDims; ++d) h.inv_nd[d] = d < grid.nd ? 1.0 / double(grid.N[d]) : 1.0;
        hostW.assign(S * S, {});
        hostMu.assign(S * S, {});
        hostWi.assign(S, {});
        std::vector<double> tkh;
        std::vector<long> tk_off(S * kMaxDims, -1);
        for (int i = 0; i < S; ++i) {
            h.unity[i] = 0;
            for (int j = 0; j < S; ++j) {
                const Elem &e = elem[i * S + j];
                std::vector<double> &W = hostW[i * S + j];
                std::vector<double> &Mu = hostMu[i * S + j];
                const Term *dip = nullptr;
                std::vector<double> kin_sum;
                for (int k : e.terms) {
                    const Term &t = terms[k];
                    if (t.kind == T_POT) {
                        if (W.empty()) W.assign(n, 0.0);
                        for (long x = 0; x < n; ++x) W[x] += t.v[x];
                    } else if (t.kind == T_NIP) {
                        std::vector<double> &Wi = hostWi[i];
                        if (Wi.empty()) Wi.assign(n, 0.0);
                        for (long x = 0; x < n; ++x) Wi[x] += t.v[x];
                    } else if (t.kind == T_DIP) {
                        if (dip && (dip->laser != t.laser || dip->scale != t.scale))
                            fail(QDX_ERR_INCOMPAT, "two dipole terms with different fields on one element");
                        if (Mu.empty()) Mu.assign(n, 0.0);
                        for (long x = 0; x < n; ++x) Mu[x] += t.v[x];
                        dip = &t;
                    } else {
                        if (i != j) fail(QDX_ERR_INCOMPAT, "kinetic operator on an off-diagonal element");
                        // per-dimension tables, summed over kinetic terms of the state
                        long base = (long)tkh.size();
                        for (int d = 0; d < grid.nd; ++d) {
                            const int Nx = grid.N[d];
                            const std::vector<double> k1 = kspace_d2(grid, d, t.mass[d]);
                            if (tk_off[i * kMaxDims + d] < 0) {
                                tk_off[i * kMaxDims + d] = (long)tkh.size();
                                tkh.insert(tkh.end(), k1.begin(), k1.end());
                            } else {
                                for (int q = 0; q < Nx; ++q) tkh[tk_off[i * kMaxDims + d] + q] += k1[q];
                            }
                        }
                        (void)base;
                    }
                }
                if (dip) h.Edip[i][j] = 0;  // set per step
                if (i == j && e.terms.empty() && multistate && unity) h.unity[i] = 1.0;
            }
        }
        // device copies
        tk.alloc(std::max<size_t>(tkh.size(), 1) * sizeof(double));
        if (!tkh.empty()) HIPCHK(hipMemcpy(tk.p, tkh.data(), tkh.size() * sizeof(double), hipMemcpyHostToDevice));
        for (int i = 0; i < S; ++i)
            for (int d = 0; d < kMaxDims; ++d)
                h.tk[i][d] = (d < grid.nd && tk_off[i * kMaxDims + d] >= 0) ? tk.as<double>() + tk_off[i * kMaxDims + d]
                                                                           : nullptr;
        for (int i = 0; i < S; ++i) {
            h.Wi[i] = hostWi[i].empty() ? nullptr : upload64(hostWi[i]);
            if (h.tk[i][0] == nullptr)
                for (int d = 0; d < kMaxDims; ++d) h.tk[i][d] = nullptr;
            for (int j = 0; j < S; ++j) {
                h.W[i][j] = nullptr;
                h.mu[i][j] = nullptr;
                if (!hostW[i * S + j].empty()) h.W[i][j] = upload64(hostW[i * S + j]);
                if (!hostMu[i * S + j].empty()) h.mu[i][j] = upload64(hostMu[i * S + j]);
            }
        }
    }
    std::map<const std::vector<double> *, const double *> up64_cache;
    const double *upload64(const std::vector<double> &v) {
        auto b = std::make_unique<DevBuf>(v.size() * sizeof(double));
        HIPCHK(hipMemcpy(b->p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
        const double *p = b->as<double>();
        dev.bufs.push_back(std::move(b));
        return p;
    }
    const float *upload32(const std::vector<double> &v) {
        std::vector<float> f(v.begin(), v.end());
        auto b = std::make_unique<DevBuf>(f.size() * sizeof(float));
        HIPCHK(hipMemcpy(b->p, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
        const float *p = b->as<float>();
        dev.bufs.push_back(std::move(b));
        return p;
    }
    void ensure32() {
        finalize();
        if (dev.have32) return;
        HamDev<float> &f = dev.h32;
        const HamDev<double> &d = dev.h64;
        f.S = d.S;
        f.inv_n = d.inv_n;
        for (int k = 0; k < kMaxDims; ++k) f.inv_nd[k] = d.inv_nd[k];
        for (int i = 0; i < S; ++i) {
            f.unity[i] = d.unity[i];
            f.Wi[i] = hostWi[i].empty() ? nullptr : upload32(hostWi[i]);
            for (int k = 0; k < kMaxDims; ++k) f.tk[i][k] = d.tk[i][k];
            for (int j = 0; j < S; ++j) {
                f.W[i][j] = hostW[i * S + j].empty() ? nullptr : upload32(hostW[i * S + j]);
                f.mu[i][j] = hostMu[i * S + j].empty() ? nullptr : upload32(hostMu[i * S + j]);
                f.Edip[i][j] = 0;
            }
        }
        dev.have32 = true;
    }

    // -laser[step] * scale for each applied dipole element (Laser.cpp:125-128)
    template <class R>
    void set_fields(HamDev<R> &h, long step) const {
        for (int i = 0; i < S; ++i)
            for (int j = 0; j < S; ++j) {
                h.Edip[i][j] = 0;
                for (int k : elem[i * S + j].terms) {
                    const Term &t = terms[k];
                    if (t.kind != T_DIP) continue;
                    if (step < 0 || step >= (long)t.laser.size())
                        fail(QDX_ERR_PARAM, "Clock values doesn't fit the laser field");
                    h.Edip[i][j] = -t.laser[step] * t.scale;
                }
            }
    }
    bool time_dependent() const {
        for (auto &t : terms)
            if (t.kind == T_DIP) return true;
        return false;
    }
    long min_laser() const {
        long m = -1;
        for (auto &t : terms)
            if (t.kind == T_DIP) m = (m < 0) ? (long)t.laser.size() : std::min(m, (long)t.laser.size());
        return m;
    }
};

namespace qdx {

// ---------------------------------------------------------------------------
// Bessel J_k(x), k = 0..n-1 by Miller's backward recurrence in extended
// precision, normalised with J_0 + 2 sum_{k>=1} J_{2k} = 1.  QDng obtains the
// same values from SLATEC ZBESJ (Bessel.cpp:22-58); its underflow count NZ is
// reproduced as the number of trailing values below exp(-ELIM),
// ELIM = 2.303 (1021 log10 2 - 3) (zbesj.f:183-188).
inline int bessel_j0(int n, double x, std::vector<double> &out) {
    out.assign(n, 0.0);
    if (n < 1) fail(QDX_ERR_PARAM, "Bessel: n must be >= 1");
    if (x == 0) {
        out[0] = 1.0;
        return 0;
    }
    const long double ax = std::fabs((long double)x);
    // start far enough above max(n, x) that J_m is negligible
    int m = std::max(n, (int)ax) + 16;
    while ((long double)m * std::log(2.0L * m / (2.718281828459045L * ax)) < 120.0L) m += 8;
    if (m % 2) ++m;
    std::vector<long double> j(m + 2, 0.0L);
    j[m + 1] = 0.0L;
    j[m] = 1e-300L;
    for (int k = m; k >= 1; --k) j[k - 1] = (2.0L * k / ax) * j[k] - j[k + 1];
    long double norm = j[0];
    for (int k = 2; k <= m; k += 2) norm += 2.0L * j[k];
    for (int k = 0; k < n; ++k) {
        long double v = j[k] / norm;
        if (x < 0 && (k & 1)) v = -v;
        out[k] = (double)v;
    }
    const double elim = 2.303 * (1021.0 * 0.30102999566398120 - 3.0);
    const double ulim = std::exp(-elim);
    int nz = 0;
    for (int k = n - 1; k >= 0 && std::fabs(out[k]) < ulim; --k) {
        out[k] = 0.0;
        ++nz;
    }
    return nz;
}

// Modified Bessel I_k(x) exp(-x), k = 0..n-1, x >= 0, by Miller's backward
// recurrence I_{k-1} = (2k/x) I_k + I_{k+1} in extended precision,
// normalised with I_0 + 2 sum_{k>=1} I_k = e^x.  QDng obtains them from
// SLATEC ZBESI with kode = 2 (Bessel.cpp:61-98); the underflow count is
// reproduced as for bessel_j0.
inline int bessel_i0_scaled(int n, double x, std::vector<double> &out) {
    out.assign(n, 0.0);
    if (n < 1) fail(QDX_ERR_PARAM, "Bessel: n must be >= 1");
    if (!(x >= 0)) fail(QDX_ERR_PARAM, "Bessel I: argument must be >= 0");
    if (x == 0) {
        out[0] = 1.0;
        return 0;
    }
    const long double ax = x;
    // I_m / I_0 ~ exp(-m^2 / (2x)) for m < x and faster beyond: start where
    // both m^2 / (2x) and m log(2m / (e x)) exceed 100
    int m = std::max(n, (int)ax) + 16;
    while ((long double)m * m / (2 * ax) < 100.0L && (long double)m * std::log(2.0L * m / (2.718281828459045L * ax)) < 100.0L)
        m += 8;
    std::vector<long double> v(m + 2, 0.0L);
    v[m + 1] = 0.0L;
    v[m] = 1e-300L;
    for (int k = m; k >= 1; --k) v[k - 1] = (2.0L * k / ax) * v[k] + v[k + 1];
    long double norm = v[0];
    for (int k = 1; k <= m; ++k) norm += 2.0L * v[k];
    for (int k = 0; k < n; ++k) out[k] = (double)(v[k] / norm);
    const double elim = 2.303 * (1021.0 * 0.30102999566398120 - 3.0);
    const double ulim = std::exp(-elim);
    int nz = 0;
    for (int k = n - 1; k >= 0 && std::fabs(out[k]) < ulim; --k) {
        out[k] = 0.0;
        ++nz;
    }
    return nz;
}

inline constexpr double kFp32CoeffFloor = 1e-10;

struct ChebyParams {
    int order = 0, nz = 0;
    double offset = 0, scaling = 1, prec = 1e-10, emin = 0, emax = 0;
    double offset_i = 0;   // Im offset (negative imaginary potential)
    std::vector<double> coeff;
};

// OCheby::Init (OCheby.cpp:382-493), real time, forward; imag: imaginary
// time (offset Emin - 5 % |Emin|, R = 1.05 |Emax - Emin|, coefficients from
// the scaled modified Bessel functions, OCheby.cpp:406-446)
inline ChebyParams cheby_params(qdx_ham *h, double dt, int order_param, double prec, double scaling_param,
                                bool imag = false) {
    if (dt == 0) fail(QDX_ERR_PARAM, "No time step defined");
    ChebyParams c;
    double emin_i = 0, emax_i = 0;
    h->bounds(c.emin, c.emax, &emin_i, &emax_i);
    if (imag) {
        if (!(dt > 0)) fail(QDX_ERR_PARAM, "imaginary time needs dt > 0");
        if (emin_i != 0 || emax_i != 0)
            fail(QDX_ERR_INCOMPAT, "negative imaginary potential: real-time Chebyshev only");
        c.offset = c.emin - std::fabs(c.emin) * 0.05;
        c.scaling = std::fabs(c.emax - c.emin) * 1.05;
    } else {
        // complex bounds (OCheby.cpp:406-416): offset (Re(Emax + Emin) / 2,
        // Im(Emax + Emin)); with Im offset != 0 R = 1.05 |Emax - Emin| and no
        // break test (OCheby.cpp:181)
        c.offset = (c.emax + c.emin) / 2;
        c.offset_i = emax_i + emin_i;
        const double dr = c.emax - c.emin, di = emax_i - emin_i;
        c.scaling = c.offset_i != 0 ? std::sqrt(dr * dr + di * di) * 1.05 : std::fabs(dr) / 2;
    }
    int order = order_param > 0 ? order_param : 0;
    if (order <= 0) order = std::abs(int(20 * c.scaling * dt));
    if (order < 10) order = 10;
    if (scaling_param > 0) c.scaling = scaling_param;
    if (order_param > 0) order = order_param;
    if (order < 3) fail(QDX_ERR_PARAM, "Chebychev recursion must be at least 3");
    std::vector<double> b;
    int zeroes = imag ? bessel_i0_scaled(order, c.scaling * dt, b) : bessel_j0(order, c.scaling * dt, b);
    order -= zeroes;
    if (order_param <= 0) {
        int i = 2;
        while (i < order && std::fabs(b[i - 1] - b[i - 2]) > 1e-32) i++;
        order = i;
    }
    if (!imag && order > 0 && order <= int(c.scaling * dt))
        fail(QDX_ERR_PARAM, "Chebychev recursion order is to small");
    c.order = order;
    c.nz = zeroes;
    c.coeff.assign(order, 0.0);
    c.coeff[0] = b[0];
    for (int i = 1; i < order; ++i) c.coeff[i] = 2.0 * b[i];
    if (prec >= 0) c.prec = prec;
    if (c.offset_i != 0) c.prec = 0.0;   // the whole series
    return c;
}

// FP32-bulk series.  The norm of a sum of FP32 terms wanders by ~1e-7 from
// term to term, so QDng's break test (|norm0 - sum |psi|^2| / norm0 < prec)
// can neither reach 1e-10 nor be trusted at a looser threshold (it may fire
// while the series is still far from converged).  The FP32 series therefore
// runs to a fixed length instead: every term up to the last k beyond R dt
// with |c_k| >= 1e-10, the norm test switched off (prec = 0).
// Orders to launch per step.  With the norm break test on (prec > 0) the
// series ends where |norm0 - sum |psi|^2| / norm0 < prec; launches are not
// issued for the trailing orders whose coefficients are all below
// 1e-4 * prec (beyond R dt): in QDng those orders only run when the norm
// test has not yet fired, and their terms together change psi by less than
// 1e-4 * prec relative.  Returns the series length when prec == 0.
inline int launch_window(const ChebyParams &cp, double dt) {
    if (!(cp.prec > 0)) return cp.order;
    const double x = cp.scaling * dt, floor = 1e-4 * cp.prec;
    int cut = cp.order;
    while (cut > 3 && cut - 1 > x && std::fabs(cp.coeff[cut - 1]) < floor) --cut;
    return cut;
}

inline void fp32_series(ChebyParams &cp, double dt) {
    cp.prec = 0.0;
    const double x = cp.scaling * dt;
    int cut = cp.order;
    while (cut > 3 && cut - 1 > x && std::fabs(cp.coeff[cut - 1]) < kFp32CoeffFloor) --cut;
    cp.order = cut;
}

}  // namespace

// ---------------------------------------------------------------------------
// propagators (shared by the step drivers in qdx_host.cpp and the optimal
// control driver in qdx_oct.cpp)
struct qdx_prop {
    enum Kind { CHEBY, GSPO, SIL, ARNOLDI } kind;
    qdx_ctx *ctx = nullptr;
    int precision = QDX_FP64;
    double dt = 0;
    Grid grid;
    int S = 1;
    // Chebyshev
    qdx_ham *ham = nullptr;
    ChebyParams cp;
    int k_launch = 0;   // orders launched per step (launch_window)
    bool imag = false;  // imaginary time, backward (exponent -dt)
    bool backward = false;  // real time backward (exponent +i dt), any real-time kind
    DevBuf phiA, phiB, X, ctl, partials, lastorders;
    bool cmp = false;   // QDX_COMPACT on a grid with nd >= 2: compact terms
    // single-state 1D lines: the whole step in one block (k_cheby_line) with
    // the phased coefficients coeff[k] exp(-i offset dt) on the device
    bool line_step = false;
    DevBuf coefd;
    int last_nb = 0, nb_init = 0;   // partial counts of the last Chebyshev step
    // SIL (Lanczos vectors and a buffer as device wave functions; `imag`
    // above selects imaginary time)
    int sil_order = 20;
    double sil_err = 1e-12;
    std::vector<qdx_wf *> lzb;
    qdx_wf *sil_buf = nullptr;
    int sil_napply = 0;   // H applications of the last step (SIL, Arnoldi)
    // Arnoldi (the Krylov vectors in lzb, the buffer in sil_buf)
    double arn_conv = 1e-12;
    int arn_order = 0;    // converged order; kept between steps as OArnoldi does
    // GSPO
    std::vector<qdx_ham *> ops;
    // exp(-i tau W) tables of time-independent x-space operators, per
    // (operator, tau, FP32 arrays)
    std::map<std::tuple<const qdx_ham *, double, bool>, std::unique_ptr<DevBuf>> unitaries;
    // exp(-i tau T_d(k)) tables of kinetic split operators, per (operator, tau)
    std::map<std::pair<const qdx_ham *, double>, std::unique_ptr<DevBuf>> kinexp;
    DevBuf psi32, psi0_32;
    DevBuf obs;   // device observables of qdx_prop_run

    virtual ~qdx_prop() = default;
};
