/*
 * FP64 CPU oracle for the QDng grid-propagation hot path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  Nothing in qdng_amd/ links,
 * loads or calls this file.
 *
 * Every function restates the arithmetic of the QDng file:line it cites
 * (paths relative to the QDng source tree).  The FFT is this file's own
 * mixed-radix Stockham transform in double precision (QDng calls FFTW).
 */
#include "qdx_oracle.h"

#include <cmath>
#include <complex>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>
#include <algorithm>
#include <map>
#include <dlfcn.h>
#include <cstdlib>
#ifdef _OPENMP
#include <omp.h>
#endif

using cplx = std::complex<double>;

namespace {

thread_local std::string g_err;

struct OracleError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

template <class F>
int guarded(F &&f) {
    try {
        return f();
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

/* ======================================================================
 * FFT: Stockham autosort, radices 4,2,3,5 and a direct DFT for any other
 * prime factor.  Convention of QDng's FFTW plans (src/fft/fft.cpp:98-119,
 * 255, 278): out[k] = sum_j in[j] exp(sign*2*pi*i*j*k/N), unnormalised,
 * sign = -1 forward, +1 backward.
 * ==================================================================== */
struct Plan1D {
    int n = 0;
    std::vector<int> radix;
    std::vector<cplx> root;   /* root[q] = exp(+2 pi i q / n), q in [0,n) */
};

Plan1D make_plan(int n) {
    if (n < 1) throw OracleError("FFT length must be positive");
    Plan1D p;
    p.n = n;
    int m = n;
    while (m % 4 == 0) { p.radix.push_back(4); m /= 4; }
    while (m % 2 == 0) { p.radix.push_back(2); m /= 2; }
    for (int f = 3; m > 1; f += 2) {
        while (m % f == 0) { p.radix.push_back(f); m /= f; }
        if ((long)f * f > m && m > 1) { p.radix.push_back(m); m = 1; }
    }
    p.root.resize(n);
    for (int q = 0; q < n; ++q) {
        long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)q / (long double)n;
        p.root[q] = cplx((double)cosl(a), (double)sinl(a));
    }
    return p;
}

inline cplx tw(const Plan1D &p, long q, int sign) {
    q %= p.n;
    cplx w = p.root[q];
    return sign < 0 ? std::conj(w) : w;
}

/* one transform of a contiguous line; work has length n */
void fft_line(const Plan1D &p, cplx *x, cplx *work, int sign) {
    const int n = p.n;
    if (n == 1) return;
    cplx *in = x, *out = work;
    int ns = 1;
    std::vector<cplx> a(64), b(64);
    for (int r : p.radix) {
        if ((int)a.size() < r) { a.resize(r); b.resize(r); }
        const int m = n / r;
        const long step = n / ((long)ns * r);     /* root index unit for this stage */
        for (int j = 0; j < m; ++j) {
            const int k = j % ns;
            for (int t = 0; t < r; ++t) a[t] = in[j + t * m] * tw(p, (long)k * t * step, sign);
            if (r == 2) {
                b[0] = a[0] + a[1];
                b[1] = a[0] - a[1];
            } else if (r == 4) {
                cplx s02 = a[0] + a[2], d02 = a[0] - a[2];
                cplx s13 = a[1] + a[3], d13 = a[1] - a[3];
                cplx jd13 = (sign < 0) ? cplx(d13.imag(), -d13.real()) : cplx(-d13.imag(), d13.real());
                b[0] = s02 + s13;
                b[2] = s02 - s13;
                b[1] = d02 + jd13;
                b[3] = d02 - jd13;
            } else {
                /* direct r-point DFT with exact roots of unity of order r */
                const long unit = n / r;
                for (int u = 0; u < r; ++u) {
                    cplx s = 0;
                    for (int t = 0; t < r; ++t) s += a[t] * tw(p, ((long)t * u % r) * unit, sign);
                    b[u] = s;
                }
            }
            const int base = (j / ns) * ns * r + k;
            for (int t = 0; t < r; ++t) out[base + t * ns] = b[t];
        }
        ns *= r;
        std::swap(in, out);
    }
    if (in != x) std::memcpy(x, in, sizeof(cplx) * n);
}

std::map<int, Plan1D> g_plans;

const Plan1D &plan_for(int n) {
    auto it = g_plans.find(n);
    if (it != g_plans.end()) return it->second;
    return g_plans.emplace(n, make_plan(n)).first->second;
}

/* ND transform over the QDng layout (dimension 0 contiguous) */
/* Optional FFT backend for the CPU baseline timing: the public FFTW3 API
 * (QDng's FFT library, src/fft/fft.cpp) as served by MKL's libmkl_rt.so,
 * loaded at run time through these prototypes of the documented FFTW3
 * interface.  Plans are made once per shape and direction (planning is not
 * part of a step) and executed in place with fftw_execute_dft. */
namespace fftw3 {
typedef double fcomplex[2];
typedef void *plan_t;
typedef plan_t (*plan_dft_fn)(int rank, const int *n, fcomplex *in, fcomplex *out, int sign, unsigned flags);
typedef void (*execute_dft_fn)(const plan_t p, fcomplex *in, fcomplex *out);
typedef void (*destroy_fn)(plan_t p);
typedef int (*init_threads_fn)(void);
typedef void (*plan_with_nthreads_fn)(int n);
typedef void (*set_threads_fn)(int n);
constexpr unsigned kEstimate = 1u << 6;   /* FFTW_ESTIMATE */
constexpr unsigned kUnaligned = 1u << 1;  /* FFTW_UNALIGNED */
struct Api {
    void *lib = nullptr;
    plan_dft_fn plan_dft = nullptr;
    execute_dft_fn execute_dft = nullptr;
    destroy_fn destroy = nullptr;
    init_threads_fn init_threads = nullptr;
    plan_with_nthreads_fn plan_with_nthreads = nullptr;
    set_threads_fn set_threads = nullptr;
    bool on = false;
    int threads = 1;
    std::map<std::vector<int>, plan_t> plans;   /* key: sign, reversed dims */
} g_fftw;

bool load() {
    if (g_fftw.lib) return true;
    /* MKL's threads through GNU OpenMP, as the rest of this library */
    setenv("MKL_THREADING_LAYER", "GNU", 0);
    const char *names[] = {"libmkl_rt.so", "libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so", "libfftw3.so.3"};
    for (const char *n : names) {
        g_fftw.lib = dlopen(n, RTLD_NOW | RTLD_LOCAL);
        if (g_fftw.lib) break;
    }
    if (!g_fftw.lib) return false;
    g_fftw.plan_dft = (plan_dft_fn)dlsym(g_fftw.lib, "fftw_plan_dft");
    g_fftw.execute_dft = (execute_dft_fn)dlsym(g_fftw.lib, "fftw_execute_dft");
    g_fftw.destroy = (destroy_fn)dlsym(g_fftw.lib, "fftw_destroy_plan");
    g_fftw.init_threads = (init_threads_fn)dlsym(g_fftw.lib, "fftw_init_threads");
    g_fftw.plan_with_nthreads = (plan_with_nthreads_fn)dlsym(g_fftw.lib, "fftw_plan_with_nthreads");
    g_fftw.set_threads = (set_threads_fn)dlsym(g_fftw.lib, "MKL_Set_Num_Threads");
    if (!g_fftw.plan_dft || !g_fftw.execute_dft || !g_fftw.destroy) {
        dlclose(g_fftw.lib);
        g_fftw.lib = nullptr;
        return false;
    }
    if (g_fftw.init_threads) g_fftw.init_threads();
    return true;
}

void clear_plans() {
    for (auto &kv : g_fftw.plans) g_fftw.destroy(kv.second);
    g_fftw.plans.clear();
}

/* data[] <- the unnormalised transform (sign -1 forward), dimension 0 fastest
 * (FFTW's rank-nd transform with the dimensions reversed, fft.cpp:91-94) */
void run(int nd, const int *N, cplx *data, int sign) {
    std::vector<int> key{sign};
    for (int d = nd - 1; d >= 0; --d) key.push_back(N[d]);
    auto it = g_fftw.plans.find(key);
    if (it == g_fftw.plans.end()) {
        long total = 1;
        for (int d = 0; d < nd; ++d) total *= N[d];
        std::vector<cplx> tmp(total);
        if (g_fftw.plan_with_nthreads) g_fftw.plan_with_nthreads(g_fftw.threads);
        plan_t pl = g_fftw.plan_dft(nd, key.data() + 1, reinterpret_cast<fcomplex *>(tmp.data()),
                                    reinterpret_cast<fcomplex *>(tmp.data()), sign, kEstimate | kUnaligned);
        if (!pl) throw OracleError("fftw_plan_dft failed");
        it = g_fftw.plans.emplace(key, pl).first;
    }
    g_fftw.execute_dft(it->second, reinterpret_cast<fcomplex *>(data), reinterpret_cast<fcomplex *>(data));
}
}  // namespace fftw3

void fftnd(int nd, const int *N, cplx *data, int sign) {
    if (fftw3::g_fftw.on) return fftw3::run(nd, N, data, sign);
    long total = 1;
    for (int d = 0; d < nd; ++d) total *= N[d];
    long stride = 1;
    for (int d = 0; d < nd; ++d) {
        const int L = N[d];
        const Plan1D &p = plan_for(L);
        const long nlines = total / L;
        const long s = stride;
#pragma omp parallel
        {
            std::vector<cplx> line(L), work(L);
#pragma omp for schedule(static)
            for (long ln = 0; ln < nlines; ++ln) {
                const long inner = ln % s, outer = ln / s;
                cplx *base = data + outer * s * L + inner;
                if (s == 1) {
                    fft_line(p, base, work.data(), sign);
                } else {
                    for (int i = 0; i < L; ++i) line[i] = base[i * s];
                    fft_line(p, line.data(), work.data(), sign);
                    for (int i = 0; i < L; ++i) base[i * s] = line[i];
                }
            }
        }
        stride *= L;
    }
}

/* ======================================================================
 * Bessel J_k(x), k = 0..n-1 (QDng src/linalg/Bessel.cpp:22-58 calls SLATEC
 * ZBESJ with kode = 1).  Values come from libstdc++'s std::cyl_bessel_j.
 * The underflow count NZ is emulated: trailing components below
 * exp(-ELIM), ELIM = 2.303*(1021*log10(2) - 3) as set in ZBESJ, are zeroed.
 * ==================================================================== */
int bessel_j(int n, double x, std::vector<double> &out) {
    out.assign(n, 0.0);
    for (int k = 0; k < n; ++k) out[k] = std::cyl_bessel_j((double)k, x);
    const double elim = 2.303 * (1021.0 * 0.30102999566398120 - 3.0);
    const double ulim = std::exp(-elim);
    int nz = 0;
    for (int k = n - 1; k >= 0 && std::fabs(out[k]) < ulim; --k) {
        out[k] = 0.0;
        ++nz;
    }
    return nz;
}

/* ======================================================================
 * Modified Bessel I_k(x) exp(-x), k = 0..n-1 (QDng src/linalg/Bessel.cpp:
 * 61-98 calls SLATEC ZBESI with kode = 2, scaled by exp(-x)).  Values come
 * from libstdc++'s std::cyl_bessel_i; the underflow count NZ is emulated as
 * for bessel_j.
 * ==================================================================== */
int bessel_i_scaled(int n, double x, std::vector<double> &out) {
    out.assign(n, 0.0);
    if (x > 700) throw OracleError("bessel_i_scaled: argument beyond the oracle's range");
    const double ex = std::exp(-x);
    for (int k = 0; k < n; ++k) out[k] = std::cyl_bessel_i((double)k, x) * ex;
    const double elim = 2.303 * (1021.0 * 0.30102999566398120 - 3.0);
    const double ulim = std::exp(-elim);
    int nz = 0;
    for (int k = n - 1; k >= 0 && std::fabs(out[k]) < ulim; --k) {
        out[k] = 0.0;
        ++nz;
    }
    return nz;
}

/* ======================================================================
 * Grid (src/qdlib/GridSystem.cpp:207-211: dx = (xmax-xmin)/(N-1))
 * ==================================================================== */
struct Grid {
    int nd = 0;
    int N[4] = {1, 1, 1, 1};
    double xmin[4] = {0, 0, 0, 0}, xmax[4] = {0, 0, 0, 0};
    long size() const {
        long s = 1;
        for (int d = 0; d < nd; ++d) s *= N[d];
        return s;
    }
    double dx(int d) const { return (xmax[d] - xmin[d]) / double(N[d] - 1); }
    /* Kspace.h:20-23 */
    double dk(int d) const { return (2 * M_PI) / (xmax[d] - xmin[d] + dx(d)); }
    double weight() const {
        double w = 1;
        for (int d = 0; d < nd; ++d) w *= dx(d);
        return w;
    }
};

std::vector<Grid> g_grids;

/* ======================================================================
 * Operators
 * ==================================================================== */
enum Kind { KIN, POT, DIP, SUM, MS, NIP };

struct Op {
    Kind kind;
    int grid = -1;
    std::vector<double> mass;     /* KIN */
    std::vector<double> tk;       /* KIN: T(k) in FFT order (Diff.cpp:120-139) */
    std::vector<double> v;        /* POT/DIP */
    std::vector<double> laser;    /* DIP: raw field values */
    double scale = 1.0;           /* DIP: Laser::_scale */
    std::vector<int> kids;        /* SUM */
    int S = 0;                    /* MS */
    std::vector<int> mat;         /* MS: S*S op ids, -1 = empty */
    bool unity = true, hermitian = true, dmulti = false;
    bool fixed = false;           /* Gobbler: Emax/Emin fixed by the gain */
    cplx fmax = 0, fmin = 0;
};

std::vector<Op> g_ops;

Op &op_at(int id) {
    if (id < 0 || id >= (int)g_ops.size()) throw OracleError("invalid operator id");
    return g_ops[id];
}

long op_size(const Op &o);

int op_grid(const Op &o) {
    if (o.kind == SUM) return op_grid(op_at(o.kids.at(0)));
    if (o.kind == MS) {
        for (int id : o.mat)
            if (id >= 0) return op_grid(op_at(id));
        throw OracleError("empty multistate operator");
    }
    return o.grid;
}

/* Kspace.cpp:22-50 (second derivative, factor -1/(2m)), summed over
 * dimensions with the per-dimension broadcast of Diff.cpp:120-139. */
std::vector<double> kinetic_table(const Grid &g, const std::vector<double> &mass) {
    std::vector<double> t(g.size(), 0.0);
    long stride = 1;
    for (int d = 0; d < g.nd; ++d) {
        const int Nx = g.N[d];
        double dp = g.dk(d);
        dp *= dp;
        const double factor = -1. / 2. / mass[d];
        dp *= -factor;
        std::vector<double> k1(Nx);
        if (Nx % 2 == 0) {
            for (int i = 0; i < Nx / 2; ++i) {
                k1[i] = (double(i) * double(i)) * dp;
                k1[Nx - i - 1] = (double(i + 1) * double(i + 1)) * dp;
            }
        } else {
            k1[0] = 0;
            for (int i = 1; i < (Nx + 1) / 2; ++i) {
                k1[i] = (double(i) * double(i)) * dp;
                k1[Nx - i] = (double(i) * double(i)) * dp;
            }
        }
        for (long j = 0; j < g.size(); ++j) t[j] += k1[(j / stride) % Nx];
        stride *= Nx;
    }
    return t;
}

/* symmetric eigenvalues (Jacobi rotations), used where QDng calls dsyev
 * (OMultistate.cpp:163-204, ODMultistate.cpp:138-159) */
void sym_eigen(int n, std::vector<double> a, std::vector<double> &ev, std::vector<double> *vec) {
    std::vector<double> V(n * n, 0.0);
    for (int i = 0; i < n; ++i) V[i * n + i] = 1;
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
                for (int k = 0; k < n; ++k) {
                    double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    ev.resize(n);
    for (int i = 0; i < n; ++i) ev[i] = a[i * n + i];
    if (vec) *vec = V;   /* columns are eigenvectors */
}

/* planted-mistake switch of the tests (0: the restatement as written; 1:
 * transposed GSPO eigenvectors, 2: +i v for the negative imaginary
 * potential, 3: coefficient phase from Re(offset) only, 4: the Gobbler high pass
 * of a NIP multiplied instead of added) */
int g_plant = 0;

/* laser value at clock step (Laser.cpp:125-128) */
double laser_at(const Op &o, int step) {
    if (step < 0 || step >= (int)o.laser.size()) throw OracleError("clock step outside the laser field");
    return o.laser[step] * o.scale;
}

cplx op_emax(const Op &o);
cplx op_emin(const Op &o);

cplx op_emax(const Op &o) {
    if (o.fixed) return o.fmax;   /* OGobbler.cpp:183-186 */
    switch (o.kind) {
    case KIN: { /* OGridNablaSq.cpp:89-101 */
        const Grid &g = g_grids[o.grid];
        double T = 0;
        for (int i = 0; i < g.nd; ++i)
            if (o.mass[i] > 0) T += 1 / (o.mass[i] * g.dx(i) * g.dx(i));
        T *= (M_PI * M_PI / 2);
        return T;
    }
    case POT: return *std::max_element(o.v.begin(), o.v.end());   /* OGridPotential.cpp:124-127 */
    case NIP: return 0;   /* fixed bounds, returned above */
    case DIP:  /* OGridDipole.cpp:36-39: max(mu) * max(raw laser) */
        return *std::max_element(o.v.begin(), o.v.end()) * *std::max_element(o.laser.begin(), o.laser.end());
    case SUM: { /* OSum.cpp:28-39 */
        cplx d = 0;
        for (int k : o.kids) d += op_emax(op_at(k));
        return d;
    }
    case MS: { /* OMultistate.cpp:163-181 */
        std::vector<double> m(o.S * o.S, 0.0), ev;
        for (int i = 0; i < o.S; ++i)
            for (int j = 0; j < o.S; ++j)
                if (o.mat[i * o.S + j] >= 0) m[i * o.S + j] = std::abs(op_emax(op_at(o.mat[i * o.S + j])));
        sym_eigen(o.S, m, ev, nullptr);
        return *std::max_element(ev.begin(), ev.end());
    }
    }
    return 0;
}

cplx op_emin(const Op &o) {
    if (o.fixed) return o.fmin;   /* OGobbler.cpp:188-192 */
    switch (o.kind) {
    case KIN: return 0;                                              /* OGridNablaSq.cpp:103-106 */
    case POT: return *std::min_element(o.v.begin(), o.v.end());      /* OGridPotential.cpp:129-132 */
    case NIP: return 0;
    case DIP: return 0;                                              /* OGridDipole.cpp:41-44 */
    case SUM: {
        cplx d = 0;
        for (int k : o.kids) d += op_emin(op_at(k));
        return d;
    }
    case MS: { /* OMultistate.cpp:183-204: |Emin_ij| before diagonalising */
        std::vector<double> m(o.S * o.S, 0.0), ev;
        for (int i = 0; i < o.S; ++i)
            for (int j = 0; j < o.S; ++j)
                if (o.mat[i * o.S + j] >= 0) m[i * o.S + j] = std::abs(op_emin(op_at(o.mat[i * o.S + j])));
        sym_eigen(o.S, m, ev, nullptr);
        return *std::min_element(ev.begin(), ev.end());
    }
    }
    return 0;
}

int op_states(const Op &o) {
    if (o.kind == SUM) return op_states(op_at(o.kids.at(0)));
    return o.kind == MS ? o.S : 1;
}

long op_size(const Op &o) { return g_grids.at(op_grid(o)).size(); }

/* dest = O src (single-state operators act on one state vector) */
void op_apply(const Op &o, int step, const cplx *src, cplx *dst);

void op_apply_add(const Op &o, int step, const cplx *src, cplx *dst) {
    const long n = op_size(o);
    switch (o.kind) {
    case POT: /* OGridPotential.cpp:139-142: MultElementsAdd */
        for (long i = 0; i < n; ++i) dst[i] += src[i] * o.v[i];
        return;
    case DIP: { /* OGridDipole.cpp:51-54: factor -E(t) */
        const double e = -laser_at(o, step);
        for (long i = 0; i < n; ++i) dst[i] += src[i] * o.v[i] * e;
        return;
    }
    case NIP: { /* OGobbler.cpp:197-201 through Operator.h:307-316: += src * (v i) * (-1) */
        const double sgn = g_plant == 2 ? 1.0 : -1.0;   /* planted: +i v */
        for (long i = 0; i < n; ++i) dst[i] += src[i] * (o.v[i] * cplx(0, 1)) * sgn;
        return;
    }
    default: { /* Operator.h:307-316: temporary + AddElements */
        std::vector<cplx> tmp(n * op_states(o));
        op_apply(o, step, src, tmp.data());
        for (long i = 0; i < (long)tmp.size(); ++i) dst[i] += tmp[i];
        return;
    }
    }
}

void op_apply(const Op &o, int step, const cplx *src, cplx *dst) {
    const Grid &g = g_grids.at(op_grid(o));
    const long n = g.size();
    switch (o.kind) {
    case KIN: { /* Diff.cpp:203-212: IFFT[ FFT[src] * T(k) / N ] */
        std::vector<cplx> buf(src, src + n);
        fftnd(g.nd, g.N, buf.data(), -1);
        const double d = 1.0 / double(n);
        for (long i = 0; i < n; ++i) buf[i] = buf[i] * o.tk[i] * d;
        fftnd(g.nd, g.N, buf.data(), +1);
        std::copy(buf.begin(), buf.end(), dst);
        return;
    }
    case POT:
        for (long i = 0; i < n; ++i) dst[i] = src[i] * o.v[i];
        return;
    case DIP: {
        const double e = -laser_at(o, step);
        for (long i = 0; i < n; ++i) dst[i] = src[i] * o.v[i] * e;
        return;
    }
    case NIP: { /* OGobbler.cpp:197-201: dest = src * (v i) * sgn, sgn = -1 */
        const double sgn = g_plant == 2 ? 1.0 : -1.0;   /* planted: +i v */
        for (long i = 0; i < n; ++i) dst[i] = src[i] * (o.v[i] * cplx(0, 1)) * sgn;
        return;
    }
    case SUM: /* OSum.cpp:53-60 */
        op_apply(op_at(o.kids[0]), step, src, dst);
        for (size_t k = 1; k < o.kids.size(); ++k) op_apply_add(op_at(o.kids[k]), step, src, dst);
        return;
    case MS: { /* OMultistate.cpp:206-232 */
        for (int i = 0; i < o.S; ++i) {
            cplx *di = dst + i * n;
            std::fill(di, di + n, cplx(0));
            for (int j = 0; j < o.S; ++j) {
                int id = o.mat[i * o.S + j];
                if (id >= 0)
                    op_apply_add(op_at(id), step, src + j * n, di);
                else if (i == j && o.unity)
                    for (long k = 0; k < n; ++k) di[k] += src[j * n + k];
            }
        }
        return;
    }
    }
}

/* ======================================================================
 * Observables
 * ==================================================================== */
cplx overlap(const Grid &g, int S, const cplx *a, const cplx *b) {
    /* WFGridCartesian.cpp:62-116 (x-space weights) and WFMultistate.cpp:202-226 */
    cplx tot = 0;
    const long n = g.size();
    for (int s = 0; s < S; ++s) {
        cplx c = 0;
        for (long i = 0; i < n; ++i) c += std::conj(a[s * n + i]) * b[s * n + i];
        tot += c * g.weight();
    }
    return tot;
}

/* ======================================================================
 * Chebyshev (OCheby.cpp)
 * ==================================================================== */
struct Cheby {
    int ham = -1;
    int S = 1;
    double dt = 0;
    int order = 0, nz = 0;
    double prec = 1e-10;        /* OCheby.h:11 DEFAULT_PREC */
    cplx offset = 0;
    double scaling = 1;
    std::vector<double> coeff;
    double emin = 0, emax = 0;
    bool imag = false;          /* imaginary time, backward: exponent -dt */
    bool backward = false;      /* real time backward: exponent +i dt (OPropagator.h:82-100) */
};
std::vector<Cheby> g_cheby;

double raw_norm(const cplx *psi, long n) { /* OCheby.cpp:64-90 */
    double norm = 0;
    for (long j = 0; j < n; ++j) norm += (std::conj(psi[j]) * psi[j]).real();
    return norm;
}

int cheby_create(int ham, int S, double dt, int order_param, double prec, double scaling_param,
                 orc_cheby_info *info, bool imag = false) {
    Cheby c;
    c.ham = ham;
    c.S = S;
    c.dt = dt;
    c.imag = imag;
    const Op &H = op_at(ham);
    if (op_states(H) != S) throw OracleError("Hamiltonian and wave function differ in number of states");
    if (dt == 0) throw OracleError("No time step defined");
    cplx Emax = op_emax(H), Emin = op_emin(H);
    c.emax = Emax.real();
    c.emin = Emin.real();
    if (!imag) {
        /* OCheby.cpp:406-416 (real time, forward: exponent -i dt) */
        c.offset = cplx((Emax + Emin).real() / 2, (Emax + Emin).imag());
        if (c.offset.imag() != 0)
            c.scaling = std::abs(Emax - Emin) * 1.05;
        else
            c.scaling = std::abs(Emax - Emin) / 2;
    } else {
        /* OCheby.cpp:410-416 (imaginary time: exponent real, so the
         * 1.05 widening applies) */
        c.offset = Emin - std::abs(Emin) * 0.05;
        c.scaling = std::abs(Emax - Emin) * 1.05;
    }
    int order = order_param > 0 ? order_param : 0;
    if (order <= 0) order = std::abs(int(20 * c.scaling * dt));   /* OCheby.cpp:419-420 */
    if (order < 10) order = 10;                                    /* OCheby.cpp:422-423 */
    if (scaling_param > 0) c.scaling = scaling_param;              /* OCheby.cpp:430-431 */
    if (order_param > 0) order = order_param;                      /* OCheby.cpp:433-434 */
    if (order < 3) throw OracleError("Chebychev recursion must be at least 3");
    std::vector<double> bessel;
    int zeroes = imag ? bessel_i_scaled(order, c.scaling * dt, bessel)   /* OCheby.cpp:443-446 */
                      : bessel_j(order, c.scaling * dt, bessel);         /* OCheby.cpp:440-442 */
    order -= zeroes;                                               /* OCheby.cpp:451 */
    if (order_param <= 0) {                                        /* OCheby.cpp:454-464 */
        const double precission = 1e-32;                           /* OCheby.h:10 BESSEL_DELTA */
        int i = 2;
        while (i < order && std::fabs(bessel[i - 1] - bessel[i - 2]) > precission) i++;
        order = i;
    }
    if (!imag && order > 0 && order <= int(c.scaling * dt))        /* OCheby.cpp:467-468 */
        throw OracleError("Chebychev recursion order is to small");
    c.order = order;
    c.nz = zeroes;
    c.coeff.assign(order, 0.0);                                    /* OCheby.cpp:476-480 */
    c.coeff[0] = bessel[0];
    for (int i = 1; i < order; ++i) c.coeff[i] = 2.0 * bessel[i];
    if (prec >= 0) c.prec = prec;                                  /* OCheby.cpp:483-484 */
    g_cheby.push_back(c);
    if (info) {
        info->order = c.order;
        info->nz = c.nz;
        info->offset = c.offset.real();
        info->scaling = c.scaling;
        info->prec = c.prec;
        info->emin = c.emin;
        info->emax = c.emax;
    }
    return (int)g_cheby.size() - 1;
}

/* OCheby.cpp:92-197, real-time branch */
int cheby_step(Cheby &c, int step, cplx *psi) {
    const Op &H = op_at(c.ham);
    const long n = op_size(H) * c.S;
    const cplx I(0, 1);
    /* OPropagator.h:44-100: -i dt forward, +i dt backward */
    const cplx exponent = c.backward ? I * c.dt : -I * c.dt;
    const cplx exp_ = exponent / c.dt;
    std::vector<cplx> ket0(psi, psi + n), ket1(n), ket2(n), buf(n);
    int napply = 0;

    const double norm = raw_norm(psi, n);
    op_apply(H, step, psi, ket1.data());
    ++napply;
    for (long j = 0; j < n; ++j) ket1[j] += -1.0 * c.offset * psi[j];   /* AddElements */
    const cplx f1 = exp_ / c.scaling;
    for (long j = 0; j < n; ++j) ket1[j] *= f1;                       /* MultElements */
    /* planted 3: the phase of the real part of the offset only */
    const cplx ph = std::exp(exponent * (g_plant == 3 ? cplx(c.offset.real(), 0) : c.offset));
    const cplx c0 = c.coeff[0] * ph, c1 = c.coeff[1] * ph;
    for (long j = 0; j < n; ++j) psi[j] *= c0;
    for (long j = 0; j < n; ++j) buf[j] = ket1[j] * c1;
    for (long j = 0; j < n; ++j) psi[j] += buf[j];

    cplx *k0 = ket0.data(), *k1 = ket1.data(), *k2 = ket2.data(), *bf = buf.data();
    const cplx exp2 = 2.0 * exp_;
    for (int i = 2; i < c.order; ++i) {
        double accnorm = 0;
        const cplx coeff = c.coeff[i] * ph;
        op_apply(H, step, k1, bf);
        ++napply;
        for (long j = 0; j < n; ++j) {                                  /* OCheby.cpp:167-174 */
            bf[j] = (bf[j] - c.offset * k1[j]) / c.scaling;
            bf[j] *= exp2;
            k2[j] = bf[j] + k0[j];
            k0[j] = k2[j];
            psi[j] += k2[j] * coeff;
            accnorm = accnorm + (std::conj(psi[j]) * psi[j]).real();
        }
        const double nconv = std::fabs(norm - accnorm) / norm;          /* OCheby.cpp:179-182 */
        if (nconv < c.prec && c.offset.imag() == 0) break;
        std::swap(k1, k0);                                              /* OCheby.cpp:191-193 */
    }
    return napply;
}

/* OCheby.cpp:200-297, imaginary time, backward (ProgEigen.cpp:691-694:
 * ImaginaryTime() then Backward(), exponent -dt, OPropagator.h:82-94) */
int cheby_imag_step(Cheby &c, int step, cplx *psi) {
    const Op &H = op_at(c.ham);
    const long n = op_size(H) * c.S;
    const cplx exponent = -c.dt;
    const cplx exp_ = exponent / c.dt;
    std::vector<cplx> ket0(psi, psi + n), ket1(n), ket2(n), buf(n);
    int napply = 0;
    double norm_last = 0;

    op_apply(H, step, psi, ket1.data());
    ++napply;
    for (long j = 0; j < n; ++j) ket1[j] += -1.0 * c.offset * psi[j];   /* AddElements */
    const cplx f1 = exp_ / c.scaling;
    for (long j = 0; j < n; ++j) ket1[j] *= f1;                       /* MultElements */
    const cplx ph = std::exp(exponent * c.offset);
    const cplx c0 = c.coeff[0] * ph, c1 = c.coeff[1] * ph;
    for (long j = 0; j < n; ++j) psi[j] *= c0;
    for (long j = 0; j < n; ++j) buf[j] = ket1[j] * c1;
    for (long j = 0; j < n; ++j) psi[j] += buf[j];

    cplx *k0 = ket0.data(), *k1 = ket1.data(), *k2 = ket2.data(), *bf = buf.data();
    const cplx exp2 = 2.0 * exp_;
    for (int i = 2; i < c.order; ++i) {
        double accnorm = 0;
        const cplx coeff = c.coeff[i] * ph;
        op_apply(H, step, k1, bf);
        ++napply;
        for (long j = 0; j < n; ++j) {                                  /* OCheby.cpp:270-277 */
            bf[j] = (bf[j] - c.offset * k1[j]) / c.scaling;
            bf[j] *= exp2;
            k2[j] = bf[j] - k0[j];
            k0[j] = k2[j];
            psi[j] += k2[j] * coeff;
            accnorm = accnorm + (std::conj(psi[j]) * psi[j]).real();
        }
        if (std::fabs(accnorm - norm_last) / norm_last < c.prec) break;  /* OCheby.cpp:283-285 */
        norm_last = accnorm;
        std::swap(k1, k0);
    }
    return napply;
}

/* ======================================================================
 * Short iterative Lanczos (OSIL.cpp:79-190): real time forward (exponent
 * -i dt) or imaginary time backward (-dt)
 * ==================================================================== */
struct SIL {
    int ham = -1;
    int S = 1;
    double dt = 0;
    int order = 20;     /* OSIL.h:11 SIL_DEF_ORDER */
    double err = 1e-12; /* OSIL.h:12 SIL_DEF_ERR */
    bool imag = false;
    bool backward = false;   /* real time backward: exponent +i dt */
};
std::vector<SIL> g_sil;

int sil_step(const SIL &c, int step, cplx *psi) {
    const Op &H = op_at(c.ham);
    const Grid &g = g_grids.at(op_grid(H));
    const long n = op_size(H) * c.S;
    const int order = c.order;
    auto norm = [&](const cplx *v) { return std::sqrt(overlap(g, c.S, v, v).real()); };
    std::vector<std::vector<cplx>> q(order, std::vector<cplx>(n));
    std::vector<cplx> buf0(n);
    std::vector<double> alpha(order, 0.0), beta(order, 0.0);
    int napply = 0;
    const double psinorm = norm(psi);
    q[0].assign(psi, psi + n);
    op_apply(H, step, q[0].data(), buf0.data());                      /* H q_0 */
    ++napply;
    alpha[0] = overlap(g, c.S, q[0].data(), buf0.data()).real();
    q[1] = buf0;
    for (long j = 0; j < n; ++j) q[1][j] -= q[0][j] * alpha[0];
    beta[0] = norm(q[1].data());
    for (long j = 0; j < n; ++j) q[1][j] *= 1 / beta[0];
    int convorder = 2, it;
    for (it = 2; it < order; it++) {                                  /* OSIL.cpp:110-160 */
        double err = std::pow(c.dt, it - 2) * beta[0];
        for (int i = 1; i < it - 2; i++) err *= beta[i] / double(i);
        err = std::fabs(err);
        convorder++;
        op_apply(H, step, q[it - 1].data(), q[it].data());
        ++napply;
        alpha[it - 1] = overlap(g, c.S, q[it - 1].data(), q[it].data()).real();
        for (long j = 0; j < n; ++j) q[it][j] -= q[it - 1][j] * alpha[it - 1];
        for (long j = 0; j < n; ++j) q[it][j] -= q[it - 2][j] * beta[it - 2];
        beta[it - 1] = norm(q[it].data());
        for (long j = 0; j < n; ++j) q[it][j] *= 1 / beta[it - 1];
        if (err < c.err * psinorm) break;
    }
    {                                                                  /* H->Expec(q_{it-1}) */
        std::vector<cplx> hq(n);
        op_apply(H, step, q[it - 1].data(), hq.data());
        ++napply;
        alpha[it - 1] = overlap(g, c.S, q[it - 1].data(), hq.data()).real();
    }
    /* DiagLZB (OSIL.cpp:168-190): the order x order tridiagonal matrix
     * (entries not reached stay 0), Z e^{c Lambda} Z^T e_0 */
    std::vector<double> T((size_t)order * order, 0.0), ev, Z;
    for (int i = 0; i < order; ++i) {
        T[(size_t)i * order + i] = alpha[i];
        if (i + 1 < order) T[(size_t)i * order + i + 1] = T[(size_t)(i + 1) * order + i] = beta[i];
    }
    sym_eigen(order, T, ev, &Z);
    const cplx expo = c.imag ? cplx(-c.dt, 0) : cplx(0, c.backward ? c.dt : -c.dt);
    std::vector<cplx> vect(order, 0.0);
    for (int k = 0; k < order; ++k) {
        const cplx w = std::exp(expo * ev[k]) * Z[(size_t)0 * order + k];
        for (int i = 0; i < order; ++i) vect[i] += Z[(size_t)i * order + k] * w;
    }
    for (long j = 0; j < n; ++j) psi[j] *= vect[0];
    for (int i = 1; i < convorder; ++i)
        for (long j = 0; j < n; ++j) psi[j] += q[i][j] * vect[i];
    return napply;
}

/* ======================================================================
 * Arnoldi (OArnoldi.cpp:64-190)
 * ==================================================================== */
struct Arnoldi {
    int ham = -1;
    int S = 1;
    double dt = 0;
    int maxorder = 20;   /* OArnoldi.h:11 ARNOLDI_DEF_ORDER */
    double conv = 1e-12; /* OArnoldi.h:13 ARNOLDI_CONV */
    bool imag = false;
    bool backward = false;   /* real time backward: exponent +i dt */
    int order = 20;      /* _order persists between steps (OArnoldi.cpp:44) */
};
std::vector<Arnoldi> g_arnoldi;

/* exp(A) e_0 for a small complex matrix: scaling and squaring with a Taylor
 * series summed to machine precision (independent of the library's Pade) */
std::vector<cplx> expm_col0(int n, std::vector<cplx> A) {
    double nrm = 0;
    for (int j = 0; j < n; ++j) {
        double c = 0;
        for (int i = 0; i < n; ++i) c += std::abs(A[(size_t)i * n + j]);
        nrm = std::max(nrm, c);
    }
    int sq = 0;
    while (nrm > 0.5) {
        nrm /= 2;
        ++sq;
    }
    const double f = std::ldexp(1.0, -sq);
    for (auto &a : A) a *= f;
    auto mul = [&](const std::vector<cplx> &X, const std::vector<cplx> &Y) {
        std::vector<cplx> Z((size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < n; ++k)
                for (int j = 0; j < n; ++j) Z[(size_t)i * n + j] += X[(size_t)i * n + k] * Y[(size_t)k * n + j];
        return Z;
    };
    std::vector<cplx> E((size_t)n * n, 0.0), term((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) E[(size_t)i * n + i] = term[(size_t)i * n + i] = 1.0;
    for (int k = 1; k < 40; ++k) {
        term = mul(term, A);
        for (auto &t : term) t /= double(k);
        for (size_t i = 0; i < E.size(); ++i) E[i] += term[i];
    }
    for (int i = 0; i < sq; ++i) E = mul(E, E);
    std::vector<cplx> c0(n);
    for (int i = 0; i < n; ++i) c0[i] = E[(size_t)i * n];
    return c0;
}

int arnoldi_step(Arnoldi &c, int step, cplx *psi) {
    const Op &H = op_at(c.ham);
    const Grid &g = g_grids.at(op_grid(H));
    const long n = op_size(H) * c.S;
    const int M = c.maxorder;
    auto norm = [&](const cplx *v) { return std::sqrt(overlap(g, c.S, v, v).real()); };
    std::vector<std::vector<cplx>> q(M, std::vector<cplx>(n));
    std::vector<cplx> buf0(n), tHA((size_t)M * M, 0.0);
    int napply = 0;
    const double psinorm = norm(psi);
    q[0].assign(psi, psi + n);
    for (int it = 1; it < M; it++) {
        op_apply(H, step, q[it - 1].data(), buf0.data());
        ++napply;
        for (int j = 0; j <= it - 1; j++) {                           /* modified Gram-Schmidt */
            const cplx h = overlap(g, c.S, q[j].data(), buf0.data());
            tHA[(size_t)j * M + it - 1] = h;
            for (long k = 0; k < n; ++k) buf0[k] += q[j][k] * (-1.0 * h);
        }
        const double nb = norm(buf0.data());
        if (nb < 1e-11) break;                                        /* ARNOLDI_UNDERRUN */
        tHA[(size_t)it * M + it - 1] = nb;
        for (long k = 0; k < n; ++k) buf0[k] *= 1. / nb;
        q[it] = buf0;
        c.order = it + 1;
        cplx err = std::pow(c.dt, it - 2) * tHA[(size_t)1 * M + 0] * psinorm;
        for (int i = 1; i < it - 2; i++) err *= tHA[(size_t)(i + 1) * M + i] / double(i);
        if (std::abs(err) < c.conv) break;
    }
    const int order = c.order;
    op_apply(H, step, q[order - 1].data(), buf0.data());               /* last column */
    ++napply;
    for (int j = 0; j < order; j++) tHA[(size_t)j * M + order - 1] = overlap(g, c.S, q[j].data(), buf0.data());
    const cplx expo = c.imag ? cplx(-c.dt, 0) : cplx(0, c.backward ? c.dt : -c.dt);
    std::vector<cplx> A((size_t)order * order);
    for (int i = 0; i < order; ++i)
        for (int j = 0; j < order; ++j) A[(size_t)i * order + j] = expo * tHA[(size_t)i * M + j];
    const std::vector<cplx> vect = expm_col0(order, A);               /* ZR e^{c Lambda} ZR^-1 e_0 */
    for (long k = 0; k < n; ++k) psi[k] *= vect[0];
    for (int i = 1; i < order; ++i)
        for (long k = 0; k < n; ++k) psi[k] += q[i][k] * vect[i];
    return napply;
}

/* ======================================================================
 * GSPO (OGSPO.cpp)
 * ==================================================================== */
struct GSPO {
    std::vector<int> ops;
    int S = 1;
    double dt = 0;
    bool imag = false;   /* imaginary time, backward: exponent -dt */
    bool backward = false;   /* real time backward: exponent +i dt */
};
std::vector<GSPO> g_gspo;

bool has_transform(const Op &o) {
    /* OGridNablaSq -> TransformFFT; OGridsystem -> none; ODMultistate -> the
     * external (kinetic) transform if any, else per-point diagonalisation
     * (ODMultistate.cpp:35-80, ODMultistate.h:65). */
    if (o.kind == KIN) return true;
    if (o.kind == MS) {
        for (int i = 0; i < o.S; ++i) {
            int id = o.mat[i * o.S + i];
            if (id >= 0 && op_at(id).kind == KIN) return true;
        }
    }
    return false;
}

/* multiply psi by exp(c * D) of one split operator.
 * KIN:   FFT, * exp(c T(k)) / N, IFFT          (OGSPO.cpp:30-47, 114-139; the
 *        1/N of TransformFFT::Normalization is always applied, which is what
 *        QDng does after the initial energy evaluation has run an FFT)
 * POT:   * exp(c V)                               (ODSpace.h:65-69)
 * DIP:   * exp(-c E(t) mu)                        (OGridDipole.cpp:83-87)
 * MS:    diagonal elements per state; coupled potentials: per-point
 *        X exp(c Lambda) X^T (ODMultistate.cpp:109-159, TransformMSD.cpp:31-101) */
void split_apply(const Op &o, int step, cplx c, cplx *psi, int S) {
    const Grid &g = g_grids.at(op_grid(o));
    const long n = g.size();
    auto diag_exp = [&](const Op &d, long i) -> double {
        if (d.kind == POT) return d.v[i];
        if (d.kind == DIP) return -laser_at(d, step) * d.v[i];
        throw OracleError("GSPO: operator has no diagonal representation");
    };
    if (o.kind == KIN) {
        std::vector<cplx> buf(psi, psi + n);
        fftnd(g.nd, g.N, buf.data(), -1);
        const double norm = 1.0 / double(n);
        for (long i = 0; i < n; ++i) buf[i] *= std::exp(c * o.tk[i]) * norm;
        fftnd(g.nd, g.N, buf.data(), +1);
        std::copy(buf.begin(), buf.end(), psi);
        return;
    }
    if (o.kind == POT || o.kind == DIP) {
        for (long i = 0; i < n; ++i) psi[i] *= std::exp(c * diag_exp(o, i));
        return;
    }
    if (o.kind == SUM) throw OracleError("GSPO: Sum is not a diagonal operator");
    /* MS / DMultistate */
    bool fulldiag = true;
    for (int i = 0; i < o.S; ++i)
        for (int j = 0; j < o.S; ++j)
            if (i != j && o.mat[i * o.S + j] >= 0) fulldiag = false;
    if (has_transform(o) || fulldiag) {
        for (int s = 0; s < o.S; ++s) {
            int id = o.mat[s * o.S + s];
            if (id < 0) continue;  /* empty diagonal keeps exp(0) = 1 */
            split_apply(op_at(id), step, c, psi + s * n, 1);
        }
        return;
    }
    /* per grid point diagonalisation; g_plant == 1 swaps X and X^T (a
     * planted mistake for the tests, never set otherwise) */
    const bool swapx = g_plant == 1;
    const int M = o.S;
    std::vector<double> W(M * M), ev, X;
    std::vector<cplx> a(M), b(M);
    for (long i = 0; i < n; ++i) {
        for (int k = 0; k < M; ++k)
            for (int l = 0; l < M; ++l) {
                int id = o.mat[k * M + l];
                W[k * M + l] = id >= 0 ? diag_exp(op_at(id), i) : 0.0;
            }
        sym_eigen(M, W, ev, &X);
        for (int s = 0; s < M; ++s) a[s] = psi[s * n + i];
        for (int r = 0; r < M; ++r) { /* forward: X^T psi */
            cplx t = 0;
            for (int s = 0; s < M; ++s) t += (swapx ? X[r * M + s] : X[s * M + r]) * a[s];
            b[r] = t * std::exp(c * ev[r]);
        }
        for (int r = 0; r < M; ++r) { /* backward: X psi */
            cplx t = 0;
            for (int s = 0; s < M; ++s) t += (swapx ? X[s * M + r] : X[r * M + s]) * b[s];
            psi[r * n + i] = t;
        }
    }
    (void)S;
}

void gspo_step(const GSPO &p, int step, cplx *psi) {
    const cplx I(0, 1);
    const cplx exponent = p.imag ? cplx(-p.dt, 0) : (p.backward ? I : -I) * p.dt;   /* OPropagator.h:62-108 */
    const int L = (int)p.ops.size();
    /* OGSPO.cpp:119-138: ops n-1..1 with c/2, then 0 (full) .. n-1 (c/2) */
    for (int i = L - 1; i > 0; --i) split_apply(op_at(p.ops[i]), step, exponent * 0.5, psi, p.S);
    for (int i = 0; i < L; ++i) split_apply(op_at(p.ops[i]), step, i == 0 ? exponent : exponent * 0.5, psi, p.S);
}

}  // namespace

/* ======================================================================
 * C ABI
 * ==================================================================== */
extern "C" {

const char *orc_last_error(void) { return g_err.c_str(); }

/* Test hook: plant a known mistake in the restatement so that a test can
 * show that its check catches it.  1: transposed per-point eigenvectors in
 * the coupled GSPO split.  0 restores the restatement. */
int orc_plant_mistake(int code) {
    g_plant = code;
    return 0;
}

/* the operator holds a negative imaginary potential */
bool has_nip(const Op &o) {
    if (o.kind == NIP) return true;
    for (int k : o.kids)
        if (has_nip(op_at(k))) return true;
    for (int k : o.mat)
        if (k >= 0 && has_nip(op_at(k))) return true;
    return false;
}

/* OPropagator::Backward() / Forward() (OPropagator.h:62-100) on a real-time
 * propagator: kind 0 Chebyshev, 1 SIL, 2 Arnoldi, 3 GSPO.  Backward real time
 * has the exponent +i dt; the operators here are real, so their complex
 * conjugates (Operator::Conj, set by Backward) are themselves.  Hamiltonians
 * with a negative imaginary potential (not self-conjugate) are refused. */
int orc_set_backward(int kind, int id, int backward) {
    return guarded([&] {
        auto chk = [&](bool imag) {
            if (imag) throw OracleError("imaginary-time propagators keep their direction");
        };
        auto chk_h = [&](int ham) {
            if (has_nip(op_at(ham)))
                throw OracleError("backward propagation with a negative imaginary potential is not supported");
        };
        switch (kind) {
        case 0: chk(g_cheby.at(id).imag); chk_h(g_cheby.at(id).ham); g_cheby.at(id).backward = backward != 0; break;
        case 1: chk(g_sil.at(id).imag); chk_h(g_sil.at(id).ham); g_sil.at(id).backward = backward != 0; break;
        case 2: chk(g_arnoldi.at(id).imag); chk_h(g_arnoldi.at(id).ham); g_arnoldi.at(id).backward = backward != 0; break;
        case 3:
            chk(g_gspo.at(id).imag);
            for (int o : g_gspo.at(id).ops) chk_h(o);
            g_gspo.at(id).backward = backward != 0;
            break;
        default: throw OracleError("unknown propagator kind");
        }
        return 0;
    });
}

/* FFT backend of the oracle: 0 its own transform, 1 the FFTW3 API through
 * MKL's libmkl_rt.so (or libfftw3) when it loads; returns the backend in
 * use.  For the CPU baseline timing only; the checks use backend 0. */
int orc_set_fft_backend(int which) {
    return guarded([&] {
        fftw3::clear_plans();
        fftw3::g_fftw.on = which == 1 && fftw3::load();
        return fftw3::g_fftw.on ? 1 : 0;
    });
}

int orc_set_threads(int n) {
    if (n > 0) {
        fftw3::g_fftw.threads = n;
        if (fftw3::g_fftw.set_threads) fftw3::g_fftw.set_threads(n);
        if (fftw3::g_fftw.lib) fftw3::clear_plans();   /* plans carry their thread count */
    }
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

int orc_fft(int ndims, const int32_t *N, double *data, int sign) {
    return guarded([&] {
        if (ndims < 1 || ndims > 4) throw OracleError("ndims out of range");
        int n[4];
        for (int d = 0; d < ndims; ++d) n[d] = N[d];
        fftnd(ndims, n, reinterpret_cast<cplx *>(data), sign < 0 ? -1 : 1);
        return 0;
    });
}

int orc_bessel_j(int n, double x, double *out, int *nz) {
    return guarded([&] {
        if (n < 1) throw OracleError("n must be >= 1");
        std::vector<double> b;
        int z = bessel_j(n, x, b);
        std::copy(b.begin(), b.end(), out);
        if (nz) *nz = z;
        return 0;
    });
}

int orc_grid(int ndims, const int32_t *N, const double *xmin, const double *xmax) {
    return guarded([&] {
        if (ndims < 1 || ndims > 4) throw OracleError("ndims out of range");
        Grid g;
        g.nd = ndims;
        for (int d = 0; d < ndims; ++d) {
            if (N[d] < 2) throw OracleError("grid needs at least 2 points per dimension");
            if (xmax[d] - xmin[d] <= 0) throw OracleError("Zero length grid defined");
            g.N[d] = N[d];
            g.xmin[d] = xmin[d];
            g.xmax[d] = xmax[d];
        }
        g_grids.push_back(g);
        return (int)g_grids.size() - 1;
    });
}

void orc_reset(void) {
    g_grids.clear();
    g_ops.clear();
    g_cheby.clear();
    g_gspo.clear();
}

int orc_op_kinetic(int grid, const double *mass) {
    return guarded([&] {
        const Grid &g = g_grids.at(grid);
        Op o;
        o.kind = KIN;
        o.grid = grid;
        bool eff = false;
        for (int d = 0; d < g.nd; ++d) {    /* OGridNablaSq.cpp:46-55 */
            if (mass[d] == 0) throw OracleError("Zero mass defined");
            if (mass[d] > 0) eff = true;
            o.mass.push_back(mass[d]);
        }
        if (!eff) throw OracleError("Nabla^2 Operator is empty (no masses defined)");
        o.tk = kinetic_table(g, o.mass);
        g_ops.push_back(o);
        return (int)g_ops.size() - 1;
    });
}

int orc_op_potential(int grid, const double *V) {
    return guarded([&] {
        const Grid &g = g_grids.at(grid);
        Op o;
        o.kind = POT;
        o.grid = grid;
        o.v.assign(V, V + g.size());
        g_ops.push_back(o);
        return (int)g_ops.size() - 1;
    });
}

/* OGobbler::_Init (OGobbler.cpp:32-82) with its parameters of
 * OGobbler::Init (OGobbler.cpp:85-164): per dimension d an optional
 * Butterworth low pass of edge lp[d] and high pass of edge rp[d] (NaN =
 * absent; edges in grid points x = (edge - xmin) / dx, checked against the
 * grid as there), filters f(i) = 1 / (1 + (i / x)^(2 order)) and
 * 1 / (1 + (x / i)^(2 order)) (math_functions.h:40-69) applied along the
 * dimension: the low pass multiplies, the high pass multiplies (nip = 0,
 * start value 1) or adds (nip = 1, start value 0); finally times gain.
 * nip = 1 gives the negative imaginary potential -i v (OGobbler.cpp:195-204);
 * the bounds are Emax = gain, Emin = -i gain (nip) or 0 (OGobbler.cpp:183-192). */
int orc_op_gobbler(int grid, const double *lp, const double *rp, int order, double gain, int nip) {
    return guarded([&] {
        const Grid &g = g_grids.at(grid);
        if (order < 1) throw OracleError("Filter order to small");
        Op o;
        o.kind = nip ? NIP : POT;
        o.grid = grid;
        o.fixed = true;
        o.fmax = cplx(gain, 0);
        o.fmin = nip ? cplx(0, -gain) : cplx(0, 0);
        o.hermitian = !nip;
        const long n = g.size();
        o.v.assign(n, nip ? 0.0 : 1.0);
        long stride = 1;
        for (int d = 0; d < g.nd; ++d) {
            const int Nd = g.N[d];
            for (int side = 0; side < 2; ++side) {
                const double edge = side == 0 ? lp[d] : rp[d];
                if (std::isnan(edge)) continue;
                const double x = (edge - g.xmin[d]) / g.dx(d);
                if (int(x) > Nd || int(x) < 0) throw OracleError("Center of filter edge out of the grid");
                std::vector<double> f(Nd);
                for (int i = 0; i < Nd; ++i)
                    f[i] = side == 0 ? 1 / (1 + std::pow(double(i) / x, 2 * double(order)))
                                     : 1 / (1 + std::pow(x / double(i), 2 * double(order)));
                for (long j = 0; j < n; ++j) {
                    const double fv = f[(j / stride) % Nd];
                    if (side == 1 && nip && g_plant != 4)   /* planted: the high pass multiplies */
                        o.v[j] += fv;
                    else
                        o.v[j] *= fv;
                }
            }
            stride *= Nd;
        }
        if (gain != 1)
            for (double &x : o.v) x *= gain;
        g_ops.push_back(o);
        return (int)g_ops.size() - 1;
    });
}

int orc_op_dipole(int grid, const double *mu, const double *laser, int nt, double scale) {
    return guarded([&] {
        const Grid &g = g_grids.at(grid);
        if (nt < 1) throw OracleError("Laser field has zero size");
        Op o;
        o.kind = DIP;
        o.grid = grid;
        o.v.assign(mu, mu + g.size());
        o.laser.assign(laser, laser + nt);
        o.scale = scale;
        g_ops.push_back(o);
        return (int)g_ops.size() - 1;
    });
}

/* Laser::Set over the whole field (ProgOCT.cpp:566-567): the raw samples of
 * a GridDipole operator, nt at most its length */
int orc_op_set_laser(int op, const double *laser, int nt) {
    return guarded([&] {
        Op &o = op_at(op);
        if (o.kind != DIP) throw OracleError("not a GridDipole operator");
        if (nt < 0 || nt > (int)o.laser.size()) throw OracleError("laser longer than the operator's field");
        std::copy(laser, laser + nt, o.laser.begin());
        return 0;
    });
}

int orc_op_sum(const int *ops, int n) {
    return guarded([&] {
        if (n < 1) throw OracleError("Sum Operator is empty");
        Op o;
        o.kind = SUM;
        for (int i = 0; i < n; ++i) {
            op_at(ops[i]);
            o.kids.push_back(ops[i]);
        }
        g_ops.push_back(o);
        return (int)g_ops.size() - 1;
    });
}

int orc_op_multistate(int nstates, const int *rows, const int *cols, const int *ops, int n, int unity,
                      int dmulti) {
    return guarded([&] {
        if (nstates < 1 || nstates > 64) throw OracleError("Multistate: invalid number of states");
        Op o;
        o.kind = MS;
        o.S = nstates;
        o.unity = unity != 0;
        o.dmulti = dmulti != 0;
        o.mat.assign(nstates * nstates, -1);
        o.hermitian = true;   /* no "nonhermitian" parameter */
        for (int k = 0; k < n; ++k) {    /* OMultistate.cpp:34-53 */
            int r = rows[k], c = cols[k];
            if (r < 0 || c < 0 || r >= nstates || c >= nstates) throw OracleError("Multistate index out of range");
            if (o.mat[r * nstates + c] >= 0) throw OracleError("Multistate matrix element is already in use");
            op_at(ops[k]);
            o.mat[r * nstates + c] = ops[k];
            if (o.mat[r * nstates + c] >= 0 && o.mat[c * nstates + r] >= 0 && r != c) o.hermitian = false;
        }
        if (o.hermitian) /* OMultistate.cpp:120-131: mirror by pointer */
            for (int i = 0; i < nstates; ++i)
                for (int j = 0; j <= i; ++j) {
                    if (i != j && o.mat[j * nstates + i] >= 0) o.mat[i * nstates + j] = o.mat[j * nstates + i];
                    if (o.mat[i * nstates + j] >= 0 && i != j) o.mat[j * nstates + i] = o.mat[i * nstates + j];
                }
        g_ops.push_back(o);
        return (int)g_ops.size() - 1;
    });
}

int orc_op_bounds(int op, double *emin, double *emax) {
    return guarded([&] {
        const Op &o = op_at(op);
        *emin = op_emin(o).real();
        *emax = op_emax(o).real();
        return 0;
    });
}

int orc_op_bounds_c(int op, double *emin, double *emax) {
    return guarded([&] {
        const Op &o = op_at(op);
        const cplx a = op_emin(o), b = op_emax(o);
        emin[0] = a.real();
        emin[1] = a.imag();
        emax[0] = b.real();
        emax[1] = b.imag();
        return 0;
    });
}

int orc_op_apply(int op, int step, const double *src, double *dest) {
    return guarded([&] {
        const Op &o = op_at(op);
        op_apply(o, step, reinterpret_cast<const cplx *>(src), reinterpret_cast<cplx *>(dest));
        return 0;
    });
}

int orc_op_expec(int op, int step, const double *psi, int nstates, double *out) {
    return guarded([&] {
        const Op &o = op_at(op);
        const Grid &g = g_grids.at(op_grid(o));
        if (nstates != op_states(o)) throw OracleError("operator and wave function differ in number of states");
        std::vector<cplx> hp(g.size() * nstates);
        const cplx *p = reinterpret_cast<const cplx *>(psi);
        op_apply(o, step, p, hp.data());
        *out = overlap(g, nstates, p, hp.data()).real();   /* Operator.h:226-252 */
        return 0;
    });
}

int orc_overlap(int grid, int nstates, const double *a, const double *b, double out[2]) {
    return guarded([&] {
        cplx c = overlap(g_grids.at(grid), nstates, reinterpret_cast<const cplx *>(a),
                         reinterpret_cast<const cplx *>(b));
        out[0] = c.real();
        out[1] = c.imag();
        return 0;
    });
}

int orc_expec_position(int grid, int dim, const double *psi, double *out) {
    return guarded([&] { /* OGridPosition.cpp: x = xmin + dx*n along dim */
        const Grid &g = g_grids.at(grid);
        const cplx *p = reinterpret_cast<const cplx *>(psi);
        long stride = 1;
        for (int d = 0; d < dim; ++d) stride *= g.N[d];
        double s = 0;
        for (long i = 0; i < g.size(); ++i) {
            double x = g.xmin[dim] + g.dx(dim) * double((i / stride) % g.N[dim]);
            s += std::norm(p[i]) * x;
        }
        *out = s * g.weight();
        return 0;
    });
}

int orc_expec_momentum(int grid, int dim, const double *psi, double *out) {
    return guarded([&] { /* OGridNabla.cpp:76-95 (momentum), Kspace.cpp:64-88 */
        const Grid &g = g_grids.at(grid);
        const long n = g.size();
        const cplx *p = reinterpret_cast<const cplx *>(psi);
        std::vector<cplx> buf(p, p + n);
        fftnd(g.nd, g.N, buf.data(), -1);
        const int Nx = g.N[dim];
        std::vector<double> k1(Nx);
        const double dp = g.dk(dim);
        if (Nx % 2 == 0) {
            for (int i = 0; i < Nx / 2; ++i) {
                k1[i] = double(i) * dp;
                k1[Nx - i - 1] = -1 * double(i + 1) * dp;
            }
        } else {
            k1[0] = 0;
            for (int i = 1; i < (Nx + 1) / 2; ++i) {
                k1[i] = double(i) * dp;
                k1[Nx - i] = -1 * double(i) * dp;
            }
        }
        long stride = 1;
        for (int d = 0; d < dim; ++d) stride *= g.N[d];
        for (long i = 0; i < n; ++i) buf[i] *= k1[(i / stride) % Nx] / double(n);
        fftnd(g.nd, g.N, buf.data(), +1);
        *out = overlap(g, 1, p, buf.data()).real();
        return 0;
    });
}

int orc_cheby_create(int ham, int nstates, double dt, int order, double prec, double scaling,
                     orc_cheby_info *info) {
    return guarded([&] { return cheby_create(ham, nstates, dt, order, prec, scaling, info); });
}

int orc_cheby_create_imag(int ham, int nstates, double dt, int order, double prec, double scaling,
                          orc_cheby_info *info) {
    return guarded([&] { return cheby_create(ham, nstates, dt, order, prec, scaling, info, true); });
}

int orc_bessel_i_scaled(int n, double x, double *out, int *nz) {
    return guarded([&] {
        std::vector<double> v;
        const int z = bessel_i_scaled(n, x, v);
        std::copy(v.begin(), v.end(), out);
        if (nz) *nz = z;
        return 0;
    });
}

int orc_sil_create(int ham, int nstates, double dt, int order, double err, int imag) {
    return guarded([&] {
        SIL c;
        c.ham = ham;
        c.S = nstates;
        c.dt = dt;
        if (order > 0) c.order = order;
        if (c.order < 3) throw OracleError("SIL: order doesn't make sense");
        if (err > 0) c.err = err;
        c.imag = imag != 0;
        if (op_states(op_at(ham)) != nstates) throw OracleError("Hamiltonian and wave function differ in number of states");
        g_sil.push_back(c);
        return (int)g_sil.size() - 1;
    });
}

int orc_sil_step(int prop, int step, double *psi, int *n_hpsi) {
    return guarded([&] {
        int n = sil_step(g_sil.at(prop), step, reinterpret_cast<cplx *>(psi));
        if (n_hpsi) *n_hpsi = n;
        return 0;
    });
}

int orc_arnoldi_create(int ham, int nstates, double dt, int order, double conv, int imag) {
    return guarded([&] {
        Arnoldi c;
        c.ham = ham;
        c.S = nstates;
        c.dt = dt;
        if (order > 0) c.maxorder = order;
        if (c.maxorder < 3) throw OracleError("Arnoldi: order doesn't make sense");
        c.order = c.maxorder;
        if (conv > 0) c.conv = conv;
        c.imag = imag != 0;
        g_arnoldi.push_back(c);
        return (int)g_arnoldi.size() - 1;
    });
}

int orc_arnoldi_step(int prop, int step, double *psi, int *n_hpsi) {
    return guarded([&] {
        int n = arnoldi_step(g_arnoldi.at(prop), step, reinterpret_cast<cplx *>(psi));
        if (n_hpsi) *n_hpsi = n;
        return 0;
    });
}

int orc_cheby_coeffs(int prop, double *coeff) {
    return guarded([&] {
        const Cheby &c = g_cheby.at(prop);
        std::copy(c.coeff.begin(), c.coeff.end(), coeff);
        return 0;
    });
}

int orc_cheby_step(int prop, int step, double *psi, int *n_hpsi) {
    return guarded([&] {
        Cheby &c = g_cheby.at(prop);
        int n = c.imag ? cheby_imag_step(c, step, reinterpret_cast<cplx *>(psi))
                       : cheby_step(c, step, reinterpret_cast<cplx *>(psi));
        if (n_hpsi) *n_hpsi = n;
        return 0;
    });
}

int orc_gspo_create_ex(const int *ops, int nops, int nstates, double dt, int imag) {
    return guarded([&] {
        if (nops < 1) throw OracleError("GSPO is empty!");
        if (dt == 0) throw OracleError("No time step defined");
        GSPO p;
        for (int i = 0; i < nops; ++i) {
            const Op &o = op_at(ops[i]);
            if (o.kind == SUM) throw OracleError("GSPO: Operator has no diagonal representation");
            if (op_states(o) != nstates) throw OracleError("GSPO operator and wave function differ in states");
            p.ops.push_back(ops[i]);
        }
        p.S = nstates;
        p.dt = dt;
        p.imag = imag != 0;
        g_gspo.push_back(p);
        return (int)g_gspo.size() - 1;
    });
}

int orc_gspo_create(const int *ops, int nops, int nstates, double dt) {
    return orc_gspo_create_ex(ops, nops, nstates, dt, 0);
}

int orc_gspo_step(int prop, int step, double *psi) {
    return guarded([&] {
        gspo_step(g_gspo.at(prop), step, reinterpret_cast<cplx *>(psi));
        return 0;
    });
}

}  // extern "C"
