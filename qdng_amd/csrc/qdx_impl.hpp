// Internal declarations shared by the host runtime (qdx_host.cpp) and the
// gfx950 kernels (qdx_kernels.hip).  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace qdx {

constexpr int kMaxStates = 4;      // states handled by the fused kernels
constexpr int kMaxStages = 16;     // radix stages of one line transform
constexpr int kMaxLine = 4096;     // longest line transform held in LDS
constexpr int kMaxDims = 4;        // grid dimensions (GridSystem's MAX_DIMS)

// ---------------------------------------------------------------------------
// Radix plan of one line length L = 2^a 3^b: radix-8 stages first, then one
// radix-4 or radix-2 stage, then radix-3 stages.  `tw` points to the device
// table exp(-2 pi i q / L), q in [0, L), in the precision of the kernel.
struct LinePlan {
    int L = 0;
    int nst = 0;
    int rad[kMaxStages] = {};
    const void *tw = nullptr;
    int tw_lds = 1;    // kernels stage the table in LDS (set by the launcher)
    int ppt = 8;       // points per thread of the line kernels (8, or 4: first
                       // and last radix 4 or 2)
};

// Grid as the kernels see it: N[d] points in dimension d, dimension 0
// contiguous.  Unused dimensions have N = 1.
struct GridDev {
    int nd = 1;
    int N[kMaxDims] = {1, 1, 1, 1};
    long size = 1;
};

// Element stride of dimension d (the product of the sizes before it).
__host__ __device__ inline long dim_stride(const GridDev &g, int d) {
    long s = 1;
    for (int e = 0; e < d; ++e) s *= g.N[e];
    return s;
}
// Lines along dimension dim >= 1 of a column tile: besides dimension 0 (the
// tile's columns) they run over the other dimensions, whose index o in
// [0, col_nother) counts them in increasing dimension order.
__host__ __device__ inline int col_nother(const GridDev &g, int dim) {
    int n = 1;
    for (int e = 1; e < g.nd; ++e)
        if (e != dim) n *= g.N[e];
    return n;
}
// element offset of "other" index o, and the index of each other dimension in k[]
// (loops unrolled over kMaxDims: k[] keeps constant indices, in registers)
__host__ __device__ inline long col_other(const GridDev &g, int dim, int o, int (&k)[kMaxDims]) {
    long off = 0, stride = g.N[0];
#pragma unroll
    for (int e = 1; e < kMaxDims; ++e) {
        if (e < g.nd && e != dim) {
            const int i = o % g.N[e];
            o /= g.N[e];
            k[e] = i;
            off += (long)i * stride;
        }
        stride *= g.N[e];
    }
    return off;
}

// Hamiltonian H = T + W(x) - E(t) mu(x) as the fused kernels apply it to a
// multistate vector phi[S][N] (OSum / OMultistate of grid terms):
//   (H phi)_i = T_i phi_i + sum_j (W_ij(x) + unity_i delta_ij) phi_j
//               + sum_j (-E_ij mu_ij(x)) phi_j
// tk[i][d] is the per-dimension kinetic table of state i (nullptr: no kinetic
// term), its value at k summed over dimensions and multiplied by inv_n is
// T_i(k)/N as in Diff::DnDxn.  W[i][j] / mu[i][j] are device arrays of the
// kernel precision or nullptr.  Edip[i][j] = -laser[step] * scale.
template <class R>
struct HamDev {
    int S = 1;
    double inv_n = 1.0;
    double inv_nd[kMaxDims] = {1.0, 1.0, 1.0, 1.0};   // 1 / N_d: T = sum_d T_d acts as 1D filters
    const double *tk[kMaxStates][kMaxDims] = {};
    const R *W[kMaxStates][kMaxStates] = {};
    const R *mu[kMaxStates][kMaxStates] = {};
    double Edip[kMaxStates][kMaxStates] = {};
    double unity[kMaxStates] = {};
    // negative imaginary potential of each diagonal element: H gets -i Wi(x)
    // (OGobbler with nip, OGobbler.cpp:195-204); null when absent
    const R *Wi[kMaxStates] = {};
};

// Per-propagator device control block for the Chebyshev series.
struct ChebyCtl {
    double norm0;        // raw_norm(psi) at the start of the step
    double accnorm;      // sum |psi|^2 of the order that ended the series
    int done;            // series converged (OCheby.cpp:179-182)
    int last_order;      // index i of the last executed recurrence iteration
    unsigned pad0;
    // batched jobs: 1 + the step whose start state had a norm that is not a
    // number (Reporter.cpp:112-113); 0 while the job is sound.  A failed job
    // is skipped by every kernel from then on.
    int failed;
};

// Break test of the previous order, evaluated at the start of a kernel from
// the per-block partial sums that the previous row kernel stored.  Every
// kernel sums the partials in the same fixed order (lane-strided in wave 0,
// then a fixed shuffle tree), so all kernels reach the same decision.
//
// Imaginary time (OCheby.cpp:283-285) tests the relative change against the
// order before, |acc(k_prev) - acc(k_prev-1)| / acc(k_prev-1) < |prec|; it
// is marked by prec < 0, and norm0 then points at the partials of order
// k_prev - 1 (same count nb).  (The struct keeps its size: kernel arguments
// that grew by a few bytes slowed the hot kernels measurably.)
struct ConvCheck {
    const double *part = nullptr;   // partials of order k_prev
    const double *norm0 = nullptr;  // sum |psi|^2 at step start (ChebyCtl::norm0)
    int nb = 0;
    int k_prev = 0;                 // < 2 (< 3 relative): nothing to test
    double prec = 0;
};
constexpr int kMaxPartials = 8192;

// Scratch of the allocation-free expectation value (qdx_ham_expec_noalloc):
// a file-scope device array of kNoAllocBytes per GPU holding H psi, the
// kinetic pass buffer X (grids with nd >= 2), reduction partials and the
// per-state results, for at most kNoAllocPoints = S * N grid points.
constexpr long kNoAllocPoints = 524288;
constexpr size_t kNoAllocBytes = (size_t)2 * kNoAllocPoints * 16 + (size_t)kMaxPartials * 8 + 4096;
// Looks up the array's device address on `device` (which must be current)
// once and keeps it; throws std::runtime_error if the lookup fails.
void *noalloc_scratch_setup(int device);
// The address found by noalloc_scratch_setup, or nullptr before it ran.
void *noalloc_scratch(int device);

// Chebyshev scalars for one order (OCheby.cpp:97-122).
struct ChebyOrder {
    int k;               // order index (1 for the prologue)
    double offset;       // Re offset (energy shift)
    double inv_scaling;  // 1 / R
    double c0r, c0i;     // order 1: coeff[0] * exp(-i offset dt)
    double ckr, cki;     // coeff[k] * exp(-i offset dt)
    double prec;         // break threshold
    // psi update at this order: the last `nterm` terms (k-2, k-1, k), 0 for
    // none (the fixed-length FP32 series updates psi every third order)
    int nterm = 1;
    double cm1r = 0, cm1i = 0;   // coeff[k-1] * exp(-i offset dt)
    double cm2r = 0, cm2i = 0;   // coeff[k-2] * exp(-i offset dt)
    // Im offset (OCheby.cpp:407-408, non-zero with a negative imaginary
    // potential): b = (H - offset_r - i offset_i) phi_{k-1} / R; the kernel
    // instance with the imaginary diagonal (Wi, offset_i) is chosen on the host
    double offset_i = 0;
    // (imaginary time, backward, exponent -dt, OCheby.cpp:200-297, is a
    // separate kernel instance: phi_1 = -b, phi_k = -2 b - phi_{k-2} with
    // b = (H - offset) phi_{k-1} / R; the coefficients carry exp(-dt offset))
};

// One x-space split operator of GSPO: exp(c * W(x)) with c = -i tau, W the
// S x S real symmetric matrix of this operator at each grid point.
template <class R>
struct XSplitDev {
    int S = 1;
    int coupled = 0;     // off-diagonal elements present
    double tau = 0;      // dt * sf
    const R *V[kMaxStates][kMaxStates] = {};   // potential arrays
    const R *mu[kMaxStates][kMaxStates] = {};  // dipole arrays
    double Edip[kMaxStates][kMaxStates] = {};  // -laser[step]*scale
    // exp(-i tau W) per grid point, computed once for a time-independent
    // operator (no dipole), FP64 complex as (re, im) pairs, component-major
    // (U + 2 (c n + x)): coupled two states c = 0, 1, 2 for U00, U01, U11,
    // otherwise c = state (the diagonal phases).  nullptr: evaluated per point.
    const double *U = nullptr;
    long n = 0;
};

constexpr int kMaxXOps = 4;
template <class R>
struct XGroupDev {
    int n = 0;
    XSplitDev<R> op[kMaxXOps];
};

// Events attached to the dispatch packet of the next kernel launch (set by the
// host's profiling timer, consumed and cleared by that launch).
struct LaunchTiming {
    hipEvent_t start = nullptr, stop = nullptr;
};
extern thread_local LaunchTiming g_launch_timing;

// ---------------------------------------------------------------------------
// Launchers implemented in qdx_kernels.hip.  All run on `stream`.
template <class R>
struct Launch {
    // psi (complex128) -> phi (complex R); norm0 into ctl; resets ctl.
    // For nd >= 2 also X = FFT_0(phi).
    // returns the number of partial sums written (one per block)
    // cmp (below): phi, X in the compact 40-bit term storage (R double,
    // nd >= 2, S <= 2)
    static int cheby_init(hipStream_t s, const GridDev &g, int S, const LinePlan &p0,
                          const double *psi, R *phi0, R *X, ChebyCtl *ctl, double *partials, bool cmp = false);
    // One order of the series; prologue when ord.k == 1.  T phi_{k-1} is
    // the dimension-0 filter IFFT_0(T_0/N_0 FFT_0 phi_{k-1}) in registers,
    // plus (nd >= 2) the other dimensions' part Y that kin_pass left in X.
    // init_part/nb_init: partials of the init kernel, summed into
    // ctl->norm0 by the prologue (ord.k == 1)
    static int cheby_row(hipStream_t s, const GridDev &g, const HamDev<R> &h, const LinePlan &p0,
                         const ChebyOrder &ord, const ConvCheck &cc, bool imag, const R *phim1, R *phim2, R *X,
                         double *psi, ChebyCtl *ctl, double *partials, const double *init_part,
                         int nb_init, bool cmp = false);
    // a whole Chebyshev step of a single-state 1D line in one block (see
    // k_cheby_line); coef: (re, im) of coeff[k] exp(-i offset dt), k < order
    // offset_i: Im offset (negative imaginary potential)
    static void cheby_line(hipStream_t s, const GridDev &g, const HamDev<R> &h, const LinePlan &p0,
                           const double *coef, int order, double offset, double offset_i, double inv_scaling,
                           double prec, double *psi, ChebyCtl *ctl, double *partials);
    // strided line pass along dimension `dim` (1 or 2) on X, in place.
    // mode 0: forward, 1: inverse, 2: forward * T/N inverse (kinetic),
    // 3: forward * exp(-i tau T)/N inverse (GSPO kinetic split operator),
    // or exp(-tau T)/N with imag (imaginary time)
    // ek (mode 3, real time; nullable): per-state tables exp(-i tau T_d(k))
    // of every dimension (see k_col)
    static void col_pass(hipStream_t s, const GridDev &g, int S, int dim, int mode,
                         const LinePlan &p, const HamDev<R> &h, double tau, R *X,
                         ChebyCtl *ctl, const ConvCheck &cc, int imag = 0, const double *ek = nullptr);
    // kinetic part along dimension `dim` of the separable T = sum_d T_d:
    // Y = IFFT_d(T_d/N_d FFT_d(src)), or Y += that when `accumulate`.  The
    // first pass of an order also evaluates the break test of the previous
    // order and records its outcome in ctl (done, accnorm), which the row
    // kernel of the order then only reads
    static void kin_pass(hipStream_t s, const GridDev &g, int S, int dim, bool accumulate,
                         const LinePlan &p, const HamDev<R> &h, const R *src, R *Y,
                         ChebyCtl *ctl, const ConvCheck &cc, bool cmp = false);
    // dst = H src (both complex128) through the same kernels; X scratch
    static void ham_apply(hipStream_t s, const GridDev &g, const HamDev<R> &h,
                          const LinePlan *plans, const double *src, double *dst, R *X);
    // GSPO: one step; psi complex R (state-major)
    // imag (here and below): imaginary time, exp(-tau D) (FP64 only)
    static void gspo_row_pre(hipStream_t s, const GridDev &g, int S, const LinePlan &p0,
                             const XGroupDev<R> &pre, const R *psi_in, R *X, int imag = 0);
    static void gspo_row_post(hipStream_t s, const GridDev &g, int S, const LinePlan &p0,
                              const XGroupDev<R> &post, const R *X, R *psi_out, int imag = 0);
    // between two steps of a run (psi held in X, FFT_0 applied): IFFT_0,
    // post of step n, pre of step n + 1, FFT_0.  opart (nullable): per block
    // and state sum |psi|^2 and sum conj(psi) psi0 of the state between the
    // steps (2 S nb (re, im) pairs, see k_gspo_row); returns nb
    static int gspo_row_bridge(hipStream_t s, const GridDev &g, int S, const LinePlan &p0,
                               const XGroupDev<R> &post, const XGroupDev<R> &pre, R *X, const R *psi0,
                               double *opart, int imag = 0);
    static void gspo_1d(hipStream_t s, const GridDev &g, int S, const LinePlan &p0,
                        const HamDev<R> &kin, double tau_kin, const XGroupDev<R> &pre,
                        const XGroupDev<R> &post, R *psi, int imag = 0);
    static void gspo_xonly(hipStream_t s, const GridDev &g, int S, const XGroupDev<R> &grp,
                           R *psi, int imag = 0);
    // the table XSplitDev::U of a real-time operator op (3 n or S n complex)
    static void gspo_unitary(hipStream_t s, const GridDev &g, int S, const XSplitDev<R> &op, double *U);
    // X = the unnormalised forward transform of every state's lines along
    // `dim` (src complex128)
    static void line_fft(hipStream_t s, const GridDev &g, int S, int dim, const LinePlan &p, const double *src,
                         R *X);
};

// precision-independent helpers
void launch_convert_c128_to_c64(hipStream_t s, const double *src, float *dst, long n);
// out[2 s] = sum_i w(i) |a_s(i)|^2 (out[2 s + 1] = 0), w the coordinate along
// dim: ktab[i_dim] if given, else base + step i_dim
void launch_wnorm(hipStream_t s, const double *a, long n, int S, const GridDev &g, int dim, double base, double step,
                  const double *ktab, double *partials, double *out);
// y = sa * y + (br + i bi) * x over n complex128 values (x may be null)
void launch_scale_axpy(hipStream_t s, double *y, const double *x, long n, double sa, double br, double bi);
void launch_convert_c64_to_c128(hipStream_t s, const float *src, double *dst, long n);
// Deterministic FP64 reductions over a complex128 vector set [S][n]:
// out[s] = sum_j conj(a_sj) b_sj  (re, im pairs), S <= kMaxStates
void launch_overlap(hipStream_t s, const double *a, const double *b, long n, int S, double *partials,
                    double *out);
// data[i] <- conj(data[i]) for n complex values (complex64 when fp32)
void launch_conj(hipStream_t s, void *data, long n, bool fp32);
// out[2 s], out[2 s + 1] = the sums over b < nbx of partials[2 (s nbx + b)],
// partials[2 (s nbx + b) + 1], in the fixed order of launch_overlap
void launch_sum_partials(hipStream_t s, const double *partials, int nbx, int S, double *out);
void launch_overlap_c64(hipStream_t s, const float *a, const float *b, long n, int S,
                        double *partials, double *out);

}  // namespace qdx
