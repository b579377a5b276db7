// gfx950 kernels of the grid-propagation hot path.
//
// Chebyshev (OCheby::real_time_step, OCheby.cpp:92-197) runs per order as
//   nd == 1 : one kernel  -- phi_{k-1} line into registers, FFT, *T/N, IFFT,
//             fused potential/dipole/coupling multiply, recurrence, FP64
//             accumulation of psi and of sum |psi|^2, device-side break test;
//   nd >= 2 : line passes -- strided FFT/T/IFFT passes along dimensions
//             >= 1 on the scratch X (which holds FFT_0(phi_{k-1})), then the
//             row kernel: IFFT_0, the same fused epilogue, and FFT_0(phi_k)
//             back into X for the next order.
// GSPO (OGSPO.cpp:114-139) uses the same line machinery with the x-space
// split operators exp(-i tau W(x)) evaluated per point in registers.
//
// Thread layout of every line kernel: a line of L = P*TL points is handled by
// TL threads, thread lt owning points lt + q*TL, q = 0..P-1 (see fft_line);
// P = 8, or 4 for the Chebyshev row kernel and kinetic passes of grids too
// small to give every SIMD two waves at 8 points per thread.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "qdx_device.hpp"
#include "qdx_impl.hpp"

namespace qdx {

thread_local LaunchTiming g_launch_timing;


// ---------------------------------------------------------------------------
// helpers

// kinetic value T_i(k) at k-space indices (k0, k1, k2), summed over dimensions
// in dimension order as Diff::_InitKspace does (Diff.cpp:120-139)
template <class R>
__device__ __forceinline__ double kin_sum(const HamDev<R> &h, int st, const GridDev &g, int k0, int k1, int k2) {
    const double *const *tk = h.tk[st];
    double t = 0.0 + tk[0][k0];
    if (g.nd > 1) {
        t += tk[1][k1];
        if (g.nd > 2) t += tk[2][k2];
    }
    return t;
}

// Row-block geometry: rb rows of dimension 0 for each of S states; line
// l = st * rb + rr; P points per thread.  Two padded LDS buffers per line.
template <class R, int P = 8>
struct RowBlock {
    int L, TL, rb, nl, line, lt, st, rr;
    long nrows, row;
    bool active;
    LdsLine<R> ln;
    cx<R> *tw;       // LDS twiddle table after the line buffers (if twl)
    cx<R> *after;    // first free LDS element after the table
    bool twl;        // twiddle table staged in LDS (else read from the plan)
    __device__ RowBlock(const GridDev &g, int rb_, int S, cx<R> *smem, const LinePlan &p) {
        L = g.N[0];
        TL = L / P;
        rb = rb_;
        nl = S * rb;
        nrows = g.size / L;
        line = threadIdx.x / TL;
        lt = threadIdx.x % TL;
        st = line / rb;
        rr = line % rb;
        row = (long)blockIdx.x * rb + rr;
        active = line < nl && row < nrows;
        if (!active) {
            st = 0;
            rr = 0;
        }
        const int lp = L + (L >> 3);
        const int l = active ? line : 0;
        ln.a = smem + (long)l * lp;
        ln.b = smem + (long)(nl + l) * lp;
        ln.es = 1;
        ln.pad = true;
        tw = smem + (long)2 * nl * lp;
        twl = p.tw_lds;
        after = twl ? tw + tw_lds_len(L) : tw;
    }
    // stage the twiddle table (when it lives in LDS)
    __device__ void stage(const LinePlan &p) {
        if (twl) stage_twiddles(tw, p);
    }
    // transform with the LDS or the global twiddle table; both calls inline
    __device__ void fft(cx<R> (&v)[P], const LinePlan &p, int sign) {
        if (twl) fft_line<R, P>(v, ln, p, tw, sign, lt, active, kTwLds);
        else fft_line<R, P>(v, ln, p, reinterpret_cast<const cx<R> *>(p.tw), sign, lt, active, kTwMem);
    }
    __device__ long gidx(int q) const { return row * L + lt + q * TL; }
};

template <class R>
static size_t row_lds_bytes(int L, int nl, bool twl) {
    return ((size_t)2 * nl * (L + L / 8) + (twl ? tw_lds_len(L) : 0)) * sizeof(cx<R>);
}

// ---------------------------------------------------------------------------
// Chebyshev init: phi0 = psi, norm0 = sum |psi|^2, X = FFT_0(phi0) (nd >= 2)
template <class R, int S, bool HAS_X>
__global__ void __launch_bounds__(512) k_cheby_init(GridDev g, LinePlan p0, int rb, const cx<double> *__restrict__ psi,
                                                     cx<R> *__restrict__ phi0, cx<R> *__restrict__ X, ChebyCtl *ctl,
                                                     double *partials) {
    extern __shared__ __align__(16) char smem_raw[];
    __shared__ double red[16];
    RowBlock<R> B(g, rb, S, reinterpret_cast<cx<R> *>(smem_raw), p0);
    const long base = (long)B.st * g.size;
    double acc = 0;
    cx<R> v[8];
    if (B.active) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const cx<double> x = psi[base + B.gidx(q)];
            acc += x.r * x.r + x.i * x.i;
            v[q] = from_d<R>(x);
            phi0[base + B.gidx(q)] = v[q];
        }
    }
    if (HAS_X) {
        B.stage(p0);
        B.fft(v, p0, -1);
        if (B.active) {
#pragma unroll
            for (int q = 0; q < 8; ++q) X[base + B.gidx(q)] = v[q];
        }
    }
    const double part = block_sum(acc, red);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = part;
        if (blockIdx.x == 0) {
            ctl->done = 0;
            ctl->last_order = 1;
        }
    }
}

// ---------------------------------------------------------------------------
// Chebyshev row kernel (see file header).  PRO: prologue (order 1).
// NT: the largest block the instance is launched with (256 leaves twice the
// registers per thread of 512).  P: points per thread.  IMAG: imaginary
// time (the recurrence signs of OCheby::imag_time_step).
template <class R, int S, bool ONE_D, bool PRO, int NT, int P = 8, bool IMAG = false>
__global__ void __launch_bounds__(NT) k_cheby_row(GridDev g, HamDev<R> h, LinePlan p0, ChebyOrder ord,
                                                    ConvCheck cc, int rb, const cx<R> *__restrict__ phim1,
                                                    cx<R> *__restrict__ phim2, cx<R> *__restrict__ X,
                                                    cx<double> *__restrict__ psi, ChebyCtl *ctl,
                                                    double *__restrict__ partials, const double *init_part,
                                                    int nb_init) {
    if (ctl->done) return;
    extern __shared__ __align__(16) char smem_raw[];
    __shared__ double red[16];
    RowBlock<R, P> B(g, rb, S, reinterpret_cast<cx<R> *>(smem_raw), p0);
    const int st = B.st;
    const long base = (long)st * g.size;

    // the line of phi_{k-1} this thread transforms; the epilogue operands
    // (psi, phi_{k-2}, Y, phi_{k-1} of every state, W) are read after the
    // transforms, so that only the line stays live across them
    cx<R> v[P];
    if (B.active) {
#pragma unroll
        for (int q = 0; q < P; ++q) v[q] = phim1[base + B.gidx(q)];
    }
    // this thread's row of the operator matrix (runtime state index, read once)
    const double *tk0 = h.tk[st][0];
    const bool has_t = tk0 != nullptr;
    double tq[P];
    if (B.active) {
#pragma unroll
        for (int q = 0; q < P; ++q) tq[q] = has_t ? (0.0 + tk0[B.lt + q * B.TL]) * h.inv_nd[0] : 0.0;
    }
    B.stage(p0);
    if (PRO) {
        // sum |psi|^2 of the step start, in the fixed order every kernel uses
        if (blockIdx.x == 0) {
            const double n0 = fixed_sum(init_part, nb_init, red);
            if (threadIdx.x == 0) {
                ctl->norm0 = n0;
                ctl->last_order = 1;
            }
        }
    } else {
        // break test of the previous order (OCheby.cpp:179-182).  On grids
        // with nd >= 2 the order's first column pass has evaluated it and
        // recorded a break in ctl->done, read at the kernel's start.
        if (ONE_D) {
            double acc_prev;
            if (prev_converged(cc, red, &acc_prev)) {
                if (blockIdx.x == 0 && threadIdx.x == 0) {
                    ctl->done = 1;
                    ctl->accnorm = acc_prev;
                }
                return;
            }
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) ctl->last_order = ord.k;
    }
    const R *Wp[S];
    const R *Mp[S];
    double Ed[S];
#pragma unroll
    for (int j = 0; j < S; ++j) {
        Wp[j] = h.W[st][j];
        Mp[j] = h.mu[st][j];
        Ed[j] = h.Edip[st][j];
    }
    const double uni = h.unity[st];
    // fused epilogue: H phi = T phi + sum_j (W_ij + unity - E mu_ij) phi_j,
    // in chunks of four points.  The first chunk's operands are requested
    // before the transforms (their latency hides behind them), the second
    // (P = 8) right after.
    const int nterm = PRO ? 1 : ord.nterm;
    const bool upd = nterm > 0;
    struct Half {
        cx<double> ps[4];
        cx<R> m2[4], yv[4], pm[S][4];
        double wv[S][4];
    };
    auto load_half = [&](Half &H, int h0) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const long gi = B.gidx(h0 + u);
            if (upd) H.ps[u] = psi[base + gi];
            if (!PRO) H.m2[u] = phim2[base + gi];
            if (!ONE_D) H.yv[u] = X[base + gi];
#pragma unroll
            for (int j = 0; j < S; ++j) {
                H.pm[j][u] = phim1[(long)j * g.size + gi];
                double w = (j == st) ? uni : 0.0;
                if (Wp[j]) w += (double)Wp[j][gi];
                if (Mp[j]) w += Ed[j] * (double)Mp[j][gi];
                H.wv[j][u] = w;
            }
        }
    };
    double acc = 0;
    const cx<double> ck = mk<double>(ord.ckr, ord.cki);
    const cx<double> c0 = mk<double>(ord.c0r, ord.c0i);
    const cx<double> cm1 = mk<double>(ord.cm1r, ord.cm1i), cm2 = mk<double>(ord.cm2r, ord.cm2i);
    auto finish_half = [&](const Half &H, int h0) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = h0 + u;
            const long gi = B.gidx(q);
            cx<double> hp = mk<double>(0, 0);
            if (has_t) hp = ONE_D ? to_d(v[q]) : to_d(v[q]) + to_d(H.yv[u]);
            cx<double> p = to_d(H.pm[0][u]);
#pragma unroll
            for (int j = 0; j < S; ++j) {
                const double w = H.wv[j][u];
                const cx<double> pj = to_d(H.pm[j][u]);
                if (j == st) p = pj;
                hp.r += w * pj.r;
                hp.i += w * pj.i;
            }
            // b = (H phi - offset phi) / R
            const cx<double> b = mk<double>((hp.r - ord.offset * p.r) * ord.inv_scaling,
                                            (hp.i - ord.offset * p.i) * ord.inv_scaling);
            cx<double> pk;
            if (PRO) {
                pk = IMAG ? mk<double>(-b.r, -b.i)                        // * (-1)
                              : mk<double>(b.i, -b.r);                    // * (-i)
                psi[base + gi] = cmul(c0, H.ps[u]) + cmul(ck, pk);
            } else {
                const cx<double> mm = to_d(H.m2[u]);
                pk = IMAG ? mk<double>(-2.0 * b.r - mm.r, -2.0 * b.i - mm.i)       // * (-2) - phi_{k-2}
                              : mk<double>(2.0 * b.i + mm.r, -2.0 * b.r + mm.i);   // * (-2i) + phi_{k-2}
                if (upd) {
                    cx<double> sv = H.ps[u];
                    if (nterm >= 3) sv = sv + cmul(cm2, mm);
                    if (nterm >= 2) sv = sv + cmul(cm1, p);
                    sv = sv + cmul(ck, pk);
                    acc += sv.r * sv.r + sv.i * sv.i;
                    psi[base + gi] = sv;
                }
            }
            phim2[base + gi] = from_d<R>(pk);
        }
    };
    Half H0, H1;
    if (B.active) load_half(H0, 0);
    // T_0 part along dimension 0 (Kspace filter, 1/N_0 folded into tq)
    B.fft(v, p0, -1);
    if (B.active) {
#pragma unroll
        for (int q = 0; q < P; ++q) v[q] = mk<R>((R)((double)v[q].r * tq[q]), (R)((double)v[q].i * tq[q]));
    }
    B.fft(v, p0, +1);
    if (B.active) {
        if constexpr (P == 8) load_half(H1, 4);
        finish_half(H0, 0);
        if constexpr (P == 8) finish_half(H1, 4);
    }
    if (PRO) return;
    // partial sum |psi|^2 of this block; the next kernels test convergence
    const double part = block_sum(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = part;
}

// ---------------------------------------------------------------------------
// A whole Chebyshev step of a single-state 1D line (L <= 4096) in one block:
// the order-by-order arithmetic of k_cheby_init and k_cheby_row (same
// operands in the same order: prologue, recurrence, FP64 psi accumulation,
// the fixed-length FP32 series' update every third order, one block sum of
// |psi|^2 per order and the break test on it, OCheby.cpp:92-197), with psi
// and the recurrence terms in registers and only the FFT stages in LDS: one
// launch per step instead of one per order.  coef[2k], coef[2k+1]: coeff[k]
// exp(-i offset dt).  partials: the per-order slots of the multi-kernel path
// ([k % 2] * kMaxPartials, step start at 2 * kMaxPartials), one sum each.
template <class R, int NT>
__global__ void __launch_bounds__(NT) k_cheby_line(GridDev g, HamDev<R> h, LinePlan p0, const double *coef,
                                                   int order, double offset, double inv_scaling, double prec,
                                                   cx<double> *__restrict__ psi, ChebyCtl *ctl,
                                                   double *__restrict__ partials) {
    extern __shared__ __align__(16) char smem_raw[];
    __shared__ double red[16];
    __shared__ double bc;
    RowBlock<R> B(g, 1, 1, reinterpret_cast<cx<R> *>(smem_raw), p0);
    const double *tk0 = h.tk[0][0];
    const bool has_t = tk0 != nullptr;
    const R *Wp = h.W[0][0];
    const R *Mp = h.mu[0][0];
    const double Ed = h.Edip[0][0];
    const double uni = h.unity[0];
    cx<double> ps[8];
    cx<R> pm1[8], pm2[8], v[8];
    double wv[8], tq[8];
    double acc = 0;
    if (B.active) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const long gi = B.gidx(q);
            ps[q] = psi[gi];
            acc += ps[q].r * ps[q].r + ps[q].i * ps[q].i;
            pm1[q] = from_d<R>(ps[q]);
            double w = uni;
            if (Wp) w += (double)Wp[gi];
            if (Mp) w += Ed * (double)Mp[gi];
            wv[q] = w;
            tq[q] = has_t ? (0.0 + tk0[B.lt + q * B.TL]) * h.inv_nd[0] : 0.0;
        }
    }
    // sum |psi|^2 of the step start (the init kernel's single partial)
    double norm0 = block_sum(acc, red);
    if (threadIdx.x == 0) {
        partials[2 * kMaxPartials] = norm0;
        bc = norm0;
        ctl->norm0 = norm0;
        ctl->done = 0;
        ctl->last_order = 1;
    }
    __syncthreads();
    norm0 = bc;
    B.stage(p0);
    const cx<double> c0 = mk<double>(coef[0], coef[1]);
    for (int k = 1; k < order; ++k) {
        if (k >= 3 && prec > 0) {
            // break test of order k - 1 on its sum (stored in bc below)
            const double a = bc;
            if (fabs(norm0 - a) / norm0 < prec) {   // uniform over the block
                if (threadIdx.x == 0) {
                    ctl->done = 1;
                    ctl->accnorm = a;
                }
                break;   // psi holds the series through order k - 1
            }
        }
        if (threadIdx.x == 0) ctl->last_order = k;
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = pm1[q];
        B.fft(v, p0, -1);
        if (B.active) {
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] = mk<R>((R)((double)v[q].r * tq[q]), (R)((double)v[q].i * tq[q]));
        }
        B.fft(v, p0, +1);
        // psi update at this order: every order with the break test, the
        // last three terms every third order of the fixed-length series
        int nterm = 1;
        if (k >= 2 && !(prec > 0)) {
            const int m = (k - 1) % 3;
            nterm = m == 0 ? 3 : (k == order - 1 ? m : 0);
        }
        const cx<double> ck = mk<double>(coef[2 * k], coef[2 * k + 1]);
        const cx<double> cm1 = k >= 2 ? mk<double>(coef[2 * (k - 1)], coef[2 * (k - 1) + 1]) : mk<double>(0, 0);
        const cx<double> cm2 = k >= 3 ? mk<double>(coef[2 * (k - 2)], coef[2 * (k - 2) + 1]) : mk<double>(0, 0);
        acc = 0;
        if (B.active) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                cx<double> hp = mk<double>(0, 0);
                if (has_t) hp = to_d(v[q]);
                const cx<double> p = to_d(pm1[q]);
                hp.r += wv[q] * p.r;
                hp.i += wv[q] * p.i;
                const cx<double> b = mk<double>((hp.r - offset * p.r) * inv_scaling, (hp.i - offset * p.i) * inv_scaling);
                cx<double> pk;
                if (k == 1) {
                    pk = mk<double>(b.i, -b.r);
                    ps[q] = cmul(c0, ps[q]) + cmul(ck, pk);
                } else {
                    const cx<double> mm = to_d(pm2[q]);
                    pk = mk<double>(2.0 * b.i + mm.r, -2.0 * b.r + mm.i);
                    if (nterm > 0) {
                        cx<double> sv = ps[q];
                        if (nterm >= 3) sv = sv + cmul(cm2, mm);
                        if (nterm >= 2) sv = sv + cmul(cm1, p);
                        sv = sv + cmul(ck, pk);
                        acc += sv.r * sv.r + sv.i * sv.i;
                        ps[q] = sv;
                    }
                }
                pm2[q] = pm1[q];
                pm1[q] = from_d<R>(pk);
            }
        }
        if (k >= 2) {
            const double a = block_sum(acc, red);
            if (threadIdx.x == 0) {
                partials[(k % 2) * kMaxPartials] = a;
                bc = a;
            }
            __syncthreads();
        }
    }
    if (B.active) {
#pragma unroll
        for (int q = 0; q < 8; ++q) psi[B.gidx(q)] = ps[q];
    }
}

// ---------------------------------------------------------------------------
// Strided line pass along dimension `dim` (1 or 2) on X, in place.  A block
// owns a tile of W adjacent dimension-0 positions for one state and one index
// of the remaining dimension; the W lanes of a butterfly touch W adjacent
// words in LDS and W adjacent complex values in memory.
// MODE 0: forward, 1: inverse, 2: forward * T/N inverse,
//      3: forward * exp(-i tau T)/N inverse,
//      4: X = IFFT(T_dim/N_dim FFT(src)), the dimension's part of the
//         separable kinetic operator, 5: X += the same.
template <class R, int MODE, int P = 8, bool IMAG = false>
__global__ void __launch_bounds__(512) k_col(GridDev g, int S, int dim, int W, LinePlan p, HamDev<R> h, double tau,
                                              cx<R> *__restrict__ X, const cx<R> *__restrict__ src,
                                              ChebyCtl *ctl, ConvCheck cc) {
    if (ctl && ctl->done) return;
    extern __shared__ __align__(16) char smem_raw[];
    cx<R> *smem = reinterpret_cast<cx<R> *>(smem_raw);
    __shared__ double bc;
    const int N0 = g.N[0];
    const int L = g.N[dim];
    const int TL = L / P;
    const long estride = dim == 1 ? N0 : (long)N0 * g.N[1];
    const int nother = g.nd == 3 ? (dim == 1 ? g.N[2] : g.N[1]) : 1;
    const long ostride = dim == 1 ? (long)N0 * g.N[1] : N0;
    const int tiles0 = N0 / W;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int c0 = (b % tiles0) * W;
    const int rest = b / tiles0;
    const int oth = rest % nother;
    const int st = rest / nother;
    const int c = threadIdx.x % W;
    const int lt = threadIdx.x / W;
    const bool active = lt < TL;
    cx<R> *base = X + (long)st * g.size + (long)oth * ostride + c0 + c;
    LdsLine<R> ln;
    ln.a = smem + c;
    ln.b = smem + (long)W * tw_lds_len(L) + c;
    ln.es = W;
    ln.pad = true;
    cx<R> *twl = smem + (long)2 * W * tw_lds_len(L);
    cx<R> v[P], yv[P];
    const bool tw_in_lds = p.tw_lds;
    auto FFT = [&](int sign) {
        if (tw_in_lds) fft_line<R, P>(v, ln, p, twl, sign, lt, active, kTwLds);
        else fft_line<R, P>(v, ln, p, reinterpret_cast<const cx<R> *>(p.tw), sign, lt, active, kTwMem);
    };

    if (active) {
        const cx<R> *in = MODE >= 4 ? src + (base - X) : base;
#pragma unroll
        for (int q = 0; q < P; ++q) v[q] = in[(long)(lt + q * TL) * estride];
        if (MODE == 5) {
#pragma unroll
            for (int q = 0; q < P; ++q) yv[q] = base[(long)(lt + q * TL) * estride];
        }
    }
    // kinetic table values of this thread's elements, requested early
    double tq[P];
    if (MODE >= 2 && active) {
        const bool has_t = h.tk[st][0] != nullptr;
        const int k0 = c0 + c;
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int kd = lt + q * TL;
            if (MODE >= 4) tq[q] = has_t ? (0.0 + h.tk[st][dim][kd]) * h.inv_nd[dim] : 0.0;
            else tq[q] = has_t ? kin_sum<R>(h, st, g, k0, dim == 1 ? kd : oth, dim == 2 ? kd : oth) : 0.0;
        }
    }
    if (tw_in_lds) stage_twiddles(twl, p);
    {
        double a;
        if (prev_converged(cc, &bc, &a)) {
            // the previous order ended the series (OCheby.cpp:179-182); the
            // row kernel of this order reads the outcome from ctl.  Every
            // block reaches the same decision, block 0 records it.
            if (MODE == 4 && ctl && b == 0 && threadIdx.x == 0) {
                ctl->done = 1;
                ctl->accnorm = a;
            }
            return;
        }
    }
    FFT(MODE == 1 ? +1 : -1);
    if (MODE >= 2) {
        if (active) {
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const double tt = tq[q];
                if (MODE >= 4) {
                    v[q] = mk<R>((R)((double)v[q].r * tt), (R)((double)v[q].i * tt));
                } else if (MODE == 2) {
                    const double t = tt * h.inv_n;
                    v[q] = mk<R>((R)((double)v[q].r * t), (R)((double)v[q].i * t));
                } else if (IMAG) {
                    // exp(-tau T) / N, imaginary time
                    const double f = exp(-tau * tt) * h.inv_n;
                    v[q] = mk<R>((R)((double)v[q].r * f), (R)((double)v[q].i * f));
                } else {
                    // exp(-i tau T) / N (OGSPO.cpp:37-41 folds 1/N into the exponential)
                    double sn, cs;
                    sincos(-tau * tt, &sn, &cs);
                    v[q] = from_d<R>(cmul(to_d(v[q]), mk<double>(cs * h.inv_n, sn * h.inv_n)));
                }
            }
        }
        FFT(+1);
    }
    if (active) {
        if (MODE == 5) {
#pragma unroll
            for (int q = 0; q < P; ++q) v[q] = v[q] + yv[q];
        }
#pragma unroll
        for (int q = 0; q < P; ++q) base[(long)(lt + q * TL) * estride] = v[q];
    }
}

// ---------------------------------------------------------------------------
// H applied to a complex128 vector (Operator::Apply): forward row pass,
// strided passes, row epilogue.
template <class R, int S>
__global__ void __launch_bounds__(512) k_row_fwd(GridDev g, LinePlan p0, int rb, const cx<double> *__restrict__ src,
                                                  cx<R> *__restrict__ X) {
    extern __shared__ __align__(16) char smem_raw[];
    RowBlock<R> B(g, rb, S, reinterpret_cast<cx<R> *>(smem_raw), p0);
    const long base = (long)B.st * g.size;
    cx<R> v[8];
    if (B.active) {
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = from_d<R>(src[base + B.gidx(q)]);
    }
    B.stage(p0);
    B.fft(v, p0, -1);
    if (B.active) {
#pragma unroll
        for (int q = 0; q < 8; ++q) X[base + B.gidx(q)] = v[q];
    }
}

template <class R, int S, bool ONE_D>
__global__ void __launch_bounds__(512) k_apply_row(GridDev g, HamDev<R> h, LinePlan p0, int rb,
                                                    const cx<double> *__restrict__ src, const cx<R> *__restrict__ X,
                                                    cx<double> *__restrict__ dst) {
    extern __shared__ __align__(16) char smem_raw[];
    RowBlock<R> B(g, rb, S, reinterpret_cast<cx<R> *>(smem_raw), p0);
    const int st = B.st;
    const long base = (long)st * g.size;
    cx<R> v[8];
    if (B.active) {
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = ONE_D ? from_d<R>(src[base + B.gidx(q)]) : X[base + B.gidx(q)];
    }
    B.stage(p0);
    const bool has_t = h.tk[st][0] != nullptr;
    if (ONE_D) {
        B.fft(v, p0, -1);
        if (B.active) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const double t = has_t ? (0.0 + h.tk[st][0][B.lt + q * B.TL]) * h.inv_n : 0.0;
                v[q] = mk<R>((R)((double)v[q].r * t), (R)((double)v[q].i * t));
            }
        }
    }
    B.fft(v, p0, +1);
    if (!B.active) return;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const long gi = B.gidx(q);
        cx<double> hp = has_t ? to_d(v[q]) : mk<double>(0, 0);
#pragma unroll
        for (int j = 0; j < S; ++j) {
            double w = (j == st) ? h.unity[st] : 0.0;
            if (h.W[st][j]) w += (double)h.W[st][j][gi];
            if (h.mu[st][j]) w += h.Edip[st][j] * (double)h.mu[st][j][gi];
            const cx<double> pj = src[(long)j * g.size + gi];
            hp.r += w * pj.r;
            hp.i += w * pj.i;
        }
        dst[base + gi] = hp;
    }
}

// ---------------------------------------------------------------------------
// GSPO x-space split operators on one grid point (all states):
//   psi <- exp(-i tau W(x)) psi, W real symmetric (ODSpace.h:65-69,
//   OGridDipole.cpp:83-87, ODMultistate.cpp:109-159 / TransformMSD.cpp:31-101)
template <class R, int S, bool IMAG = false>
__device__ __forceinline__ void xsplit_point(const XSplitDev<R> &op, long gi, cx<double> (&v)[S]) {
    double Wm[S][S];
#pragma unroll
    for (int i = 0; i < S; ++i)
#pragma unroll
        for (int j = 0; j < S; ++j) {
            double w = 0;
            if (op.V[i][j]) w += (double)op.V[i][j][gi];
            if (op.mu[i][j]) w += op.Edip[i][j] * (double)op.mu[i][j][gi];
            Wm[i][j] = w;
        }
    if (!op.coupled || S == 1) {
#pragma unroll
        for (int i = 0; i < S; ++i) {
            if (IMAG) {   // exp(-tau w), imaginary time
                const double f = exp(-op.tau * Wm[i][i]);
                v[i] = mk<double>(v[i].r * f, v[i].i * f);
                continue;
            }
            double sn, cs;
            sincos(-op.tau * Wm[i][i], &sn, &cs);
            v[i] = cmul(v[i], mk<double>(cs, sn));
        }
        return;
    }
    if constexpr (S == 2) {
        if (IMAG) {
            // exp(-tau W) = e^{-tau m} [cosh(tau O) - sinh(tau O)/O (W - m)]
            const double m = 0.5 * (Wm[0][0] + Wm[1][1]);
            const double d = 0.5 * (Wm[0][0] - Wm[1][1]);
            const double b = Wm[0][1];
            const double O = sqrt(d * d + b * b);
            const double ch = cosh(op.tau * O);
            const double sh = O > 0 ? sinh(op.tau * O) / O : op.tau;
            const double e = exp(-op.tau * m);
            const double u00 = e * (ch - sh * d), u11 = e * (ch + sh * d), u01 = -e * sh * b;
            const cx<double> a0 = mk<double>(u00 * v[0].r + u01 * v[1].r, u00 * v[0].i + u01 * v[1].i);
            const cx<double> a1 = mk<double>(u01 * v[0].r + u11 * v[1].r, u01 * v[0].i + u11 * v[1].i);
            v[0] = a0;
            v[1] = a1;
            return;
        }
        // exp(-i tau W) = e^{-i tau m} [cos(tau O) - i sin(tau O)/O (W - m)]
        const double m = 0.5 * (Wm[0][0] + Wm[1][1]);
        const double d = 0.5 * (Wm[0][0] - Wm[1][1]);
        const double b = Wm[0][1];
        const double O = sqrt(d * d + b * b);
        const double c = cos(op.tau * O);
        const double sc = O > 0 ? sin(op.tau * O) / O : op.tau;
        double sm, cm;
        sincos(-op.tau * m, &sm, &cm);
        const cx<double> ph = mk<double>(cm, sm);
        const cx<double> u00 = mk<double>(c, -sc * d), u11 = mk<double>(c, sc * d);
        const cx<double> u01 = mk<double>(0, -sc * b);
        const cx<double> a0 = cmul(u00, v[0]) + cmul(u01, v[1]);
        const cx<double> a1 = cmul(u01, v[0]) + cmul(u11, v[1]);
        v[0] = cmul(ph, a0);
        v[1] = cmul(ph, a1);
    }
}

template <class R, int S, bool IMAG = false>
__device__ __forceinline__ void xgroup_point(const XGroupDev<R> &grp, long gi, cx<double> (&v)[S]) {
    for (int o = 0; o < grp.n; ++o) xsplit_point<R, S, IMAG>(grp.op[o], gi, v);
}

// GSPO kinetic split step over rows (P points per thread).  ONE_D: psi ->
// pre -> FFT -> exp(-i tau
// T)/N -> IFFT -> post -> psi in one pass.  Otherwise PHASE 0: psi -> pre ->
// FFT_0 -> X; PHASE 1: X -> IFFT_0 -> post -> psi.  The x-space operators act
// on all states of a point at once, so the S lines of a row are exchanged
// through LDS around them.
template <class R, int S, bool ONE_D, int PHASE, bool IMAG = false, int P = 8>
__global__ void __launch_bounds__(512) k_gspo_row(GridDev g, HamDev<R> kin, double tau, LinePlan p0, int rb,
                                                   XGroupDev<R> pre, XGroupDev<R> post, cx<R> *__restrict__ psi,
                                                   cx<R> *__restrict__ X) {
    extern __shared__ __align__(16) char smem_raw[];
    cx<R> *smem = reinterpret_cast<cx<R> *>(smem_raw);
    RowBlock<R, P> B(g, rb, S, smem, p0);
    const int st = B.st;
    const long base = (long)st * g.size;
    cx<R> v[P];
    B.stage(p0);
    // exchange area after the transform buffers and the twiddle table: [S][rb][L]
    cx<R> *xa = B.after;
    auto xspace = [&](const XGroupDev<R> &grp) {
        if (grp.n == 0) return;
        if (S == 1) {
            if (B.active) {
#pragma unroll
                for (int q = 0; q < P; ++q) {
                    cx<double> w[S];
                    w[0] = to_d(v[q]);
                    xgroup_point<R, S, IMAG>(grp, B.gidx(q), w);
                    v[q] = from_d<R>(w[0]);
                }
            }
            return;
        }
        __syncthreads();
        if (B.active) {
#pragma unroll
            for (int q = 0; q < P; ++q) xa[((long)st * B.rb + B.rr) * B.L + B.lt + q * B.TL] = v[q];
        }
        __syncthreads();
        // the points of row rr are shared out among the S lines' threads
        if (B.active) {
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const int x = B.lt + q * B.TL;
                if (x % S != st) continue;
                cx<double> w[S];
#pragma unroll
                for (int s2 = 0; s2 < S; ++s2) w[s2] = to_d(xa[((long)s2 * B.rb + B.rr) * B.L + x]);
                xgroup_point<R, S, IMAG>(grp, B.row * B.L + x, w);
#pragma unroll
                for (int s2 = 0; s2 < S; ++s2) xa[((long)s2 * B.rb + B.rr) * B.L + x] = from_d<R>(w[s2]);
            }
        }
        __syncthreads();
        if (B.active) {
#pragma unroll
            for (int q = 0; q < P; ++q) v[q] = xa[((long)st * B.rb + B.rr) * B.L + B.lt + q * B.TL];
        }
    };
    if (PHASE == 0 || ONE_D) {
        if (B.active) {
#pragma unroll
            for (int q = 0; q < P; ++q) v[q] = psi[base + B.gidx(q)];
        }
        xspace(pre);
        B.fft(v, p0, -1);
        if (!ONE_D) {
            if (B.active) {
#pragma unroll
                for (int q = 0; q < P; ++q) X[base + B.gidx(q)] = v[q];
            }
            return;
        }
        if (B.active) {
            const bool has_t = kin.tk[st][0] != nullptr;
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const double tt = has_t ? 0.0 + kin.tk[st][0][B.lt + q * B.TL] : 0.0;
                if (IMAG) {   // exp(-tau T)/N, imaginary time
                    const double f = exp(-tau * tt) * kin.inv_n;
                    v[q] = mk<R>((R)((double)v[q].r * f), (R)((double)v[q].i * f));
                    continue;
                }
                double sn, cs;
                sincos(-tau * tt, &sn, &cs);
                v[q] = from_d<R>(cmul(to_d(v[q]), mk<double>(cs * kin.inv_n, sn * kin.inv_n)));
            }
        }
    } else {
        if (B.active) {
#pragma unroll
            for (int q = 0; q < P; ++q) v[q] = X[base + B.gidx(q)];
        }
    }
    B.fft(v, p0, +1);
    xspace(post);
    if (B.active) {
#pragma unroll
        for (int q = 0; q < P; ++q) psi[base + B.gidx(q)] = v[q];
    }
}

template <class R, int S, bool IMAG = false>
__global__ void k_gspo_xonly(GridDev g, XGroupDev<R> grp, cx<R> *__restrict__ psi) {
    const long gi = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= g.size) return;
    cx<double> v[S];
#pragma unroll
    for (int st = 0; st < S; ++st) v[st] = to_d(psi[(long)st * g.size + gi]);
    xgroup_point<R, S, IMAG>(grp, gi, v);
#pragma unroll
    for (int st = 0; st < S; ++st) psi[(long)st * g.size + gi] = from_d<R>(v[st]);
}

// ---------------------------------------------------------------------------
// deterministic per-state complex overlap: stage 1 writes partials, stage 2
// (one block) combines them in index order
template <class R>
__global__ void k_overlap_stage1(const cx<R> *__restrict__ a, const cx<R> *__restrict__ b, long n,
                                 double *partials) {
    __shared__ double red[16];
    const int st = blockIdx.y;
    double sr = 0, si = 0;
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        cx<double> x = to_d(a[(long)st * n + j]), y = to_d(b[(long)st * n + j]);
        sr += x.r * y.r + x.i * y.i;
        si += x.r * y.i - x.i * y.r;
    }
    double pr = block_sum(sr, red);
    double pi = block_sum(si, red);
    if (threadIdx.x == 0) {
        partials[2 * (st * gridDim.x + blockIdx.x)] = pr;
        partials[2 * (st * gridDim.x + blockIdx.x) + 1] = pi;
    }
}

// sum_i w(i) |a_s(i)|^2 per state s (blockIdx.y), w the coordinate of i
// along `dim`: ktab[i_dim] when ktab is given (k-space), else base + step
// i_dim (x-space).  Partials as (sum, 0) pairs for k_overlap_stage2.
__global__ void k_wnorm_stage1(const cx<double> *__restrict__ a, long n, int stride, int Nd, double base,
                               double step, const double *__restrict__ ktab, double *partials) {
    __shared__ double red[16];
    const int st = blockIdx.y;
    double sum = 0;
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        const int id = (int)((j / stride) % Nd);
        const double w = ktab ? ktab[id] : base + step * (double)id;
        const cx<double> x = a[(long)st * n + j];
        sum += w * (x.r * x.r + x.i * x.i);
    }
    const double ps = block_sum(sum, red);
    if (threadIdx.x == 0) {
        partials[2 * (st * gridDim.x + blockIdx.x)] = ps;
        partials[2 * (st * gridDim.x + blockIdx.x) + 1] = 0.0;
    }
}

__global__ void k_overlap_stage2(const double *partials, int nbx, int S, double *out) {
    __shared__ double red[16];
    for (int st = 0; st < S; ++st) {
        double sr = 0, si = 0;
        for (int i = threadIdx.x; i < nbx; i += blockDim.x) {
            sr += partials[2 * (st * nbx + i)];
            si += partials[2 * (st * nbx + i) + 1];
        }
        double tr = block_sum(sr, red);
        double ti = block_sum(si, red);
        if (threadIdx.x == 0) {
            out[2 * st] = tr;
            out[2 * st + 1] = ti;
        }
        __syncthreads();
    }
}

__global__ void k_c128_to_c64(const double *__restrict__ s, float *__restrict__ d, long n2) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (long)gridDim.x * blockDim.x)
        d[i] = (float)s[i];
}
// y = sa * y + b * x over n complex128 values (x may be null): scaling and
// linear combinations of wave functions (cVec MultElements / MultElementsAdd)
__global__ void k_scale_axpy(double *__restrict__ y, const double *__restrict__ x, long n, double sa, double br,
                             double bi) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double yr = y[2 * i], yi = y[2 * i + 1];
        if (sa != 1.0) {
            yr *= sa;
            yi *= sa;
        }
        if (x) {
            const double xr = x[2 * i], xi = x[2 * i + 1];
            yr += br * xr - bi * xi;
            yi += br * xi + bi * xr;
        }
        y[2 * i] = yr;
        y[2 * i + 1] = yi;
    }
}

__global__ void k_c64_to_c128(const float *__restrict__ s, double *__restrict__ d, long n2) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (long)gridDim.x * blockDim.x)
        d[i] = (double)s[i];
}

// ---------------------------------------------------------------------------
// launch helpers

static int round64(int n) {
    const int r = ((n + 63) / 64) * 64;
    if (r > 512) throw std::runtime_error("line block needs more than 512 threads");
    return r;
}

// rows per block: grow the block towards 256 threads while the grid keeps at
// least 512 blocks (P points per thread)
static int rows_per_block(const GridDev &g, int S, int L, int P = 8) {
    const long nrows = g.size / L;
    const int TL = L / P;
    int rb = 1;
    while (rb * 2 <= nrows && S * rb * 2 * TL <= 256 && nrows / (rb * 2) >= 512) rb *= 2;
    return rb;
}

constexpr size_t kLdsBytes = 160 * 1024;

template <class K>
static void set_lds(K kernel, size_t bytes) {
    if (bytes > kLdsBytes) throw std::runtime_error("line block exceeds the 160 KiB LDS of a gfx950 CU");
    if (bytes > 65536)
        QDX_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

// LDS of a strided pass: two padded buffers of W lines and the twiddle copy
static size_t col_lds_bytes(int W, int L, size_t elem) {
    return ((size_t)2 * W * tw_lds_len(L) + tw_lds_len(L)) * elem;
}

// tile width of the strided passes: adjacent dimension-0 positions per block.
// Prefer >= 64-byte segments, keep >= 256 blocks when possible.
static int col_tile(const GridDev &g, int S, int dim, int L, size_t elem, int P = 8) {
    const int nother = g.nd == 3 ? (dim == 1 ? g.N[2] : g.N[1]) : 1;
    int W = 16;
    auto fits = [&](int w) {
        return g.N[0] % w == 0 && col_lds_bytes(w, L, elem) <= 96 * 1024 && w * (L / P) <= 512;
    };
    while (W > 1 && !fits(W)) W /= 2;
    const int wmin = std::max(1, (int)(64 / elem));
    while (W > wmin && (long)(g.N[0] / W) * nother * S < 256 && fits(W / 2)) W /= 2;
    return W;
}

struct RowLaunch {
    int rb, nb, threads;
    size_t lds;
};

template <class R>
static RowLaunch row_launch(const GridDev &g, int S, LinePlan &p0, bool extra_exchange = false) {
    const int L = g.N[0];
    RowLaunch r;
    r.rb = rows_per_block(g, S, L, p0.ppt);
    const long nrows = g.size / L;
    r.nb = (int)((nrows + r.rb - 1) / r.rb);
    r.threads = std::max(64, round64(S * r.rb * (L / p0.ppt)));
    const size_t extra = extra_exchange ? (size_t)S * r.rb * L * sizeof(cx<R>) : 0;
    // keep the twiddle table in LDS when it fits next to the line buffers
    p0.tw_lds = row_lds_bytes<R>(L, S * r.rb, true) + extra <= kLdsBytes ? 1 : 0;
    r.lds = row_lds_bytes<R>(L, S * r.rb, p0.tw_lds != 0) + extra;
    return r;
}

template <class R, int S>
static int cheby_init_S(hipStream_t st, const GridDev &g, const LinePlan &plan, const double *psi, R *phi0,
                        R *X, ChebyCtl *ctl, double *partials) {
    LinePlan p0 = plan;
    p0.ppt = 8;   // the init kernel holds 8 points per thread
    RowLaunch r = row_launch<R>(g, S, p0);
    if (r.nb > kMaxPartials) throw std::runtime_error("grid too large for the partial-sum buffer");
    auto cpsi = reinterpret_cast<const cx<double> *>(psi);
    auto cphi = reinterpret_cast<cx<R> *>(phi0);
    auto cx_ = reinterpret_cast<cx<R> *>(X);
    if (X) {
        set_lds(k_cheby_init<R, S, true>, r.lds);
        QDX_LAUNCH((k_cheby_init<R, S, true>), dim3(r.nb), dim3(r.threads), r.lds, st, g, p0, r.rb, cpsi,
                           cphi, cx_, ctl, partials);
    } else {
        QDX_LAUNCH((k_cheby_init<R, S, false>), dim3(r.nb), dim3(r.threads), 0, st, g, p0, r.rb, cpsi, cphi,
                           cx_, ctl, partials);
    }
    QDX_HIP(hipGetLastError());
    return r.nb;
}

#define QDX_DISPATCH_S(S_, CALL)                                              \
    switch (S_) {                                                             \
    case 1: { constexpr int SS = 1; CALL; } break;                            \
    case 2: { constexpr int SS = 2; CALL; } break;                            \
    case 3: { constexpr int SS = 3; CALL; } break;                            \
    case 4: { constexpr int SS = 4; CALL; } break;                            \
    default: throw std::runtime_error("unsupported number of states");        \
    }

template <class R>
int Launch<R>::cheby_init(hipStream_t s, const GridDev &g, int S, const LinePlan &p0, const double *psi, R *phi0,
                          R *X, ChebyCtl *ctl, double *partials) {
    int nb = 0;
    QDX_DISPATCH_S(S, (nb = cheby_init_S<R, SS>(s, g, p0, psi, phi0, X, ctl, partials)));
    return nb;
}

template <class R, int S, bool ONE_D, bool PRO, int P>
static int cheby_row_SKP(hipStream_t st, const GridDev &g, const HamDev<R> &h, const LinePlan &plan,
                         const ChebyOrder &ord, const ConvCheck &cc, bool imag, const R *phim1, R *phim2, R *X, double *psi,
                         ChebyCtl *ctl, double *partials, const double *init_part, int nb_init) {
    LinePlan p0 = plan;
    RowLaunch r = row_launch<R>(g, S, p0);
    if (r.nb > kMaxPartials) throw std::runtime_error("grid too large for the partial-sum buffer");
    auto cphim1 = reinterpret_cast<const cx<R> *>(phim1);
    auto cphim2 = reinterpret_cast<cx<R> *>(phim2);
    auto cX = reinterpret_cast<cx<R> *>(X);
    auto cpsi = reinterpret_cast<cx<double> *>(psi);
    if (imag) {   // imaginary time: FP64 only, the 512-thread instance
        if constexpr (std::is_same<R, double>::value) {
            set_lds(k_cheby_row<R, S, ONE_D, PRO, 512, P, true>, r.lds);
            QDX_LAUNCH((k_cheby_row<R, S, ONE_D, PRO, 512, P, true>), dim3(r.nb), dim3(r.threads), r.lds, st, g, h,
                       p0, ord, cc, r.rb, cphim1, cphim2, cX, cpsi, ctl, partials, init_part, nb_init);
        } else {
            throw std::runtime_error("imaginary time: FP64 only");
        }
    } else if (r.threads <= 256) {
        set_lds(k_cheby_row<R, S, ONE_D, PRO, 256, P>, r.lds);
        QDX_LAUNCH((k_cheby_row<R, S, ONE_D, PRO, 256, P>), dim3(r.nb), dim3(r.threads), r.lds, st, g, h, p0, ord,
                   cc, r.rb, cphim1, cphim2, cX, cpsi, ctl, partials, init_part, nb_init);
    } else {
        set_lds(k_cheby_row<R, S, ONE_D, PRO, 512, P>, r.lds);
        QDX_LAUNCH((k_cheby_row<R, S, ONE_D, PRO, 512, P>), dim3(r.nb), dim3(r.threads), r.lds, st, g, h, p0, ord,
                   cc, r.rb, cphim1, cphim2, cX, cpsi, ctl, partials, init_part, nb_init);
    }
    QDX_HIP(hipGetLastError());
    return r.nb;
}

// 4 points per thread only on grids with nd >= 2 (single-state 1D lines run
// the whole-step kernel)
template <class R, int S, bool ONE_D, bool PRO>
static int cheby_row_SK(hipStream_t st, const GridDev &g, const HamDev<R> &h, const LinePlan &plan,
                        const ChebyOrder &ord, const ConvCheck &cc, bool imag, const R *phim1, R *phim2, R *X, double *psi,
                        ChebyCtl *ctl, double *partials, const double *init_part, int nb_init) {
    if (plan.ppt == 4) {
        if constexpr (!ONE_D)
            return cheby_row_SKP<R, S, ONE_D, PRO, 4>(st, g, h, plan, ord, cc, imag, phim1, phim2, X, psi, ctl, partials,
                                                      init_part, nb_init);
        throw std::runtime_error("4 points per thread: grids with nd >= 2 only");
    }
    return cheby_row_SKP<R, S, ONE_D, PRO, 8>(st, g, h, plan, ord, cc, imag, phim1, phim2, X, psi, ctl, partials, init_part,
                                              nb_init);
}

template <class R, int S>
static int cheby_row_S(hipStream_t st, const GridDev &g, const HamDev<R> &h, const LinePlan &p0,
                       const ChebyOrder &ord, const ConvCheck &cc, bool imag, const R *phim1, R *phim2, R *X, double *psi,
                       ChebyCtl *ctl, double *partials, const double *ip, int ni) {
    const bool one = g.nd == 1;
    const bool pro = ord.k == 1;
    if (one && pro)
        return cheby_row_SK<R, S, true, true>(st, g, h, p0, ord, cc, imag, phim1, phim2, X, psi, ctl, partials, ip, ni);
    if (one) return cheby_row_SK<R, S, true, false>(st, g, h, p0, ord, cc, imag, phim1, phim2, X, psi, ctl, partials, ip, ni);
    if (pro) return cheby_row_SK<R, S, false, true>(st, g, h, p0, ord, cc, imag, phim1, phim2, X, psi, ctl, partials, ip, ni);
    return cheby_row_SK<R, S, false, false>(st, g, h, p0, ord, cc, imag, phim1, phim2, X, psi, ctl, partials, ip, ni);
}

template <class R>
int Launch<R>::cheby_row(hipStream_t s, const GridDev &g, const HamDev<R> &h, const LinePlan &p0,
                         const ChebyOrder &ord, const ConvCheck &cc, bool imag, const R *phim1, R *phim2, R *X, double *psi,
                         ChebyCtl *ctl, double *partials, const double *init_part, int nb_init) {
    int nb = 0;
    QDX_DISPATCH_S(h.S, (nb = cheby_row_S<R, SS>(s, g, h, p0, ord, cc, imag, phim1, phim2, X, psi, ctl, partials,
                                                 init_part, nb_init)));
    return nb;
}

template <class R>
void Launch<R>::cheby_line(hipStream_t s, const GridDev &g, const HamDev<R> &h, const LinePlan &p0,
                           const double *coef, int order, double offset, double inv_scaling, double prec,
                           double *psi, ChebyCtl *ctl, double *partials) {
    const int L = g.N[0], TL = L / 8;
    if (g.nd != 1 || h.S != 1 || L > kMaxLine) throw std::runtime_error("whole-step line kernel: one 1D line");
    LinePlan p = p0;
    size_t lds = row_lds_bytes<R>(L, 1, true);
    p.tw_lds = lds <= kLdsBytes ? 1 : 0;
    if (!p.tw_lds) lds = row_lds_bytes<R>(L, 1, false);
    const int threads = std::max(64, round64(TL));
    auto cpsi = reinterpret_cast<cx<double> *>(psi);
    if (threads <= 256) {
        set_lds(k_cheby_line<R, 256>, lds);
        QDX_LAUNCH((k_cheby_line<R, 256>), dim3(1), dim3(threads), lds, s, g, h, p, coef, order, offset, inv_scaling,
                   prec, cpsi, ctl, partials);
    } else {
        set_lds(k_cheby_line<R, 512>, lds);
        QDX_LAUNCH((k_cheby_line<R, 512>), dim3(1), dim3(threads), lds, s, g, h, p, coef, order, offset, inv_scaling,
                   prec, cpsi, ctl, partials);
    }
    QDX_HIP(hipGetLastError());
}

template <class R, int P>
static void col_pass_P4(hipStream_t s, const GridDev &g, int S, int dim, int mode, const LinePlan &plan,
                        const HamDev<R> &h, double tau, R *X, ChebyCtl *ctl, const ConvCheck &cc, int imag) {
    const int L = g.N[dim];
    const int W = col_tile(g, S, dim, L, sizeof(cx<R>), P);
    const int nother = g.nd == 3 ? (dim == 1 ? g.N[2] : g.N[1]) : 1;
    const int nb = (g.N[0] / W) * nother * S;
    const int threads = std::max(64, round64(W * (L / P)));
    const size_t lds = col_lds_bytes(W, L, sizeof(cx<R>));
    cx<R> *x = reinterpret_cast<cx<R> *>(X);
    LinePlan pl = plan;
    pl.tw_lds = 1;
    const cx<R> *none = nullptr;
    switch (mode) {
    case 0:
        set_lds(k_col<R, 0, P>, lds);
        QDX_LAUNCH((k_col<R, 0, P>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, tau, x, none, ctl, cc);
        break;
    case 1:
        set_lds(k_col<R, 1, P>, lds);
        QDX_LAUNCH((k_col<R, 1, P>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, tau, x, none, ctl, cc);
        break;
    case 3:
        if (imag) {
            if constexpr (std::is_same<R, double>::value) {
                set_lds(k_col<R, 3, P, true>, lds);
                QDX_LAUNCH((k_col<R, 3, P, true>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, tau, x,
                           none, ctl, cc);
            } else {
                throw std::runtime_error("imaginary time: FP64 only");
            }
            break;
        }
        set_lds(k_col<R, 3, P>, lds);
        QDX_LAUNCH((k_col<R, 3, P>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, tau, x, none, ctl, cc);
        break;
    default:
        throw std::runtime_error("line pass: mode 0, 1 or 3 at 4 points per thread");
    }
    QDX_HIP(hipGetLastError());
}

template <class R>
void Launch<R>::col_pass(hipStream_t s, const GridDev &g, int S, int dim, int mode, const LinePlan &plan,
                         const HamDev<R> &h, double tau, R *X, ChebyCtl *ctl, const ConvCheck &cc, int imag) {
    if (plan.ppt == 4) return col_pass_P4<R, 4>(s, g, S, dim, mode, plan, h, tau, X, ctl, cc, imag);
    const int L = g.N[dim];
    const int W = col_tile(g, S, dim, L, sizeof(cx<R>));
    const int nother = g.nd == 3 ? (dim == 1 ? g.N[2] : g.N[1]) : 1;
    const int nb = (g.N[0] / W) * nother * S;
    const int threads = std::max(64, round64(W * (L / 8)));
    const size_t lds = col_lds_bytes(W, L, sizeof(cx<R>));
    cx<R> *x = reinterpret_cast<cx<R> *>(X);
    LinePlan pl = plan;
    pl.tw_lds = 1;
    const cx<R> *none = nullptr;
    switch (mode) {
    case 0:
        set_lds(k_col<R, 0>, lds);
        QDX_LAUNCH((k_col<R, 0>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, tau, x, none, ctl, cc);
        break;
    case 1:
        set_lds(k_col<R, 1>, lds);
        QDX_LAUNCH((k_col<R, 1>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, tau, x, none, ctl, cc);
        break;
    case 2:
        set_lds(k_col<R, 2>, lds);
        QDX_LAUNCH((k_col<R, 2>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, tau, x, none, ctl, cc);
        break;
    default:
        if (imag) {   // imaginary time: FP64 only
            if constexpr (std::is_same<R, double>::value) {
                set_lds(k_col<R, 3, 8, true>, lds);
                QDX_LAUNCH((k_col<R, 3, 8, true>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, tau, x, none,
                           ctl, cc);
            } else {
                throw std::runtime_error("imaginary time: FP64 only");
            }
            break;
        }
        set_lds(k_col<R, 3>, lds);
        QDX_LAUNCH((k_col<R, 3>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, tau, x, none, ctl, cc);
        break;
    }
    QDX_HIP(hipGetLastError());
}

template <class R, int P>
static void kin_pass_P(hipStream_t s, const GridDev &g, int S, int dim, bool accumulate, const LinePlan &plan,
                       const HamDev<R> &h, const R *src, R *Y, ChebyCtl *ctl, const ConvCheck &cc) {
    const int L = g.N[dim];
    const int W = col_tile(g, S, dim, L, sizeof(cx<R>), P);
    const int nother = g.nd == 3 ? (dim == 1 ? g.N[2] : g.N[1]) : 1;
    const int nb = (g.N[0] / W) * nother * S;
    const int threads = std::max(64, round64(W * (L / P)));
    const size_t lds = col_lds_bytes(W, L, sizeof(cx<R>));
    cx<R> *y = reinterpret_cast<cx<R> *>(Y);
    const cx<R> *in = reinterpret_cast<const cx<R> *>(src);
    LinePlan pl = plan;
    pl.tw_lds = 1;
    if (accumulate) {
        set_lds(k_col<R, 5, P>, lds);
        QDX_LAUNCH((k_col<R, 5, P>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, 0.0, y, in, ctl, cc);
    } else {
        set_lds(k_col<R, 4, P>, lds);
        QDX_LAUNCH((k_col<R, 4, P>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, 0.0, y, in, ctl, cc);
    }
    QDX_HIP(hipGetLastError());
}

template <class R>
void Launch<R>::kin_pass(hipStream_t s, const GridDev &g, int S, int dim, bool accumulate, const LinePlan &plan,
                         const HamDev<R> &h, const R *src, R *Y, ChebyCtl *ctl, const ConvCheck &cc) {
    if (plan.ppt == 4) kin_pass_P<R, 4>(s, g, S, dim, accumulate, plan, h, src, Y, ctl, cc);
    else kin_pass_P<R, 8>(s, g, S, dim, accumulate, plan, h, src, Y, ctl, cc);
}

template <class R, int S>
static void ham_apply_S(hipStream_t st, const GridDev &g, const HamDev<R> &h, const LinePlan *plans_in,
                        const double *src, double *dst, R *X) {
    LinePlan plans[3] = {plans_in[0], g.nd > 1 ? plans_in[1] : LinePlan(), g.nd > 2 ? plans_in[2] : LinePlan()};
    RowLaunch r = row_launch<R>(g, S, plans[0]);
    auto s = reinterpret_cast<const cx<double> *>(src);
    auto d = reinterpret_cast<cx<double> *>(dst);
    auto x = reinterpret_cast<cx<R> *>(X);
    if (g.nd == 1) {
        set_lds(k_apply_row<R, S, true>, r.lds);
        QDX_LAUNCH((k_apply_row<R, S, true>), dim3(r.nb), dim3(r.threads), r.lds, st, g, h, plans[0], r.rb, s,
                           x, d);
        QDX_HIP(hipGetLastError());
        return;
    }
    set_lds(k_row_fwd<R, S>, r.lds);
    QDX_LAUNCH((k_row_fwd<R, S>), dim3(r.nb), dim3(r.threads), r.lds, st, g, plans[0], r.rb, s, x);
    QDX_HIP(hipGetLastError());
    if (g.nd == 2) {
        Launch<R>::col_pass(st, g, S, 1, 2, plans[1], h, 0.0, X, nullptr, ConvCheck());
    } else {
        Launch<R>::col_pass(st, g, S, 1, 0, plans[1], h, 0.0, X, nullptr, ConvCheck());
        Launch<R>::col_pass(st, g, S, 2, 2, plans[2], h, 0.0, X, nullptr, ConvCheck());
        Launch<R>::col_pass(st, g, S, 1, 1, plans[1], h, 0.0, X, nullptr, ConvCheck());
    }
    set_lds(k_apply_row<R, S, false>, r.lds);
    QDX_LAUNCH((k_apply_row<R, S, false>), dim3(r.nb), dim3(r.threads), r.lds, st, g, h, plans[0], r.rb, s, x,
                       d);
    QDX_HIP(hipGetLastError());
}

template <class R>
void Launch<R>::ham_apply(hipStream_t s, const GridDev &g, const HamDev<R> &h, const LinePlan *plans,
                          const double *src, double *dst, R *X) {
    QDX_DISPATCH_S(h.S, (ham_apply_S<R, SS>(s, g, h, plans, src, dst, X)));
}

template <class R, bool ONE_D, int PHASE, bool IMAG, int P>
static void gspo_row_launch_IP(hipStream_t s, const GridDev &g, int S, const HamDev<R> &kin, double tau,
                               const LinePlan &plan, const XGroupDev<R> &pre, const XGroupDev<R> &post, R *psi,
                               R *X) {
    QDX_DISPATCH_S(S, {
        LinePlan p0 = plan;
        RowLaunch r = row_launch<R>(g, SS, p0, SS > 1);
        set_lds(k_gspo_row<R, SS, ONE_D, PHASE, IMAG, P>, r.lds);
        QDX_LAUNCH((k_gspo_row<R, SS, ONE_D, PHASE, IMAG, P>), dim3(r.nb), dim3(r.threads), r.lds, s, g, kin, tau,
                           p0, r.rb, pre, post, reinterpret_cast<cx<R> *>(psi), reinterpret_cast<cx<R> *>(X));
    });
    QDX_HIP(hipGetLastError());
}

// 4 points per thread on grids with nd >= 2 when the plan asks for it
template <class R, bool ONE_D, int PHASE, bool IMAG>
static void gspo_row_launch_I(hipStream_t s, const GridDev &g, int S, const HamDev<R> &kin, double tau,
                              const LinePlan &plan, const XGroupDev<R> &pre, const XGroupDev<R> &post, R *psi,
                              R *X) {
    if constexpr (!ONE_D) {
        if (plan.ppt == 4) return gspo_row_launch_IP<R, ONE_D, PHASE, IMAG, 4>(s, g, S, kin, tau, plan, pre, post, psi, X);
    }
    if (plan.ppt != 8) throw std::runtime_error("GSPO rows: 8 or 4 points per thread (4 for nd >= 2)");
    gspo_row_launch_IP<R, ONE_D, PHASE, IMAG, 8>(s, g, S, kin, tau, plan, pre, post, psi, X);
}

// imag: imaginary time (FP64 only)
template <class R, bool ONE_D, int PHASE>
static void gspo_row_launch(hipStream_t s, const GridDev &g, int S, const HamDev<R> &kin, double tau,
                            const LinePlan &plan, const XGroupDev<R> &pre, const XGroupDev<R> &post, R *psi, R *X,
                            int imag = 0) {
    if (!imag) return gspo_row_launch_I<R, ONE_D, PHASE, false>(s, g, S, kin, tau, plan, pre, post, psi, X);
    if constexpr (std::is_same<R, double>::value)
        return gspo_row_launch_I<R, ONE_D, PHASE, true>(s, g, S, kin, tau, plan, pre, post, psi, X);
    throw std::runtime_error("imaginary time: FP64 only");
}

template <class R>
void Launch<R>::gspo_row_pre(hipStream_t s, const GridDev &g, int S, const LinePlan &p0, const XGroupDev<R> &pre,
                             const R *psi, R *X, int imag) {
    HamDev<R> none;
    XGroupDev<R> empty;
    gspo_row_launch<R, false, 0>(s, g, S, none, 0.0, p0, pre, empty, const_cast<R *>(psi), X, imag);
}

template <class R>
void Launch<R>::gspo_row_post(hipStream_t s, const GridDev &g, int S, const LinePlan &p0, const XGroupDev<R> &post,
                              const R *X, R *psi, int imag) {
    HamDev<R> none;
    XGroupDev<R> empty;
    gspo_row_launch<R, false, 1>(s, g, S, none, 0.0, p0, empty, post, psi, const_cast<R *>(X), imag);
}

template <class R>
void Launch<R>::gspo_1d(hipStream_t s, const GridDev &g, int S, const LinePlan &p0, const HamDev<R> &kin,
                        double tau, const XGroupDev<R> &pre, const XGroupDev<R> &post, R *psi, int imag) {
    gspo_row_launch<R, true, 0>(s, g, S, kin, tau, p0, pre, post, psi, nullptr, imag);
}

template <class R>
void Launch<R>::gspo_xonly(hipStream_t s, const GridDev &g, int S, const XGroupDev<R> &grp, R *psi, int imag) {
    const int threads = 256;
    const int nb = (int)((g.size + threads - 1) / threads);
    if (imag) {
        if constexpr (std::is_same<R, double>::value) {
            QDX_DISPATCH_S(S, {
                QDX_LAUNCH((k_gspo_xonly<R, SS, true>), dim3(nb), dim3(threads), 0, s, g, grp,
                           reinterpret_cast<cx<R> *>(psi));
            });
            QDX_HIP(hipGetLastError());
            return;
        }
        throw std::runtime_error("imaginary time: FP64 only");
    }
    QDX_DISPATCH_S(S, {
        QDX_LAUNCH((k_gspo_xonly<R, SS>), dim3(nb), dim3(threads), 0, s, g, grp,
                           reinterpret_cast<cx<R> *>(psi));
    });
    QDX_HIP(hipGetLastError());
}

template <class R>
void Launch<R>::line_fft(hipStream_t s, const GridDev &g, int S, int dim, const LinePlan &plan, const double *src,
                         R *X) {
    if (dim == 0) {
        QDX_DISPATCH_S(S, {
            LinePlan p0 = plan;
            RowLaunch r = row_launch<R>(g, SS, p0);
            set_lds(k_row_fwd<R, SS>, r.lds);
            QDX_LAUNCH((k_row_fwd<R, SS>), dim3(r.nb), dim3(r.threads), r.lds, s, g, p0, r.rb,
                       reinterpret_cast<const cx<double> *>(src), reinterpret_cast<cx<R> *>(X));
        });
        QDX_HIP(hipGetLastError());
        return;
    }
    if constexpr (!std::is_same<R, double>::value) throw std::runtime_error("line_fft: FP64 only");
    QDX_HIP(hipMemcpyAsync(X, src, (size_t)S * g.size * 2 * sizeof(double), hipMemcpyDeviceToDevice, s));
    col_pass(s, g, S, dim, 0, plan, HamDev<R>(), 0.0, X, nullptr, ConvCheck());
}

template struct Launch<double>;
template struct Launch<float>;

void launch_convert_c128_to_c64(hipStream_t s, const double *src, float *dst, long n) {
    const long n2 = 2 * n;
    const int nb = (int)std::min<long>((n2 + 255) / 256, 8192);
    QDX_LAUNCH(k_c128_to_c64, dim3(nb), dim3(256), 0, s, src, dst, n2);
    QDX_HIP(hipGetLastError());
}

void launch_scale_axpy(hipStream_t s, double *y, const double *x, long n, double sa, double br, double bi) {
    const int nb = (int)std::min<long>((n + 255) / 256, 8192);
    QDX_LAUNCH(k_scale_axpy, dim3(std::max(nb, 1)), dim3(256), 0, s, y, x, n, sa, br, bi);
    QDX_HIP(hipGetLastError());
}

void launch_convert_c64_to_c128(hipStream_t s, const float *src, double *dst, long n) {
    const long n2 = 2 * n;
    const int nb = (int)std::min<long>((n2 + 255) / 256, 8192);
    QDX_LAUNCH(k_c64_to_c128, dim3(nb), dim3(256), 0, s, src, dst, n2);
    QDX_HIP(hipGetLastError());
}

template <class R>
static void overlap_impl(hipStream_t s, const R *a, const R *b, long n, int S, double *partials, double *out) {
    const int threads = 256;
    int nbx = (int)std::min<long>((n + threads - 1) / threads, 1024);
    if (nbx < 1) nbx = 1;
    QDX_LAUNCH((k_overlap_stage1<R>), dim3(nbx, S), dim3(threads), 0, s,
                       reinterpret_cast<const cx<R> *>(a), reinterpret_cast<const cx<R> *>(b), n, partials);
    QDX_HIP(hipGetLastError());
    QDX_LAUNCH(k_overlap_stage2, dim3(1), dim3(256), 0, s, partials, nbx, S, out);
    QDX_HIP(hipGetLastError());
}

void launch_wnorm(hipStream_t s, const double *a, long n, int S, const GridDev &g, int dim, double base, double step,
                  const double *ktab, double *partials, double *out) {
    if (dim < 0 || dim >= g.nd) throw std::runtime_error("dimension out of range");
    int stride = 1;
    for (int d = 0; d < dim; ++d) stride *= g.N[d];
    const int threads = 256;
    int nbx = (int)std::min<long>((n + threads - 1) / threads, 1024);
    if (nbx < 1) nbx = 1;
    QDX_LAUNCH(k_wnorm_stage1, dim3(nbx, S), dim3(threads), 0, s, reinterpret_cast<const cx<double> *>(a), n, stride,
               g.N[dim], base, step, ktab, partials);
    QDX_HIP(hipGetLastError());
    QDX_LAUNCH(k_overlap_stage2, dim3(1), dim3(256), 0, s, partials, nbx, S, out);
    QDX_HIP(hipGetLastError());
}

void launch_overlap(hipStream_t s, const double *a, const double *b, long n, int S, double *partials, double *out) {
    overlap_impl<double>(s, a, b, n, S, partials, out);
}

void launch_overlap_c64(hipStream_t s, const float *a, const float *b, long n, int S, double *partials,
                        double *out) {
    overlap_impl<float>(s, a, b, n, S, partials, out);
}

}  // namespace qdx
