// gfx950 strided line passes (kinetic parts along dimensions >= 1, GSPO
// kinetic split operator, forward/inverse transforms) and H applied to a
// complex128 vector (Operator::Apply).
#include "qdx_kernels_common.hpp"

namespace qdx {

// ---------------------------------------------------------------------------
// Strided line pass along dimension `dim` (1 or 2) on X, in place.  A block
// owns a tile of W adjacent dimension-0 positions for one state and one index
// of the remaining dimension; the W lanes of a butterfly touch W adjacent
// words in LDS and W adjacent complex values in memory.
// MODE 0: forward, 1: inverse, 2: forward * T/N inverse,
//      3: forward * exp(-i tau T)/N inverse,
//      4: X = IFFT(T_dim/N_dim FFT(src)), the dimension's part of the
//         separable kinetic operator, 5: X += the same.
// CMP (modes 4 and 5): src and X in the compact 40-bit term storage.
template <class R, int MODE, int P = 8, bool IMAG = false, bool CMP = false>
__global__ void __launch_bounds__(512) k_col(GridDev g, int S, int dim, int W, LinePlan p, HamDev<R> h, double tau,
                                              cx<R> *__restrict__ X, const cx<R> *__restrict__ src,
                                              ChebyCtl *ctl, ConvCheck cc, const double *__restrict__ ek) {
    if (ctl && ctl->done) return;
    extern __shared__ __align__(16) char smem_raw[];
    cx<R> *smem = reinterpret_cast<cx<R> *>(smem_raw);
    __shared__ double bc;
    const int N0 = g.N[0];
    const int L = g.N[dim];
    const int TL = L / P;
    const long estride = dim_stride(g, dim);
    const int nother = col_nother(g, dim);
    const int tiles0 = N0 / W;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int c0 = (b % tiles0) * W;
    const int rest = b / tiles0;
    const int oth = rest % nother;
    const int st = rest / nother;
    const int c = threadIdx.x % W;
    const int lt = threadIdx.x / W;
    const bool active = lt < TL;
    int kk[kMaxDims] = {0, 0, 0, 0};   // k-space index per dimension (dimensions 0 and dim set below)
    const long boff = (long)st * g.size + col_other(g, dim, oth, kk) + c0 + c;   // element of this thread's column
    const long nT = (long)S * g.size;
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
        const cx<R> *in = MODE >= 4 ? src : X;
#pragma unroll
        for (int q = 0; q < P; ++q) v[q] = tld<CMP>(in, boff + (long)(lt + q * TL) * estride, nT);
        if (MODE == 5) {
#pragma unroll
            for (int q = 0; q < P; ++q) yv[q] = tld<CMP>(X, boff + (long)(lt + q * TL) * estride, nT);
        }
    }
    // kinetic table values of this thread's elements, requested early.
    // Real-time GSPO (mode 3) with ek: the factor exp(-i tau T(k)) / N as
    // inv_n times the product over dimensions of exp(-i tau T_d(k_d)) from
    // the per-state tables ek (state st at 2 st sum_d N_d doubles, dimension
    // d after the tables of the dimensions before it, (re, im) pairs)
    double tq[P];
    cx<double> fq[P];
    const bool etab = MODE == 3 && !IMAG && ek != nullptr;
    if (etab && active) {
        long tot = 0, offd = 0;
        long off[kMaxDims] = {0, 0, 0, 0};
#pragma unroll
        for (int e = 0; e < kMaxDims; ++e) {
            off[e] = tot;
            if (e < g.nd) tot += g.N[e];
            if (e == dim) offd = off[e];
        }
        const double *es = ek + 2 * (long)st * tot;
        kk[0] = c0 + c;
        cx<double> fo = mk<double>(h.inv_n, 0.0);
#pragma unroll
        for (int e = 0; e < kMaxDims; ++e) {
            if (e < g.nd && e != dim) {
                const double *q = es + 2 * (off[e] + kk[e]);
                fo = cmul(fo, mk<double>(q[0], q[1]));
            }
        }
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const double *t = es + 2 * (offd + lt + q * TL);
            fq[q] = cmul(fo, mk<double>(t[0], t[1]));
        }
    } else if (MODE >= 2 && active) {
        const bool has_t = h.tk[st][0] != nullptr;
        kk[0] = c0 + c;
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int kd = lt + q * TL;
#pragma unroll
            for (int e = 1; e < kMaxDims; ++e)
                if (e == dim) kk[e] = kd;
            if (MODE >= 4) tq[q] = has_t ? (0.0 + h.tk[st][dim][kd]) * h.inv_nd[dim] : 0.0;
            else tq[q] = has_t ? kin_sum<R>(h, st, g, kk) : 0.0;
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
                } else if (etab) {
                    v[q] = from_d<R>(cmul(to_d(v[q]), fq[q]));
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
        for (int q = 0; q < P; ++q) tst<CMP>(X, boff + (long)(lt + q * TL) * estride, nT, v[q]);
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
        if (h.Wi[st]) {   // negative imaginary potential: -i Wi psi
            const double c = (double)h.Wi[st][gi];
            const cx<double> ps = src[base + gi];
            hp.r += c * ps.i;
            hp.i -= c * ps.r;
        }
        dst[base + gi] = hp;
    }
}

template <class R, int P>
static void col_pass_P4(hipStream_t s, const GridDev &g, int S, int dim, int mode, const LinePlan &plan,
                        const HamDev<R> &h, double tau, R *X, ChebyCtl *ctl, const ConvCheck &cc, int imag,
                        const double *ek) {
    const int L = g.N[dim];
    const int W = col_tile(g, S, dim, L, sizeof(cx<R>), P);
    const int nb = (g.N[0] / W) * col_nother(g, dim) * S;
    const int threads = std::max(64, round64(W * (L / P)));
    const size_t lds = col_lds_bytes(W, L, sizeof(cx<R>));
    cx<R> *x = reinterpret_cast<cx<R> *>(X);
    LinePlan pl = plan;
    pl.tw_lds = 1;
    const cx<R> *none = nullptr;
    const double *ekp = mode == 3 && !imag ? ek : nullptr;
    switch (mode) {
    case 0:
        set_lds(k_col<R, 0, P>, lds);
        QDX_LAUNCH((k_col<R, 0, P>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, tau, x, none, ctl, cc, ekp);
        break;
    case 1:
        set_lds(k_col<R, 1, P>, lds);
        QDX_LAUNCH((k_col<R, 1, P>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, tau, x, none, ctl, cc, ekp);
        break;
    case 3:
        if (imag) {
            if constexpr (std::is_same<R, double>::value) {
                set_lds(k_col<R, 3, P, true>, lds);
                QDX_LAUNCH((k_col<R, 3, P, true>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, tau, x,
                           none, ctl, cc, ekp);
            } else {
                throw std::runtime_error("imaginary time: FP64 only");
            }
            break;
        }
        set_lds(k_col<R, 3, P>, lds);
        QDX_LAUNCH((k_col<R, 3, P>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, tau, x, none, ctl, cc, ekp);
        break;
    default:
        throw std::runtime_error("line pass: mode 0, 1 or 3 at 4 points per thread");
    }
    QDX_HIP(hipGetLastError());
}

template <class R>
void Launch<R>::col_pass(hipStream_t s, const GridDev &g, int S, int dim, int mode, const LinePlan &plan,
                         const HamDev<R> &h, double tau, R *X, ChebyCtl *ctl, const ConvCheck &cc, int imag,
                         const double *ek) {
    if (plan.ppt == 4) return col_pass_P4<R, 4>(s, g, S, dim, mode, plan, h, tau, X, ctl, cc, imag, ek);
    const int L = g.N[dim];
    const int W = col_tile(g, S, dim, L, sizeof(cx<R>));
    const int nb = (g.N[0] / W) * col_nother(g, dim) * S;
    const int threads = std::max(64, round64(W * (L / 8)));
    const size_t lds = col_lds_bytes(W, L, sizeof(cx<R>));
    cx<R> *x = reinterpret_cast<cx<R> *>(X);
    LinePlan pl = plan;
    pl.tw_lds = 1;
    const cx<R> *none = nullptr;
    const double *ekp = mode == 3 && !imag ? ek : nullptr;
    switch (mode) {
    case 0:
        set_lds(k_col<R, 0>, lds);
        QDX_LAUNCH((k_col<R, 0>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, tau, x, none, ctl, cc, ekp);
        break;
    case 1:
        set_lds(k_col<R, 1>, lds);
        QDX_LAUNCH((k_col<R, 1>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, tau, x, none, ctl, cc, ekp);
        break;
    case 2:
        set_lds(k_col<R, 2>, lds);
        QDX_LAUNCH((k_col<R, 2>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, tau, x, none, ctl, cc, ekp);
        break;
    default:
        if (imag) {   // imaginary time: FP64 only
            if constexpr (std::is_same<R, double>::value) {
                set_lds(k_col<R, 3, 8, true>, lds);
                QDX_LAUNCH((k_col<R, 3, 8, true>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, tau, x, none,
                           ctl, cc, ekp);
            } else {
                throw std::runtime_error("imaginary time: FP64 only");
            }
            break;
        }
        set_lds(k_col<R, 3>, lds);
        QDX_LAUNCH((k_col<R, 3>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, tau, x, none, ctl, cc, ekp);
        break;
    }
    QDX_HIP(hipGetLastError());
}

template <class R, int P>
static void kin_pass_P(hipStream_t s, const GridDev &g, int S, int dim, bool accumulate, const LinePlan &plan,
                       const HamDev<R> &h, const R *src, R *Y, ChebyCtl *ctl, const ConvCheck &cc, bool cmp) {
    const int L = g.N[dim];
    const int W = col_tile(g, S, dim, L, sizeof(cx<R>), P);
    const int nb = (g.N[0] / W) * col_nother(g, dim) * S;
    const int threads = std::max(64, round64(W * (L / P)));
    const size_t lds = col_lds_bytes(W, L, sizeof(cx<R>));
    cx<R> *y = reinterpret_cast<cx<R> *>(Y);
    const cx<R> *in = reinterpret_cast<const cx<R> *>(src);
    const double *ekp = nullptr;
    LinePlan pl = plan;
    pl.tw_lds = 1;
    if (cmp) {
        if constexpr (std::is_same<R, double>::value) {
            if (accumulate) {
                set_lds(k_col<R, 5, P, false, true>, lds);
                QDX_LAUNCH((k_col<R, 5, P, false, true>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, 0.0, y,
                           in, ctl, cc, ekp);
            } else {
                set_lds(k_col<R, 4, P, false, true>, lds);
                QDX_LAUNCH((k_col<R, 4, P, false, true>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, 0.0, y,
                           in, ctl, cc, ekp);
            }
        } else {
            throw std::runtime_error("compact terms: FP64 only");
        }
    } else if (accumulate) {
        set_lds(k_col<R, 5, P>, lds);
        QDX_LAUNCH((k_col<R, 5, P>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, 0.0, y, in, ctl, cc, ekp);
    } else {
        set_lds(k_col<R, 4, P>, lds);
        QDX_LAUNCH((k_col<R, 4, P>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, pl, h, 0.0, y, in, ctl, cc, ekp);
    }
    QDX_HIP(hipGetLastError());
}

template <class R>
void Launch<R>::kin_pass(hipStream_t s, const GridDev &g, int S, int dim, bool accumulate, const LinePlan &plan,
                         const HamDev<R> &h, const R *src, R *Y, ChebyCtl *ctl, const ConvCheck &cc, bool cmp) {
    if (plan.ppt == 4) kin_pass_P<R, 4>(s, g, S, dim, accumulate, plan, h, src, Y, ctl, cc, cmp);
    else kin_pass_P<R, 8>(s, g, S, dim, accumulate, plan, h, src, Y, ctl, cc, cmp);
}

template <class R, int S>
static void ham_apply_S(hipStream_t st, const GridDev &g, const HamDev<R> &h, const LinePlan *plans_in,
                        const double *src, double *dst, R *X) {
    LinePlan plans[kMaxDims];
    for (int d = 0; d < g.nd; ++d) plans[d] = plans_in[d];
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
    // forward transforms along dimensions 1 .. nd-2, the last dimension's
    // FFT * T(k)/N * IFFT, inverse transforms back down to dimension 1
    for (int d = 1; d < g.nd - 1; ++d) Launch<R>::col_pass(st, g, S, d, 0, plans[d], h, 0.0, X, nullptr, ConvCheck());
    Launch<R>::col_pass(st, g, S, g.nd - 1, 2, plans[g.nd - 1], h, 0.0, X, nullptr, ConvCheck());
    for (int d = g.nd - 2; d >= 1; --d) Launch<R>::col_pass(st, g, S, d, 1, plans[d], h, 0.0, X, nullptr, ConvCheck());
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

#define QDX_INST_LINES(R)                                                                                      \
    template void Launch<R>::col_pass(hipStream_t, const GridDev &, int, int, int, const LinePlan &,               \
                                      const HamDev<R> &, double, R *, ChebyCtl *, const ConvCheck &, int,           \
                                      const double *);                                                              \
    template void Launch<R>::kin_pass(hipStream_t, const GridDev &, int, int, bool, const LinePlan &,              \
                                      const HamDev<R> &, const R *, R *, ChebyCtl *, const ConvCheck &, bool);      \
    template void Launch<R>::ham_apply(hipStream_t, const GridDev &, const HamDev<R> &, const LinePlan *,          \
                                       const double *, double *, R *);                                              \
    template void Launch<R>::line_fft(hipStream_t, const GridDev &, int, int, const LinePlan &, const double *, R *);
QDX_INST_LINES(double)
QDX_INST_LINES(float)

}  // namespace qdx
