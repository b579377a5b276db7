// gfx950 Chebyshev kernels of the single propagation (OCheby::real_time_step
// and imag_time_step, OCheby.cpp:92-297): the init pass, the per-order row
// kernel with the fused epilogue and the whole-step kernel of 1D lines.  The
// thread layout of the line kernels is described in qdx_kernels.hip.  The
// launchers are instantiated per precision in qdx_cheby64.hip and
// qdx_cheby32.hip (QDX_INST_CHEBY).
#pragma once

#include "qdx_kernels_common.hpp"

namespace qdx {

// ---------------------------------------------------------------------------
// Chebyshev init: phi0 = psi, norm0 = sum |psi|^2, X = FFT_0(phi0) (nd >= 2)
// CMP: phi0 in the compact term storage (tld/tst)
template <class R, int S, bool HAS_X, bool CMP = false>
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
            tst<CMP>(phi0, base + B.gidx(q), (long)S * g.size, v[q]);
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
// time (the recurrence signs of OCheby::imag_time_step).  CMP: the terms
// phi and the kinetic part Y in the compact 40-bit storage (R double).
// NIP: single-state real time with a negative imaginary potential: H adds
// -i Wi(x) and the offset its imaginary part (b = (H - offset) phi / R with
// complex offset, OCheby.cpp:104-105, 168).
template <class R, int S, bool ONE_D, bool PRO, int NT, int P = 8, bool IMAG = false, bool CMP = false,
          bool NIP = false>
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
    const long nT = (long)S * g.size;   // elements of each term buffer

    // the line of phi_{k-1} this thread transforms; the epilogue operands
    // (psi, phi_{k-2}, Y, phi_{k-1} of every state, W) are read after the
    // transforms, so that only the line stays live across them
    cx<R> v[P];
    if (B.active) {
#pragma unroll
        for (int q = 0; q < P; ++q) v[q] = tld<CMP>(phim1, base + B.gidx(q), nT);
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
    const R *Wip = NIP ? h.Wi[st] : nullptr;
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
        double wi[4];   // NIP: Wi + Im offset
    };
    auto load_half = [&](Half &H, int h0) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const long gi = B.gidx(h0 + u);
            if (upd) H.ps[u] = psi[base + gi];
            if (!PRO) H.m2[u] = tld<CMP>(phim2, base + gi, nT);
            if (!ONE_D) H.yv[u] = tld<CMP>(X, base + gi, nT);
#pragma unroll
            for (int j = 0; j < S; ++j) {
                H.pm[j][u] = tld<CMP>(phim1, (long)j * g.size + gi, nT);
                double w = (j == st) ? uni : 0.0;
                if (Wp[j]) w += (double)Wp[j][gi];
                if (Mp[j]) w += Ed[j] * (double)Mp[j][gi];
                H.wv[j][u] = w;
            }
            if (NIP) H.wi[u] = ord.offset_i + (Wip ? (double)Wip[gi] : 0.0);
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
            // b = (H phi - offset phi) / R; NIP: H - offset carries
            // -i (Wi + Im offset) on the diagonal
            cx<double> b = mk<double>((hp.r - ord.offset * p.r) * ord.inv_scaling,
                                      (hp.i - ord.offset * p.i) * ord.inv_scaling);
            if (NIP) {
                const double c = H.wi[u] * ord.inv_scaling;
                b.r += c * p.i;
                b.i -= c * p.r;
            }
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
            tst<CMP>(phim2, base + gi, nT, from_d<R>(pk));
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
// exp(-i offset dt), staged into LDS (after the line buffers and the twiddle
// table) before the first order.  Nothing is written to global memory inside
// the order loop (a global store before a barrier makes the barrier wait for
// it): the last order, its sum of |psi|^2 (slot [k % 2] * kMaxPartials of the
// multi-kernel layout) and the break are recorded after the loop; the step
// start's sum goes to slot 2 * kMaxPartials.  A block of one wave reduces
// with shuffles alone.
template <class R, int NT, bool NIP = false, int P = 8>
__global__ void __launch_bounds__(NT) k_cheby_line(GridDev g, HamDev<R> h, LinePlan p0, const double *coef,
                                                   int order, double offset, double offset_i, double inv_scaling,
                                                   double prec, cx<double> *__restrict__ psi, ChebyCtl *ctl,
                                                   double *__restrict__ partials) {
    extern __shared__ __align__(16) char smem_raw[];
    __shared__ double red[16];
    __shared__ double bc;
    RowBlock<R, P> B(g, 1, 1, reinterpret_cast<cx<R> *>(smem_raw), p0);
    double *cs = reinterpret_cast<double *>(B.after);   // 2 * order coefficients
    const bool one_wave = blockDim.x <= 64;
    const double *tk0 = h.tk[0][0];
    const bool has_t = tk0 != nullptr;
    const R *Wp = h.W[0][0];
    const R *Mp = h.mu[0][0];
    const R *Wip = NIP ? h.Wi[0] : nullptr;
    const double Ed = h.Edip[0][0];
    const double uni = h.unity[0];
    for (int i = threadIdx.x; i < 2 * order; i += blockDim.x) cs[i] = coef[i];
    cx<double> ps[P];
    cx<R> pm1[P], pm2[P], v[P];
    double wv[P], tq[P], wi[P];
    double acc = 0;
    if (B.active) {
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const long gi = B.gidx(q);
            ps[q] = psi[gi];
            acc += ps[q].r * ps[q].r + ps[q].i * ps[q].i;
            pm1[q] = from_d<R>(ps[q]);
            double w = uni;
            if (Wp) w += (double)Wp[gi];
            if (Mp) w += Ed * (double)Mp[gi];
            wv[q] = w;
            if (NIP) wi[q] = (offset_i + (Wip ? (double)Wip[gi] : 0.0)) * inv_scaling;
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
    }
    __syncthreads();   // also publishes the coefficients
    norm0 = bc;
    B.stage(p0);
    const cx<double> c0 = mk<double>(cs[0], cs[1]);
    int last = 1;           // the last order whose terms are in psi
    double a_last = 0;      // its sum of |psi|^2 (orders >= 2)
    bool broke = false;
    for (int k = 1; k < order; ++k) {
        if (k >= 3 && prec > 0) {
            // break test of order k - 1 on its sum (uniform over the block)
            if (fabs(norm0 - a_last) / norm0 < prec) {
                broke = true;
                break;   // psi holds the series through order k - 1
            }
        }
        last = k;
#pragma unroll
        for (int q = 0; q < P; ++q) v[q] = pm1[q];
        B.fft(v, p0, -1);
        if (B.active) {
#pragma unroll
            for (int q = 0; q < P; ++q) v[q] = mk<R>((R)((double)v[q].r * tq[q]), (R)((double)v[q].i * tq[q]));
        }
        B.fft(v, p0, +1);
        // psi update at this order: every order with the break test, the
        // last three terms every third order of the fixed-length series
        int nterm = 1;
        if (k >= 2 && !(prec > 0)) {
            const int m = (k - 1) % 3;
            nterm = m == 0 ? 3 : (k == order - 1 ? m : 0);
        }
        const cx<double> ck = mk<double>(cs[2 * k], cs[2 * k + 1]);
        const cx<double> cm1 = k >= 2 ? mk<double>(cs[2 * (k - 1)], cs[2 * (k - 1) + 1]) : mk<double>(0, 0);
        const cx<double> cm2 = k >= 3 ? mk<double>(cs[2 * (k - 2)], cs[2 * (k - 2) + 1]) : mk<double>(0, 0);
        acc = 0;
        if (B.active) {
#pragma unroll
            for (int q = 0; q < P; ++q) {
                cx<double> hp = mk<double>(0, 0);
                if (has_t) hp = to_d(v[q]);
                const cx<double> p = to_d(pm1[q]);
                hp.r += wv[q] * p.r;
                hp.i += wv[q] * p.i;
                cx<double> b = mk<double>((hp.r - offset * p.r) * inv_scaling, (hp.i - offset * p.i) * inv_scaling);
                if (NIP) {
                    b.r += wi[q] * p.i;
                    b.i -= wi[q] * p.r;
                }
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
            if (one_wave) {
                a_last = __shfl(wave_sum(acc), 0, 64);   // the block sum, in every lane
            } else {
                const double a = block_sum(acc, red);
                if (threadIdx.x == 0) bc = a;
                __syncthreads();
                a_last = bc;
            }
        }
    }
    if (threadIdx.x == 0) {
        ctl->last_order = last;
        if (last >= 2) partials[(last % 2) * kMaxPartials] = a_last;
        if (broke) {
            ctl->done = 1;
            ctl->accnorm = a_last;
        }
    }
    if (B.active) {
#pragma unroll
        for (int q = 0; q < P; ++q) psi[B.gidx(q)] = ps[q];
    }
}

template <class R, int S>
static int cheby_init_S(hipStream_t st, const GridDev &g, const LinePlan &plan, const double *psi, R *phi0,
                        R *X, ChebyCtl *ctl, double *partials, bool cmp) {
    LinePlan p0 = plan;
    p0.ppt = 8;   // the init kernel holds 8 points per thread
    RowLaunch r = row_launch<R>(g, S, p0);
    if (r.nb > kMaxPartials) throw std::runtime_error("grid too large for the partial-sum buffer");
    auto cpsi = reinterpret_cast<const cx<double> *>(psi);
    auto cphi = reinterpret_cast<cx<R> *>(phi0);
    auto cx_ = reinterpret_cast<cx<R> *>(X);
    if (cmp) {
        if constexpr (std::is_same<R, double>::value && S <= 2) {
            if (X) throw std::runtime_error("compact terms: the init kernel writes phi_0 only");
            QDX_LAUNCH((k_cheby_init<R, S, false, true>), dim3(r.nb), dim3(r.threads), 0, st, g, p0, r.rb, cpsi,
                       cphi, cx_, ctl, partials);
        } else {
            throw std::runtime_error("compact terms: FP64 with one or two states");
        }
    } else if (X) {
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

template <class R>
int Launch<R>::cheby_init(hipStream_t s, const GridDev &g, int S, const LinePlan &p0, const double *psi, R *phi0,
                          R *X, ChebyCtl *ctl, double *partials, bool cmp) {
    int nb = 0;
    QDX_DISPATCH_S(S, (nb = cheby_init_S<R, SS>(s, g, p0, psi, phi0, X, ctl, partials, cmp)));
    return nb;
}

template <class R, int S, bool ONE_D, bool PRO, int P>
static int cheby_row_SKP(hipStream_t st, const GridDev &g, const HamDev<R> &h, const LinePlan &plan,
                         const ChebyOrder &ord, const ConvCheck &cc, bool imag, const R *phim1, R *phim2, R *X, double *psi,
                         ChebyCtl *ctl, double *partials, const double *init_part, int nb_init, bool cmp) {
    LinePlan p0 = plan;
    RowLaunch r = row_launch<R>(g, S, p0);
    if (r.nb > kMaxPartials) throw std::runtime_error("grid too large for the partial-sum buffer");
    auto cphim1 = reinterpret_cast<const cx<R> *>(phim1);
    auto cphim2 = reinterpret_cast<cx<R> *>(phim2);
    auto cX = reinterpret_cast<cx<R> *>(X);
    auto cpsi = reinterpret_cast<cx<double> *>(psi);
    if (ord.offset_i != 0 || h.Wi[0] != nullptr) {
        // negative imaginary potential: one state, real time (512-thread
        // instances; the block never exceeds 512 threads)
        if constexpr (S == 1) {
            if (imag) throw std::runtime_error("negative imaginary potential: real time only");
            if (cmp) {
                if constexpr (std::is_same<R, double>::value && !ONE_D) {
                    set_lds(k_cheby_row<R, 1, ONE_D, PRO, 512, P, false, true, true>, r.lds);
                    QDX_LAUNCH((k_cheby_row<R, 1, ONE_D, PRO, 512, P, false, true, true>), dim3(r.nb),
                               dim3(r.threads), r.lds, st, g, h, p0, ord, cc, r.rb, cphim1, cphim2, cX, cpsi, ctl,
                               partials, init_part, nb_init);
                } else {
                    throw std::runtime_error("compact terms: FP64, nd >= 2");
                }
            } else {
                set_lds(k_cheby_row<R, 1, ONE_D, PRO, 512, P, false, false, true>, r.lds);
                QDX_LAUNCH((k_cheby_row<R, 1, ONE_D, PRO, 512, P, false, false, true>), dim3(r.nb), dim3(r.threads),
                           r.lds, st, g, h, p0, ord, cc, r.rb, cphim1, cphim2, cX, cpsi, ctl, partials, init_part,
                           nb_init);
            }
        } else {
            throw std::runtime_error("negative imaginary potential: single-state Hamiltonians only");
        }
    } else if (cmp) {   // compact terms: FP64 arithmetic, real time, nd >= 2, S <= 2
        if constexpr (std::is_same<R, double>::value && !ONE_D && S <= 2) {
            if (imag) throw std::runtime_error("compact terms: real time only");
            if (r.threads <= 256) {
                set_lds(k_cheby_row<R, S, ONE_D, PRO, 256, P, false, true>, r.lds);
                QDX_LAUNCH((k_cheby_row<R, S, ONE_D, PRO, 256, P, false, true>), dim3(r.nb), dim3(r.threads), r.lds,
                           st, g, h, p0, ord, cc, r.rb, cphim1, cphim2, cX, cpsi, ctl, partials, init_part, nb_init);
            } else {
                set_lds(k_cheby_row<R, S, ONE_D, PRO, 512, P, false, true>, r.lds);
                QDX_LAUNCH((k_cheby_row<R, S, ONE_D, PRO, 512, P, false, true>), dim3(r.nb), dim3(r.threads), r.lds,
                           st, g, h, p0, ord, cc, r.rb, cphim1, cphim2, cX, cpsi, ctl, partials, init_part, nb_init);
            }
        } else {
            throw std::runtime_error("compact terms: FP64, nd >= 2, one or two states");
        }
    } else if (imag) {   // imaginary time: FP64 only, the 512-thread instance
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
                        ChebyCtl *ctl, double *partials, const double *init_part, int nb_init, bool cmp) {
    if (plan.ppt == 4) {
        if constexpr (!ONE_D)
            return cheby_row_SKP<R, S, ONE_D, PRO, 4>(st, g, h, plan, ord, cc, imag, phim1, phim2, X, psi, ctl, partials,
                                                      init_part, nb_init, cmp);
        throw std::runtime_error("4 points per thread: grids with nd >= 2 only");
    }
    return cheby_row_SKP<R, S, ONE_D, PRO, 8>(st, g, h, plan, ord, cc, imag, phim1, phim2, X, psi, ctl, partials, init_part,
                                              nb_init, cmp);
}

template <class R, int S>
static int cheby_row_S(hipStream_t st, const GridDev &g, const HamDev<R> &h, const LinePlan &p0,
                       const ChebyOrder &ord, const ConvCheck &cc, bool imag, const R *phim1, R *phim2, R *X, double *psi,
                       ChebyCtl *ctl, double *partials, const double *ip, int ni, bool cmp) {
    const bool one = g.nd == 1;
    const bool pro = ord.k == 1;
    if (one && pro)
        return cheby_row_SK<R, S, true, true>(st, g, h, p0, ord, cc, imag, phim1, phim2, X, psi, ctl, partials, ip, ni,
                                              cmp);
    if (one)
        return cheby_row_SK<R, S, true, false>(st, g, h, p0, ord, cc, imag, phim1, phim2, X, psi, ctl, partials, ip, ni,
                                               cmp);
    if (pro)
        return cheby_row_SK<R, S, false, true>(st, g, h, p0, ord, cc, imag, phim1, phim2, X, psi, ctl, partials, ip, ni,
                                               cmp);
    return cheby_row_SK<R, S, false, false>(st, g, h, p0, ord, cc, imag, phim1, phim2, X, psi, ctl, partials, ip, ni,
                                            cmp);
}

template <class R>
int Launch<R>::cheby_row(hipStream_t s, const GridDev &g, const HamDev<R> &h, const LinePlan &p0,
                         const ChebyOrder &ord, const ConvCheck &cc, bool imag, const R *phim1, R *phim2, R *X, double *psi,
                         ChebyCtl *ctl, double *partials, const double *init_part, int nb_init, bool cmp) {
    int nb = 0;
    QDX_DISPATCH_S(h.S, (nb = cheby_row_S<R, SS>(s, g, h, p0, ord, cc, imag, phim1, phim2, X, psi, ctl, partials,
                                                 init_part, nb_init, cmp)));
    return nb;
}

template <class R>
void Launch<R>::cheby_line(hipStream_t s, const GridDev &g, const HamDev<R> &h, const LinePlan &p0,
                           const double *coef, int order, double offset, double offset_i, double inv_scaling,
                           double prec, double *psi, ChebyCtl *ctl, double *partials) {
    const int L = g.N[0], P = p0.ppt == 4 ? 4 : 8, TL = L / P;
    if (g.nd != 1 || h.S != 1 || L > kMaxLine) throw std::runtime_error("whole-step line kernel: one 1D line");
    LinePlan p = p0;
    const size_t lcoef = (size_t)2 * order * sizeof(double);   // coefficients staged in LDS
    size_t lds = row_lds_bytes<R>(L, 1, true) + lcoef;
    p.tw_lds = lds <= kLdsBytes ? 1 : 0;
    if (!p.tw_lds) lds = row_lds_bytes<R>(L, 1, false) + lcoef;
    const int threads = std::max(64, round64(TL));
    auto cpsi = reinterpret_cast<cx<double> *>(psi);
    const bool nip = offset_i != 0 || h.Wi[0] != nullptr;   // negative imaginary potential
    const double oi = nip ? offset_i : 0.0;
#define QDX_LINE(NT_, NIP_, P_)                                                                              \
    do {                                                                                                     \
        set_lds(k_cheby_line<R, NT_, NIP_, P_>, lds);                                                        \
        QDX_LAUNCH((k_cheby_line<R, NT_, NIP_, P_>), dim3(1), dim3(threads), lds, s, g, h, p, coef, order,   \
                   offset, oi, inv_scaling, prec, cpsi, ctl, partials);                                      \
    } while (0)
    if (P == 4) {   // short lines: 4 points per thread, at most one 256-thread block
        if (threads > 256) throw std::runtime_error("whole-step line kernel: 4 points per thread up to 1024 points");
        if (nip) QDX_LINE(256, true, 4);
        else QDX_LINE(256, false, 4);
    } else if (nip && threads <= 256) {
        QDX_LINE(256, true, 8);
    } else if (nip) {
        QDX_LINE(512, true, 8);
    } else if (threads <= 256) {
        QDX_LINE(256, false, 8);
    } else {
        QDX_LINE(512, false, 8);
    }
#undef QDX_LINE
    QDX_HIP(hipGetLastError());
}

#define QDX_INST_CHEBY(R)                                                                                      \
    template int Launch<R>::cheby_init(hipStream_t, const GridDev &, int, const LinePlan &, const double *, R *, R *, \
                                       ChebyCtl *, double *, bool);                                                    \
    template int Launch<R>::cheby_row(hipStream_t, const GridDev &, const HamDev<R> &, const LinePlan &,             \
                                      const ChebyOrder &, const ConvCheck &, bool, const R *, R *, R *, double *,    \
                                      ChebyCtl *, double *, const double *, int, bool);                             \
    template void Launch<R>::cheby_line(hipStream_t, const GridDev &, const HamDev<R> &, const LinePlan &,           \
                                        const double *, int, double, double, double, double, double *, ChebyCtl *,  \
                                        double *);

}  // namespace qdx
