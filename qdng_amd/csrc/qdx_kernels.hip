// gfx950 kernels of the grid-propagation hot path.
//
// Chebyshev (OCheby::real_time_step, OCheby.cpp:92-197) runs per order as
//   nd == 1 : one kernel  -- load phi_{k-1} lines into LDS, FFT, *T/N, IFFT,
//             fused potential/dipole/coupling multiply, recurrence, FP64
//             accumulation of psi and of sum |psi|^2, device-side break test;
//   nd >= 2 : line passes -- strided FFT/T/IFFT passes along dimensions
//             >= 1 on the scratch X (which holds FFT_0(phi_{k-1})), then the
//             row kernel: IFFT_0, the same fused epilogue, and FFT_0(phi_k)
//             back into X for the next order.
// GSPO (OGSPO.cpp:114-139) uses the same line machinery with the x-space
// split operators exp(-i tau W(x)) evaluated per point in registers.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <stdexcept>

#include "qdx_device.hpp"
#include "qdx_impl.hpp"

namespace qdx {

#define QDX_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") +         \
                                                       hipGetErrorString(e_) + " at " #call); \
    } while (0)

// ---------------------------------------------------------------------------
// helpers

// kinetic value T_i(k) at a flat k-space index (dimension 0 fastest), summed
// over dimensions in dimension order as Diff::_InitKspace does (Diff.cpp:120-139)
template <class R>
__device__ __forceinline__ double kin_sum(const HamDev<R> &h, int st, const GridDev &g, long idx) {
    const double *const *tk = h.tk[st];
    const int k0 = (int)(idx % g.N[0]);
    double t = 0.0 + tk[0][k0];
    if (g.nd > 1) {
        const long r = idx / g.N[0];
        t += tk[1][(int)(r % g.N[1])];
        if (g.nd > 2) t += tk[2][(int)(r / g.N[1])];
    }
    return t;
}

// (H_x phi)_i: potential, unity and dipole part at grid index gi
template <class R, int S>
__device__ __forceinline__ void xspace_apply(const HamDev<R> &h, long gi, const cx<double> (&p)[S],
                                             cx<double> (&hp)[S]) {
#pragma unroll
    for (int i = 0; i < S; ++i) {
#pragma unroll
        for (int j = 0; j < S; ++j) {
            double w = (i == j) ? h.unity[i] : 0.0;
            if (h.W[i][j]) w += (double)h.W[i][j][gi];
            if (h.mu[i][j]) w += h.Edip[i][j] * (double)h.mu[i][j][gi];
            hp[i].r += w * p[j].r;
            hp[i].i += w * p[j].i;
        }
    }
}

__device__ __forceinline__ int xcd_remap(int b, int nb) {
    // 8 XCDs receive blocks round robin; give each XCD a contiguous range of
    // tiles so that neighbouring tiles share one L2.
    if (nb % 8 != 0) return b;
    return (b % 8) * (nb / 8) + (b / 8);
}

// ---------------------------------------------------------------------------
// Chebyshev init: phi0 = psi, norm0 = sum |psi|^2, X = FFT_0(phi0) (nd >= 2)
template <class R, int S>
__global__ void k_cheby_init(GridDev g, LinePlan p0, int rb, const cx<double> *__restrict__ psi,
                             cx<R> *__restrict__ phi0, cx<R> *__restrict__ X, ChebyCtl *ctl,
                             double *partials) {
    extern __shared__ __align__(16) char smem_raw[];
    cx<R> *s = reinterpret_cast<cx<R> *>(smem_raw);
    __shared__ int flag;
    const int L = g.N[0];
    const long nrows = g.size / L;
    const long row0 = (long)blockIdx.x * rb;
    const int nl = S * rb;
    double acc = 0;
    for (int e = threadIdx.x; e < nl * L; e += blockDim.x) {
        const int line = e / L, x = e % L;
        const int st = line / rb, rr = line % rb;
        const long row = row0 + rr;
        if (row < nrows) {
            const long gi = (long)st * g.size + row * L + x;
            cx<double> v = psi[gi];
            acc += v.r * v.r + v.i * v.i;
            cx<R> vr = from_d<R>(v);
            phi0[gi] = vr;
            s[line * L + x] = vr;
        }
    }
    if (X) {
        __syncthreads();
        fft_lines<R>(s, p0, nl, true, -1);
        for (int e = threadIdx.x; e < nl * L; e += blockDim.x) {
            const int line = e / L, x = e % L;
            const int st = line / rb, rr = line % rb;
            const long row = row0 + rr;
            if (row < nrows) X[(long)st * g.size + row * L + x] = s[line * L + x];
        }
    }
    __shared__ double red[16];
    double part = block_sum(acc, red);
    double total = 0;
    if (last_block_sum(part, partials, &ctl->ticket, total, &flag) && threadIdx.x == 0) {
        ctl->norm0 = total;
        ctl->accnorm = total;
        ctl->done = 0;
        ctl->last_order = 1;
    }
}

// ---------------------------------------------------------------------------
// Chebyshev row kernel (see file header).  PRO: prologue (order 1).
template <class R, int S, bool ONE_D, bool PRO>
__global__ void k_cheby_row(GridDev g, HamDev<R> h, LinePlan p0, ChebyOrder ord, int rb,
                            const cx<R> *__restrict__ phim1, cx<R> *__restrict__ phim2,
                            cx<R> *__restrict__ X, cx<double> *__restrict__ psi, ChebyCtl *ctl,
                            double *partials) {
    if (ctl->done) return;
    extern __shared__ __align__(16) char smem_raw[];
    cx<R> *s = reinterpret_cast<cx<R> *>(smem_raw);
    __shared__ int flag;
    const int L = g.N[0];
    const long nrows = g.size / L;
    const long row0 = (long)blockIdx.x * rb;
    const int nl = S * rb;

    // 1. T phi_{k-1} into LDS
    const cx<R> *src = ONE_D ? phim1 : X;
    for (int e = threadIdx.x; e < nl * L; e += blockDim.x) {
        const int line = e / L, x = e % L;
        const int st = line / rb, rr = line % rb;
        const long row = row0 + rr;
        cx<R> v = mk<R>(0, 0);
        if (row < nrows) v = src[(long)st * g.size + row * L + x];
        s[line * L + x] = v;
    }
    __syncthreads();
    if (ONE_D) {
        fft_lines<R>(s, p0, nl, true, -1);
        for (int e = threadIdx.x; e < nl * L; e += blockDim.x) {
            const int line = e / L, x = e % L;
            const int st = line / rb;
            const double t = h.tk[st][0] ? (0.0 + h.tk[st][0][x]) * h.inv_n : 0.0;
            cx<R> v = s[line * L + x];
            s[line * L + x] = mk<R>((R)((double)v.r * t), (R)((double)v.i * t));
        }
        __syncthreads();
    }
    fft_lines<R>(s, p0, nl, true, +1);

    // 2. fused epilogue per grid point, all states
    double acc = 0;
    const cx<double> ck = mk<double>(ord.ckr, ord.cki);
    const cx<double> c0 = mk<double>(ord.c0r, ord.c0i);
    for (int e = threadIdx.x; e < rb * L; e += blockDim.x) {
        const int rr = e / L, x = e % L;
        const long row = row0 + rr;
        if (row >= nrows) continue;
        const long gi = row * L + x;
        cx<double> p[S], hp[S];
#pragma unroll
        for (int st = 0; st < S; ++st) {
            p[st] = to_d(phim1[(long)st * g.size + gi]);
            hp[st] = to_d(s[(st * rb + rr) * L + x]);
            if (!h.tk[st][0]) hp[st] = mk<double>(0, 0);
        }
        xspace_apply<R, S>(h, gi, p, hp);
#pragma unroll
        for (int st = 0; st < S; ++st) {
            const long si = (long)st * g.size + gi;
            // b = (H phi - offset phi) / R
            cx<double> b = mk<double>((hp[st].r - ord.offset * p[st].r) * ord.inv_scaling,
                                      (hp[st].i - ord.offset * p[st].i) * ord.inv_scaling);
            cx<double> pk;
            cx<double> ps = psi[si];
            if (PRO) {
                pk = mk<double>(b.i, -b.r);                       // * (-i)
                ps = cmul(c0, ps) + cmul(ck, pk);
            } else {
                const cx<double> pm2 = to_d(phim2[si]);
                pk = mk<double>(2.0 * b.i + pm2.r, -2.0 * b.r + pm2.i);   // * (-2i) + phi_{k-2}
                ps = ps + cmul(ck, pk);
                acc += ps.r * ps.r + ps.i * ps.i;
            }
            psi[si] = ps;
            const cx<R> pkr = from_d<R>(pk);
            phim2[si] = pkr;
            s[(st * rb + rr) * L + x] = pkr;
        }
    }
    // 3. X <- FFT_0(phi_k) for the next order's strided passes
    if (!ONE_D) {
        __syncthreads();
        fft_lines<R>(s, p0, nl, true, -1);
        for (int e = threadIdx.x; e < nl * L; e += blockDim.x) {
            const int line = e / L, x = e % L;
            const int st = line / rb, rr = line % rb;
            const long row = row0 + rr;
            if (row < nrows) X[(long)st * g.size + row * L + x] = s[line * L + x];
        }
    }
    if (PRO) return;
    // 4. convergence test on sum |psi|^2 (OCheby.cpp:179-182)
    __shared__ double red[16];
    double part = block_sum(acc, red);
    double total = 0;
    if (last_block_sum(part, partials, &ctl->ticket, total, &flag) && threadIdx.x == 0) {
        ctl->accnorm = total;
        ctl->last_order = ord.k;
        const double nconv = fabs(ctl->norm0 - total) / ctl->norm0;
        if (nconv < ord.prec) ctl->done = 1;
    }
}

// ---------------------------------------------------------------------------
// Strided line pass along dimension `dim` on X (in place).  A block owns a
// tile of W adjacent dimension-0 positions for one state and one index of the
// remaining dimension; lines are stored column-major in LDS so that the W
// lanes of one butterfly read adjacent words.
template <class R, int MODE>
__global__ void k_col(GridDev g, int S, int dim, int W, LinePlan p, HamDev<R> h, double tau,
                      cx<R> *__restrict__ X, const ChebyCtl *ctl) {
    if (ctl && ctl->done) return;
    extern __shared__ __align__(16) char smem_raw[];
    cx<R> *s = reinterpret_cast<cx<R> *>(smem_raw);
    const int N0 = g.N[0];
    const int L = g.N[dim];
    const long estride = dim == 1 ? N0 : (long)N0 * g.N[1];
    // remaining dimension (the one that is neither 0 nor dim)
    const int nother = g.nd == 3 ? (dim == 1 ? g.N[2] : g.N[1]) : 1;
    const long ostride = dim == 1 ? (long)N0 * g.N[1] : N0;
    const int tiles0 = N0 / W;
    const int nb = gridDim.x;
    const int b = xcd_remap(blockIdx.x, nb);
    const int c0 = (b % tiles0) * W;
    const int rest = b / tiles0;
    const int oth = rest % nother;
    const int st = rest / nother;
    cx<R> *base = X + (long)st * g.size + (long)oth * ostride + c0;

    for (int e = threadIdx.x; e < W * L; e += blockDim.x) {
        const int q = e / W, c = e % W;
        s[q * W + c] = base[q * estride + c];
    }
    __syncthreads();
    const int sign1 = MODE == 1 ? +1 : -1;
    fft_lines<R>(s, p, W, false, sign1);
    if (MODE >= 2) {
        for (int e = threadIdx.x; e < W * L; e += blockDim.x) {
            const int q = e / W, c = e % W;
            // k-space indices of this element
            long kidx;
            if (g.nd == 2) kidx = (long)q * N0 + c0 + c;
            else kidx = (dim == 2) ? ((long)q * g.N[1] + oth) * N0 + c0 + c
                                   : ((long)oth * g.N[1] + q) * N0 + c0 + c;
            const double tt = h.tk[st][0] ? kin_sum<R>(h, st, g, kidx) : 0.0;
            cx<R> v = s[q * W + c];
            if (MODE == 2) {
                const double t = tt * h.inv_n;
                s[q * W + c] = mk<R>((R)((double)v.r * t), (R)((double)v.i * t));
            } else {
                // exp(-i tau T) / N (OGSPO.cpp:37-41 folds 1/N into the exponential)
                double sn, cs;
                sincos(-tau * tt, &sn, &cs);
                cx<double> w = mk<double>(cs * h.inv_n, sn * h.inv_n);
                s[q * W + c] = from_d<R>(cmul(to_d(v), w));
            }
        }
        __syncthreads();
        fft_lines<R>(s, p, W, false, +1);
    }
    for (int e = threadIdx.x; e < W * L; e += blockDim.x) {
        const int q = e / W, c = e % W;
        base[q * estride + c] = s[q * W + c];
    }
}

// ---------------------------------------------------------------------------
// H applied to a complex128 vector: forward row pass, strided passes, row
// epilogue.  Reuses the Chebyshev row machinery with plain output.
template <class R, int S, bool ONE_D>
__global__ void k_apply_row(GridDev g, HamDev<R> h, LinePlan p0, int rb, const cx<double> *__restrict__ src,
                            cx<R> *__restrict__ X, cx<double> *__restrict__ dst) {
    extern __shared__ __align__(16) char smem_raw[];
    cx<R> *s = reinterpret_cast<cx<R> *>(smem_raw);
    const int L = g.N[0];
    const long nrows = g.size / L;
    const long row0 = (long)blockIdx.x * rb;
    const int nl = S * rb;
    for (int e = threadIdx.x; e < nl * L; e += blockDim.x) {
        const int line = e / L, x = e % L;
        const int st = line / rb, rr = line % rb;
        const long row = row0 + rr;
        cx<R> v = mk<R>(0, 0);
        if (row < nrows) v = ONE_D ? from_d<R>(src[(long)st * g.size + row * L + x]) : X[(long)st * g.size + row * L + x];
        s[line * L + x] = v;
    }
    __syncthreads();
    if (ONE_D) {
        fft_lines<R>(s, p0, nl, true, -1);
        for (int e = threadIdx.x; e < nl * L; e += blockDim.x) {
            const int line = e / L, x = e % L;
            const int st = line / rb;
            const double t = h.tk[st][0] ? (0.0 + h.tk[st][0][x]) * h.inv_n : 0.0;
            cx<R> v = s[line * L + x];
            s[line * L + x] = mk<R>((R)((double)v.r * t), (R)((double)v.i * t));
        }
        __syncthreads();
    }
    fft_lines<R>(s, p0, nl, true, +1);
    for (int e = threadIdx.x; e < rb * L; e += blockDim.x) {
        const int rr = e / L, x = e % L;
        const long row = row0 + rr;
        if (row >= nrows) continue;
        const long gi = row * L + x;
        cx<double> p[S], hp[S];
#pragma unroll
        for (int st = 0; st < S; ++st) {
            p[st] = src[(long)st * g.size + gi];
            hp[st] = h.tk[st][0] ? to_d(s[(st * rb + rr) * L + x]) : mk<double>(0, 0);
        }
        xspace_apply<R, S>(h, gi, p, hp);
#pragma unroll
        for (int st = 0; st < S; ++st) dst[(long)st * g.size + gi] = hp[st];
    }
}

template <class R, int S>
__global__ void k_row_fwd(GridDev g, LinePlan p0, int rb, const cx<double> *__restrict__ src,
                          cx<R> *__restrict__ X) {
    extern __shared__ __align__(16) char smem_raw[];
    cx<R> *s = reinterpret_cast<cx<R> *>(smem_raw);
    const int L = g.N[0];
    const long nrows = g.size / L;
    const long row0 = (long)blockIdx.x * rb;
    const int nl = S * rb;
    for (int e = threadIdx.x; e < nl * L; e += blockDim.x) {
        const int line = e / L, x = e % L;
        const int st = line / rb, rr = line % rb;
        const long row = row0 + rr;
        cx<R> v = mk<R>(0, 0);
        if (row < nrows) v = from_d<R>(src[(long)st * g.size + row * L + x]);
        s[line * L + x] = v;
    }
    __syncthreads();
    fft_lines<R>(s, p0, nl, true, -1);
    for (int e = threadIdx.x; e < nl * L; e += blockDim.x) {
        const int line = e / L, x = e % L;
        const int st = line / rb, rr = line % rb;
        const long row = row0 + rr;
        if (row < nrows) X[(long)st * g.size + row * L + x] = s[line * L + x];
    }
}

// ---------------------------------------------------------------------------
// GSPO x-space split operators on one grid point (all states):
//   psi <- exp(-i tau W(x)) psi, W real symmetric (ODSpace.h:65-69,
//   OGridDipole.cpp:83-87, ODMultistate.cpp:109-159 / TransformMSD.cpp:31-101)
template <class R, int S>
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
            double sn, cs;
            sincos(-op.tau * Wm[i][i], &sn, &cs);
            v[i] = cmul(v[i], mk<double>(cs, sn));
        }
        return;
    }
    if constexpr (S == 2) {
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
        // U = ph * [[c - i sc d, -i sc b], [-i sc b, c + i sc d]]
        const cx<double> u00 = mk<double>(c, -sc * d), u11 = mk<double>(c, sc * d);
        const cx<double> u01 = mk<double>(0, -sc * b);
        cx<double> a0 = cmul(u00, v[0]) + cmul(u01, v[1]);
        cx<double> a1 = cmul(u01, v[0]) + cmul(u11, v[1]);
        v[0] = cmul(ph, a0);
        v[1] = cmul(ph, a1);
    }
}

template <class R, int S>
__device__ __forceinline__ void xgroup_point(const XGroupDev<R> &grp, long gi, cx<double> (&v)[S]) {
    for (int o = 0; o < grp.n; ++o) xsplit_point<R, S>(grp.op[o], gi, v);
}

// pre-ops then FFT_0 into X (nd >= 2)
template <class R, int S>
__global__ void k_gspo_row_pre(GridDev g, LinePlan p0, int rb, XGroupDev<R> pre,
                               const cx<R> *__restrict__ psi, cx<R> *__restrict__ X) {
    extern __shared__ __align__(16) char smem_raw[];
    cx<R> *s = reinterpret_cast<cx<R> *>(smem_raw);
    const int L = g.N[0];
    const long nrows = g.size / L;
    const long row0 = (long)blockIdx.x * rb;
    const int nl = S * rb;
    for (int e = threadIdx.x; e < rb * L; e += blockDim.x) {
        const int rr = e / L, x = e % L;
        const long row = row0 + rr;
        cx<double> v[S];
        if (row < nrows) {
            const long gi = row * L + x;
#pragma unroll
            for (int st = 0; st < S; ++st) v[st] = to_d(psi[(long)st * g.size + gi]);
            xgroup_point<R, S>(pre, gi, v);
        } else {
#pragma unroll
            for (int st = 0; st < S; ++st) v[st] = mk<double>(0, 0);
        }
#pragma unroll
        for (int st = 0; st < S; ++st) s[(st * rb + rr) * L + x] = from_d<R>(v[st]);
    }
    __syncthreads();
    fft_lines<R>(s, p0, nl, true, -1);
    for (int e = threadIdx.x; e < nl * L; e += blockDim.x) {
        const int line = e / L, x = e % L;
        const int st = line / rb, rr = line % rb;
        const long row = row0 + rr;
        if (row < nrows) X[(long)st * g.size + row * L + x] = s[line * L + x];
    }
}

// IFFT_0 of X then post-ops into psi (nd >= 2)
template <class R, int S>
__global__ void k_gspo_row_post(GridDev g, LinePlan p0, int rb, XGroupDev<R> post,
                                const cx<R> *__restrict__ X, cx<R> *__restrict__ psi) {
    extern __shared__ __align__(16) char smem_raw[];
    cx<R> *s = reinterpret_cast<cx<R> *>(smem_raw);
    const int L = g.N[0];
    const long nrows = g.size / L;
    const long row0 = (long)blockIdx.x * rb;
    const int nl = S * rb;
    for (int e = threadIdx.x; e < nl * L; e += blockDim.x) {
        const int line = e / L, x = e % L;
        const int st = line / rb, rr = line % rb;
        const long row = row0 + rr;
        cx<R> v = mk<R>(0, 0);
        if (row < nrows) v = X[(long)st * g.size + row * L + x];
        s[line * L + x] = v;
    }
    __syncthreads();
    fft_lines<R>(s, p0, nl, true, +1);
    for (int e = threadIdx.x; e < rb * L; e += blockDim.x) {
        const int rr = e / L, x = e % L;
        const long row = row0 + rr;
        if (row >= nrows) continue;
        const long gi = row * L + x;
        cx<double> v[S];
#pragma unroll
        for (int st = 0; st < S; ++st) v[st] = to_d(s[(st * rb + rr) * L + x]);
        xgroup_point<R, S>(post, gi, v);
#pragma unroll
        for (int st = 0; st < S; ++st) psi[(long)st * g.size + gi] = from_d<R>(v[st]);
    }
}

// whole GSPO step for nd == 1 (pre, FFT, exp(-i tau T)/N, IFFT, post)
template <class R, int S>
__global__ void k_gspo_1d(GridDev g, LinePlan p0, HamDev<R> kin, double tau, XGroupDev<R> pre,
                          XGroupDev<R> post, cx<R> *__restrict__ psi) {
    extern __shared__ __align__(16) char smem_raw[];
    cx<R> *s = reinterpret_cast<cx<R> *>(smem_raw);
    const int L = g.N[0];
    for (int x = threadIdx.x; x < L; x += blockDim.x) {
        cx<double> v[S];
#pragma unroll
        for (int st = 0; st < S; ++st) v[st] = to_d(psi[(long)st * g.size + x]);
        xgroup_point<R, S>(pre, x, v);
#pragma unroll
        for (int st = 0; st < S; ++st) s[st * L + x] = from_d<R>(v[st]);
    }
    __syncthreads();
    fft_lines<R>(s, p0, S, true, -1);
    for (int e = threadIdx.x; e < S * L; e += blockDim.x) {
        const int st = e / L, x = e % L;
        if (!kin.tk[st][0]) {
            // no kinetic term on this state: exp(0)/N keeps the round trip exact
            cx<R> v = s[e];
            s[e] = from_d<R>(mk<double>((double)v.r * kin.inv_n, (double)v.i * kin.inv_n));
            continue;
        }
        const double tt = 0.0 + kin.tk[st][0][x];
        double sn, cs;
        sincos(-tau * tt, &sn, &cs);
        s[e] = from_d<R>(cmul(to_d(s[e]), mk<double>(cs * kin.inv_n, sn * kin.inv_n)));
    }
    __syncthreads();
    fft_lines<R>(s, p0, S, true, +1);
    for (int x = threadIdx.x; x < L; x += blockDim.x) {
        cx<double> v[S];
#pragma unroll
        for (int st = 0; st < S; ++st) v[st] = to_d(s[st * L + x]);
        xgroup_point<R, S>(post, x, v);
#pragma unroll
        for (int st = 0; st < S; ++st) psi[(long)st * g.size + x] = from_d<R>(v[st]);
    }
}

template <class R, int S>
__global__ void k_gspo_xonly(GridDev g, XGroupDev<R> grp, cx<R> *__restrict__ psi) {
    const long gi = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= g.size) return;
    cx<double> v[S];
#pragma unroll
    for (int st = 0; st < S; ++st) v[st] = to_d(psi[(long)st * g.size + gi]);
    xgroup_point<R, S>(grp, gi, v);
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
__global__ void k_c64_to_c128(const float *__restrict__ s, double *__restrict__ d, long n2) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (long)gridDim.x * blockDim.x)
        d[i] = (double)s[i];
}

// ---------------------------------------------------------------------------
// launch helpers

static int round64(int n) {
    const int r = ((n + 63) / 64) * 64;
    if (r > 1024) throw std::runtime_error("line block needs more than 1024 threads");
    return r;
}

// rows per block so that a row block has >= 256 threads of butterflies, but
// the grid keeps at least ~1024 blocks where the problem allows it
static int rows_per_block(const GridDev &g, int S, int L) {
    const long nrows = g.size / L;
    const int TL = (L + 7) / 8;
    int rb = 1;
    while (rb * 2 <= nrows && S * rb * 2 * TL <= 256 && nrows / (rb * 2) >= 1024) rb *= 2;
    return rb;
}

template <class K>
static void set_lds(K kernel, size_t bytes) {
    if (bytes > 160 * 1024) throw std::runtime_error("line block exceeds the 160 KiB LDS of a gfx950 CU");
    if (bytes > 65536)
        QDX_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

static int col_tile(const GridDev &g, int L, size_t elem) {
    // widest tile (<= 16 lanes, <= 64 KiB LDS) that still leaves >= 2048 blocks
    int W = 16;
    while (W > 1 && ((size_t)W * L * elem > 65536 || g.N[0] % W != 0)) W /= 2;
    while (W > 1 && (g.size / L) / W < 2048) W /= 2;
    if (g.N[0] % W != 0) W = 1;
    return W;
}

template <class R, int S>
static void cheby_init_S(hipStream_t st, const GridDev &g, const LinePlan &p0, const double *psi, R *phi0,
                         R *X, ChebyCtl *ctl, double *partials) {
    const int L = g.N[0];
    const int rb = rows_per_block(g, S, L);
    const long nrows = g.size / L;
    const int nb = (int)((nrows + rb - 1) / rb);
    if (nb > kMaxPartials) throw std::runtime_error("grid too large for the partial-sum buffer");
    const int threads = std::max(64, round64(S * rb * ((L + 7) / 8)));
    const size_t lds = (size_t)S * rb * L * sizeof(cx<R>);
    set_lds(k_cheby_init<R, S>, lds);
    hipLaunchKernelGGL((k_cheby_init<R, S>), dim3(nb), dim3(threads), lds, st, g, p0, rb,
                       reinterpret_cast<const cx<double> *>(psi), reinterpret_cast<cx<R> *>(phi0),
                       reinterpret_cast<cx<R> *>(X), ctl, partials);
    QDX_HIP(hipGetLastError());
}

template <class R>
void Launch<R>::cheby_init(hipStream_t s, const GridDev &g, int S, const LinePlan &p0, const double *psi,
                           R *phi0, R *X, ChebyCtl *ctl, double *partials) {
    switch (S) {
    case 1: cheby_init_S<R, 1>(s, g, p0, psi, phi0, X, ctl, partials); break;
    case 2: cheby_init_S<R, 2>(s, g, p0, psi, phi0, X, ctl, partials); break;
    case 3: cheby_init_S<R, 3>(s, g, p0, psi, phi0, X, ctl, partials); break;
    case 4: cheby_init_S<R, 4>(s, g, p0, psi, phi0, X, ctl, partials); break;
    default: throw std::runtime_error("unsupported number of states");
    }
}

template <class R, int S, bool ONE_D, bool PRO>
static void cheby_row_SK(hipStream_t st, const GridDev &g, const HamDev<R> &h, const LinePlan &p0,
                         const ChebyOrder &ord, const R *phim1, R *phim2, R *X, double *psi, ChebyCtl *ctl,
                         double *partials) {
    const int L = g.N[0];
    const int rb = rows_per_block(g, S, L);
    const long nrows = g.size / L;
    const int nb = (int)((nrows + rb - 1) / rb);
    if (nb > kMaxPartials) throw std::runtime_error("grid too large for the partial-sum buffer");
    const int threads = std::max(64, round64(S * rb * ((L + 7) / 8)));
    const size_t lds = (size_t)S * rb * L * sizeof(cx<R>);
    set_lds(k_cheby_row<R, S, ONE_D, PRO>, lds);
    hipLaunchKernelGGL((k_cheby_row<R, S, ONE_D, PRO>), dim3(nb), dim3(threads), lds, st, g, h, p0, ord, rb,
                       reinterpret_cast<const cx<R> *>(phim1), reinterpret_cast<cx<R> *>(phim2),
                       reinterpret_cast<cx<R> *>(X), reinterpret_cast<cx<double> *>(psi), ctl, partials);
    QDX_HIP(hipGetLastError());
}

template <class R, int S>
static void cheby_row_S(hipStream_t st, const GridDev &g, const HamDev<R> &h, const LinePlan &p0,
                        const ChebyOrder &ord, const R *phim1, R *phim2, R *X, double *psi, ChebyCtl *ctl,
                        double *partials) {
    const bool one = g.nd == 1;
    const bool pro = ord.k == 1;
    if (one && pro) cheby_row_SK<R, S, true, true>(st, g, h, p0, ord, phim1, phim2, X, psi, ctl, partials);
    else if (one) cheby_row_SK<R, S, true, false>(st, g, h, p0, ord, phim1, phim2, X, psi, ctl, partials);
    else if (pro) cheby_row_SK<R, S, false, true>(st, g, h, p0, ord, phim1, phim2, X, psi, ctl, partials);
    else cheby_row_SK<R, S, false, false>(st, g, h, p0, ord, phim1, phim2, X, psi, ctl, partials);
}

template <class R>
void Launch<R>::cheby_row(hipStream_t s, const GridDev &g, const HamDev<R> &h, const LinePlan &p0,
                          const ChebyOrder &ord, const R *phim1, R *phim2, R *X, double *psi, ChebyCtl *ctl,
                          double *partials) {
    switch (h.S) {
    case 1: cheby_row_S<R, 1>(s, g, h, p0, ord, phim1, phim2, X, psi, ctl, partials); break;
    case 2: cheby_row_S<R, 2>(s, g, h, p0, ord, phim1, phim2, X, psi, ctl, partials); break;
    case 3: cheby_row_S<R, 3>(s, g, h, p0, ord, phim1, phim2, X, psi, ctl, partials); break;
    case 4: cheby_row_S<R, 4>(s, g, h, p0, ord, phim1, phim2, X, psi, ctl, partials); break;
    default: throw std::runtime_error("unsupported number of states");
    }
}

template <class R>
void Launch<R>::col_pass(hipStream_t s, const GridDev &g, int S, int dim, int mode, const LinePlan &p,
                         const HamDev<R> &h, double tau, R *X, const ChebyCtl *ctl) {
    const int L = g.N[dim];
    const int W = col_tile(g, L, sizeof(cx<R>));
    const int nother = g.nd == 3 ? (dim == 1 ? g.N[2] : g.N[1]) : 1;
    const int nb = (g.N[0] / W) * nother * S;
    const int threads = std::max(64, round64(W * ((L + 7) / 8)));
    const size_t lds = (size_t)W * L * sizeof(cx<R>);
    cx<R> *x = reinterpret_cast<cx<R> *>(X);
    switch (mode) {
    case 0:
        set_lds(k_col<R, 0>, lds);
        hipLaunchKernelGGL((k_col<R, 0>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, p, h, tau, x, ctl);
        break;
    case 1:
        set_lds(k_col<R, 1>, lds);
        hipLaunchKernelGGL((k_col<R, 1>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, p, h, tau, x, ctl);
        break;
    case 2:
        set_lds(k_col<R, 2>, lds);
        hipLaunchKernelGGL((k_col<R, 2>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, p, h, tau, x, ctl);
        break;
    default:
        set_lds(k_col<R, 3>, lds);
        hipLaunchKernelGGL((k_col<R, 3>), dim3(nb), dim3(threads), lds, s, g, S, dim, W, p, h, tau, x, ctl);
        break;
    }
    QDX_HIP(hipGetLastError());
}

template <class R, int S>
static void ham_apply_S(hipStream_t st, const GridDev &g, const HamDev<R> &h, const LinePlan *plans,
                        const double *src, double *dst, R *X) {
    const int L = g.N[0];
    const int rb = rows_per_block(g, S, L);
    const long nrows = g.size / L;
    const int nb = (int)((nrows + rb - 1) / rb);
    const int threads = std::max(64, round64(S * rb * ((L + 7) / 8)));
    const size_t lds = (size_t)S * rb * L * sizeof(cx<R>);
    const cx<double> *s = reinterpret_cast<const cx<double> *>(src);
    cx<double> *d = reinterpret_cast<cx<double> *>(dst);
    if (g.nd == 1) {
        set_lds(k_apply_row<R, S, true>, lds);
        hipLaunchKernelGGL((k_apply_row<R, S, true>), dim3(nb), dim3(threads), lds, st, g, h, plans[0], rb, s,
                           reinterpret_cast<cx<R> *>(X), d);
        QDX_HIP(hipGetLastError());
        return;
    }
    set_lds(k_row_fwd<R, S>, lds);
    hipLaunchKernelGGL((k_row_fwd<R, S>), dim3(nb), dim3(threads), lds, st, g, plans[0], rb, s,
                       reinterpret_cast<cx<R> *>(X));
    QDX_HIP(hipGetLastError());
    if (g.nd == 2) {
        Launch<R>::col_pass(st, g, S, 1, 2, plans[1], h, 0.0, X, nullptr);
    } else {
        Launch<R>::col_pass(st, g, S, 1, 0, plans[1], h, 0.0, X, nullptr);
        Launch<R>::col_pass(st, g, S, 2, 2, plans[2], h, 0.0, X, nullptr);
        Launch<R>::col_pass(st, g, S, 1, 1, plans[1], h, 0.0, X, nullptr);
    }
    set_lds(k_apply_row<R, S, false>, lds);
    hipLaunchKernelGGL((k_apply_row<R, S, false>), dim3(nb), dim3(threads), lds, st, g, h, plans[0], rb, s,
                       reinterpret_cast<cx<R> *>(X), d);
    QDX_HIP(hipGetLastError());
}

template <class R>
void Launch<R>::ham_apply(hipStream_t s, const GridDev &g, const HamDev<R> &h, const LinePlan *plans,
                          const double *src, double *dst, R *X) {
    switch (h.S) {
    case 1: ham_apply_S<R, 1>(s, g, h, plans, src, dst, X); break;
    case 2: ham_apply_S<R, 2>(s, g, h, plans, src, dst, X); break;
    case 3: ham_apply_S<R, 3>(s, g, h, plans, src, dst, X); break;
    case 4: ham_apply_S<R, 4>(s, g, h, plans, src, dst, X); break;
    default: throw std::runtime_error("unsupported number of states");
    }
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
void Launch<R>::gspo_row_pre(hipStream_t s, const GridDev &g, int S, const LinePlan &p0, const XGroupDev<R> &pre,
                             const R *psi, R *X) {
    const int L = g.N[0];
    QDX_DISPATCH_S(S, {
        const int rb = rows_per_block(g, SS, L);
        const int nb = (int)((g.size / L + rb - 1) / rb);
        const int threads = std::max(64, round64(SS * rb * ((L + 7) / 8)));
        const size_t lds = (size_t)SS * rb * L * sizeof(cx<R>);
        set_lds(k_gspo_row_pre<R, SS>, lds);
        hipLaunchKernelGGL((k_gspo_row_pre<R, SS>), dim3(nb), dim3(threads), lds, s, g, p0, rb, pre,
                           reinterpret_cast<const cx<R> *>(psi), reinterpret_cast<cx<R> *>(X));
    });
    QDX_HIP(hipGetLastError());
}

template <class R>
void Launch<R>::gspo_row_post(hipStream_t s, const GridDev &g, int S, const LinePlan &p0, const XGroupDev<R> &post,
                              const R *X, R *psi) {
    const int L = g.N[0];
    QDX_DISPATCH_S(S, {
        const int rb = rows_per_block(g, SS, L);
        const int nb = (int)((g.size / L + rb - 1) / rb);
        const int threads = std::max(64, round64(SS * rb * ((L + 7) / 8)));
        const size_t lds = (size_t)SS * rb * L * sizeof(cx<R>);
        set_lds(k_gspo_row_post<R, SS>, lds);
        hipLaunchKernelGGL((k_gspo_row_post<R, SS>), dim3(nb), dim3(threads), lds, s, g, p0, rb, post,
                           reinterpret_cast<const cx<R> *>(X), reinterpret_cast<cx<R> *>(psi));
    });
    QDX_HIP(hipGetLastError());
}

template <class R>
void Launch<R>::gspo_1d(hipStream_t s, const GridDev &g, int S, const LinePlan &p0, const HamDev<R> &kin,
                        double tau, const XGroupDev<R> &pre, const XGroupDev<R> &post, R *psi) {
    const int L = g.N[0];
    QDX_DISPATCH_S(S, {
        const int threads = std::max(64, round64(SS * ((L + 7) / 8)));
        const size_t lds = (size_t)SS * L * sizeof(cx<R>);
        set_lds(k_gspo_1d<R, SS>, lds);
        hipLaunchKernelGGL((k_gspo_1d<R, SS>), dim3(1), dim3(threads), lds, s, g, p0, kin, tau, pre, post,
                           reinterpret_cast<cx<R> *>(psi));
    });
    QDX_HIP(hipGetLastError());
}

template <class R>
void Launch<R>::gspo_xonly(hipStream_t s, const GridDev &g, int S, const XGroupDev<R> &grp, R *psi) {
    const int threads = 256;
    const int nb = (int)((g.size + threads - 1) / threads);
    QDX_DISPATCH_S(S, {
        hipLaunchKernelGGL((k_gspo_xonly<R, SS>), dim3(nb), dim3(threads), 0, s, g, grp,
                           reinterpret_cast<cx<R> *>(psi));
    });
    QDX_HIP(hipGetLastError());
}

template struct Launch<double>;
template struct Launch<float>;

void launch_convert_c128_to_c64(hipStream_t s, const double *src, float *dst, long n) {
    const long n2 = 2 * n;
    const int nb = (int)std::min<long>((n2 + 255) / 256, 8192);
    hipLaunchKernelGGL(k_c128_to_c64, dim3(nb), dim3(256), 0, s, src, dst, n2);
    QDX_HIP(hipGetLastError());
}

void launch_convert_c64_to_c128(hipStream_t s, const float *src, double *dst, long n) {
    const long n2 = 2 * n;
    const int nb = (int)std::min<long>((n2 + 255) / 256, 8192);
    hipLaunchKernelGGL(k_c64_to_c128, dim3(nb), dim3(256), 0, s, src, dst, n2);
    QDX_HIP(hipGetLastError());
}

template <class R>
static void overlap_impl(hipStream_t s, const R *a, const R *b, long n, int S, double *partials, double *out) {
    const int threads = 256;
    int nbx = (int)std::min<long>((n + threads - 1) / threads, 1024);
    if (nbx < 1) nbx = 1;
    hipLaunchKernelGGL((k_overlap_stage1<R>), dim3(nbx, S), dim3(threads), 0, s,
                       reinterpret_cast<const cx<R> *>(a), reinterpret_cast<const cx<R> *>(b), n, partials);
    QDX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_overlap_stage2, dim3(1), dim3(256), 0, s, partials, nbx, S, out);
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
