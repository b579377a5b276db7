// gfx950 GSPO kernels of the single propagation (OGSPO.cpp:114-139): the
// x-space split operators as per-point S x S unitaries fused with the
// dimension-0 transforms.
#include "qdx_kernels_common.hpp"

namespace qdx {

// ---------------------------------------------------------------------------
// GSPO x-space split operators on one grid point (all states):
//   psi <- exp(-i tau W(x)) psi, W real symmetric (ODSpace.h:65-69,
//   OGridDipole.cpp:83-87, ODMultistate.cpp:109-159 / TransformMSD.cpp:31-101)
template <class R, int S, bool IMAG = false>
__device__ __forceinline__ void xsplit_point(const XSplitDev<R> &op, long gi, cx<double> (&v)[S]) {
    if (!IMAG && op.U) {   // the table of a time-independent operator
        auto u = [&](int c) {
            const double *q = op.U + 2 * ((long)c * op.n + gi);
            return mk<double>(q[0], q[1]);
        };
        if (S == 2 && op.coupled) {
            const cx<double> u00 = u(0), u01 = u(1), u11 = u(2);
            const cx<double> a0 = cmul(u00, v[0]) + cmul(u01, v[1]);
            const cx<double> a1 = cmul(u01, v[0]) + cmul(u11, v[1]);
            v[0] = a0;
            v[1] = a1;
        } else {
#pragma unroll
            for (int i = 0; i < S; ++i) v[i] = cmul(v[i], u(i));
        }
        return;
    }
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
        double so, c;
        sincos(op.tau * O, &so, &c);
        const double sc = O > 0 ? so / O : op.tau;
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
// FFT_0 -> X; PHASE 1: X -> IFFT_0 -> post -> psi; PHASE 2 (between two
// steps of a run): X -> IFFT_0 -> post of step n -> pre of step n + 1 ->
// FFT_0 -> X, so that psi stays in X between the steps.  With opart, PHASE 2
// also stores this block's sums of the state between the steps, per state s:
// opart[2 (s nb + b)] = sum |psi|^2 and, after 2 S nb entries, the
// (re, im) of sum conj(psi) psi0 (Reporter observables, Reporter.cpp:101-137).
// The x-space operators act on all states of a point at once, so the S lines
// of a row are exchanged through LDS around them.
template <class R, int S, bool ONE_D, int PHASE, bool IMAG = false, int P = 8>
__global__ void __launch_bounds__(512) k_gspo_row(GridDev g, HamDev<R> kin, double tau, LinePlan p0, int rb,
                                                   XGroupDev<R> pre, XGroupDev<R> post, cx<R> *__restrict__ psi,
                                                   cx<R> *__restrict__ X, const cx<R> *__restrict__ psi0,
                                                   double *__restrict__ opart) {
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
        // the points of row rr are shared out among the S lines' threads:
        // point lt + q TL goes to thread lt of line q mod S, so each thread
        // takes P / S of them (x mod S would give a thread all of its P
        // points or none when TL is a multiple of S)
        if (B.active) {
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const int x = B.lt + q * B.TL;
                if (q % S != st) continue;
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
    if (PHASE == 2) {
        // step n's post operators, the observables of the state between the
        // steps (rounded to R, as a stored psi) and step n + 1's pre
        // operators on each point in one pass: for S > 1 the S lines of a
        // row are exchanged through LDS once around all three
        double n2[S], ar[S], ai[S];
#pragma unroll
        for (int s2 = 0; s2 < S; ++s2) n2[s2] = ar[s2] = ai[s2] = 0;
        auto point = [&](long gi, cx<double> (&w)[S]) {
            xgroup_point<R, S, IMAG>(post, gi, w);
#pragma unroll
            for (int s2 = 0; s2 < S; ++s2) {
                w[s2] = to_d(from_d<R>(w[s2]));
                if (opart) {
                    const cx<double> a = w[s2], b = to_d(psi0[(long)s2 * g.size + gi]);
                    n2[s2] += a.r * a.r + a.i * a.i;
                    ar[s2] += a.r * b.r + a.i * b.i;
                    ai[s2] += a.r * b.i - a.i * b.r;
                }
            }
            xgroup_point<R, S, IMAG>(pre, gi, w);
        };
        if (S == 1) {
            if (B.active) {
#pragma unroll
                for (int q = 0; q < P; ++q) {
                    cx<double> w[S];
                    w[0] = to_d(v[q]);
                    point(B.gidx(q), w);
                    v[q] = from_d<R>(w[0]);
                }
            }
        } else {
            __syncthreads();
            if (B.active) {
#pragma unroll
                for (int q = 0; q < P; ++q) xa[((long)st * B.rb + B.rr) * B.L + B.lt + q * B.TL] = v[q];
            }
            __syncthreads();
            if (B.active) {
#pragma unroll
                for (int q = 0; q < P; ++q) {
                    const int x = B.lt + q * B.TL;
                    if (q % S != st) continue;
                    cx<double> w[S];
#pragma unroll
                    for (int s2 = 0; s2 < S; ++s2) w[s2] = to_d(xa[((long)s2 * B.rb + B.rr) * B.L + x]);
                    point(B.row * B.L + x, w);
#pragma unroll
                    for (int s2 = 0; s2 < S; ++s2) xa[((long)s2 * B.rb + B.rr) * B.L + x] = from_d<R>(w[s2]);
                }
            }
            __syncthreads();
            if (B.active) {
#pragma unroll
                for (int q = 0; q < P; ++q) v[q] = xa[((long)st * B.rb + B.rr) * B.L + B.lt + q * B.TL];
            }
        }
        if (opart) {
            // the 3 S block sums in one reduction
            __shared__ double red[16 * 3 * S];
            double t[3 * S];
#pragma unroll
            for (int s2 = 0; s2 < S; ++s2) {
                t[3 * s2] = n2[s2];
                t[3 * s2 + 1] = ar[s2];
                t[3 * s2 + 2] = ai[s2];
            }
            block_sum_n<3 * S>(t, red);
            if (threadIdx.x == 0) {
                const long nb = gridDim.x;
#pragma unroll
                for (int s2 = 0; s2 < S; ++s2) {
                    opart[2 * (s2 * nb + blockIdx.x)] = t[3 * s2];
                    opart[2 * (s2 * nb + blockIdx.x) + 1] = 0.0;
                    opart[2 * S * nb + 2 * (s2 * nb + blockIdx.x)] = t[3 * s2 + 1];
                    opart[2 * S * nb + 2 * (s2 * nb + blockIdx.x) + 1] = t[3 * s2 + 2];
                }
            }
        }
        B.fft(v, p0, -1);
        if (B.active) {
#pragma unroll
            for (int q = 0; q < P; ++q) X[base + B.gidx(q)] = v[q];
        }
        return;
    }
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

// The table XSplitDev::U: the per-point arithmetic above applied to unit
// vectors (exact zeros and ones, so each entry is that arithmetic's product
// rounded once): coupled two states U00, U01 from e_0 and U11 from e_1,
// otherwise the S diagonal phases from (1, ..., 1).
template <class R, int S>
__global__ void k_xsplit_unitary(GridDev g, XSplitDev<R> op, double *__restrict__ U) {
    const long gi = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= g.size) return;
    op.U = nullptr;
    auto put = [&](int c, cx<double> z) {
        U[2 * ((long)c * g.size + gi)] = z.r;
        U[2 * ((long)c * g.size + gi) + 1] = z.i;
    };
    cx<double> v[S];
    if (S == 2 && op.coupled) {
        v[0] = mk<double>(1, 0);
        v[1] = mk<double>(0, 0);
        xsplit_point<R, S>(op, gi, v);
        put(0, v[0]);
        put(1, v[1]);
        v[0] = mk<double>(0, 0);
        v[1] = mk<double>(1, 0);
        xsplit_point<R, S>(op, gi, v);
        put(2, v[1]);
        return;
    }
#pragma unroll
    for (int i = 0; i < S; ++i) v[i] = mk<double>(1, 0);
    xsplit_point<R, S>(op, gi, v);
#pragma unroll
    for (int i = 0; i < S; ++i) put(i, v[i]);
}

template <class R>
void Launch<R>::gspo_unitary(hipStream_t s, const GridDev &g, int S, const XSplitDev<R> &op, double *U) {
    const int threads = 256;
    const int nb = (int)((g.size + threads - 1) / threads);
    QDX_DISPATCH_S(S, {
        QDX_LAUNCH((k_xsplit_unitary<R, SS>), dim3(nb), dim3(threads), 0, s, g, op, U);
    });
    QDX_HIP(hipGetLastError());
}

// returns the number of row blocks launched
template <class R, bool ONE_D, int PHASE, bool IMAG, int P>
static int gspo_row_launch_IP(hipStream_t s, const GridDev &g, int S, const HamDev<R> &kin, double tau,
                              const LinePlan &plan, const XGroupDev<R> &pre, const XGroupDev<R> &post, R *psi,
                              R *X, const R *psi0, double *opart) {
    int nb = 0;
    QDX_DISPATCH_S(S, {
        LinePlan p0 = plan;
        RowLaunch r = row_launch<R>(g, SS, p0, SS > 1);
        if (opart && (long)4 * SS * r.nb > kMaxPartials)
            throw std::runtime_error("grid too large for the observables' partial sums");
        set_lds(k_gspo_row<R, SS, ONE_D, PHASE, IMAG, P>, r.lds);
        QDX_LAUNCH((k_gspo_row<R, SS, ONE_D, PHASE, IMAG, P>), dim3(r.nb), dim3(r.threads), r.lds, s, g, kin, tau,
                   p0, r.rb, pre, post, reinterpret_cast<cx<R> *>(psi), reinterpret_cast<cx<R> *>(X),
                   reinterpret_cast<const cx<R> *>(psi0), opart);
        nb = r.nb;
    });
    QDX_HIP(hipGetLastError());
    return nb;
}

// 4 points per thread on grids with nd >= 2 when the plan asks for it
template <class R, bool ONE_D, int PHASE, bool IMAG>
static int gspo_row_launch_I(hipStream_t s, const GridDev &g, int S, const HamDev<R> &kin, double tau,
                             const LinePlan &plan, const XGroupDev<R> &pre, const XGroupDev<R> &post, R *psi,
                             R *X, const R *psi0, double *opart) {
    if constexpr (!ONE_D) {
        if (plan.ppt == 4)
            return gspo_row_launch_IP<R, ONE_D, PHASE, IMAG, 4>(s, g, S, kin, tau, plan, pre, post, psi, X, psi0,
                                                                 opart);
    }
    if (plan.ppt != 8) throw std::runtime_error("GSPO rows: 8 or 4 points per thread (4 for nd >= 2)");
    return gspo_row_launch_IP<R, ONE_D, PHASE, IMAG, 8>(s, g, S, kin, tau, plan, pre, post, psi, X, psi0, opart);
}

// imag: imaginary time (FP64 only)
template <class R, bool ONE_D, int PHASE>
static int gspo_row_launch(hipStream_t s, const GridDev &g, int S, const HamDev<R> &kin, double tau,
                           const LinePlan &plan, const XGroupDev<R> &pre, const XGroupDev<R> &post, R *psi, R *X,
                           int imag = 0, const R *psi0 = nullptr, double *opart = nullptr) {
    if (!imag) return gspo_row_launch_I<R, ONE_D, PHASE, false>(s, g, S, kin, tau, plan, pre, post, psi, X, psi0, opart);
    if constexpr (std::is_same<R, double>::value)
        return gspo_row_launch_I<R, ONE_D, PHASE, true>(s, g, S, kin, tau, plan, pre, post, psi, X, psi0, opart);
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
int Launch<R>::gspo_row_bridge(hipStream_t s, const GridDev &g, int S, const LinePlan &p0, const XGroupDev<R> &post,
                               const XGroupDev<R> &pre, R *X, const R *psi0, double *opart, int imag) {
    HamDev<R> none;
    return gspo_row_launch<R, false, 2>(s, g, S, none, 0.0, p0, pre, post, nullptr, X, imag, psi0, opart);
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

#define QDX_INST_GSPO(R)                                                                                       \
    template void Launch<R>::gspo_row_pre(hipStream_t, const GridDev &, int, const LinePlan &, const XGroupDev<R> &, \
                                          const R *, R *, int);                                                     \
    template void Launch<R>::gspo_row_post(hipStream_t, const GridDev &, int, const LinePlan &,                    \
                                           const XGroupDev<R> &, const R *, R *, int);                              \
    template void Launch<R>::gspo_1d(hipStream_t, const GridDev &, int, const LinePlan &, const HamDev<R> &,       \
                                     double, const XGroupDev<R> &, const XGroupDev<R> &, R *, int);                 \
    template void Launch<R>::gspo_xonly(hipStream_t, const GridDev &, int, const XGroupDev<R> &, R *, int);       \
    template void Launch<R>::gspo_unitary(hipStream_t, const GridDev &, int, const XSplitDev<R> &, double *);     \
    template int Launch<R>::gspo_row_bridge(hipStream_t, const GridDev &, int, const LinePlan &,                   \
                                            const XGroupDev<R> &, const XGroupDev<R> &, R *, const R *, double *,  \
                                            int);
QDX_INST_GSPO(double)
QDX_INST_GSPO(float)

}  // namespace qdx
