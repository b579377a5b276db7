// gfx950 kernels of the batched-jobs runtime.
//
// A tick advances every active job by one Chebyshev time step
// (OCheby::real_time_step, OCheby.cpp:92-197).  Each pass is one launch over
// the tiles of all jobs: a block reads its tile, then its job's descriptor,
// and runs the same register-resident line transforms and fused epilogue as
// the single-propagation kernels, on that job's arrays.  Jobs keep their own
// series length and break test: blocks of a job that has converged, or whose
// series is shorter than the order of the launch, return at once.
//
// The kinetic operator is separable, T(k) = T_0(k_0) + T_1(k_1), so T phi is
// the sum of 1D filters IFFT_d(T_d/N_d FFT_d phi) along each dimension.
// Row tiles: rb lines of dimension 0 of one job; the dimension-0 filter of
// phi_{k-1} runs in registers/LDS, the dimension-1 part Y comes from the
// column pass.  Column tiles (2D jobs): W adjacent columns of phi_{k-1},
// filtered along dimension 1 into Y (held in X).
#include <hip/hip_runtime.h>

#include "qdx_batch.hpp"
#include "qdx_device.hpp"

namespace qdx {

// geometry of one row tile; LB LDS buffers per line (2: ping-pong, 1: the
// FFT's middle stages in place)
template <class R, int LB>
struct BRowTile {
    int L, TL, line, lt;
    long row;
    bool active;
    LdsLine<R> ln;
    cx<R> *tw;
    __device__ BRowTile(const BJob<R> &J, int row0, cx<R> *smem) {
        L = J.N0;
        TL = L >> 3;
        line = threadIdx.x / TL;
        lt = threadIdx.x - line * TL;
        row = (long)row0 + line;
        active = line < J.rb && row < J.N1;
        const int lp = L + (L >> 3);
        const int l = active ? line : 0;
        ln.a = smem + (long)l * lp;
        ln.b = LB == 2 ? smem + (long)(J.rb + l) * lp : ln.a;
        ln.es = 1;
        ln.pad = true;
        tw = smem + (long)LB * J.rb * lp;
    }
    __device__ long gidx(int q) const { return row * L + lt + q * TL; }
};

template <class R, int LB>
__device__ __forceinline__ void bfft(cx<R> (&v)[8], const BRowTile<R, LB> &B, const LinePlan &p, bool twl,
                                     int sign) {
    if (twl) fft8<R, LB == 1>(v, B.ln, p, B.tw, sign, B.lt, B.active, kTwLds);
    else fft8<R, LB == 1>(v, B.ln, p, reinterpret_cast<const cx<R> *>(p.tw), sign, B.lt, B.active, kTwMem);
}

// PHASE 0: init (phi_0 = psi, sum |psi|^2 of the step start, X = FFT_0 phi_0)
// PHASE 1: prologue (order 1); PHASE 2: order k >= 2
// CMP: phi and Y in the compact 40-bit term storage (R double)
template <class R, int NT, int PHASE, int LB, bool CMP = false>
__global__ void __launch_bounds__(NT) k_brow(const BJob<R> *__restrict__ jobs, const BTile *__restrict__ tiles,
                                             const LinePlan *__restrict__ plans, int k, const int *tick_p, int twl) {
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const BTile t = tiles[b];
    const BJob<R> J = jobs[t.job];
    const int step = *tick_p - J.start_tick;
    if (step < 0 || step >= J.nsteps || J.ctl->failed) return;
    if (PHASE == 2 && (k >= J.order || J.ctl->done)) return;
    extern __shared__ __align__(16) char smem_raw[];
    __shared__ double red[16];
    BRowTile<R, LB> B(J, t.first, reinterpret_cast<cx<R> *>(smem_raw));
    const LinePlan &p0 = plans[J.plan0];
    const int local = b - J.rt0;   // tile index within the job
    cx<double> *psi = reinterpret_cast<cx<double> *>(J.psi);
    cx<R> *A = reinterpret_cast<cx<R> *>(J.phiA);
    cx<R> *Bf = reinterpret_cast<cx<R> *>(J.phiB);
    cx<R> *X = reinterpret_cast<cx<R> *>(J.X);
    const bool two_d = J.nd == 2;

    if (PHASE == 0) {
        double acc = 0;
        if (B.active) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const cx<double> x = psi[B.gidx(q)];
                acc += x.r * x.r + x.i * x.i;
                tst<CMP>(A, B.gidx(q), J.size, from_d<R>(x));
            }
        }
        const double part = block_sum(acc, red);
        if (threadIdx.x == 0) {
            J.part[2 * J.nrt + local] = part;
            if (t.first == 0) {
                J.ctl->done = 0;
                J.ctl->last_order = 1;
            }
        }
        return;
    }

    const cx<R> *phim1 = (k - 1) % 2 == 0 ? A : Bf;
    cx<R> *phik = k % 2 == 0 ? A : Bf;   // phi_k replaces phi_{k-2}
    // only the transformed line stays live across the transforms; the
    // epilogue operands are read afterwards
    cx<R> v[8];
    if (B.active) {
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = tld<CMP>(phim1, B.gidx(q), J.size);
    }
    const bool has_t = J.tk0 != nullptr;
    if (twl) stage_twiddles(B.tw, p0);
    if (PHASE == 1) {
        if (t.first == 0) {
            const double n0 = fixed_sum(J.part + 2 * J.nrt, J.nrt, red);
            if (threadIdx.x == 0) {
                J.ctl->norm0 = n0;
                J.ctl->last_order = 1;
            }
        }
    } else {
        ConvCheck cc;
        cc.part = J.part + ((k - 1) % 2) * J.nrt;
        cc.norm0 = &J.ctl->norm0;
        cc.nb = J.nrt;
        cc.k_prev = k - 1;
        cc.prec = J.prec;
        double acc_prev;
        if (prev_converged(cc, red, &acc_prev)) {
            if (t.first == 0 && threadIdx.x == 0) {
                J.ctl->done = 1;
                J.ctl->accnorm = acc_prev;
            }
            return;
        }
        if (t.first == 0 && threadIdx.x == 0) J.ctl->last_order = k;
    }
    const double Ed = J.nt > 0 ? -J.laser[step] * J.lscale : 0.0;
    bfft(v, B, p0, twl, -1);
    // T_0/N_0 read after the forward transform (not held across it: the
    // registers bound the waves per SIMD; other waves hide the load)
    if (B.active) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const double t = has_t ? (0.0 + J.tk0[B.lt + q * B.TL]) * J.inv_n0 : 0.0;
            v[q] = mk<R>((R)((double)v[q].r * t), (R)((double)v[q].i * t));
        }
    }
    bfft(v, B, p0, twl, +1);

    // psi accumulation.  With the norm break test (prec > 0) every order adds
    // its term and stores psi.  The fixed-length FP32 series (prec == 0)
    // updates psi only at orders 1, 4, 7, ... and at the last order, adding
    // the terms since the previous update from the stored phi_{k-2},
    // phi_{k-1} and the new phi_k, in order.
    const bool defer = !(J.prec > 0);
    int nterm = 1;                                  // terms added to psi at this order
    if (PHASE == 2 && defer) {
        const int m = (k - 1) % 3;
        nterm = m == 0 ? 3 : (k == J.order - 1 ? m : 0);
    }
    const bool upd = nterm > 0;
    double acc = 0;
    if (B.active) {
        const double *cf = J.coef;
        const cx<double> ck = mk<double>(cf[2 * (PHASE == 1 ? 1 : k)], cf[2 * (PHASE == 1 ? 1 : k) + 1]);
        const cx<double> c0 = mk<double>(cf[0], cf[1]);
        cx<double> cm1 = mk<double>(0, 0), cm2 = mk<double>(0, 0);
        if (PHASE == 2 && nterm >= 2) cm1 = mk<double>(cf[2 * (k - 1)], cf[2 * (k - 1) + 1]);
        if (PHASE == 2 && nterm >= 3) cm2 = mk<double>(cf[2 * (k - 2)], cf[2 * (k - 2) + 1]);
        // two points at a time: fewer operands live, more waves per SIMD
        constexpr int kEp = 2;
#pragma unroll
        for (int h0 = 0; h0 < 8; h0 += kEp) {
            cx<double> ps[kEp];
            cx<R> m2[kEp], yv[kEp], pm[kEp];
            double wv[kEp];
#pragma unroll
            for (int u = 0; u < kEp; ++u) {
                const long gi = B.gidx(h0 + u);
                if (upd) ps[u] = psi[gi];
                pm[u] = tld<CMP>(phim1, gi, J.size);
                if (PHASE == 2) m2[u] = tld<CMP>(phik, gi, J.size);
                if (two_d) yv[u] = tld<CMP>(X, gi, J.size);
                double w = 0.0;
                if (J.V) w += (double)J.V[gi];
                if (J.mu) w += Ed * (double)J.mu[gi];
                wv[u] = w;
            }
#pragma unroll
            for (int u = 0; u < kEp; ++u) {
                const int q = h0 + u;
                const long gi = B.gidx(q);
                cx<double> hp = mk<double>(0, 0);
                if (has_t) hp = two_d ? to_d(v[q]) + to_d(yv[u]) : to_d(v[q]);
                const cx<double> p = to_d(pm[u]);
                hp.r += wv[u] * p.r;
                hp.i += wv[u] * p.i;
                const cx<double> bb = mk<double>((hp.r - J.offset * p.r) * J.inv_scaling,
                                                 (hp.i - J.offset * p.i) * J.inv_scaling);
                cx<double> pk;
                if (PHASE == 1) {
                    pk = mk<double>(bb.i, -bb.r);
                    psi[gi] = cmul(c0, ps[u]) + cmul(ck, pk);
                } else {
                    const cx<double> mm = to_d(m2[u]);
                    pk = mk<double>(2.0 * bb.i + mm.r, -2.0 * bb.r + mm.i);
                    if (upd) {
                        cx<double> sv = ps[u];
                        if (nterm >= 3) sv = sv + cmul(cm2, mm);
                        if (nterm >= 2) sv = sv + cmul(cm1, p);
                        sv = sv + cmul(ck, pk);
                        acc += sv.r * sv.r + sv.i * sv.i;
                        psi[gi] = sv;
                    }
                }
                tst<CMP>(phik, gi, J.size, from_d<R>(pk));
            }
        }
    }
    if (PHASE == 1) return;
    const double part = block_sum(acc, red);
    if (threadIdx.x == 0) J.part[(k % 2) * J.nrt + local] = part;
}

// dimension-1 part of the kinetic operator of a 2D job:
// Y = IFFT_1(T_1/N_1 FFT_1 phi_{k-1}), held in X; LB LDS buffers per line;
// CMP: phi and Y in the compact 40-bit term storage (R double)
template <class R, int LB, bool CMP = false>
__global__ void __launch_bounds__(256) k_bcol(const BJob<R> *__restrict__ jobs, const BTile *__restrict__ tiles,
                                              const LinePlan *__restrict__ plans, int k, const int *tick_p, int twl) {
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const BTile t = tiles[b];
    const BJob<R> J = jobs[t.job];
    const int step = *tick_p - J.start_tick;
    if (step < 0 || step >= J.nsteps || J.ctl->failed) return;
    if (k >= 2 && (k >= J.order || J.ctl->done)) return;
    extern __shared__ __align__(16) char smem_raw[];
    cx<R> *smem = reinterpret_cast<cx<R> *>(smem_raw);
    __shared__ double bc;
    const LinePlan &p = plans[J.plan1];
    const int N0 = J.N0, L = J.N1, TL = L >> 3, W = J.W;
    const int c = threadIdx.x % W;
    const int lt = threadIdx.x / W;
    const int col = t.first + c;
    const bool active = lt < TL && col < N0;
    cx<R> *Xb = reinterpret_cast<cx<R> *>(J.X);
    const cx<R> *in = reinterpret_cast<const cx<R> *>((k - 1) % 2 == 0 ? J.phiA : J.phiB);
    LdsLine<R> ln;
    ln.a = smem + c;
    ln.b = LB == 2 ? smem + (long)W * tw_lds_len(L) + c : ln.a;
    ln.es = W;
    ln.pad = true;
    cx<R> *twt = smem + (long)LB * W * tw_lds_len(L);
    cx<R> v[8];
    if (active) {
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = tld<CMP>(in, col + (long)(lt + q * TL) * N0, J.size);
    }
    if (twl) stage_twiddles(twt, p);
    if (k >= 3) {
        ConvCheck cc;
        cc.part = J.part + ((k - 1) % 2) * J.nrt;
        cc.norm0 = &J.ctl->norm0;
        cc.nb = J.nrt;
        cc.k_prev = k - 1;
        cc.prec = J.prec;
        double a;
        if (prev_converged(cc, &bc, &a)) return;
    }
    const cx<R> *tws = twl ? twt : reinterpret_cast<const cx<R> *>(p.tw);
    const int tsh = twl ? kTwLds : kTwMem;
    fft8<R, LB == 1>(v, ln, p, tws, -1, lt, active, tsh);
    if (active) {   // T_1/N_1 read after the forward transform, as in k_brow
        const bool has_t = J.tk1 != nullptr;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const double t = has_t ? (0.0 + J.tk1[lt + q * TL]) * J.inv_n1 : 0.0;
            v[q] = mk<R>((R)((double)v[q].r * t), (R)((double)v[q].i * t));
        }
    }
    fft8<R, LB == 1>(v, ln, p, tws, +1, lt, active, tsh);
    if (active) {
#pragma unroll
        for (int q = 0; q < 8; ++q) tst<CMP>(Xb, col + (long)(lt + q * TL) * N0, J.size, v[q]);
    }
}

template <class R>
__global__ void __launch_bounds__(256) k_btick_end(const BJob<R> *__restrict__ jobs, int njobs, int *tick) {
    const int tk = *tick;
    for (int j = threadIdx.x; j < njobs; j += blockDim.x) {
        const BJob<R> &J = jobs[j];
        const int step = tk - J.start_tick;
        if (step >= 0 && step < J.nsteps && !J.ctl->failed) {
            // a step that started from (Chebyshev) or produced (GSPO) a state
            // whose norm is not a number ends the job there; its batch-mates
            // are untouched
            if (!isfinite(J.ctl->norm0)) {
                J.ctl->failed = step + 1;
                continue;
            }
            J.stats[0] += J.kind == 1 ? 1 : J.ctl->last_order;
            J.stats[1] += 1;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) *tick = tk + 1;
}

// One whole time step of a 1D job in one block: the line (L <= 4096 points,
// 8 per thread) stays in registers through all orders of the series; only the
// FFT stages exchange through LDS.  Same arithmetic, order by order, as the
// per-order kernels (init, prologue, recurrence with the break test after
// each order), so the results agree with them bit for bit.
template <class R, int NT>
__global__ void __launch_bounds__(NT) k_bline(const BJob<R> *__restrict__ jobs, const int *__restrict__ list,
                                              const LinePlan *__restrict__ plans, const int *tick_p, int twl) {
    const BJob<R> J = jobs[list[blockIdx.x]];
    const int step = *tick_p - J.start_tick;
    if (step < 0 || step >= J.nsteps || J.ctl->failed) return;
    extern __shared__ __align__(16) char smem_raw[];
    __shared__ double red[16];
    __shared__ double bc;
    cx<R> *smem = reinterpret_cast<cx<R> *>(smem_raw);
    const LinePlan &p = plans[J.plan0];
    const int L = J.N0, TL = L >> 3, lt = threadIdx.x;
    const bool active = lt < TL;
    LdsLine<R> ln;
    ln.a = smem;
    ln.b = smem + (L + (L >> 3));
    ln.es = 1;
    ln.pad = true;
    cx<R> *twt = smem + 2 * (L + (L >> 3));
    if (twl) stage_twiddles(twt, p);
    const cx<R> *tws = twl ? twt : reinterpret_cast<const cx<R> *>(p.tw);
    const int tsh = twl ? kTwLds : kTwMem;
    cx<double> *psi = reinterpret_cast<cx<double> *>(J.psi);

    cx<double> ps[8];
    cx<R> pm1[8], pm2[8], v[8];
    double wv[8], tq[8];
    const bool has_t = J.tk0 != nullptr;
    const double Ed = J.nt > 0 ? -J.laser[step] * J.lscale : 0.0;
    double acc = 0;
    if (active) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int x = lt + q * TL;
            ps[q] = psi[x];
            acc += ps[q].r * ps[q].r + ps[q].i * ps[q].i;
            pm1[q] = from_d<R>(ps[q]);
            double w = 0.0;
            if (J.V) w += (double)J.V[x];
            if (J.mu) w += Ed * (double)J.mu[x];
            wv[q] = w;
            tq[q] = has_t ? (0.0 + J.tk0[x]) * J.inv_n : 0.0;
        }
    }
    // sum |psi|^2 at the step start, as the init kernel's single partial
    double norm0 = block_sum(acc, red);
    if (threadIdx.x == 0) bc = norm0;
    __syncthreads();
    norm0 = bc;
    const cx<double> c0 = mk<double>(J.coef[0], J.coef[1]);
    int last = 1;
    for (int k = 1; k < J.order; ++k) {
        // T phi_{k-1} (OGridNablaSq through Kspace: FFT, *T/N, IFFT)
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = pm1[q];
        fft8<R>(v, ln, p, tws, -1, lt, active, tsh);
        if (active) {
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] = mk<R>((R)((double)v[q].r * tq[q]), (R)((double)v[q].i * tq[q]));
        }
        fft8<R>(v, ln, p, tws, +1, lt, active, tsh);
        const cx<double> ck = mk<double>(J.coef[2 * k], J.coef[2 * k + 1]);
        acc = 0;
        if (active) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                cx<double> hp = has_t ? to_d(v[q]) : mk<double>(0, 0);
                const cx<double> ph = to_d(pm1[q]);
                hp.r += wv[q] * ph.r;
                hp.i += wv[q] * ph.i;
                const cx<double> bb = mk<double>((hp.r - J.offset * ph.r) * J.inv_scaling,
                                                 (hp.i - J.offset * ph.i) * J.inv_scaling);
                cx<double> pk;
                if (k == 1) {
                    pk = mk<double>(bb.i, -bb.r);
                    ps[q] = cmul(c0, ps[q]) + cmul(ck, pk);
                } else {
                    const cx<double> mm = to_d(pm2[q]);
                    pk = mk<double>(2.0 * bb.i + mm.r, -2.0 * bb.r + mm.i);
                    ps[q] = ps[q] + cmul(ck, pk);
                    acc += ps[q].r * ps[q].r + ps[q].i * ps[q].i;
                }
                pm2[q] = pm1[q];
                pm1[q] = from_d<R>(pk);
            }
        }
        last = k;
        // break test after order k (OCheby.cpp:179-182); off when prec == 0
        if (k >= 2 && J.prec > 0) {
            double a = block_sum(acc, red);
            if (threadIdx.x == 0) bc = a;
            __syncthreads();
            a = bc;
            if (fabs(norm0 - a) / norm0 < J.prec) break;
        }
    }
    if (active) {
#pragma unroll
        for (int q = 0; q < 8; ++q) psi[lt + q * TL] = ps[q];
    }
    if (threadIdx.x == 0) {
        J.ctl->norm0 = norm0;
        J.ctl->last_order = last;
        J.ctl->done = 1;
    }
}


// ---------------------------------------------------------------------------
// GSPO jobs

// psi <- exp(-i tau W) psi on the S <= 2 states of one grid point, W real
// symmetric: per-state phases, or for coupled states
// e^{-i tau m} [cos(tau O) - i sin(tau O)/O (W - m)], m = (W00 + W11)/2,
// O = sqrt(((W00 - W11)/2)^2 + W01^2) (ODMultistate.cpp:109-159 diagonalises W
// per point through TransformMSD.cpp:31-101; same operator)
__device__ __forceinline__ void wsplit(int S, bool coupled, double tau, double w00, double w11, double w01,
                                       cx<double> &a, cx<double> &b) {
    if (S == 1 || !coupled) {
        double sn, cs;
        sincos(-tau * w00, &sn, &cs);
        a = cmul(a, mk<double>(cs, sn));
        if (S == 2) {
            sincos(-tau * w11, &sn, &cs);
            b = cmul(b, mk<double>(cs, sn));
        }
        return;
    }
    const double m = 0.5 * (w00 + w11);
    const double d = 0.5 * (w00 - w11);
    const double O = sqrt(d * d + w01 * w01);
    double so, c;
    sincos(tau * O, &so, &c);
    const double sc = O > 0 ? so / O : tau;
    double sm, cm;
    sincos(-tau * m, &sm, &cm);
    const cx<double> ph = mk<double>(cm, sm);
    const cx<double> u00 = mk<double>(c, -sc * d), u11 = mk<double>(c, sc * d);
    const cx<double> u01 = mk<double>(0, -sc * w01);
    const cx<double> a0 = cmul(u00, a) + cmul(u01, b);
    const cx<double> a1 = cmul(u01, a) + cmul(u11, b);
    a = cmul(ph, a0);
    b = cmul(ph, a1);
}

template <class R>
__device__ __forceinline__ void wsplit_point(const BJob<R> &J, long gi, double Ed, double tau, cx<R> &v0, cx<R> &v1) {
    double w00 = 0.0, w11 = 0.0, w01 = 0.0;
    if (J.V) w00 += (double)J.V[gi];
    if (J.mu) w00 += Ed * (double)J.mu[gi];
    if (J.S == 2) {
        if (J.V11) w11 += (double)J.V11[gi];
        if (J.V01) w01 += (double)J.V01[gi];
    }
    cx<double> a = to_d(v0), b = J.S == 2 ? to_d(v1) : mk<double>(0, 0);
    wsplit(J.S, J.V01 != nullptr, tau, w00, w11, w01, a, b);
    v0 = from_d<R>(a);
    if (J.S == 2) v1 = from_d<R>(b);
}

// exp(-i tau T(k))/N with T summed over dimensions in dimension order; the
// 1/N of the inverse transform folded in (OGSPO.cpp:37-41)
__device__ __forceinline__ cx<double> kin_phase(double tt, double tau, double inv_n) {
    double sn, cs;
    sincos(-tau * tt, &sn, &cs);
    return mk<double>(cs * inv_n, sn * inv_n);
}

// PHASE 0: psi -> exp(-i tau/2 W) -> FFT_0 -> X (2D), or the whole step of a
// 1D job; PHASE 1: X -> IFFT_0 -> exp(-i tau/2 W) -> psi (2D jobs).  A thread
// carries both states of its points, transforming them one after the other.
template <class R, int NT, int PHASE>
__global__ void __launch_bounds__(NT) k_bgspo_row(const BJob<R> *__restrict__ jobs, const BTile *__restrict__ tiles,
                                                  const LinePlan *__restrict__ plans, const int *tick_p, int twl) {
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const BTile t = tiles[b];
    const BJob<R> J = jobs[t.job];
    const int step = *tick_p - J.start_tick;
    if (step < 0 || step >= J.nsteps || J.ctl->failed) return;
    if (PHASE == 1 && J.nd == 1) return;
    extern __shared__ __align__(16) char smem_raw[];
    BRowTile<R, 2> B(J, t.first, reinterpret_cast<cx<R> *>(smem_raw));
    const LinePlan &p0 = plans[J.plan0];
    if (twl) stage_twiddles(B.tw, p0);
    const bool two = J.S == 2;
    cx<R> *psi = reinterpret_cast<cx<R> *>(J.psiR);
    cx<R> *X = reinterpret_cast<cx<R> *>(J.X);
    const double Ed = J.nt > 0 ? -J.laser[step] * J.lscale : 0.0;
    const double th = 0.5 * J.tau;
    cx<R> v0[8], v1[8];
    if (PHASE == 0) {
        if (B.active) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                v0[q] = psi[B.gidx(q)];
                if (two) v1[q] = psi[J.size + B.gidx(q)];
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) wsplit_point(J, B.gidx(q), Ed, th, v0[q], v1[q]);
        }
        bfft(v0, B, p0, twl, -1);
        if (two) bfft(v1, B, p0, twl, -1);
        if (J.nd == 2) {
            if (B.active) {
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    X[B.gidx(q)] = v0[q];
                    if (two) X[J.size + B.gidx(q)] = v1[q];
                }
            }
            return;
        }
        if (B.active) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int x = B.lt + q * B.TL;
                const double t0 = J.tk0 ? 0.0 + J.tk0[x] : 0.0;
                v0[q] = from_d<R>(cmul(to_d(v0[q]), kin_phase(t0, J.tau, J.inv_n)));
                if (two) {
                    const double t1 = J.tk0s1 ? 0.0 + J.tk0s1[x] : 0.0;
                    v1[q] = from_d<R>(cmul(to_d(v1[q]), kin_phase(t1, J.tau, J.inv_n)));
                }
            }
        }
    } else {
        if (B.active) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                v0[q] = X[B.gidx(q)];
                if (two) v1[q] = X[J.size + B.gidx(q)];
            }
        }
    }
    bfft(v0, B, p0, twl, +1);
    if (two) bfft(v1, B, p0, twl, +1);
    if (B.active) {
        bool bad = false;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            wsplit_point(J, B.gidx(q), Ed, th, v0[q], v1[q]);
            psi[B.gidx(q)] = v0[q];
            if (two) psi[J.size + B.gidx(q)] = v1[q];
            bad |= !isfinite((double)v0[q].r + (double)v0[q].i);
            if (two) bad |= !isfinite((double)v1[q].r + (double)v1[q].i);
        }
        // a step that leaves a value that is not a number marks the job
        // (read at the tick's end); every writer stores the same NaN
        if (bad) J.ctl->norm0 = __builtin_nan("");
    }
}

// GSPO kinetic split operator along dimension 1 of a 2D job's X (one state
// per tile): forward * exp(-i tau T)/N * inverse
template <class R>
__global__ void __launch_bounds__(256) k_bgspo_col(const BJob<R> *__restrict__ jobs, const BTile *__restrict__ tiles,
                                                   const LinePlan *__restrict__ plans, const int *tick_p, int twl) {
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const BTile t = tiles[b];
    const BJob<R> J = jobs[t.job];
    const int step = *tick_p - J.start_tick;
    if (step < 0 || step >= J.nsteps || J.ctl->failed) return;
    extern __shared__ __align__(16) char smem_raw[];
    cx<R> *smem = reinterpret_cast<cx<R> *>(smem_raw);
    const LinePlan &p = plans[J.plan1];
    const int N0 = J.N0, L = J.N1, TL = L >> 3, W = J.W;
    const int st = t.first / N0;
    const int c = threadIdx.x % W;
    const int lt = threadIdx.x / W;
    const int col = t.first - st * N0 + c;
    const bool active = lt < TL && col < N0;
    cx<R> *base = reinterpret_cast<cx<R> *>(J.X) + (long)st * J.size + col;
    LdsLine<R> ln;
    ln.a = smem + c;
    ln.b = smem + (long)W * tw_lds_len(L) + c;
    ln.es = W;
    ln.pad = true;
    cx<R> *twt = smem + (long)2 * W * tw_lds_len(L);
    const double *tk0 = st == 0 ? J.tk0 : J.tk0s1;
    const double *tk1 = st == 0 ? J.tk1 : J.tk1s1;
    cx<R> v[8];
    double tq[8];
    if (active) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            v[q] = base[(long)(lt + q * TL) * N0];
            double tt = 0.0;
            if (tk0) {
                tt = 0.0 + tk0[col];
                tt += tk1[lt + q * TL];
            }
            tq[q] = tt;
        }
    }
    if (twl) stage_twiddles(twt, p);
    const cx<R> *tws = twl ? twt : reinterpret_cast<const cx<R> *>(p.tw);
    const int tsh = twl ? kTwLds : kTwMem;
    fft8<R>(v, ln, p, tws, -1, lt, active, tsh);
    if (active) {
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = from_d<R>(cmul(to_d(v[q]), kin_phase(tq[q], J.tau, J.inv_n)));
    }
    fft8<R>(v, ln, p, tws, +1, lt, active, tsh);
    if (active) {
#pragma unroll
        for (int q = 0; q < 8; ++q) base[(long)(lt + q * TL) * N0] = v[q];
    }
}

// ---------------------------------------------------------------------------
// launchers

// raise the kernel's dynamic-LDS limit when a launch needs more than 64 KiB
// (one attribute call per kernel instance and new maximum)
template <class K>
static void allow_lds(K kernel, size_t bytes) {
    static size_t allowed = 65536;
    if (bytes <= allowed) return;
    QDX_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)bytes));
    allowed = bytes;
}

template <class R, int NT, int PHASE, int LB, bool CMP>
static void row_launch(hipStream_t s, int k, const BJob<R> *jobs, const BTile *tiles, int ntiles,
                       const LinePlan *plans, const int *tick, size_t lds, bool twl) {
    allow_lds(k_brow<R, NT, PHASE, LB, CMP>, lds);
    QDX_LAUNCH((k_brow<R, NT, PHASE, LB, CMP>), dim3(ntiles), dim3(NT), lds, s, jobs, tiles, plans, k, tick,
               twl ? 1 : 0);
    QDX_HIP(hipGetLastError());
}

template <class R, int NT, bool CMP>
static void row_phase(hipStream_t s, int lbuf, int phase, int k, const BJob<R> *jobs, const BTile *tiles,
                      int ntiles, const LinePlan *plans, const int *tick, size_t lds, bool twl) {
    // the init phase runs no transform: one instance serves both layouts
    if (phase == 0) row_launch<R, NT, 0, 2, CMP>(s, k, jobs, tiles, ntiles, plans, tick, lds, twl);
    else if (phase == 1 && lbuf == 2) row_launch<R, NT, 1, 2, CMP>(s, k, jobs, tiles, ntiles, plans, tick, lds, twl);
    else if (phase == 1) row_launch<R, NT, 1, 1, CMP>(s, k, jobs, tiles, ntiles, plans, tick, lds, twl);
    else if (lbuf == 2) row_launch<R, NT, 2, 2, CMP>(s, k, jobs, tiles, ntiles, plans, tick, lds, twl);
    else row_launch<R, NT, 2, 1, CMP>(s, k, jobs, tiles, ntiles, plans, tick, lds, twl);
}

template <class R>
void BatchLaunch<R>::row(hipStream_t s, int NT, int lbuf, int phase, int k, const BJob<R> *jobs,
                         const BTile *tiles, int ntiles, const LinePlan *plans, const int *tick, size_t lds,
                         bool tw_lds, bool cmp) {
    if (ntiles <= 0) return;
    if (lds > 160 * 1024) throw std::runtime_error("row tile exceeds the 160 KiB LDS of a gfx950 CU");
    if (lbuf != 1 && lbuf != 2) throw std::runtime_error("row tiles hold 1 or 2 LDS buffers per line");
    if (NT != 256 && NT != 512) throw std::runtime_error("row tiles use 256 or 512 threads");
    if (cmp) {
        if constexpr (std::is_same<R, double>::value) {
            if (NT == 256) row_phase<R, 256, true>(s, lbuf, phase, k, jobs, tiles, ntiles, plans, tick, lds, tw_lds);
            else row_phase<R, 512, true>(s, lbuf, phase, k, jobs, tiles, ntiles, plans, tick, lds, tw_lds);
        } else {
            throw std::runtime_error("compact terms: FP64 batches only");
        }
        return;
    }
    if (NT == 256) row_phase<R, 256, false>(s, lbuf, phase, k, jobs, tiles, ntiles, plans, tick, lds, tw_lds);
    else row_phase<R, 512, false>(s, lbuf, phase, k, jobs, tiles, ntiles, plans, tick, lds, tw_lds);
}

template <class R, int LB, bool CMP>
static void col_launch(hipStream_t s, int k, const BJob<R> *jobs, const BTile *tiles, int ntiles,
                       const LinePlan *plans, const int *tick, size_t lds, bool tw_lds) {
    allow_lds(k_bcol<R, LB, CMP>, lds);
    QDX_LAUNCH((k_bcol<R, LB, CMP>), dim3(ntiles), dim3(256), lds, s, jobs, tiles, plans, k, tick, tw_lds ? 1 : 0);
    QDX_HIP(hipGetLastError());
}

template <class R>
void BatchLaunch<R>::col(hipStream_t s, int lbuf, int k, const BJob<R> *jobs, const BTile *tiles, int ntiles,
                         const LinePlan *plans, const int *tick, size_t lds, bool tw_lds, bool cmp) {
    if (ntiles <= 0) return;
    if (lds > 160 * 1024) throw std::runtime_error("column tile exceeds the 160 KiB LDS of a gfx950 CU");
    if (lbuf != 1 && lbuf != 2) throw std::runtime_error("column tiles hold 1 or 2 LDS buffers per line");
    if (cmp) {
        if constexpr (std::is_same<R, double>::value) {
            if (lbuf == 2) col_launch<R, 2, true>(s, k, jobs, tiles, ntiles, plans, tick, lds, tw_lds);
            else col_launch<R, 1, true>(s, k, jobs, tiles, ntiles, plans, tick, lds, tw_lds);
        } else {
            throw std::runtime_error("compact terms: FP64 batches only");
        }
        return;
    }
    if (lbuf == 2) col_launch<R, 2, false>(s, k, jobs, tiles, ntiles, plans, tick, lds, tw_lds);
    else col_launch<R, 1, false>(s, k, jobs, tiles, ntiles, plans, tick, lds, tw_lds);
}

template <class R>
void BatchLaunch<R>::line(hipStream_t s, int NT, const BJob<R> *jobs, const int *list, int n, const LinePlan *plans,
                          const int *tick, size_t lds, bool tw_lds) {
    if (n <= 0) return;
    if (lds > 160 * 1024) throw std::runtime_error("1D job exceeds the 160 KiB LDS of a gfx950 CU");
    if (NT == 256) {
        allow_lds(k_bline<R, 256>, lds);
        QDX_LAUNCH((k_bline<R, 256>), dim3(n), dim3(256), lds, s, jobs, list, plans, tick, tw_lds ? 1 : 0);
    } else if (NT == 512) {
        allow_lds(k_bline<R, 512>, lds);
        QDX_LAUNCH((k_bline<R, 512>), dim3(n), dim3(512), lds, s, jobs, list, plans, tick, tw_lds ? 1 : 0);
    } else {
        throw std::runtime_error("1D jobs use 256 or 512 threads");
    }
    QDX_HIP(hipGetLastError());
}

template <class R>
void BatchLaunch<R>::gspo_row(hipStream_t s, int NT, int phase, const BJob<R> *jobs, const BTile *tiles, int ntiles,
                              const LinePlan *plans, const int *tick, size_t lds, bool tw_lds) {
    if (ntiles <= 0) return;
    if (lds > 160 * 1024) throw std::runtime_error("row tile exceeds the 160 KiB LDS of a gfx950 CU");
    const int twl = tw_lds ? 1 : 0;
    if (NT == 256) {
        if (phase == 0) {
            allow_lds(k_bgspo_row<R, 256, 0>, lds);
            QDX_LAUNCH((k_bgspo_row<R, 256, 0>), dim3(ntiles), dim3(256), lds, s, jobs, tiles, plans, tick, twl);
        } else {
            allow_lds(k_bgspo_row<R, 256, 1>, lds);
            QDX_LAUNCH((k_bgspo_row<R, 256, 1>), dim3(ntiles), dim3(256), lds, s, jobs, tiles, plans, tick, twl);
        }
    } else if (NT == 512) {
        if (phase == 0) {
            allow_lds(k_bgspo_row<R, 512, 0>, lds);
            QDX_LAUNCH((k_bgspo_row<R, 512, 0>), dim3(ntiles), dim3(512), lds, s, jobs, tiles, plans, tick, twl);
        } else {
            allow_lds(k_bgspo_row<R, 512, 1>, lds);
            QDX_LAUNCH((k_bgspo_row<R, 512, 1>), dim3(ntiles), dim3(512), lds, s, jobs, tiles, plans, tick, twl);
        }
    } else {
        throw std::runtime_error("row tiles use 256 or 512 threads");
    }
    QDX_HIP(hipGetLastError());
}

template <class R>
void BatchLaunch<R>::gspo_col(hipStream_t s, const BJob<R> *jobs, const BTile *tiles, int ntiles,
                              const LinePlan *plans, const int *tick, size_t lds, bool tw_lds) {
    if (ntiles <= 0) return;
    if (lds > 160 * 1024) throw std::runtime_error("column tile exceeds the 160 KiB LDS of a gfx950 CU");
    allow_lds(k_bgspo_col<R>, lds);
    QDX_LAUNCH((k_bgspo_col<R>), dim3(ntiles), dim3(256), lds, s, jobs, tiles, plans, tick, tw_lds ? 1 : 0);
    QDX_HIP(hipGetLastError());
}

template <class R>
void BatchLaunch<R>::tick_end(hipStream_t s, const BJob<R> *jobs, int njobs, int *tick) {
    hipLaunchKernelGGL((k_btick_end<R>), dim3(1), dim3(256), 0, s, jobs, njobs, tick);
    QDX_HIP(hipGetLastError());
}

template struct BatchLaunch<double>;
template struct BatchLaunch<float>;

}  // namespace qdx
