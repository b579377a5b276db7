// Device building blocks for gfx950: complex arithmetic, radix-2/3/4/8
// Stockham stages on LDS-resident lines, wave64 reductions.
#pragma once

#include <hip/hip_runtime.h>

#include "qdx_impl.hpp"

namespace qdx {

template <class R>
struct alignas(2 * sizeof(R)) cx {
    R r, i;
};

template <class R>
__device__ __forceinline__ cx<R> mk(R r, R i) {
    cx<R> c;
    c.r = r;
    c.i = i;
    return c;
}
template <class R>
__device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) { return mk<R>(a.r + b.r, a.i + b.i); }
template <class R>
__device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) { return mk<R>(a.r - b.r, a.i - b.i); }
template <class R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
    return mk<R>(a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r);
}
template <class R>
__device__ __forceinline__ cx<R> scale(cx<R> a, R s) { return mk<R>(a.r * s, a.i * s); }
// multiply by sign * i
template <class R>
__device__ __forceinline__ cx<R> mul_si(cx<R> a, int sign) {
    return sign < 0 ? mk<R>(a.i, -a.r) : mk<R>(-a.i, a.r);
}
template <class R>
__device__ __forceinline__ cx<double> to_d(cx<R> a) { return mk<double>((double)a.r, (double)a.i); }
template <class R>
__device__ __forceinline__ cx<R> from_d(cx<double> a) { return mk<R>((R)a.r, (R)a.i); }

// ---------------------------------------------------------------------------
// r-point DFTs, y[u] = sum_t x[t] exp(sign 2 pi i t u / r)
template <int RR, class R>
struct Bfly;

template <class R>
struct Bfly<2, R> {
    static __device__ __forceinline__ void run(cx<R> *x, int) {
        cx<R> a = x[0], b = x[1];
        x[0] = a + b;
        x[1] = a - b;
    }
};

template <class R>
struct Bfly<4, R> {
    static __device__ __forceinline__ void run(cx<R> *x, int sign) {
        cx<R> s02 = x[0] + x[2], d02 = x[0] - x[2];
        cx<R> s13 = x[1] + x[3], d13 = mul_si(x[1] - x[3], sign);
        x[0] = s02 + s13;
        x[2] = s02 - s13;
        x[1] = d02 + d13;
        x[3] = d02 - d13;
    }
};

template <class R>
struct Bfly<3, R> {
    static __device__ __forceinline__ void run(cx<R> *x, int sign) {
        const R h = (R)0.86602540378443864676372317075293618;  // sqrt(3)/2
        cx<R> t1 = x[1] + x[2];
        cx<R> t2 = mul_si(scale(x[1] - x[2], h), sign);
        cx<R> m = x[0] - scale(t1, (R)0.5);
        x[0] = x[0] + t1;
        x[1] = m + t2;
        x[2] = m - t2;
    }
};

template <class R>
struct Bfly<8, R> {
    static __device__ __forceinline__ void run(cx<R> *x, int sign) {
        const R c = (R)0.70710678118654752440084436210484904;  // 1/sqrt(2)
        cx<R> e[4] = {x[0], x[2], x[4], x[6]};
        cx<R> o[4] = {x[1], x[3], x[5], x[7]};
        Bfly<4, R>::run(e, sign);
        Bfly<4, R>::run(o, sign);
        // o[u] *= exp(sign 2 pi i u / 8)
        cx<R> o1 = o[1], o3 = o[3];
        o[1] = scale(sign < 0 ? mk<R>(o1.r + o1.i, o1.i - o1.r) : mk<R>(o1.r - o1.i, o1.i + o1.r), c);
        o[2] = mul_si(o[2], sign);
        o[3] = scale(sign < 0 ? mk<R>(-o3.r + o3.i, -o3.i - o3.r) : mk<R>(-o3.r - o3.i, -o3.i + o3.r), c);
        for (int u = 0; u < 4; ++u) {
            x[u] = e[u] + o[u];
            x[u + 4] = e[u] - o[u];
        }
    }
};

// ---------------------------------------------------------------------------
// One Stockham stage on `nlines` lines of length L held in LDS.  Element q of
// line l sits at s[l * lstr + q * estr].  Thread (line, lt) handles
// butterflies j = lt + p * TL.  The stage reads all its inputs into
// registers, synchronises, and writes its outputs in place.
template <int RR, class R>
__device__ __forceinline__ void fft_stage(cx<R> *s, int L, int Ns, int sign, const cx<R> *__restrict__ tw,
                                          int line, int lt, int TL, int lstr, int estr, bool active) {
    constexpr int Q = RR == 2 ? 4 : RR == 3 ? 3 : RR == 4 ? 2 : 1;
    const int m = L / RR;
    cx<R> v[Q][RR];
    cx<R> *base = s + line * lstr;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = lt + q * TL;
        if (active && j < m) {
#pragma unroll
            for (int t = 0; t < RR; ++t) v[q][t] = base[(j + t * m) * estr];
        }
    }
    __syncthreads();
    const int tstep = L / (Ns * RR);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = lt + q * TL;
        if (active && j < m) {
            const int k = j % Ns;
            if (k != 0) {
#pragma unroll
                for (int t = 1; t < RR; ++t) {
                    cx<R> w = tw[k * t * tstep];
                    if (sign > 0) w.i = -w.i;
                    v[q][t] = cmul(v[q][t], w);
                }
            }
            Bfly<RR, R>::run(v[q], sign);
            const int ob = (j - k) * RR + k;
#pragma unroll
            for (int t = 0; t < RR; ++t) base[(ob + t * Ns) * estr] = v[q][t];
        }
    }
    __syncthreads();
}

// Full transform of LDS-resident lines (unnormalised, sign -1 forward).
// rowmajor: line l occupies s[l*L .. l*L+L); otherwise element q of line l is
// at s[q*nlines + l] (a tile of adjacent columns).
// Requires blockDim.x >= nlines * ceil(L/8).  All threads must call it.
template <class R>
__device__ void fft_lines(cx<R> *s, const LinePlan &p, int nlines, bool rowmajor, int sign) {
    const int L = p.L;
    if (L == 1) return;
    const int TL = (L + 7) / 8;
    const int tid = threadIdx.x;
    int line, lt, lstr, estr;
    bool active;
    if (rowmajor) {
        line = tid / TL;
        lt = tid % TL;
        active = line < nlines;
        lstr = L;
        estr = 1;
    } else {
        line = tid % nlines;
        lt = tid / nlines;
        active = lt < TL;
        lstr = 1;
        estr = nlines;
    }
    if (!active) line = 0;
    const cx<R> *tw = reinterpret_cast<const cx<R> *>(p.tw);
    int Ns = 1;
    for (int st = 0; st < p.nst; ++st) {
        const int r = p.rad[st];
        switch (r) {
        case 8: fft_stage<8, R>(s, L, Ns, sign, tw, line, lt, TL, lstr, estr, active); break;
        case 4: fft_stage<4, R>(s, L, Ns, sign, tw, line, lt, TL, lstr, estr, active); break;
        case 2: fft_stage<2, R>(s, L, Ns, sign, tw, line, lt, TL, lstr, estr, active); break;
        default: fft_stage<3, R>(s, L, Ns, sign, tw, line, lt, TL, lstr, estr, active); break;
        }
        Ns *= r;
    }
}

// ---------------------------------------------------------------------------
// FP64 block reduction (wave64 shuffles, then LDS across waves).  Returns the
// block sum in thread 0.  Summation order depends only on the thread layout,
// so results are reproducible launch to launch.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

__device__ __forceinline__ double block_sum(double v, double *red /* >= 16 doubles */) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < nw; ++i) t += red[i];
    return t;
}

// Last-block completion of a grid-wide reduction: every block stores its
// partial, the block that arrives last sums the partials in index order.
// Returns true (in all threads of that block) for the last block; `total`
// then holds the sum in thread 0.
__device__ __forceinline__ bool last_block_sum(double part, double *partials, unsigned *ticket,
                                               double &total, int *flag_lds) {
    const int nb = gridDim.x * gridDim.y;
    const int bid = blockIdx.y * gridDim.x + blockIdx.x;
    if (threadIdx.x == 0) {
        __hip_atomic_store(&partials[bid], part, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        unsigned t = atomicAdd(ticket, 1u);
        *flag_lds = (t == (unsigned)(nb - 1));
    }
    __syncthreads();
    const bool last = *flag_lds != 0;
    if (last) {
        __threadfence();
        double v = 0;
        for (int i = threadIdx.x; i < nb; i += blockDim.x)
            v += __hip_atomic_load(&partials[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // fixed-order combination: per-thread strided sums, then block_sum
        __shared__ double red[16];
        total = block_sum(v, red);
        if (threadIdx.x == 0) *ticket = 0u;
    }
    return last;
}

}  // namespace qdx
