// Device building blocks for gfx950: complex arithmetic, radix-2/3/4/8
// Stockham line transforms with register-resident first/last stages, wave64
// reductions.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "qdx_impl.hpp"

namespace qdx {

#define QDX_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") +         \
                                                       hipGetErrorString(e_) + " at " #call); \
    } while (0)

// kernel launch that carries the profiling events of g_launch_timing, if any
#define QDX_LAUNCH(kernel, grid, block, lds, stream, ...)                                          \
    do {                                                                                          \
        if (g_launch_timing.start) {                                                              \
            hipExtLaunchKernelGGL(kernel, grid, block, lds, stream, g_launch_timing.start,        \
                                  g_launch_timing.stop, 0, __VA_ARGS__);                          \
            g_launch_timing = LaunchTiming();                                                     \
        } else {                                                                                  \
            hipLaunchKernelGGL(kernel, grid, block, lds, stream, __VA_ARGS__);                    \
        }                                                                                         \
    } while (0)

__device__ __forceinline__ int xcd_remap(int b, int nb) {
    // 8 XCDs receive blocks round robin (block b on XCD b % 8); give each XCD
    // a contiguous range of tiles so that neighbouring tiles share one L2.
    // XCD x runs nb/8 blocks, plus one when x < nb % 8.
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}


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
// Compact (40-bit) storage of complex128 terms.  Each component keeps the
// upper 40 bits of its FP64 pattern (sign, exponent and the 28 leading
// mantissa bits), rounded to nearest on the dropped bits: relative error at
// most 2^-29 (FP32 storage: 2^-24); encoding and decoding are integer moves.
// A buffer of n complex values holds the n pairs of upper words (8 n bytes:
// bits 63..32 of re and im) followed by the n pairs of the next bytes (2 n
// bytes: bits 31..24).
struct alignas(2) u8x2 {
    unsigned char x, y;
};

__device__ __forceinline__ void t40_split(double x, unsigned &hi, unsigned char &lo) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x) + 0x800000ull;   // round at bit 23
    hi = (unsigned)(b >> 32);
    lo = (unsigned char)(b >> 24);
}
__device__ __forceinline__ double t40_join(unsigned hi, unsigned char lo) {
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | ((unsigned long long)lo << 24)));
}

// element i of a term buffer of n complex values: plain cx<R>, or (CMP, R
// double) the compact layout above
template <bool CMP, class R>
__device__ __forceinline__ cx<R> tld(const cx<R> *p, long i, long n) {
    if constexpr (!CMP) {
        return p[i];
    } else {
        static_assert(sizeof(R) == 8, "compact storage holds complex128 terms");
        const uint2 h = reinterpret_cast<const uint2 *>(p)[i];
        const u8x2 l = reinterpret_cast<const u8x2 *>(reinterpret_cast<const uint2 *>(p) + n)[i];
        return mk<R>(t40_join(h.x, l.x), t40_join(h.y, l.y));
    }
}
template <bool CMP, class R>
__device__ __forceinline__ void tst(cx<R> *p, long i, long n, cx<R> v) {
    if constexpr (!CMP) {
        p[i] = v;
    } else {
        static_assert(sizeof(R) == 8, "compact storage holds complex128 terms");
        uint2 h;
        u8x2 l;
        t40_split(v.r, h.x, l.x);
        t40_split(v.i, h.y, l.y);
        reinterpret_cast<uint2 *>(p)[i] = h;
        reinterpret_cast<u8x2 *>(reinterpret_cast<uint2 *>(p) + n)[i] = l;
    }
}

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
// Register-resident Stockham transform of one line of length L = P*TL
// (P = 8 or 4 points per thread).
//
// Thread lt (0 <= lt < TL) of a line owns elements lt + q*TL, q = 0..P-1, in
// v[q].  In the Stockham formulation the first stage (span 1) reads exactly
// those elements and the last stage (span L/r) writes exactly those elements,
// so both run on registers; only the middle stages go through LDS, using two
// buffers alternately (one barrier per stage).  The plan's first and last
// radices are 8, 4 or 2 and divide P; middle stages may also be radix 3 or
// (two-buffer schedule only) another prime (mid_stage_gen).
//
// LDS addressing of one line: element i of buffer a / b lives at
// a[idx(i)] / b[idx(i)], idx(i) = (pad ? i + i/8 : i) * es.
template <class R>
struct LdsLine {
    cx<R> *a;
    cx<R> *b;
    int es;
    bool pad;
    __device__ __forceinline__ int idx(int i) const { return (pad ? i + (i >> 3) : i) * es; }
};

// twiddle q of a table; tsh = 3 for the padded LDS copy (entry q at
// q + q/8, see stage_twiddles), 31 for the plain table in memory
template <class R>
__device__ __forceinline__ cx<R> twiddle(const cx<R> *tw, int q, int sign, int tsh) {
    cx<R> w = tw[q + (q >> tsh)];
    if (sign > 0) w.i = -w.i;
    return w;
}

// butterflies of one middle stage per thread: ceil(P / RR), the last one
// guarded by j < L / RR
template <int RR, int P>
constexpr int stage_bflies() { return (P + RR - 1) / RR; }

template <int RR, class R, int P = 8>
__device__ __forceinline__ void mid_stage(const cx<R> *in, cx<R> *out, const LdsLine<R> &ln, int L, int Ns,
                                          int TL, int lt, int sign, const cx<R> *tw, int tsh) {
    constexpr int Q = stage_bflies<RR, P>();
    const int m = L / RR;
    const int tstep = L / (Ns * RR);
    const bool pow2 = (Ns & (Ns - 1)) == 0;
#pragma unroll
    for (int p = 0; p < Q; ++p) {
        const int j = lt + p * TL;
        if (j < m) {
            const int k = pow2 ? (j & (Ns - 1)) : (j % Ns);
            cx<R> x[RR];
#pragma unroll
            for (int t = 0; t < RR; ++t) x[t] = in[ln.idx(j + t * m)];
            if (k != 0) {
#pragma unroll
                for (int t = 1; t < RR; ++t) x[t] = cmul(x[t], twiddle(tw, k * t * tstep, sign, tsh));
            }
            Bfly<RR, R>::run(x, sign);
            const int ob = (j - k) * RR + k;
#pragma unroll
            for (int t = 0; t < RR; ++t) out[ln.idx(ob + t * Ns)] = x[t];
        }
    }
}

// A middle stage of an odd prime radix r > 3 (5, 7, 11, ...; lengths with
// such factors only).  Output u of butterfly j sums its r inputs, each read
// again from LDS, times one table twiddle: exp(sign 2 pi i (k t / (Ns r) +
// (t u mod r) / r)) = table entry k t L/(Ns r) + (t u mod r) L/r (mod L).
// Only one accumulator stays live, so the kernels keep the registers of
// their radix 8/4/3/2 stages; the cost is r^2 LDS reads per butterfly.
template <class R>
__device__ __forceinline__ void mid_stage_gen(const cx<R> *in, cx<R> *out, const LdsLine<R> &ln, int L, int Ns,
                                              int TL, int lt, int sign, const cx<R> *tw, int tsh, int r) {
    const int m = L / r;
    const int tstep = L / (Ns * r);
    const int ustep = L / r;
#pragma unroll 1
    for (int j = lt; j < m; j += TL) {
        const int k = j % Ns;
        const int ob = (j - k) * r + k;
#pragma unroll 1
        for (int u = 0; u < r; ++u) {
            cx<R> acc = in[ln.idx(j)];
            int tu = 0;   // t u mod r
#pragma unroll 1
            for (int t = 1; t < r; ++t) {
                tu += u;
                if (tu >= r) tu -= r;
                int q = k * t * tstep + tu * ustep;
                if (q >= L) q -= L;
                acc = acc + cmul(in[ln.idx(j + t * m)], twiddle(tw, q, sign, tsh));
            }
            out[ln.idx(ob + u * Ns)] = acc;
        }
    }
}

// The same stage on one buffer (in place): every thread reads its inputs,
// the block synchronises, then the outputs are written.  Called by all
// threads of the block (it synchronises); threads without a line pass
// active = false.
template <int RR, class R, int P = 8>
__device__ __forceinline__ void mid_stage_ip(cx<R> *buf, const LdsLine<R> &ln, int L, int Ns, int TL, int lt,
                                             int sign, const cx<R> *tw, int tsh, bool active) {
    constexpr int Q = stage_bflies<RR, P>();
    const int m = L / RR;
    cx<R> x[Q][RR];
    if (active) {
#pragma unroll
        for (int p = 0; p < Q; ++p) {
            const int j = lt + p * TL;
            if (j < m) {
#pragma unroll
                for (int t = 0; t < RR; ++t) x[p][t] = buf[ln.idx(j + t * m)];
            }
        }
    }
    __syncthreads();
    if (!active) return;
    const int tstep = L / (Ns * RR);
    const bool pow2 = (Ns & (Ns - 1)) == 0;
#pragma unroll
    for (int p = 0; p < Q; ++p) {
        const int j = lt + p * TL;
        if (j < m) {
            const int k = pow2 ? (j & (Ns - 1)) : (j % Ns);
            if (k != 0) {
#pragma unroll
                for (int t = 1; t < RR; ++t) x[p][t] = cmul(x[p][t], twiddle(tw, k * t * tstep, sign, tsh));
            }
            Bfly<RR, R>::run(x[p], sign);
            const int ob = (j - k) * RR + k;
#pragma unroll
            for (int t = 0; t < RR; ++t) buf[ln.idx(ob + t * Ns)] = x[p][t];
        }
    }
}

// copy the twiddle table of plan p into LDS (all threads of the block),
// entry q at q + q/8: the strided reads of a stage (q = k t L/(Ns r)) then
// fall on different banks.  The copy takes L + L/8 entries.
template <class R>
__device__ __forceinline__ void stage_twiddles(cx<R> *dst, const LinePlan &p) {
    const cx<R> *src = reinterpret_cast<const cx<R> *>(p.tw);
    for (int i = threadIdx.x; i < p.L; i += blockDim.x) dst[i + (i >> 3)] = src[i];
}
constexpr int kTwLds = 3;    // twiddle shift of the padded LDS copy
constexpr int kTwMem = 31;   // twiddle shift of the table in memory
__host__ __device__ constexpr int tw_lds_len(int L) { return L + (L >> 3); }

// Full unnormalised transform (sign -1 forward).  Every thread of the block
// must call it (it synchronises); threads without a line pass active=false.
// IP: the middle stages run in place on buffer a (half the LDS, one more
// barrier per middle stage); otherwise they alternate between a and b.
// tw: the table exp(-2 pi i q / L), q in [0, L): the padded LDS copy made by
// stage_twiddles (tsh = kTwLds; the barrier at the start of fft_line
// publishes it) or the table in memory (tsh = kTwMem).
template <class R, int P, bool IP = false>
__device__ __forceinline__ void fft_line(cx<R> (&v)[P], const LdsLine<R> &ln, const LinePlan &p, const cx<R> *tw,
                                         int sign, int lt, bool active, int tsh) {
    static_assert(P == 8 || P == 4, "8 or 4 points per thread");
    const int L = p.L, TL = L / P;
    if (p.nst == 1) {   // L == P
        if (active) Bfly<P, R>::run(v, sign);
        return;
    }
    __syncthreads();   // previous readers of the LDS buffers are done
    // first stage: span 1, no twiddles, registers -> buffer a.  Butterfly j =
    // lt + q*TL (q < P/r) takes the thread's elements q + t*(P/r).
    const int r0 = p.rad[0];
    if (active) {
        if (P == 8 && r0 == 8) {
            Bfly<P, R>::run(v, sign);
#pragma unroll
            for (int t = 0; t < P; ++t) ln.a[ln.idx(lt * P + t)] = v[t];
        } else if (r0 == 4) {
#pragma unroll
            for (int q = 0; q < P / 4; ++q) {
                cx<R> x[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) x[t] = v[q + t * (P / 4)];
                Bfly<4, R>::run(x, sign);
                const int j = lt + q * TL;
#pragma unroll
                for (int t = 0; t < 4; ++t) ln.a[ln.idx(j * 4 + t)] = x[t];
            }
        } else {
#pragma unroll
            for (int q = 0; q < P / 2; ++q) {
                cx<R> x[2] = {v[q], v[q + P / 2]};
                Bfly<2, R>::run(x, sign);
                const int j = lt + q * TL;
#pragma unroll
                for (int t = 0; t < 2; ++t) ln.a[ln.idx(j * 2 + t)] = x[t];
            }
        }
    }
    __syncthreads();
    int Ns = r0;
    const cx<R> *in = ln.a;
    cx<R> *out = ln.b;
    if constexpr (IP) {
        for (int st = 1; st < p.nst - 1; ++st) {
            const int r = p.rad[st];
            // the lane index is made opaque per stage: otherwise every radix
            // case's lane-dependent addresses are hoisted out of the loop and
            // stay live together (tens of VGPRs, a lower occupancy)
            int lv = lt;
            asm volatile("" : "+v"(lv));
            switch (r) {
            case 8: mid_stage_ip<8, R, P>(ln.a, ln, L, Ns, TL, lv, sign, tw, tsh, active); break;
            case 4: mid_stage_ip<4, R, P>(ln.a, ln, L, Ns, TL, lv, sign, tw, tsh, active); break;
            case 2: mid_stage_ip<2, R, P>(ln.a, ln, L, Ns, TL, lv, sign, tw, tsh, active); break;
            // radix 3 (the hosts keep other primes out of in-place plans)
            default: mid_stage_ip<3, R, P>(ln.a, ln, L, Ns, TL, lv, sign, tw, tsh, active); break;
            }
            __syncthreads();
            Ns *= r;
        }
    }
    for (int st = 1; !IP && st < p.nst - 1; ++st) {
        const int r = p.rad[st];
        int lv = lt;   // opaque per stage, as above
        asm volatile("" : "+v"(lv));
        if (active) {
            switch (r) {
            case 8: mid_stage<8, R, P>(in, out, ln, L, Ns, TL, lv, sign, tw, tsh); break;
            case 4: mid_stage<4, R, P>(in, out, ln, L, Ns, TL, lv, sign, tw, tsh); break;
            case 2: mid_stage<2, R, P>(in, out, ln, L, Ns, TL, lv, sign, tw, tsh); break;
            case 3: mid_stage<3, R, P>(in, out, ln, L, Ns, TL, lv, sign, tw, tsh); break;
            default: mid_stage_gen<R>(in, out, ln, L, Ns, TL, lv, sign, tw, tsh, r); break;
            }
        }
        __syncthreads();
        const cx<R> *tmp = in;
        in = out;
        out = const_cast<cx<R> *>(tmp);
        Ns *= r;
    }
    // last stage: span L/r, buffer -> registers in the owned layout.
    // Butterfly j = lt + q*TL (q < P/r) yields the elements q + t*(P/r).
    const int rl = p.rad[p.nst - 1];
    if (!active) return;
    if (P == 8 && rl == 8) {
        const int j = lt;
        cx<R> x[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) x[t] = in[ln.idx(j + t * Ns)];
        if (j != 0) {
#pragma unroll
            for (int t = 1; t < 8; ++t) x[t] = cmul(x[t], twiddle(tw, j * t, sign, tsh));
        }
        Bfly<8, R>::run(x, sign);
#pragma unroll
        for (int t = 0; t < 8; ++t) v[t] = x[t];
    } else if (rl == 4) {
#pragma unroll
        for (int q = 0; q < P / 4; ++q) {
            const int j = lt + q * TL;
            cx<R> x[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) x[t] = in[ln.idx(j + t * Ns)];
            if (j != 0) {
#pragma unroll
                for (int t = 1; t < 4; ++t) x[t] = cmul(x[t], twiddle(tw, j * t, sign, tsh));
            }
            Bfly<4, R>::run(x, sign);
#pragma unroll
            for (int t = 0; t < 4; ++t) v[q + (P / 4) * t] = x[t];
        }
    } else {
#pragma unroll
        for (int q = 0; q < P / 2; ++q) {
            const int j = lt + q * TL;
            cx<R> x[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) x[t] = in[ln.idx(j + t * Ns)];
            if (j != 0) x[1] = cmul(x[1], twiddle(tw, j, sign, tsh));
            Bfly<2, R>::run(x, sign);
            v[q] = x[0];
            v[q + P / 2] = x[1];
        }
    }
}

// the 8-points-per-thread transform of the batched kernels
template <class R, bool IP = false>
__device__ __forceinline__ void fft8(cx<R> (&v)[8], const LdsLine<R> &ln, const LinePlan &p, const cx<R> *tw,
                                     int sign, int lt, bool active, int tsh) {
    fft_line<R, 8, IP>(v, ln, p, tw, sign, lt, active, tsh);
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

// N block sums at once (one LDS exchange); the sums land in v[] of thread 0,
// each in the order of block_sum
template <int N>
__device__ __forceinline__ void block_sum_n(double (&v)[N], double *red /* >= 16 N doubles */) {
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = wave_sum(v[i]);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) red[w * N + i] = v[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            double t = 0;
            for (int k = 0; k < nw; ++k) t += red[k * N + i];
            v[i] = t;
        }
    }
}

// Sum of n partials in a fixed order that does not depend on the block size:
// lane l of wave 0 adds partials l, l+64, ..., then a fixed shuffle tree.
// The result is broadcast to the whole block through *bcast.
__device__ __forceinline__ double fixed_sum(const double *p, int n, double *bcast) {
    if (threadIdx.x < 64) {
        double v = 0;
        for (int i = threadIdx.x; i < n; i += 64) v += p[i];
        v = wave_sum(v);
        if (threadIdx.x == 0) *bcast = v;
    }
    __syncthreads();
    const double r = *bcast;
    __syncthreads();
    return r;
}

// true when the previous order met |norm0 - accnorm| / norm0 < prec (real
// time, OCheby.cpp:179-182) or, with prec < 0, |accnorm - accnorm of the
// order before| / the latter < -prec (imaginary time, OCheby.cpp:283-285;
// QDng's first comparison, against 0, never passes)
__device__ __forceinline__ bool prev_converged(const ConvCheck &c, double *bcast, double *acc_out) {
    const bool rel = c.prec < 0;   // imaginary time (see ConvCheck)
    if (c.k_prev < (rel ? 3 : 2) || c.prec == 0) return false;   // prec 0: fixed-length series
    const double acc = fixed_sum(c.part, c.nb, bcast);
    *acc_out = acc;
    if (rel) {
        const double last = fixed_sum(c.norm0, c.nb, bcast);
        return fabs(acc - last) / last < -c.prec;
    }
    const double n0 = *c.norm0;
    return fabs(n0 - acc) / n0 < c.prec;
}

}  // namespace qdx
