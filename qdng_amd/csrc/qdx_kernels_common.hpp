// Shared pieces of the gfx950 line kernels (included by qdx_kernels.hip,
// qdx_cheby.hip, qdx_lines.hip and qdx_gspo.hip): the row-block geometry, the
// launch-shape helpers and the state-count dispatch.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <stdexcept>
#include <string>

#include "qdx_device.hpp"
#include "qdx_impl.hpp"

namespace qdx {

// ---------------------------------------------------------------------------
// helpers

// kinetic value T_i(k) at k-space indices k[0 .. nd-1], summed over
// dimensions in dimension order as Diff::_InitKspace does (Diff.cpp:120-139)
template <class R>
__device__ __forceinline__ double kin_sum(const HamDev<R> &h, int st, const GridDev &g, const int (&k)[kMaxDims]) {
    const double *const *tk = h.tk[st];
    double t = 0.0 + tk[0][k[0]];
#pragma unroll
    for (int d = 1; d < kMaxDims; ++d)
        if (d < g.nd) t += tk[d][k[d]];
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
    const int nother = col_nother(g, dim);
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

#define QDX_DISPATCH_S(S_, CALL)                                              \
    switch (S_) {                                                             \
    case 1: { constexpr int SS = 1; CALL; } break;                            \
    case 2: { constexpr int SS = 2; CALL; } break;                            \
    case 3: { constexpr int SS = 3; CALL; } break;                            \
    case 4: { constexpr int SS = 4; CALL; } break;                            \
    default: throw std::runtime_error("unsupported number of states");        \
    }

}  // namespace qdx
