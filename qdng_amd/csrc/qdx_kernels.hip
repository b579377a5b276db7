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
//
// Translation units: qdx_cheby.hip (Chebyshev init, row and line kernels),
// qdx_lines.hip (strided line passes, H applied to a vector), qdx_gspo.hip
// (GSPO), this file (reductions, conversions, the allocation-free entry
// point's device scratch); shared pieces in qdx_kernels_common.hpp.
#include "qdx_kernels_common.hpp"

namespace qdx {

thread_local LaunchTiming g_launch_timing;

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

template <class T>
__global__ void k_conj(T *__restrict__ d, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        d[2 * i + 1] = -d[2 * i + 1];
}

void launch_conj(hipStream_t s, void *data, long n, bool fp32) {
    const int threads = 256;
    const int nb = (int)std::min<long>((n + threads - 1) / threads, 4096);
    if (nb < 1) return;
    if (fp32) QDX_LAUNCH((k_conj<float>), dim3(nb), dim3(threads), 0, s, static_cast<float *>(data), n);
    else QDX_LAUNCH((k_conj<double>), dim3(nb), dim3(threads), 0, s, static_cast<double *>(data), n);
    QDX_HIP(hipGetLastError());
}

void launch_sum_partials(hipStream_t s, const double *partials, int nbx, int S, double *out) {
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

// ---------------------------------------------------------------------------
// Device scratch of the allocation-free entry point.  Host code never passes
// the array's name to a kernel or a copy: it uses the address that
// hipGetSymbolAddress returns for the current device, kept per device.
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

void *noalloc_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = ptr;
    }
    return g_scratch_ptr[device];
}

void *noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

void launch_overlap(hipStream_t s, const double *a, const double *b, long n, int S, double *partials, double *out) {
    overlap_impl<double>(s, a, b, n, S, partials, out);
}

void launch_overlap_c64(hipStream_t s, const float *a, const float *b, long n, int S, double *partials,
                        double *out) {
    overlap_impl<float>(s, a, b, n, S, partials, out);
}

}  // namespace qdx
