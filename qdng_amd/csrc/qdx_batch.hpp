// Device-side tables of the batched-jobs runtime: many independent
// single-state Chebyshev propagations (1D or 2D grids of different sizes,
// each with its own potential, dipole and laser field) advanced together, one
// time step per tick, every pass of the series covering all active jobs in a
// single launch.  Shared by qdx_batch.cpp (host) and qdx_batch.hip (kernels).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "qdx_impl.hpp"

namespace qdx {

// One job as the kernels see it.  H = T + V(x) - E(t) mu(x) on an nd = 1 or
// 2 grid; the Chebyshev scalars follow OCheby::Init / real_time_step
// (OCheby.cpp:92-197, 358-493) for this job alone.
template <class R>
struct BJob {
    int kind;               // 0: Chebyshev (one state), 1: GSPO (S = 1 or 2)
    int S;                  // states
    int nd;                 // 1 or 2
    int N0, N1;             // N1 = 1 when nd == 1
    int plan0, plan1;       // indices into the plan table
    int rb;                 // rows (lines of dimension 0) per row tile
    int rt0, nrt;           // first row tile of the job, row tiles of the job
    int W;                  // columns per column tile (nd == 2)
    int order;              // series length (OCheby order after truncation)
    int start_tick;         // tick of the job's step 0
    int nsteps;
    int nt;                 // laser samples (0: no dipole term)
    long size;              // N0 * N1
    double inv_n;           // 1 / size
    double inv_n0, inv_n1;  // 1 / N0, 1 / N1: T = T_0 + T_1 acts as 1D filters
    double offset;          // (Emax + Emin) / 2
    double inv_scaling;     // 1 / R
    double prec;            // break threshold
    double lscale;          // dipole scale
    const double *coef;     // (re, im): coeff[k] * exp(-i offset dt), k < order
    const double *tk0;      // kinetic table of dimension 0 (FP64), nullptr: no T
    const double *tk1;      // dimension 1
    const R *V;             // potential, nullptr: none
    const R *mu;            // dipole moment, nullptr: none
    const double *laser;    // field per step
    double *psi;            // complex128 [size]
    R *phiA, *phiB;         // phi_k for even / odd k (complex R)
    R *X;                   // FFT_0 scratch (nd == 2)
    double *part;           // 3 * nrt partial sums: [k % 2], [2] = step start
    ChebyCtl *ctl;
    long long *stats;       // [0] H applications, [1] steps done
    // GSPO (OGSPO.cpp:114-139 with split operators T and W): one step is
    // exp(-i dt/2 W) exp(-i dt T) exp(-i dt/2 W); W = [[V, V01], [V01, V11]]
    // (minus E(t) mu on a single state); psi is held as complex R in psiR.
    double tau;             // dt
    R *psiR;                // complex R [S][size]
    const R *V11, *V01;     // second diagonal element and coupling (nullptr: none)
    const double *tk0s1;    // kinetic tables of state 1
    const double *tk1s1;
};

// a row tile: rows row0 .. row0 + rb - 1 of job `job`; a column tile:
// columns c0 .. c0 + W - 1 of state first / N0 (c0 = first % N0)
struct BTile {
    int job;
    int first;
};

// Launchers (qdx_batch.hip).  `tick` is a device counter advanced by
// batch_tick_end, so a tick's launch sequence does not change from tick to
// tick and can be replayed from a captured graph.
template <class R>
struct BatchLaunch {
    // init (phase 0), prologue (phase 1) or order k >= 2 (phase 2) over the
    // row tiles [0, ntiles); NT = 256 or 512 threads per block; lbuf LDS
    // buffers per line (2: ping-pong, 1: in place)
    // (cmp: phi and X of the 2D jobs in the compact 40-bit term storage,
    // R double)
    static void row(hipStream_t s, int NT, int lbuf, int phase, int k, const BJob<R> *jobs, const BTile *tiles,
                    int ntiles, const LinePlan *plans, const int *tick, size_t lds, bool tw_lds, bool cmp = false);
    // dimension-1 part of the kinetic operator of the 2D jobs (into X)
    static void col(hipStream_t s, int lbuf, int k, const BJob<R> *jobs, const BTile *tiles, int ntiles,
                    const LinePlan *plans, const int *tick, size_t lds, bool tw_lds, bool cmp = false);
    // one whole time step of each 1D job in `list` (one block per job)
    static void line(hipStream_t s, int NT, const BJob<R> *jobs, const int *list, int n, const LinePlan *plans,
                     const int *tick, size_t lds, bool tw_lds);
    // GSPO jobs: phase 0 = exp(-i tau/2 W) then FFT_0 into X (2D) or the
    // whole step (1D); phase 1 = IFFT_0 of X then exp(-i tau/2 W) (2D)
    static void gspo_row(hipStream_t s, int NT, int phase, const BJob<R> *jobs, const BTile *tiles, int ntiles,
                         const LinePlan *plans, const int *tick, size_t lds, bool tw_lds);
    // GSPO kinetic split operator along dimension 1 of the 2D jobs' X:
    // forward * exp(-i tau T)/N * inverse
    static void gspo_col(hipStream_t s, const BJob<R> *jobs, const BTile *tiles, int ntiles, const LinePlan *plans,
                         const int *tick, size_t lds, bool tw_lds);
    // per-job bookkeeping at the end of a tick; advances *tick
    static void tick_end(hipStream_t s, const BJob<R> *jobs, int njobs, int *tick);
};

}  // namespace qdx
