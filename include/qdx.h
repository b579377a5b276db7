/*
 * qdx.h - C ABI of the MI355X grid-propagation library (libqdx.so).
 *
 * The library replaces, for one GPU, the part of QDng that propagates grid
 * wave functions: the Chebyshev (OCheby) and general split-operator (OGSPO)
 * propagators, the Hamiltonian they apply (OSum / OMultistate of
 * OGridNablaSq, OGridPotential, OGridDipole) and the per-step observables of
 * the Reporter.  Every entry point names the QDng interface it stands in for
 * (paths relative to the QDng source tree).
 *
 * Conventions
 *   - All functions return QDX_OK (0) or a negative qdx_status; the message of
 *     the last failure on the calling thread is qdx_last_error().
 *   - Host arrays are borrowed for the duration of the call only.  Complex
 *     arrays are interleaved (re, im) doubles in QDng's layout: dimension 0
 *     fastest (GridSystem.h:26), multistate vectors state-major [S][N]
 *     (WFMultistate.cpp:43-140).
 *   - The library owns all device memory.  Handles are destroyed explicitly.
 *   - One host thread per context at a time; each context owns its stream.
 */
#ifndef QDX_H
#define QDX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum qdx_status {
    QDX_OK = 0,
    QDX_ERR_PARAM = -1,       /* QDng EParamProblem (tools/Exception.h:115-146)  */
    QDX_ERR_INCOMPAT = -2,    /* QDng EIncompatible (tools/Exception.h:97-104)   */
    QDX_ERR_OVERFLOW = -3,    /* QDng EOverflow (tools/Exception.h:148-154)      */
    QDX_ERR_MEMORY = -4,      /* QDng EMemError (tools/Exception.h:76-84)        */
    QDX_ERR_DEVICE = -5,      /* HIP runtime failure                            */
    QDX_ERR_STATE = -6,       /* call order violated (Operator.h:48-58)          */
    QDX_ERR_IO = -7           /* QDng EIOError (tools/Exception.h:86-92)         */
} qdx_status;

typedef enum qdx_precision {
    QDX_FP64 = 0,        /* plain path: every array complex128 / float64        */
    QDX_FP32_BULK = 1,   /* FFT, recurrence and operator tables in FP32; Bessel
                            coefficients, reductions and the accumulated psi in
                            FP64                                                 */
    QDX_COMPACT = 2      /* FP64 arithmetic throughout; the Chebyshev terms
                            phi_k and the kinetic pass buffer of grids with
                            nd >= 2 (S <= 2) stored in 40 bits per component
                            (the upper 40 bits of the FP64 pattern, rounded:
                            relative error 2^-29), fixed-length series;
                            everything else as QDX_FP64                      */
} qdx_precision;

typedef struct qdx_ctx qdx_ctx;
typedef struct qdx_wf qdx_wf;
typedef struct qdx_ham qdx_ham;
typedef struct qdx_prop qdx_prop;

/* Grid geometry, QDng GridSystem (GridSystem.h:29-118, GridSystem.proto:6-17):
 * dx = (xmax - xmin) / (N - 1), both end points are grid points. */
typedef struct qdx_grid {
    int32_t ndims;          /* 1..4 (QDng's MAX_DIMS)                            */
    int32_t N[4];           /* multiples of 8 up to 4096 without prime factors
                               above 31 (QDng's FFTW takes any length)        */
    double xmin[4];
    double xmax[4];
} qdx_grid;

/* ---- library / context --------------------------------------------------- */
const char *qdx_last_error(void);
const char *qdx_version(void);
/* Bind a context (stream, plans, twiddle tables) to HIP device `device`. */
int qdx_ctx_create(int device, qdx_ctx **out);
int qdx_ctx_synchronize(qdx_ctx *ctx);
int qdx_ctx_destroy(qdx_ctx *ctx);

/* Per-launch kernel timing with HIP events on the context's stream.  QDng
 * itself only reports per-program wall/CPU seconds (CmdHandler.cpp:225-233);
 * these counters serve the roofline accounting of the benchmark.
 * qdx_ctx_profile(ctx, 1) clears the counters and starts recording,
 * qdx_ctx_profile(ctx, 0) stops; qdx_ctx_kernel_time sums one class. */
typedef enum qdx_kernel_class {
    QDX_K_CHEBY_ROW = 0,    /* Chebyshev order: IFFT_0 + fused recurrence + FFT_0 */
    QDX_K_LINE_PASS = 1,    /* strided FFT passes along dimensions >= 1          */
    QDX_K_CHEBY_INIT = 2,   /* step start: psi -> phi_0, FFT_0, sum |psi|^2        */
    QDX_K_GSPO_ROW = 3,     /* GSPO x-space split operators + FFT_0 / IFFT_0       */
    QDX_K_BATCH_ROW = 4,    /* batched jobs: row tiles of every job (one launch)  */
    QDX_K_BATCH_COL = 5,    /* batched jobs: dimension-1 kinetic pass, 2D jobs     */
    QDX_K_BATCH_LINE = 6,   /* batched jobs: whole time step of each 1D job        */
    QDX_NUM_KERNEL_CLASSES = 7
} qdx_kernel_class;
int qdx_ctx_profile(qdx_ctx *ctx, int32_t on);
int qdx_ctx_kernel_time(qdx_ctx *ctx, int32_t kernel_class, int64_t *launches, double *total_ms);

/* ---- wave functions (WaveFunction.h:34-314, WFGridCartesian.cpp:16-178,
 *      WFMultistate.cpp:43-300) ------------------------------------------- */
int qdx_wf_create(qdx_ctx *ctx, const qdx_grid *grid, int32_t nstates, qdx_wf **out);
/* state < 0: all states ([S][N] array); else one state ([N] array) */
int qdx_wf_upload(qdx_wf *wf, int32_t state, const double *c128);
int qdx_wf_download(qdx_wf *wf, int32_t state, double *c128);
int qdx_wf_copy(qdx_wf *dst, const qdx_wf *src);
/* QDng WaveFunction::Norm() = sqrt(<psi|psi>) with prod(dx) weights; per_state
 * (nullable, S entries) receives each state's squared norm (Reporter.cpp:115). */
int qdx_wf_norm(qdx_wf *wf, double *norm, double *per_state_pop);
/* <a|b> = sum conj(a) b prod(dx), summed over states (WFGridCartesian.cpp:62-116) */
int qdx_wf_overlap(qdx_wf *a, qdx_wf *b, double out[2]);
/* psi *= 1/Norm() (WaveFunction.h:103-106) */
int qdx_wf_normalize(qdx_wf *wf);
/* Post-step expectation filters of FiltersList (src/FiltersList.cpp):
 * per state s, <x_d> = sum_x x_d |psi_s(x)|^2 prod(dx) (OGridPosition,
 * x = xmin + n dx) and <p_d> = <psi_s| -i d/dx_d |psi_s> (OGridNabla as
 * momentum with Kspace::Init1Dddx, OGridNabla.cpp:76-95), the latter through
 * the transform along d on the device (sum_k k |psi~(k)|^2 / N_d).
 * per_state: S values. */
int qdx_wf_expec_position(qdx_wf *wf, int32_t dim, double *per_state);
int qdx_wf_expec_momentum(qdx_wf *wf, int32_t dim, double *per_state);
/* y += (re + i im) x (cVec MultElementsAdd; same grid and states) */
int qdx_wf_axpy(qdx_wf *y, double re, double im, const qdx_wf *x);
/* the grid and state count of a wave function (WaveFunction::Params,
 * WFMultistate::States) */
int qdx_wf_info(const qdx_wf *wf, qdx_grid *grid, int32_t *nstates);
int qdx_wf_destroy(qdx_wf *wf);

/* ---- Hamiltonian ----------------------------------------------------------
 * multistate = 0: QDng OSum of grid terms acting on one state
 *                 (OSum.cpp:28-60); nstates must be 1.
 * multistate = 1: QDng OMultistate (OMultistate.cpp:34-232): element (r,c) is
 *                 the sum of the terms added at (r,c); a lone off-diagonal
 *                 element is mirrored to (c,r); a missing diagonal acts as the
 *                 identity when unity != 0 (QDng default).                     */
int qdx_ham_create(qdx_ctx *ctx, const qdx_grid *grid, int32_t nstates, int32_t multistate,
                   int32_t unity, qdx_ham **out);
/* ctx may be NULL: a host-only Hamiltonian that supports qdx_ham_bounds and
 * qdx_cheby_params but no device operation. */
/* OGridNablaSq (OGridNablaSq.cpp:28-111): -1/(2 m_d) d^2/dx_d^2 on state s */
int qdx_ham_add_kinetic(qdx_ham *h, int32_t state, const double *mass_per_dim);
/* OGridPotential (OGridPotential.cpp:124-148): V(x) */
int qdx_ham_add_potential(qdx_ham *h, int32_t row, int32_t col, const double *V);
/* OGridDipole (OGridDipole.cpp:36-60) + Laser (Laser.cpp:125-128):
 * -mu(x) * laser[step] * scale */
int qdx_ham_add_dipole(qdx_ham *h, int32_t row, int32_t col, const double *mu,
                       const double *laser, int32_t nt, double scale);
/* OGobbler (OGobbler.cpp:32-204): Butterworth filter edges lp[d] / rp[d] (x
 * units, one per dimension, NaN = none) of the given order, times gain.
 * nip = 0: a real potential (bounds gain and 0); nip = 1: the negative
 * imaginary potential -i v(x) (bounds gain and -i gain), single-state Sum
 * Hamiltonians only, real-time Chebyshev and Arnoldi propagators (not GSPO,
 * SIL, imaginary time, backward propagation or the batch). */
int qdx_ham_add_gobbler(qdx_ham *h, int32_t row, int32_t col, const double *lp, const double *rp,
                        int32_t order, double gain, int32_t nip);
/* Emin / Emax with QDng's rules (OSum.cpp:28-51, OMultistate.cpp:163-204) */
int qdx_ham_bounds(qdx_ham *h, double *emin, double *emax);
/* the complex Emin / Emax (Operator::Emin / Emax, Operator.h:259-266) as
 * (re, im) pairs; imaginary parts come from negative imaginary potentials */
int qdx_ham_bounds_c(qdx_ham *h, double emin[2], double emax[2]);
/* the grid the Hamiltonian was created on */
int qdx_ham_grid(const qdx_ham *h, qdx_grid *out);
/* dst = H(step) src (Operator::Apply, Operator.h:276) */
int qdx_ham_apply(qdx_ham *h, qdx_wf *dst, qdx_wf *src, int64_t step);
/* Re <psi|H(step)|psi> (Operator::Expec, Operator.h:249-252) */
int qdx_ham_expec(qdx_ham *h, qdx_wf *psi, int64_t step, double *out);
int qdx_ham_destroy(qdx_ham *h);
/* <psi|H|psi> as qdx_ham_expec, without allocating: no counterpart in QDng.
 * It is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  H psi, the kinetic pass buffer and
 * the reduction partials live in one fixed device array per GPU (about
 * 16 MiB, looked up by qdx_ctx_create), which limits S * N to 524288 grid
 * points (QDX_ERR_MEMORY beyond).  Set-up first: the Hamiltonian must have
 * been applied once (qdx_ham_apply or qdx_ham_expec), which uploads its
 * arrays and builds the transform tables (QDX_ERR_STATE otherwise).  Every
 * call on one GPU shares that array: calls that overlap on one GPU (other
 * host threads, other contexts) race. */
int qdx_ham_expec_noalloc(qdx_ham *h, qdx_wf *psi, int64_t step, double *out);

/* ---- host-side parameter rules (no device needed) ------------------------ */
/* BesselJ0 (linalg/Bessel.cpp:22-58): J_k(x), k = 0..n-1, and the number of
 * trailing values zeroed by underflow. */
int qdx_bessel_j0(int32_t n, double x, double *out, int32_t *nz);
/* BesselI0 (linalg/Bessel.cpp:61-98, ZBESI kode 2): I_k(x) exp(-x), x >= 0 */
int qdx_bessel_i0(int32_t n, double x, double *out, int32_t *nz);

/* ---- propagators --------------------------------------------------------- */
typedef struct qdx_cheby_info {
    int32_t order;        /* series length after QDng's truncation              */
    int32_t nz;           /* Bessel underflow count removed (OCheby.cpp:451)   */
    double offset;        /* (Emax+Emin)/2   (OCheby.cpp:406-408)               */
    double scaling;       /* |Emax-Emin|/2   (OCheby.cpp:413-416)               */
    double prec;          /* break criterion (OCheby.cpp:179-182)               */
    double emin, emax;
    double offset_i;      /* Im(Emax+Emin) (OCheby.cpp:408): non-zero with a
                             negative imaginary potential (then R = 1.05
                             |Emax-Emin| and no break test)                     */
} qdx_cheby_info;

/* OCheby::Init (OCheby.cpp:358-493), real-time forward propagation.
 * order <= 0: QDng's automatic order; prec < 0: QDng default 1e-10;
 * scaling <= 0: automatic.  The Hamiltonian is borrowed and must outlive the
 * propagator. */
int qdx_cheby_create(qdx_ham *h, double dt, int32_t order, double prec, double scaling,
                     int32_t precision, qdx_prop **out);
/* OCheby in imaginary time, backward (exponent -dt), as ProgEigen sets it up
 * (ImaginaryTime() + Backward(), ProgEigen.cpp:691-694; OPropagator.h:82-108):
 * offset Emin - 0.05 |Emin|, scaling 1.05 |Emax - Emin|, coefficients from
 * I_k(R dt) exp(-R dt) (OCheby.cpp:406-446), the recurrence and relative
 * break test of imag_time_step (OCheby.cpp:200-297).  A step maps psi to
 * exp(-R dt) exp(-H dt) psi (the series' constant factor; normalise).  FP64
 * only: the series cancels by exp(-R dt), and rounding grows accordingly. */
int qdx_cheby_create_imag(qdx_ham *h, double dt, int32_t order, double prec, double scaling, qdx_prop **out);
int qdx_cheby_info_get(qdx_prop *p, qdx_cheby_info *info);
/* OSIL (OSIL.cpp:79-190): short iterative Lanczos with at most `order`
 * Lanczos vectors (<= 0: 20) and error bound err (<= 0: 1e-12), OSIL.h:11-12;
 * imag != 0: imaginary time, backward (exponent -dt), as ProgEigen sets it
 * up.  The Lanczos recursion runs on the device (H applications, scalar
 * products, vector updates), the tridiagonal matrix is diagonalised on the
 * host.  FP64. */
int qdx_sil_create(qdx_ham *h, double dt, int32_t order, double err, int32_t imag, qdx_prop **out);
/* OArnoldi (OArnoldi.cpp:64-190): Arnoldi with at most `order` Krylov vectors
 * (<= 0: 20), convergence bound conv (<= 0: 1e-12), OArnoldi.h:11-13;
 * imag as for qdx_sil_create.  Modified Gram-Schmidt on the device; the
 * Hessenberg matrix exponential exp(c H_A) e_0 on the host.  FP64. */
int qdx_arnoldi_create(qdx_ham *h, double dt, int32_t order, double conv, int32_t imag, qdx_prop **out);
/* Eigenstates by imaginary-time relaxation (ProgEigen::ImagTimeEF,
 * ProgEigen.cpp:318-412): for i < nef, psi = psi0; repeat at most max_steps
 * times while |1 - <psi_old|psi>| > conv: one step of U (from
 * qdx_cheby_create_imag, qdx_gspo_create_imag, or qdx_sil_create /
 * qdx_arnoldi_create with imag), remove the states found so far (OProjection, psi -= sum_j ef_j
 * <ef_j|psi>), normalise; the raw energy is <ef_i|H|ef_i> with H the given
 * Hamiltonian, or U's own when H is NULL (required for GSPO, whose split
 * operators form no Hamiltonian).  diag != 0 (QDng's default, DiagBasis,
 * ProgEigen.cpp:93-138): H is diagonalised in the found basis and the
 * returned states and energies (ascending) are those of the diagonalised
 * basis.  QDng's defaults: conv 1e-5 with diag, 1e-8 without; max_steps 512
 * (ProgEigen.h:15-18).
 * efs receives nef new wave functions (the caller destroys them); energies
 * nef values; energies_raw and steps (nullable) the relaxation's own. */
int qdx_eigen_imag(qdx_prop *U, qdx_ham *H, const qdx_wf *psi0, int32_t nef, int32_t max_steps, double conv,
                   int32_t diag, qdx_wf **efs, double *energies, double *energies_raw, int32_t *steps);
/* ProgEigen method "ac" (ProgEigen::AutoCorrEF, src/ProgEigen.cpp:414-600):
 * nsteps - 1 real-time steps of U from psi0 and c(t) = <psi(t)|psi0>; the Hann window (win != 0, math_functions.h:
 * 108-113), the forward FFT and PeakFinder on |FFT c| (tolerance tol, QDng
 * default 1; width > 0: the windowed search, else the region search,
 * tools/PeakFinder.cpp:83-170) give one energy E = index * 2 pi/(dt nsteps)
 * (index - nsteps above nsteps/2) per peak; each eigenfunction is
 * psi0 + sum_s psi(s dt) e^{i E s dt}, normalised, with the raw energy
 * <psi|H|psi>; diag != 0 then diagonalises H in that basis (DiagBasis).
 * *nef receives the number of peaks (QDX_ERR_PARAM when above max_ef; efs,
 * energies and energies_raw hold max_ef entries); autocorr (nullable,
 * 2*nsteps doubles) receives c(t) before the window.  QDng keeps every
 * step's function (WFBuffer); here the propagation runs a second time to add
 * each step into the peak functions, so the device holds n_ef + 2 wave
 * functions (the propagation is deterministic: both passes see the same
 * psi(t)). */
int qdx_eigen_ac(qdx_prop *U, qdx_ham *H, const qdx_wf *psi0, int32_t nsteps, int32_t win, double tol,
                 int32_t width, int32_t diag, int32_t max_ef, qdx_wf **efs, double *energies, double *energies_raw,
                 int32_t *nef, double *autocorr);
/* The scalars and coefficients OCheby::Init would choose for this Hamiltonian,
 * without allocating device memory (h may have been created with ctx NULL).
 * coeff (nullable) receives min(coeff_len, order) values. */
int qdx_cheby_params(qdx_ham *h, double dt, int32_t order, double prec, double scaling,
                     qdx_cheby_info *info, double *coeff, int32_t coeff_len);
/* Bessel-weighted coefficients c_k (order doubles, OCheby.cpp:476-480) */
int qdx_cheby_coeffs(qdx_prop *p, double *coeff);

/* OGSPO (OGSPO.cpp:30-139): split operators A, B, C, ... each a qdx_ham
 * holding one kind of term (kinetic, potential or dipole) per matrix element,
 * QDng's GridNablaSq / GridPotential / GridDipole / DMultistate. */
int qdx_gspo_create(qdx_ham *const *ops, int32_t nops, double dt, int32_t precision,
                    qdx_prop **out);
/* OGSPO in imaginary time, backward (exponent -dt, as ProgEigen sets it up):
 * every split operator becomes exp(-dt sf D) (OGSPO.cpp:30-47 with
 * ODSpace::InitExponential, ODSpace.h:65-69) -- real decays, the coupled
 * 2 x 2 potential through cosh / sinh.  FP64. */
int qdx_gspo_create_imag(qdx_ham *const *ops, int32_t nops, double dt, qdx_prop **out);

typedef struct qdx_step_stats {
    int64_t n_hpsi;        /* H applications executed (Chebyshev: 1 + recurrence
                              iterations actually run; GSPO: steps)              */
    double norm2;          /* <psi|psi> after the last step                      */
    double nconv;          /* Chebyshev: |norm0 - accnorm| / norm0 of the last
                              executed order of the last step (OCheby.cpp:179)  */
} qdx_step_stats;

/* OPropagator::Apply(psi) at clock step `step` (OCheby.cpp:306-322,
 * OGSPO.cpp:114-139). */
int qdx_prop_apply(qdx_prop *p, qdx_wf *psi, int64_t step, qdx_step_stats *stats);

/* ProgPropa::Run step loop (ProgPropa.cpp:317-326) for steps
 * first_step .. first_step+nsteps-1 without host round trips.  Before each
 * step the Reporter observables are recorded (Reporter.cpp:65-158):
 *   norm2[i]     = <psi|psi>            (nullable)
 *   autocorr[2i] = <psi(t)|psi0>        (nullable; needs psi0)
 * n_hpsi_per_step (nullable) receives the H-application count of each step. */
int qdx_prop_run(qdx_prop *p, qdx_wf *psi, int64_t first_step, int64_t nsteps, qdx_wf *psi0,
                 double *norm2, double *autocorr, int32_t *n_hpsi_per_step,
                 qdx_step_stats *stats);
/* OPropagator::Forward() / Backward() (src/qdlib/OPropagator.h:62-100) of a
 * real-time propagator: backward = 1 gives the exponent +i dt (the operators
 * conjugated, as QDng's Backward sets them); imaginary-time propagators run
 * backward (exponent -dt) only and reject forward with QDX_ERR_INCOMPAT. */
int qdx_prop_set_direction(qdx_prop *p, int32_t backward);
/* OPropagator::Exponent() (src/qdlib/OPropagator.h:148-155): c of
 * exp(c H) for one step, as (re, im). */
int qdx_prop_exponent(qdx_prop *p, double out[2]);
int qdx_prop_destroy(qdx_prop *p);

/* ---- optimal control (ProgOCT, src/ProgOCT.cpp) ----------------------------
 * The laser samples of a Hamiltonian's dipole terms are QDng's Laser objects
 * (Laser.cpp:125-136): the field applied at clock step s is laser[s] * scale.
 * qdx_ham_laser_get reads the raw samples of the first dipole term (n values,
 * n <= its length); qdx_ham_laser_set writes them into every dipole term
 * (the identical lasers ProgOCT optimises, ProgOCT.cpp:925-950), n at most
 * the shortest term's length.  A propagator created on the Hamiltonian sees
 * the new samples from its next step; its Chebyshev scaling stays the one of
 * its creation (OCheby::Init runs once, as in ProgOCT::Run). */
int qdx_ham_laser_len(qdx_ham *h, int64_t *n);
int qdx_ham_laser_get(qdx_ham *h, double *laser, int64_t n);
int qdx_ham_laser_set(qdx_ham *h, const double *laser, int64_t n);

/* OProjection (OProjection.cpp:93-125): P psi = sign sum_k <wf_k|psi> wf_k
 * (sign -1 when positive = 0), or psi - that when inv != 0.  The functions
 * are copied; a projection onto one state of a multistate function is a set
 * of multistate functions that are zero in the other states (QDng's
 * OMultistate of projections with unity off, as in OCT operator targets). */
typedef struct qdx_proj qdx_proj;
int qdx_proj_create(qdx_wf *const *wfs, int32_t n, int32_t positive, int32_t inv, qdx_proj **out);
int qdx_proj_apply(qdx_proj *p, qdx_wf *dst, qdx_wf *src);
/* Re <psi|P|psi> (Operator::Expec, Operator.h:249-252) */
int qdx_proj_expec(qdx_proj *p, qdx_wf *psi, double *out);
int qdx_proj_destroy(qdx_proj *p);

typedef enum qdx_oct_method {
    QDX_OCT_KROTOV = 0,      /* ProgOCT::Iterate, new field = old - correction   */
    QDX_OCT_RABITZ = 1,      /* ProgOCT::Iterate, new field = -correction        */
    QDX_OCT_RABITZFB = 2     /* ProgOCT::IterateRabitzFB (Zhu and Rabitz 1998)   */
} qdx_oct_method;

typedef struct qdx_oct_opts {
    int32_t method;          /* qdx_oct_method                                  */
    int32_t operator_target; /* 0: overlap targets psi_t, 1: operator targets   */
    int32_t phase;           /* phase-sensitive objective                       */
    int32_t ntargets;        /* 1..16 (MAX_TARGETS, ProgOCT.h:19)               */
    int32_t iterations;      /* maximum iterations (QDng default 50)            */
    int32_t nsteps;          /* clock steps of one propagation                   */
    double conv;             /* convergence (QDng default 1e-5)                  */
    double alpha;            /* field restriction, != 0                          */
} qdx_oct_opts;

/* Columns of one report row (ProgOCT::Report, ProgOCT.cpp:363-425): the
 * iteration, the laser objective (dt * sum of the update terms), the
 * objective (yield - laser objective), the pulse energy of the applied field
 * (Laser::PulseEnergy, Laser.cpp:174-184) and the returned yield (|sum of
 * |overlap|^2|, |sum of overlaps| when phase-sensitive, sum |<O>|), then per
 * target: norm of the propagated initial state, norm of the target state,
 * |overlap|^2, phase of the overlap / pi, <O> (NaN where QDng prints
 * nothing; the iteration-0 row has NaN laser objective and objective). */
#define QDX_OCT_ROW_FIXED 5
#define QDX_OCT_ROW_PER_TARGET 5

/* ProgOCT::Run (ProgOCT.cpp:868-1092) over the propagator U, whose
 * Hamiltonian's dipole terms (or, for GSPO, its split operators' dipole
 * terms) carry the field being optimised.  The coupling operator is the sum
 * of those dipole arrays (single state) or the multistate operator of the
 * off-diagonal ones, with unity off (ChainLoader.h:48-160, ProgOCT.cpp:
 * 954-961).  psi_i: ntargets initial states; psi_t: ntargets target states
 * (operator_target = 0); targets: ntargets projections (operator_target =
 * 1); shape: the field's shape function (at least nsteps samples); gobbler
 * (nullable): applied to every propagated state after each step, which is
 * then normalised.  Device memory: nsteps + 1 wave functions per target (the
 * backward-propagation buffer).  On return the Hamiltonian's lasers hold the
 * optimised field (QDng writes it as the final laser file); report
 * (nullable) receives (iterations + 1) rows, rows_done the rows written,
 * converged whether the change of the yield fell to conv. */
int qdx_oct_run(qdx_prop *U, const qdx_oct_opts *opts, const double *shape, int64_t nshape,
                qdx_wf *const *psi_i, qdx_wf *const *psi_t, qdx_proj *const *targets, qdx_ham *gobbler,
                double *report, int32_t *rows_done, int32_t *converged);

/* ---- batched jobs ---------------------------------------------------------
 * Many independent propagations (1D or 2D grids of different sizes, each
 * with its own potential, dipole and laser field): single-state Chebyshev
 * jobs and GSPO jobs on one or two coupled states, kept resident in HBM and
 * advanced together.  One tick is one time step of every active job
 * (OCheby::real_time_step, OCheby.cpp:92-197; OGSPO::Apply,
 * OGSPO.cpp:114-139); each pass is one launch over the tiles of all jobs.  QDng runs one
 * propagation per program (ProgPropa::Run, ProgPropa.cpp:317-326); a job here
 * is exactly that propagation with its own series length and break test, so
 * its result does not depend on the other jobs in the batch.
 *
 * Device memory of the jobs comes from one arena of `hbm_budget` bytes
 * allocated by qdx_batch_create; qdx_batch_submit fails with QDX_ERR_MEMORY
 * (nothing allocated) when the job does not fit, so callers can queue it.
 * Grids: 1D or 2D, each N <= 4096 and of the form 8 * 2^a * 3^b (other
 * prime factors are rejected with QDX_ERR_INCOMPAT).  One block
 * runs a 1D job's whole step with the line held in registers. */
typedef struct qdx_batch qdx_batch;
typedef struct qdx_job_desc {
    qdx_ham *ham;          /* Chebyshev job: one state, kinetic, potential and
                              dipole terms (OSum of OGridNablaSq,
                              OGridPotential, OGridDipole)                     */
    qdx_ham *const *ops;   /* GSPO job (ham NULL): split operators {T, W}, T
                              kinetic terms only, W potentials on 1 or 2
                              states (+ a dipole on one state); OGSPO with
                              operators A = T, B = W (OGSPO.cpp:30-139)        */
    int32_t nops;          /* 2 for a GSPO job, 0 otherwise                    */
    int32_t nsteps;
    double dt;
    int32_t order;         /* Chebyshev: 0 = OCheby's automatic order          */
    int32_t first_step;    /* clock step of the job's first step (its laser
                              sample index); 0 for a new job.  A snapshot
                              (qdx_batch_fetch after k steps) resubmitted with
                              first_step = k and the remaining steps continues
                              the propagation bit for bit (ProgPropa's restart
                              from a written wave function, `start`)       */
    double prec;           /* Chebyshev: < 0 = OCheby's default 1e-10          */
    const double *psi0;    /* complex128 initial state [S][N]                  */
} qdx_job_desc;

int qdx_batch_create(qdx_ctx *ctx, int32_t precision, int64_t hbm_budget, qdx_batch **out);
/* device bytes the job would take from the arena */
int qdx_batch_footprint(qdx_batch *b, const qdx_job_desc *job, int64_t *bytes);
/* copy the job's data to the device; it starts at the next tick */
int qdx_batch_submit(qdx_batch *b, const qdx_job_desc *job, int64_t *id);
/* enqueue `nticks` ticks (asynchronous; qdx_ctx_synchronize waits) */
int qdx_batch_run(qdx_batch *b, int32_t nticks);
/* steps still to run (0: finished once the enqueued ticks complete), and,
 * after a synchronisation, H applications (GSPO: steps) and steps executed */
int qdx_batch_job_state(qdx_batch *b, int64_t id, int64_t *steps_left, int64_t *n_hpsi, int64_t *steps_done);
/* wait for the enqueued ticks and copy the job's psi (complex128, [S][N]) */
int qdx_batch_fetch(qdx_batch *b, int64_t id, double *psi);
/* Per-job failure isolation (Reporter.cpp:112-113, "Norm is not a number"):
 * *failed_at = -1 for a sound job, else the step whose state had a norm that
 * is not a number (Chebyshev: the state the step started from; GSPO: the
 * state it produced).  A failed job stops there: its kernels are skipped,
 * steps_done and n_hpsi stay at the last sound step, and its batch-mates
 * are unaffected. */
int qdx_batch_job_status(qdx_batch *b, int64_t id, int64_t *failed_at);
/* wait for the enqueued ticks and return the job's memory to the arena */
int qdx_batch_release(qdx_batch *b, int64_t id);
/* jobs held, arena bytes in use, ticks enqueued so far */
int qdx_batch_usage(qdx_batch *b, int32_t *jobs, int64_t *bytes_used, int64_t *ticks);
/* Tick replay through hipGraphs (default on).  The second tick of one shape
 * (same job mix, same active job classes, same longest series) is captured
 * into a graph and later ticks of that shape replay it; the kernels read the
 * tick from a device counter, so one graph serves every such tick.  A new
 * job mix (submit, release) starts new graphs.  Off: every tick is launched
 * kernel by kernel.  Results are the same either way.  No counterpart in
 * QDng (its propagators run on the host). */
int qdx_batch_set_graphs(qdx_batch *b, int32_t enable);
/* graphs captured and graph replays so far (either may be NULL) */
int qdx_batch_graph_stats(qdx_batch *b, int64_t *captured, int64_t *replayed);

/* ---- schedules and the autotuner ----------------------------------------
 * QDng's counterpart is FFTW's measured planning and its wisdom file
 * (src/fft/FFTGlobal.cpp:41-80: "fftwOptimize" selects FFTW_PATIENT, plans
 * are read from "wisdom" and written to "wisdomsave"; src/fft/fft.cpp:95-110):
 * the autotuner measures schedules on the device, the tuning file plays the
 * part of the wisdom file. */

/* Radix plan of the line transforms of length L in one precision on a
 * context: the sequence of radices (product L; first and last 8, 4 or 2;
 * middle 8, 4, 3, 2 or a prime 5 .. 31; at most 16).  A plan set here is used by the
 * propagators and Hamiltonians created and the batch jobs submitted
 * afterwards on this context (a batch job keeps the plans it was submitted
 * with).  Different plans give the same transform up to rounding. */
int qdx_ctx_set_radix_plan(qdx_ctx *ctx, int32_t precision, int32_t L, const int32_t *radix, int32_t nst);
/* the plan in use for length L (radix: room for 16) */
int qdx_ctx_get_radix_plan(qdx_ctx *ctx, int32_t precision, int32_t L, int32_t *radix, int32_t *nst);

/* Points per thread of the Chebyshev propagators' line kernels on grids with
 * two or more dimensions (the row kernel and the kinetic passes along
 * dimensions >= 1), for steps taken afterwards: 8, 4, or 0 (the default) to
 * choose 4 where the lines fit (states * N0 / 4 <= 512 threads, N_d <= 2048)
 * and, for the FP64 kernels, the grid gives fewer than two waves per SIMD
 * at 8.  At 4 a line
 * uses the context's plan for its length when that plan's first and last
 * radix are 4 or 2, else first and last radix 4.  Both give the same
 * transform up to rounding.  Single-state 1D lines run a whole step in one
 * block, with 4 points per thread up to 1024 points and 8 above, unless 8
 * is set.  No
 * QDng counterpart: FFTW lays out its own threads (src/fft/fft.cpp:95-110). */
int qdx_ctx_set_line_points(qdx_ctx *ctx, int32_t points);
int qdx_ctx_get_line_points(qdx_ctx *ctx, int32_t *points);

/* Tile shape of the batched kernels for 2D Chebyshev jobs of one grid
 * shape.  Row tiles: row_lines lines of dimension 0 (N0/8 threads each) in a
 * block of row_threads (256 or 512) threads.  Column tiles: col_width
 * columns (N1/8 threads each, col_width * N1/8 <= 256, col_width divides
 * N0). */
typedef struct qdx_tiling {
    int32_t row_threads;
    int32_t row_lines;
    int32_t col_width;
} qdx_tiling;
/* the tiling jobs of shape n0 x n1 get when submitted */
int qdx_batch_get_tiling(qdx_batch *b, int32_t n0, int32_t n1, qdx_tiling *t);
/* set it for jobs of that shape submitted afterwards (checked: thread and
 * LDS limits of the kernels) */
int qdx_batch_set_tiling(qdx_batch *b, int32_t n0, int32_t n1, const qdx_tiling *t);

/* Autotuner for 2D jobs of shape n0 x n1 in the batch's precision.  Runs
 * synthetic fixed-length jobs of that shape (about 2 M grid points) in a
 * context and batch of its own on the same device, and times the row and
 * column kernels with events on their dispatch packets: first the radix
 * plans of N0 and N1 (up to 12 with no more stages than the default plan,
 * fewest first, and 4 with one stage more), then the column widths (row
 * tiles keep the default block size: another one would give the shape a
 * launch of its own).  The winners are
 * installed: the plans on the batch's context (qdx_ctx_set_radix_plan), the
 * tiling on the batch (qdx_batch_set_tiling), for jobs submitted afterwards.
 * Times are per tick of the synthetic jobs, in microseconds. */
typedef struct qdx_tune_result {
    qdx_tiling tiling;
    int32_t radix0[16], nst0;      /* plan of N0 */
    int32_t radix1[16], nst1;      /* plan of N1 */
    double row_us_default, row_us; /* row kernels: default schedule, chosen */
    double col_us_default, col_us; /* column kernels */
    int32_t candidates;            /* schedules timed */
} qdx_tune_result;
int qdx_batch_autotune(qdx_batch *b, int32_t n0, int32_t n1, qdx_tune_result *out);
/* tune each new 2D job shape when its first job is submitted (default off) */
int qdx_batch_set_autotune(qdx_batch *b, int32_t enable);
/* LDS schedule of the batch's 2D Chebyshev row and column tiles: LDS
 * buffers per line, 2 (the FFT's middle stages alternate between two
 * buffers, one barrier per stage) or 1 (in place: half the LDS, one more
 * barrier per middle stage).  Results are bit-identical either way.
 * Default: 1 for rows and columns. */
int qdx_batch_set_lds_schedule(qdx_batch *b, int32_t row_bufs, int32_t col_bufs);
int qdx_batch_get_lds_schedule(qdx_batch *b, int32_t *row_bufs, int32_t *col_bufs);
/* Time the four LDS schedules on the batch's current jobs (three ticks each;
 * they are ordinary steps of the jobs, which advance by 12 steps) and keep
 * the fastest.  ms_per_tick (nullable, 4 entries) receives the tick times of
 * (rows, columns) = (2,2), (1,2), (2,1), (1,1).  QDX_ERR_STATE when no job
 * is active, an active job has fewer than 12 steps left (every schedule is
 * timed on the same work) or the context is profiling. */
int qdx_batch_autotune_lds(qdx_batch *b, double *ms_per_tick);
/* Write / read the batch's tilings, its LDS schedule and its context's
 * radix plans of the batch's precision as text ("tile N0 N1 row_threads
 * row_lines col_width", "lds row_bufs col_bufs", "plan L r1 r2 ..."), so a
 * later process can skip the tuning.  A file with a malformed or invalid
 * line is rejected whole. */
int qdx_batch_tuning_save(qdx_batch *b, const char *path);
int qdx_batch_tuning_load(qdx_batch *b, const char *path);

/* Continuous batching over a stream of requests.  Request i arrives at
 * arrival_s[i] seconds after the call starts (non-decreasing).  Arrived
 * requests are admitted in arrival order as soon as the arena has room for
 * them (and fewer than max_active jobs are held, if max_active > 0); every
 * tick advances all admitted jobs by one step, so the work of jobs admitted
 * at different times overlaps on the device; a job leaves after its last
 * step and its memory is reused.  psi_out (nullable, entries nullable)
 * receives each final psi (complex128).  Returns when all requests are done;
 * a request larger than the whole budget fails with QDX_ERR_MEMORY. */
typedef struct qdx_serve_result {
    double admit_s;        /* admission, seconds after the start (host clock)  */
    double finish_s;       /* end of the job's last step (device clock)        */
    int64_t n_hpsi;        /* H applications (GSPO: steps)                      */
    int64_t failed_step;   /* -1, or as qdx_batch_job_status                     */
} qdx_serve_result;
int qdx_batch_serve(qdx_batch *b, const qdx_job_desc *jobs, const double *arrival_s, int32_t n,
                    int32_t max_active, double *const *psi_out, qdx_serve_result *out);
int qdx_batch_destroy(qdx_batch *b);

#ifdef __cplusplus
}
#endif
#endif
