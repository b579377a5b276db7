/*
 * FP64 CPU oracle for the QDng grid-propagation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * it may be loaded by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg, never by the product library (qdng_amd/).
 *
 * It restates, in plain C++17 and double precision, what QDng computes on
 * its CPU path (file:line citations are relative to the QDng source tree):
 *   - grid geometry           src/qdlib/GridSystem.cpp:207-211
 *   - kinetic operator        src/qdlib/OGridNablaSq.cpp:28-111, Diff.cpp:52-140,203-220,
 *                             Kspace.cpp:22-50, Kspace.h:20-23
 *   - FFT convention          src/fft/fft.cpp:91-119,250-285 (FFTW, unnormalised,
 *                             forward e^{-i}, dimension 0 fastest)
 *   - potential / dipole      src/qdlib/OGridPotential.cpp:124-148, OGridDipole.cpp:36-60
 *   - laser value per step    src/qdlib/Laser.cpp:125-128
 *   - sums / multistate       src/qdlib/OSum.cpp:28-60, OMultistate.cpp:99-232
 *   - Chebyshev propagator    src/qdlib/OCheby.cpp:64-197,358-493
 *   - GSPO propagator         src/qdlib/OGSPO.cpp:30-139, ODMultistate.cpp:109-183,
 *                             TransformMSD.cpp:31-101
 *   - observables             src/qdlib/WFGridCartesian.cpp:62-116, WFMultistate.cpp:158-226,
 *                             src/Reporter.cpp:65-158
 *
 * The C ABI below exists so that Python tests can drive the oracle through
 * ctypes.  Complex vectors are interleaved (re, im) doubles; a multistate
 * vector is state-major [S][N] with dimension 0 fastest inside a state.
 */
#ifndef QDX_ORACLE_H
#define QDX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- plumbing -------------------------------------------------------- */
const char *orc_last_error(void);
int orc_set_threads(int n);
int orc_set_fft_backend(int which);   /* 0 own FFT, 1 FFTW3 API (MKL) if it loads */
/* test hook: 1 transposed GSPO eigenvectors, 2 +i v NIP, 3 coefficient phase
 * from Re(offset) only, 4 NIP high pass multiplied */
int orc_plant_mistake(int code);
/* real-time propagator direction: kind 0 Cheby, 1 SIL, 2 Arnoldi, 3 GSPO */
int orc_set_backward(int kind, int id, int backward);

/* ---- FFT (dimension 0 fastest, unnormalised, sign -1 forward) -------- */
int orc_fft(int ndims, const int32_t *N, double *data, int sign);

/* ---- Bessel J_k(x), k = 0..n-1, plus the emulated underflow count ---- */
int orc_bessel_j(int n, double x, double *out, int *nz);
/* I_k(x) exp(-x), k = 0..n-1 (Bessel.cpp:61-98, ZBESI kode = 2) */
int orc_bessel_i_scaled(int n, double x, double *out, int *nz);

/* ---- grid ------------------------------------------------------------ */
int orc_grid(int ndims, const int32_t *N, const double *xmin, const double *xmax);

/* ---- operator tree (ids are small integers, freed by orc_reset) ------ */
void orc_reset(void);
int orc_op_kinetic(int grid, const double *mass);            /* GridNablaSq  */
int orc_op_potential(int grid, const double *V);              /* GridPotential */
int orc_op_dipole(int grid, const double *mu, const double *laser, int nt,
                  double scale);                               /* GridDipole   */
/* the raw samples of a GridDipole's field (Laser::Set, Laser.cpp:133-136) */
int orc_op_set_laser(int op, const double *laser, int nt);
int orc_op_sum(const int *ops, int n);                         /* Sum          */
/* Gobbler: lp, rp hold one edge per dimension, NaN = none; nip = 1: -i v */
int orc_op_gobbler(int grid, const double *lp, const double *rp, int order, double gain, int nip);
/* Multistate (dmulti = 0) or DMultistate (dmulti = 1). rows/cols/ops give the
 * matrix elements; unity: missing diagonal acts as identity (QDng default). */
int orc_op_multistate(int nstates, const int *rows, const int *cols,
                      const int *ops, int n, int unity, int dmulti);

int orc_op_bounds(int op, double *emin, double *emax);
/* complex Emin / Emax: (re, im) pairs */
int orc_op_bounds_c(int op, double *emin, double *emax);
/* dest = O(step) src ; vectors hold nstates*size complex values */
int orc_op_apply(int op, int step, const double *src, double *dest);
/* Re <psi|O|psi> with grid weights (QDng Operator::Expec) */
int orc_op_expec(int op, int step, const double *psi, int nstates, double *out);

/* ---- observables ----------------------------------------------------- */
/* <a|b> with prod dx weights, summed over states */
int orc_overlap(int grid, int nstates, const double *a, const double *b, double out[2]);
/* <x> (dim d) and <p> (dim d) expectations, single state */
int orc_expec_position(int grid, int dim, const double *psi, double *out);
int orc_expec_momentum(int grid, int dim, const double *psi, double *out);

/* ---- Chebyshev ------------------------------------------------------- */
typedef struct orc_cheby_info {
    int order;          /* series length after truncation */
    int nz;             /* Bessel underflow count removed */
    double offset;      /* Re part of the energy offset   */
    double scaling;     /* R = |Emax-Emin|/2               */
    double prec;
    double emin, emax;
} orc_cheby_info;

/* order <= 0: automatic; prec < 0: QDng default 1e-10; scaling <= 0: automatic */
int orc_cheby_create(int ham, int nstates, double dt, int order, double prec,
                     double scaling, orc_cheby_info *info);
/* imaginary time, backward (exponent -dt): OCheby with ImaginaryTime() +
 * Backward() as ProgEigen sets it up (ProgEigen.cpp:691-694) */
int orc_cheby_create_imag(int ham, int nstates, double dt, int order, double prec,
                          double scaling, orc_cheby_info *info);
int orc_cheby_coeffs(int prop, double *coeff /* order doubles */);
/* one step at clock step `step` (real or imaginary time, as created);
 * returns number of H applications */
int orc_cheby_step(int prop, int step, double *psi, int *n_hpsi);

/* ---- short iterative Lanczos (OSIL.cpp) ------------------------------- */
/* order <= 0: 20; err <= 0: 1e-12 (OSIL.h:11-12); imag: imaginary time,
 * backward (exponent -dt) */
int orc_sil_create(int ham, int nstates, double dt, int order, double err, int imag);
int orc_sil_step(int prop, int step, double *psi, int *n_hpsi);

/* ---- Arnoldi (OArnoldi.cpp) -------------------------------------------- */
/* order <= 0: 20; conv <= 0: 1e-12 (OArnoldi.h:11-13) */
int orc_arnoldi_create(int ham, int nstates, double dt, int order, double conv, int imag);
int orc_arnoldi_step(int prop, int step, double *psi, int *n_hpsi);

/* ---- GSPO ------------------------------------------------------------ */
int orc_gspo_create(const int *ops, int nops, int nstates, double dt);
/* imag: imaginary time, backward (exponent -dt) */
int orc_gspo_create_ex(const int *ops, int nops, int nstates, double dt, int imag);
int orc_gspo_step(int prop, int step, double *psi);

#ifdef __cplusplus
}
#endif
#endif
