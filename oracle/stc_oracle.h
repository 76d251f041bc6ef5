/*
 * stc_oracle.h -- CPU oracle for the ROM frequency sweep.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (staccato_amd/, include/)
 * includes, links or loads this file or the library built from it.  It is the
 * checker the HIP path is compared against, used only by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg.
 *
 * It restates, in plain C99 complex arithmetic, what the reference does per
 * excitation frequency:
 *   src/analysis/KrylovROMSubstructure.cpp:1403-1431  (assemble, zgetrf, zgetrs)
 *   srcPerformanceTest/main.cpp:375-402             (OpenMP loop over frequencies)
 *   srcPerformanceTest/solver/assembly.cpp:8-12      (A = K - f^2 M, no damping)
 * with LAPACK zgetf2/zgetrs semantics (partial pivoting by izamax on
 * cabs1 = |re|+|im|, first index on ties, 1-based ipiv, full-row swaps,
 * unit-lower L and U stored in place).
 *
 * Layout: column-major, interleaved complex128 ({double re, im}, the layout of
 * MKL_Complex16 / STACCATOComplexDouble, src/tools/AuxiliaryParameters.h:52).
 */
#ifndef STC_ORACLE_H
#define STC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* A = (1 + i*eta) K + i*omega D - omega^2 M  (n x n, lda = n).
 * K, M are n x n complex; D may be NULL (no damping).
 * Follows KrylovROMSubstructure.cpp:1412-1420 (K, then -w^2 M, then i w D). */
void orc_assemble(int n, const double* K, const double* M, const double* D,
                  double eta, double omega, double* A);

/* LAPACK zgetf2 semantics.  Returns info (0, or first zero pivot, 1-based). */
int orc_zgetrf(int n, double* A, int lda, int* ipiv);

/* LAPACK zgetrs semantics, trans = 'N', 'T' or 'C'. */
void orc_zgetrs(char trans, int n, int nrhs, const double* A, int lda,
                const int* ipiv, double* B, int ldb);

/* Single-precision counterparts (LAPACK cgetf2/cgetrs), interleaved complex64. */
int orc_cgetrf(int n, float* A, int lda, int* ipiv);
void orc_cgetrs(char trans, int n, int nrhs, const float* A, int lda,
                const int* ipiv, float* B, int ldb);

/* Normwise backward error of one column:
 *   ||f - A x||_inf / (||A||_inf ||x||_inf + ||f||_inf),  |z| = modulus. */
double orc_backward_error(int n, const double* A, int lda, const double* x,
                          const double* f);

/* The reference sweep: for every frequency f_j (Hz, omega = 2 pi f_j)
 * assemble A, zgetrf, zgetrs with nrhs columns of F (ldf), write X_j (n x nrhs,
 * ldx) at X + j*ldx*nrhs, info[j] = zgetrf info.  berr (nullable) receives
 * nfreq*nrhs backward errors.  nthreads > 1 runs the frequency loop with
 * OpenMP (srcPerformanceTest/main.cpp:365-375); returns wall seconds. */
double orc_sweep(int n, const double* K, const double* M, const double* D,
                 double eta, int nfreq, const double* freq_hz, int nrhs,
                 const double* F, int ldf, double* X, int ldx, int* info,
                 double* berr, int nthreads);

/* Mixed precision restatement: cgetrf on A rounded to complex64, cgetrs,
 * then FP64 iterative refinement r = F - A x (A in FP64), A d = r in FP32,
 * x += d, until every column's backward error <= tol or max_iter steps.
 * iters[j] (nullable) = refinement steps taken, info[j] = cgetrf info or -1
 * when not converged. */
double orc_sweep_refined(int n, const double* K, const double* M, const double* D,
                         double eta, int nfreq, const double* freq_hz, int nrhs,
                         const double* F, int ldf, double* X, int ldx, int* info,
                         int* iters, double* berr, double tol, int max_iter,
                         int nthreads);

/* FP64 adjoint of J = sum_w sum_r |c^H X_r(w)|^2 with respect to the modal
 * stiffness scales s_k (dA/ds_k = (1+i eta) kdiag_k phi_k phi_k^T) and to eta
 * (dA/deta = i K).  Phi is n x p real column-major (ldphi).  Per frequency:
 * zgetrf, zgetrs('N') for X, g = c^H X, mu = A^{-H} (c g) by zgetrs('C'),
 * dJ/dtheta += -2 Re(mu^H dA/dtheta X).  Returns info count of singular w. */
int orc_adjoint(int n, const double* K, const double* M, const double* D,
                double eta, int nfreq, const double* freq_hz, int nrhs,
                const double* F, int ldf, const double* c, int p,
                const double* Phi, int ldphi, const double* kdiag,
                double* J, double* dJ_ds, double* dJ_deta, int nthreads);

/* CPU baseline with an optimised LAPACK: the same loop as orc_sweep but with
 * zgetrf_/zgetrs_ taken from a shared library opened at run time (dlopen),
 * symbol names given by the caller (e.g. the scipy-bundled OpenBLAS exports
 * scipy_zgetrf_ / scipy_zgetrs_).  Returns wall seconds, or -1 when the
 * library or symbols are missing. */
double orc_sweep_lapack(const char* libpath, const char* getrf_sym,
                        const char* getrs_sym, const char* setthreads_sym,
                        int n, const double* K, const double* M, const double* D,
                        double eta, int nfreq, const double* freq_hz, int nrhs,
                        const double* F, int ldf, double* X, int ldx, int* info,
                        int nthreads);

#ifdef __cplusplus
}
#endif
#endif
