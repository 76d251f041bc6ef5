/*
 * stc_oracle.c -- CPU oracle for the ROM frequency sweep (test infrastructure).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * this code, and only as the checker / the timed CPU baseline.  The product
 * path never calls it.  Interface and citations: stc_oracle.h.
 *
 * Every routine is plain C99 complex arithmetic written from the LAPACK
 * definitions (zgetf2, zlaswp, ztrsm, zgetrs) and from the reference loop
 * src/analysis/KrylovROMSubstructure.cpp:1403-1431.
 */
#include "stc_oracle.h"

#include <complex.h>
#include <dlfcn.h>
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex zc;
typedef float complex cc;

static double wall_now(void) {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

/* ---- assembly: KrylovROMSubstructure.cpp:1412-1420 ------------------------ */
void orc_assemble(int n, const double* K, const double* M, const double* D,
                  double eta, double omega, double* A) {
    const zc* k = (const zc*)K;
    const zc* m = (const zc*)M;
    const zc* d = (const zc*)D;
    zc* a = (zc*)A;
    const zc keta = 1.0 + I * eta;
    const double w2 = omega * omega;
    const size_t nn = (size_t)n * n;
    for (size_t e = 0; e < nn; ++e) {
        zc v = keta * k[e] - w2 * m[e];
        if (d) v += (I * omega) * d[e];
        a[e] = v;
    }
}

/* ---- LAPACK izamax: first index of max cabs1 --------------------------------- */
static double cabs1(zc z) { return fabs(creal(z)) + fabs(cimag(z)); }
static float cabs1f(cc z) { return fabsf(crealf(z)) + fabsf(cimagf(z)); }

/* ---- LAPACK zgetf2 ----------------------------------------------------------- */
int orc_zgetrf(int n, double* Ap, int lda, int* ipiv) {
    zc* A = (zc*)Ap;
    int info = 0;
    const double sfmin = DBL_MIN;
#define a(i, j) A[(size_t)(j) * lda + (i)]
    for (int j = 0; j < n; ++j) {
        int jp = j;
        double best = cabs1(a(j, j));
        for (int i = j + 1; i < n; ++i) {
            double v = cabs1(a(i, j));
            if (v > best) { best = v; jp = i; }
        }
        ipiv[j] = jp + 1;
        if (a(jp, j) != 0.0) {
            if (jp != j)
                for (int c = 0; c < n; ++c) { zc t = a(j, c); a(j, c) = a(jp, c); a(jp, c) = t; }
            zc piv = a(j, j);
            if (cabs(piv) >= sfmin) {
                zc r = 1.0 / piv;
                for (int i = j + 1; i < n; ++i) a(i, j) *= r;
            } else {
                for (int i = j + 1; i < n; ++i) a(i, j) /= piv;
            }
        } else if (info == 0) {
            info = j + 1;
        }
        for (int c = j + 1; c < n; ++c) {
            zc u = a(j, c);
            if (u == 0.0) continue;
            for (int i = j + 1; i < n; ++i) a(i, c) -= a(i, j) * u;
        }
    }
#undef a
    return info;
}

int orc_cgetrf(int n, float* Ap, int lda, int* ipiv) {
    cc* A = (cc*)Ap;
    int info = 0;
    const float sfmin = FLT_MIN;
#define a(i, j) A[(size_t)(j) * lda + (i)]
    for (int j = 0; j < n; ++j) {
        int jp = j;
        float best = cabs1f(a(j, j));
        for (int i = j + 1; i < n; ++i) {
            float v = cabs1f(a(i, j));
            if (v > best) { best = v; jp = i; }
        }
        ipiv[j] = jp + 1;
        if (a(jp, j) != 0.0f) {
            if (jp != j)
                for (int c = 0; c < n; ++c) { cc t = a(j, c); a(j, c) = a(jp, c); a(jp, c) = t; }
            cc piv = a(j, j);
            if (cabsf(piv) >= sfmin) {
                cc r = 1.0f / piv;
                for (int i = j + 1; i < n; ++i) a(i, j) *= r;
            } else {
                for (int i = j + 1; i < n; ++i) a(i, j) /= piv;
            }
        } else if (info == 0) {
            info = j + 1;
        }
        for (int c = j + 1; c < n; ++c) {
            cc u = a(j, c);
            if (u == 0.0f) continue;
            for (int i = j + 1; i < n; ++i) a(i, c) -= a(i, j) * u;
        }
    }
#undef a
    return info;
}

/* ---- LAPACK zgetrs ------------------------------------------------------------
 * 'N': B := P B (zlaswp forward), L y = B (unit), U x = y.
 * 'T'/'C': A = P^T L U, so A^T = U^T L^T P:  U^T y = B, L^T z = y,
 *          B := P^T z (zlaswp backward); 'C' conjugates the factors. */
void orc_zgetrs(char trans, int n, int nrhs, const double* Ap, int lda,
                const int* ipiv, double* Bp, int ldb) {
    const zc* A = (const zc*)Ap;
    zc* B = (zc*)Bp;
#define a(i, j) A[(size_t)(j) * lda + (i)]
    for (int r = 0; r < nrhs; ++r) {
        zc* b = B + (size_t)r * ldb;
        if (trans == 'N' || trans == 'n') {
            for (int i = 0; i < n; ++i) {
                int p = ipiv[i] - 1;
                if (p != i) { zc t = b[i]; b[i] = b[p]; b[p] = t; }
            }
            for (int k = 0; k < n; ++k) {
                zc x = b[k];
                if (x == 0.0) continue;
                for (int i = k + 1; i < n; ++i) b[i] -= x * a(i, k);
            }
            for (int k = n - 1; k >= 0; --k) {
                if (b[k] == 0.0) continue;
                b[k] /= a(k, k);
                zc x = b[k];
                for (int i = 0; i < k; ++i) b[i] -= x * a(i, k);
            }
        } else {
            const int cj = (trans == 'C' || trans == 'c');
            /* U^T y = b : forward, row k of U^T is column k of U */
            for (int k = 0; k < n; ++k) {
                zc s = b[k];
                for (int i = 0; i < k; ++i) {
                    zc u = cj ? conj(a(i, k)) : a(i, k);
                    s -= u * b[i];
                }
                zc d = cj ? conj(a(k, k)) : a(k, k);
                b[k] = s / d;
            }
            /* L^T z = y : backward, unit diagonal */
            for (int k = n - 1; k >= 0; --k) {
                zc s = b[k];
                for (int i = k + 1; i < n; ++i) {
                    zc l = cj ? conj(a(i, k)) : a(i, k);
                    s -= l * b[i];
                }
                b[k] = s;
            }
            for (int i = n - 1; i >= 0; --i) {
                int p = ipiv[i] - 1;
                if (p != i) { zc t = b[i]; b[i] = b[p]; b[p] = t; }
            }
        }
    }
#undef a
}

void orc_cgetrs(char trans, int n, int nrhs, const float* Ap, int lda,
                const int* ipiv, float* Bp, int ldb) {
    const cc* A = (const cc*)Ap;
    cc* B = (cc*)Bp;
#define a(i, j) A[(size_t)(j) * lda + (i)]
    for (int r = 0; r < nrhs; ++r) {
        cc* b = B + (size_t)r * ldb;
        if (trans == 'N' || trans == 'n') {
            for (int i = 0; i < n; ++i) {
                int p = ipiv[i] - 1;
                if (p != i) { cc t = b[i]; b[i] = b[p]; b[p] = t; }
            }
            for (int k = 0; k < n; ++k) {
                cc x = b[k];
                if (x == 0.0f) continue;
                for (int i = k + 1; i < n; ++i) b[i] -= x * a(i, k);
            }
            for (int k = n - 1; k >= 0; --k) {
                if (b[k] == 0.0f) continue;
                b[k] /= a(k, k);
                cc x = b[k];
                for (int i = 0; i < k; ++i) b[i] -= x * a(i, k);
            }
        } else {
            const int cj = (trans == 'C' || trans == 'c');
            for (int k = 0; k < n; ++k) {
                cc s = b[k];
                for (int i = 0; i < k; ++i) {
                    cc u = cj ? conjf(a(i, k)) : a(i, k);
                    s -= u * b[i];
                }
                cc d = cj ? conjf(a(k, k)) : a(k, k);
                b[k] = s / d;
            }
            for (int k = n - 1; k >= 0; --k) {
                cc s = b[k];
                for (int i = k + 1; i < n; ++i) {
                    cc l = cj ? conjf(a(i, k)) : a(i, k);
                    s -= l * b[i];
                }
                b[k] = s;
            }
            for (int i = n - 1; i >= 0; --i) {
                int p = ipiv[i] - 1;
                if (p != i) { cc t = b[i]; b[i] = b[p]; b[p] = t; }
            }
        }
    }
#undef a
}

/* ---- normwise backward error --------------------------------------------------- */
double orc_backward_error(int n, const double* Ap, int lda, const double* xp,
                          const double* fp) {
    const zc* A = (const zc*)Ap;
    const zc* x = (const zc*)xp;
    const zc* f = (const zc*)fp;
    double rmax = 0.0, anorm = 0.0, xmax = 0.0, fmax = 0.0;
    for (int i = 0; i < n; ++i) {
        zc s = f[i];
        double rs = 0.0;
        for (int j = 0; j < n; ++j) {
            zc aij = A[(size_t)j * lda + i];
            s -= aij * x[j];
            rs += cabs(aij);
        }
        if (cabs(s) > rmax) rmax = cabs(s);
        if (rs > anorm) anorm = rs;
        if (cabs(x[i]) > xmax) xmax = cabs(x[i]);
        if (cabs(f[i]) > fmax) fmax = cabs(f[i]);
    }
    double den = anorm * xmax + fmax;
    if (den == 0.0) return rmax == 0.0 ? 0.0 : INFINITY;
    return rmax / den;
}

/* ---- the reference loop (KrylovROMSubstructure.cpp:1403-1431) ------------------ */
double orc_sweep(int n, const double* K, const double* M, const double* D,
                 double eta, int nfreq, const double* freq_hz, int nrhs,
                 const double* F, int ldf, double* X, int ldx, int* info,
                 double* berr, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    double t0 = wall_now();
#pragma omp parallel num_threads(nthreads)
    {
        double* A = (double*)malloc(sizeof(zc) * (size_t)n * n);
        double* A0 = berr ? (double*)malloc(sizeof(zc) * (size_t)n * n) : NULL;
        int* ipiv = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
#pragma omp for schedule(dynamic, 1)
        for (int j = 0; j < nfreq; ++j) {
            double omega = 2.0 * M_PI * freq_hz[j];
            orc_assemble(n, K, M, D, eta, omega, A);
            if (A0) memcpy(A0, A, sizeof(zc) * (size_t)n * n);
            zc* xj = (zc*)X + (size_t)j * ldx * nrhs;
            for (int r = 0; r < nrhs; ++r)
                memcpy(xj + (size_t)r * ldx, (const zc*)F + (size_t)r * ldf, sizeof(zc) * n);
            int inf = orc_zgetrf(n, A, n, ipiv);
            info[j] = inf;
            if (inf == 0) orc_zgetrs('N', n, nrhs, A, n, ipiv, (double*)xj, ldx);
            if (berr)
                for (int r = 0; r < nrhs; ++r)
                    berr[(size_t)j * nrhs + r] =
                        inf ? INFINITY
                            : orc_backward_error(n, A0, n, (const double*)(xj + (size_t)r * ldx),
                                                 (const double*)((const zc*)F + (size_t)r * ldf));
        }
        free(A); free(A0); free(ipiv);
    }
    return wall_now() - t0;
}

/* ---- FP32 factor + FP64 refinement (no reference counterpart; the north-star
 *      mixed-precision variant of the same loop) ------------------------------ */
static void residual64(int n, const zc* A, const zc* x, const zc* f, zc* r) {
    for (int i = 0; i < n; ++i) r[i] = f[i];
    for (int j = 0; j < n; ++j) {
        zc xj = x[j];
        for (int i = 0; i < n; ++i) r[i] -= A[(size_t)j * n + i] * xj;
    }
}

double orc_sweep_refined(int n, const double* K, const double* M, const double* D,
                         double eta, int nfreq, const double* freq_hz, int nrhs,
                         const double* F, int ldf, double* X, int ldx, int* info,
                         int* iters, double* berr, double tol, int max_iter,
                         int nthreads) {
    if (nthreads < 1) nthreads = 1;
    double t0 = wall_now();
#pragma omp parallel num_threads(nthreads)
    {
        const size_t nn = (size_t)n * n;
        zc* A = (zc*)malloc(sizeof(zc) * nn);
        cc* A32 = (cc*)malloc(sizeof(cc) * nn);
        cc* v32 = (cc*)malloc(sizeof(cc) * (size_t)(n > 0 ? n : 1));
        zc* r = (zc*)malloc(sizeof(zc) * (size_t)(n > 0 ? n : 1));
        int* ipiv = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
#pragma omp for schedule(dynamic, 1)
        for (int j = 0; j < nfreq; ++j) {
            double omega = 2.0 * M_PI * freq_hz[j];
            orc_assemble(n, K, M, D, eta, omega, (double*)A);
            for (size_t e = 0; e < nn; ++e) A32[e] = (cc)A[e];
            int inf = orc_cgetrf(n, (float*)A32, n, ipiv);
            info[j] = inf;
            int used = 0;
            zc* xj = (zc*)X + (size_t)j * ldx * nrhs;
            for (int c = 0; c < nrhs; ++c) {
                const zc* f = (const zc*)F + (size_t)c * ldf;
                zc* x = xj + (size_t)c * ldx;
                double be = INFINITY;
                if (inf) {
                    for (int i = 0; i < n; ++i) x[i] = 0.0;
                } else {
                    for (int i = 0; i < n; ++i) v32[i] = (cc)f[i];
                    orc_cgetrs('N', n, 1, (float*)A32, n, ipiv, (float*)v32, n);
                    for (int i = 0; i < n; ++i) x[i] = (zc)v32[i];
                    int it = 0;
                    for (;;) {
                        be = orc_backward_error(n, (const double*)A, n, (const double*)x,
                                                (const double*)f);
                        if (be <= tol || it >= max_iter) break;
                        residual64(n, A, x, f, r);
                        double s = 0.0;
                        for (int i = 0; i < n; ++i) if (cabs1(r[i]) > s) s = cabs1(r[i]);
                        if (s == 0.0) break;
                        for (int i = 0; i < n; ++i) v32[i] = (cc)(r[i] / s);
                        orc_cgetrs('N', n, 1, (float*)A32, n, ipiv, (float*)v32, n);
                        for (int i = 0; i < n; ++i) x[i] += s * (zc)v32[i];
                        ++it;
                    }
                    if (it > used) used = it;
                    if (be > tol) info[j] = -1;
                }
                if (berr) berr[(size_t)j * nrhs + c] = be;
            }
            if (iters) iters[j] = used;
        }
        free(A); free(A32); free(v32); free(r); free(ipiv);
    }
    return wall_now() - t0;
}

/* ---- FP64 adjoint ---------------------------------------------------------------- */
int orc_adjoint(int n, const double* K, const double* M, const double* D,
                double eta, int nfreq, const double* freq_hz, int nrhs,
                const double* F, int ldf, const double* cvec, int p,
                const double* Phi, int ldphi, const double* kdiag,
                double* J, double* dJ_ds, double* dJ_deta, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    const zc* c = (const zc*)cvec;
    const zc* k = (const zc*)K;
    double Jt = 0.0, deta = 0.0;
    int nsing = 0;
    double* ds = (double*)calloc((size_t)(p > 0 ? p : 1), sizeof(double));
#pragma omp parallel num_threads(nthreads)
    {
        const size_t nn = (size_t)n * n;
        double* A = (double*)malloc(sizeof(zc) * nn);
        zc* x = (zc*)malloc(sizeof(zc) * (size_t)(n > 0 ? n : 1));
        zc* mu = (zc*)malloc(sizeof(zc) * (size_t)(n > 0 ? n : 1));
        int* ipiv = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
        double* dsl = (double*)calloc((size_t)(p > 0 ? p : 1), sizeof(double));
        double Jl = 0.0, detal = 0.0;
        int nsl = 0;
#pragma omp for schedule(dynamic, 1)
        for (int j = 0; j < nfreq; ++j) {
            double omega = 2.0 * M_PI * freq_hz[j];
            orc_assemble(n, K, M, D, eta, omega, A);
            if (orc_zgetrf(n, A, n, ipiv)) { ++nsl; continue; }
            for (int r = 0; r < nrhs; ++r) {
                memcpy(x, (const zc*)F + (size_t)r * ldf, sizeof(zc) * n);
                orc_zgetrs('N', n, 1, A, n, ipiv, (double*)x, n);
                zc g = 0.0;
                for (int i = 0; i < n; ++i) g += conj(c[i]) * x[i];
                Jl += creal(g) * creal(g) + cimag(g) * cimag(g);
                for (int i = 0; i < n; ++i) mu[i] = c[i] * g;
                orc_zgetrs('C', n, 1, A, n, ipiv, (double*)mu, n);
                /* dJ/deta = -2 Re(mu^H (i K) x) */
                zc s = 0.0;
                for (int i = 0; i < n; ++i) {
                    zc kx = 0.0;
                    for (int jj = 0; jj < n; ++jj) kx += k[(size_t)jj * n + i] * x[jj];
                    s += conj(mu[i]) * kx;
                }
                detal += -2.0 * creal(I * s);
                /* dJ/ds_k = -2 Re((1+i eta) kdiag_k conj(phi_k^T mu) (phi_k^T x)) */
                for (int q = 0; q < p; ++q) {
                    const double* ph = Phi + (size_t)q * ldphi;
                    zc pm = 0.0, px = 0.0;
                    for (int i = 0; i < n; ++i) { pm += ph[i] * mu[i]; px += ph[i] * x[i]; }
                    dsl[q] += -2.0 * creal((1.0 + I * eta) * kdiag[q] * conj(pm) * px);
                }
            }
        }
#pragma omp critical
        {
            Jt += Jl; deta += detal; nsing += nsl;
            for (int q = 0; q < p; ++q) ds[q] += dsl[q];
        }
        free(A); free(x); free(mu); free(ipiv); free(dsl);
    }
    *J = Jt;
    *dJ_deta = deta;
    for (int q = 0; q < p; ++q) dJ_ds[q] = ds[q];
    free(ds);
    return nsing;
}

/* ---- CPU baseline with a run-time-loaded LAPACK ----------------------------------- */
typedef void (*getrf_fn)(const int*, const int*, void*, const int*, int*, int*);
typedef void (*getrs_fn)(const char*, const int*, const int*, const void*, const int*,
                         const int*, void*, const int*, int*, size_t);
typedef void (*setthreads_fn)(int);

double orc_sweep_lapack(const char* libpath, const char* getrf_sym,
                        const char* getrs_sym, const char* setthreads_sym,
                        int n, const double* K, const double* M, const double* D,
                        double eta, int nfreq, const double* freq_hz, int nrhs,
                        const double* F, int ldf, double* X, int ldx, int* info,
                        int nthreads) {
    void* lib = dlopen(libpath, RTLD_NOW | RTLD_LOCAL);
    if (!lib) return -1.0;
    getrf_fn getrf = (getrf_fn)dlsym(lib, getrf_sym);
    getrs_fn getrs = (getrs_fn)dlsym(lib, getrs_sym);
    if (!getrf || !getrs) return -1.0;
    if (setthreads_sym) {
        setthreads_fn st = (setthreads_fn)dlsym(lib, setthreads_sym);
        if (st) st(1); /* one LAPACK thread per OpenMP thread */
    }
    if (nthreads < 1) nthreads = 1;
    double t0 = wall_now();
#pragma omp parallel num_threads(nthreads)
    {
        double* A = (double*)malloc(sizeof(zc) * (size_t)n * n);
        int* ipiv = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
#pragma omp for schedule(dynamic, 1)
        for (int j = 0; j < nfreq; ++j) {
            double omega = 2.0 * M_PI * freq_hz[j];
            orc_assemble(n, K, M, D, eta, omega, A);
            zc* xj = (zc*)X + (size_t)j * ldx * nrhs;
            for (int r = 0; r < nrhs; ++r)
                memcpy(xj + (size_t)r * ldx, (const zc*)F + (size_t)r * ldf, sizeof(zc) * n);
            int inf = 0;
            getrf(&n, &n, A, &n, ipiv, &inf);
            info[j] = inf;
            if (inf == 0) {
                int inf2 = 0;
                getrs("N", &n, &nrhs, A, &n, ipiv, xj, &ldx, &inf2, 1);
            }
        }
        free(A); free(ipiv);
    }
    return wall_now() - t0;
}
