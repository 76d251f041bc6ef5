/*
 * staccato_sweep.h -- C ABI of the MI355X frequency-sweep library
 * (libstaccato_sweep.so, built from staccato_amd/csrc/).
 *
 * For every excitation frequency f (omega = 2 pi f) the library forms the
 * reduced dynamic stiffness
 *
 *     A(omega) = (1 + i eta) K + i omega D - omega^2 M
 *
 * from the ROM operators K, M, D and solves A(omega) X = F by LU with partial
 * pivoting, entirely on one AMD Instinct MI355X (gfx950).  A(omega) is
 * assembled inside the factorisation kernel: small systems are factored in
 * LDS; larger ones in a per-workgroup factor slot of device memory that is
 * reused point after point (its traffic reaches HBM once the resident slots
 * exceed the caches), so no buffer per frequency is ever allocated.
 *
 * What this replaces in the reference (DrStS/STACCATO):
 *   - the body of the frequency loop of KrylovROMSubstructure::backTransformKMOR
 *     (src/analysis/KrylovROMSubstructure.cpp:1403-1431: three cblas_zaxpy
 *     assembly calls, LAPACKE_zgetrf, LAPACKE_zgetrs),
 *   - the dense-mode frequency loop of the benchmark
 *     (srcPerformanceTest/main.cpp:375-402 with
 *      srcPerformanceTest/solver/assembly.cpp:8-12),
 *   - the cuBLAS getrfBatched / cuSOLVER getrf tests
 *     (test/krylovMidsizeTest/gpu/concurrent/batchedLU/src/main.cu:248-259,
 *      test/krylovTestGPU/matrixInSeparateLoop/gpu/src/main.cu:205-222),
 *     which factor without pivoting; this library always pivots.
 *
 * Conventions (as in the reference):
 *   - column-major, leading dimensions given explicitly;
 *   - complex numbers interleaved {double re, im}: stc_z64 is bit-identical
 *     to MKL_Complex16 = STACCATOComplexDouble
 *     (src/tools/AuxiliaryParameters.h:52);
 *   - pivots are 1-based 32-bit ints with LAPACK zgetrf meaning
 *     (LP64 lapack_int, cmake/modules/FindMKL.cmake:57-72);
 *   - per-frequency info: 0 = solved, k > 0 = U(k,k) is exactly zero
 *     (LAPACK zgetrf info), -1 = iterative refinement did not reach tol,
 *     -2 = the point could not be run (internal; X is NaN and the call
 *     returns STC_EDEVICE).
 *     The reference ignores zgetrf's info (KrylovROMSubstructure.cpp:1428);
 *     this library reports it and never calls exit().
 *
 * Threading: a handle owns one device, one HIP stream and its device buffers.
 * Use one handle per host thread; calls on one handle are not thread safe.
 * Calls taking host pointers are synchronous.  The *_dev calls take device
 * pointers and a stream (a hipStream_t passed as void*, NULL = the handle's
 * stream) and are asynchronous with respect to the host.  Launches of one
 * handle share its factor slots, so each waits (on the device) for the
 * handle's previous launch when it goes to a different stream.  A *_dev call
 * that needs larger factor slots than any earlier call on the handle waits on
 * the host for the previous launch and reallocates them.
 */
#ifndef STACCATO_SWEEP_H
#define STACCATO_SWEEP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } stc_z64; /* == MKL_Complex16 */
typedef struct { float re, im; } stc_c32;  /* == MKL_Complex8  */

typedef struct stc_ctx* stc_handle;
typedef struct stc_plan_s* stc_plan;

/* status codes returned by every call */
enum {
    STC_OK = 0,
    STC_EINVAL = -1,   /* bad argument (size, pointer, rom id, mode)      */
    STC_ENOMEM = -2,   /* device allocation failed                        */
    STC_EDEVICE = -3,  /* HIP runtime error                               */
    STC_ENOROM = -4,   /* rom_id not set on this handle                   */
    STC_EBUSY = -5,    /* rom or handle still used by a live plan         */
    STC_ETOOBIG = -6,  /* system too large for the fixed scratch of
                          stc_sweep_dev_noalloc                           */
    STC_ESINGULAR = 1, /* at least one frequency had info > 0             */
    STC_ENOCONV = 2    /* at least one frequency had info == -1           */
};

/* factorisation modes */
enum {
    STC_PLAIN_FP64 = 0,   /* complex128 LU with partial pivoting (zgetrf)      */
    STC_REFINED_FP32 = 1  /* complex64 LU + FP64 iterative refinement        */
};

typedef struct {
    double seconds;            /* device time of the solve, hipEvent clock     */
    long long points;          /* (system, frequency) points solved            */
    int max_refine_steps;      /* largest refinement step count over points    */
    int n_singular;            /* points with info > 0                         */
    int n_not_converged;       /* points with info == -1                       */
    double max_backward_error; /* max over points and columns (if computed)    */
} stc_stats;

/* ---- handle --------------------------------------------------------------- */

/* Create a handle on HIP device `device` (ordinal). */
int stc_create(stc_handle* h, int device);
/* STC_EBUSY while plans created on the handle are alive. */
int stc_destroy(stc_handle h);
const char* stc_strerror(int code);
/* 1 when the library was built for and runs on gfx950, else 0 (no GPU). */
int stc_device_ok(stc_handle h);

/* ---- launch policy (set by an autotuner; optional) -----------------------------
 * How systems that do not fit the LDS whole are factored.  kernel: 0 =
 * automatic, 1 = 16-column register panel (three workgroups per CU for
 * complex64, two for complex128), 2 = wide panel held in LDS (16..64 columns,
 * one or two workgroups per CU, by lds_budget), 3 = as 2 with the trailing
 * update's complex products in Gauss 3M form (three real MFMA products instead
 * of four; complex64 only, complex128 launches run 2; the automatic choice
 * for complex64 from n = 129 to 256), 4 = as 2 with a two-level
 * update (the LDS panels of each 64-column step, 32 for complex128, update only
 * that step; the columns right of it get one deferred rank-64 / rank-32
 * update; the automatic choice above n = 256).  lds_budget: bytes of LDS per
 * workgroup, 0 = automatic, else 16384 .. 163584 (the LDS the panel may use
 * sets the workgroups per CU and the panel width).  jb_max: widest panel
 * (0 = no cap).  An entry applies to launches whose largest system has order
 * n in that mode, or the nearest order with an entry; NULL removes it.  The
 * results do not depend on the policy beyond rounding (pivots follow LAPACK
 * in every kernel).  No counterpart in the reference (its only tuning knob is a
 * fixed batch size, test/krylovMidsizeTest/gpu/concurrent/batchedLU/src/config/config.cu:29-30). */
typedef struct {
    int kernel;
    int lds_budget;
    int jb_max;
} stc_policy;
int stc_set_policy(stc_handle h, int n, int mode, const stc_policy* p);
int stc_clear_policies(stc_handle h);

/* ---- ROM operators ------------------------------------------------------------
 * Copies K, M, D (n x n, leading dimension ld, host memory) to the device.
 * D may be NULL (undamped).  eta is the structural loss factor applied to K;
 * eta = 0 reproduces the reference, whose K_R already carries (1 + i eta)
 * (KrylovROMSubstructure.cpp:332-333, :996-997).  Replaces the members
 * myKComplexReduced / myMComplexReduced / myDComplexReduced
 * (src/analysis/KrylovROMSubstructure.h:212-216).  rom_id >= 0; setting an id
 * again replaces it (and drops its B_R / C_R when n changes).  Non-finite
 * entries are rejected (STC_EINVAL); a ROM read by a live plan cannot be
 * replaced or cleared (STC_EBUSY). */
int stc_rom_set(stc_handle h, int rom_id, int n, const stc_z64* K,
                const stc_z64* M, const stc_z64* D, int ld, double eta);

/* Input/output projections for the full back-transformation:
 * B_R (n x nin, ldb) and C_R (nout x n, ldc), myBReduced / myCReduced
 * (KrylovROMSubstructure.h:217-220). */
int stc_rom_set_io(stc_handle h, int rom_id, int nin, const stc_z64* B, int ldb,
                   int nout, const stc_z64* C, int ldc);

int stc_rom_clear(stc_handle h, int rom_id);

/* ---- the sweep --------------------------------------------------------------
 * For j in [0, nfreq): X_j = A(2 pi freq_hz[j])^{-1} F.
 *   F   : n x nrhs (ldf), shared by all frequencies (the reference reuses the
 *         projected load B_R F for every frequency, KrylovROMSubstructure.cpp:1423)
 *   X   : nfreq blocks of n x nrhs, block j at X + j * ldx * nrhs (frequency-
 *         major, as `results` is appended per frequency, :1451)
 *   info: nfreq ints (see above)
 *   berr: nullable; nfreq * nrhs normwise backward errors
 *         ||F - A X||_inf / (||A||_inf ||X||_inf + ||F||_inf) computed in FP64
 *   mode: STC_PLAIN_FP64 or STC_REFINED_FP32; tol / max_iter apply to the
 *         refined mode (tol <= 0 selects 1e-12, max_iter <= 0 selects 10)
 * Host memory in, host memory out; synchronous.
 * Returns STC_OK, STC_ESINGULAR or STC_ENOCONV (results still written), or an
 * error < 0. */
int stc_sweep(stc_handle h, int rom_id, int nfreq, const double* freq_hz,
              int nrhs, const stc_z64* F, int ldf, stc_z64* X, int ldx,
              int* info, double* berr, int mode, double tol, int max_iter,
              stc_stats* stats);

/* Same with every array already in device memory (d_ prefix), enqueued on
 * `stream` (hipStream_t as void*, NULL = handle stream).  d_berr nullable.
 * Does not synchronise; stats (nullable) are not filled. */
int stc_sweep_dev(stc_handle h, int rom_id, int nfreq, const double* d_freq_hz,
                  int nrhs, const stc_z64* d_F, int ldf, stc_z64* d_X, int ldx,
                  int* d_info, double* d_berr, int mode, double tol,
                  int max_iter, void* stream);

/* Allocation-free variant of stc_sweep_dev, same arguments and results.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory during the call (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  It calls no hipMalloc / hipFree and no
 * allocating library routine: the factor slots and the work-queue head live
 * in a fixed device array of the library (64 MiB per GPU, held from the
 * library's first stc_create on that GPU until the process ends).  Points run
 * on as many workgroups as that array has factor slots for, so very large n
 * runs on fewer workgroups (n = 1024 in complex128: three); a system whose
 * single slot would exceed the array is rejected with STC_ETOOBIG.  Every call on one GPU shares that GPU's array, whatever the handle:
 * calls that overlap in time on one GPU (other streams, other handles, other
 * host threads) race and corrupt each other's results; order them.  The queue
 * head is cleared by a kernel at the start of each call. */
int stc_sweep_dev_noalloc(stc_handle h, int rom_id, int nfreq,
                          const double* d_freq_hz, int nrhs, const stc_z64* d_F,
                          int ldf, stc_z64* d_X, int ldx, int* d_info,
                          double* d_berr, int mode, double tol, int max_iter,
                          void* stream);

/* ---- LU factors (LAPACK zgetrf / cgetrf layout) ---------------------------------
 * Factor A(2 pi freq_hz[j]) for nfreq frequencies and return the in-place
 * L\U factors (n x n, column-major, lda = n, block j at LU + j*n*n) and the
 * 1-based pivots (block j at ipiv + j*n).  mode STC_PLAIN_FP64 writes stc_z64,
 * STC_REFINED_FP32 writes stc_c32 (the factors the refined path uses).
 * info as above.  Host memory; synchronous. */
int stc_getrf(stc_handle h, int rom_id, int nfreq, const double* freq_hz,
              int mode, void* LU, int* ipiv, int* info);

/* ---- uneven batch: many ROMs and frequency sets in one persistent launch ----- */
typedef struct {
    int rom_id;
    int nfreq;
    const double* freq_hz; /* nfreq                                    */
    int nrhs;
    const stc_z64* F;      /* n x nrhs, ldf                             */
    int ldf;
    stc_z64* X;            /* nfreq blocks of n x nrhs (ldx)            */
    int ldx;
    int* info;             /* nfreq                                     */
    double* berr;          /* nullable, nfreq * nrhs                    */
} stc_sys_desc;

/* Host-memory batch: copies in, one persistent launch over every
 * (system, frequency) point, copies out.  Synchronous. */
int stc_sweep_batch(stc_handle h, int nsys, const stc_sys_desc* sys, int mode,
                    double tol, int max_iter, stc_stats* stats);

/* A plan holds a device-side work queue for a batch whose stc_sys_desc arrays
 * are DEVICE pointers, ordered longest-first, its own factor slots and queue
 * head, and a hipGraph of the launch.  stc_plan_execute replays it on `stream`
 * (NULL = handle stream); replays of one plan are ordered with each other.
 * The plan keeps its ROMs: stc_rom_set / stc_rom_clear on them, and
 * stc_destroy of the handle, return STC_EBUSY until stc_plan_destroy. */
int stc_plan_create(stc_handle h, int nsys, const stc_sys_desc* dev_sys,
                    int mode, double tol, int max_iter, stc_plan* plan);
int stc_plan_execute(stc_plan plan, void* stream);
int stc_plan_destroy(stc_plan plan);

/* ---- adjoint -------------------------------------------------------------------
 * Objective J = sum_j sum_r |c^H X_r(omega_j)|^2 (c: n complex).  Parameters:
 * per-mode stiffness scales s_k with dA/ds_k = (1 + i eta) kdiag[k] phi_k phi_k^T
 * (Phi: n x p real, ldphi; D held fixed) and the loss factor, dA/deta = i K.
 * Per frequency the kernel factors A, solves A X = F, then A^H mu = c (c^H X)
 * with the same factors and accumulates dJ/dtheta = -2 Re(mu^H dA/dtheta X).
 * The factors live only in the workgroup's factor slot between the two solves
 * of one frequency: nothing is stored per frequency.
 * mode as for stc_sweep (refined: both solves are refined to tol; the
 * gradient's error then follows kappa(A) * tol: tol = 1e-12 gives ~1e-7
 * relative on the n = 128 workload, tol = 1e-10 only ~1e-5).
 * No counterpart in the reference. */
int stc_adjoint(stc_handle h, int rom_id, int nfreq, const double* freq_hz,
                int nrhs, const stc_z64* F, int ldf, const stc_z64* c, int p,
                const double* Phi, int ldphi, const double* kdiag, int mode,
                double tol, int max_iter, double* J, double* dJ_ds,
                double* dJ_deta, int* info);

/* ---- full back-transformation (KrylovROMSubstructure.cpp:1393-1460) ----------
 * For every frequency: X = A^{-1} (B_R inputLoad), Y_j = C_R X (nout x nrhs,
 * block j at Y + j*nout*nrhs, the order of `results`, :1451) and, when H is
 * not NULL, H_j = C_R A^{-1} B_R (nout x nin, column-major, block j at
 * H + j*nout*nin; the reference writes it as one .mtx file per frequency,
 * :1433-1445).  Requires stc_rom_set_io.  B_R inputLoad is formed on the
 * device once per call; the sweep kernel projects each point's solution with
 * C_R itself, so only Y and H (nout rows per column) reach device memory, never
 * X.  Host memory; synchronous. */
int stc_back_transform(stc_handle h, int rom_id, int nfreq, const double* freq_hz,
                       const stc_z64* input_load, int nrhs, stc_z64* Y,
                       stc_z64* H, int* info, int mode, double tol,
                       int max_iter, stc_stats* stats);

/* ---- explicit matrices (batched getrf + getrs) ------------------------------
 * Solve A_j X_j = B_j for `batch` given n x n matrices with LU and partial
 * pivoting (zgetrf/zgetrs semantics; STC_REFINED_FP32 factors in complex64
 * and refines to tol as for stc_sweep).  A_j at A + j*strideA (column-major,
 * lda), B_j at B + j*strideB (n x nrhs, ldb; strideB = 0: one B for every j),
 * X_j written at X + j*strideX (ldx).  info[j] and berr[j*nrhs + c]
 * (nullable) as for stc_sweep.  n <= 1024.  Replaces the batched solves of
 * the interface Jacobian, cublasZgetrfBatched + cublasZgetrsBatched in
 * test/IJCSA/src/IJSA.cu:1201-1208 (there unpivoted: pivots NULL) and
 * test/krylovMidsizeTest/gpu/concurrent/batchedLU/src/main.cu:254-259.
 * Host memory; synchronous. */
int stc_gesv_batched(stc_handle h, int n, int batch, const stc_z64* A, int lda,
                     long long strideA, int nrhs, const stc_z64* B, int ldb,
                     long long strideB, stc_z64* X, int ldx, long long strideX,
                     int* info, double* berr, int mode, double tol, int max_iter,
                     stc_stats* stats);

/* Same with every array in device memory, enqueued on `stream` (NULL = handle
 * stream): A_j at d_A + j*strideA with lda = n (strideA >= n*n), X_j at
 * d_X + j*ldx*nrhs.  Does not synchronise. */
int stc_gesv_batched_dev(stc_handle h, int n, int batch, const stc_z64* d_A,
                         long long strideA, int nrhs, const stc_z64* d_B, int ldb,
                         long long strideB, stc_z64* d_X, int ldx, int* d_info,
                         double* d_berr, int mode, double tol, int max_iter,
                         void* stream);

#ifdef __cplusplus
}
#endif
#endif /* STACCATO_SWEEP_H */
