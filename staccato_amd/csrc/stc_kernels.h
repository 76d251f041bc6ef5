// stc_kernels.h -- gfx950 kernels of the frequency sweep.
//
// One workgroup (256 lanes = 4 wavefronts) owns one point (system, frequency)
// at a time and takes points from a device work queue until it is empty
// (persistent launch).  Per point it
//   1. assembles A(w) = (1 + i eta) K + i w D - w^2 M in FP64 from the ROM
//      operators (KrylovROMSubstructure.cpp:1412-1420) and rounds it to the
//      factor type T (double: plain path, float: refined path),
//   2. factors P A = L U with partial pivoting, LAPACK zgetrf semantics
//      (KrylovROMSubstructure.cpp:1428): pivot = first row of largest
//      |re| + |im|, full-row interchanges, unit-lower L and U in place,
//      - small n: the whole matrix lives in LDS, unblocked right-looking;
//      - larger n: the matrix lives in a per-workgroup slot of global memory
//        (cache resident) and is factored blocked right-looking: an LDS panel
//        of jb columns, deferred row interchanges, U12 = L11^{-1} A12 and the
//        trailing update A22 -= L21 U12 per chunk of columns,
//   3. solves for all right-hand sides (KrylovROMSubstructure.cpp:1430),
//   4. optionally refines in FP64: r = F - A x with A re-assembled in FP64,
//      A d = r with the T factors, x += d, until the normwise backward error
//      ||r||_inf / (||A||_inf ||x||_inf + ||F||_inf) <= tol,
//   5. optionally runs the adjoint with the same factors.
// A(w) is never written to global memory on the small path; on the blocked
// path the only copy is the workgroup's factor slot.
#pragma once

#include "stc_internal.h"

namespace stc {

template <typename T>
struct cx {
    T re, im;
};
using z64 = cx<double>;

constexpr double kTwoPi = 6.283185307179586476925286766559;

template <typename T>
__device__ __forceinline__ cx<T> cmk(T r, T i) {
    cx<T> z;
    z.re = r;
    z.im = i;
    return z;
}
template <typename T>
__device__ __forceinline__ cx<T> czero() { return cmk<T>(T(0), T(0)); }
template <typename T>
__device__ __forceinline__ cx<T> cconj(cx<T> a) { return cmk<T>(a.re, -a.im); }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
    return cmk<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
// acc -= a * b
template <typename T>
__device__ __forceinline__ void cfms(cx<T>& acc, cx<T> a, cx<T> b) {
    acc.re = fma(-a.re, b.re, acc.re);
    acc.re = fma(a.im, b.im, acc.re);
    acc.im = fma(-a.re, b.im, acc.im);
    acc.im = fma(-a.im, b.re, acc.im);
}
// acc += a * b
template <typename T>
__device__ __forceinline__ void cfma(cx<T>& acc, cx<T> a, cx<T> b) {
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(-a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(a.im, b.re, acc.im);
}
// 1 / z by Smith's method (no overflow of |z|^2)
template <typename T>
__device__ __forceinline__ cx<T> crecip(cx<T> z) {
    if (fabs(z.re) >= fabs(z.im)) {
        T r = z.im / z.re;
        T d = z.re + z.im * r;
        return cmk<T>(T(1) / d, -r / d);
    }
    T r = z.re / z.im;
    T d = z.im + z.re * r;
    return cmk<T>(r / d, T(-1) / d);
}
template <typename T>
__device__ __forceinline__ cx<T> cdiv(cx<T> a, cx<T> b) { return cmul(a, crecip(b)); }
template <typename T>
__device__ __forceinline__ T cabs1(cx<T> a) { return fabs(a.re) + fabs(a.im); }
template <typename T>
__device__ __forceinline__ cx<T> cshfl(cx<T> z, int src) {
    return cmk<T>(__shfl(z.re, src, 64), __shfl(z.im, src, 64));
}
// Broadcast from a wavefront-uniform lane: v_readlane into a scalar register
// (no LDS crossbar traffic, a few cycles of latency).
__device__ __forceinline__ float lane_read(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_read(double v, int l) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <typename T>
__device__ __forceinline__ cx<T> cbcast(cx<T> z, int l) {
    return cmk<T>(lane_read(z.re, l), lane_read(z.im, l));
}
// Replace lane l of a register with the (uniform) value v.
template <typename T>
__device__ __forceinline__ cx<T> cput(cx<T> v, int l, cx<T> old) {
    return ((int)(threadIdx.x & 63) == l) ? v : old;
}
template <typename T>
__device__ __forceinline__ cx<T> to_t(z64 a) { return cmk<T>((T)a.re, (T)a.im); }
template <typename T>
__device__ __forceinline__ z64 to_z(cx<T> a) { return cmk<double>((double)a.re, (double)a.im); }
__device__ __forceinline__ z64 ld_z(const double2* p) {
    double2 v = *p;
    return cmk<double>(v.x, v.y);
}
__device__ __forceinline__ void st_z(double2* p, z64 a) { *p = make_double2(a.re, a.im); }
// NaN-propagating max
__device__ __forceinline__ double nmax(double a, double b) { return (a > b || a != a) ? a : b; }
__device__ __forceinline__ double zabs(z64 a) { return hypot(a.re, a.im); }

// ---- assembly: element e (column-major, ld = n) of A(w) in FP64 ---------------
__device__ __forceinline__ z64 assemble_at(const SysDesc& s, size_t e, double w, double w2) {
    if (s.real_ops) {
        const double k = s.K[e];
        const double m = s.M[e];
        const double d = s.has_D ? s.D[e] : 0.0;
        return cmk<double>(k - w2 * m, s.eta * k + w * d);
    }
    const double2 k = reinterpret_cast<const double2*>(s.K)[e];
    const double2 m = reinterpret_cast<const double2*>(s.M)[e];
    double2 d = make_double2(0.0, 0.0);
    if (s.has_D) d = reinterpret_cast<const double2*>(s.D)[e];
    return cmk<double>(k.x - s.eta * k.y - w2 * m.x - w * d.y,
                       k.y + s.eta * k.x - w2 * m.y + w * d.x);
}

// ---- pivot search: first index of max |re|+|im| over rows [start, m) -------------
// Executed by one whole wavefront; every lane returns the same row.
template <typename T>
__device__ __forceinline__ int wave_argmax(const cx<T>* col, int start, int m) {
    const int lane = threadIdx.x & 63;
    T best = T(-1);
    int bi = 0x7fffffff;
    for (int i = start + lane; i < m; i += 64) {
        T v = cabs1(col[i]);
        if (v > best) { best = v; bi = i; }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        T ob = __shfl_xor(best, off, 64);
        int oi = __shfl_xor(bi, off, 64);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    return bi == 0x7fffffff ? start : bi;
}

// ---- unblocked LU of an LDS panel ---------------------------------------------------
// P: m x jb (column-major, ld ldp) in LDS; local row 0 is global row j0.
// Two barriers per column: phase S (wave 0 swaps + scales the pivot column,
// the other waves swap the pivot rows in the remaining panel columns) and
// phase U (rank-1 update, columns dealt round-robin to the waves; wave 0
// updates column jj+1 first and searches its pivot, a one-column look-ahead).
// ipiv[j0+jj] = global 0-based pivot row; rmap (logical row -> physical row,
// the composite row permutation) gets the same interchange; sc[0] = pivot
// slot, sc[1] = first zero pivot (global, 1-based) or 0.
template <typename T, int MAXR>
__device__ void panel_factor(cx<T>* P, int ldp, int m, int jb, int j0, int* ipiv,
                             int* rmap, int* sc) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (wave == 0) {
        int p = wave_argmax(P, 0, m);
        if (lane == 0) sc[0] = p;
    }
    __syncthreads();
    for (int jj = 0; jj < jb; ++jj) {
        const int p = sc[0];
        cx<T>* colj = P + (size_t)jj * ldp;
        if (wave == 0) {
            const cx<T> pv = colj[p];
            if (pv.re != T(0) || pv.im != T(0)) {
                const cx<T> r = crecip(pv);
                cx<T> v[MAXR];
#pragma unroll
                for (int t = 0; t < MAXR; ++t) {
                    const int i = jj + 1 + lane + 64 * t;
                    if (i < m) v[t] = colj[i == p ? jj : i];
                }
#pragma unroll
                for (int t = 0; t < MAXR; ++t) {
                    const int i = jj + 1 + lane + 64 * t;
                    if (i < m) colj[i] = cmul(v[t], r);
                }
                if (lane == 0) colj[jj] = pv;
            } else if (lane == 0 && sc[1] == 0) {
                sc[1] = j0 + jj + 1;
            }
            if (lane == 0) ipiv[j0 + jj] = j0 + p;
        } else {
            if (p != jj) {
                for (int c = tid - 64; c < jb; c += kThreads - 64) {
                    if (c == jj) continue;
                    cx<T>* col = P + (size_t)c * ldp;
                    cx<T> a = col[jj];
                    col[jj] = col[p];
                    col[p] = a;
                }
                if (tid == 64) {
                    const int t = rmap[j0 + jj];
                    rmap[j0 + jj] = rmap[j0 + p];
                    rmap[j0 + p] = t;
                }
            }
        }
        __syncthreads();
        if (jj + 1 < jb) {
            cx<T> l[MAXR];
#pragma unroll
            for (int t = 0; t < MAXR; ++t) {
                const int i = jj + 1 + lane + 64 * t;
                l[t] = i < m ? colj[i] : czero<T>();
            }
            for (int c = jj + 1 + wave; c < jb; c += kWaves) {
                cx<T>* colc = P + (size_t)c * ldp;
                const cx<T> u = colc[jj];
                if (u.re != T(0) || u.im != T(0)) {
#pragma unroll
                    for (int t = 0; t < MAXR; ++t) {
                        const int i = jj + 1 + lane + 64 * t;
                        if (i < m) {
                            cx<T> a = colc[i];
                            cfms(a, l[t], u);
                            colc[i] = a;
                        }
                    }
                }
                if (c == jj + 1) {  // wave 0 only: look-ahead pivot search
                    int pn = wave_argmax(colc, jj + 1, m);
                    if (lane == 0) sc[0] = pn;
                }
            }
        }
        __syncthreads();
    }
}

// ---- MFMA 16x16x4 (one wavefront): D[i][j] += sum_k A[i][k] B[k][j] ------------------
// Lane l supplies A[l%16][l/16] and B[l/16][l%16].  Accumulator element r of
// lane l holds D[i][l%16] with i = 4*(l/16) + r for f32 and i = l/16 + 4*r for
// f64 (both measured on gfx950); Mfma<T>::row gives i.
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));
template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
    using acc = v4f;
    static __device__ __forceinline__ int row(int lg, int r) { return 4 * lg + r; }
    static __device__ __forceinline__ acc run(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <>
struct Mfma<double> {
    using acc = v4d;
    static __device__ __forceinline__ int row(int lg, int r) { return lg + 4 * r; }
    static __device__ __forceinline__ acc run(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

constexpr int kChunk = 32;      // trailing columns per TRSM/update chunk
constexpr int kChunkLd = kChunk + 2;
constexpr int kSolveBlock = 32; // diagonal block of the blocked triangular solves

// ---- U12 chunk = L11^{-1} A12 (wave substitution) -----------------------------------------
// A12: logical rows [j0, j0+jb) of global columns [cb, cb+cw) (row r lives at
// physical row rmap[r]); L11 = unit-lower P[0:jb, 0:jb] (LDS, ld ldp).  Each
// wave solves columns wave, wave+4, ... (lanes = rows, jb <= 64) and the result
// goes to Uc (LDS, row-major, ld kChunkLd) and back to S.
template <typename T>
__device__ void trsm_chunk(cx<T>* __restrict__ S, int ld, const int* rmap, int j0, int jb,
                           const cx<T>* P, int ldp, cx<T>* Uc, int cb, int cw) {
    constexpr int CPW = kChunk / kWaves;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const size_t prow = lane < jb ? (size_t)rmap[j0 + lane] : 0;
    cx<T> y[CPW];
#pragma unroll
    for (int q = 0; q < CPW; ++q) {
        const int c = wave + kWaves * q;
        y[q] = (lane < jb && c < cw) ? S[prow + (size_t)(cb + c) * ld] : czero<T>();
    }
    for (int k = 0; k < jb - 1; ++k) {
        const cx<T> lk = (lane > k && lane < jb) ? P[lane + (size_t)k * ldp] : czero<T>();
#pragma unroll
        for (int q = 0; q < CPW; ++q) {
            const cx<T> vk = cbcast(y[q], k);
            cfms(y[q], lk, vk);
        }
    }
#pragma unroll
    for (int q = 0; q < CPW; ++q) {
        const int c = wave + kWaves * q;
        if (lane < jb && c < cw) {
            Uc[lane * kChunkLd + c] = y[q];
            S[prow + (size_t)(cb + c) * ld] = y[q];
        }
    }
}

// ---- trailing update of one chunk on MFMA -------------------------------------------------
// C = logical rows [j0+jb, n) x columns [cb, cb+cw) of S (physical rows via rmap)
//   -= L21 U12,  L21 = P[jb:m, 0:jb] (LDS), U12 = Uc (jb x cw, LDS, row-major).
// Computed transposed (C^T -= U^T L21^T) so that a lane's accumulators are 16
// consecutive rows of one column: wave tile 32 rows x 32 columns (2 x 2 MFMA
// tiles, real and imaginary accumulators, 4 real MFMAs per complex product).
template <typename T>
__device__ void trailing_mfma(cx<T>* __restrict__ S, int ld, const int* rmap, int j0, int jb, int m,
                              const cx<T>* P, int ldp, const cx<T>* Uc, int cb, int cw) {
    using M = Mfma<T>;
    using acc_t = typename M::acc;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int m2 = m - jb;
    for (int r0 = wave * 32; r0 < m2; r0 += kWaves * 32) {
        acc_t cre[2][2], cim[2][2];
        size_t prow[2];
        bool rok[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int row = r0 + 16 * a + li;
            rok[a] = row < m2;
            prow[a] = rok[a] ? (size_t)rmap[j0 + jb + row] : 0;
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = 16 * b + M::row(lg, r);
                    cx<T> v = czero<T>();
                    if (rok[a] && col < cw) v = S[prow[a] + (size_t)(cb + col) * ld];
                    cre[a][b][r] = v.re;
                    cim[a][b][r] = v.im;
                }
#pragma unroll 2
        for (int k0 = 0; k0 < jb; k0 += 4) {
            const int k = k0 + lg;
            const bool kok = k < jb;
            cx<T> lv[2], uv[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const int row = min(r0 + 16 * a + li, m2 - 1);
                lv[a] = kok ? P[(jb + row) + (size_t)k * ldp] : czero<T>();
            }
#pragma unroll
            for (int b = 0; b < 2; ++b) uv[b] = kok ? Uc[k * kChunkLd + 16 * b + li] : czero<T>();
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    cre[a][b] = M::run(-uv[b].re, lv[a].re, cre[a][b]);
                    cre[a][b] = M::run(uv[b].im, lv[a].im, cre[a][b]);
                    cim[a][b] = M::run(-uv[b].re, lv[a].im, cim[a][b]);
                    cim[a][b] = M::run(-uv[b].im, lv[a].re, cim[a][b]);
                }
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = 16 * b + M::row(lg, r);
                    if (rok[a] && col < cw)
                        S[prow[a] + (size_t)(cb + col) * ld] = cmk<T>(cre[a][b][r], cim[a][b][r]);
                }
    }
}

// ---- trailing update of one chunk on FP64 FMA (VALU) ----------------------------------------
// Same wave tile and lane ownership as trailing_mfma (lane: rows r0+16a+li,
// columns 16b+4lg+r); on MI355X FP64 FMA and FP64 MFMA share one peak rate.
template <typename T>
__device__ void trailing_valu(cx<T>* __restrict__ S, int ld, const int* rmap, int j0, int jb, int m,
                              const cx<T>* P, int ldp, const cx<T>* Uc, int cb, int cw) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int m2 = m - jb;
    for (int r0 = wave * 32; r0 < m2; r0 += kWaves * 32) {
        cx<T> c[2][2][4];
        size_t prow[2];
        bool rok[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int row = r0 + 16 * a + li;
            rok[a] = row < m2;
            prow[a] = rok[a] ? (size_t)rmap[j0 + jb + row] : 0;
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = 16 * b + 4 * lg + r;
                    c[a][b][r] = (rok[a] && col < cw) ? S[prow[a] + (size_t)(cb + col) * ld]
                                                      : czero<T>();
                }
        for (int k = 0; k < jb; ++k) {
            cx<T> lv[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) lv[a] = P[(jb + min(r0 + 16 * a + li, m2 - 1)) + (size_t)k * ldp];
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const cx<T> u = Uc[k * kChunkLd + 16 * b + 4 * lg + r];
#pragma unroll
                    for (int a = 0; a < 2; ++a) cfms(c[a][b][r], lv[a], u);
                }
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = 16 * b + 4 * lg + r;
                    if (rok[a] && col < cw) S[prow[a] + (size_t)(cb + col) * ld] = c[a][b][r];
                }
    }
}

// ---- LDS carve-up ---------------------------------------------------------------------
struct Lds {
    int* ipiv;   // pivot row (0-based, logical) per column
    int* rmap;   // logical row -> physical row (the composite row permutation)
    int* sc;     // scalars: [0] pivot slot, [1] first zero pivot (1-based) or 0
    double* red; // reduction scratch
    char* main;
};

__device__ __forceinline__ Lds carve(char* smem, int n) {
    Lds L;
    int off = 0;
    L.ipiv = reinterpret_cast<int*>(smem);
    L.rmap = L.ipiv + n;
    off += align16(2LL * n * 4);
    L.sc = reinterpret_cast<int*>(smem + off);
    off += align16(16 * 4);
    L.red = reinterpret_cast<double*>(smem + off);
    off += align16((long long)kRedDoubles * 8);
    L.main = smem + off;
    return L;
}

// ---- register-resident panel factorisation (blocked path) --------------------------------
// Wave w owns panel columns w, w+4, ... (CPW of them) and lane l owns local
// rows l + 64 t (t < MAXR) of those columns, so the rank-1 updates and the row
// interchanges never touch memory.  Per column jj: the owning wave searches
// the pivot (first row of largest |re|+|im|), swaps and scales its column and
// publishes the multipliers in lcol (LDS); after one barrier every wave swaps
// rows jj <-> p in its other columns (lane shuffles) and applies the rank-1
// update.  The factored panel ends in P (LDS, m x jb, ld m) and in S through
// the final row map.
template <typename T, int MAXR>
__device__ __forceinline__ cx<T> reg_get(const cx<T> (&col)[MAXR], int row) {
    const int slot = row >> 6;
    cx<T> v = col[0];
#pragma unroll
    for (int t = 1; t < MAXR; ++t)
        if (slot == t) v = col[t];
    return cshfl(v, row & 63);
}

// Register panel without row interchanges.  Rows stay where they were
// gathered (gather row g = lane + 64 t); every lane tracks lpos[t], the LAPACK
// logical position of its rows (a pivot swaps the positions jj <-> p only).
// Pivot candidates at step jj are the rows with lpos >= jj; ties go to the
// smallest lpos (first index, as izamax).  At the end the panel is scattered
// once into logical order (P, LDS) and into S at the rows' unchanged physical
// addresses, and rmap is rewritten: rmap[j0 + lpos] = physical row.
template <typename T, int MAXR, int CPW>
__device__ void panel_factor_reg(cx<T>* __restrict__ S, int ld, int n, int j0, int jb, int* ipiv,
                                 int* rmap, int* sc, cx<T>* lcol, cx<T>* P) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m = n - j0;
    cx<T> x[CPW][MAXR];
    int lpos[MAXR];
    int prow[MAXR];
#pragma unroll
    for (int t = 0; t < MAXR; ++t) {
        const int r = lane + 64 * t;
        lpos[t] = r < m ? r : 0x7fffffff;
        prow[t] = r < m ? rmap[j0 + r] : 0;
    }
#pragma unroll
    for (int q = 0; q < CPW; ++q)
#pragma unroll
        for (int t = 0; t < MAXR; ++t) {
            const int c = wave + kWaves * q, r = lane + 64 * t;
            x[q][t] = (c < jb && r < m) ? S[(size_t)prow[t] + (size_t)(j0 + c) * ld] : czero<T>();
        }
    for (int jj = 0; jj < jb; ++jj) {
        const int ow = jj & (kWaves - 1), oq = jj / kWaves;
        if (wave == ow) {
            cx<T> col[MAXR];
#pragma unroll
            for (int t = 0; t < MAXR; ++t) {
                col[t] = x[0][t];
#pragma unroll
                for (int q = 1; q < CPW; ++q)
                    if (q == oq) col[t] = x[q][t];
            }
            // argmax of |re|+|im| over rows with lpos >= jj; ties: smallest lpos
            T best = T(-1);
            int bpos = 0x7fffffff, bg = 0;
#pragma unroll
            for (int t = 0; t < MAXR; ++t) {
                if (lpos[t] >= jj && lpos[t] != 0x7fffffff) {
                    const T v = cabs1(col[t]);
                    if (v > best || (v == best && lpos[t] < bpos)) {
                        best = v;
                        bpos = lpos[t];
                        bg = lane + 64 * t;
                    }
                }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const T ob = __shfl_xor(best, off, 64);
                const int op = __shfl_xor(bpos, off, 64);
                const int og = __shfl_xor(bg, off, 64);
                if (ob > best || (ob == best && op < bpos)) { best = ob; bpos = op; bg = og; }
            }
            const cx<T> pv = reg_get<T, MAXR>(col, bg);
            const bool zero = pv.re == T(0) && pv.im == T(0);
            int pl = bpos, pg = bg;
            if (zero) {
                // no pivot: the row at logical position jj stays (LAPACK: ipiv = jj)
                pl = jj;
                int g = 0x7fffffff;
#pragma unroll
                for (int t = 0; t < MAXR; ++t)
                    if (lpos[t] == jj) g = lane + 64 * t;
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) g = min(g, __shfl_xor(g, off, 64));
                pg = g;
            } else {
                const cx<T> rp = crecip(pv);
#pragma unroll
                for (int t = 0; t < MAXR; ++t) {
                    // rows below the pivot after this step: lpos > jj, except the pivot row
                    const bool below = lpos[t] >= jj && lpos[t] != 0x7fffffff && (lane + 64 * t) != pg;
                    if (below) col[t] = cmul(col[t], rp);
                }
            }
#pragma unroll
            for (int t = 0; t < MAXR; ++t) {
                const bool below = lpos[t] >= jj && lpos[t] != 0x7fffffff && (lane + 64 * t) != pg;
                lcol[lane + 64 * t] = below ? col[t] : czero<T>();
#pragma unroll
                for (int q = 0; q < CPW; ++q)
                    if (q == oq) x[q][t] = col[t];
            }
            if (lane == 0) {
                sc[0] = pg;
                sc[2] = pl;
                ipiv[j0 + jj] = j0 + pl;
                if (zero && sc[1] == 0) sc[1] = j0 + jj + 1;
            }
        }
        __syncthreads();
        const int pg = sc[0], pl = sc[2];
        // logical positions: jj <-> pl
#pragma unroll
        for (int t = 0; t < MAXR; ++t) {
            if (lpos[t] == jj) lpos[t] = pl;
            else if (lpos[t] == pl) lpos[t] = jj;
        }
        cx<T> l[MAXR];
#pragma unroll
        for (int t = 0; t < MAXR; ++t) l[t] = lcol[lane + 64 * t];
#pragma unroll
        for (int q = 0; q < CPW; ++q) {
            const int c = wave + kWaves * q;
            if (c > jj && c < jb) {
                const cx<T> u = reg_get<T, MAXR>(x[q], pg);
#pragma unroll
                for (int t = 0; t < MAXR; ++t) cfms(x[q][t], l[t], u);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < MAXR; ++t) {
        const int r = lane + 64 * t;
        if (r < m) rmap[j0 + lpos[t]] = prow[t];
    }
#pragma unroll
    for (int q = 0; q < CPW; ++q)
#pragma unroll
        for (int t = 0; t < MAXR; ++t) {
            const int c = wave + kWaves * q, r = lane + 64 * t;
            if (c < jb && r < m) {
                P[lpos[t] + (size_t)c * m] = x[q][t];
                S[(size_t)prow[t] + (size_t)(j0 + c) * ld] = x[q][t];
            }
        }
}

// ---- inverses of the diagonal block L11 (unit lower) and U11 (upper) ----------------------
// From P[0:jb, 0:jb] (LDS, ld ldp).  Wave w computes columns w, w+4, ... of
// both, all of its columns interleaved (lanes = rows, jb <= 64).  Linv goes
// to LDS (ld jb) for the TRSM and, with Uinv, to the slot's block store
// (block column-major, ld jb) for the triangular solves.
template <typename T, int CPW>
__device__ void invert_diag(const cx<T>* P, int ldp, int jb, cx<T>* Linv_lds, cx<T>* __restrict__ Lg,
                            cx<T>* __restrict__ Ug) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    cx<T> xl[CPW], xu[CPW];
#pragma unroll
    for (int q = 0; q < CPW; ++q) {
        const int j = wave + kWaves * q;
        xl[q] = lane == j ? cmk<T>(T(1), T(0)) : czero<T>();
        xu[q] = xl[q];
    }
    for (int k = 0; k < jb - 1; ++k) {
        const cx<T> lk = (lane > k && lane < jb) ? P[lane + (size_t)k * ldp] : czero<T>();
#pragma unroll
        for (int q = 0; q < CPW; ++q) cfms(xl[q], lk, cbcast(xl[q], k));
    }
    for (int k = jb - 1; k >= 0; --k) {
        const cx<T> rk = crecip(P[k + (size_t)k * ldp]);
        const cx<T> uk = (lane < k) ? P[lane + (size_t)k * ldp] : czero<T>();
#pragma unroll
        for (int q = 0; q < CPW; ++q) {
            if (lane == k) xu[q] = cmul(xu[q], rk);
            cfms(xu[q], uk, cbcast(xu[q], k));
        }
    }
#pragma unroll
    for (int q = 0; q < CPW; ++q) {
        const int j = wave + kWaves * q;
        if (j < jb && lane < jb) {
            Linv_lds[lane + j * jb] = xl[q];
            Lg[lane + (size_t)j * jb] = xl[q];
            Ug[lane + (size_t)j * jb] = xu[q];
        }
    }
}

// ---- U12 chunk = Linv A12 (parallel product) --------------------------------------------------
// Thread t: column c = t % 32 of the chunk, rows t/32 + 8 s.  A12 comes
// straight from S (logical rows j0.. via rmap), all jb loads in flight.  The
// result goes to Uc only: the threads of one column read every row of A12,
// so S is updated from Uc after the next barrier (uc_to_slot).
template <typename T, int JB8>
__device__ void trsm_linv(cx<T>* __restrict__ S, int ld, const int* rmap, int j0, int jb,
                          const cx<T>* Linv, cx<T>* Uc, int cb, int cw) {
    const int tid = threadIdx.x, c = tid & 31, i0 = tid >> 5;
    cx<T> acc[JB8];
#pragma unroll
    for (int q = 0; q < JB8; ++q) acc[q] = czero<T>();
    if (c < cw) {
        const size_t col = (size_t)(cb + c) * ld;
        for (int k0 = 0; k0 < jb; k0 += 8) {
            cx<T> a[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                a[u] = (k0 + u < jb) ? S[(size_t)rmap[j0 + k0 + u] + col] : czero<T>();
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int k = k0 + u;
#pragma unroll
                for (int q = 0; q < JB8; ++q) {
                    const int i = i0 + 8 * q;
                    if (k < jb && i < jb && k <= i) cfma(acc[q], Linv[i + k * jb], a[u]);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < JB8; ++q) {
            const int i = i0 + 8 * q;
            if (i < jb) Uc[i * kChunkLd + c] = acc[q];
        }
    }
}

template <typename T>
__device__ __forceinline__ void uc_to_slot(cx<T>* __restrict__ S, int ld, const int* rmap, int j0,
                                           int jb, const cx<T>* Uc, int cb, int cw) {
    for (int e = threadIdx.x; e < jb * kChunk; e += kThreads) {
        const int i = e / kChunk, c = e % kChunk;
        if (c < cw) S[(size_t)rmap[j0 + i] + (size_t)(cb + c) * ld] = Uc[i * kChunkLd + c];
    }
}

// ---- blocked right-looking factorisation in a global slot --------------------------------
// Rows are never moved in S: a pivot swaps two entries of rmap and the panel
// rows (registers), so no row-interchange pass touches global memory.  Row r
// of the LAPACK-layout factors is physical row rmap[r] of S.  The slot also
// holds the inverted diagonal blocks (Linv, Uinv; jb x jb each per panel),
// used by the triangular solves.
template <typename T, int MAXR>
__device__ void factor_blocked(const SysDesc& s, double w, double w2, cx<T>* __restrict__ S,
                               const Layout& lay, const Lds& L) {
    constexpr int CPW = panel_cols_per_wave((int)sizeof(cx<T>), MAXR);
    const int n = s.n, tid = threadIdx.x;
    const size_t nn = (size_t)n * n;
#pragma unroll 4
    for (size_t e = tid; e < nn; e += kThreads) S[e] = to_t<T>(assemble_at(s, e, w, w2));
    for (int r = tid; r < n; r += kThreads) L.rmap[r] = r;
    if (tid == 0) L.sc[1] = 0;
    __syncthreads();

    const int jbmax = lay.jb;
    cx<T>* P = reinterpret_cast<cx<T>*>(L.main);
    cx<T>* Uc = reinterpret_cast<cx<T>*>(L.main + align16((long long)n * jbmax * sizeof(cx<T>)));
    cx<T>* Linv = reinterpret_cast<cx<T>*>(reinterpret_cast<char*>(Uc) +
                                          align16((long long)jbmax * kChunkLd * sizeof(cx<T>)));
    cx<T>* lcol = P;  // the multiplier column lives in the panel region while the panel is in registers
    const size_t nblk = (size_t)((n + jbmax - 1) / jbmax);
    cx<T>* Lg = S + nn;
    cx<T>* Ug = Lg + nblk * jbmax * jbmax;

    for (int j0 = 0, blk = 0; j0 < n; j0 += jbmax, ++blk) {
        const int jb = min(jbmax, n - j0), m = n - j0;
        panel_factor_reg<T, MAXR, CPW>(S, n, n, j0, jb, L.ipiv, L.rmap, L.sc, lcol, P);
        __syncthreads();
        invert_diag<T, CPW>(P, m, jb, Linv, Lg + (size_t)blk * jbmax * jbmax,
                            Ug + (size_t)blk * jbmax * jbmax);
        __syncthreads();
        for (int cb = j0 + jb; cb < n; cb += kChunk) {
            const int cw = min(kChunk, n - cb);
            trsm_linv<T, 8>(S, n, L.rmap, j0, jb, Linv, Uc, cb, cw);
            __syncthreads();
            uc_to_slot(S, n, L.rmap, j0, jb, Uc, cb, cw);
            trailing_mfma(S, n, L.rmap, j0, jb, m, P, m, Uc, cb, cw);
            __syncthreads();
        }
    }
}

// ---- triangular solves on LDS vectors w (n x gc, ld n) -----------------------------------
// LU: factors with logical row r at physical row rmap[r] (rmap == nullptr:
// identity).  Dg: LDS staging for one kSolveBlock^2 diagonal block, or nullptr
// when LU itself is in LDS.  Off-diagonal updates issue 8 loads per round.
template <typename T>
struct TriCtx {
    const cx<T>* LU;
    int ld;
    int n;
    const int* rmap;
    cx<T>* Dg;
    // blocked path: inverted diagonal blocks (block b at b*bs*bs, ld = its width)
    const cx<T>* Linv;
    const cx<T>* Uinv;
    int bs;
    int inv_u;  // use Uinv for the U solve (refined mode); else substitution
    __device__ __forceinline__ size_t row(int i) const { return rmap ? (size_t)rmap[i] : (size_t)i; }
    __device__ __forceinline__ cx<T> at(int i, int j) const { return LU[row(i) + (size_t)j * ld]; }
};

// Copy the diagonal block [k0, k0+kb)^2 into Dg (when staging); returns the
// accessor base and leading dimension for that block.
template <typename T>
__device__ __forceinline__ const cx<T>* stage_diag(const TriCtx<T>& t, int k0, int kb, int* ldd,
                                                   bool* mapped) {
    if (t.Dg) {
        for (int e = threadIdx.x; e < kb * kb; e += kThreads) {
            const int i = e % kb, j = e / kb;
            t.Dg[i + j * kSolveBlock] = t.at(k0 + i, k0 + j);
        }
        __syncthreads();
        *ldd = kSolveBlock;
        *mapped = false;
        return t.Dg;
    }
    *ldd = t.ld;
    *mapped = false;
    return t.LU + k0 + (size_t)k0 * t.ld;
}

template <typename T, bool CONJ, bool TRANS>
__device__ __forceinline__ void rect_update(const TriCtx<T>& t, int i, int k0, int kb, cx<T>* w,
                                            int gc) {
    // w[i] -= sum_{k<kb} op(LU)(i, k0+k) w[k0+k];  op = (TRANS ? element (k0+k, i) : (i, k0+k)),
    // conjugated when CONJ
    cx<T> acc[kMaxGroup];
#pragma unroll
    for (int g = 0; g < kMaxGroup; ++g) acc[g] = czero<T>();
    const size_t pi = t.row(i);
    for (int kk = 0; kk < kb; kk += 8) {
        cx<T> l[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int k = k0 + kk + u;
            l[u] = czero<T>();
            if (kk + u < kb) l[u] = TRANS ? t.LU[t.row(k) + (size_t)i * t.ld] : t.LU[pi + (size_t)k * t.ld];
            if (CONJ) l[u] = cconj(l[u]);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (kk + u < kb)
#pragma unroll
                for (int g = 0; g < kMaxGroup; ++g)
                    if (g < gc) cfma(acc[g], l[u], w[k0 + kk + u + (size_t)g * t.n]);
    }
#pragma unroll
    for (int g = 0; g < kMaxGroup; ++g)
        if (g < gc) {
            cx<T> a = w[i + (size_t)g * t.n];
            a.re -= acc[g].re;
            a.im -= acc[g].im;
            w[i + (size_t)g * t.n] = a;
        }
}

// y_b = Binv_b w_b for one diagonal block (lanes = rows, one wave per column).
template <typename T, bool LOWER>
__device__ __forceinline__ void block_inv_apply(const cx<T>* __restrict__ Bi, int k0, int kb, cx<T>* w,
                                                int n, int gc) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int g = wave; g < gc; g += kWaves) {
        cx<T>* y = w + (size_t)g * n + k0;
        cx<T> acc = czero<T>();
        for (int k8 = 0; k8 < kb; k8 += 8) {
            cx<T> bv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int k = k8 + u;
                const bool in = lane < kb && k < kb && (LOWER ? k <= lane : k >= lane);
                bv[u] = in ? Bi[lane + (size_t)k * kb] : czero<T>();
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (k8 + u < kb) cfma(acc, bv[u], y[k8 + u]);
        }
        if (lane < kb) y[lane] = acc;
    }
}

// L unit lower, forward, with the inverted diagonal blocks.
template <typename T>
__device__ void lower_unit_solve_inv(const TriCtx<T>& t, cx<T>* w, int gc) {
    const int tid = threadIdx.x, n = t.n;
    for (int k0 = 0, b = 0; k0 < n; k0 += t.bs, ++b) {
        const int kb = min(t.bs, n - k0);
        block_inv_apply<T, true>(t.Linv + (size_t)b * t.bs * t.bs, k0, kb, w, n, gc);
        __syncthreads();
        for (int i = k0 + kb + tid; i < n; i += kThreads) rect_update<T, false, false>(t, i, k0, kb, w, gc);
        __syncthreads();
    }
}

// U upper, backward, with the inverted diagonal blocks.
template <typename T>
__device__ void upper_solve_inv(const TriCtx<T>& t, cx<T>* w, int gc) {
    const int tid = threadIdx.x, n = t.n;
    const int nb = (n + t.bs - 1) / t.bs;
    for (int b = nb - 1; b >= 0; --b) {
        const int k0 = b * t.bs, kb = min(t.bs, n - k0);
        block_inv_apply<T, false>(t.Uinv + (size_t)b * t.bs * t.bs, k0, kb, w, n, gc);
        __syncthreads();
        for (int i = tid; i < k0; i += kThreads) rect_update<T, false, false>(t, i, k0, kb, w, gc);
        __syncthreads();
    }
}

// L unit lower, forward.
template <typename T>
__device__ void lower_unit_solve(const TriCtx<T>& t, cx<T>* w, int gc) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, n = t.n;
    for (int k0 = 0; k0 < n; k0 += kSolveBlock) {
        const int kb = min(kSolveBlock, n - k0);
        int ldd;
        bool mp;
        const cx<T>* Db = stage_diag(t, k0, kb, &ldd, &mp);
        for (int g = wave; g < gc; g += kWaves) {
            cx<T>* y = w + (size_t)g * n + k0;
            cx<T> v = lane < kb ? y[lane] : czero<T>();
            for (int k = 0; k < kb - 1; ++k) {
                const cx<T> vk = cbcast(v, k);
                if (lane > k && lane < kb) cfms(v, Db[lane + (size_t)k * ldd], vk);
            }
            if (lane < kb) y[lane] = v;
        }
        __syncthreads();
        for (int i = k0 + kb + tid; i < n; i += kThreads) rect_update<T, false, false>(t, i, k0, kb, w, gc);
        __syncthreads();
    }
}

// U upper, backward.
template <typename T>
__device__ void upper_solve(const TriCtx<T>& t, cx<T>* w, int gc) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, n = t.n;
    const int nb = (n + kSolveBlock - 1) / kSolveBlock;
    for (int b = nb - 1; b >= 0; --b) {
        const int k0 = b * kSolveBlock, kb = min(kSolveBlock, n - k0);
        int ldd;
        bool mp;
        const cx<T>* Db = stage_diag(t, k0, kb, &ldd, &mp);
        for (int g = wave; g < gc; g += kWaves) {
            cx<T>* y = w + (size_t)g * n + k0;
            cx<T> v = lane < kb ? y[lane] : czero<T>();
            for (int k = kb - 1; k >= 0; --k) {
                if (lane == k) v = cdiv(v, Db[k + (size_t)k * ldd]);
                const cx<T> vk = cbcast(v, k);
                if (lane < k) cfms(v, Db[lane + (size_t)k * ldd], vk);
            }
            if (lane < kb) y[lane] = v;
        }
        __syncthreads();
        for (int i = tid; i < k0; i += kThreads) rect_update<T, false, false>(t, i, k0, kb, w, gc);
        __syncthreads();
    }
}

// U^H y = b, forward (adjoint).
template <typename T>
__device__ void upper_h_solve(const TriCtx<T>& t, cx<T>* w, int gc) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, n = t.n;
    for (int k0 = 0; k0 < n; k0 += kSolveBlock) {
        const int kb = min(kSolveBlock, n - k0);
        int ldd;
        bool mp;
        const cx<T>* Db = stage_diag(t, k0, kb, &ldd, &mp);
        for (int g = wave; g < gc; g += kWaves) {
            cx<T>* y = w + (size_t)g * n + k0;
            cx<T> v = lane < kb ? y[lane] : czero<T>();
            for (int k = 0; k < kb; ++k) {
                if (lane == k) v = cdiv(v, cconj(Db[k + (size_t)k * ldd]));
                const cx<T> vk = cbcast(v, k);
                if (lane > k && lane < kb) cfms(v, cconj(Db[k + (size_t)lane * ldd]), vk);
            }
            if (lane < kb) y[lane] = v;
        }
        __syncthreads();
        for (int i = k0 + kb + tid; i < n; i += kThreads) rect_update<T, true, true>(t, i, k0, kb, w, gc);
        __syncthreads();
    }
}

// L^H v = y, backward, unit diagonal (adjoint).
template <typename T>
__device__ void lower_h_unit_solve(const TriCtx<T>& t, cx<T>* w, int gc) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, n = t.n;
    const int nb = (n + kSolveBlock - 1) / kSolveBlock;
    for (int b = nb - 1; b >= 0; --b) {
        const int k0 = b * kSolveBlock, kb = min(kSolveBlock, n - k0);
        int ldd;
        bool mp;
        const cx<T>* Db = stage_diag(t, k0, kb, &ldd, &mp);
        for (int g = wave; g < gc; g += kWaves) {
            cx<T>* y = w + (size_t)g * n + k0;
            cx<T> v = lane < kb ? y[lane] : czero<T>();
            for (int k = kb - 1; k > 0; --k) {
                const cx<T> vk = cbcast(v, k);
                if (lane < k) cfms(v, cconj(Db[k + (size_t)lane * ldd]), vk);
            }
            if (lane < kb) y[lane] = v;
        }
        __syncthreads();
        for (int i = tid; i < k0; i += kThreads) rect_update<T, true, true>(t, i, k0, kb, w, gc);
        __syncthreads();
    }
}

// ---- block-wide max of the kRedSlots values (NaN propagating) --------------------------
__device__ void block_max(double (&vals)[kRedSlots], double* red) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int q = 0; q < kRedSlots; ++q) {
        double v = vals[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v = nmax(v, __shfl_xor(v, off, 64));
        vals[q] = v;
    }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < kRedSlots; ++q) red[wave * kRedSlots + q] = vals[q];
    }
    __syncthreads();
    if (tid < kRedSlots) {
        double v = red[tid];
        for (int w = 1; w < kWaves; ++w) v = nmax(v, red[w * kRedSlots + tid]);
        red[kWaves * kRedSlots + tid] = v;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kRedSlots; ++q) vals[q] = red[kWaves * kRedSlots + q];
}

// ---- block-wide sum of NS values (fixed order, deterministic) -------------------------
template <int NS>
__device__ void block_sum(double (&vals)[NS], double* red) {
    static_assert(NS <= kRedSlots, "reduction scratch");
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        double v = vals[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        vals[q] = v;
    }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NS; ++q) red[wave * kRedSlots + q] = vals[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        double v = red[q];
        for (int w = 1; w < kWaves; ++w) v += red[w * kRedSlots + q];
        vals[q] = v;
    }
    __syncthreads();
}

// ---- FP64 residual r = b - op(A) x with the norms of the backward error -----------------
// ADJ = false: op(A) = A; ADJ = true: op(A) = A^H.  out[0] = ||op(A)||_inf (when
// need_anorm), then per column g: out[1+4g] = ||r||_inf, out[2+4g] = ||x||_inf,
// out[3+4g] = ||b||_inf, out[4+4g] = max |re r| + |im r|.
// b is global (ldb) when bg != nullptr, else LDS (bl, ld n).
template <bool ADJ>
__device__ void residual(const SysDesc& s, double w, double w2, const double2* bg, int ldb,
                         const z64* bl, const z64* xs, z64* rs, int gc, bool need_anorm,
                         double* red, double (&out)[kRedSlots]) {
    const int n = s.n, tid = threadIdx.x;
    // pass 1: r = b - op(A) x (and the row sums of |op(A)|), 4 columns of the
    // operators in flight per round
    double amax = 0.0;
    for (int i = tid; i < n; i += kThreads) {
        z64 acc[kMaxGroup];
#pragma unroll
        for (int g = 0; g < kMaxGroup; ++g)
            acc[g] = g < gc ? (bg ? ld_z(bg + i + (size_t)g * ldb) : bl[i + (size_t)g * n])
                            : czero<double>();
        double asum = 0.0;
        for (int j0 = 0; j0 < n; j0 += 4) {
            z64 a[4];
            if (s.real_ops) {
                const double* __restrict__ Kp = s.K;
                const double* __restrict__ Mp = s.M;
                const double* __restrict__ Dp = s.D;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int j = min(j0 + u, n - 1);
                    const size_t e = ADJ ? (size_t)j + (size_t)i * n : (size_t)i + (size_t)j * n;
                    const double k = Kp[e], mm = Mp[e];
                    const double d = s.has_D ? Dp[e] : 0.0;
                    a[u] = cmk<double>(k - w2 * mm, (ADJ ? -1.0 : 1.0) * (s.eta * k + w * d));
                }
            } else {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int j = min(j0 + u, n - 1);
                    a[u] = ADJ ? cconj(assemble_at(s, (size_t)j + (size_t)i * n, w, w2))
                               : assemble_at(s, (size_t)i + (size_t)j * n, w, w2);
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (j0 + u < n) {
                    if (need_anorm) asum += zabs(a[u]);
#pragma unroll
                    for (int g = 0; g < kMaxGroup; ++g)
                        if (g < gc) cfms(acc[g], a[u], xs[j0 + u + (size_t)g * n]);
                }
            }
        }
        amax = nmax(amax, asum);
#pragma unroll
        for (int g = 0; g < kMaxGroup; ++g)
            if (g < gc) rs[i + (size_t)g * n] = acc[g];
    }
    __syncthreads();
    // pass 2: the norms of the backward error
    double v[kRedSlots];
#pragma unroll
    for (int q = 0; q < kRedSlots; ++q) v[q] = 0.0;
    v[0] = amax;
    for (int i = tid; i < n; i += kThreads) {
#pragma unroll
        for (int g = 0; g < kMaxGroup; ++g)
            if (g < gc) {
                const z64 r = rs[i + (size_t)g * n];
                const z64 b = bg ? ld_z(bg + i + (size_t)g * ldb) : bl[i + (size_t)g * n];
                v[1 + 4 * g] = nmax(v[1 + 4 * g], zabs(r));
                v[2 + 4 * g] = nmax(v[2 + 4 * g], zabs(xs[i + (size_t)g * n]));
                v[3 + 4 * g] = nmax(v[3 + 4 * g], zabs(b));
                v[4 + 4 * g] = nmax(v[4 + 4 * g], cabs1(r));
            }
    }
    block_max(v, red);
#pragma unroll
    for (int q = 0; q < kRedSlots; ++q) out[q] = v[q];
}

__device__ __forceinline__ double backward_error(double rmax, double anorm, double xmax,
                                                 double bmax) {
    const double den = anorm * xmax + bmax;
    if (den == 0.0) return rmax == 0.0 ? 0.0 : __builtin_inf();
    return rmax / den;
}

// ---- solve (+ refine) one group of right-hand sides ---------------------------------------
// On entry wv holds P b (permuted, type T) for gc columns and xs/rs are free.
// On exit xs holds the solution (FP64).  Returns refinement steps; conv[] = 1
// for columns meeting tol; berr_out[] their backward errors (when computed).
template <typename T, bool ADJ>
__device__ int solve_group(const SysDesc& s, double w, double w2, const TriCtx<T>& tri,
                           const int* perm, cx<T>* wv, z64* xs, z64* rs, int gc,
                           const double2* bg, int ldb, const z64* bl, bool refine, bool want_berr,
                           double tol, int max_iter, double* anorm,
                           double (&berr_out)[kMaxGroup], bool* allconv, double* red) {
    const int n = s.n, tid = threadIdx.x;
    auto solve_t = [&]() {
        if (!ADJ) {
            if (tri.Linv) lower_unit_solve_inv(tri, wv, gc);
            else lower_unit_solve(tri, wv, gc);
            if (tri.Linv && tri.inv_u) upper_solve_inv(tri, wv, gc);
            else upper_solve(tri, wv, gc);
        } else {
            upper_h_solve(tri, wv, gc);
            lower_h_unit_solve(tri, wv, gc);
        }
    };
    // x = op(A)^{-1} b in T; ADJ leaves the result un-permuted (x[perm[r]] = wv[r])
    solve_t();
    for (int e = tid; e < n * gc; e += kThreads) {
        const int r = e % n, g = e / n;
        const int dst = ADJ ? perm[r] : r;
        xs[dst + (size_t)g * n] = to_z(wv[e]);
    }
    __syncthreads();
    int it = 0;
    *allconv = true;
    if (!want_berr) return 0;
    for (;;) {
        double out[kRedSlots];
        residual<ADJ>(s, w, w2, bg, ldb, bl, xs, rs, gc, *anorm < 0.0, red, out);
        if (*anorm < 0.0) *anorm = out[0];
        bool conv = true;
#pragma unroll
        for (int g = 0; g < kMaxGroup; ++g) {
            if (g < gc) {
                const double be =
                    backward_error(out[1 + 4 * g], *anorm, out[2 + 4 * g], out[3 + 4 * g]);
                berr_out[g] = be;
                if (!(be <= tol)) conv = false;
            }
        }
        if (!refine || conv || it >= max_iter) {
            *allconv = conv;
            break;
        }
        // d = op(A)^{-1} (r / s) in T, x += s d  (s = max |re|+|im| of r per column)
        double* scl = red + (kWaves + 1) * kRedSlots;
        if (tid < kMaxGroup) {
            double sv = 1.0;
#pragma unroll
            for (int g = 0; g < kMaxGroup; ++g)
                if (g == tid && out[4 + 4 * g] > 0.0) sv = out[4 + 4 * g];
            scl[tid] = sv;
        }
        __syncthreads();
        for (int e = tid; e < n * gc; e += kThreads) {
            const int r = e % n, g = e / n;
            const double sg = scl[g];
            const z64 rv = rs[(ADJ ? r : perm[r]) + (size_t)g * n];
            wv[e] = to_t<T>(cmk<double>(rv.re / sg, rv.im / sg));
        }
        __syncthreads();
        solve_t();
        for (int e = tid; e < n * gc; e += kThreads) {
            const int r = e % n, g = e / n;
            const double sg = scl[g];
            const int dst = ADJ ? perm[r] : r;
            z64 x = xs[dst + (size_t)g * n];
            const z64 d = to_z(wv[e]);
            x.re = fma(sg, d.re, x.re);
            x.im = fma(sg, d.im, x.im);
            xs[dst + (size_t)g * n] = x;
        }
        __syncthreads();
        ++it;
    }
    return it;
}

// ---- fused adjoint of one point ------------------------------------------------------------
// J_w = sum_r |c^H x_r|^2; mu_r = A^{-H} (c g_r) with the factors already in
// LU; dJ/deta_w = -2 Re(mu^H (i K) x) and dJ/ds_k,w = -2 Re((1+i eta) kdiag_k
// conj(phi_k^T mu) (phi_k^T x)), summed over r.  Writes adj_out[fidx*(p+2) + .]
// = {J_w, dJ/deta_w, dJ/ds_0,w, ...} and info[fidx] (0, or -1 when a refined
// solve missed tol).
template <typename T>
__device__ void adjoint_point(const SysDesc& s, int fidx, double w, double w2, const TriCtx<T>& tri,
                              const Lds& L, cx<T>* wv, z64* xs, z64* rs, z64* xk, z64* bv,
                              bool refine, double tol, int max_iter) {
    const int n = s.n, tid = threadIdx.x, p = s.p;
    double* out = s.adj_out + (size_t)fidx * (p + 2);
    double anorm_f = -1.0, anorm_a = -1.0;
    double accJ = 0.0, accE = 0.0;
    bool ok = true;
    for (int r = 0; r < s.nrhs; ++r) {
        const double2* Fr = s.F + (size_t)r * s.ldf;
        for (int i = tid; i < n; i += kThreads) wv[i] = to_t<T>(ld_z(Fr + L.rmap[i]));
        __syncthreads();
        double be[kMaxGroup];
        bool conv;
        solve_group<T, false>(s, w, w2, tri, L.rmap, wv, xs, rs, 1, Fr, s.ldf, nullptr, refine,
                              refine, tol, max_iter, &anorm_f, be, &conv, L.red);
        if (!conv) ok = false;
        // g = c^H x
        double gs[2] = {0.0, 0.0};
        for (int i = tid; i < n; i += kThreads) {
            const z64 x = xs[i];
            xk[i] = x;
            const z64 c = ld_z(s.c + i);
            gs[0] += c.re * x.re + c.im * x.im;
            gs[1] += c.re * x.im - c.im * x.re;
        }
        block_sum<2>(gs, L.red);
        const z64 g = cmk<double>(gs[0], gs[1]);
        if (s.X) {
            double2* X = s.X + (size_t)fidx * s.ldx * s.nrhs + (size_t)r * s.ldx;
            for (int i = tid; i < n; i += kThreads) st_z(X + i, xk[i]);
        }
        // adjoint right-hand side b = c g (not permuted for the A^H solve)
        for (int i = tid; i < n; i += kThreads) {
            const z64 b = cmul(ld_z(s.c + i), g);
            bv[i] = b;
            wv[i] = to_t<T>(b);
        }
        __syncthreads();
        solve_group<T, true>(s, w, w2, tri, L.rmap, wv, xs, rs, 1, nullptr, n, bv, refine,
                             refine, tol, max_iter, &anorm_a, be, &conv, L.red);
        if (!conv) ok = false;
        // mu^H (K x)
        double ks[2] = {0.0, 0.0};
        for (int i = tid; i < n; i += kThreads) {
            z64 kx = czero<double>();
            for (int j = 0; j < n; ++j) {
                const size_t e = (size_t)i + (size_t)j * n;
                z64 kij;
                if (s.real_ops) kij = cmk<double>(s.K[e], 0.0);
                else kij = ld_z(reinterpret_cast<const double2*>(s.K) + e);
                cfma(kx, kij, xk[j]);
            }
            cfma(*reinterpret_cast<z64*>(ks), cconj(xs[i]), kx);
        }
        block_sum<2>(ks, L.red);
        // -2 Re(i * s) = 2 Im(s)
        accE += 2.0 * ks[1];
        accJ += g.re * g.re + g.im * g.im;
        const z64 keta = cmk<double>(1.0, s.eta);
        for (int k = tid; k < p; k += kThreads) {
            const double* ph = s.Phi + (size_t)k * n;
            z64 pm = czero<double>(), px = czero<double>();
            for (int i = 0; i < n; ++i) {
                const double f = ph[i];
                pm.re = fma(f, xs[i].re, pm.re);
                pm.im = fma(f, xs[i].im, pm.im);
                px.re = fma(f, xk[i].re, px.re);
                px.im = fma(f, xk[i].im, px.im);
            }
            const z64 t = cmul(cmul(keta, cconj(pm)), px);
            const double d = -2.0 * s.kdiag[k] * t.re;
            out[2 + k] = (r == 0) ? d : out[2 + k] + d;
        }
        __syncthreads();
    }
    if (tid == 0) {
        out[0] = accJ;
        out[1] = accE;
        s.info[fidx] = ok ? 0 : -1;
        if (s.iters) s.iters[fidx] = 0;
    }
    __syncthreads();
}

// ---- one point ---------------------------------------------------------------------------
template <typename T, int MAXR>
__device__ void process_point(const SysDesc& s, int fidx, const KParams& kp, char* smem) {
    const int n = s.n, tid = threadIdx.x, nrhs = s.nrhs;
    const double w = kTwoPi * s.freq[fidx], w2 = w * w;
    Layout lay;
    lay.small = s.lay_small;
    lay.jb = s.lay_jb;
    lay.cw = s.lay_cw;
    lay.G = s.lay_G;
    lay.lds = 0;
    const Lds L = carve(smem, n);
    const bool refine = (kp.flags & kRefine) != 0;
    const bool want_berr = refine || (kp.flags & kWantBerr) || (s.berr != nullptr);

    const cx<T>* LU;
    char* solvep;
    if (lay.small) {
        cx<T>* A = reinterpret_cast<cx<T>*>(L.main);
        const size_t nn = (size_t)n * n;
        for (size_t e = tid; e < nn; e += kThreads) A[e] = to_t<T>(assemble_at(s, e, w, w2));
        for (int r = tid; r < n; r += kThreads) L.rmap[r] = r;
        if (tid == 0) L.sc[1] = 0;
        __syncthreads();
        panel_factor<T, MAXR>(A, n, n, n, 0, L.ipiv, L.rmap, L.sc);
        LU = A;
        solvep = L.main + align16((long long)nn * sizeof(cx<T>));
    } else {
        if (kp.scratch == nullptr || kp.slot_elems < slot_elems_for(n, lay.jb) ||
            lay.jb > kWaves * panel_cols_per_wave((int)sizeof(cx<T>), MAXR) || lay.jb > 64) {
            // no factor slot, or a panel wider than the register panel: report, touch nothing
            if (tid == 0) s.info[fidx] = -2;
            __syncthreads();
            return;
        }
        cx<T>* S = reinterpret_cast<cx<T>*>(kp.scratch) + (size_t)blockIdx.x * kp.slot_elems;
        factor_blocked<T, MAXR>(s, w, w2, S, lay, L);
        LU = S;
        solvep = L.main;
    }
    const int info = L.sc[1];

    if (kp.flags & kFactorOut) {
        // LAPACK layout: row r of the factors is physical row rmap[r] (blocked path)
        cx<T>* out = reinterpret_cast<cx<T>*>(s.factor_out) + (size_t)fidx * n * n;
        for (size_t e = tid; e < (size_t)n * n; e += kThreads) {
            const int r = (int)(e % n);
            const size_t c = e / n;
            out[e] = lay.small ? LU[e] : LU[(size_t)L.rmap[r] + c * n];
        }
        for (int r = tid; r < n; r += kThreads) s.ipiv_out[(size_t)fidx * n + r] = L.ipiv[r] + 1;
        if (tid == 0) s.info[fidx] = info;
        __syncthreads();
        return;
    }

    double2* X = s.X + (size_t)fidx * s.ldx * nrhs;
    if (info != 0) {
        const double qnan = __builtin_nan("");
        if (X)
            for (int e = tid; e < n * nrhs; e += kThreads) {
                const int r = e % n, g = e / n;
                X[r + (size_t)g * s.ldx] = make_double2(qnan, qnan);
            }
        if (s.berr)
            for (int g = tid; g < nrhs; g += kThreads) s.berr[(size_t)fidx * nrhs + g] = __builtin_inf();
        if (s.adj_out && tid == 0)
            for (int q = 0; q < s.p + 2; ++q) s.adj_out[(size_t)fidx * (s.p + 2) + q] = 0.0;
        if (tid == 0) {
            s.info[fidx] = info;
            if (s.iters) s.iters[fidx] = 0;
        }
        __syncthreads();
        return;
    }

    const int G = lay.G;
    const int nvec = (kp.flags & kAdjoint) ? 4 : 2;
    cx<T>* wv = reinterpret_cast<cx<T>*>(solvep);
    z64* xs = reinterpret_cast<z64*>(solvep + align16((long long)n * G * sizeof(cx<T>)));
    z64* rs = reinterpret_cast<z64*>(reinterpret_cast<char*>(xs) + align16((long long)n * G * 16));
    char* after = reinterpret_cast<char*>(xs) + nvec * align16((long long)n * G * 16);
    TriCtx<T> tri;
    tri.LU = LU;
    tri.ld = n;
    tri.n = n;
    tri.rmap = lay.small ? nullptr : L.rmap;
    tri.Dg = lay.small ? nullptr : reinterpret_cast<cx<T>*>(after);
    tri.bs = lay.jb;
    tri.Linv = lay.small ? nullptr : LU + (size_t)n * n;
    tri.Uinv = lay.small ? nullptr : LU + (size_t)n * n + (size_t)((n + lay.jb - 1) / lay.jb) * lay.jb * lay.jb;
    tri.inv_u = refine ? 1 : 0;
    if (kp.flags & kAdjoint) {
        z64* xk = reinterpret_cast<z64*>(reinterpret_cast<char*>(rs) + align16((long long)n * G * 16));
        z64* bv = reinterpret_cast<z64*>(reinterpret_cast<char*>(xk) + align16((long long)n * G * 16));
        adjoint_point<T>(s, fidx, w, w2, tri, L, wv, xs, rs, xk, bv, refine, kp.tol, kp.max_iter);
        return;
    }
    double anorm = -1.0;
    int it_max = 0;
    bool ok = true;
    for (int g0 = 0; g0 < nrhs; g0 += G) {
        const int gc = min(G, nrhs - g0);
        const double2* Fg = s.F + (size_t)g0 * s.ldf;
        for (int e = tid; e < n * gc; e += kThreads) {
            const int r = e % n, g = e / n;
            wv[e] = to_t<T>(ld_z(Fg + L.rmap[r] + (size_t)g * s.ldf));
        }
        __syncthreads();
        double be[kMaxGroup];
        bool conv;
        const int it = solve_group<T, false>(s, w, w2, tri, L.rmap, wv, xs, rs, gc, Fg, s.ldf,
                                             nullptr, refine, want_berr, kp.tol, kp.max_iter,
                                             &anorm, be, &conv, L.red);
        if (refine && !conv) ok = false;
        if (it > it_max) it_max = it;
        for (int e = tid; e < n * gc; e += kThreads) {
            const int r = e % n, g = e / n;
            st_z(X + r + (size_t)(g0 + g) * s.ldx, xs[e]);
        }
        if (s.berr && want_berr && tid == 0) {
#pragma unroll
            for (int g = 0; g < kMaxGroup; ++g)
                if (g < gc) s.berr[(size_t)fidx * nrhs + g0 + g] = be[g];
        }
        __syncthreads();
    }
    if (tid == 0) {
        s.info[fidx] = ok ? 0 : -1;
        if (s.iters) s.iters[fidx] = it_max;
    }
}

// ---- descriptor in scalar registers --------------------------------------------------------
// Every dword passes through readfirstlane, so the compiler knows the fields
// are wavefront-uniform and keeps them (and the addresses built from them) in
// SGPRs instead of spending vector registers on them.
__device__ __forceinline__ SysDesc load_uniform(const SysDesc* p) {
    static_assert(sizeof(SysDesc) % 4 == 0, "descriptor size");
    SysDesc d;
    const int* src = reinterpret_cast<const int*>(p);
    int* dst = reinterpret_cast<int*>(&d);
#pragma unroll
    for (int q = 0; q < (int)(sizeof(SysDesc) / 4); ++q) dst[q] = __builtin_amdgcn_readfirstlane(src[q]);
    return d;
}

// ---- the persistent kernel -------------------------------------------------------------
// Two workgroups (8 wavefronts) per CU: the register file is split two ways,
// so each wavefront gets at most 256 registers.
template <typename T, int MAXR>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2, 2)))
sweep_kernel(KParams kp) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    __shared__ int s_item;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) s_item = atomicAdd(kp.counter, 1);
        __syncthreads();
        // uniform across the workgroup: keeps the descriptor in scalar registers
        const int item = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (item >= kp.nitems) break;
        int sysi = 0, fidx = item;
        if (kp.items) {
            const int2 it = kp.items[item];
            sysi = __builtin_amdgcn_readfirstlane(it.x);
            fidx = __builtin_amdgcn_readfirstlane(it.y);
        }
        const SysDesc s = kp.sys ? load_uniform(kp.sys + sysi) : load_uniform(&kp.one);
        process_point<T, MAXR>(s, fidx, kp, smem);
        __syncthreads();
    }
}

__global__ void reset_counter_kernel(int* counter) {
    if (threadIdx.x == 0) counter[0] = 0;
}

}  // namespace stc
