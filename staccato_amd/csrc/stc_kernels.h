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
//        (reused point after point; its traffic goes through L2 and, once the
//        resident slots exceed the caches, HBM) and is factored blocked
//        right-looking: a panel of jb columns (registers, or LDS in the wide
//        path of stc_wide.h), deferred row interchanges, U12 = L11^{-1} A12
//        and the trailing update A22 -= L21 U12 per chunk of columns,
//   3. solves for all right-hand sides (KrylovROMSubstructure.cpp:1430),
//   4. optionally refines in FP64: r = F - A x with A re-assembled in FP64,
//      A d = r with the T factors, x += d, until the normwise backward error
//      ||r||_inf / (||A||_inf ||x||_inf + ||F||_inf) <= tol,
//   5. optionally runs the adjoint with the same factors,
//   6. optionally projects the solution, Y = C X (KrylovROMSubstructure.cpp:1447-1451),
//      so that only Y leaves the workgroup.
// A(w) is never written to global memory on the small path; on the blocked
// path the only copy is the workgroup's factor slot.
#pragma once

#include "stc_internal.h"

#include <type_traits>
#include <utility>

namespace stc {

template <typename T>
struct cx {
    T re, im;
};
using z64 = cx<double>;

// ---- explicit address spaces ---------------------------------------------------------------
// Pointers reach device code through the system descriptor as generic
// pointers, which the compiler turns into flat instructions with 64-bit
// address arithmetic.  These accessors state the address space: 1 = global
// memory (global_load/store, scalar base + 32-bit offset), 3 = LDS (ds_*).
template <typename T>
struct vec2;
template <>
struct vec2<float> {
    typedef float type __attribute__((ext_vector_type(2)));
};
template <>
struct vec2<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};
template <int AS, typename T>
__device__ __forceinline__ cx<T> ldc(const cx<T>* p) {
    using V = typename vec2<T>::type;
    const V v = *(const __attribute__((address_space(AS))) V*)(p);
    cx<T> c;
    c.re = v.x;
    c.im = v.y;
    return c;
}
template <int AS, typename T>
__device__ __forceinline__ void stc(cx<T>* p, cx<T> a) {
    using V = typename vec2<T>::type;
    V v;
    v.x = a.re;
    v.y = a.im;
    *(__attribute__((address_space(AS))) V*)(p) = v;
}
template <int AS, typename T>
__device__ __forceinline__ T lds1(const T* p) {
    return *(const __attribute__((address_space(AS))) T*)(p);
}

constexpr double kTwoPi = 6.283185307179586476925286766559;

template <typename T>
__device__ __forceinline__ cx<T> cmk(T r, T i);
// Load that every lane issues, its address clamped into range by the caller,
// zeroed afterwards where the element does not exist.  A load under a
// divergent branch makes the compiler wait for it before the branch joins;
// these issue as one batch and are waited for once.
template <int AS, typename T>
__device__ __forceinline__ cx<T> ldz(const cx<T>* p, bool ok) {
    const cx<T> v = ldc<AS>(p);
    return cmk<T>(ok ? v.re : T(0), ok ? v.im : T(0));
}

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
// complex64 (the factor of the refined mode, whose rounding the FP64
// refinement absorbs): Smith's scaling with one hardware reciprocal
// (v_rcp_f32, 1 ulp) instead of three correctly rounded divisions.
template <>
__device__ __forceinline__ cx<float> crecip(cx<float> z) {
    const bool re_big = fabsf(z.re) >= fabsf(z.im);
    const float a = re_big ? z.re : z.im, b = re_big ? z.im : z.re;
    const float ia = __builtin_amdgcn_rcpf(a);
    const float r = b * ia;
    const float id = __builtin_amdgcn_rcpf(a + b * r);
    return re_big ? cmk<float>(id, -r * id) : cmk<float>(r * id, -id);
}
template <typename T>
__device__ __forceinline__ cx<T> cdiv(cx<T> a, cx<T> b) { return cmul(a, crecip(b)); }
template <typename T>
__device__ __forceinline__ T cabs1(cx<T> a) { return fabs(a.re) + fabs(a.im); }
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
// z from lane src of the wavefront (one ds_bpermute per part).
// lane K of each row of 16 lanes to the whole row (DPP row_newbcast)
template <int K>
__device__ __forceinline__ unsigned rowbcast_u(unsigned v) {
    static_assert(K >= 0 && K < 16, "lane within a row of 16");
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x150 + K, 0xF, 0xF, false);
}
template <int K>
__device__ __forceinline__ float rowbcast(float v) { return __uint_as_float(rowbcast_u<K>(__float_as_uint(v))); }
template <int K>
__device__ __forceinline__ double rowbcast(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return __longlong_as_double((long long)(((unsigned long long)rowbcast_u<K>((unsigned)(b >> 32)) << 32) |
                                            rowbcast_u<K>((unsigned)b)));
}
template <int K, typename T>
__device__ __forceinline__ cx<T> crowbcast(cx<T> z) { return cmk<T>(rowbcast<K>(z.re), rowbcast<K>(z.im)); }
// f(integral_constant<int, J>) for each J of the sequence, in order
template <typename F, int... J>
__device__ __forceinline__ void static_for(F&& f, std::integer_sequence<int, J...>) {
    (f(std::integral_constant<int, J>{}), ...);
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
// |a| for norms: sqrt(re^2 + im^2), falling back to hypot's scaling only
// when the squares overflow or underflow
__device__ __forceinline__ double zabs(z64 a) {
    const double q = fma(a.re, a.re, a.im * a.im);
    return (q < 1e300 && q > 1e-300) || q == 0.0 ? sqrt(q) : hypot(a.re, a.im);
}

// ---- assembly: element e (column-major, ld = n) of A(w) in FP64 ---------------
__device__ __forceinline__ z64 assemble_at(const SysDesc& s, size_t e, double w, double w2) {
    if (s.real_ops) {
        const double k = lds1<1>(s.K + e);
        const double m = lds1<1>(s.M + e);
        const double d = s.has_D ? lds1<1>(s.D + e) : 0.0;
        return cmk<double>(k - w2 * m, s.eta * k + w * d);
    }
    const z64 k = ldc<1>(reinterpret_cast<const z64*>(s.K) + e);
    const z64 m = ldc<1>(reinterpret_cast<const z64*>(s.M) + e);
    z64 d = czero<double>();
    if (s.has_D) d = ldc<1>(reinterpret_cast<const z64*>(s.D) + e);
    return cmk<double>(k.re - s.eta * k.im - w2 * m.re - w * d.im,
                       k.im + s.eta * k.re - w2 * m.im + w * d.re);
}

// ---- pivot key of a candidate row: |re| + |im|, with LAPACK izamax's NaN rule ------
// izamax starts from the first element and replaces it only by a strictly
// larger one, so a NaN in the first eligible row keeps that row, and a NaN
// anywhere else is never chosen.  As a key: NaN in the first row -> +inf (it
// wins every comparison, ties go to the first row), NaN elsewhere -> -0.5
// (below every number, above the "no candidate" start value -1).
template <typename T>
__device__ __forceinline__ T pivot_key(cx<T> a, bool first) {
    const T v = cabs1(a);
    return v == v ? v : (first ? T(__builtin_inf()) : T(-0.5));
}

// ---- pivot search: first index of max |re|+|im| over rows [start, m) -------------
// Executed by one whole wavefront; every lane returns the same row.
template <typename T>
__device__ __forceinline__ int wave_argmax(const cx<T>* col, int start, int m) {
    const int lane = threadIdx.x & 63;
    T best = T(-1);
    int bi = 0x7fffffff;
    for (int i = start + lane; i < m; i += 64) {
        T v = pivot_key(col[i], i == start);
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
__device__ __forceinline__ void panel_factor(cx<T>* P, int ldp, int m, int jb, int j0, int* ipiv,
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

constexpr int kSolveBlock = 32; // diagonal block of the blocked triangular solves
// Operator columns per load round of the residual: 8 in the complex64 kernel
// (168 registers per lane), 16 in the complex128 one (256).
template <typename T>
struct ResCols {
    static constexpr int K = sizeof(T) == 4 ? 8 : 16;
};
constexpr int kUw = kUwCols;    // columns of one wave's U12 chunk
constexpr int kUwLd = kUwLdCols;

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
// interchanges never touch memory.  The factored panel ends in P (LDS, m x jb,
// ld m, LAPACK row order) and in S through the final row map.
// ---- DPP lane permutations inside rows of 16 lanes (gfx9 encodings) ----------------------
constexpr int kDppXor1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141; // lane i <-> 7-i within 8
constexpr int kDppMirror = 0x140;     // lane i <-> 15-i within 16

template <int CTRL>
__device__ __forceinline__ int dpp(int v) { return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xF, 0xF, false); }
template <int CTRL>
__device__ __forceinline__ float dpp(float v) { return __int_as_float(dpp<CTRL>(__float_as_int(v))); }
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)dpp<CTRL>((int)b), hi = (unsigned)dpp<CTRL>((int)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float uniform(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double uniform(double v) { return lane_read(v, 0); }
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Pivot search reduction over one wave: keeps (best, bpos, bg) of the lane
// with the largest best, ties to the smallest bpos (the first row in LAPACK
// order).  Four DPP steps reduce each row of 16 lanes, two lane shuffles
// join the rows; the result is read from lane 0, so it is wave-uniform.
template <typename T>
__device__ __forceinline__ void wave_argmax3(T& best, int& bpos, int& bg) {
    auto comb = [&](T ob, int op, int og) {
        if (ob > best || (ob == best && op < bpos)) {
            best = ob;
            bpos = op;
            bg = og;
        }
    };
    comb(dpp<kDppXor1>(best), dpp<kDppXor1>(bpos), dpp<kDppXor1>(bg));
    comb(dpp<kDppXor2>(best), dpp<kDppXor2>(bpos), dpp<kDppXor2>(bg));
    comb(dpp<kDppHalfMirror>(best), dpp<kDppHalfMirror>(bpos), dpp<kDppHalfMirror>(bg));
    comb(dpp<kDppMirror>(best), dpp<kDppMirror>(bpos), dpp<kDppMirror>(bg));
    comb(__shfl_xor(best, 16, 64), __shfl_xor(bpos, 16, 64), __shfl_xor(bg, 16, 64));
    comb(__shfl_xor(best, 32, 64), __shfl_xor(bpos, 32, 64), __shfl_xor(bg, 32, 64));
    best = uniform(best);
    bpos = uniform(bpos);
    bg = uniform(bg);
}

// Register panel without row interchanges.  Rows stay where they were
// gathered (gather row g = lane + 64 t); every lane tracks lpos[t], the LAPACK
// logical position of its rows (a pivot swaps the positions jj <-> p only).
// Column jj belongs to wave jj % 4.  A wave keeps its columns in x[] rotated
// so that its next unfactored column is always x[0] (one rotation after each
// of its columns, undone at the end), which reaches the pivot column without
// a register select.  The owner publishes the multipliers (lcol) and the
// pivot (sc) in one of two LDS buffers, alternating by column: one barrier
// per column suffices, since a buffer is rewritten only two columns later,
// after every wave has passed the barrier of the column in between.  Each
// wave then reads the pivot row's entries of its own columns with v_readlane
// (the pivot lane is uniform) and applies the rank-1 update.
template <typename T, int MAXR, int CPW>
__device__ __forceinline__ void rotate_cols(cx<T> (&x)[CPW][MAXR]) {
#pragma unroll
    for (int t = 0; t < MAXR; ++t) {
        const cx<T> x0 = x[0][t];
#pragma unroll
        for (int q = 0; q + 1 < CPW; ++q) x[q][t] = x[q + 1][t];
        x[CPW - 1][t] = x0;
    }
}

template <typename T, int MAXR, int CPW>
__device__ __forceinline__ void panel_factor_reg(cx<T>* __restrict__ S, int ld, int n, int j0, int jb, int* ipiv,
                                 int* rmap, int* sc, cx<T>* lcolbuf, cx<T>* P) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m = n - j0;
    constexpr int kLcolStride = 64 * MAXR;  // second buffer: lcolbuf + kLcolStride (<= n jb / 2)
    cx<T> x[CPW][MAXR];
    int lpos[MAXR];
    int prow[MAXR];
#pragma unroll
    for (int t = 0; t < MAXR; ++t) {
        const int r = lane + 64 * t;
        lpos[t] = r < m ? r : 0x7fffffff;
        prow[t] = r < m ? lds1<3>(rmap + j0 + r) : 0;
    }
#pragma unroll
    for (int q = 0; q < CPW; ++q)
#pragma unroll
        for (int t = 0; t < MAXR; ++t) {
            const int c = wave + kWaves * q, r = lane + 64 * t;
            x[q][t] = (c < jb && r < m) ? ldc<1>(S + (prow[t] + (j0 + c) * ld)) : czero<T>();
        }
    int done = 0;  // columns of this wave already factored (= rotations applied)
    for (int jj = 0; jj < jb; ++jj) {
        const int lb = (jj & 1) * kLcolStride, sb = 4 + 4 * (jj & 1);
        if (wave == (jj & (kWaves - 1))) {
            // argmax of |re|+|im| over rows with lpos >= jj; ties: smallest lpos.
            // All-zero column: the smallest eligible lpos is jj itself, so the
            // row stays (LAPACK: ipiv = jj) and no scaling happens.  NaN keys
            // follow izamax (pivot_key), so some eligible row always wins.
            T best = T(-1);
            int bpos = 0x7fffffff, bg = 0;
#pragma unroll
            for (int t = 0; t < MAXR; ++t) {
                if (lpos[t] >= jj && lpos[t] != 0x7fffffff) {
                    const T v = pivot_key(x[0][t], lpos[t] == jj);
                    if (v > best || (v == best && lpos[t] < bpos)) {
                        best = v;
                        bpos = lpos[t];
                        bg = lane + 64 * t;
                    }
                }
            }
            wave_argmax3(best, bpos, bg);
            const int ps = bg >> 6, pln = bg & 63;
            cx<T> pv = x[0][0];
#pragma unroll
            for (int t = 1; t < MAXR; ++t)
                if (ps == t) pv = x[0][t];
            pv = cbcast(pv, pln);
            const bool zero = pv.re == T(0) && pv.im == T(0);
            const cx<T> rp = zero ? czero<T>() : crecip(pv);
#pragma unroll
            for (int t = 0; t < MAXR; ++t) {
                // rows below the pivot after this step: lpos >= jj, except the pivot row
                // (slot ps of lane pln; both uniform)
                const bool below = lpos[t] >= jj && lpos[t] != 0x7fffffff && (t != ps || lane != pln);
                if (below && !zero) x[0][t] = cmul(x[0][t], rp);
                stc<3>(lcolbuf + lb + lane + 64 * t, below ? x[0][t] : czero<T>());
            }
            if (lane == 0) {
                sc[sb] = bg;
                sc[sb + 2] = bpos;
                ipiv[j0 + jj] = j0 + bpos;
                if (zero && sc[1] == 0) sc[1] = j0 + jj + 1;
            }
        }
        __syncthreads();
        const int pg = uniform(sc[sb]), pl = uniform(sc[sb + 2]);
        // logical positions: jj <-> pl
#pragma unroll
        for (int t = 0; t < MAXR; ++t) {
            if (lpos[t] == jj) lpos[t] = pl;
            else if (lpos[t] == pl) lpos[t] = jj;
        }
        cx<T> l[MAXR];
#pragma unroll
        for (int t = 0; t < MAXR; ++t) l[t] = ldc<3>(lcolbuf + lb + lane + 64 * t);
        const int ps = pg >> 6, pln = pg & 63;
#pragma unroll
        for (int q = 0; q < CPW; ++q) {
            const int c = wave + kWaves * (q + done);  // x[q] holds this column while unfactored
            if (c > jj && c < jb) {
                cx<T> v = x[q][0];
#pragma unroll
                for (int t = 1; t < MAXR; ++t)
                    if (ps == t) v = x[q][t];
                const cx<T> u = cbcast(v, pln);
#pragma unroll
                for (int t = 0; t < MAXR; ++t) cfms(x[q][t], l[t], u);
            }
        }
        if (wave == (jj & (kWaves - 1))) {
            rotate_cols(x);
            ++done;
        }
    }
    for (int r = done; r < CPW; ++r) rotate_cols(x);  // back to x[q] = column wave + 4 q
    // every wave has read the last multipliers before P (which holds the
    // lcol buffers) is overwritten
    __syncthreads();
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
                stc<3>(P + (lpos[t] + c * m), x[q][t]);
                stc<1>(S + (prow[t] + (j0 + c) * ld), x[q][t]);
            }
        }
}

// ---- inverse of the diagonal block L11 (unit lower) ----------------------------------------
// From P[0:jb, 0:jb] (LDS, ld ldp).  Wave w computes columns w, w+4, ... (lanes
// = rows, jb <= 64).  Linv goes to LDS (ld jb) for the TRSM and to the slot's
// block store (block column-major, ld jb) for the triangular solves.
template <typename T, int CPW>
__device__ __forceinline__ void invert_diag(const cx<T>* P, int ldp, int jb, cx<T>* Linv_lds, cx<T>* __restrict__ Lg) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    cx<T> xl[CPW];
#pragma unroll
    for (int q = 0; q < CPW; ++q) {
        const int j = wave + kWaves * q;
        xl[q] = lane == j ? cmk<T>(T(1), T(0)) : czero<T>();
    }
    for (int k = 0; k < jb - 1; ++k) {
        const cx<T> lk = (lane > k && lane < jb) ? P[lane + (size_t)k * ldp] : czero<T>();
#pragma unroll
        for (int q = 0; q < CPW; ++q) cfms(xl[q], lk, cbcast(xl[q], k));
    }
#pragma unroll
    for (int q = 0; q < CPW; ++q) {
        const int j = wave + kWaves * q;
        if (j < jb && lane < jb) {
            Linv_lds[lane + j * jb] = xl[q];
            Lg[lane + (size_t)j * jb] = xl[q];
        }
    }
}

// ---- inverses of the diagonal blocks of U, after the factorisation -----------------------
// Only the refined mode's block solves use them.  Wave w inverts blocks w,
// w+4, ... (lanes = rows, bs <= 16) from the final factors in the slot, so no
// block barrier is needed; the result goes to the block store (ld = block width).
// Inverses of the bs x bs (bs <= 16) diagonal blocks of U (logical rows of the
// slot): block b to Ug + b bs^2, kb x kb column-major (ld kb).  Back
// substitution X = U^-1, one block per wavefront at a time: lane (i, g) =
// (lane % 16, lane / 16) holds X[i][4g .. 4g+3]; step k scales row k by
// 1 / U[k][k] and the rows above take row k from lane 16 g + k (one cross-lane
// read per complex entry, the four column groups in parallel).
template <typename T>
__device__ __forceinline__ void uinv_blocks(const cx<T>* __restrict__ S, int n, const int* rmap, int bs,
                                            cx<T>* __restrict__ Ug) {
    constexpr int B = 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 15, g = lane >> 4;
    const int nb = (n + bs - 1) / bs;
    for (int b = wave; b < nb; b += kWaves) {
        const int k0 = b * bs, kb = min(bs, n - k0);
        // rows beyond kb read row kb - 1 and are never stored or read by other rows
        const int pr = lds1<3>(rmap + k0 + min(i, kb - 1));
        cx<T> u[B], x[4];
#pragma unroll
        for (int k = 0; k < B; ++k) u[k] = ldc<1>(S + (pr + (size_t)(k0 + min(k, kb - 1)) * n));
        const cx<T> rd = crecip(ldc<1>(S + (pr + (size_t)(k0 + min(i, kb - 1)) * n)));
#pragma unroll
        for (int q = 0; q < 4; ++q) x[q] = i == 4 * g + q ? cmk<T>(T(1), T(0)) : czero<T>();
        // steps k = 15 .. 0; row k of each column group comes from lane k of its row of 16
        static_for([&](auto kc) {
            constexpr int k = B - 1 - decltype(kc)::value;
            if (k < kb) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (i == k) x[q] = cmul(x[q], rd);
                    const cx<T> xk = crowbcast<k>(x[q]);
                    if (i < k) cfms(x[q], u[k], xk);
                }
            }
        }, std::make_integer_sequence<int, B>{});
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = 4 * g + q;
            if (c < kb && i < kb) stc<1>(Ug + ((size_t)b * bs * bs + i + c * kb), x[q]);
        }
    }
}

// ---- block row and trailing update after one panel, wave-parallel ---------------------
// The trailing columns [j0+jb, n) are cut into chunks of kUw columns; wave w
// takes chunks w, w+4, ... and needs no block barrier:
//   U12 = Linv A12 on MFMA (A12 = logical rows [j0, j0+jb) of the chunk, all
//   of its loads in flight together), written to S and to the wave's LDS
//   buffer Uw (jb x kUw, row-major, ld kUwLd);
//   C -= L21 U12 for logical rows [j0+jb, n) in tiles of 64 rows, the next
//   tile's C loaded before the current tile's products.
// Both products are computed transposed so that the MFMA operands come
// straight from the column-major slot and LDS panel: D = A12^T Linv^T = U12^T
// and D = U12^T L21^T = (L21 U12)^T.  Complex products take 4 real MFMAs.
// jb <= 32.
// ASM: first block step, whose A12 and C have not been written to the slot
// yet: they are assembled from the operators (physical row r, column c of A(w)
// is element r + c n) and the results stored as usual.
template <typename T, bool ASM>
__device__ __forceinline__ void block_update(const SysDesc& sd, double w, double w2, cx<T>* __restrict__ S,
                                             int ld, const int* rmap, int j0, int jb, int n, int ncol,
                                             const cx<T>* P, int ldp, const cx<T>* Linv, cx<T>* Uw_all) {
    using M = Mfma<T>;
    using acc_t = typename M::acc;
    constexpr int KS = 8;  // k-steps of 4 for jb <= 32
    constexpr int kTA = 4; // 16-row MFMA tiles per C tile (64 rows)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int c0 = j0 + jb, m2 = n - c0;            // rows of the trailing update
    const int nch = (ncol - c0 + kUw - 1) / kUw;    // columns: matrix, then right-hand sides
    cx<T>* Uw = Uw_all + wave * jb * kUwLd;
    auto load_a12 = [&](int cb, int cw, cx<T> (&av)[KS]) {
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int k = 4 * s + lg;
            if (ASM && cb + li < n)
                av[s] = (k < jb && li < cw)
                            ? to_t<T>(assemble_at(sd, (size_t)lds1<3>(rmap + j0 + k) + (size_t)(cb + li) * ld, w, w2))
                            : czero<T>();
            else
                av[s] = (k < jb && li < cw) ? ldc<1>(S + (lds1<3>(rmap + j0 + k) + (cb + li) * ld))
                                            : czero<T>();
        }
    };
    auto load_c = [&](int r0, int cb, int cw, acc_t (&re)[kTA], acc_t (&im)[kTA], int (&pr)[kTA]) {
#pragma unroll
        for (int a = 0; a < kTA; ++a) {
            const int row = r0 + 16 * a + li;
            const bool ok = row < m2;
            pr[a] = ok ? lds1<3>(rmap + c0 + row) : 0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = M::row(lg, r);
                cx<T> v = czero<T>();
                if (ok && col < cw)
                    v = (ASM && cb + col < n)
                            ? to_t<T>(assemble_at(sd, (size_t)pr[a] + (size_t)(cb + col) * ld, w, w2))
                            : ldc<1>(S + (pr[a] + (cb + col) * ld));
                re[a][r] = v.re;
                im[a][r] = v.im;
            }
        }
    };
    cx<T> av[KS];
    for (int ch = wave; ch < nch; ch += kWaves) {
        const int cb = c0 + ch * kUw, cw = min(kUw, ncol - cb);
        load_a12(cb, cw, av);
        // ---- U12 = Linv A12 ----
        {
            acc_t ure[2], uim[2];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int r = 0; r < 4; ++r) ure[a][r] = uim[a][r] = T(0);
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                if (4 * s < jb) {
                    const int k = 4 * s + lg;
#pragma unroll
                    for (int a = 0; a < 2; ++a) {
                        const int j = 16 * a + li;
                        const cx<T> yv = (k < jb && j < jb) ? ldc<3>(Linv + j + k * jb) : czero<T>();
                        ure[a] = M::run(av[s].re, yv.re, ure[a]);
                        ure[a] = M::run(-av[s].im, yv.im, ure[a]);
                        uim[a] = M::run(av[s].re, yv.im, uim[a]);
                        uim[a] = M::run(av[s].im, yv.re, uim[a]);
                    }
                }
            }
            // lane holds U12[16a + li][M::row(lg, r)]
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const int row = 16 * a + li;
                if (row < jb) {
                    const int prow = lds1<3>(rmap + j0 + row);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int col = M::row(lg, r);
                        const cx<T> v = cmk<T>(ure[a][r], uim[a][r]);
                        stc<3>(Uw + row * kUwLd + col, v);
                        if (col < cw) stc<1>(S + (prow + (cb + col) * ld), v);
                    }
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        // ---- C -= L21 U12, tiles of 16 kTA rows, the next tile loaded before
        // the current tile's products ----
        acc_t cre[kTA], cim[kTA];
        int prow[kTA];
        load_c(0, cb, cw, cre, cim, prow);
        for (int r0 = 0; r0 < m2; r0 += 16 * kTA) {
            acc_t nre[kTA], nim[kTA];
            int nprow[kTA];
            const bool more = r0 + 16 * kTA < m2;
            if (more && !ASM) load_c(r0 + 16 * kTA, cb, cw, nre, nim, nprow);
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                if (4 * s < jb) {
                    const int k = 4 * s + lg;
                    const bool kok = k < jb;
                    const cx<T> uv = kok ? ldc<3>(Uw + k * kUwLd + li) : czero<T>();
#pragma unroll
                    for (int a = 0; a < kTA; ++a) {
                        if (r0 + 16 * a < m2) {
                            const int row = min(r0 + 16 * a + li, m2 - 1);
                            const cx<T> lv = kok ? ldc<3>(P + ((jb + row) + k * ldp)) : czero<T>();
                            cre[a] = M::run(-uv.re, lv.re, cre[a]);
                            cre[a] = M::run(uv.im, lv.im, cre[a]);
                            cim[a] = M::run(-uv.re, lv.im, cim[a]);
                            cim[a] = M::run(-uv.im, lv.re, cim[a]);
                        }
                    }
                }
            }
#pragma unroll
            for (int a = 0; a < kTA; ++a) {
                const int row = r0 + 16 * a + li;
                if (row < m2) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int col = M::row(lg, r);
                        if (col < cw) stc<1>(S + (prow[a] + (cb + col) * ld), cmk<T>(cre[a][r], cim[a][r]));
                    }
                }
            }
            if (more) {
                if (ASM) {
                    load_c(r0 + 16 * kTA, cb, cw, cre, cim, prow);
                } else {
#pragma unroll
                    for (int a = 0; a < kTA; ++a) {
                        cre[a] = nre[a];
                        cim[a] = nim[a];
                        prow[a] = nprow[a];
                    }
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ---- blocked right-looking factorisation in a global slot --------------------------------
// Rows are never moved in S: a pivot swaps two entries of rmap and the panel
// rows (registers), so no row-interchange pass touches global memory.  Row r
// of the LAPACK-layout factors is physical row rmap[r] of S.  The slot also
// holds the inverted diagonal blocks (Linv, Uinv; jb x jb each per panel),
// used by the triangular solves.
template <typename T, int MAXR>
__device__ __forceinline__ void factor_blocked(const SysDesc& s, double w, double w2, cx<T>* __restrict__ S,
                                               const Layout& lay, const Lds& L, int nr) {
    constexpr int CPW = panel_cols_per_wave((int)sizeof(cx<T>), MAXR);
    const int n = s.n, tid = threadIdx.x;
    const size_t nn = (size_t)n * n;
    // A(w) is assembled in FP64 and rounded to T where it is first needed: the
    // first panel's columns here, every other column by the first block update
    // (its A12 and C tiles come straight from the operators), so the matrix is
    // never written to the slot and read back unmodified
    const int nfirst = n * min(lay.jb, n);
#pragma unroll 16
    for (int e = tid; e < nfirst; e += kThreads) stc<1>(S + e, to_t<T>(assemble_at(s, e, w, w2)));
    // the first nr right-hand sides ride along as columns n .. n+nr-1: the
    // block updates perform their forward elimination (L y = P b)
    for (int e = tid; e < n * nr; e += kThreads) {
        const int r = e % n, g = e / n;
        stc<1>(S + (nn + (size_t)g * n + r), to_t<T>(ldc<1>(reinterpret_cast<const z64*>(s.F) + r + (size_t)g * s.ldf)));
    }
    for (int r = tid; r < n; r += kThreads) L.rmap[r] = r;
    if (tid == 0) L.sc[1] = 0;
    __syncthreads();

    const int jbmax = lay.jb;
    cx<T>* P = reinterpret_cast<cx<T>*>(L.main);
    cx<T>* Uw = reinterpret_cast<cx<T>*>(L.main + align16(panel_region_elems(n, jbmax, MAXR) * sizeof(cx<T>)));
    cx<T>* Linv = reinterpret_cast<cx<T>*>(reinterpret_cast<char*>(Uw) +
                                          align16((long long)kWaves * jbmax * kUwLd * sizeof(cx<T>)));
    cx<T>* lcol = P;  // the multiplier column lives in the panel region while the panel is in registers
    const size_t nblk = (size_t)((n + jbmax - 1) / jbmax);
    cx<T>* Lg = S + (size_t)n * (n + kMaxGroup);
    cx<T>* Ug = Lg + nblk * jbmax * jbmax;

    for (int j0 = 0, blk = 0; j0 < n; j0 += jbmax, ++blk) {
        const int jb = min(jbmax, n - j0), m = n - j0;
        panel_factor_reg<T, MAXR, CPW>(S, n, n, j0, jb, L.ipiv, L.rmap, L.sc, lcol, P);
        __syncthreads();
        invert_diag<T, CPW>(P, m, jb, Linv, Lg + (size_t)blk * jbmax * jbmax);
        __syncthreads();
        if (j0 == 0) block_update<T, true>(s, w, w2, S, n, L.rmap, j0, jb, n, n + nr, P, m, Linv, Uw);
        else block_update<T, false>(s, w, w2, S, n, L.rmap, j0, jb, n, n + nr, P, m, Linv, Uw);
        __syncthreads();
    }
    // the complex64 (refined) solves apply inverted U blocks; complex128 solves substitute
    if (sizeof(T) == 4) {
        uinv_blocks(S, n, L.rmap, jbmax, Ug);
        __syncthreads();
    }
}

// ---- triangular solves on LDS vectors w (n x gc, ld n) -----------------------------------
// LU: factors with logical row r at physical row rmap[r] (rmap == nullptr:
// identity).  Dg: LDS staging for one kSolveBlock^2 diagonal block, or nullptr
// when LU itself is in LDS.  Off-diagonal updates issue SolveRound loads per round.
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

// Loads per round of the triangular-solve products (16 complex values: what
// the register budget of either kernel leaves room for).
template <typename T>
struct SolveRound {
    static constexpr int K = 16;
};

template <int LUAS, typename T, bool CONJ, bool TRANS>
__device__ __forceinline__ void rect_update_as(const TriCtx<T>& t, int i, int k0, int kb, cx<T>* w,
                                               int gc) {
    // w[i] -= sum_{k<kb} op(LU)(i, k0+k) w[k0+k];  op = (TRANS ? element (k0+k, i) : (i, k0+k)),
    // conjugated when CONJ.  Every load of a round is issued before the first use.
    // LUAS: address space of the factors (1 = global slot, 0 = generic).
    constexpr int KR = SolveRound<T>::K;
    k0 = __builtin_amdgcn_readfirstlane(k0);
    kb = __builtin_amdgcn_readfirstlane(kb);
    gc = __builtin_amdgcn_readfirstlane(gc);
    const int n = __builtin_amdgcn_readfirstlane(t.n);
    cx<T> acc[kMaxGroup];
#pragma unroll
    for (int g = 0; g < kMaxGroup; ++g) acc[g] = czero<T>();
    const size_t pi = t.row(i);
    for (int kk = 0; kk < kb; kk += KR) {
        cx<T> l[KR];
#pragma unroll
        for (int u = 0; u < KR; ++u) {
            const int k = k0 + min(kk + u, kb - 1);
            l[u] = ldc<LUAS>(TRANS ? t.LU + t.row(k) + (size_t)i * t.ld : t.LU + pi + (size_t)k * t.ld);
        }
#pragma unroll
        for (int u = 0; u < KR; ++u) {
            if (kk + u < kb) {
                const cx<T> lu = CONJ ? cconj(l[u]) : l[u];
#pragma unroll
                for (int g = 0; g < kMaxGroup; ++g)
                    if (g < gc) cfma(acc[g], lu, ldc<3>(w + k0 + kk + u + (size_t)g * n));
            }
        }
    }
#pragma unroll
    for (int g = 0; g < kMaxGroup; ++g)
        if (g < gc) {
            cx<T> a = ldc<3>(w + i + (size_t)g * n);
            a.re -= acc[g].re;
            a.im -= acc[g].im;
            stc<3>(w + i + (size_t)g * n, a);
        }
}

// Factors in the global slot (blocked path, rmap set) or in LDS (small path).
template <typename T, bool CONJ, bool TRANS>
__device__ __forceinline__ void rect_update(const TriCtx<T>& t, int i, int k0, int kb, cx<T>* w,
                                            int gc) {
    if (t.rmap) rect_update_as<1, T, CONJ, TRANS>(t, i, k0, kb, w, gc);
    else rect_update_as<0, T, CONJ, TRANS>(t, i, k0, kb, w, gc);
}

// L unit lower, forward.
template <typename T>
__device__ __forceinline__ void lower_unit_solve(const TriCtx<T>& t, cx<T>* w, int gc) {
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
__device__ __forceinline__ void upper_solve(const TriCtx<T>& t, cx<T>* w, int gc) {
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
__device__ __forceinline__ void upper_h_solve(const TriCtx<T>& t, cx<T>* w, int gc) {
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
__device__ __forceinline__ void lower_h_unit_solve(const TriCtx<T>& t, cx<T>* w, int gc) {
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

// ---- pipelined blocked triangular solves (factors in the global slot) ----------------------
// Solves op(LU) on the LDS vectors w (n x gc, ld n) in 16-row blocks.  No
// global load of a block depends on the solution (neither its diagonal block
// or stored inverse nor the rows of its off-diagonal update), so block b+1's
// loads are issued before block b is solved and the factor slot's latency
// overlaps the LDS work.  Per block: the diagonal block (or its inverse) is
// staged in LDS from one prefetched element per thread, then solved (INV:
// y = Binv y, one lane per (row, column); else substitution, one wave per
// column, unit diagonal when UNIT), then every thread updates its row of the
// rows outside the block from its prefetched 16 entries (further rows of
// n > 256 load theirs on demand).  FWD: blocks in increasing order (lower
// triangular op), else decreasing.  TRANS: op(LU)(i, k) = conj(LU(k, i)).
// BF: the inverse-apply and off-diagonal loops without branches (entries
// beyond the block contribute zero, right-hand sides g >= gc repeat column
// gc - 1 and are discarded, all 16 entries of a row are loaded); more
// registers, used by the complex64 wide-panel kernel only.
template <typename T, bool FWD, bool INV, bool UNIT, bool TRANS, bool BF = false>
__device__ __forceinline__ void tri_solve_pipe(const TriCtx<T>& t, const cx<T>* __restrict__ binv, cx<T>* w,
                                               int gc) {
    constexpr int B = 16;
    constexpr bool CONJ = TRANS;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = __builtin_amdgcn_readfirstlane(t.n);
    gc = __builtin_amdgcn_readfirstlane(gc);
    const int ld = __builtin_amdgcn_readfirstlane(t.ld);
    const int bs = INV ? __builtin_amdgcn_readfirstlane(t.bs) : B;
    const int nb = (n + bs - 1) / bs;
    const cx<T>* __restrict__ LU = t.LU;
    cx<T>* Dg = t.Dg;
    const int i0 = tid;  // this thread's prefetched row
    const int i0c = min(i0, n - 1);
    const size_t pr0 = (size_t)lds1<3>(t.rmap + i0c);
    auto in_rows = [&](int i, int k0, int kb) { return FWD ? (i >= k0 + kb && i < n) : (i < k0); };
    auto load_diag = [&](int b) -> cx<T> {
        const int k0 = b * bs, kb = min(bs, n - k0);
        const int i = tid & (B - 1), j = tid >> 4;
        const int ic = min(i, kb - 1), jc = min(j, kb - 1);
        const bool ok = i < kb && j < kb;
        (void)ok;  // entries outside the kb x kb block are staged but never read
        if (INV) return ldc<1>(binv + ((size_t)b * bs * bs + ic + (size_t)jc * kb));
        return ldc<1>(LU + ((size_t)lds1<3>(t.rmap + k0 + ic) + (size_t)(k0 + jc) * ld));
    };
    auto load_rect = [&](int b, cx<T> (&r)[B]) {
        const int k0 = b * bs, kb = min(bs, n - k0);
#pragma unroll
        for (int u = 0; u < B; ++u) {
            // rows outside the update are loaded (clamped) but never used; entries
            // beyond kb (narrow blocks of the register-panel kernel) are not loaded
            if (BF || u < kb) {
                const int k = k0 + min(u, kb - 1);
                r[u] = TRANS ? ldc<1>(LU + ((size_t)lds1<3>(t.rmap + k) + (size_t)i0c * ld))
                             : ldc<1>(LU + (pr0 + (size_t)k * ld));
            }
        }
    };
    cx<T> dc, dn = czero<T>();
    cx<T> rc[B], rn[B];
    {
        const int b = FWD ? 0 : nb - 1;
        dc = load_diag(b);
        load_rect(b, rc);
    }
    for (int st = 0; st < nb; ++st) {
        const int b = FWD ? st : nb - 1 - st;
        const int k0 = b * bs, kb = min(bs, n - k0);
        stc<3>(Dg + tid, dc);  // element (tid % 16, tid / 16), ld 16
        if (st + 1 < nb) {
            const int bn = FWD ? b + 1 : b - 1;
            dn = load_diag(bn);
            load_rect(bn, rn);
        }
        __syncthreads();
        if (INV) {
            if (tid < B * gc) {
                const int r = tid & (B - 1), g = tid >> 4;
                cx<T>* y = w + (size_t)g * n + k0;
                cx<T> acc = czero<T>();
#pragma unroll
                for (int k = 0; k < B; ++k) {
                    if (BF) {
                        cx<T> d = ldc<3>(Dg + (r + k * B));
                        d.re = k < kb ? d.re : T(0);
                        d.im = k < kb ? d.im : T(0);
                        cfma(acc, d, ldc<3>(y + min(k, kb - 1)));
                    } else if (k < kb) {
                        cfma(acc, ldc<3>(Dg + (r + k * B)), ldc<3>(y + k));
                    }
                }
                if (r < kb) stc<3>(y + r, acc);  // one wave: every read above precedes the writes
            }
        } else {
            for (int g = wave; g < gc; g += kWaves) {
                cx<T>* y = w + (size_t)g * n + k0;
                cx<T> v = lane < kb ? ldc<3>(y + lane) : czero<T>();
                if (FWD) {
                    for (int k = 0; k < kb; ++k) {
                        if (!UNIT && lane == k) {
                            const cx<T> d = ldc<3>(Dg + (k + k * B));
                            v = cdiv(v, CONJ ? cconj(d) : d);
                        }
                        const cx<T> vk = cbcast(v, k);
                        if (lane > k && lane < kb) {
                            const cx<T> a = TRANS ? ldc<3>(Dg + (k + lane * B)) : ldc<3>(Dg + (lane + k * B));
                            cfms(v, CONJ ? cconj(a) : a, vk);
                        }
                    }
                } else {
                    for (int k = kb - 1; k >= 0; --k) {
                        if (!UNIT && lane == k) {
                            const cx<T> d = ldc<3>(Dg + (k + k * B));
                            v = cdiv(v, CONJ ? cconj(d) : d);
                        }
                        const cx<T> vk = cbcast(v, k);
                        if (lane < k) {
                            const cx<T> a = TRANS ? ldc<3>(Dg + (k + lane * B)) : ldc<3>(Dg + (lane + k * B));
                            cfms(v, CONJ ? cconj(a) : a, vk);
                        }
                    }
                }
                if (lane < kb) stc<3>(y + lane, v);
            }
        }
        __syncthreads();
        // off-diagonal rows: w[i] -= op(LU)(i, k0:k0+kb) w[k0:k0+kb]
        if (in_rows(i0, k0, kb)) {
            cx<T> acc[kMaxGroup];
#pragma unroll
            for (int g = 0; g < kMaxGroup; ++g) acc[g] = czero<T>();
#pragma unroll
            for (int u = 0; u < B; ++u) {
                if (BF) {
                    cx<T> a = CONJ ? cconj(rc[u]) : rc[u];
                    a.re = u < kb ? a.re : T(0);
                    a.im = u < kb ? a.im : T(0);
                    const int uc = min(u, kb - 1);
#pragma unroll
                    for (int g = 0; g < kMaxGroup; ++g)
                        cfma(acc[g], a, ldc<3>(w + (k0 + uc + (size_t)min(g, gc - 1) * n)));
                } else if (u < kb) {
                    const cx<T> a = CONJ ? cconj(rc[u]) : rc[u];
#pragma unroll
                    for (int g = 0; g < kMaxGroup; ++g)
                        if (g < gc) cfma(acc[g], a, ldc<3>(w + (k0 + u + (size_t)g * n)));
                }
            }
#pragma unroll
            for (int g = 0; g < kMaxGroup; ++g)
                if (g < gc) {
                    cx<T> a = ldc<3>(w + (i0 + (size_t)g * n));
                    a.re -= acc[g].re;
                    a.im -= acc[g].im;
                    stc<3>(w + (i0 + (size_t)g * n), a);
                }
        }
        for (int i = i0 + kThreads; i < n; i += kThreads)
            if (in_rows(i, k0, kb)) rect_update_as<1, T, CONJ, TRANS>(t, i, k0, kb, w, gc);
        dc = dn;
#pragma unroll
        for (int u = 0; u < B; ++u) rc[u] = rn[u];
    }
    __syncthreads();
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
// BF: the real-operator rounds without branches (columns beyond n contribute
// zero, right-hand sides g >= gc repeat column gc - 1 and are discarded), so
// the compiler interleaves a round's loads with its arithmetic; it needs more
// registers and is used by the complex64 wide-panel kernel only.
// RG: rs may be global memory (generic stores), else it is LDS.
template <bool ADJ, int RC, bool BF, bool RG = false>
__device__ __forceinline__ void residual(const SysDesc& s, double w, double w2, const double2* bg, int ldb,
                         const z64* bl, const z64* xs, z64* rs, int gc, bool need_anorm,
                         double* red, double (&out)[kRedSlots]) {
    const int n = __builtin_amdgcn_readfirstlane(s.n), tid = threadIdx.x;
    gc = __builtin_amdgcn_readfirstlane(gc);
    // pass 1: r = b - op(A) x (and the row sums of |op(A)|), RC columns of
    // the operators in flight per round
    double amax = 0.0;
    for (int i = tid; i < n; i += kThreads) {
        z64 acc[kMaxGroup];
#pragma unroll
        for (int g = 0; g < kMaxGroup; ++g)
            acc[g] = g < gc ? (bg ? ldc<1>(reinterpret_cast<const z64*>(bg) + i + (size_t)g * ldb)
                                  : ldc<3>(bl + i + (size_t)g * n))
                            : czero<double>();
        double asum = 0.0;
        for (int j0 = 0; j0 < n; j0 += RC) {
            if (s.real_ops) {
                const double* __restrict__ Kp = s.K;
                const double* __restrict__ Mp = s.M;
                const double* __restrict__ Dp = s.D;
                double kv[RC], mv[RC], dv[RC];
#pragma unroll
                for (int u = 0; u < RC; ++u) {
                    const int j = min(j0 + u, n - 1);
                    const int e = ADJ ? j + i * n : i + j * n;
                    kv[u] = lds1<1>(Kp + e);
                    mv[u] = lds1<1>(Mp + e);
                    dv[u] = s.has_D ? lds1<1>(Dp + e) : 0.0;
                }
                if (BF) {
                    // branch-free: columns beyond n contribute zero, and columns
                    // g >= gc repeat column gc - 1 (their sums are discarded)
    #pragma unroll
                    for (int u = 0; u < RC; ++u) {
                        const bool in = j0 + u < n;
                        const int j = min(j0 + u, n - 1);
                        z64 au = cmk<double>(kv[u] - w2 * mv[u], (ADJ ? -1.0 : 1.0) * (s.eta * kv[u] + w * dv[u]));
                        au.re = in ? au.re : 0.0;
                        au.im = in ? au.im : 0.0;
                        if (need_anorm) asum += zabs(au);
    #pragma unroll
                        for (int g = 0; g < kMaxGroup; ++g)
                            cfms(acc[g], au, ldc<3>(xs + j + (size_t)min(g, gc - 1) * n));
                    }
                } else {
    #pragma unroll
                    for (int u = 0; u < RC; ++u) {
                        if (j0 + u < n) {
                            const z64 au = cmk<double>(kv[u] - w2 * mv[u], (ADJ ? -1.0 : 1.0) * (s.eta * kv[u] + w * dv[u]));
                            if (need_anorm) asum += zabs(au);
    #pragma unroll
                            for (int g = 0; g < kMaxGroup; ++g)
                                if (g < gc) cfms(acc[g], au, ldc<3>(xs + j0 + u + (size_t)g * n));
                        }
                    }
                }
                continue;
            }
            // complex operators: 4 columns per round
            for (int j1 = j0; j1 < j0 + RC && j1 < n; j1 += 4) {
                z64 a[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int j = min(j1 + u, n - 1);
                    a[u] = ADJ ? cconj(assemble_at(s, (size_t)j + (size_t)i * n, w, w2))
                               : assemble_at(s, (size_t)i + (size_t)j * n, w, w2);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (j1 + u < n) {
                        if (need_anorm) asum += zabs(a[u]);
#pragma unroll
                        for (int g = 0; g < kMaxGroup; ++g)
                            if (g < gc) cfms(acc[g], a[u], ldc<3>(xs + j1 + u + (size_t)g * n));
                    }
                }
            }
        }
        amax = nmax(amax, asum);
#pragma unroll
        for (int g = 0; g < kMaxGroup; ++g)
            if (g < gc) {
                if (RG) rs[i + (size_t)g * n] = acc[g];
                else stc<3>(rs + i + (size_t)g * n, acc[g]);
            }
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
template <typename T, bool ADJ, bool BF = false, bool RG = false>
__device__ __forceinline__ int solve_group(const SysDesc& s, double w, double w2, const TriCtx<T>& tri,
                           const int* perm, cx<T>* wv, z64* xs, z64* rs, int gc,
                           const double2* bg, int ldb, const z64* bl, bool refine, bool want_berr,
                           double tol, int max_iter, double* anorm,
                           double (&berr_out)[kMaxGroup], bool* allconv, double* red,
                           bool fwd_done = false) {
    const int n = s.n, tid = threadIdx.x;
    // fwd_done: wv already holds L^-1 P b (eliminated with the factorisation),
    // so the first solve is the U solve only
    auto solve_t = [&]() {
        if (tri.rmap) {  // factors in the global slot: pipelined block solves
            if (!ADJ) {
                if (fwd_done) fwd_done = false;
                else tri_solve_pipe<T, true, true, true, false, BF>(tri, tri.Linv, wv, gc);
                if (tri.inv_u) tri_solve_pipe<T, false, true, false, false, BF>(tri, tri.Uinv, wv, gc);
                else tri_solve_pipe<T, false, false, false, false>(tri, nullptr, wv, gc);
            } else {
                tri_solve_pipe<T, true, false, false, true>(tri, nullptr, wv, gc);
                tri_solve_pipe<T, false, false, true, true>(tri, nullptr, wv, gc);
            }
        } else if (!ADJ) {
            if (fwd_done) fwd_done = false;
            else lower_unit_solve(tri, wv, gc);
            upper_solve(tri, wv, gc);
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
        residual<ADJ, ResCols<T>::K, BF, RG>(s, w, w2, bg, ldb, bl, xs, rs, gc, *anorm < 0.0, red, out);
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
__device__ __forceinline__ void adjoint_point(const SysDesc& s, int fidx, double w, double w2, const TriCtx<T>& tri,
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

}  // namespace stc

#include "stc_wide.h"

namespace stc {

// ---- fused output projection -------------------------------------------------------------
// Y(:, g0 + g) = C x_g for the gc solved columns x_g = xs + g n (FP64) of one
// point: nout x gc outputs, each a length-n complex dot product with the
// nout x n matrix C (ld nout).  Many outputs: one per thread, the reads of C
// coalesced over o.  Few outputs (nout << n): one per wavefront, its lanes
// split the sum over i and reduce by butterfly.  Only Y leaves the workgroup.
__device__ __forceinline__ void project_cols(const SysDesc& s, int fidx, const z64* xs, int g0, int gc) {
    const int n = s.n, nout = s.nout, tid = threadIdx.x, q = nout * gc;
    double2* Y = s.Y + ((size_t)fidx * s.nrhs + g0) * nout;
    if (q >= kThreads / 2) {
        for (int e = tid; e < q; e += kThreads) {
            const int o = e % nout;
            const z64* x = xs + (size_t)(e / nout) * n;
            double re = 0.0, im = 0.0;
            for (int i = 0; i < n; ++i) {
                const double2 c = s.C[o + (size_t)i * nout];
                const z64 v = x[i];
                re = fma(c.x, v.re, fma(-c.y, v.im, re));
                im = fma(c.x, v.im, fma(c.y, v.re, im));
            }
            Y[e] = make_double2(re, im);
        }
    } else {
        const int lane = tid & 63;
        for (int e = tid >> 6; e < q; e += kWaves) {
            const int o = e % nout;
            const z64* x = xs + (size_t)(e / nout) * n;
            double re = 0.0, im = 0.0;
            for (int i = lane; i < n; i += 64) {
                const double2 c = s.C[o + (size_t)i * nout];
                const z64 v = x[i];
                re = fma(c.x, v.re, fma(-c.y, v.im, re));
                im = fma(c.x, v.im, fma(c.y, v.re, im));
            }
            for (int off = 32; off > 0; off >>= 1) {
                re += __shfl_xor(re, off, 64);
                im += __shfl_xor(im, off, 64);
            }
            if (lane == 0) Y[e] = make_double2(re, im);
        }
    }
}

// NaN into X and Y of a point that could not be solved.
__device__ __forceinline__ void nan_outputs(const SysDesc& s, int fidx) {
    const int tid = threadIdx.x;
    const double2 qn = make_double2(__builtin_nan(""), __builtin_nan(""));
    if (s.X) {
        double2* X = s.X + (size_t)fidx * s.ldx * s.nrhs;
        for (int e = tid; e < s.n * s.nrhs; e += kThreads) X[e % s.n + (size_t)(e / s.n) * s.ldx] = qn;
    }
    if (s.Y) {
        double2* Y = s.Y + (size_t)fidx * s.nout * s.nrhs;
        for (int e = tid; e < s.nout * s.nrhs; e += kThreads) Y[e] = qn;
    }
}

// ---- one point ---------------------------------------------------------------------------
// WIDE != 0: blocked systems use the wide-panel factorisation (factor_wide, LDS
// panel of lay.jb columns, 16 x 16 inverted diagonal blocks), WIDE == 2 with
// Gauss 3M products in its trailing update, WIDE == 3 with the two-level
// update (lay.jo columns per step); WIDE == 0: the register panel
// (factor_blocked, lay.jb x lay.jb blocks).
template <typename T, int MAXR, int WIDE>
__device__ __forceinline__ void process_point(const SysDesc& s, int fidx, const KParams& kp, char* smem) {
    const int n = s.n, tid = threadIdx.x, nrhs = s.nrhs;
    const double w = s.freq ? kTwoPi * s.freq[fidx] : 0.0, w2 = w * w;
    Layout lay;
    lay.small = s.lay_small;
    lay.jb = s.lay_jb;
    lay.cw = s.lay_cw;
    lay.G = s.lay_G;
    lay.lds = 0;
    lay.jo = s.lay_jo;
    const Lds L = carve(smem, n);
    const bool refine = (kp.flags & kRefine) != 0;
    const bool want_berr = refine || (kp.flags & kWantBerr) || (s.berr != nullptr);

    const cx<T>* LU;
    char* solvep;
    int fused = 0;  // right-hand sides whose forward elimination the factorisation did
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
        const bool bad_jb = WIDE ? (lay.jb < kSB || lay.jb % kSB != 0 || lay.jb > kSB * WideCfg<T>::NB ||
                                    (WIDE == 3 && lay.jo > lay.jb &&
                                     (lay.jo % lay.jb != 0 || lay.jo > kSB * WideCfg<T>::NB)))
                                 : (lay.jb > kWaves * panel_cols_per_wave((int)sizeof(cx<T>), MAXR) || lay.jb > 32);
        if (kp.scratch == nullptr ||
            kp.slot_elems < slot_elems_for(n, WIDE ? kSB : lay.jb) +
                                (WIDE && s.lay_rsg ? slot_rs_elems(n, sizeof(cx<T>)) : 0) ||
            bad_jb) {
            // no factor slot, or a panel wider than the register panel (the host
            // plans neither): report, write NaN, touch nothing else
            nan_outputs(s, fidx);
            if (tid == 0) s.info[fidx] = -2;
            __syncthreads();
            return;
        }
        cx<T>* S = reinterpret_cast<cx<T>*>(kp.scratch) + (size_t)blockIdx.x * kp.slot_elems;
        // right-hand sides eliminated with the factorisation (the first group of
        // a forward sweep)
        const int nr = (kp.flags & (kAdjoint | kFactorOut)) ? 0 : min(nrhs, lay.G);
        if (WIDE) factor_wide<T, MAXR, WIDE == 2, WIDE == 3>(s, w, w2, S, lay, L, nr);
        else factor_blocked<T, MAXR>(s, w, w2, S, lay, L, nr);
        fused = nr;
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

    double2* X = s.X ? s.X + (size_t)fidx * s.ldx * nrhs : nullptr;
    if (info != 0) {
        nan_outputs(s, fidx);
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
    // wide path: the FP64 residuals of a group go to the slot's tail when the
    // host planned them there (lay_rsg: the group would not fit the LDS with them)
    // (only the kernels for n > 256 carry the second copy of the solve loop)
    constexpr bool kRsg = WIDE != 0 && MAXR >= 8;
    const bool rsg = kRsg && !lay.small && s.lay_rsg;
    const int nvec = ((kp.flags & kAdjoint) ? 4 : 2) - (rsg ? 1 : 0);
    cx<T>* wv = reinterpret_cast<cx<T>*>(solvep);
    z64* xs = reinterpret_cast<z64*>(solvep + align16((long long)n * G * sizeof(cx<T>)));
    z64* rs = reinterpret_cast<z64*>(reinterpret_cast<char*>(xs) + align16((long long)n * G * 16));
    z64* rs_slot = nullptr;
    if (rsg) {
        const size_t nb16 = (size_t)((n + kSB - 1) / kSB);
        const size_t off = ((size_t)n * (n + kMaxGroup) + 2 * nb16 * kSB * kSB + 1) & ~(size_t)1;
        rs_slot = reinterpret_cast<z64*>(const_cast<cx<T>*>(LU) + off);
    }
    char* after = reinterpret_cast<char*>(xs) + nvec * align16((long long)n * G * 16);
    TriCtx<T> tri;
    tri.LU = LU;
    tri.ld = n;
    tri.n = n;
    tri.rmap = lay.small ? nullptr : L.rmap;
    tri.Dg = lay.small ? nullptr : reinterpret_cast<cx<T>*>(after);
    const int bs = WIDE ? kSB : lay.jb;  // inverted diagonal blocks of the slot
    tri.bs = bs;
    tri.Linv = lay.small ? nullptr : LU + (size_t)n * (n + kMaxGroup);
    tri.Uinv = lay.small ? nullptr : LU + (size_t)n * (n + kMaxGroup) + (size_t)((n + bs - 1) / bs) * bs * bs;
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
        const bool fwd = g0 == 0 && fused == gc;
        for (int e = tid; e < n * gc; e += kThreads) {
            const int r = e % n, g = e / n;
            // forward-eliminated columns: y = L^-1 P b, row r at physical row rmap[r]
            wv[e] = fwd ? LU[(size_t)L.rmap[r] + (size_t)(n + g) * n]
                        : to_t<T>(ld_z(Fg + L.rmap[r] + (size_t)g * s.ldf));
        }
        __syncthreads();
        double be[kMaxGroup];
        bool conv;
        // branch-free residual rounds in the complex64 wide-panel kernel (the
        // others spill with them); residuals in LDS, or in the slot (a copy of
        // the loop with generic accesses to them)
        constexpr bool BF = WIDE != 0 && sizeof(T) == 4;
        if (BF) __builtin_amdgcn_s_setprio(2);
        int it;
        if (kRsg && rsg)
            it = solve_group<T, false, BF, kRsg>(s, w, w2, tri, L.rmap, wv, xs, rs_slot, gc, Fg, s.ldf,
                                                 nullptr, refine, want_berr, kp.tol, kp.max_iter,
                                                 &anorm, be, &conv, L.red, fwd);
        else
            it = solve_group<T, false, BF, false>(s, w, w2, tri, L.rmap, wv, xs, rs, gc, Fg, s.ldf,
                                                  nullptr, refine, want_berr, kp.tol, kp.max_iter,
                                                  &anorm, be, &conv, L.red, fwd);
        if (BF) __builtin_amdgcn_s_setprio(0);
        if (refine && !conv) ok = false;
        if (it > it_max) it_max = it;
        if (X)
            for (int e = tid; e < n * gc; e += kThreads) {
                const int r = e % n, g = e / n;
                st_z(X + r + (size_t)(g0 + g) * s.ldx, xs[e]);
            }
        if (s.C) project_cols(s, fidx, xs, g0, gc);
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
// wg_per_cu(sizeof(cx<T>)) workgroups of 4 wavefronts per CU: the register
// file is split that many ways (at most 168 registers per wavefront for three,
// 256 for two).  The wide-panel kernel (WIDE) runs one or two per CU (its LDS
// panel decides) with up to 256 registers.
template <typename T, int MAXR, int WIDE>
__global__ void __launch_bounds__(kThreads)
    __attribute__((amdgpu_waves_per_eu(WIDE ? 2 : wg_per_cu(sizeof(cx<T>)), WIDE ? 2 : wg_per_cu(sizeof(cx<T>)))))
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
        SysDesc s = kp.sys ? load_uniform(kp.sys + sysi) : load_uniform(&kp.one);
        if (s.op_stride) {  // one explicit matrix per frequency index
            const long long o = (long long)fidx * s.op_stride;
            s.K += o;
            s.M += o;
            if (s.has_D) s.D += o;
        }
        if (s.f_stride) s.F += (long long)fidx * s.f_stride;
        process_point<T, MAXR, WIDE>(s, fidx, kp, smem);
        __syncthreads();
    }
}

}  // namespace stc
