// stc_wide.h -- wide-panel blocked LU of the sweep kernel (gfx950).
//
// The blocked path of sweep_kernel<T, MAXR, 1> (<T, MAXR, 2>: Gauss 3M
// products in the trailing update; <T, MAXR, 3>: two-level update).  It factors A(w) in the workgroup's
// global factor slot, right-looking, with panels of jb = 16..64 columns held
// in LDS (one or two workgroups per CU), and cuts the trailing-matrix traffic
// by jb / 16 against a 16-column panel:
//
//   panel (m = n - j0 rows x jb columns, LDS, logical row order, assembled
//   from the operators at j0 = 0):
//     per 16-column sub-panel: pivoted LU with the rows distributed over the
//     threads (thread t holds rows t + 256 u in registers, one barrier per
//     column), rows permuted in the LDS panel afterwards; the 16 x 16
//     unit-lower diagonal block inverted (one entry per thread); the panel's
//     remaining columns updated on MFMA (U rows = Linv A, rows below -= L21 U);
//   trailing columns (matrix, then the first right-hand sides), 16 per chunk,
//     one wavefront per chunk, no block barrier:
//     U12 = L11^{-1} A12 by block forward substitution on MFMA in registers
//     (16-row blocks: acc_b -= L_ba U_a, U_b = Linv_b acc_b), U12 to the slot,
//     then C -= L21 U12 over tiles of 16 TA rows, the next tile's C loaded
//     before the current tile's products (assembled from the operators at the
//     first panel).
//
// Rows never move in the slot (row map, as in the 16-column path); only the
// LDS panel is kept in logical order.  Complex products take four real MFMAs
// (three in the Gauss form).
//
// MFMA operand layout used throughout (16x16x4, one wavefront; li = lane % 16,
// lg = lane / 16): lane supplies A[li][k-slot lg] and B[k-slot lg][li]; the
// accumulator register r holds D[row(lg, r)][li].  The k index of step s of a
// 16-block is kk(lg, s) = row(lg, s), so a block's accumulator is directly the
// B (or A) operand of the next product: no register shuffles.
#pragma once

namespace stc {

// Largest number of 16-row blocks of a wide panel kept in registers.
constexpr int kPB = 8;  // panel columns loaded per batch

template <typename T>
struct WideCfg {
    static constexpr int NB = wide_jb_max(2 * sizeof(T)) / kSB;  // jb <= 64 (complex64), 32 (complex128)
    static constexpr int TA = sizeof(T) == 4 ? 4 : 2;  // 16-row MFMA tiles per C tile
};

// d += a * b (complex; a from the lane's A operand, b its B operand).
template <typename T>
__device__ __forceinline__ void cmfma(typename Mfma<T>::acc& dre, typename Mfma<T>::acc& dim, cx<T> a,
                                      T bre, T bim) {
    using M = Mfma<T>;
    dre = M::run(a.re, bre, dre);
    dre = M::run(-a.im, bim, dre);
    dim = M::run(a.re, bim, dim);
    dim = M::run(a.im, bre, dim);
}

// Gauss (3M) form of d += a * b over three accumulators: x += a_r b_r,
// y += a_i b_i, z += (a_r + a_i)(b_r + b_i); with x, y, z started at (d_r, 0,
// d_r + d_i) the result is d_r = x - y, d_i = z - x - y.  Three real products
// instead of four.
template <typename T>
__device__ __forceinline__ void cmfma3(typename Mfma<T>::acc& x, typename Mfma<T>::acc& y,
                                       typename Mfma<T>::acc& z, cx<T> a, T asum, T bre, T bim) {
    using M = Mfma<T>;
    x = M::run(a.re, bre, x);
    y = M::run(a.im, bim, y);
    z = M::run(asum, bre + bim, z);
}

template <typename T>
__device__ __forceinline__ void acc_zero(typename Mfma<T>::acc& a) {
#pragma unroll
    for (int r = 0; r < 4; ++r) a[r] = T(0);
}


// ---- branch-free pivot search of one wavefront -----------------------------------------------
// A candidate row is the pair (hk, lo), compared lexicographically: hk orders
// |re| + |im| with LAPACK izamax's NaN rule (as pivot_key: a NaN in the first
// eligible row wins, a NaN elsewhere is never chosen; 0 = not a candidate),
// lo = (1023 - logical position) << 10 | gather position, so ties go to the
// first row in LAPACK order and the winner's location rides along.
__device__ __forceinline__ unsigned long long pivot_ukey(float v, bool first) {
    return v == v ? (unsigned long long)__float_as_uint(v) + 2ull : (first ? 0x7F800002ull : 1ull);
}
__device__ __forceinline__ unsigned long long pivot_ukey(double v, bool first) {
    return v == v ? (unsigned long long)__double_as_longlong(v) + 2ull : (first ? 0x7FF0000000000002ull : 1ull);
}
template <int CTRL, int RM>
__device__ __forceinline__ unsigned dppu(unsigned v) {
    return (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, RM, 0xF, false);
}
template <int CTRL, int RM>
__device__ __forceinline__ void argmax_step(unsigned long long& hk, unsigned& lo) {
    const unsigned long long ok = ((unsigned long long)dppu<CTRL, RM>((unsigned)(hk >> 32)) << 32) |
                                  dppu<CTRL, RM>((unsigned)hk);
    const unsigned ol = dppu<CTRL, RM>(lo);
    const bool better = ok > hk || (ok == hk && ol > lo);
    hk = better ? ok : hk;
    lo = better ? ol : lo;
}
constexpr int kDppRowBcast15 = 0x142;  // lane 15 of each row -> the next row
constexpr int kDppRowBcast31 = 0x143;  // lane 31 -> rows 2 and 3
// Maximum over the wave, returned in every lane: four DPP steps give each
// row of 16 lanes its maximum, two row broadcasts carry it to lane 63.
__device__ __forceinline__ void wave_argmax_packed(unsigned long long& hk, unsigned& lo) {
    argmax_step<kDppXor1, 0xF>(hk, lo);
    argmax_step<kDppXor2, 0xF>(hk, lo);
    argmax_step<kDppHalfMirror, 0xF>(hk, lo);
    argmax_step<kDppMirror, 0xF>(hk, lo);
    argmax_step<kDppRowBcast15, 0xA>(hk, lo);
    argmax_step<kDppRowBcast31, 0xC>(hk, lo);
    const unsigned h = (unsigned)__builtin_amdgcn_readlane((int)(hk >> 32), 63);
    const unsigned l = (unsigned)__builtin_amdgcn_readlane((int)hk, 63);
    hk = ((unsigned long long)h << 32) | l;
    lo = (unsigned)__builtin_amdgcn_readlane((int)lo, 63);
}
// complex64 keys fit one 64-bit word: the 32-bit magnitude key above the
// 20-bit position word (lo < 2^20), so the lexicographic order is the integer
// order and a DPP step moves two dwords instead of three.
template <int CTRL, int RM>
__device__ __forceinline__ void argmax_step64(unsigned long long& v) {
    const unsigned long long o = ((unsigned long long)dppu<CTRL, RM>((unsigned)(v >> 32)) << 32) |
                                 dppu<CTRL, RM>((unsigned)v);
    v = o > v ? o : v;
}
__device__ __forceinline__ void wave_argmax64(unsigned long long& v) {
    argmax_step64<kDppXor1, 0xF>(v);
    argmax_step64<kDppXor2, 0xF>(v);
    argmax_step64<kDppHalfMirror, 0xF>(v);
    argmax_step64<kDppMirror, 0xF>(v);
    argmax_step64<kDppRowBcast15, 0xA>(v);
    argmax_step64<kDppRowBcast31, 0xC>(v);
    const unsigned h = (unsigned)__builtin_amdgcn_readlane((int)(v >> 32), 63);
    const unsigned l = (unsigned)__builtin_amdgcn_readlane((int)v, 63);
    v = ((unsigned long long)h << 32) | l;
}
__device__ __forceinline__ void stc1i(int* p, int v) { *(__attribute__((address_space(3))) int*)(p) = v; }
__device__ __forceinline__ void stc1u(unsigned* p, unsigned v) { *(__attribute__((address_space(3))) unsigned*)(p) = v; }

// Rows moved by a sub-panel's pivoting (the list mvl, sc[12] entries): their
// entries in the other panel columns and in the row map follow; every source
// is read before any destination is written.
template <typename T>
__device__ __forceinline__ void subpanel_moves(cx<T>* __restrict__ P, int ldp, int jbl, int c0, int sbw, int j0,
                                               int* rmap, int* sc, int2* mvl) {
    const int tid = threadIdx.x;
    __syncthreads();
    // permute the moved rows of the other panel columns and of the row map:
    // read every source, then write every destination
    const int cnt = uniform(lds1<3>(sc + 12));
    const int nother = jbl - sbw;  // panel columns outside the sub-panel
    const int per = nother + 1;    // + the row-map entry
    const int total = cnt * per;
    constexpr int kMv = 8;         // cnt <= 2 sbw = 32, per <= 49: <= 1568 = 6.1 per lane
    cx<T> tv[kMv];
    int ti[kMv];
#pragma unroll
    for (int u = 0; u < kMv; ++u) {
        const int e = tid + kThreads * u;
        if (e < total) {
            const int i = e / per, k = e - i * per;
            const int2 mvp = mvl[i];
            if (k == nother) {
                ti[u] = rmap[j0 + mvp.x];
            } else {
                const int c = k < c0 ? k : k + sbw;
                tv[u] = ldc<3>(P + (mvp.x + (size_t)c * ldp));
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kMv; ++u) {
        const int e = tid + kThreads * u;
        if (e < total) {
            const int i = e / per, k = e - i * per;
            const int2 mvp = mvl[i];
            if (k == nother) {
                rmap[j0 + mvp.y] = ti[u];
            } else {
                const int c = k < c0 ? k : k + sbw;
                stc<3>(P + (mvp.y + (size_t)c * ldp), tv[u]);
            }
        }
    }
    __syncthreads();
}

// ---- 16-column sub-panel of the LDS panel ------------------------------------------------
// Columns [c0, c0 + sbw) of P (m x jb, ld m, logical row order), rows
// [c0, m).  Every thread holds whole rows of the sub-panel (rows
// tid + 256 t) in registers, x[c] = column c.
// Per column one barrier: each wave finds its best candidate row (packed
// keys, DPP), the winning lane publishes that row's 16 values and the wave's
// key in one of two alternating LDS buffers; after the barrier every thread
// picks the best of the four candidates (the same in all), takes the pivot
// row from LDS, scales its own rows' multipliers and applies the rank-1
// update.  Rows stay in their registers; pivots only change their logical
// positions lpos.  Afterwards the sub-panel is stored at its new rows and the
// moved rows of the other panel columns and of rmap[j0 + .] are permuted to
// match (moves listed in mvl).  LAPACK zgetf2 semantics: pivot = first row of
// largest |re| + |im| (izamax, NaN rule of pivot_key), zero pivot: no scaling,
// info = first zero column.  ipiv[j0 + c0 + jj] = 0-based pivot row;
// sc[1] = first zero pivot (1-based).
template <typename T, int MAXR>
__device__ __forceinline__ void subpanel_factor(cx<T>* __restrict__ P, int m, int ldp, int jbl, int c0, int sbw, int j0,
                                                int* ipiv, int* rmap, int* sc, cx<T>* rowbuf, int2* mvl) {
    constexpr int RM = MAXR >= 4 ? MAXR / 4 : 1;  // rows per thread
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int mq = m - c0;
    cx<T>* Q = P + c0 + (size_t)c0 * ldp;  // sub-panel origin, ld ldp
    // candidate keys after the row buffers: 2 buffers x 4 waves x {key lo, key hi, lo, -}
    unsigned* cand = reinterpret_cast<unsigned*>(rowbuf + 2 * kWaves * kSB);
    cx<T> x[kSB][RM];
    int lpos[RM];
#pragma unroll
    for (int t = 0; t < RM; ++t) {
        const int r = tid + kThreads * t;
        lpos[t] = r < mq ? r : 0x7fffffff;
#pragma unroll
        for (int c = 0; c < kSB; ++c)
            x[c][t] = ldz<3>(Q + (min(r, mq - 1) + (size_t)min(c, sbw - 1) * ldp), c < sbw && r < mq);
    }
    if (tid == 0) stc1i(sc + 12, 0);
    // at most two rows a thread: the column loop unrolled (static register
    // indices, no rotation; n = 256: +3.3 %); four rows: a loop that rotates x
    // so that x[0] is the current column (unrolled, it spills)
    if constexpr (RM <= 2) {
        auto column = [&](auto jc) {
            constexpr int jj = decltype(jc)::value;
            if (jj >= sbw) return;
            const int pb = jj & 1;
            cx<T>* rb = rowbuf + pb * kWaves * kSB;
            unsigned* cb = cand + pb * kWaves * 4;
            unsigned long long hk = 0;
            unsigned lo = 0;
#pragma unroll
            for (int t = 0; t < RM; ++t) {
                const bool elig = lpos[t] >= jj && lpos[t] != 0x7fffffff;
                const unsigned long long k = elig ? pivot_ukey(cabs1(x[jj][t]), lpos[t] == jj) : 0ull;
                const unsigned l = elig ? ((unsigned)(1023 - lpos[t]) << 10) | (unsigned)(tid + kThreads * t) : 0u;
                const bool better = k > hk || (k == hk && l > lo);
                hk = better ? k : hk;
                lo = better ? l : lo;
            }
            unsigned long long pk = 0;
            if (sizeof(T) == 4) {
                pk = (hk << 20) | lo;
                wave_argmax64(pk);
                hk = pk >> 20;
                lo = (unsigned)(pk & 0xFFFFFull);
            } else {
                wave_argmax_packed(hk, lo);
            }
            const int wg = (int)(lo & 1023u);
            if (hk != 0ull && tid == (wg & (kThreads - 1))) {
                const int tw = wg / kThreads;
#pragma unroll
                for (int c = jj; c < kSB; ++c) {
                    cx<T> v = x[c][0];
#pragma unroll
                    for (int t = 1; t < RM; ++t)
                        if (tw == t) v = x[c][t];
                    stc<3>(rb + (wave * kSB + c), v);
                }
            }
            if (lane == 0) {
                if (sizeof(T) == 4) {
                    *(__attribute__((address_space(3))) unsigned long long*)(cb + wave * 4) = pk;
                } else {
                    stc1u(cb + wave * 4, (unsigned)hk);
                    stc1u(cb + wave * 4 + 1, (unsigned)(hk >> 32));
                    stc1u(cb + wave * 4 + 2, lo);
                }
            }
            __syncthreads();
            unsigned long long bk = 0;
            unsigned bl = 0;
            int bw = 0;
            if (sizeof(T) == 4) {
                unsigned long long bp = 0;
#pragma unroll
                for (int v = 0; v < kWaves; ++v) {
                    const unsigned long long p = *(const __attribute__((address_space(3))) unsigned long long*)(cb + v * 4);
                    const bool better = p > bp;
                    bp = better ? p : bp;
                    bw = better ? v : bw;
                }
                bl = (unsigned)(bp & 0xFFFFFull);
            } else {
#pragma unroll
                for (int v = 0; v < kWaves; ++v) {
                    const unsigned long long k = ((unsigned long long)lds1<3>(cb + v * 4 + 1) << 32) | lds1<3>(cb + v * 4);
                    const unsigned l = lds1<3>(cb + v * 4 + 2);
                    const bool better = k > bk || (k == bk && l > bl);
                    bk = better ? k : bk;
                    bl = better ? l : bl;
                    bw = better ? v : bw;
                }
            }
            bw = __builtin_amdgcn_readfirstlane(bw);
            const int bpos = 1023 - (int)(__builtin_amdgcn_readfirstlane(bl) >> 10);
            const cx<T>* u = rb + bw * kSB;
            const cx<T> pv = ldc<3>(u + jj);
            const bool zero = pv.re == T(0) && pv.im == T(0);
            const cx<T> rp = zero ? czero<T>() : crecip(pv);
            if (tid == 0) {
                stc1i(ipiv + j0 + c0 + jj, j0 + c0 + bpos);
                if (zero && lds1<3>(sc + 1) == 0) stc1i(sc + 1, j0 + c0 + jj + 1);
            }
#pragma unroll
            for (int t = 0; t < RM; ++t) {
                if (lpos[t] == jj) lpos[t] = bpos;
                else if (lpos[t] == bpos) lpos[t] = jj;
            }
            cx<T> uv[kSB];
#pragma unroll
            for (int c = jj + 1; c < kSB; ++c) uv[c] = ldc<3>(u + c);
#pragma unroll
            for (int t = 0; t < RM; ++t) {
                if (lpos[t] > jj && lpos[t] != 0x7fffffff) {
                    const cx<T> l = zero ? x[jj][t] : cmul(x[jj][t], rp);
                    x[jj][t] = l;
#pragma unroll
                    for (int c = jj + 1; c < kSB; ++c) cfms(x[c][t], l, uv[c]);
                }
            }
        };
        static_for(column, std::make_integer_sequence<int, kSB>{});
    } else {
        for (int jj = 0; jj < sbw; ++jj) {
            const int pb = jj & 1;
            cx<T>* rb = rowbuf + pb * kWaves * kSB;
            unsigned* cb = cand + pb * kWaves * 4;
            // ---- this wave's candidate
            unsigned long long hk = 0;
            unsigned lo = 0;
#pragma unroll
            for (int t = 0; t < RM; ++t) {
                const bool elig = lpos[t] >= jj && lpos[t] != 0x7fffffff;
                const unsigned long long k = elig ? pivot_ukey(cabs1(x[0][t]), lpos[t] == jj) : 0ull;
                const unsigned l = elig ? ((unsigned)(1023 - lpos[t]) << 10) | (unsigned)(tid + kThreads * t) : 0u;
                const bool better = k > hk || (k == hk && l > lo);
                hk = better ? k : hk;
                lo = better ? l : lo;
            }
            unsigned long long pk = 0;  // complex64: key and position in one word
            if (sizeof(T) == 4) {
                pk = (hk << 20) | lo;  // hk < 2^32 for complex64
                wave_argmax64(pk);
                hk = pk >> 20;
                lo = (unsigned)(pk & 0xFFFFFull);
            } else {
                wave_argmax_packed(hk, lo);
            }
            const int wg = (int)(lo & 1023u);
            if (hk != 0ull && tid == (wg & (kThreads - 1))) {
                const int tw = wg / kThreads;
#pragma unroll
                for (int c = 0; c < kSB; ++c) {
                    cx<T> v = x[c][0];
#pragma unroll
                    for (int t = 1; t < RM; ++t)
                        if (tw == t) v = x[c][t];
                    stc<3>(rb + (wave * kSB + c), v);
                }
            }
            if (lane == 0) {
                if (sizeof(T) == 4) {
                    *(__attribute__((address_space(3))) unsigned long long*)(cb + wave * 4) = pk;
                } else {
                    stc1u(cb + wave * 4, (unsigned)hk);
                    stc1u(cb + wave * 4 + 1, (unsigned)(hk >> 32));
                    stc1u(cb + wave * 4 + 2, lo);
                }
            }
            __syncthreads();
            // ---- the pivot: best of the four waves' candidates
            unsigned long long bk = 0;
            unsigned bl = 0;
            int bw = 0;
            if (sizeof(T) == 4) {
                unsigned long long bp = 0;
#pragma unroll
                for (int v = 0; v < kWaves; ++v) {
                    const unsigned long long p = *(const __attribute__((address_space(3))) unsigned long long*)(cb + v * 4);
                    const bool better = p > bp;
                    bp = better ? p : bp;
                    bw = better ? v : bw;
                }
                bl = (unsigned)(bp & 0xFFFFFull);
            } else {
#pragma unroll
                for (int v = 0; v < kWaves; ++v) {
                    const unsigned long long k = ((unsigned long long)lds1<3>(cb + v * 4 + 1) << 32) | lds1<3>(cb + v * 4);
                    const unsigned l = lds1<3>(cb + v * 4 + 2);
                    const bool better = k > bk || (k == bk && l > bl);
                    bk = better ? k : bk;
                    bl = better ? l : bl;
                    bw = better ? v : bw;
                }
            }
            bw = __builtin_amdgcn_readfirstlane(bw);
            const int bpos = 1023 - (int)(__builtin_amdgcn_readfirstlane(bl) >> 10);
            const cx<T>* u = rb + bw * kSB;
            const cx<T> pv = ldc<3>(u);
            const bool zero = pv.re == T(0) && pv.im == T(0);
            const cx<T> rp = zero ? czero<T>() : crecip(pv);
            if (tid == 0) {
                stc1i(ipiv + j0 + c0 + jj, j0 + c0 + bpos);
                if (zero && lds1<3>(sc + 1) == 0) stc1i(sc + 1, j0 + c0 + jj + 1);
            }
#pragma unroll
            for (int t = 0; t < RM; ++t) {
                if (lpos[t] == jj) lpos[t] = bpos;
                else if (lpos[t] == bpos) lpos[t] = jj;
            }
            // ---- multipliers and rank-1 update of the rows below the pivot
            cx<T> uv[kSB];
#pragma unroll
            for (int c = 1; c < kSB; ++c) uv[c] = ldc<3>(u + c);
            const int ncr = sbw - jj;  // columns of the sub-panel from the current one on (uniform)
#pragma unroll
            for (int t = 0; t < RM; ++t) {
                if (lpos[t] > jj && lpos[t] != 0x7fffffff) {
                    const cx<T> l = zero ? x[0][t] : cmul(x[0][t], rp);
                    x[0][t] = l;
#pragma unroll
                    for (int c = 1; c < kSB; ++c)
                        if (c < ncr) cfms(x[c][t], l, uv[c]);  // finished columns sit at the back
                }
            }
            // ---- next column into x[0]; the finished one to the back
#pragma unroll
            for (int t = 0; t < RM; ++t) {
                const cx<T> x0 = x[0][t];
#pragma unroll
                for (int c = 0; c + 1 < kSB; ++c) x[c][t] = x[c + 1][t];
                x[kSB - 1][t] = x0;
            }
        }
        // back to x[c] = column c (kSB rotations in all)
        for (int r = sbw; r < kSB; ++r) {
#pragma unroll
            for (int t = 0; t < RM; ++t) {
                const cx<T> x0 = x[0][t];
#pragma unroll
                for (int c = 0; c + 1 < kSB; ++c) x[c][t] = x[c + 1][t];
                x[kSB - 1][t] = x0;
            }
        }
    }
    // the sub-panel at its new rows (every element to a distinct position;
    // nothing reads these columns now) and the list of moved rows
#pragma unroll
    for (int t = 0; t < RM; ++t) {
        const int r = tid + kThreads * t;
        if (r < mq) {
#pragma unroll
            for (int c = 0; c < kSB; ++c)
                if (c < sbw) stc<3>(Q + (lpos[t] + (size_t)c * ldp), x[c][t]);
            if (lpos[t] != r) {
                const int idx = atomicAdd(sc + 12, 1);
                mvl[idx] = make_int2(c0 + r, c0 + lpos[t]);
            }
        }
    }
    subpanel_moves<T>(P, ldp, jbl, c0, sbw, j0, rmap, sc, mvl);
}

// ---- one-wavefront sub-panel (complex64, at most 64 RW rows) --------------------------------
// The same factorisation as subpanel_factor, with every row of the sub-panel
// in wave 0's registers: lane l holds rows l + 64 t (t < RW), x[c][t] =
// column c.  The column loop is unrolled, so no rotation is needed; per column
// the argmax is one DPP reduction (the same packed keys and tie-breaking), the
// winner's lane and register come back uniform, and the pivot row reaches
// every lane by readlane -- no LDS exchange and no barrier per column.  The
// other waves wait at subpanel_moves' first barrier.
template <int RW>
__device__ __forceinline__ void subpanel_factor_1w(cx<float>* __restrict__ P, int m, int ldp, int jbl, int c0, int sbw,
                                                   int j0, int* ipiv, int* rmap, int* sc, int2* mvl) {
    static_assert(RW >= 1 && RW <= 4, "one wavefront holds at most 256 rows");
    const int tid = threadIdx.x;
    const int mq = m - c0;
    cx<float>* Q = P + c0 + (size_t)c0 * ldp;  // sub-panel origin, ld ldp
    if (tid < 64) {
        const int lane = tid;
        cx<float> x[kSB][RW];
        int lpos[RW];  // logical position in the sub-panel; -1: no row
#pragma unroll
        for (int t = 0; t < RW; ++t) {
            const int r = lane + 64 * t;
            lpos[t] = r < mq ? r : -1;
#pragma unroll
            for (int c = 0; c < kSB; ++c)
                x[c][t] = ldz<3>(Q + (min(r, mq - 1) + (size_t)min(c, sbw - 1) * ldp), c < sbw && r < mq);
        }
        if (lane == 0) stc1i(sc + 12, 0);
        // column jj of the sub-panel; instantiated for jj = 0 .. 15 so that every
        // register index is static
        auto column = [&](auto jc) {
            constexpr int jj = decltype(jc)::value;
            if (jj >= sbw) return;
            unsigned long long pk = 0;
#pragma unroll
            for (int t = 0; t < RW; ++t) {
                const bool elig = lpos[t] >= jj;
                const unsigned long long k =
                    elig ? (pivot_ukey(cabs1(x[jj][t]), lpos[t] == jj) << 20) |
                               ((unsigned)(1023 - lpos[t]) << 10) | (unsigned)(lane + 64 * t)
                         : 0ull;
                pk = k > pk ? k : pk;
            }
            wave_argmax64(pk);
            const unsigned lo = (unsigned)(pk & 0xFFFFFull);
            const int lw = (int)(lo & 63u), tw = (int)((lo >> 6) & 15u);
            const int bpos = 1023 - (int)(lo >> 10);
            // the pivot row, columns jj .. 15, uniform in every lane
            cx<float> uv[kSB];
#pragma unroll
            for (int c = 0; c < kSB; ++c) uv[c] = czero<float>();
#pragma unroll
            for (int t = 0; t < RW; ++t) {
                if (tw == t) {
#pragma unroll
                    for (int c = jj; c < kSB; ++c) {
                        uv[c].re = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x[c][t].re), lw));
                        uv[c].im = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x[c][t].im), lw));
                    }
                }
            }
            const cx<float> pv = uv[jj];
            const bool zero = pv.re == 0.f && pv.im == 0.f;
            const cx<float> rp = zero ? czero<float>() : crecip(pv);
            if (lane == 0) {
                stc1i(ipiv + j0 + c0 + jj, j0 + c0 + bpos);
                if (zero && lds1<3>(sc + 1) == 0) stc1i(sc + 1, j0 + c0 + jj + 1);
            }
#pragma unroll
            for (int t = 0; t < RW; ++t) {
                if (lpos[t] == jj) lpos[t] = bpos;
                else if (lpos[t] == bpos) lpos[t] = jj;
            }
            // multipliers and rank-1 update of the rows below the pivot
#pragma unroll
            for (int t = 0; t < RW; ++t) {
                if (lpos[t] > jj) {
                    const cx<float> l = zero ? x[jj][t] : cmul(x[jj][t], rp);
                    x[jj][t] = l;
#pragma unroll
                    for (int c = jj + 1; c < kSB; ++c) cfms(x[c][t], l, uv[c]);
                }
            }
        };
        static_for(column, std::make_integer_sequence<int, kSB>{});
        // the sub-panel at its new rows and the list of moved rows
#pragma unroll
        for (int t = 0; t < RW; ++t) {
            const int r = lane + 64 * t;
            if (r < mq) {
#pragma unroll
                for (int c = 0; c < kSB; ++c)
                    if (c < sbw) stc<3>(Q + (lpos[t] + (size_t)c * ldp), x[c][t]);
                if (lpos[t] != r) {
                    const int idx = atomicAdd(sc + 12, 1);
                    mvl[idx] = make_int2(c0 + r, c0 + lpos[t]);
                }
            }
        }
    }
    subpanel_moves<float>(P, ldp, jbl, c0, sbw, j0, rmap, sc, mvl);
}

// ---- inverse of a unit-lower 16 x 16 diagonal block ----------------------------------------
// From Q (LDS, ld ldq; kb <= 16 valid rows/columns), by forward substitution
// X[i][c] -= L[i][k] X[k][c] (i > k) over all 256 threads, one entry each;
// X[k][c] comes from the same wavefront by one cross-lane read per part.
// Linv (LDS, 16 x 16, ld 16) is zero outside the kb x kb block; Lg (slot,
// ld kb) gets the kb x kb block for the triangular solves.
template <typename T>
__device__ __forceinline__ void invert_unit_lower16(const cx<T>* Q, int ldq, int kb, cx<T>* Linv,
                                                    cx<T>* __restrict__ Lg) {
    static_assert(kThreads == kSB * kSB, "one thread per entry of the block");
    const int tid = threadIdx.x;
    const int i = tid & 15, c = tid >> 4;  // thread (i, c) holds X[i][c]
    cx<T> lrow[kSB - 1];                   // L[i][k] for k < i (zero for rows >= kb)
#pragma unroll
    for (int k = 0; k < kSB - 1; ++k)
        lrow[k] = ldz<3>(Q + (min(i, kb - 1) + (size_t)min(k, kb - 1) * ldq), k < i && i < kb);
    cx<T> x = (i == c && c < kb) ? cmk<T>(T(1), T(0)) : czero<T>();
    static_for([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if (k < kb - 1) {
            // X[k][c] sits on lane k + 16 (c % 4) of this wavefront: lane k of the row
            const cx<T> xk = crowbcast<k>(x);
            if (i > k) cfms(x, lrow[k], xk);
        }
    }, std::make_integer_sequence<int, kSB - 1>{});
    stc<3>(Linv + (i + c * kSB), x);
    if (i < kb && c < kb) stc<1>(Lg + (i + (size_t)c * kb), x);
}

// ---- update of the panel's columns right of sub-panel c0 (MFMA, LDS) ------------------------
// Rows [c0, c0+16): U = Linv A (the block row of U); rows [c0+16, m):
// A -= L U.  Columns [c0 + 16, jbl) in chunks of 16.
template <typename T>
__device__ __forceinline__ void inpanel_update(cx<T>* __restrict__ P, int m, int ldp, int jbl, int c0,
                                               const cx<T>* Linv) {
    using M = Mfma<T>;
    using acc_t = typename M::acc;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int cr = c0 + kSB;              // first column (and row) right of / below the block
    const int nchk = (jbl - cr + kSB - 1) / kSB;
    // block row of U, one chunk per wave
    for (int ch = wave; ch < nchk; ch += kWaves) {
        const int cb = cr + kSB * ch, cw = min(kSB, jbl - cb);
        acc_t ure, uim;
        acc_zero<T>(ure);
        acc_zero<T>(uim);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int k = M::row(lg, s);
            const cx<T> a = ldc<3>(Linv + (li + k * kSB));
            const cx<T> b = ldz<3>(P + ((c0 + k) + (size_t)(cb + min(li, cw - 1)) * ldp), li < cw);
            cmfma<T>(ure, uim, a, b.re, b.im);
        }
        if (li < cw)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                stc<3>(P + ((c0 + M::row(lg, r)) + (size_t)(cb + li) * ldp), cmk<T>(ure[r], uim[r]));
    }
    __syncthreads();
    // rows below: (row tile, chunk) units round-robin over the waves
    const int ntile = (m - cr + kSB - 1) / kSB;
    for (int u = wave; u < ntile * nchk; u += kWaves) {
        const int it = u / nchk, ch = u - it * nchk;
        const int cb = cr + kSB * ch, cw = min(kSB, jbl - cb);
        const int r0 = cr + kSB * it;
        acc_t dre, dim;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = r0 + M::row(lg, r);
            const cx<T> v = ldz<3>(P + (min(row, m - 1) + (size_t)(cb + min(li, cw - 1)) * ldp), row < m && li < cw);
            dre[r] = v.re;
            dim[r] = v.im;
        }
        const int arow = min(r0 + li, m - 1);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int k = M::row(lg, s);
            cx<T> a = ldc<3>(P + (arow + (size_t)(c0 + k) * ldp));
            a.re = -a.re;
            a.im = -a.im;
            const cx<T> b = ldz<3>(P + ((c0 + k) + (size_t)(cb + min(li, cw - 1)) * ldp), li < cw);
            cmfma<T>(dre, dim, a, b.re, b.im);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = r0 + M::row(lg, r);
            if (row < m && li < cw) stc<3>(P + (row + (size_t)(cb + li) * ldp), cmk<T>(dre[r], dim[r]));
        }
    }
    __syncthreads();
}

// ---- trailing update after a wide panel ------------------------------------------------------
// Columns [j0 + jbl, ncol) (the matrix, then right-hand sides), chunks of 16
// columns dealt to the waves.  P: the factored panel (LDS, ld ldp >= n - j0,
// logical order); Linv: its inverted 16 x 16 diagonal blocks (ld 16).
// ASM: first panel, whose trailing entries are assembled from the operators
// (physical row r, column c of A(w) is element r + c n) instead of loaded.
// G3: the C update uses the Gauss 3M product (cmfma3) instead of four real
// products per complex one.
// LG: the factored columns [j0, j0 + jbl) are read from the slot (logical row
// r at physical row rmap[r]) instead of the LDS panel, and Linv points to
// their inverted diagonal blocks in the slot: the deferred update of a
// two-level step, whose jbl = 16..64 columns were factored as several LDS
// panels (every 16-block full).
// NBL > 0: the panel has exactly NBL 16-column blocks (jbl = 16 NBL), known at
// compile time, so the block loops carry no branches and no unused registers.
// NOC: U12 only (stored to the slot), no C update (the staged deferred update
// applies C afterwards).
template <typename T, bool ASM, bool G3, bool LG, int NBL = 0, bool NOC = false>
__device__ __forceinline__ void wide_trailing(const SysDesc& sd, double w, double w2, cx<T>* __restrict__ S, int ld,
                                              const int* rmap, int j0, int jbl, int n, int ncol, const cx<T>* P,
                                              int ldp, const cx<T>* Linv) {
    using M = Mfma<T>;
    using acc_t = typename M::acc;
    constexpr int NB = NBL ? NBL : WideCfg<T>::NB;
    if (NBL) jbl = kSB * NBL;
    // the assembling variant keeps half the C tile (its FP64 operator loads need the registers)
    constexpr int TA = ASM ? (WideCfg<T>::TA > 2 ? 2 : WideCfg<T>::TA) : WideCfg<T>::TA;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int c0 = j0 + jbl, m2 = n - c0;
    const int nbl = (jbl + kSB - 1) / kSB;
    const int nch = (ncol - c0 + kSB - 1) / kSB;
    // LG: physical rows of this lane's row li in each 16-row block of L11
    int plr[LG ? NB : 1];
#pragma unroll
    for (int b = 0; b < (LG ? NB : 1); ++b) plr[b] = LG ? lds1<3>(rmap + j0 + min(kSB * b + li, jbl - 1)) : 0;
    auto ldL = [&](const cx<T>* p) -> cx<T> { return LG ? ldc<1>(p) : ldc<3>(p); };
    // element (physical row prow, column col) of the trailing matrix: the
    // matrix part assembled from the operators at the first panel (ASM), the
    // right-hand-side columns (col >= n) from the slot; no branch around the loads
    auto load_el = [&](int prow, int col) -> cx<T> {
        if (!ASM) return ldc<1>(S + (prow + (size_t)col * ld));
        const cx<T> a = to_t<T>(assemble_at(sd, (size_t)prow + (size_t)min(col, n - 1) * ld, w, w2));
        const cx<T> f = ldc<1>(S + (prow + (size_t)max(col, n) * ld));
        return col < n ? a : f;
    };
    for (int ch = wave; ch < nch; ch += kWaves) {
        const int cb = c0 + ch * kSB, cw = min(kSB, ncol - cb);
        // ---- A12 in operand layout: block b, register r = logical row j0 + 16 b + row(lg, r)
        acc_t are[NB], aim[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = kSB * b + M::row(lg, r);
                // rows beyond the panel must be zero (they meet the zeroed panel columns
                // of L11 and Linv); columns beyond cw only feed unstored results
                const cx<T> v = load_el(lds1<3>(rmap + j0 + min(k, jbl - 1)), cb + min(li, cw - 1));
                const bool ok = b < nbl && k < jbl;
                are[b][r] = ok ? v.re : T(0);
                aim[b][r] = ok ? v.im : T(0);
            }
        // ---- U12 = L11^{-1} A12, block forward substitution
        acc_t ure[NB], uim[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            acc_zero<T>(ure[b]);
            acc_zero<T>(uim[b]);
            if (b < nbl) {
#pragma unroll
                for (int a = 0; a < b; ++a) {
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const int k = kSB * a + M::row(lg, s);
                        const int row = kSB * b + li;
                        // rows beyond: unstored
                        cx<T> l = LG ? ldc<1>(S + (plr[b] + (size_t)(j0 + k) * ld))
                                     : ldc<3>(P + (min(row, jbl - 1) + (size_t)k * ldp));
                        l.re = -l.re;
                        l.im = -l.im;
                        cmfma<T>(are[b], aim[b], l, ure[a][s], uim[a][s]);
                    }
                }
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const cx<T> li_v = ldL(Linv + (kSB * kSB * b + li + M::row(lg, s) * kSB));
                    cmfma<T>(ure[b], uim[b], li_v, are[b][s], aim[b][s]);
                }
            }
        }
        // ---- U12 to the slot
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = kSB * b + M::row(lg, r);
                if (b < nbl && k < jbl && li < cw)
                    stc<1>(S + (lds1<3>(rmap + j0 + k) + (size_t)(cb + li) * ld), cmk<T>(ure[b][r], uim[b][r]));
            }
        if (NOC || m2 <= 0) continue;
        // ---- C -= L21 U12, transposed: D = U12^T L21^T; register r of tile a holds
        // C[row 16 a + li][column row(lg, r)]
        __builtin_amdgcn_wave_barrier();
        auto load_c = [&](int r0, acc_t (&cre)[TA], acc_t (&cim)[TA], int (&pr)[TA]) {
#pragma unroll
            for (int a = 0; a < TA; ++a) {
                const int row = r0 + kSB * a + li;
                pr[a] = lds1<3>(rmap + c0 + min(row, m2 - 1));
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = M::row(lg, r);
                    // rows beyond m2 and columns beyond cw are computed but never stored
                    const cx<T> v = load_el(pr[a], cb + min(col, cw - 1));
                    cre[a][r] = v.re;
                    cim[a][r] = v.im;
                }
            }
        };
        acc_t cre[TA], cim[TA];
        acc_t cy[G3 ? TA : 1];  // G3: sum of a_i b_i
        int prow[TA];
        load_c(0, cre, cim, prow);
        for (int r0 = 0; r0 < m2; r0 += kSB * TA) {
            acc_t nre[TA], nim[TA];
            int nprow[TA];
            const bool more = r0 + kSB * TA < m2;
            if (more && !ASM) load_c(r0 + kSB * TA, nre, nim, nprow);
            if (G3) {
#pragma unroll
                for (int a = 0; a < TA; ++a) {
                    acc_zero<T>(cy[a]);
#pragma unroll
                    for (int r = 0; r < 4; ++r) cim[a][r] += cre[a][r];
                }
            }
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                if (b < nbl) {
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const int k = kSB * b + M::row(lg, s);
                        // A operand: -U12^T (lane: U12[k][li], zero for k beyond the panel, whose
                        // LDS columns up to the next multiple of 16 are zero)
                        const cx<T> ua = cmk<T>(-ure[b][s], -uim[b][s]);
                        cx<T> lv[TA];
#pragma unroll
                        for (int a = 0; a < TA; ++a) {
                            const int row = min(r0 + kSB * a + li, m2 - 1);
                            // LG: L21 shares its physical rows with the C tile
                            lv[a] = LG ? ldc<1>(S + (prow[a] + (size_t)(j0 + k) * ld))
                                       : ldc<3>(P + ((jbl + row) + (size_t)k * ldp));
                        }
                        if (G3) {
                            const T us = -(ure[b][s] + uim[b][s]);
#pragma unroll
                            for (int a = 0; a < TA; ++a)
                                cmfma3<T>(cre[a], cy[a], cim[a], ua, us, lv[a].re, lv[a].im);
                        } else {
#pragma unroll
                            for (int a = 0; a < TA; ++a) cmfma<T>(cre[a], cim[a], ua, lv[a].re, lv[a].im);
                        }
                    }
                }
            }
            if (G3) {
#pragma unroll
                for (int a = 0; a < TA; ++a)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const T x = cre[a][r], y = cy[a][r];
                        cre[a][r] = x - y;
                        cim[a][r] = cim[a][r] - x - y;
                    }
            }
#pragma unroll
            for (int a = 0; a < TA; ++a) {
                const int row = r0 + kSB * a + li;
                if (row < m2) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int col = M::row(lg, r);
                        if (col < cw) stc<1>(S + (prow[a] + (size_t)(cb + col) * ld), cmk<T>(cre[a][r], cim[a][r]));
                    }
                }
            }
            if (more) {
                if (ASM) {
                    load_c(r0 + kSB * TA, cre, cim, prow);
                } else {
#pragma unroll
                    for (int a = 0; a < TA; ++a) {
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

// ---- deferred update of a two-level step, L21 staged through LDS -------------------------------
// Columns [c0, ncol) with c0 = j0 + jo, rows [c0, n): C -= L21 U12 where the
// step's columns [j0, j0 + jo) are factored in the slot.  U12 = L11^{-1} A12 is
// computed per chunk first (wide_trailing without its C update); then, per
// band of BR rows, the band's L21 (BR x jo) is staged in the LDS region Pb
// (pcap complex elements, free between panels, odd leading dimension) and
// every wavefront updates its chunks' C tiles of the band with U12 read back
// from the slot.  Each C tile is read and written once per step instead of
// once per LDS panel, and L21 is read once.  Needs pcap >= 17 jo.
template <typename T, bool ASM, bool G3>
__device__ __forceinline__ void wide_trailing_staged(const SysDesc& sd, double w, double w2, cx<T>* __restrict__ S,
                                                     int ld, const int* rmap, int j0, int jo, int n, int ncol,
                                                     cx<T>* Pb, long long pcap, const cx<T>* Linv) {
    using M = Mfma<T>;
    using acc_t = typename M::acc;
    constexpr int NB = WideCfg<T>::NB;
    constexpr int TA = ASM ? (WideCfg<T>::TA > 2 ? 2 : WideCfg<T>::TA) : WideCfg<T>::TA;
    wide_trailing<T, ASM, G3, true, 0, true>(sd, w, w2, S, ld, rmap, j0, jo, n, ncol, Pb, 0, Linv);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int c0 = j0 + jo, m2 = n - c0;
    const int nbl = jo / kSB;
    const int nch = (ncol - c0 + kSB - 1) / kSB;
    if (m2 <= 0) return;
    // band height: a multiple of 64 rows when one fits, else of 16 (the tile
    // loop clamps a partial band)
    int BR = (int)(pcap / jo) / 16 * 16;
    if (BR >= 64) BR = BR / 64 * 64;
    while (BR > 16 && (long long)(BR | 1) * jo > pcap) BR -= (BR > 64 ? 64 : 16);
    const int ldb = BR | 1;
    auto load_el = [&](int prow, int col) -> cx<T> {
        if (!ASM) return ldc<1>(S + (prow + (size_t)col * ld));
        const cx<T> a = to_t<T>(assemble_at(sd, (size_t)prow + (size_t)min(col, n - 1) * ld, w, w2));
        const cx<T> f = ldc<1>(S + (prow + (size_t)max(col, n) * ld));
        return col < n ? a : f;
    };
    for (int rb = 0; rb < m2; rb += BR) {
        const int nr = min(BR, m2 - rb);
        __syncthreads();  // U12 stored, the previous band consumed
        for (int e = tid; e < nr * jo; e += kThreads) {
            const int r = e % nr, k = e / nr;
            stc<3>(Pb + (r + (size_t)k * ldb), ldc<1>(S + (lds1<3>(rmap + c0 + rb + r) + (size_t)(j0 + k) * ld)));
        }
        __syncthreads();
        for (int ch = wave; ch < nch; ch += kWaves) {
            const int cb = c0 + ch * kSB, cw = min(kSB, ncol - cb);
            // U12 of the chunk back from the slot, operand layout (lane: U12[k][li])
            acc_t ure[NB], uim[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int k = kSB * b + M::row(lg, q);
                    const cx<T> v = ldc<1>(S + (lds1<3>(rmap + j0 + min(k, jo - 1)) + (size_t)(cb + min(li, cw - 1)) * ld));
                    const bool ok = b < nbl;
                    ure[b][q] = ok ? v.re : T(0);
                    uim[b][q] = ok ? v.im : T(0);
                }
            auto load_c = [&](int r0, acc_t (&cre)[TA], acc_t (&cim)[TA], int (&pr)[TA]) {
#pragma unroll
                for (int a = 0; a < TA; ++a) {
                    const int row = rb + min(r0 + kSB * a + li, nr - 1);
                    pr[a] = lds1<3>(rmap + c0 + row);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const cx<T> v = load_el(pr[a], cb + min(M::row(lg, r), cw - 1));
                        cre[a][r] = v.re;
                        cim[a][r] = v.im;
                    }
                }
            };
            acc_t cre[TA], cim[TA];
            acc_t cy[G3 ? TA : 1];
            int prow[TA];
            load_c(0, cre, cim, prow);
            for (int r0 = 0; r0 < nr; r0 += kSB * TA) {
                acc_t nre[TA], nim[TA];
                int nprow[TA];
                const bool more = r0 + kSB * TA < nr;
                if (more && !ASM) load_c(r0 + kSB * TA, nre, nim, nprow);
                if (G3) {
#pragma unroll
                    for (int a = 0; a < TA; ++a) {
                        acc_zero<T>(cy[a]);
#pragma unroll
                        for (int r = 0; r < 4; ++r) cim[a][r] += cre[a][r];
                    }
                }
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    if (b < nbl) {
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const int k = kSB * b + M::row(lg, q);
                            const cx<T> ua = cmk<T>(-ure[b][q], -uim[b][q]);
                            cx<T> lv[TA];
#pragma unroll
                            for (int a = 0; a < TA; ++a)
                                lv[a] = ldc<3>(Pb + (min(r0 + kSB * a + li, nr - 1) + (size_t)k * ldb));
                            if (G3) {
                                const T us = -(ure[b][q] + uim[b][q]);
#pragma unroll
                                for (int a = 0; a < TA; ++a)
                                    cmfma3<T>(cre[a], cy[a], cim[a], ua, us, lv[a].re, lv[a].im);
                            } else {
#pragma unroll
                                for (int a = 0; a < TA; ++a) cmfma<T>(cre[a], cim[a], ua, lv[a].re, lv[a].im);
                            }
                        }
                    }
                }
                if (G3) {
#pragma unroll
                    for (int a = 0; a < TA; ++a)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const T x = cre[a][r], y = cy[a][r];
                            cre[a][r] = x - y;
                            cim[a][r] = cim[a][r] - x - y;
                        }
                }
#pragma unroll
                for (int a = 0; a < TA; ++a) {
                    if (r0 + kSB * a + li < nr) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int col = M::row(lg, r);
                            if (col < cw) stc<1>(S + (prow[a] + (size_t)(cb + col) * ld), cmk<T>(cre[a][r], cim[a][r]));
                        }
                    }
                }
                if (more) {
                    if (ASM) {
                        load_c(r0 + kSB * TA, cre, cim, prow);
                    } else {
#pragma unroll
                        for (int a = 0; a < TA; ++a) {
                            cre[a] = nre[a];
                            cim[a] = nim[a];
                            prow[a] = nprow[a];
                        }
                    }
                }
            }
        }
    }
}

// wide_trailing with the panel's block count fixed at compile time for the
// common widths (16 and 32 columns).
template <typename T, bool ASM, bool G3>
__device__ __forceinline__ void wide_trailing_w(const SysDesc& sd, double w, double w2, cx<T>* __restrict__ S, int ld,
                                                const int* rmap, int j0, int jbl, int n, int ncol, const cx<T>* P,
                                                int ldp, const cx<T>* Linv) {
    if (jbl == kSB) wide_trailing<T, ASM, G3, false, 1>(sd, w, w2, S, ld, rmap, j0, jbl, n, ncol, P, ldp, Linv);
    else if (jbl == 2 * kSB) wide_trailing<T, ASM, G3, false, 2>(sd, w, w2, S, ld, rmap, j0, jbl, n, ncol, P, ldp, Linv);
    else wide_trailing<T, ASM, G3, false>(sd, w, w2, S, ld, rmap, j0, jbl, n, ncol, P, ldp, Linv);
}

// ---- the wide-panel factorisation ------------------------------------------------------------
// TWO: two-level update (lay.jo columns per step); otherwise one level.
template <typename T, int MAXR, bool G3, bool TWO>
__device__ __forceinline__ void factor_wide(const SysDesc& s, double w, double w2, cx<T>* __restrict__ S,
                                            const Layout& lay, const Lds& L, int nr) {
    const int n = s.n, tid = threadIdx.x;
    const size_t nn = (size_t)n * n;
    const int jb = lay.jb;
    char* base = L.main;
    cx<T>* P = reinterpret_cast<cx<T>*>(base);
    base += align16((long long)(n | 1) * jb * sizeof(cx<T>));
    cx<T>* Linv = reinterpret_cast<cx<T>*>(base);
    base += align16((long long)jb * kSB * sizeof(cx<T>));
    cx<T>* cand = reinterpret_cast<cx<T>*>(base);  // sub-panel pivot candidates
    base += align16(wide_cand_bytes((int)sizeof(cx<T>)));
    int2* mvl = reinterpret_cast<int2*>(base);
    const size_t nblk = (size_t)((n + kSB - 1) / kSB);
    cx<T>* Lg = S + (size_t)n * (n + kMaxGroup);
    cx<T>* Ug = Lg + nblk * kSB * kSB;

    // right-hand sides eliminated along with the factorisation (columns n .. n+nr-1)
    for (int e = tid; e < n * nr; e += kThreads) {
        const int r = e % n, g = e / n;
        stc<1>(S + (nn + (size_t)g * n + r),
               to_t<T>(ldc<1>(reinterpret_cast<const z64*>(s.F) + r + (size_t)g * s.ldf)));
    }
    for (int r = tid; r < n; r += kThreads) L.rmap[r] = r;
    if (tid == 0) L.sc[1] = 0;
    __syncthreads();

    // two-level steps of jo > jb columns: the LDS panels of a step update only
    // the step's own columns, and the columns right of it get one deferred
    // rank-jo update with L read back from the slot after the step's last
    // panel; the step holding the matrix's last columns updates everything
    // panel by panel.  jo <= jb: one level.
    const int jo = (TWO && lay.jo > jb) ? lay.jo : jb;
    for (int J0 = 0; J0 < n; J0 += jo) {
        const int jend = min(J0 + jo, n);
        const bool defer = TWO && jo > jb && jend < n;
        for (int j0 = J0; j0 < jend; j0 += jb) {
            const int jbl = min(jb, n - j0), m = n - j0;
            // complex64: odd leading dimension, so the columns an MFMA operand read
            // spans start in different LDS banks (complex128 elements spread already)
            const int ldp = sizeof(T) == 4 ? (m | 1) : m;
            // ---- panel into LDS (logical rows), assembled from the operators at j0 = 0:
            // items of kPB columns of one row (loads in flight together), spread over
            // all threads, so short panels keep every thread loading
            const int nbat = (jbl + kPB - 1) / kPB;
            for (int e = tid; e < m * nbat; e += kThreads) {
                const int r = e % m, cb = (e / m) * kPB;
                const int pr = j0 == 0 ? r : lds1<3>(L.rmap + j0 + r);
                cx<T> v[kPB];
#pragma unroll
                for (int u = 0; u < kPB; ++u) {
                    const int c = min(cb + u, jbl - 1);
                    v[u] = j0 == 0 ? to_t<T>(assemble_at(s, (size_t)pr + (size_t)c * n, w, w2))
                                   : ldc<1>(S + (pr + (size_t)(j0 + c) * n));
                }
#pragma unroll
                for (int u = 0; u < kPB; ++u)
                    if (cb + u < jbl) stc<3>(P + (r + (size_t)(cb + u) * ldp), v[u]);
                // columns from jbl to the next multiple of 16 read as zero by the
                // trailing update's last k-block
                if (cb == 0)
                    for (int c = jbl; c < ((jbl + kSB - 1) / kSB) * kSB; ++c) stc<3>(P + (r + (size_t)c * ldp), czero<T>());
            }
            __syncthreads();
            // ---- sub-panels
            for (int c0 = 0; c0 < jbl; c0 += kSB) {
                const int sbw = min(kSB, jbl - c0);
                // the sub-panel's chain of barriers is latency-bound: its waves
                // take issue priority over the other workgroup's update on the CU
                // (complex64: +0.6 %; complex128 measured -0.6 % and keeps 0)
                if (sizeof(T) == 4) __builtin_amdgcn_s_setprio(2);
                // complex64 sub-panels of at most 64 rows: one wavefront, no
                // barrier per column (measured at n = 256: +2.3 %; with up to 256
                // rows in one wavefront its serial rank-1 updates cost about what
                // the barriers save)
                bool one_wave = false;
                if constexpr (sizeof(T) == 4) {
                    one_wave = m - c0 <= 64;
                    if (one_wave) subpanel_factor_1w<1>(P, m, ldp, jbl, c0, sbw, j0, L.ipiv, L.rmap, L.sc, mvl);
                }
                if (!one_wave) subpanel_factor<T, MAXR>(P, m, ldp, jbl, c0, sbw, j0, L.ipiv, L.rmap, L.sc, cand, mvl);
                if (sizeof(T) == 4) __builtin_amdgcn_s_setprio(0);
                invert_unit_lower16<T>(P + c0 + (size_t)c0 * ldp, ldp, sbw, Linv + (size_t)(c0 / kSB) * kSB * kSB,
                                       Lg + (size_t)((j0 + c0) / kSB) * kSB * kSB);
                __syncthreads();
                if (c0 + kSB < jbl) inpanel_update<T>(P, m, ldp, jbl, c0, Linv + (size_t)(c0 / kSB) * kSB * kSB);
            }
            // ---- the factored panel to the slot (rows at their physical positions)
            for (int r = tid; r < m; r += kThreads) {
                const int pr = lds1<3>(L.rmap + j0 + r);
                for (int c = 0; c < jbl; ++c) stc<1>(S + (pr + (size_t)(j0 + c) * n), ldc<3>(P + (r + (size_t)c * ldp)));
            }
            // ---- trailing columns (only the step's own when deferring)
            const int tcol = defer ? jend : n + nr;
            // (the two-level kernel keeps one variant of its narrow within-step update:
            // less code, fewer spills)
            if (TWO) {
                if (j0 == 0) wide_trailing<T, true, G3, false>(s, w, w2, S, n, L.rmap, j0, jbl, n, tcol, P, ldp, Linv);
                else wide_trailing<T, false, G3, false>(s, w, w2, S, n, L.rmap, j0, jbl, n, tcol, P, ldp, Linv);
            } else {
                if (j0 == 0) wide_trailing_w<T, true, G3>(s, w, w2, S, n, L.rmap, j0, jbl, n, tcol, P, ldp, Linv);
                else wide_trailing_w<T, false, G3>(s, w, w2, S, n, L.rmap, j0, jbl, n, tcol, P, ldp, Linv);
            }
            __syncthreads();
        }
        if (TWO && defer) {
            // ---- deferred update of columns [jend, n + nr) by the step's jo columns
            const cx<T>* Lgs = Lg + (size_t)(J0 / kSB) * kSB * kSB;
            // L21 staged through the panel's LDS region ((n | 1) jb >= 17 jo always:
            // jb >= 16, jo <= 4 jb, n >= jo)
            const long long pcap = (long long)(n | 1) * jb;
            if (J0 == 0) wide_trailing_staged<T, true, G3>(s, w, w2, S, n, L.rmap, J0, jo, n, n + nr, P, pcap, Lgs);
            else wide_trailing_staged<T, false, G3>(s, w, w2, S, n, L.rmap, J0, jo, n, n + nr, P, pcap, Lgs);
            __syncthreads();
        }
    }
    if (sizeof(T) == 4) {
        uinv_blocks(S, n, L.rmap, kSB, Ug);
        __syncthreads();
    }
}

}  // namespace stc
