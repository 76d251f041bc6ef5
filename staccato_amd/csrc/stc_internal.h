// stc_internal.h -- types shared by the host API and the gfx950 kernels.
//
// A launch processes "points": one (system, frequency) pair each.  A system
// is one ROM (K, M, D, eta, n) with its frequency list, load F and outputs.
// The descriptor layout is private to the library; the public ABI is
// include/staccato_sweep.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace stc {

constexpr int kThreads = 256;   // 4 wavefronts of 64 lanes per workgroup
constexpr int kWaves = kThreads / 64;
constexpr int kMaxGroup = 4;    // right-hand sides solved together
constexpr int kRedSlots = 1 + 5 * kMaxGroup;  // block reductions in the residual
// reduction scratch: one row of kRedSlots per wave, the result row, scale factors
constexpr int kRedDoubles = (kWaves + 1) * kRedSlots + kMaxGroup;

// flags of a launch
enum : int {
    kRefine = 1,      // complex64 factors + FP64 refinement
    kWantBerr = 2,    // compute the backward error (always on when refining)
    kAdjoint = 4,     // fused adjoint after the forward solve
    kFactorOut = 8,   // copy factors + pivots out (stc_getrf)
    kRhsLoad = 16,    // F given per point (the right-hand side for the adjoint pass)
};

struct SysDesc {
    int n;
    int nrhs;
    int nfreq;
    int ldf;
    int ldx;
    int real_ops;        // K, M, D stored as n*n doubles (imaginary parts zero)
    int has_D;
    int p;               // adjoint: number of modal parameters
    double eta;
    const double* K;     // n*n (real) or 2*n*n (interleaved complex), ld = n
    const double* M;
    const double* D;     // may be null
    const double* freq;  // Hz
    const double2* F;    // n x nrhs (ldf)
    double2* X;          // nfreq blocks of n x nrhs (ldx)
    int* info;           // nfreq
    double* berr;        // nfreq * nrhs, nullable
    int* iters;          // nfreq, nullable (refinement steps used)
    void* factor_out;    // nfreq blocks of n*n (complex of the factor type), nullable
    int* ipiv_out;       // nfreq blocks of n, nullable
    // adjoint
    const double2* c;    // n
    const double* Phi;   // n x p, ld = n
    const double* kdiag; // p
    double* adj_out;     // nfreq blocks of (p + 2): J_w, dJ/deta_w, dJ/ds_k,w
    // fused output projection: C nout x n (ld nout), Y nfreq blocks of
    // nout x nrhs (ld nout); X may then be null (not written)
    const double2* C;
    double2* Y;
    int nout;
    int pad_;
    long long item_base; // first point index of this system in the queue
    // explicit matrices: > 0 makes K, M and D advance by op_stride doubles and F by
    // f_stride complex entries per frequency index (freq == nullptr: w = 0)
    long long op_stride;
    long long f_stride;
    // LDS plan chosen by the host (plan_layout) for this system
    int lay_small;
    int lay_jb;
    int lay_cw;
    int lay_G;
    int lay_jo;          // wide path: outer width of the two-level update (<= lay_jb: none)
    int lay_rsg;         // wide path: FP64 residuals in the factor slot instead of LDS
};

struct KParams {
    SysDesc one;             // used when sys == nullptr (single system)
    const SysDesc* sys;      // device array of nsys descriptors, or nullptr
    int nsys;
    const int2* items;       // (system, frequency) per point, or nullptr: (0, idx)
    int nitems;
    int* counter;            // work-queue head (zeroed by a reset kernel)
    void* scratch;           // per-workgroup factor slots (blocked path)
    size_t slot_elems;       // complex elements per slot
    int flags;
    int max_iter;
    double tol;
    int lds_budget;          // bytes of dynamic LDS given to the launch
};

// Partition of the dynamic LDS for one point.
struct Layout {
    int small;      // 1: whole matrix in LDS; 0: blocked, matrix in a global slot
    int jb;         // panel width (blocked)
    int cw;         // trailing column chunk width (blocked)
    int G;          // right-hand sides per solve group
    int lds;        // bytes needed
    int jo = 0;     // wide path: columns per two-level step (<= jb: one level)
    int rsg = 0;    // wide path: FP64 residuals in the factor slot (frees LDS for the group)
};

__host__ __device__ inline int align16(long long b) { return (int)((b + 15) & ~15LL); }

constexpr int kUwCols = 16;       // trailing columns per wave chunk (kernel: kUw)
constexpr int kUwLdCols = 17;     // padded row length of a wave's U12 chunk in LDS
constexpr int kSolveBlk = 32;     // diagonal block of the blocked solves

// Panel columns each wavefront keeps in registers (x MAXR rows per lane):
// 16 complex values per lane for complex64 (the kernel runs three workgroups
// per CU, at most 168 registers), 16 for complex128 up to n = 512 (two per CU).
__host__ __device__ constexpr int panel_cols_per_wave(int elem_bytes, int maxr) {
    return elem_bytes == 8 ? (maxr <= 8 ? 4 : 2)
                           : (maxr <= 4 ? 4 : (maxr <= 8 ? 2 : 1));
}

// Workgroups (of four wavefronts) per CU the kernel is compiled for: three for
// the complex64 factorisation, two for complex128 (its panel and products need
// up to 256 registers).
__host__ __device__ constexpr int wg_per_cu(int elem_bytes) { return elem_bytes == 8 ? 3 : 2; }

// Elements of a blocked-path slot: the n x n factors, up to kMaxGroup
// right-hand-side columns eliminated along with them (columns n .. n+3, ld n),
// and the inverted diagonal blocks of L and U (ceil(n/jb) blocks of jb x jb each).
__host__ __device__ inline size_t slot_elems_for(int n, int jb) {
    const size_t nb = (size_t)((n + jb - 1) / jb);
    return (size_t)n * (n + kMaxGroup) + 2 * nb * jb * jb;
}

// Fixed region at the start of LDS: ipiv and the row map (n ints each),
// scalars (16 ints), reduction scratch.
__host__ __device__ inline int lds_fixed_bytes(int n) {
    return align16(2LL * n * 4) + align16(16 * 4) + align16((long long)kRedDoubles * 8);
}

// Solve vectors: w (factor type), xs, rs (FP64) and, for the adjoint, the
// kept forward solution and the adjoint right-hand side (FP64).
__host__ __device__ inline int lds_solve_bytes(int n, int G, int ts, int adj, int rsg = 0) {
    return align16((long long)n * G * ts) + ((adj ? 4 : 2) - (rsg ? 1 : 0)) * align16((long long)n * G * 16);
}

// Factor-slot elements (of ts bytes) kept for the FP64 residuals of one solve
// group when they do not fit the LDS (wide path, lay_rsg).
__host__ __device__ inline size_t slot_rs_elems(int n, int ts) {
    return (size_t)n * kMaxGroup * 16 / ts + 1;  // + 1: 16-byte alignment for complex64
}

// Complex elements of the blocked path's LDS panel region: the n x jb panel,
// which also holds the two alternating multiplier columns of the register
// panel (64 * maxr elements each, maxr = rows per lane of the launch, which a
// small system of a mixed batch shares with the largest one).
__host__ __device__ inline long long panel_region_elems(int n, int jb, int maxr) {
    const long long a = (long long)n * jb, b = 2LL * 64 * maxr;
    return a > b ? a : b;
}

// Choose the path and blocking for a system of order n.  ts = sizeof(complex
// factor element) (8 or 16).  budget = dynamic LDS bytes of the launch; maxr =
// rows per lane of the kernel the launch uses.
inline Layout plan_layout(int n, int nrhs, int ts, int budget, int jb_cap, int adj, int maxr) {
    Layout L{0, 0, 0, 1, 0};
    const int fixed = lds_fixed_bytes(n);
    int Gmax = adj ? 1 : (nrhs < kMaxGroup ? nrhs : kMaxGroup);
    if (Gmax < 1) Gmax = 1;
    for (int G = Gmax; G >= 1; --G) {
        int need = fixed + align16((long long)n * n * ts) + lds_solve_bytes(n, G, ts, adj);
        if (need <= budget) {
            L.small = 1; L.G = G; L.lds = need;
            return L;
        }
    }
    // blocked: LDS panel (n x jb) + one U12 chunk per wave (jb x kUwLdCols) +
    // the inverted diagonal block (jb x jb); the solve phase reuses the same
    // region for its vectors + a staged diagonal block
    const int jbs[] = {32, 28, 24, 20, 16, 12, 8, 4};
    for (int a = 0; a < 8; ++a) {
        int jb = jbs[a];
        if (jb > jb_cap) continue;
        int need = fixed + align16(panel_region_elems(n, jb, maxr) * ts) +
                   align16((long long)kWaves * jb * kUwLdCols * ts) + align16((long long)jb * jb * ts);
        if (need <= budget) {
            L.jb = jb; L.cw = kUwCols; L.lds = need;
            for (int G = Gmax; G >= 1; --G) {
                int ns = fixed + lds_solve_bytes(n, G, ts, adj) +
                         align16((long long)kSolveBlk * kSolveBlk * ts);
                if (ns <= budget) { L.G = G; if (ns > L.lds) L.lds = ns; return L; }
            }
            L.lds = 0x7fffffff;  // cannot even hold one solve column
            return L;
        }
    }
    L.lds = 0x7fffffff;
    return L;
}

// ---- wide-panel blocked path (sweep_kernel<T, MAXR, true>) ----------------------------
constexpr int kSB = 16;  // sub-panel width and inverted diagonal block size

// Widest panel the wide kernel keeps in registers for the trailing update:
// 64 columns for complex64, 32 for complex128.
__host__ __device__ constexpr int wide_jb_max(int elem_bytes) { return elem_bytes == 8 ? 64 : 32; }

// LDS of the wide path beyond lds_fixed_bytes: the n x jb panel (leading
// dimension rounded up to odd, so its columns start in different banks), its inverted
// 16 x 16 diagonal blocks, the sub-panel's two alternating pivot-candidate
// buffers (per wave one row of 16 and its 16-byte key) and the moved-row
// list (2 x 16).  maxr is not needed (kept for the callers' signature).
__host__ __device__ inline long long wide_cand_bytes(int ts) { return 2LL * kWaves * kSB * ts + 2LL * kWaves * 16; }
__host__ __device__ inline long long wide_lds_bytes(int n, int jb, int ts, int maxr) {
    (void)maxr;
    return align16((long long)(n | 1) * jb * ts) + align16((long long)jb * kSB * ts) + align16(wide_cand_bytes(ts)) +
           align16(2LL * kSB * 8);
}

// Layout of a system in a wide-panel launch: the small path when the matrix
// fits the LDS budget whole, else the widest panel (a multiple of 16 up to
// jb_cap) that fits.
inline Layout plan_layout_wide(int n, int nrhs, int ts, int budget, int jb_cap, int adj, int maxr) {
    Layout L{0, 0, 0, 1, 0};
    const int fixed = lds_fixed_bytes(n);
    int Gmax = adj ? 1 : (nrhs < kMaxGroup ? nrhs : kMaxGroup);
    if (Gmax < 1) Gmax = 1;
    for (int G = Gmax; G >= 1; --G) {
        int need = fixed + align16((long long)n * n * ts) + lds_solve_bytes(n, G, ts, adj);
        if (need <= budget) {
            L.small = 1; L.G = G; L.lds = need;
            return L;
        }
    }
    for (int jb = wide_jb_max(ts); jb >= kSB; jb -= kSB) {
        if (jb > jb_cap) continue;
        const long long need = fixed + wide_lds_bytes(n, jb, ts, maxr);
        if (need <= budget) {
            L.jb = jb; L.cw = kSB; L.lds = (int)need;
            // the largest solve group, with the FP64 residuals in LDS if that
            // group fits so, else in the factor slot
            for (int G = Gmax; G >= 1; --G) {
                // the adjoint keeps them in LDS, and so do the kernels for n <= 256
                // (maxr < 8), whose groups always fit
                for (int rsg = 0; rsg <= ((adj || maxr < 8) ? 0 : 1); ++rsg) {
                    const int ns = fixed + lds_solve_bytes(n, G, ts, adj, rsg) +
                                   align16((long long)kSolveBlk * kSolveBlk * ts);
                    if (ns <= budget) {
                        L.G = G; L.rsg = rsg;
                        if (ns > L.lds) L.lds = ns;
                        return L;
                    }
                }
            }
            L.lds = 0x7fffffff;
            return L;
        }
    }
    L.lds = 0x7fffffff;
    return L;
}

}  // namespace stc
