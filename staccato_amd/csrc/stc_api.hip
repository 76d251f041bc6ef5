// stc_api.hip -- C ABI of libstaccato_sweep.so (declared in include/staccato_sweep.h).
//
// Host side of the MI355X frequency sweep: a handle owns one device, one HIP
// stream, the ROM operators in device memory and the per-workgroup factor
// slots.  Every sweep is one persistent launch of sweep_kernel over a device
// work queue of (system, frequency) points, preceded by a one-lane reset of
// the queue head (a kernel, so the pair can be captured in a hipGraph).
//
// Launches of one handle share its factor slots and queue head, so they are
// ordered: each waits for the previous one (an event) when it goes to another
// stream.  A plan owns its own slots and queue head and keeps the ROMs it
// reads alive (rom_set / rom_clear refuse them while the plan exists).
#include "../../include/staccato_sweep.h"
#include "stc_launch.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

using namespace stc;

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int ensure(size_t b) {
        if (b <= bytes) return STC_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (hipMalloc(&p, b) != hipSuccess) return STC_ENOMEM;
        bytes = b;
        return STC_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename U>
    U* as() const { return reinterpret_cast<U*>(p); }
};

struct RomDev {
    int n = 0;
    double eta = 0.0;
    int real_ops = 1;
    int has_D = 0;
    DevBuf K, M, D;
    int nin = 0, nout = 0;
    DevBuf B, C;  // n x nin, nout x n (complex, ld = n / nout)
    int plan_refs = 0;  // live plans whose graph reads K, M, D
};

// Orders the launches that share one set of device buffers: a launch on a
// stream other than the previous one's waits for it first.
struct LaunchOrder {
    hipEvent_t ev = nullptr;
    hipStream_t last = nullptr;
    bool used = false;
    int before(hipStream_t st) {
        if (used && st != last && hipStreamWaitEvent(st, ev, 0) != hipSuccess) return STC_EDEVICE;
        return STC_OK;
    }
    int after(hipStream_t st) {
        if (hipEventRecord(ev, st) != hipSuccess) return STC_EDEVICE;
        last = st;
        used = true;
        return STC_OK;
    }
    // host wait for the last launch (before its buffers are freed or resized)
    void drain() {
        if (used) (void)hipEventSynchronize(ev);
    }
};

}  // namespace

// How the blocked path of a launch runs (stc_policy): kernel 0 = automatic,
// 1 = 16-column register panel, 2 = wide LDS panel, 3 = wide LDS panel with
// Gauss 3M products in the trailing update (complex64; else as 2), 4 = wide LDS
// panel with the two-level update (several LDS panels per step, one deferred
// rank-64 / rank-32 update of the columns right of the step); lds_budget bytes per
// workgroup (0 = automatic); jb_max caps the panel width (0 = none).
struct Policy {
    int kernel = 0;
    int lds_budget = 0;
    int jb_max = 0;
};

struct stc_ctx {
    int device = 0;
    int num_cu = 0;
    int ok_arch = 0;
    void* na_scratch = nullptr;  // device address of this GPU's `scratch` array (noalloc entry)
    hipStream_t stream = nullptr;
    std::map<int, RomDev> roms;
    DevBuf scratch, counter;
    LaunchOrder order;
    int live_plans = 0;
    std::map<std::pair<int, int>, Policy> policy;  // (order n, refine) -> policy
    // staging for host-memory calls
    DevBuf freq, F, X, info, berr, iters, fac, ipiv, sys, items, adj, c, phi, kd, A, load;
};

struct stc_plan_s {
    stc_ctx* h = nullptr;
    DevBuf sys, items, scratch, counter;
    std::vector<int> roms;  // rom ids whose plan_refs this plan holds
    LaunchOrder order;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

#define HIP_TRY(x)                                       \
    do {                                                 \
        if ((x) != hipSuccess) return STC_EDEVICE;       \
    } while (0)

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int d) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != d) (void)hipSetDevice(d);
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

int upload(DevBuf& b, const void* src, size_t bytes, hipStream_t st) {
    int rc = b.ensure(bytes ? bytes : 16);
    if (rc) return rc;
    if (bytes) HIP_TRY(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, st));
    return STC_OK;
}

// ---- fixed scratch of the allocation-free entry (stc_sweep_dev_noalloc) -----------------
// One array per GPU, part of the library's code object: the queue head in its
// first 256 bytes, the factor slots after it.  In host code `scratch` names a
// host-side stand-in, never the device memory: the device address comes from
// hipGetSymbolAddress, once per GPU (at its first stc_create, with that GPU
// current), and only that address is passed to kernels.
constexpr size_t kNoAllocBytes = size_t(64) << 20;
constexpr size_t kNoAllocHead = 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_na_mu;
void* g_na_ptr[kMaxDevices] = {};

// Device address of the current GPU's `scratch` (device == current ordinal).
void* noalloc_scratch_of(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_na_ptr[device] = p;
    }
    return g_na_ptr[device];
}

constexpr int kBudgetThird = 53 * 1024;       // three workgroups per CU
constexpr int kBudgetSmall = 80 * 1024;       // two workgroups per CU
constexpr int kBudgetLarge = 160 * 1024 - 256;  // one workgroup per CU

// Pick the LDS budget of a launch (forced_budget > 0: that one) and fill each
// system's layout for the register-panel kernel: the largest number of
// workgroups per CU the kernel is built for at which every blocked system
// still gets a panel of at least 12 columns, or as many as the registers
// hold when that is fewer (a thin panel starves the trailing update; at
// n = 1024 complex128 the registers cap it at 4, and two workgroups per CU
// measured 1126 against 857 points/s for one).
int plan_systems(SysDesc* sys, int nsys, int ts, int adj, int jb_cap, int maxr, int forced_budget,
                 int* lds_out) {
    int budget = kBudgetLarge;
    if (forced_budget > 0) {
        budget = forced_budget;
    } else {
        const int cands[] = {kBudgetThird, kBudgetSmall};
        for (int b : cands) {
            if (b == kBudgetThird && wg_per_cu(ts) < 3) continue;
            bool ok = true;
            for (int i = 0; i < nsys; ++i) {
                const SysDesc& s = sys[i];
                Layout L = plan_layout(s.n, s.nrhs, ts, b, jb_cap, adj, maxr);
                if (L.lds > b || (!L.small && L.jb < std::min(12, jb_cap))) ok = false;
            }
            if (ok) {
                budget = b;
                break;
            }
        }
    }
    int lds = 0;
    for (int i = 0; i < nsys; ++i) {
        SysDesc& s = sys[i];
        Layout L = plan_layout(s.n, s.nrhs, ts, budget, jb_cap, adj, maxr);
        if (L.lds > budget) return STC_EINVAL;
        s.lay_small = L.small;
        s.lay_jb = L.jb;
        s.lay_cw = L.cw;
        s.lay_G = L.G;
        lds = std::max(lds, L.lds);
    }
    *lds_out = lds;
    return STC_OK;
}

// Wide-panel layout of every system within `budget` bytes of LDS; false when
// one of them does not fit.
bool layout_wide(SysDesc* sys, int nsys, int ts, int adj, int jb_cap, int maxr, int budget,
                 bool two_level, int* lds_out) {
    int lds = 0;
    for (int i = 0; i < nsys; ++i) {
        const Layout L = plan_layout_wide(sys[i].n, sys[i].nrhs, ts, budget, jb_cap, adj, maxr);
        if (L.lds > budget) return false;
        lds = std::max(lds, L.lds);
    }
    for (int i = 0; i < nsys; ++i) {
        const Layout L = plan_layout_wide(sys[i].n, sys[i].nrhs, ts, budget, jb_cap, adj, maxr);
        sys[i].lay_small = L.small;
        sys[i].lay_jb = L.jb;
        sys[i].lay_cw = L.cw;
        sys[i].lay_G = L.G;
        sys[i].lay_rsg = L.rsg;
        // two-level update: steps of the widest panel the trailing update holds
        // (64 columns complex64, 32 complex128), a multiple of the LDS panel
        const int jmax = wide_jb_max(ts);
        sys[i].lay_jo = (two_level && !L.small && L.jb >= kSB && L.jb < jmax) ? (jmax / L.jb) * L.jb : 0;
    }
    *lds_out = lds;
    return true;
}

// Policy for a launch whose largest system has order n: the entry for n, else
// the one with the nearest order, else automatic.
Policy policy_for(const stc_ctx* h, int n, int refine) {
    const Policy* best = nullptr;
    int bd = 0x7fffffff;
    for (const auto& kv : h->policy) {
        if (kv.first.second != refine) continue;
        const int d = std::abs(kv.first.first - n);
        if (d < bd) {
            bd = d;
            best = &kv.second;
        }
    }
    return best ? *best : Policy{};
}

struct LaunchCfg {
    int refine = 0;
    int maxr = 1;
    int wide = 0;   // sweep_kernel<., ., wide>: 1 wide LDS panels on the blocked path, 2 with 3M products, 3 two-level
    int lds = 0;
    int grid = 1;
    size_t slot_elems = 0;
    size_t slot_bytes = 0;  // bytes of one factor slot (0: no blocked system)
};

// Layout and grid of a launch, no allocation (host or device): grid = resident
// workgroups, capped by the points and, when slot_budget > 0, by the factor
// slots that fit in slot_budget bytes (STC_ETOOBIG when not even one does).
int plan_launch(stc_ctx* h, SysDesc* sys, int nsys, long long nitems, int refine, int adj,
                size_t slot_budget, LaunchCfg* cfg) {
    const int ts = refine ? 8 : 16;
    int maxn = 0;
    for (int i = 0; i < nsys; ++i) maxn = std::max(maxn, sys[i].n);
    if (maxn < 1 || maxn > 1024) return STC_EINVAL;
    int r = (maxn + 63) / 64, maxr = 1;
    while (maxr < r) maxr <<= 1;
    cfg->refine = refine;
    cfg->maxr = maxr;
    const Policy pol = policy_for(h, maxn, refine);
    // the register-resident panel holds 4 (or 2, 1) columns per wavefront
    int jb_cap = kWaves * panel_cols_per_wave(ts, maxr);
    if (pol.jb_max > 0) jb_cap = std::max(4, std::min(pol.jb_max, jb_cap));
    int rc = plan_systems(sys, nsys, ts, adj, jb_cap, maxr, pol.kernel == 1 ? pol.lds_budget : 0, &cfg->lds);
    if (rc) return rc;
    // automatic: systems that need a factor slot go to the wide-panel kernel
    // when the largest order exceeds 128 (below, the register panel at three
    // workgroups per CU measured faster), the kernel is built for these rows
    // per lane and every system's panel fits its LDS; policy kernels 2 and 3
    // take it for every launch it fits (3: Gauss 3M products, complex64 only)
    bool blocked = false;
    for (int i = 0; i < nsys; ++i) blocked = blocked || !sys[i].lay_small;
    cfg->wide = 0;
    // automatic: the two-level update above n = 256 (measured on MI355X, both
    // modes: n = 384 +11 % / +7 %, 512 +24 % / +12 %, 1024 +36 % complex64),
    // one level up to n = 256 (there two-level loses for complex64: 374k
    // against 457k points/s at n = 256), with Gauss 3M products for complex64
    // (n = 160 / 192 / 224 / 256: +1.5 / +5.3 / +4.8 / +3.5 % over 4M)
    const bool two_auto = pol.kernel == 0 && maxn > 256;
    const bool g3_auto = pol.kernel == 0 && refine && maxn <= 256;
    const int wkind = ((pol.kernel == 3 || g3_auto) && sweep_kernel_fn(refine, maxr, 2)) ? 2
                      : ((pol.kernel == 4 || two_auto) && sweep_kernel_fn(refine, maxr, 3)) ? 3 : 1;
    if (((pol.kernel == 0 && blocked && maxn > 128) || pol.kernel >= 2) && sweep_kernel_fn(refine, maxr, wkind)) {
        int wcap = wide_jb_max(ts);
        if (pol.jb_max > 0) wcap = std::max(kSB, std::min(pol.jb_max, wcap));
        // automatic LDS budget (measured on MI355X, n = 160 .. 1024): two
        // workgroups per CU up to n = 512 for complex64 and n = 256 for
        // complex128, one per CU (wider panels, half the slots) beyond
        const int auto_budget = maxn > (ts == 8 ? 512 : 256) ? kBudgetLarge : kBudgetSmall;
        const int budget = pol.lds_budget > 0 ? pol.lds_budget : auto_budget;
        int lds = 0;
        if (layout_wide(sys, nsys, ts, adj, wcap, maxr, budget, wkind == 3, &lds)) {  // writes only when all fit
            cfg->wide = wkind;
            cfg->lds = lds;
        }
    }
    void* kfun = sweep_kernel_fn(refine, maxr, cfg->wide);
    if (!kfun) return STC_EINVAL;
    HIP_TRY(hipFuncSetAttribute(kfun, hipFuncAttributeMaxDynamicSharedMemorySize, cfg->lds));
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfun, kThreads, cfg->lds));
    if (per_cu < 1) per_cu = 1;
    long long g = (long long)h->num_cu * per_cu;
    // every blocked system factors in a per-workgroup slot (also for stc_getrf,
    // which copies the slot out in LAPACK layout afterwards)
    size_t slot = 0;
    for (int i = 0; i < nsys; ++i)
        if (!sys[i].lay_small)
            slot = std::max(slot, slot_elems_for(sys[i].n, cfg->wide ? kSB : sys[i].lay_jb) +
                                      (cfg->wide && sys[i].lay_rsg ? slot_rs_elems(sys[i].n, ts) : 0));
    slot = (slot + 1) & ~(size_t)1;  // every slot 16-byte aligned (complex64 elements)
    cfg->slot_elems = slot;
    cfg->slot_bytes = slot * ts;
    if (slot && slot_budget > 0) {
        const long long fit = (long long)(slot_budget / cfg->slot_bytes);
        if (fit < 1) return STC_ETOOBIG;
        g = std::min(g, fit);
    }
    cfg->grid = (int)std::max(1LL, std::min(g, nitems));
    return STC_OK;
}

// Queue-head reset + the persistent kernel on `st` (nothing else: capturable).
__global__ void reset_counter_kernel(int* counter) {
    if (threadIdx.x == 0) counter[0] = 0;
}

int enqueue_raw(const LaunchCfg& cfg, KParams& kp, int* counter, void* scratch, hipStream_t st) {
    kp.counter = counter;
    kp.scratch = scratch;
    kp.slot_elems = cfg.slot_elems;
    kp.lds_budget = cfg.lds;
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, st, kp.counter);
    HIP_TRY(hipGetLastError());
    HIP_TRY(sweep_kernel_launch(cfg.refine, cfg.maxr, cfg.wide, cfg.grid, cfg.lds, st, kp));
    return STC_OK;
}

// Plan a launch and size the handle's factor slots for it.  Growing the slots
// frees the old ones, so the handle's previous launch is waited for first.
int configure(stc_ctx* h, std::vector<SysDesc>& sys, long long nitems, int refine, int adj,
              LaunchCfg* cfg) {
    int rc = plan_launch(h, sys.data(), (int)sys.size(), nitems, refine, adj, 0, cfg);
    if (rc) return rc;
    const size_t need = cfg->slot_bytes * (size_t)cfg->grid;
    if (need > h->scratch.bytes) {
        h->order.drain();
        if ((rc = h->scratch.ensure(need))) return rc;
    }
    return STC_OK;
}

// A launch on the handle's slots and queue head, ordered after its previous one.
int enqueue(stc_ctx* h, const LaunchCfg& cfg, KParams& kp, hipStream_t st) {
    int rc = h->order.before(st);
    if (rc) return rc;
    if ((rc = enqueue_raw(cfg, kp, h->counter.as<int>(), h->scratch.p, st))) return rc;
    return h->order.after(st);
}

SysDesc base_desc(const RomDev& r) {
    SysDesc s;
    memset(&s, 0, sizeof(s));
    s.n = r.n;
    s.eta = r.eta;
    s.real_ops = r.real_ops;
    s.has_D = r.has_D;
    s.K = r.K.as<double>();
    s.M = r.M.as<double>();
    s.D = r.has_D ? r.D.as<double>() : nullptr;
    return s;
}

KParams base_params(int mode, double tol, int max_iter) {
    KParams kp;
    memset(&kp, 0, sizeof(kp));
    kp.flags = (mode == STC_REFINED_FP32) ? kRefine : 0;
    kp.tol = tol > 0.0 ? tol : 1e-12;
    kp.max_iter = max_iter > 0 ? max_iter : 10;
    return kp;
}

int check_mode(int mode) { return (mode == STC_PLAIN_FP64 || mode == STC_REFINED_FP32) ? 0 : 1; }

// info -1: refinement missed tol (soft); info -2: the kernel could not run the
// point (no factor slot or a layout it cannot hold; the host never plans one),
// X is NaN there: a hard error.
int status_from_info(const int* info, long long cnt, stc_stats* st) {
    int nsing = 0, nconv = 0, nfail = 0;
    for (long long j = 0; j < cnt; ++j) {
        if (info[j] > 0) ++nsing;
        else if (info[j] == -1) ++nconv;
        else if (info[j] < 0) ++nfail;
    }
    if (st) {
        st->n_singular = nsing;
        st->n_not_converged = nconv;
    }
    if (nfail) return STC_EDEVICE;
    if (nsing) return STC_ESINGULAR;
    if (nconv) return STC_ENOCONV;
    return STC_OK;
}

// Copy a host n x ncol block with leading dimension ld into a packed device buffer.
int upload_block(DevBuf& b, const stc_z64* src, int n, int ncol, int ld, hipStream_t st) {
    std::vector<stc_z64> tmp((size_t)n * ncol);
    for (int c = 0; c < ncol; ++c)
        memcpy(&tmp[(size_t)c * n], src + (size_t)c * ld, sizeof(stc_z64) * n);
    int rc = upload(b, tmp.data(), tmp.size() * sizeof(stc_z64), st);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(st));
    return STC_OK;
}

// Right-hand sides of the back-transformation, formed on the device:
// F(:, j) = B L(:, j) for j < nrhs (B n x nin, ld n; L nin x nrhs, ld nin) and,
// when withB, F(:, nrhs + k) = B(:, k) (the columns of the transfer function).
__global__ void load_rhs_kernel(const double2* B, int n, int nin, const double2* L, int nrhs,
                                int withB, double2* F) {
    const long long ncol = nrhs + (withB ? nin : 0);
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)n * ncol) return;
    const int i = (int)(e % n), j = (int)(e / n);
    if (j >= nrhs) {
        F[e] = B[i + (size_t)(j - nrhs) * n];
        return;
    }
    double re = 0.0, im = 0.0;
    for (int k = 0; k < nin; ++k) {
        const double2 b = B[i + (size_t)k * n], l = L[k + (size_t)j * nin];
        re = fma(b.x, l.x, fma(-b.y, l.y, re));
        im = fma(b.x, l.y, fma(b.y, l.x, im));
    }
    F[e] = make_double2(re, im);
}

}  // namespace

// ============================================================================
extern "C" {

const char* stc_strerror(int code) {
    switch (code) {
        case STC_OK: return "ok";
        case STC_EINVAL: return "invalid argument";
        case STC_ENOMEM: return "device allocation failed";
        case STC_EDEVICE: return "HIP runtime error";
        case STC_ENOROM: return "rom_id not set";
        case STC_EBUSY: return "in use by a live plan (destroy the plan first)";
        case STC_ETOOBIG: return "system too large for the 64 MiB scratch of stc_sweep_dev_noalloc";
        case STC_ESINGULAR: return "singular A(omega) at one or more frequencies";
        case STC_ENOCONV: return "iterative refinement did not reach tol at one or more frequencies";
        default: return "unknown status";
    }
}

int stc_create(stc_handle* out, int device) {
    if (!out) return STC_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return STC_EDEVICE;
    stc_ctx* h = new stc_ctx();
    h->device = device;
    DeviceGuard g(device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete h; return STC_EDEVICE; }
    h->num_cu = prop.multiProcessorCount;
    h->ok_arch = strncmp(prop.gcnArchName, "gfx950", 6) == 0;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return STC_EDEVICE;
    }
    if (h->counter.ensure(256) != STC_OK) {
        (void)hipStreamDestroy(h->stream);
        delete h;
        return STC_ENOMEM;
    }
    if (hipEventCreateWithFlags(&h->order.ev, hipEventDisableTiming) != hipSuccess) {
        h->counter.release();
        (void)hipStreamDestroy(h->stream);
        delete h;
        return STC_EDEVICE;
    }
    // set-up of the allocation-free entry (null when unavailable: that entry
    // then returns STC_EDEVICE; every other call works without it)
    if (h->ok_arch) h->na_scratch = noalloc_scratch_of(device);
    *out = h;
    return STC_OK;
}

int stc_destroy(stc_handle h) {
    if (!h) return STC_EINVAL;
    if (h->live_plans > 0) return STC_EBUSY;
    {
        DeviceGuard g(h->device);
        h->order.drain();
        (void)hipStreamSynchronize(h->stream);
        (void)hipEventDestroy(h->order.ev);
        for (auto& kv : h->roms) {
            kv.second.K.release(); kv.second.M.release(); kv.second.D.release();
            kv.second.B.release(); kv.second.C.release();
        }
        DevBuf* bufs[] = {&h->scratch, &h->counter, &h->freq, &h->F, &h->X, &h->info, &h->berr,
                          &h->iters, &h->fac, &h->ipiv, &h->sys, &h->items, &h->adj, &h->c,
                          &h->phi, &h->kd, &h->A, &h->load};
        for (DevBuf* b : bufs) b->release();
        (void)hipStreamDestroy(h->stream);
    }
    delete h;
    return STC_OK;
}

int stc_device_ok(stc_handle h) { return h ? h->ok_arch : 0; }

int stc_set_policy(stc_handle h, int n, int mode, const stc_policy* p) {
    if (!h || n < 1 || n > 1024 || check_mode(mode)) return STC_EINVAL;
    const auto key = std::make_pair(n, mode == STC_REFINED_FP32 ? 1 : 0);
    if (!p) {
        h->policy.erase(key);
        return STC_OK;
    }
    if (p->kernel < 0 || p->kernel > 4 || p->jb_max < 0 ||
        (p->lds_budget != 0 && (p->lds_budget < 16384 || p->lds_budget > kBudgetLarge)))
        return STC_EINVAL;
    Policy q;
    q.kernel = p->kernel;
    q.lds_budget = p->lds_budget;
    q.jb_max = p->jb_max;
    h->policy[key] = q;
    return STC_OK;
}

int stc_clear_policies(stc_handle h) {
    if (!h) return STC_EINVAL;
    h->policy.clear();
    return STC_OK;
}

int stc_rom_set(stc_handle h, int rom_id, int n, const stc_z64* K, const stc_z64* M,
                const stc_z64* D, int ld, double eta) {
    if (!h || rom_id < 0 || n < 1 || n > 1024 || ld < n || !K || !M || !std::isfinite(eta))
        return STC_EINVAL;
    const stc_z64* ops[3] = {K, M, D};
    bool real = true;
    for (int o = 0; o < 3; ++o) {
        if (!ops[o]) continue;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const stc_z64 v = ops[o][(size_t)j * ld + i];
                if (!std::isfinite(v.re) || !std::isfinite(v.im)) return STC_EINVAL;
                if (v.im != 0.0) real = false;
            }
    }
    auto old = h->roms.find(rom_id);
    if (old != h->roms.end() && old->second.plan_refs > 0) return STC_EBUSY;
    DeviceGuard g(h->device);
    // the operators may be read by a launch still in flight
    h->order.drain();
    RomDev& r = h->roms[rom_id];
    if (r.n != n) {  // input/output projections belong to the old order
        r.B.release();
        r.C.release();
        r.nin = r.nout = 0;
    }
    r.n = n;
    r.eta = eta;
    r.has_D = D != nullptr;
    r.real_ops = real;
    DevBuf* dst[3] = {&r.K, &r.M, &r.D};
    for (int o = 0; o < 3; ++o) {
        if (!ops[o]) continue;
        if (real) {
            std::vector<double> tmp((size_t)n * n);
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) tmp[(size_t)j * n + i] = ops[o][(size_t)j * ld + i].re;
            int rc = upload(*dst[o], tmp.data(), tmp.size() * sizeof(double), h->stream);
            if (rc) return rc;
            HIP_TRY(hipStreamSynchronize(h->stream));
        } else {
            int rc = upload_block(*dst[o], ops[o], n, n, ld, h->stream);
            if (rc) return rc;
        }
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    return STC_OK;
}

int stc_rom_set_io(stc_handle h, int rom_id, int nin, const stc_z64* B, int ldb, int nout,
                   const stc_z64* C, int ldc) {
    if (!h) return STC_EINVAL;
    auto it = h->roms.find(rom_id);
    if (it == h->roms.end()) return STC_ENOROM;
    RomDev& r = it->second;
    if (nin < 1 || nout < 1 || !B || !C || ldb < r.n || ldc < nout) return STC_EINVAL;
    DeviceGuard g(h->device);
    int rc = upload_block(r.B, B, r.n, nin, ldb, h->stream);
    if (rc) return rc;
    rc = upload_block(r.C, C, nout, r.n, ldc, h->stream);
    if (rc) return rc;
    r.nin = nin;
    r.nout = nout;
    return STC_OK;
}

int stc_rom_clear(stc_handle h, int rom_id) {
    if (!h) return STC_EINVAL;
    auto it = h->roms.find(rom_id);
    if (it == h->roms.end()) return STC_ENOROM;
    if (it->second.plan_refs > 0) return STC_EBUSY;
    DeviceGuard g(h->device);
    h->order.drain();
    (void)hipStreamSynchronize(h->stream);
    it->second.K.release(); it->second.M.release(); it->second.D.release();
    it->second.B.release(); it->second.C.release();
    h->roms.erase(it);
    return STC_OK;
}

int stc_sweep_dev(stc_handle h, int rom_id, int nfreq, const double* d_freq_hz, int nrhs,
                  const stc_z64* d_F, int ldf, stc_z64* d_X, int ldx, int* d_info,
                  double* d_berr, int mode, double tol, int max_iter, void* stream) {
    if (!h || nfreq < 0 || nrhs < 1 || check_mode(mode)) return STC_EINVAL;
    auto it = h->roms.find(rom_id);
    if (it == h->roms.end()) return STC_ENOROM;
    const RomDev& r = it->second;
    if (ldf < r.n || ldx < r.n || (nfreq && (!d_freq_hz || !d_F || !d_X || !d_info)))
        return STC_EINVAL;
    if (nfreq == 0) return STC_OK;
    DeviceGuard g(h->device);
    hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : h->stream;
    KParams kp = base_params(mode, tol, max_iter);
    std::vector<SysDesc> sys(1, base_desc(r));
    SysDesc& s = sys[0];
    s.nrhs = nrhs;
    s.nfreq = nfreq;
    s.ldf = ldf;
    s.ldx = ldx;
    s.freq = d_freq_hz;
    s.F = reinterpret_cast<const double2*>(d_F);
    s.X = reinterpret_cast<double2*>(d_X);
    s.info = d_info;
    s.berr = d_berr;
    LaunchCfg cfg;
    int rc = configure(h, sys, nfreq, mode == STC_REFINED_FP32, 0, &cfg);
    if (rc) return rc;
    kp.one = sys[0];
    kp.nitems = nfreq;
    if (d_berr) kp.flags |= kWantBerr;
    return enqueue(h, cfg, kp, st);
}

// Same launch as stc_sweep_dev on the GPU's fixed `scratch` array (through its
// device address, h->na_scratch): no device or host allocation, no handle
// buffer, no ordering with the handle's other launches.
int stc_sweep_dev_noalloc(stc_handle h, int rom_id, int nfreq, const double* d_freq_hz, int nrhs,
                          const stc_z64* d_F, int ldf, stc_z64* d_X, int ldx, int* d_info,
                          double* d_berr, int mode, double tol, int max_iter, void* stream) {
    if (!h || nfreq < 0 || nrhs < 1 || check_mode(mode)) return STC_EINVAL;
    auto it = h->roms.find(rom_id);
    if (it == h->roms.end()) return STC_ENOROM;
    const RomDev& r = it->second;
    if (ldf < r.n || ldx < r.n || (nfreq && (!d_freq_hz || !d_F || !d_X || !d_info)))
        return STC_EINVAL;
    if (nfreq == 0) return STC_OK;
    if (!h->na_scratch) return STC_EDEVICE;
    DeviceGuard g(h->device);
    hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : h->stream;
    KParams kp = base_params(mode, tol, max_iter);
    SysDesc s = base_desc(r);
    s.nrhs = nrhs;
    s.nfreq = nfreq;
    s.ldf = ldf;
    s.ldx = ldx;
    s.freq = d_freq_hz;
    s.F = reinterpret_cast<const double2*>(d_F);
    s.X = reinterpret_cast<double2*>(d_X);
    s.info = d_info;
    s.berr = d_berr;
    LaunchCfg cfg;
    int rc = plan_launch(h, &s, 1, nfreq, mode == STC_REFINED_FP32, 0, kNoAllocBytes - kNoAllocHead, &cfg);
    if (rc) return rc;
    kp.one = s;
    kp.nitems = nfreq;
    if (d_berr) kp.flags |= kWantBerr;
    char* base = static_cast<char*>(h->na_scratch);
    return enqueue_raw(cfg, kp, reinterpret_cast<int*>(base), base + kNoAllocHead, st);
}

int stc_sweep(stc_handle h, int rom_id, int nfreq, const double* freq_hz, int nrhs,
              const stc_z64* F, int ldf, stc_z64* X, int ldx, int* info, double* berr, int mode,
              double tol, int max_iter, stc_stats* stats) {
    if (!h || nfreq < 0 || nrhs < 1 || check_mode(mode)) return STC_EINVAL;
    auto it = h->roms.find(rom_id);
    if (it == h->roms.end()) return STC_ENOROM;
    const int n = it->second.n;
    if (ldf < n || ldx < n || (nfreq && (!freq_hz || !F || !X || !info))) return STC_EINVAL;
    if (stats) memset(stats, 0, sizeof(*stats));
    if (nfreq == 0) return STC_OK;
    DeviceGuard g(h->device);
    hipStream_t st = h->stream;
    int rc = upload(h->freq, freq_hz, sizeof(double) * nfreq, st);
    if (rc) return rc;
    rc = upload_block(h->F, F, n, nrhs, ldf, st);
    if (rc) return rc;
    const size_t xcount = (size_t)nfreq * n * nrhs;
    if ((rc = h->X.ensure(xcount * sizeof(stc_z64)))) return rc;
    if ((rc = h->info.ensure(sizeof(int) * nfreq))) return rc;
    if ((rc = h->iters.ensure(sizeof(int) * nfreq))) return rc;
    if (berr && (rc = h->berr.ensure(sizeof(double) * nfreq * nrhs))) return rc;

    const auto& r = it->second;
    KParams kp = base_params(mode, tol, max_iter);
    std::vector<SysDesc> sys(1, base_desc(r));
    SysDesc& s = sys[0];
    s.nrhs = nrhs; s.nfreq = nfreq; s.ldf = n; s.ldx = n;
    s.freq = h->freq.as<double>();
    s.F = h->F.as<double2>();
    s.X = h->X.as<double2>();
    s.info = h->info.as<int>();
    s.berr = berr ? h->berr.as<double>() : nullptr;
    s.iters = h->iters.as<int>();
    LaunchCfg cfg;
    rc = configure(h, sys, nfreq, mode == STC_REFINED_FP32, 0, &cfg);
    if (rc) return rc;
    kp.one = sys[0];
    kp.nitems = nfreq;
    if (berr) kp.flags |= kWantBerr;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, st));
    rc = enqueue(h, cfg, kp, st);
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipStreamSynchronize(st));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    // copy out (X packed n x nrhs per frequency -> caller's ldx)
    std::vector<stc_z64> xh(xcount);
    HIP_TRY(hipMemcpy(xh.data(), h->X.p, xcount * sizeof(stc_z64), hipMemcpyDeviceToHost));
    for (int f = 0; f < nfreq; ++f)
        for (int c = 0; c < nrhs; ++c)
            memcpy(X + ((size_t)f * nrhs + c) * ldx, &xh[((size_t)f * nrhs + c) * n],
                   sizeof(stc_z64) * n);
    HIP_TRY(hipMemcpy(info, h->info.p, sizeof(int) * nfreq, hipMemcpyDeviceToHost));
    std::vector<int> its(nfreq);
    HIP_TRY(hipMemcpy(its.data(), h->iters.p, sizeof(int) * nfreq, hipMemcpyDeviceToHost));
    if (berr)
        HIP_TRY(hipMemcpy(berr, h->berr.p, sizeof(double) * nfreq * nrhs, hipMemcpyDeviceToHost));
    if (stats) {
        stats->seconds = ms * 1e-3;
        stats->points = nfreq;
        stats->max_refine_steps = *std::max_element(its.begin(), its.end());
        if (berr) {
            double m = 0.0;
            for (long long q = 0; q < (long long)nfreq * nrhs; ++q)
                if (std::isfinite(berr[q])) m = std::max(m, berr[q]);
            stats->max_backward_error = m;
        }
    }
    return status_from_info(info, nfreq, stats);
}

int stc_getrf(stc_handle h, int rom_id, int nfreq, const double* freq_hz, int mode, void* LU,
              int* ipiv, int* info) {
    if (!h || nfreq < 0 || check_mode(mode)) return STC_EINVAL;
    auto it = h->roms.find(rom_id);
    if (it == h->roms.end()) return STC_ENOROM;
    if (nfreq == 0) return STC_OK;
    if (!freq_hz || !LU || !ipiv || !info) return STC_EINVAL;
    const int n = it->second.n;
    const bool refine = mode == STC_REFINED_FP32;
    const size_t ts = refine ? sizeof(stc_c32) : sizeof(stc_z64);
    DeviceGuard g(h->device);
    hipStream_t st = h->stream;
    int rc = upload(h->freq, freq_hz, sizeof(double) * nfreq, st);
    if (rc) return rc;
    if ((rc = h->fac.ensure(ts * (size_t)n * n * nfreq))) return rc;
    if ((rc = h->ipiv.ensure(sizeof(int) * (size_t)n * nfreq))) return rc;
    if ((rc = h->info.ensure(sizeof(int) * nfreq))) return rc;
    KParams kp = base_params(mode, 0.0, 0);
    kp.flags |= kFactorOut;
    std::vector<SysDesc> sys(1, base_desc(it->second));
    SysDesc& s = sys[0];
    s.nrhs = 1; s.nfreq = nfreq; s.ldf = n; s.ldx = n;
    s.freq = h->freq.as<double>();
    s.info = h->info.as<int>();
    s.factor_out = h->fac.p;
    s.ipiv_out = h->ipiv.as<int>();
    LaunchCfg cfg;
    rc = configure(h, sys, nfreq, refine, 0, &cfg);
    if (rc) return rc;
    kp.one = sys[0];
    kp.nitems = nfreq;
    rc = enqueue(h, cfg, kp, st);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipMemcpy(LU, h->fac.p, ts * (size_t)n * n * nfreq, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(ipiv, h->ipiv.p, sizeof(int) * (size_t)n * nfreq, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(info, h->info.p, sizeof(int) * nfreq, hipMemcpyDeviceToHost));
    return status_from_info(info, nfreq, nullptr);
}

// ---- uneven batch ---------------------------------------------------------------
namespace {

// Build descriptors + a longest-first (n^3) item table.
int build_batch(stc_ctx* h, int nsys, const stc_sys_desc* in, std::vector<SysDesc>& sys,
                std::vector<int2>& items) {
    sys.clear();
    items.clear();
    for (int i = 0; i < nsys; ++i) {
        auto it = h->roms.find(in[i].rom_id);
        if (it == h->roms.end()) return STC_ENOROM;
        const RomDev& r = it->second;
        if (in[i].nfreq < 0 || in[i].nrhs < 1 || in[i].ldf < r.n || in[i].ldx < r.n)
            return STC_EINVAL;
        SysDesc s = base_desc(r);
        s.nrhs = in[i].nrhs; s.nfreq = in[i].nfreq; s.ldf = in[i].ldf; s.ldx = in[i].ldx;
        s.freq = in[i].freq_hz;
        s.F = reinterpret_cast<const double2*>(in[i].F);
        s.X = reinterpret_cast<double2*>(in[i].X);
        s.info = in[i].info;
        s.berr = in[i].berr;
        sys.push_back(s);
    }
    std::vector<int> order(nsys);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const double ca = (double)sys[a].n * sys[a].n * (sys[a].n + 3.0 * sys[a].nrhs);
        const double cb = (double)sys[b].n * sys[b].n * (sys[b].n + 3.0 * sys[b].nrhs);
        return ca > cb;
    });
    for (int i : order)
        for (int f = 0; f < sys[i].nfreq; ++f) items.push_back(make_int2(i, f));
    return STC_OK;
}

}  // namespace

int stc_sweep_batch(stc_handle h, int nsys, const stc_sys_desc* sysd, int mode, double tol,
                    int max_iter, stc_stats* stats) {
    if (!h || nsys < 0 || (nsys && !sysd) || check_mode(mode)) return STC_EINVAL;
    if (stats) memset(stats, 0, sizeof(*stats));
    if (nsys == 0) return STC_OK;
    DeviceGuard g(h->device);
    hipStream_t st = h->stream;
    // stage every system's host arrays in one device buffer each
    size_t nf = 0, nF = 0, nX = 0, nB = 0;
    for (int i = 0; i < nsys; ++i) {
        auto it = h->roms.find(sysd[i].rom_id);
        if (it == h->roms.end()) return STC_ENOROM;
        const int n = it->second.n;
        if (sysd[i].nfreq < 0 || sysd[i].nrhs < 1 || (sysd[i].nfreq && (!sysd[i].freq_hz ||
            !sysd[i].F || !sysd[i].X || !sysd[i].info)) || sysd[i].ldf < n || sysd[i].ldx < n)
            return STC_EINVAL;
        nf += sysd[i].nfreq;
        nF += (size_t)n * sysd[i].nrhs;
        nX += (size_t)n * sysd[i].nrhs * sysd[i].nfreq;
        nB += (size_t)sysd[i].nrhs * sysd[i].nfreq;
    }
    std::vector<double> hf;
    std::vector<stc_z64> hF;
    hf.reserve(nf);
    hF.reserve(nF);
    std::vector<stc_sys_desc> dd(sysd, sysd + nsys);
    std::vector<size_t> ofs_f(nsys), ofs_F(nsys), ofs_X(nsys), ofs_B(nsys);
    size_t cx_ = 0, cb_ = 0;
    for (int i = 0; i < nsys; ++i) {
        const int n = h->roms[sysd[i].rom_id].n;
        ofs_f[i] = hf.size();
        hf.insert(hf.end(), sysd[i].freq_hz, sysd[i].freq_hz + sysd[i].nfreq);
        ofs_F[i] = hF.size();
        for (int c = 0; c < sysd[i].nrhs; ++c)
            hF.insert(hF.end(), sysd[i].F + (size_t)c * sysd[i].ldf,
                      sysd[i].F + (size_t)c * sysd[i].ldf + n);
        ofs_X[i] = cx_;
        cx_ += (size_t)n * sysd[i].nrhs * sysd[i].nfreq;
        ofs_B[i] = cb_;
        cb_ += (size_t)sysd[i].nrhs * sysd[i].nfreq;
    }
    int rc;
    if ((rc = upload(h->freq, hf.data(), hf.size() * sizeof(double), st))) return rc;
    if ((rc = upload(h->F, hF.data(), hF.size() * sizeof(stc_z64), st))) return rc;
    if ((rc = h->X.ensure(std::max<size_t>(1, nX) * sizeof(stc_z64)))) return rc;
    if ((rc = h->info.ensure(std::max<size_t>(1, nf) * sizeof(int)))) return rc;
    if ((rc = h->berr.ensure(std::max<size_t>(1, nB) * sizeof(double)))) return rc;
    for (int i = 0; i < nsys; ++i) {
        const int n = h->roms[sysd[i].rom_id].n;
        dd[i].freq_hz = h->freq.as<double>() + ofs_f[i];
        dd[i].F = reinterpret_cast<const stc_z64*>(h->F.as<stc_z64>() + ofs_F[i]);
        dd[i].ldf = n;
        dd[i].X = h->X.as<stc_z64>() + ofs_X[i];
        dd[i].ldx = n;
        dd[i].info = h->info.as<int>() + ofs_f[i];
        dd[i].berr = sysd[i].berr ? h->berr.as<double>() + ofs_B[i] : nullptr;
    }
    std::vector<SysDesc> sys;
    std::vector<int2> items;
    if ((rc = build_batch(h, nsys, dd.data(), sys, items))) return rc;
    if (items.empty()) return STC_OK;
    const bool refine = mode == STC_REFINED_FP32;
    LaunchCfg cfg;
    if ((rc = configure(h, sys, (long long)items.size(), refine, 0, &cfg))) return rc;
    if ((rc = upload(h->sys, sys.data(), sys.size() * sizeof(SysDesc), st))) return rc;
    if ((rc = upload(h->items, items.data(), items.size() * sizeof(int2), st))) return rc;
    KParams kp = base_params(mode, tol, max_iter);
    kp.sys = h->sys.as<SysDesc>();
    kp.nsys = nsys;
    kp.items = h->items.as<int2>();
    kp.nitems = (int)items.size();
    kp.flags |= kWantBerr;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, st));
    rc = enqueue(h, cfg, kp, st);
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipStreamSynchronize(st));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    std::vector<stc_z64> xh(std::max<size_t>(1, nX));
    std::vector<int> ih(std::max<size_t>(1, nf));
    std::vector<double> bh(std::max<size_t>(1, nB));
    HIP_TRY(hipMemcpy(xh.data(), h->X.p, nX * sizeof(stc_z64), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(ih.data(), h->info.p, nf * sizeof(int), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(bh.data(), h->berr.p, nB * sizeof(double), hipMemcpyDeviceToHost));
    double bmax = 0.0;
    for (int i = 0; i < nsys; ++i) {
        const int n = h->roms[sysd[i].rom_id].n;
        for (int f = 0; f < sysd[i].nfreq; ++f)
            for (int c = 0; c < sysd[i].nrhs; ++c)
                memcpy(sysd[i].X + ((size_t)f * sysd[i].nrhs + c) * sysd[i].ldx,
                       &xh[ofs_X[i] + ((size_t)f * sysd[i].nrhs + c) * n], sizeof(stc_z64) * n);
        memcpy(sysd[i].info, &ih[ofs_f[i]], sizeof(int) * sysd[i].nfreq);
        if (sysd[i].berr) {
            memcpy(sysd[i].berr, &bh[ofs_B[i]], sizeof(double) * sysd[i].nfreq * sysd[i].nrhs);
            for (size_t q = 0; q < (size_t)sysd[i].nfreq * sysd[i].nrhs; ++q)
                if (std::isfinite(sysd[i].berr[q])) bmax = std::max(bmax, sysd[i].berr[q]);
        }
    }
    if (stats) {
        stats->seconds = ms * 1e-3;
        stats->points = (long long)items.size();
        stats->max_backward_error = bmax;
    }
    return status_from_info(ih.data(), (long long)nf, stats);
}

int stc_plan_create(stc_handle h, int nsys, const stc_sys_desc* dev_sys, int mode, double tol,
                    int max_iter, stc_plan* out) {
    if (!h || !out || nsys < 1 || !dev_sys || check_mode(mode)) return STC_EINVAL;
    *out = nullptr;
    DeviceGuard g(h->device);
    std::vector<SysDesc> sys;
    std::vector<int2> items;
    int rc = build_batch(h, nsys, dev_sys, sys, items);
    if (rc) return rc;
    if (items.empty()) return STC_EINVAL;
    const bool refine = mode == STC_REFINED_FP32;
    LaunchCfg cfg;
    if ((rc = plan_launch(h, sys.data(), (int)sys.size(), (long long)items.size(), refine, 0, 0, &cfg)))
        return rc;
    stc_plan_s* p = new stc_plan_s();
    p->h = h;
    auto fail = [&](int code) {
        if (p->order.ev) (void)hipEventDestroy(p->order.ev);
        p->sys.release(); p->items.release(); p->scratch.release(); p->counter.release();
        delete p;
        return code;
    };
    hipStream_t st = h->stream;
    // the plan's own factor slots and queue head: no later call on the handle
    // can resize or share them
    if ((rc = p->counter.ensure(256)) || (cfg.slot_bytes && (rc = p->scratch.ensure(cfg.slot_bytes * cfg.grid))) ||
        (rc = upload(p->sys, sys.data(), sys.size() * sizeof(SysDesc), st)) ||
        (rc = upload(p->items, items.data(), items.size() * sizeof(int2), st)))
        return fail(rc);
    if (hipEventCreateWithFlags(&p->order.ev, hipEventDisableTiming) != hipSuccess) return fail(STC_EDEVICE);
    if (hipStreamSynchronize(st) != hipSuccess) return fail(STC_EDEVICE);
    KParams kp = base_params(mode, tol, max_iter);
    kp.sys = p->sys.as<SysDesc>();
    kp.nsys = nsys;
    kp.items = p->items.as<int2>();
    kp.nitems = (int)items.size();
    for (const SysDesc& s : sys)
        if (s.berr) kp.flags |= kWantBerr;
    // the handle's previous launch is not part of the graph
    h->order.drain();
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return fail(STC_EDEVICE);
    rc = enqueue_raw(cfg, kp, p->counter.as<int>(), p->scratch.p, st);
    hipGraph_t graph = nullptr;
    const hipError_t ce = hipStreamEndCapture(st, &graph);
    if (rc || ce != hipSuccess || !graph) {
        if (graph) (void)hipGraphDestroy(graph);
        return fail(rc ? rc : STC_EDEVICE);
    }
    p->graph = graph;
    if (hipGraphInstantiate(&p->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(graph);
        p->graph = nullptr;
        return fail(STC_EDEVICE);
    }
    // the graph reads these ROMs' operators: keep them until the plan goes
    for (int i = 0; i < nsys; ++i) {
        if (std::find(p->roms.begin(), p->roms.end(), dev_sys[i].rom_id) != p->roms.end()) continue;
        p->roms.push_back(dev_sys[i].rom_id);
        h->roms[dev_sys[i].rom_id].plan_refs++;
    }
    h->live_plans++;
    *out = p;
    return STC_OK;
}

int stc_plan_execute(stc_plan p, void* stream) {
    if (!p || !p->exec) return STC_EINVAL;
    DeviceGuard g(p->h->device);
    hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : p->h->stream;
    // replays share the plan's slots and queue head: order them
    int rc = p->order.before(st);
    if (rc) return rc;
    HIP_TRY(hipGraphLaunch(p->exec, st));
    return p->order.after(st);
}

int stc_plan_destroy(stc_plan p) {
    if (!p) return STC_EINVAL;
    {
        DeviceGuard g(p->h->device);
        p->order.drain();
        (void)hipStreamSynchronize(p->h->stream);
        if (p->exec) (void)hipGraphExecDestroy(p->exec);
        if (p->graph) (void)hipGraphDestroy(p->graph);
        (void)hipEventDestroy(p->order.ev);
        p->sys.release();
        p->items.release();
        p->scratch.release();
        p->counter.release();
        for (int id : p->roms) {
            auto it = p->h->roms.find(id);
            if (it != p->h->roms.end() && it->second.plan_refs > 0) it->second.plan_refs--;
        }
        p->h->live_plans--;
    }
    delete p;
    return STC_OK;
}

// ---- adjoint ---------------------------------------------------------------------
int stc_adjoint(stc_handle h, int rom_id, int nfreq, const double* freq_hz, int nrhs,
                const stc_z64* F, int ldf, const stc_z64* c, int p, const double* Phi,
                int ldphi, const double* kdiag, int mode, double tol, int max_iter, double* J,
                double* dJ_ds, double* dJ_deta, int* info) {
    if (!h || nfreq < 1 || nrhs < 1 || p < 0 || check_mode(mode) || !freq_hz || !F || !c ||
        !J || !dJ_deta || !info || (p > 0 && (!Phi || !kdiag || !dJ_ds)))
        return STC_EINVAL;
    auto it = h->roms.find(rom_id);
    if (it == h->roms.end()) return STC_ENOROM;
    const int n = it->second.n;
    if (ldf < n || (p > 0 && ldphi < n)) return STC_EINVAL;
    DeviceGuard g(h->device);
    hipStream_t st = h->stream;
    int rc;
    if ((rc = upload(h->freq, freq_hz, sizeof(double) * nfreq, st))) return rc;
    if ((rc = upload_block(h->F, F, n, nrhs, ldf, st))) return rc;
    if ((rc = upload(h->c, c, sizeof(stc_z64) * n, st))) return rc;
    std::vector<double> ph((size_t)n * std::max(p, 1));
    for (int k = 0; k < p; ++k) memcpy(&ph[(size_t)k * n], Phi + (size_t)k * ldphi, sizeof(double) * n);
    if ((rc = upload(h->phi, ph.data(), ph.size() * sizeof(double), st))) return rc;
    if ((rc = upload(h->kd, p ? kdiag : ph.data(), sizeof(double) * std::max(p, 1), st))) return rc;
    if ((rc = h->adj.ensure(sizeof(double) * (size_t)nfreq * (p + 2)))) return rc;
    if ((rc = h->info.ensure(sizeof(int) * nfreq))) return rc;
    KParams kp = base_params(mode, tol, max_iter);
    kp.flags |= kAdjoint;
    std::vector<SysDesc> sys(1, base_desc(it->second));
    SysDesc& s = sys[0];
    s.nrhs = nrhs; s.nfreq = nfreq; s.ldf = n; s.ldx = n;
    s.freq = h->freq.as<double>();
    s.F = h->F.as<double2>();
    s.info = h->info.as<int>();
    s.p = p;
    s.c = h->c.as<double2>();
    s.Phi = h->phi.as<double>();
    s.kdiag = h->kd.as<double>();
    s.adj_out = h->adj.as<double>();
    const bool refine = mode == STC_REFINED_FP32;
    LaunchCfg cfg;
    if ((rc = configure(h, sys, nfreq, refine, 1, &cfg))) return rc;
    kp.one = sys[0];
    kp.nitems = nfreq;
    if ((rc = enqueue(h, cfg, kp, st))) return rc;
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<double> ao((size_t)nfreq * (p + 2));
    HIP_TRY(hipMemcpy(ao.data(), h->adj.p, ao.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(info, h->info.p, sizeof(int) * nfreq, hipMemcpyDeviceToHost));
    double Js = 0.0, Es = 0.0;
    std::vector<double> ds(std::max(p, 1), 0.0);
    for (int f = 0; f < nfreq; ++f) {
        if (info[f] > 0) continue;
        const double* o = &ao[(size_t)f * (p + 2)];
        Js += o[0];
        Es += o[1];
        for (int k = 0; k < p; ++k) ds[k] += o[2 + k];
    }
    *J = Js;
    *dJ_deta = Es;
    for (int k = 0; k < p; ++k) dJ_ds[k] = ds[k];
    return status_from_info(info, nfreq, nullptr);
}

// ---- full back-transformation ---------------------------------------------------------
int stc_back_transform(stc_handle h, int rom_id, int nfreq, const double* freq_hz,
                       const stc_z64* input_load, int nrhs, stc_z64* Y, stc_z64* H, int* info,
                       int mode, double tol, int max_iter, stc_stats* stats) {
    if (!h || nfreq < 0 || nrhs < 1 || check_mode(mode)) return STC_EINVAL;
    auto it = h->roms.find(rom_id);
    if (it == h->roms.end()) return STC_ENOROM;
    const RomDev& r = it->second;
    if (r.nin < 1 || r.nout < 1) return STC_EINVAL;
    if (stats) memset(stats, 0, sizeof(*stats));
    if (nfreq == 0) return STC_OK;
    if (!freq_hz || !input_load || !Y || !info) return STC_EINVAL;
    const int n = r.n, nin = r.nin, nout = r.nout;
    const int ncol = nrhs + (H ? nin : 0);
    DeviceGuard g(h->device);
    hipStream_t st = h->stream;
    int rc;
    // right-hand sides [B_R * inputLoad | B_R] on the device (the projection is
    // frequency independent, hoisted out of the loop of
    // KrylovROMSubstructure.cpp:1423-1424)
    if ((rc = upload(h->freq, freq_hz, sizeof(double) * nfreq, st))) return rc;
    if ((rc = upload(h->load, input_load, sizeof(stc_z64) * (size_t)nin * nrhs, st))) return rc;
    if ((rc = h->F.ensure(sizeof(stc_z64) * (size_t)n * ncol))) return rc;
    if ((rc = h->info.ensure(sizeof(int) * nfreq))) return rc;
    if ((rc = h->berr.ensure(sizeof(stc_z64) * (size_t)nfreq * nout * ncol))) return rc;
    {
        const long long cnt = (long long)n * ncol;
        hipLaunchKernelGGL(load_rhs_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st,
                           r.B.as<double2>(), n, nin, h->load.as<double2>(), nrhs, H ? 1 : 0,
                           h->F.as<double2>());
        HIP_TRY(hipGetLastError());
    }
    // y = C_R X and H_R = C_R A^{-1} B_R (:1438-1448), formed by the sweep kernel
    // from each point's solution; X itself is never written to memory
    double2* Yd = h->berr.as<double2>();
    KParams kp = base_params(mode, tol, max_iter);
    std::vector<SysDesc> sys(1, base_desc(r));
    SysDesc& s = sys[0];
    s.nrhs = ncol; s.nfreq = nfreq; s.ldf = n; s.ldx = n;
    s.freq = h->freq.as<double>();
    s.F = h->F.as<double2>();
    s.X = nullptr;
    s.C = r.C.as<double2>();
    s.Y = Yd;
    s.nout = nout;
    s.info = h->info.as<int>();
    LaunchCfg cfg;
    if ((rc = configure(h, sys, nfreq, mode == STC_REFINED_FP32, 0, &cfg))) return rc;
    kp.one = sys[0];
    kp.nitems = nfreq;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, st));
    if ((rc = enqueue(h, cfg, kp, st))) return rc;
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipStreamSynchronize(st));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::vector<stc_z64> yh((size_t)nfreq * nout * ncol);
    HIP_TRY(hipMemcpy(yh.data(), Yd, yh.size() * sizeof(stc_z64), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(info, h->info.p, sizeof(int) * nfreq, hipMemcpyDeviceToHost));
    for (int f = 0; f < nfreq; ++f) {
        const stc_z64* yf = &yh[(size_t)f * nout * ncol];
        memcpy(Y + (size_t)f * nout * nrhs, yf, sizeof(stc_z64) * nout * nrhs);
        if (H) memcpy(H + (size_t)f * nout * nin, yf + (size_t)nout * nrhs, sizeof(stc_z64) * nout * nin);
    }
    if (stats) {
        stats->seconds = ms * 1e-3;
        stats->points = nfreq;
    }
    return status_from_info(info, nfreq, stats);
}

// ---- explicit matrices --------------------------------------------------------
// The sweep kernel factors A = (1+i eta)K + i w D - w^2 M; with eta = 0, w = 0
// (no frequency list) and one K block per frequency index it factors the given
// matrices themselves (M aliases K and is multiplied by zero).
int stc_gesv_batched_dev(stc_handle h, int n, int batch, const stc_z64* d_A, long long strideA,
                         int nrhs, const stc_z64* d_B, int ldb, long long strideB, stc_z64* d_X,
                         int ldx, int* d_info, double* d_berr, int mode, double tol,
                         int max_iter, void* stream) {
    if (!h || n < 1 || n > 1024 || batch < 0 || nrhs < 1 || check_mode(mode)) return STC_EINVAL;
    if (strideA < (long long)n * n || ldb < n || ldx < n ||
        (strideB != 0 && strideB < (long long)ldb * nrhs))
        return STC_EINVAL;
    if (batch == 0) return STC_OK;
    if (!d_A || !d_B || !d_X || !d_info) return STC_EINVAL;
    DeviceGuard g(h->device);
    hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : h->stream;
    KParams kp = base_params(mode, tol, max_iter);
    std::vector<SysDesc> sys(1);
    SysDesc& s = sys[0];
    memset(&s, 0, sizeof(s));
    s.n = n;
    s.real_ops = 0;
    s.has_D = 0;
    s.eta = 0.0;
    s.K = reinterpret_cast<const double*>(d_A);
    s.M = s.K;
    s.op_stride = 2 * strideA;
    s.f_stride = strideB;
    s.nrhs = nrhs;
    s.nfreq = batch;
    s.ldf = ldb;
    s.ldx = ldx;
    s.freq = nullptr;
    s.F = reinterpret_cast<const double2*>(d_B);
    s.X = reinterpret_cast<double2*>(d_X);
    s.info = d_info;
    s.berr = d_berr;
    LaunchCfg cfg;
    int rc = configure(h, sys, batch, mode == STC_REFINED_FP32, 0, &cfg);
    if (rc) return rc;
    kp.one = sys[0];
    kp.nitems = batch;
    if (d_berr) kp.flags |= kWantBerr;
    return enqueue(h, cfg, kp, st);
}

int stc_gesv_batched(stc_handle h, int n, int batch, const stc_z64* A, int lda, long long strideA,
                     int nrhs, const stc_z64* B, int ldb, long long strideB, stc_z64* X, int ldx,
                     long long strideX, int* info, double* berr, int mode, double tol,
                     int max_iter, stc_stats* stats) {
    if (!h || n < 1 || n > 1024 || batch < 0 || nrhs < 1 || check_mode(mode)) return STC_EINVAL;
    if (lda < n || ldb < n || ldx < n || strideA < (long long)lda * n ||
        (strideB != 0 && strideB < (long long)ldb * nrhs) || strideX < (long long)ldx * nrhs)
        return STC_EINVAL;
    if (stats) memset(stats, 0, sizeof(*stats));
    if (batch == 0) return STC_OK;
    if (!A || !B || !X || !info) return STC_EINVAL;
    DeviceGuard g(h->device);
    hipStream_t st = h->stream;
    const size_t nn = (size_t)n * n, nb = strideB ? (size_t)batch : 1;
    int rc;
    if ((rc = h->A.ensure(sizeof(stc_z64) * nn * batch))) return rc;
    if ((rc = h->F.ensure(sizeof(stc_z64) * (size_t)n * nrhs * nb))) return rc;
    if ((rc = h->X.ensure(sizeof(stc_z64) * (size_t)n * nrhs * batch))) return rc;
    if ((rc = h->info.ensure(sizeof(int) * batch))) return rc;
    if (berr && (rc = h->berr.ensure(sizeof(double) * (size_t)batch * nrhs))) return rc;
    // pack the matrices and right-hand sides (leading dimension n); one copy
    // when the caller's layout is packed already
    if (lda == n && strideA == (long long)nn)
        HIP_TRY(hipMemcpyAsync(h->A.p, A, sizeof(stc_z64) * nn * batch, hipMemcpyHostToDevice, st));
    else
        for (int j = 0; j < batch; ++j)
            HIP_TRY(hipMemcpy2DAsync(h->A.as<stc_z64>() + (size_t)j * nn, sizeof(stc_z64) * n,
                                     A + (size_t)j * strideA, sizeof(stc_z64) * lda,
                                     sizeof(stc_z64) * n, n, hipMemcpyHostToDevice, st));
    if (ldb == n && (nb == 1 || strideB == (long long)n * nrhs))
        HIP_TRY(hipMemcpyAsync(h->F.p, B, sizeof(stc_z64) * n * nrhs * nb, hipMemcpyHostToDevice,
                               st));
    else
        for (size_t j = 0; j < nb; ++j)
            HIP_TRY(hipMemcpy2DAsync(h->F.as<stc_z64>() + j * n * nrhs, sizeof(stc_z64) * n,
                                     B + j * strideB, sizeof(stc_z64) * ldb, sizeof(stc_z64) * n,
                                     nrhs, hipMemcpyHostToDevice, st));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, st));
    rc = stc_gesv_batched_dev(h, n, batch, h->A.as<stc_z64>(), (long long)nn, nrhs,
                              h->F.as<stc_z64>(), n, strideB ? (long long)n * nrhs : 0,
                              h->X.as<stc_z64>(), n, h->info.as<int>(),
                              berr ? h->berr.as<double>() : nullptr, mode, tol, max_iter, st);
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipStreamSynchronize(st));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (ldx == n && strideX == (long long)n * nrhs)
        HIP_TRY(hipMemcpy(X, h->X.p, sizeof(stc_z64) * n * nrhs * batch, hipMemcpyDeviceToHost));
    else
        for (int j = 0; j < batch; ++j)
            HIP_TRY(hipMemcpy2D(X + (size_t)j * strideX, sizeof(stc_z64) * ldx,
                                h->X.as<stc_z64>() + (size_t)j * n * nrhs, sizeof(stc_z64) * n,
                                sizeof(stc_z64) * n, nrhs, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(info, h->info.p, sizeof(int) * batch, hipMemcpyDeviceToHost));
    if (berr)
        HIP_TRY(hipMemcpy(berr, h->berr.p, sizeof(double) * batch * nrhs, hipMemcpyDeviceToHost));
    if (stats) {
        stats->seconds = ms * 1e-3;
        stats->points = batch;
        if (berr) {
            double m = 0.0;
            for (long long q = 0; q < (long long)batch * nrhs; ++q)
                if (std::isfinite(berr[q])) m = std::max(m, berr[q]);
            stats->max_backward_error = m;
        }
    }
    return status_from_info(info, batch, stats);
}

}  // extern "C"
