// libpcube: host side of the C ABI declared in include/pcube.h.
//
// One process drives one device through one default-priority stream.  Device
// memory comes from a capped, size-class caching pool so that per-batch
// buffers do not call hipMalloc in steady state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/pcube.h"
#include "pcube_device.h"
#include "pcube_wheel.h"
#include "pcube_staged.h"
#include "pcube_cubecount.h"

using namespace pcube;

// Device scratch of pcube_cube_count_noalloc, for callers that cannot allocate
// device memory.  Layout: the counter at byte 0, the odd primes <= p (at most
// 302 for p <= 2000) at kNoAllocPrimes, the root bitset at kNoAllocGood.  Each
// device holds its own copy for the life of the process; host code reaches it
// only through the address hipGetSymbolAddress returns on that device.
constexpr uint32_t kNoAllocPrimes = 256;                                  // byte offset
constexpr uint32_t kNoAllocMaxPrimes = 512;
constexpr uint32_t kNoAllocGood = kNoAllocPrimes + kNoAllocMaxPrimes * 4;  // byte offset, 2304
constexpr uint32_t kNoAllocGoodWords = 1u << 16;                          // 2^21 odd roots
constexpr uint32_t kNoAllocBytes = kNoAllocGood + kNoAllocGoodWords * 4;   // 264,448
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace {

std::recursive_mutex g_mu;
// the last error of the calling thread: calls on other threads leave it alone
thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(PCUBE_DEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                 \
    } while (0)

#define TRY(expr)                  \
    do {                           \
        int rc_ = (expr);          \
        if (rc_ != PCUBE_OK) return rc_; \
    } while (0)

struct Context {
    int device = -1;
    hipStream_t stream = nullptr;
    uint64_t cap = 0, reserved = 0, peak = 0;
    std::unordered_map<void*, uint64_t> live;      // address -> class size
    std::multimap<uint64_t, void*> cached;          // class size -> free block
    int compute_units = 0;
    uint32_t lds_per_block = 65536;                  // largest LDS one workgroup may use
    char name[256] = {0};
    // device address of `scratch` per device, from hipGetSymbolAddress at init
    static constexpr int kMaxDevices = 64;
    void* scratch_at[kMaxDevices] = {};
} g;

// HIP's current device belongs to the calling thread: an entry point called
// from a thread other than pcube_init's makes the library's device current
// there first, so its allocations land on that device
thread_local int t_device = -1;

struct Entry {
    std::lock_guard<std::recursive_mutex> lock{g_mu};
    Entry() {
        if (g.device >= 0 && t_device != g.device && hipSetDevice(g.device) == hipSuccess) t_device = g.device;
    }
};

uint64_t size_class(uint64_t bytes) {
    if (bytes == 0) bytes = 1;
    if (bytes <= (1ull << 20)) {
        uint64_t c = 256;
        while (c < bytes) c <<= 1;
        return c;
    }
    const uint64_t mb = 1ull << 20;
    return (bytes + mb - 1) / mb * mb;
}

// Cached blocks may still be read or written by work queued on the stream
// (frees do not wait), so drain it before handing them back to the driver.
void release_cache() {
    if (g.cached.empty()) return;
    if (g.stream) (void)hipStreamSynchronize(g.stream);
    for (auto& kv : g.cached) {
        (void)hipFree(kv.second);
        g.reserved -= kv.first;
    }
    g.cached.clear();
}

int dev_alloc(uint64_t bytes, void** out) {
    if (g.device < 0) return fail(PCUBE_INVALID, "pcube_init has not been called");
    const uint64_t c = size_class(bytes);
    auto it = g.cached.find(c);
    if (it != g.cached.end()) {
        *out = it->second;
        g.cached.erase(it);
        g.live[*out] = c;
        return PCUBE_OK;
    }
    if (g.reserved + c > g.cap) release_cache();
    if (g.reserved + c > g.cap)
        return fail(PCUBE_NOMEM, "device memory cap exceeded: %llu reserved + %llu requested > cap %llu",
                    (unsigned long long)g.reserved, (unsigned long long)c, (unsigned long long)g.cap);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, c);
    if (e != hipSuccess) {
        release_cache();
        e = hipMalloc(&p, c);
        if (e != hipSuccess) return fail(PCUBE_NOMEM, "hipMalloc(%llu): %s", (unsigned long long)c, hipGetErrorString(e));
    }
    g.reserved += c;
    g.peak = std::max(g.peak, g.reserved);
    g.live[p] = c;
    *out = p;
    return PCUBE_OK;
}

int dev_free(void* p) {
    if (!p) return PCUBE_OK;
    auto it = g.live.find(p);
    if (it == g.live.end()) return fail(PCUBE_INVALID, "free of an address the pool does not own");
    g.cached.emplace(it->second, p);
    g.live.erase(it);
    return PCUBE_OK;
}

// A grow-only device buffer owned by a library object.
struct Buffer {
    void* p = nullptr;
    uint64_t bytes = 0;
    int ensure(uint64_t need) {
        if (need <= bytes && p) return PCUBE_OK;
        if (p) dev_free(p);
        p = nullptr;
        bytes = 0;
        TRY(dev_alloc(need, &p));
        bytes = need;
        return PCUBE_OK;
    }
    void release() {
        if (p) dev_free(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const { return (T*)p; }
};

uint32_t blocks_for(uint64_t threads, uint32_t per_block) {
    return (uint32_t)std::max<uint64_t>(1, (threads + per_block - 1) / per_block);
}

// ---------------------------------------------------------------------------
// Scan and bucket sort helpers

struct ScanWork {
    Buffer partial, partial_off, partial_total;
};

// out[0..n) = exclusive scan of in, out[n] = total.  n <= 4096 * 4096.
int scan(const uint32_t* in, uint32_t* out, uint64_t n, ScanWork& w) {
    const uint64_t nblk = (n + kScanSegment - 1) / kScanSegment;
    if (nblk == 0 || nblk > (uint64_t)kScanSegment) return fail(PCUBE_INVALID, "scan size %llu unsupported", (unsigned long long)n);
    TRY(w.partial.ensure(nblk * 4));
    TRY(w.partial_off.ensure(nblk * 4 + 4));
    TRY(w.partial_total.ensure(4));
    hipLaunchKernelGGL(k_scan_local, dim3((uint32_t)nblk), dim3(kScanThreads), 0, g.stream, in, out, n,
                       w.partial.as<uint32_t>());
    if (nblk == 1) {
        HIP_TRY(hipMemcpyAsync(out + n, w.partial.as<uint32_t>(), 4, hipMemcpyDeviceToDevice, g.stream));
        return PCUBE_OK;
    }
    hipLaunchKernelGGL(k_scan_local, dim3(1), dim3(kScanThreads), 0, g.stream, w.partial.as<uint32_t>(),
                       w.partial_off.as<uint32_t>(), nblk, w.partial_total.as<uint32_t>());
    hipLaunchKernelGGL(k_scan_add, dim3(blocks_for(n, 256)), dim3(256), 0, g.stream, out, n,
                       w.partial_off.as<uint32_t>());
    HIP_TRY(hipMemcpyAsync(out + n, w.partial_total.as<uint32_t>(), 4, hipMemcpyDeviceToDevice, g.stream));
    return PCUBE_OK;
}

int fill32(uint32_t* p, uint64_t n, uint32_t v) {
    if (!n) return PCUBE_OK;
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min<uint64_t>(blocks_for(n, 256), 4096)), dim3(256), 0, g.stream, p, n, v);
    return PCUBE_OK;
}
int fill64(uint64_t* p, uint64_t n, uint64_t v) {
    if (!n) return PCUBE_OK;
    hipLaunchKernelGGL(k_fill_u64, dim3(std::min<uint64_t>(blocks_for(n, 256), 4096)), dim3(256), 0, g.stream, p, n, v);
    return PCUBE_OK;
}

// Smallest shift with ((span - 1) >> shift) + 1 <= target buckets.
uint32_t bucket_shift(uint64_t span, uint64_t target) {
    if (target < 1) target = 1;
    uint32_t s = 0;
    while (s < 63 && ((span - 1) >> s) + 1 > target) ++s;
    return s;
}

struct BucketWork {
    Buffer hist, offsets;
    ScanWork scan;
};

// Sort n device values lying in [low, low + span) into out.
int bucket_sort(const uint64_t* in, uint64_t n, uint64_t low, uint64_t span, uint64_t* out, BucketWork& w) {
    if (!n) return PCUBE_OK;
    uint64_t target = 1;
    while (target * 2 <= n / 2 && target < (1ull << 22)) target *= 2;
    const uint32_t shift = bucket_shift(span, target);
    const uint64_t nb = ((span - 1) >> shift) + 1;
    TRY(w.hist.ensure(nb * 8));
    TRY(w.offsets.ensure(nb * 4 + 4));
    uint32_t* hist = w.hist.as<uint32_t>();
    uint32_t* cursor = hist + nb;
    TRY(fill32(hist, 2 * nb, 0));
    hipLaunchKernelGGL(k_bucket_count, dim3(blocks_for(n, 256)), dim3(256), 0, g.stream, in, n, low, shift, hist);
    TRY(scan(hist, w.offsets.as<uint32_t>(), nb, w.scan));
    hipLaunchKernelGGL(k_bucket_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, g.stream, in, n, low, shift,
                       w.offsets.as<uint32_t>(), cursor, out);
    hipLaunchKernelGGL(k_bucket_sort, dim3(blocks_for(nb, 256)), dim3(256), 0, g.stream, out,
                       w.offsets.as<uint32_t>(), nb, n);
    return PCUBE_OK;
}

// ---------------------------------------------------------------------------
// Focus objects

// CRT wheel of a focus: inner ring prefix with modulus M1 < 2^40 and its
// sorted residues R1, outer rings tested per candidate (pcube_wheel.h).
// Pinned host staging for the small uploads and downloads of one call (plans,
// queries, counters).  Every entry point that uses it synchronises before it
// returns, and calls are serialised, so one grow-only pair serves them all.
struct Pinned {
    uint8_t* p = nullptr;
    uint64_t bytes = 0;
    unsigned flags = hipHostMallocDefault;
    int ensure(uint64_t need) {
        if (need <= bytes && p) return PCUBE_OK;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
        const uint64_t n = std::max<uint64_t>(need, 1ull << 16);
        HIP_TRY(hipHostMalloc((void**)&p, n, flags));
        bytes = n;
        return PCUBE_OK;
    }
    template <class T>
    T* as() const { return (T*)p; }
};
Pinned g_up_host, g_down_host;
Pinned g_lease_up, g_lease_down;   // a lease pass's subtile queries; its counters and status words
// Coherent pinned memory that kernels write results into directly: the
// stores reach host memory before the kernel's completion is signalled, so a
// stream synchronisation makes them visible without a copy.
Pinned g_result_host{nullptr, 0, hipHostMallocCoherent};
Buffer g_up_dev;   // device image of the current call's uploads

struct Wheel {
    bool ready = false;
    uint64_t M1 = 1;
    uint32_t n1 = 0, nouter = 0, allow_words = 0, packed = 0;
    std::vector<uint64_t> R1host;
    Modulus outer[kMaxOuterRings];
    double rate = 1.0;      // expected fraction of candidates that pass the outer rings
    Buffer R1, allow, slots, wave_count, seg_excl, chunk_total, win_info;
    Buffer flag;            // overflow flag of one-window launches (zero except while a retry is due)
    void release() {
        R1.release(); allow.release(); slots.release(); wave_count.release(); seg_excl.release();
        chunk_total.release(); win_info.release(); flag.release();
    }
};

struct Focus {
    uint64_t M = 0, nc = 0, nd = 0, nbk = 0;
    uint32_t dshift = 0;
    Buffer C, D, dir;
    BucketWork work;        // window buckets and scan scratch
    Buffer slots, flag;     // single-pass window slots and their overflow flag
    Wheel wheel;
    ~Focus() {
        C.release(); D.release(); dir.release();
        work.hist.release(); work.offsets.release(); slots.release(); flag.release();
        work.scan.partial.release(); work.scan.partial_off.release(); work.scan.partial_total.release();
        wheel.release();
    }
};
std::vector<std::unique_ptr<Focus>> g_focus;

Focus* focus_at(int id) {
    if (id < 0 || id >= (int)g_focus.size() || !g_focus[id]) return nullptr;
    return g_focus[id].get();
}

int build_half(uint64_t M, const uint64_t* terms, const uint32_t* sizes, uint32_t nrings, Buffer& out, uint64_t* n,
               BucketWork& work) {
    uint64_t count = 1, nterms = 0;
    for (uint32_t k = 0; k < nrings; ++k) {
        if (sizes[k] == 0) return fail(PCUBE_INVALID, "empty ring");
        count *= sizes[k];
        nterms += sizes[k];
        if (count > (1ull << 31)) return fail(PCUBE_INVALID, "focus half too large");
    }
    *n = count;
    Buffer dterms, dsizes, unsorted;
    TRY(dterms.ensure(std::max<uint64_t>(8, nterms * 8)));
    TRY(dsizes.ensure(std::max<uint64_t>(4, nrings * 4)));
    TRY(unsorted.ensure(count * 8));
    TRY(out.ensure(count * 8));
    if (nterms) HIP_TRY(hipMemcpyAsync(dterms.p, terms, nterms * 8, hipMemcpyHostToDevice, g.stream));
    if (nrings) HIP_TRY(hipMemcpyAsync(dsizes.p, sizes, nrings * 4, hipMemcpyHostToDevice, g.stream));
    hipLaunchKernelGGL(k_partial, dim3(blocks_for(count, 256)), dim3(256), 0, g.stream, dterms.as<uint64_t>(),
                       dsizes.as<uint32_t>(), nrings, M, count, unsorted.as<uint64_t>());
    TRY(bucket_sort(unsorted.as<uint64_t>(), count, 0, M, out.as<uint64_t>(), work));
    HIP_TRY(hipStreamSynchronize(g.stream));
    dterms.release(); dsizes.release(); unsorted.release();
    return PCUBE_OK;
}

// Expected number of coordinates of focus f in a window of the given span.
double expected_count(const Focus& f, uint64_t span) {
    return (double)span / (double)f.M * (double)f.nc * (double)f.nd;
}

// Enqueue a window [low, high) of focus f into out (capacity cap).  The
// device total is left at f.work.offsets[nb] (*nb_out receives nb).  Small
// windows take the single-pass slot path (*single = true) whose overflow flag
// the caller must check (f.flag); exact = true forces the two-pass path.
int window_enqueue(Focus& f, uint64_t low, uint64_t high, uint64_t* out, uint64_t cap, bool exact,
                   uint64_t* nb_out, bool* single) {
    const uint64_t per0 = low / f.M, per1 = (high - 1) / f.M, nper = per1 - per0 + 1;
    if (nper > (1u << 20)) return fail(PCUBE_INVALID, "window spans too many periods");
    const uint64_t span = high - low;
    const double expected = expected_count(f, span);
    WindowArgs w;
    w.C = f.C.as<uint64_t>();
    w.D = f.D.as<uint64_t>();
    w.dir = f.dir.as<uint32_t>();
    w.nc = f.nc; w.nd = f.nd; w.nbk = f.nbk; w.M = f.M;
    w.low = low; w.high = high; w.per0 = per0; w.cap = cap;
    w.dshift = f.dshift; w.nper = (uint32_t)nper;
    w.nchunks = (uint32_t)((f.nd + kChunk - 1) / kChunk);
    const uint64_t threads = (uint64_t)w.nper * w.nchunks;
    *single = !exact && expected <= (double)(1u << 20);
    uint64_t target = 1;
    if (*single) {
        while ((double)target < 2 * expected && target < (1ull << 22)) target *= 2;
    } else {
        while ((double)(target * 2) <= expected / 2 && target < (1ull << 22)) target *= 2;
    }
    w.bshift = bucket_shift(span, target);
    const uint64_t nb = ((span - 1) >> w.bshift) + 1;
    TRY(f.work.hist.ensure(nb * 8));
    TRY(f.work.offsets.ensure(nb * 4 + 4));
    uint32_t* hist = f.work.hist.as<uint32_t>();
    uint32_t* cursor = hist + nb;
    uint32_t* offsets = f.work.offsets.as<uint32_t>();
    *nb_out = nb;
    if (*single) {
        TRY(f.slots.ensure(nb * kSlotsPerBucket * 8));
        TRY(f.flag.ensure(4));
        TRY(fill32(hist, nb, 0));
        TRY(fill32(f.flag.as<uint32_t>(), 1, 0));
        hipLaunchKernelGGL(k_window<kSlots>, dim3(blocks_for(threads, 256)), dim3(256), 0, g.stream, w, hist,
                           (const uint32_t*)offsets, cursor, f.slots.as<uint64_t>(), f.flag.as<uint32_t>());
        TRY(scan(hist, offsets, nb, f.work.scan));
        hipLaunchKernelGGL(k_slots_compact, dim3(blocks_for(nb, 256)), dim3(256), 0, g.stream,
                           (const uint64_t*)f.slots.as<uint64_t>(), (const uint32_t*)offsets, nb, out, cap);
        return PCUBE_OK;
    }
    TRY(fill32(hist, 2 * nb, 0));
    hipLaunchKernelGGL(k_window<kCount>, dim3(blocks_for(threads, 256)), dim3(256), 0, g.stream, w, hist,
                       (const uint32_t*)offsets, cursor, out, (uint32_t*)nullptr);
    TRY(scan(hist, offsets, nb, f.work.scan));
    hipLaunchKernelGGL(k_window<kEmit>, dim3(blocks_for(threads, 256)), dim3(256), 0, g.stream, w, hist,
                       (const uint32_t*)offsets, cursor, out, (uint32_t*)nullptr);
    hipLaunchKernelGGL(k_bucket_sort, dim3(blocks_for(nb, 256)), dim3(256), 0, g.stream, out, offsets, nb, cap);
    return PCUBE_OK;
}

// Run a window to completion and return its size; reruns the exact path when
// a single-pass bucket overflowed its slots.
int window_run(Focus& f, uint64_t low, uint64_t high, uint64_t* out, uint64_t cap, uint64_t* count) {
    uint64_t nb = 0;
    bool single = false;
    TRY(window_enqueue(f, low, high, out, cap, false, &nb, &single));
    uint32_t total = 0, flag = 0;
    HIP_TRY(hipMemcpyAsync(&total, f.work.offsets.as<uint32_t>() + nb, 4, hipMemcpyDeviceToHost, g.stream));
    if (single) HIP_TRY(hipMemcpyAsync(&flag, f.flag.p, 4, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    if (flag) {
        TRY(window_enqueue(f, low, high, out, cap, true, &nb, &single));
        HIP_TRY(hipMemcpyAsync(&total, f.work.offsets.as<uint32_t>() + nb, 4, hipMemcpyDeviceToHost, g.stream));
        HIP_TRY(hipStreamSynchronize(g.stream));
    }
    *count = total;
    return PCUBE_OK;
}

uint64_t modinv(uint64_t a, uint64_t m) {
    // inverse of a modulo m (gcd 1), extended Euclid in signed 128-bit
    __int128 t = 0, nt = 1, r = m, nr = a % m;
    while (nr) {
        __int128 q = r / nr, x;
        x = t - q * nt; t = nt; nt = x;
        x = r - q * nr; r = nr; nr = x;
    }
    if (t < 0) t += m;
    return (uint64_t)t;
}

Modulus make_modulus(uint32_t q) {
    Modulus m;
    memset(&m, 0, sizeof m);
    m.q = q;
    m.inv = (uint32_t)((1ull << 32) / q);
    m.r32 = (uint32_t)((1ull << 32) % q);
    m.r64 = (uint32_t)(((unsigned __int128)1 << 64) % q);
    return m;
}

constexpr uint32_t kMaxInnerResidues = 1u << 27;           // 8-byte entries: at most 1 GiB
constexpr uint64_t kInnerModulusLimit = 1ull << kResidueBits;   // inner residues fit an entry's low bits

// Plan the segments of a set of windows; returns the total segment count.
uint64_t wheel_plan(const Wheel& wh, uint32_t n, const uint64_t* lows, const uint64_t* highs,
                    std::vector<WheelWindow>& plan, std::vector<uint64_t>& seg_first) {
    plan.resize(n);
    seg_first.resize(n + 1);
    uint64_t total = 0;
    const uint64_t S = kWheelSegment;
    for (uint32_t w = 0; w < n; ++w) {
        WheelWindow ww;
        memset(&ww, 0, sizeof ww);
        ww.low = lows[w];
        ww.high = highs[w];
        ww.seg_first = total;
        seg_first[w] = total;
        if (highs[w] > lows[w]) {
            const uint64_t t0 = lows[w] / wh.M1, t1 = (highs[w] - 1) / wh.M1;
            const uint64_t lo_r = lows[w] - t0 * wh.M1, hi_r = highs[w] - t1 * wh.M1;
            ww.t0 = t0;
            ww.nt = (uint32_t)(t1 - t0 + 1);
            ww.ib = (uint32_t)(std::lower_bound(wh.R1host.begin(), wh.R1host.end(), lo_r) - wh.R1host.begin());
            ww.ie = hi_r >= wh.M1 ? wh.n1
                                  : (uint32_t)(std::lower_bound(wh.R1host.begin(), wh.R1host.end(), hi_r) - wh.R1host.begin());
            uint64_t segs;
            if (ww.nt == 1) {
                ww.s0 = ww.ie > ww.ib ? (uint32_t)((ww.ie - ww.ib + S - 1) / S) : 0;
                segs = ww.s0;
            } else {
                ww.s0 = (uint32_t)((wh.n1 - ww.ib + S - 1) / S);
                ww.sm = (uint32_t)((wh.n1 + S - 1) / S);
                ww.sl = (uint32_t)((ww.ie + S - 1) / S);
                segs = ww.s0 + (uint64_t)(ww.nt - 2) * ww.sm + ww.sl;
            }
            if (segs >= (1ull << 32)) segs = 1ull << 40;  // rejected by wheel_check
            total += segs;
        }
        plan[w] = ww;
    }
    seg_first[n] = total;
    return total;
}

// Slot entries per wave for a wheel: twice the expected survivors of 1024
// candidates plus a margin, in steps of 32; 1024 cannot overflow.
uint32_t wheel_slot_cap(const Wheel& wh) {
    const double expect = wh.rate * kWheelWaveSpan;
    const uint64_t want = ((uint64_t)(2.0 * expect) + 32 + 31) / 32 * 32;
    return (uint32_t)std::min<uint64_t>(kWheelWaveSpan, want);
}

// Words of the device image of one wheel launch's inputs: the window plans,
// seg_first (n + 1) and a zero overflow flag.
uint64_t wheel_image_words(uint32_t n) {
    static_assert(sizeof(WheelWindow) % 8 == 0, "WheelWindow must be 8-byte words");
    return (uint64_t)n * (sizeof(WheelWindow) / 8) + (n + 1) + 1;
}
void wheel_image(const std::vector<WheelWindow>& plan, const std::vector<uint64_t>& seg_first, uint64_t* img) {
    const uint32_t n = (uint32_t)plan.size();
    if (n) memcpy(img, plan.data(), n * sizeof(WheelWindow));
    memcpy(img + (uint64_t)n * (sizeof(WheelWindow) / 8), seg_first.data(), (n + 1) * 8);
    img[wheel_image_words(n) - 1] = 0;
}

struct WheelInputs {
    const WheelWindow* windows;   // null: one window, win0
    const uint64_t* seg_first;    // null with win0
    uint32_t* flag;
    uint32_t n;
    uint64_t nseg;
    WheelWindow win0;
};
WheelInputs wheel_inputs(uint64_t* img_dev, uint32_t n, uint64_t nseg) {
    WheelInputs in{};
    in.windows = (const WheelWindow*)img_dev;
    in.seg_first = img_dev + (uint64_t)n * (sizeof(WheelWindow) / 8);
    in.flag = (uint32_t*)(img_dev + wheel_image_words(n) - 1);
    in.n = n;
    in.nseg = nseg;
    return in;
}

int wheel_check(uint64_t nseg) {
    if (nseg >= (1ull << 31)) return fail(PCUBE_INVALID, "window set too large for one pass");
    return PCUBE_OK;
}

// Enqueue the test, segment-scan and compaction passes of a wheel launch
// whose inputs are on the device.  Coordinates at positions >= cap are
// dropped.  wh.win_info receives the n + 1 window offsets (the last is the
// total) and, at n + 1, the overflow flag.
// Scratch buffers and kernel arguments of one wheel pass (in.nseg > 0).
int wheel_setup(Wheel& wh, const WheelInputs& in, uint32_t slot_cap, WheelArgs& a, uint64_t& nchunks) {
    const uint64_t nwaves = in.nseg * kWheelWaves;
    nchunks = (in.nseg + kWheelScanChunk - 1) / kWheelScanChunk;
    TRY(wh.slots.ensure(nwaves * slot_cap * 8));
    TRY(wh.wave_count.ensure(nwaves * 4));
    TRY(wh.seg_excl.ensure(in.nseg * 4));
    TRY(wh.chunk_total.ensure(nchunks * 8));
    memset(&a, 0, sizeof a);
    a.R1 = wh.R1.as<uint64_t>();
    a.windows = in.windows;
    a.seg_first = in.seg_first;
    a.allow = wh.allow.as<uint32_t>();
    a.slots = wh.slots.as<uint64_t>();
    a.wave_count = wh.wave_count.as<uint32_t>();
    a.flag = in.flag;
    a.M1 = wh.M1; a.nseg = in.nseg;
    a.n1 = wh.n1; a.nwin = in.n; a.nouter = wh.nouter; a.allow_words = wh.allow_words;
    a.packed = wh.packed; a.slot_cap = slot_cap;
    for (uint32_t j = 0; j < wh.nouter; ++j) a.outer[j] = wh.outer[j];
    a.win0 = in.win0;
    return PCUBE_OK;
}

int wheel_enqueue(Wheel& wh, const WheelInputs& in, uint64_t* out, uint64_t cap, uint32_t slot_cap) {
    TRY(wh.win_info.ensure(((uint64_t)in.n + 2) * 8));
    if (in.nseg == 0) return fill64(wh.win_info.as<uint64_t>(), (uint64_t)in.n + 2, 0);
    WheelArgs a;
    uint64_t nchunks = 0;
    TRY(wheel_setup(wh, in, slot_cap, a, nchunks));
    void* args[] = {&a};
    HIP_TRY(hipLaunchKernel(wheel_kernel(wh.packed, wh.nouter), dim3((uint32_t)in.nseg), dim3(kWheelThreads), args, 0,
                            g.stream));
    hipLaunchKernelGGL(k_wheel_seg_scan, dim3((uint32_t)nchunks), dim3(256), 0, g.stream,
                       (const uint32_t*)wh.wave_count.as<uint32_t>(), in.nseg, wh.seg_excl.as<uint32_t>(),
                       wh.chunk_total.as<uint64_t>());
    hipLaunchKernelGGL(k_wheel_compact, dim3((uint32_t)in.nseg), dim3(kWheelThreads), 0, g.stream, a,
                       (const uint32_t*)wh.seg_excl.as<uint32_t>(), (const uint64_t*)wh.chunk_total.as<uint64_t>(),
                       out, cap, wh.win_info.as<uint64_t>());
    return PCUBE_OK;
}

// The passes of two different wheels (a batch's u and v windows) as three
// paired launches, so their latency chains overlap; two single passes when
// the wheels use different instantiations or a window is empty.
// With prepared (one window per wheel), the compaction also answers its subtile
// queries and clears prepared->clear; *fused tells whether it did (else the
// caller searches the bounds itself).
int wheel_enqueue_pair(Wheel& wa, const WheelInputs& ia, uint64_t* out_a, uint64_t cap_a, uint32_t slot_a, Wheel& wb,
                       const WheelInputs& ib, uint64_t* out_b, uint64_t cap_b, uint32_t slot_b,
                       WheelPrepared* prepared = nullptr, bool* fused = nullptr) {
    if (fused) *fused = false;
    if (&wa == &wb || ia.nseg == 0 || ib.nseg == 0 || wa.packed != wb.packed || wa.nouter != wb.nouter ||
        ia.nseg + ib.nseg > (1ull << 31)) {
        TRY(wheel_enqueue(wa, ia, out_a, cap_a, slot_a));
        return wheel_enqueue(wb, ib, out_b, cap_b, slot_b);
    }
    TRY(wa.win_info.ensure(((uint64_t)ia.n + 2) * 8));
    TRY(wb.win_info.ensure(((uint64_t)ib.n + 2) * 8));
    WheelArgs a, b;
    uint64_t chunks_a = 0, chunks_b = 0;
    TRY(wheel_setup(wa, ia, slot_a, a, chunks_a));
    TRY(wheel_setup(wb, ib, slot_b, b, chunks_b));
    const uint32_t blocks = (uint32_t)(ia.nseg + ib.nseg);
    void* args[] = {&a, &b};
    HIP_TRY(hipLaunchKernel(wheel_pair_kernel(wa.packed, wa.nouter), dim3(blocks), dim3(kWheelThreads), args, 0,
                            g.stream));
    // a batch's windows have a few hundred segments each: the compaction sums
    // the earlier segments' counts itself and the scan pass is not launched
    const uint32_t direct = ia.nseg <= kDirectPrefixSegments && ib.nseg <= kDirectPrefixSegments ? 1u : 0u;
    if (!direct) {
        const WheelScan sa{wa.wave_count.as<uint32_t>(), ia.nseg, wa.seg_excl.as<uint32_t>(),
                           wa.chunk_total.as<uint64_t>(), (uint32_t)chunks_a};
        const WheelScan sb{wb.wave_count.as<uint32_t>(), ib.nseg, wb.seg_excl.as<uint32_t>(),
                           wb.chunk_total.as<uint64_t>(), (uint32_t)chunks_b};
        hipLaunchKernelGGL(k_wheel_seg_scan_pair, dim3((uint32_t)(chunks_a + chunks_b)), dim3(256), 0, g.stream, sa,
                           sb);
    }
    const WheelOut oa{wa.seg_excl.as<uint32_t>(), wa.chunk_total.as<uint64_t>(), out_a, cap_a,
                      wa.win_info.as<uint64_t>(), direct};
    const WheelOut ob{wb.seg_excl.as<uint32_t>(), wb.chunk_total.as<uint64_t>(), out_b, cap_b,
                      wb.win_info.as<uint64_t>(), direct};
    if (prepared && ia.n == 1 && ib.n == 1) {
        prepared->oa = oa;
        prepared->ob = ob;
        hipLaunchKernelGGL(k_wheel_compact_prepared, dim3(blocks), dim3(kWheelThreads), 0, g.stream, a, b, *prepared);
        if (fused) *fused = true;
        return PCUBE_OK;
    }
    hipLaunchKernelGGL(k_wheel_compact_pair, dim3(blocks), dim3(kWheelThreads), 0, g.stream, a, oa, b, ob);
    return PCUBE_OK;
}

// n windows of a wheel into out, synchronously; offsets (n + 1) on the host.
// Returns OVERFLOW when the total exceeds cap (out then holds the first cap
// coordinates).
int wheel_run(Wheel& wh, uint32_t n, const uint64_t* lows, const uint64_t* highs, uint64_t* out, uint64_t cap,
              uint64_t* offsets) {
    std::vector<WheelWindow> plan;
    std::vector<uint64_t> seg_first;
    const uint64_t nseg = wheel_plan(wh, n, lows, highs, plan, seg_first);
    TRY(wheel_check(nseg));
    const uint64_t words = wheel_image_words(n);
    TRY(g_up_host.ensure(words * 8));
    TRY(g_down_host.ensure(((uint64_t)n + 2) * 8));
    TRY(g_up_dev.ensure(words * 8));
    wheel_image(plan, seg_first, g_up_host.as<uint64_t>());
    HIP_TRY(hipMemcpyAsync(g_up_dev.p, g_up_host.p, words * 8, hipMemcpyHostToDevice, g.stream));
    const WheelInputs in = wheel_inputs(g_up_dev.as<uint64_t>(), n, nseg);
    uint32_t slot_cap = wheel_slot_cap(wh);
    for (;;) {
        TRY(wheel_enqueue(wh, in, out, cap, slot_cap));
        HIP_TRY(hipMemcpyAsync(g_down_host.p, wh.win_info.p, ((uint64_t)n + 2) * 8, hipMemcpyDeviceToHost, g.stream));
        HIP_TRY(hipStreamSynchronize(g.stream));
        const uint64_t* info = g_down_host.as<uint64_t>();
        if (!info[n + 1] || slot_cap >= (uint32_t)kWheelWaveSpan) break;
        slot_cap = kWheelWaveSpan;   // a slot overflowed: once more with room for every candidate
        TRY(fill32(in.flag, 1, 0));
    }
    memcpy(offsets, g_down_host.p, ((uint64_t)n + 1) * 8);
    if (offsets[n] > cap) return fail(PCUBE_OVERFLOW, "focus buffer overflow: %llu coordinates, capacity %llu",
                                      (unsigned long long)offsets[n], (unsigned long long)cap);
    return PCUBE_OK;
}

// Windows of one focus into out; offsets (n + 1) on the host.  Returns
// OVERFLOW when the total exceeds cap; out's contents are then unspecified.
int focus_windows(Focus& f, uint32_t n, const uint64_t* lows, const uint64_t* highs, uint64_t* out, uint64_t cap,
                  uint64_t* offsets) {
    if (!f.wheel.ready) {
        if (n != 1) return fail(PCUBE_INVALID, "multi-window generation needs the focus wheel");
        uint64_t count = 0;
        TRY(window_run(f, lows[0], highs[0], out, cap, &count));
        offsets[0] = 0;
        offsets[1] = count;
        if (count > cap) return fail(PCUBE_OVERFLOW, "focus buffer overflow: %llu coordinates, capacity %llu",
                                     (unsigned long long)count, (unsigned long long)cap);
        return PCUBE_OK;
    }
    return wheel_run(f.wheel, n, lows, highs, out, cap, offsets);
}

// One lease-wide window: the wheel when the focus has one, else the exact
// two-pass directory walk.  offsets[1] receives the size.
int lease_window(Focus& f, uint64_t low, uint64_t high, uint64_t* out, uint64_t cap, uint64_t* offsets) {
    if (f.wheel.ready) {
        const int rc = focus_windows(f, 1, &low, &high, out, cap, offsets);
        if (rc == PCUBE_OVERFLOW) return fail(PCUBE_OVERFLOW, "lease focus window overflow (%llu > %llu)",
                                              (unsigned long long)offsets[1], (unsigned long long)cap);
        return rc;
    }
    uint64_t nb = 0;
    bool single = false;
    TRY(window_enqueue(f, low, high, out, cap, true, &nb, &single));
    uint32_t total = 0;
    HIP_TRY(hipMemcpyAsync(&total, f.work.offsets.as<uint32_t>() + nb, 4, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    offsets[0] = 0;
    offsets[1] = total;
    if (total > cap) return fail(PCUBE_OVERFLOW, "lease focus window overflow (%u > %llu)", total, (unsigned long long)cap);
    return PCUBE_OK;
}

// ---------------------------------------------------------------------------
// Matchers

// A batch enqueued by pcube_match_start and collected by pcube_match_finish.
struct Pending {
    bool active = false;
    uint64_t us = 0, nu = 0, vs = 0, nv = 0, v_first = 0, tile_width = 0, cap = 0, list_cap = 0;
    uint32_t ntiles = 0;
    int engine = 0;
    uint64_t launch = 0;   // the matcher's launch count right after this batch was enqueued
    // enqueued before the windows' sizes were known: nu and nv are estimates,
    // geo the device words that receive the true counts (staged engine only)
    const uint64_t* geo = nullptr;
    uint64_t nu_real = 0, nv_real = 0;   // with geo: the true sizes, once the host has seen them
};

struct Matcher {
    uint64_t A = 0, B = 0;
    U128 B_mu{0, 0};    // floor((2^128 - 1) / B)
    U128 lo{0, 0}, hi{0, 0};
    uint32_t nres = 0, nbits = 0, ncache = 0, ng = 0, allow_words = 0;
    uint32_t row_base[kMaxBitsetFilters] = {0, 0, 0, 0};
    uint32_t mask_rows = 0;
    // staged engine: bitset filters, their rows (sum of q), cached chain filters, LDS bytes
    uint32_t sbits = 0, slines = 0, sncache = 0;   // staged filters, lines of a chunk image (largest q)
    uint32_t s_pstr_off[kStagedFilters] = {}, s_nbeta[kStagedFilters] = {};
    std::vector<Modulus> mods;
    Buffer d_mods, d_allow, d_pstr;
    // status words: [0] survivor list length, [1] error bits, [2] unused, [3] most survivor
    // records of any wave of the staged match that overflowed its region since the last clear
    Buffer rows, masks, rho, counters, list, status, cand, items;
    Buffer geo;                            // true u and v counts of a speculatively enqueued batch
    uint64_t last_nu = 0, last_nv = 0;     // window sizes of the last prepared batch (the next estimate)
    uint64_t cand_cap = 256;                            // bitset-survivor records per wave of the staged match
    // start, bitsets built, match done, intersection kernel start and end, download done
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    Pending pending;       // the batch started by pcube_match_start
    Pinned down;           // its download: [status words, counters]
    bool kernel_span = false;   // the last batch's match kernel was timed by ev[1] and ev[2]
    uint64_t head_zero = 0;   // leading bytes of status known to be zero (cleared by a prepare's bounds kernel)
    uint64_t launches = 0;    // enqueues that used the batch buffers (status, counters, list, rows, masks)
    uint64_t* list_count() const { return status.as<uint64_t>(); }
    uint32_t* error() const { return (uint32_t*)(status.as<uint64_t>() + 1); }
    ~Matcher() {
        d_mods.release(); d_allow.release(); d_pstr.release(); rows.release(); masks.release(); rho.release();
        counters.release(); list.release(); status.release(); cand.release(); items.release();
        for (hipEvent_t& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (down.p) (void)hipHostFree(down.p);
    }
};
std::vector<std::unique_ptr<Matcher>> g_matchers;

Matcher* matcher_at(int id) {
    if (id < 0 || id >= (int)g_matchers.size() || !g_matchers[id]) return nullptr;
    return g_matchers[id].get();
}

bool cubic_residue(uint32_t q, uint32_t a) {
    for (uint32_t t = 1; t < q; ++t)
        if ((uint64_t)t * t % q * t % q == a) return true;
    return false;
}
bool allowed_residue(uint32_t q, uint32_t a) {
    if (q == 9) return a == 1 || a == 8;
    if (a == 0) return false;
    if (q % 3 != 1) return true;
    return cubic_residue(q, a);
}

// Staged engine for one batch (engine 0 when the matcher has staged filters):
// rows with per-subtile valid counts, staged bitsets, cached residues, then
// one block per (stage chunk, row piece).
int staged_enqueue(Matcher& mt, MatchArgs m, hipEvent_t ev_pre, hipEvent_t ev_match, hipEvent_t ev_k0,
                   hipEvent_t ev_k1, const uint64_t* geo = nullptr) {
    m.staged = 1;
    m.nbits = mt.sbits;
    m.ncache = mt.sncache;
    for (int k = 0; k < kStagedFilters; ++k) {
        m.nbeta[k] = mt.s_nbeta[k];
        m.pstr_off[k] = mt.s_pstr_off[k];
    }
    m.pstr = mt.d_pstr.as<uint32_t>();
    m.stage_lines = mt.slines;
    m.nstage = (m.words + kStageWords - 1) / kStageWords;
    TRY(mt.masks.ensure(std::max<uint64_t>(8, (uint64_t)m.nstage * mt.slines * kStageLineBytes)));
    TRY(mt.rho.ensure(std::max<uint64_t>(16, (uint64_t)m.nu * kCachedResidues * 2)));
    m.masks = mt.masks.as<uint64_t>();
    m.rho = mt.rho.as<uint16_t>();
    m.cand_max = mt.status.as<uint64_t>() + 3;
    m.item_cap = (uint32_t)mt.cand_cap;
    // kStagedBlocksPerCU persistent blocks per CU, no more than there can be work units
    const uint32_t npieces = (m.nv + kPieceRows - 1) / kPieceRows;
    const uint64_t items_max = (uint64_t)npieces * m.nstage;   // staged_fits bounds it
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((uint64_t)kStagedBlocksPerCU * g.compute_units,
                                                         std::max<uint64_t>(1, items_max * kUnitsPerItem));
    // the chunk table of the (piece, chunk) work list, each piece's reach and the
    // match blocks' runs
    TRY(mt.items.ensure(((uint64_t)m.nstage + 1) * sizeof(uint2) + (uint64_t)npieces * 8 +
                        ((uint64_t)blocks + 1) * sizeof(uint2)));
    uint2* chunks = mt.items.as<uint2>();
    uint32_t* reach = (uint32_t*)(chunks + m.nstage + 1);
    uint2* runs = (uint2*)(reach + 2 * (uint64_t)npieces);
    // each wave's survivor region: cand_cap survivor records
    TRY(mt.cand.ensure(std::max<uint64_t>(16, (uint64_t)blocks * kStagedWaves * mt.cand_cap * kRecordUnits * 16)));
    m.cand = mt.cand.as<uint32_t>();
    // one launch: the item plan, the rows, the bitsets and the residue cache
    PrefilterArgs pa;
    memset(&pa, 0, sizeof pa);
    pa.rows_blocks = (uint32_t)blocks_for(m.nv, 256);
    // one wave per (stage chunk, bitset filter)
    pa.mask_blocks = m.nu ? (uint32_t)blocks_for((uint64_t)mt.sbits * m.nstage * kWave, 256) : 0;
    pa.rho_blocks = m.nu && m.ncache ? (uint32_t)blocks_for((uint64_t)m.nu, 256) : 0;
    pa.npieces = npieces;
    pa.chunks = chunks;
    pa.reach = reach;
    pa.reach_lds = npieces <= kPlanLdsPieces;
    pa.grid = blocks;
    // with two blocks on every CU, the first-round blocks take the larger share
    pa.first64 = blocks == (uint32_t)(kStagedBlocksPerCU * g.compute_units) ? kFirstRoundWeight : 64u;
    pa.runs = runs;
    pa.geo = geo;
    uint32_t dir_stride, dir_entries;
    directory_shape(m.nu, dir_stride, dir_entries);
    const uint32_t dynamic_lds = (std::max<uint32_t>(1, dir_entries) + (pa.reach_lds ? npieces : 0)) * 8;
    hipLaunchKernelGGL(k_prefilter_staged, dim3(1 + pa.rows_blocks + pa.mask_blocks + pa.rho_blocks), dim3(256),
                       dynamic_lds, g.stream, m, pa);
    if (ev_pre) HIP_TRY(hipEventRecord(ev_pre, g.stream));
    // with both of these events recorded, the match kernel alone lies between
    // them and stands for its own timing events
    mt.kernel_span = ev_pre && ev_match;
    if (m.nu) {
        // persistent blocks over the runs of work units the plan assigned, then the residual chain
        const uint2* chunks_arg = chunks;
        const uint2* runs_arg = runs;
        void* args[] = {&m, &chunks_arg, &runs_arg};
        if (ev_k0 && !mt.kernel_span) HIP_TRY(hipEventRecord(ev_k0, g.stream));
        HIP_TRY(hipLaunchKernel(staged_kernel(mt.sbits, mt.nres), dim3(blocks), dim3(kStagedThreads), args, 0,
                                g.stream));
        if (ev_k1 && !mt.kernel_span) HIP_TRY(hipEventRecord(ev_k1, g.stream));
    }
    if (ev_match) HIP_TRY(hipEventRecord(ev_match, g.stream));
    return PCUBE_OK;
}

// Whether one batch of this geometry runs on the staged engine.
bool staged_fits(const Matcher& mt, uint64_t nu, uint64_t nv) {
    if (!mt.sbits || kStagedLdsBytes > g.lds_per_block) return false;
    const uint64_t words = (nu + 63) / 64, nstage = (words + kStageWords - 1) / kStageWords;
    const uint64_t npieces = (nv + kPieceRows - 1) / kPieceRows;
    return npieces * nstage <= (1ull << 22) && nu < (1ull << 31);
}

// Internal engine code: the residue-bitset engine without LDS staging.
constexpr int kUnstagedBitsets = 2;
// Survivor records (rows with bitset survivors in one chunk) one wave of the
// staged match may hold before the unstaged engine is used instead (a wave's
// region is item_cap records of 48 bytes; 4096 waves at most).
constexpr uint64_t kMaxWaveCandidates = 1ull << 11;

// Enqueue the kernels of one batch; counters must be zeroed by the caller.
int match_enqueue(Matcher& mt, const uint64_t* us, uint64_t nu, const uint64_t* vs, uint64_t nv, uint64_t v_first,
                  uint64_t tile_width, uint32_t ntiles, uint64_t cap, int engine, uint64_t* counters,
                  uint64_t tile_base, uint64_t list_cap, hipEvent_t ev_pre, hipEvent_t ev_match,
                  hipEvent_t ev_k0 = nullptr, hipEvent_t ev_k1 = nullptr, const uint64_t* geo = nullptr) {
    if (nu >= (1ull << 32) - 64 || nv >= (1ull << 32)) return fail(PCUBE_INVALID, "batch too large");
    if (geo && !(engine == 0 && staged_fits(mt, nu, nv)))
        return fail(PCUBE_INVALID, "a batch of unknown size runs on the staged engine only");
    const uint32_t words = (uint32_t)((nu + 63) / 64);
    TRY(mt.rows.ensure(std::max<uint64_t>(1, nv) * sizeof(Row)));
    if (engine == 0) {
        TRY(mt.masks.ensure(std::max<uint64_t>(8, (uint64_t)mt.mask_rows * std::max(1u, words) * 8)));
        TRY(mt.rho.ensure(std::max<uint64_t>(16, nu * kCachedResidues * 2)));
    }
    MatchArgs m;
    memset(&m, 0, sizeof m);
    m.us = us; m.vs = vs;
    m.rows = mt.rows.as<Row>();
    m.masks = mt.masks.as<uint64_t>();
    m.rho = mt.rho.as<uint16_t>();
    m.mods = mt.d_mods.as<Modulus>();
    m.allow = mt.d_allow.as<uint32_t>();
    m.counters = counters;
    m.list = mt.list.as<uint64_t>();
    m.list_count = mt.list_count();
    m.error = mt.error();
    m.A = mt.A; m.B = mt.B; m.B_mu = mt.B_mu; m.v_first = v_first; m.tile_width = tile_width; m.cap = cap;
    m.list_cap = list_cap; m.tile_base = tile_base;
    m.lo = mt.lo; m.hi = mt.hi;
    m.nu = (uint32_t)nu; m.nv = (uint32_t)nv; m.words = words;
    m.nres = mt.nres; m.nbits = engine == 0 ? mt.nbits : 0; m.ncache = engine == 0 ? mt.ncache : 0;
    m.ng = mt.ng; m.ntiles = ntiles; m.allow_words = mt.allow_words;
    for (int k = 0; k < kMaxBitsetFilters; ++k) m.row_base[k] = mt.row_base[k];
    if (nv == 0) return PCUBE_OK;
    if (engine == 0 && staged_fits(mt, nu, nv)) return staged_enqueue(mt, m, ev_pre, ev_match, ev_k0, ev_k1, geo);
    mt.kernel_span = false;
    if (engine == kUnstagedBitsets) engine = 0;
    // k_rows always computes v mod q for the bitset filters of the matcher
    MatchArgs mr = m;
    mr.nbits = mt.nbits;
    hipLaunchKernelGGL(k_rows, dim3(blocks_for(nv, 256)), dim3(256), 0, g.stream, mr);
    if (engine == 0 && nu) {
        uint32_t chunk_base[kMaxBitsetFilters + 1] = {0, 0, 0, 0, 0};
        for (uint32_t k = 0; k < kMaxBitsetFilters; ++k)
            chunk_base[k + 1] = chunk_base[k] + (k < mt.nbits ? (mt.mods[k].q + kWave - 1) / kWave : 0);
        const uint32_t chunks = chunk_base[mt.nbits];
        const uint64_t waves = (uint64_t)chunks * words;
        hipLaunchKernelGGL(k_masks, dim3(blocks_for(waves * kWave, 256)), dim3(256), 0, g.stream, m, chunks,
                           chunk_base[1], mt.nbits > 2 ? chunk_base[2] : 0xffffffffu,
                           mt.nbits > 3 ? chunk_base[3] : 0xffffffffu);
        if (mt.ncache)
            hipLaunchKernelGGL(k_rho, dim3(blocks_for(nu, 256)), dim3(256), 0, g.stream, m);
    }
    if (ev_pre) HIP_TRY(hipEventRecord(ev_pre, g.stream));
    if (ev_k0) HIP_TRY(hipEventRecord(ev_k0, g.stream));
    const uint64_t rows_per_wave = std::max<uint64_t>(1, (nv + 8191) / 8192);
    m.rows_per_wave = (uint32_t)rows_per_wave;
    const uint64_t waves = (nv + rows_per_wave - 1) / rows_per_wave;
    uint64_t blocks = (waves + kMatchWaves - 1) / kMatchWaves;
    if (engine == 0) {
        blocks = (blocks + 7) / 8 * 8;
        hipLaunchKernelGGL(k_match, dim3((uint32_t)blocks), dim3(kMatchThreads), 0, g.stream, m);
    } else {
        hipLaunchKernelGGL(k_match_direct, dim3((uint32_t)blocks), dim3(kMatchThreads), 0, g.stream, m);
    }
    if (ev_k1) HIP_TRY(hipEventRecord(ev_k1, g.stream));
    if (ev_match) HIP_TRY(hipEventRecord(ev_match, g.stream));
    return PCUBE_OK;
}

// Enqueue the pending batch of a matcher: clear its status words and counters,
// run the engine, download both to the matcher's pinned buffer and record
// ev[5] after the download.  The counters follow the four status words in the
// status buffer, so one fill clears and one copy downloads [status, counters].
int match_launch(Matcher& mt) {
    const Pending& pd = mt.pending;
    const uint64_t nc = (uint64_t)pd.ntiles * (mt.ng + 3);
    hipEvent_t e0 = mt.ev[0], e1 = mt.ev[1], e2 = mt.ev[2], k0 = mt.ev[3], k1 = mt.ev[4];
    uint64_t* head = mt.status.as<uint64_t>();
    ++mt.launches;
    // the last batch's finish cleared the buffer already, unless it has grown or been used since
    if (mt.head_zero < (nc + 4) * 8)
        hipLaunchKernelGGL(k_fill2_u64, dim3(std::min<uint64_t>(blocks_for(nc + 4, 256), 4096)), dim3(256), 0,
                           g.stream, head, nc + 4, head, (uint64_t)0);
    mt.head_zero = 0;
    HIP_TRY(hipEventRecord(e0, g.stream));
    TRY(match_enqueue(mt, (const uint64_t*)pd.us, pd.nu, (const uint64_t*)pd.vs, pd.nv, pd.v_first, pd.tile_width,
                      pd.ntiles, pd.cap, pd.engine, head + 4, 0, pd.list_cap, e1, e2, k0, k1, pd.geo));
    if (pd.nv == 0)
        for (hipEvent_t e : {e1, k0, k1, e2}) HIP_TRY(hipEventRecord(e, g.stream));
    HIP_TRY(hipMemcpyAsync(mt.down.p, head, (nc + 4) * 8, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipEventRecord(mt.ev[5], g.stream));
    return PCUBE_OK;
}

// Copy the tagged value list to the host grouped by tile, in tile order.
int collect_values(Matcher& mt, uint64_t ntiles_total, const uint64_t* host_counters, uint64_t cap,
                   uint64_t list_cap, uint64_t* values, uint64_t values_cap, int* status, uint64_t listed) {
    uint64_t expected = 0;
    std::vector<uint64_t> offset(ntiles_total + 1, 0);
    for (uint64_t t = 0; t < ntiles_total; ++t) {
        const uint64_t emitted = host_counters[t * (mt.ng + 3) + mt.ng + 1];
        offset[t + 1] = offset[t] + std::min(emitted, cap);
    }
    expected = offset[ntiles_total];
    if (expected > list_cap) {
        // more survivors than the shared list holds (no subtile above cap): capacity, not disagreement
        *status = PCUBE_OVERFLOW;
        fail(PCUBE_OVERFLOW, "survivor list capacity %llu below the %llu values the counters report",
             (unsigned long long)list_cap, (unsigned long long)expected);
        return PCUBE_OK;
    }
    if (listed != expected) {
        // every emitted value is counted and listed once: a difference means lost or extra values
        *status = PCUBE_MISMATCH;
        fail(PCUBE_MISMATCH, "survivor list holds %llu values, counters say %llu",
             (unsigned long long)listed, (unsigned long long)expected);
        return PCUBE_OK;
    }
    if (expected > values_cap) return fail(PCUBE_OVERFLOW, "host value buffer too small");
    if (!listed) return PCUBE_OK;
    TRY(g_down_host.ensure(listed * 24));
    const uint64_t* triples = g_down_host.as<uint64_t>();
    HIP_TRY(hipMemcpyAsync(g_down_host.p, mt.list.p, listed * 24, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    std::vector<uint64_t> cursor(offset.begin(), offset.end() - 1);
    for (uint64_t e = 0; e < listed; ++e) {
        const uint64_t t = triples[3 * e];
        if (t >= ntiles_total || cursor[t] >= offset[t + 1]) return fail(PCUBE_INVALID, "corrupt survivor list");
        values[2 * cursor[t]] = triples[3 * e + 1];
        values[2 * cursor[t] + 1] = triples[3 * e + 2];
        ++cursor[t];
    }
    return PCUBE_OK;
}

// Translate counters and error bits into the status code.
int batch_status(const Matcher& mt, const uint64_t* c, uint64_t ntiles, uint64_t cap, uint32_t err) {
    if (err & 1u) return fail(PCUBE_INVALID, "a v coordinate lies outside the subtiles");
    for (uint64_t t = 0; t < ntiles; ++t)
        if (c[t * (mt.ng + 3) + mt.ng + 2]) return fail(PCUBE_UNDERFLOW, "unsigned underflow flagged");
    if (err & 2u) return fail(PCUBE_RANGE, "an emitted value lies outside (lo, hi]");
    for (uint64_t t = 0; t < ntiles; ++t)
        if (c[t * (mt.ng + 3) + mt.ng + 1] > cap) return fail(PCUBE_OVERFLOW, "output overflow in subtile %llu", (unsigned long long)t);
    return PCUBE_OK;
}

// out[q] = searchsorted(data[begin[q] .. begin[q] + len[q]), key[q], 'left')
__global__ void k_searchsorted_seg(const uint64_t* __restrict__ data, const uint64_t* __restrict__ begin,
                                   const uint64_t* __restrict__ len, const uint64_t* __restrict__ key, uint64_t nq,
                                   uint64_t* __restrict__ out) {
    uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= nq) return;
    const uint64_t* a = data + begin[t];
    uint64_t l = 0, h = len[t];
    const uint64_t k = key[t];
    while (l < h) {
        const uint64_t m = l + ((h - l) >> 1);
        if (a[m] < k) l = m + 1; else h = m;
    }
    out[t] = l;
}

// Subtile bounds of one prepared batch: lower bounds of the 2*ntiles u queries
// in us and of the 2*ntiles v queries in vs (sizes from each wheel's window
// info, clamped to the capacities), then both sizes and overflow flags.  One
// wave per query: each round its 64 lanes probe evenly spaced entries of the
// remaining range [l, h) and a ballot cuts the range 64-fold, so a million
// entries take four dependent loads.  Invariant: the first index with
// a[i] >= k lies in [l, h] (h = the size when every entry is smaller).
// The same search for a lease pass: nq_u queries into the u window, then nq_v
// into the v window, read from device memory; results as k_prepare_bounds
// (positions, then both sizes and overflow flags at nq_u + nq_v).
__global__ __launch_bounds__(256) void k_lease_bounds(const uint64_t* __restrict__ us,
                                                      const uint64_t* __restrict__ uinfo, uint64_t us_cap,
                                                      const uint64_t* __restrict__ vs,
                                                      const uint64_t* __restrict__ vinfo, uint64_t vs_cap,
                                                      const uint64_t* __restrict__ queries, uint64_t nq_u,
                                                      uint64_t nq_v, uint64_t* __restrict__ out) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t q = blockIdx.x * (uint64_t)(blockDim.x / kWave) + threadIdx.x / kWave;
    const uint64_t nq = nq_u + nq_v;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        out[nq] = uinfo[1];
        out[nq + 1] = vinfo[1];
        out[nq + 2] = uinfo[2];
        out[nq + 3] = vinfo[2];
    }
    if (q >= nq) return;
    const bool on_u = q < nq_u;
    const uint64_t* a = on_u ? us : vs;
    uint64_t l = 0, h = on_u ? min(uinfo[1], us_cap) : min(vinfo[1], vs_cap);
    const uint64_t k = queries[q];
    while (l < h) {
        const uint64_t step = (h - l + kWave - 1) / kWave;
        const uint64_t at = l + lane * step;
        const uint32_t below = (uint32_t)__popcll(__ballot(at < h && a[at] < k));
        if (below == 0) break;   // a[l] >= k
        const uint64_t top = l + below * step;
        l += (below - 1) * step + 1;
        h = min(h, top);
    }
    if (lane == 0) out[q] = l;
}

// The queries of a prepared batch, passed with the kernel arguments: (u_min, u_end)
// per subtile, then (v_first, v_end) per subtile, at most 64 subtiles.
constexpr uint32_t kMaxPrepareTiles = 64;
struct PrepareQueries {
    uint64_t q[4 * kMaxPrepareTiles];
};

__global__ __launch_bounds__(256) void k_prepare_bounds(const uint64_t* __restrict__ us,
                                                        const uint64_t* __restrict__ uinfo, uint64_t us_cap,
                                                        const uint64_t* __restrict__ vs,
                                                        const uint64_t* __restrict__ vinfo, uint64_t vs_cap,
                                                        const PrepareQueries queries, uint32_t ntiles,
                                                        uint64_t* __restrict__ out, uint64_t* __restrict__ clear,
                                                        uint64_t nclear) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t q = blockIdx.x * (uint64_t)(blockDim.x / kWave) + threadIdx.x / kWave;
    const uint64_t nq = 4 * (uint64_t)ntiles;
    // the matcher's status words and counters, cleared for the match this
    // batch will start (stream-ordered after the last batch's download)
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nclear; i += (uint64_t)gridDim.x * blockDim.x)
        clear[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        out[nq] = uinfo[1];
        out[nq + 1] = vinfo[1];
        out[nq + 2] = uinfo[2];
        out[nq + 3] = vinfo[2];
    }
    if (q >= nq) return;
    const bool on_u = q < 2 * (uint64_t)ntiles;
    const uint64_t* a = on_u ? us : vs;
    uint64_t l = 0, h = on_u ? min(uinfo[1], us_cap) : min(vinfo[1], vs_cap);
    const uint64_t k = queries.q[q];
    while (l < h) {
        const uint64_t step = (h - l + kWave - 1) / kWave;
        const uint64_t at = l + lane * step;
        const uint32_t below = (uint32_t)__popcll(__ballot(at < h && a[at] < k));
        if (below == 0) break;   // a[l] >= k
        const uint64_t top = l + below * step;
        l += (below - 1) * step + 1;
        h = min(h, top);
    }
    if (lane == 0) out[q] = l;
}

// Batched engine: configurations, focus wheels and the grow-only buffers one
// batched launch needs, plus the captured launch sequence.
struct BatchEngine {
    std::vector<int> matchers;
    int focus_u = -1, focus_v = -1;
    uint64_t cap = 8192, list_cap = 1u << 22;
    uint32_t blocks_match = 1024;
    bool use_graph = true;
    BatchArgs args;
    Buffer win, totals, rows, counters, list, list_count, error, cursor, masks, rho, lus, lvs, cfg;
    Buffer q_begin, q_len, q_key, q_out;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<const void*> key;
    uint64_t launches = 0, graph_launches = 0;
    ~BatchEngine() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (Buffer* b : {&win, &totals, &rows, &counters, &list, &list_count, &error, &cursor, &masks, &rho, &lus,
                          &lvs, &cfg, &q_begin, &q_len, &q_key, &q_out})
            b->release();
    }
};
std::vector<std::unique_ptr<BatchEngine>> g_batch;

constexpr uint32_t kFillBlocks = 1024, kRowBlocks = 2048, kMaskBlocks = 4096, kRhoBlocks = 2048;

// The fixed-grid launch sequence of one batched pass.
int batch_sequence(BatchEngine& e) {
    hipLaunchKernelGGL(k_batch_fill, dim3(kFillBlocks), dim3(256), 0, g.stream, e.args);
    hipLaunchKernelGGL(k_batch_rows, dim3(kRowBlocks), dim3(256), 0, g.stream, e.args);
    hipLaunchKernelGGL(k_batch_masks, dim3(kMaskBlocks), dim3(256), 0, g.stream, e.args);
    hipLaunchKernelGGL(k_batch_rho, dim3(kRhoBlocks), dim3(256), 0, g.stream, e.args);
    hipLaunchKernelGGL(k_batch_match, dim3(e.blocks_match), dim3(kMatchThreads), 0, g.stream, e.args);
    return PCUBE_OK;
}

// Launch the sequence, through the captured graph when enabled; the graph is
// (re)captured whenever a buffer it refers to has moved.
int batch_launch(BatchEngine& e) {
    ++e.launches;
    if (!e.use_graph) return batch_sequence(e);
    std::vector<const void*> key = {e.win.p, e.totals.p, e.rows.p, e.counters.p, e.list.p, e.list_count.p,
                                    e.error.p, e.cursor.p};
    if (!e.exec || key != e.key) {
        if (e.exec) (void)hipGraphExecDestroy(e.exec);
        if (e.graph) (void)hipGraphDestroy(e.graph);
        e.exec = nullptr;
        e.graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
        int rc = batch_sequence(e);
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamEndCapture(g.stream, &graph));
        if (rc) return rc;
        e.graph = graph;
        HIP_TRY(hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0));
        e.key = key;
    }
    ++e.graph_launches;
    HIP_TRY(hipGraphLaunch(e.exec, g.stream));
    return PCUBE_OK;
}

struct Kernel {
    const char* name;
    const void* fn;
};
const Kernel kKernels[] = {
    {"partial", (const void*)k_partial},
    {"window_count", (const void*)k_window<kCount>},
    {"window_emit", (const void*)k_window<kEmit>},
    {"window_slots", (const void*)k_window<kSlots>},
    {"slots_compact", (const void*)k_slots_compact},
    {"wheel_slots", (const void*)k_wheel_slots<4>},
    {"wheel_compact", (const void*)k_wheel_compact},
    {"bucket_sort", (const void*)k_bucket_sort},
    {"scan_local", (const void*)k_scan_local},
    {"rows", (const void*)k_rows},
    {"mask_build", (const void*)k_masks},
    {"prefilter_staged", (const void*)k_prefilter_staged},
    {"match_staged", (const void*)k_match_staged<kStagedFilters, false>},
    {"residues", (const void*)k_rho},
    {"match", (const void*)k_match},
    {"match_direct", (const void*)k_match_direct},
    {"arithmetic", (const void*)k_arithmetic},
    {"copy_probe", (const void*)k_copy_probe},
    {"lds_probe", (const void*)k_lds_probe},
    {"valu_probe", (const void*)k_valu_probe},
    {"cube_count", (const void*)k_cube_count},
    {"batch_rows", (const void*)k_batch_rows},
    {"batch_masks", (const void*)k_batch_masks},
    {"batch_rho", (const void*)k_batch_rho},
    {"batch_match", (const void*)k_batch_match},
};

double now_seconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Wait for an event by polling it: a blocking wait sleeps and costs a wake-up of
// tens of microseconds, twice per batch on the plain path.  After kPollSeconds the
// wait blocks as hipEventSynchronize does, so long work does not hold a core.
constexpr double kPollSeconds = 0.02;
int event_wait(hipEvent_t e) {
    const double t0 = now_seconds();
    for (uint32_t k = 1;; ++k) {
        const hipError_t rc = hipEventQuery(e);
        if (rc == hipSuccess) return PCUBE_OK;
        if (rc != hipErrorNotReady) return fail(PCUBE_DEVICE, "hipEventQuery: %s", hipGetErrorString(rc));
        if (k % 64 == 0 && now_seconds() - t0 > kPollSeconds) break;
        __builtin_ia32_pause();
    }
    HIP_TRY(hipEventSynchronize(e));
    return PCUBE_OK;
}

// event_wait on everything queued on the library's stream so far.
int stream_wait() {
    static hipEvent_t done = nullptr;
    if (!done) HIP_TRY(hipEventCreateWithFlags(&done, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(done, g.stream));
    return event_wait(done);
}

}  // namespace

// ===========================================================================
// C ABI

extern "C" {

const char* pcube_last_error(void) { return g_err.c_str(); }

int pcube_device_count(int* count) {
    Entry lock;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return PCUBE_OK;
}

int pcube_init(int device, uint64_t memory_cap, double fraction, uint64_t* cap, uint64_t* free_bytes,
               uint64_t* total_bytes) {
    Entry lock;
    if (g.device >= 0 && g.device != device) return fail(PCUBE_INVALID, "already initialised on device %d", g.device);
    if (g.device < 0) {
        HIP_TRY(hipSetDevice(device));
        // host waits spin instead of sleeping: the library synchronises twice per
        // batch and each wake-up from a sleep costs tens of microseconds.  The
        // flag can only be set before the device is in use; when another
        // component initialised it first, its setting stays.
        if (hipSetDeviceFlags(hipDeviceScheduleSpin) != hipSuccess) (void)hipGetLastError();
        HIP_TRY(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
        hipDeviceProp_t prop;
        HIP_TRY(hipGetDeviceProperties(&prop, device));
        g.compute_units = prop.multiProcessorCount;
        snprintf(g.name, sizeof g.name, "%s (%s)", prop.name, prop.gcnArchName);
        int lds = 0;
        HIP_TRY(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
        g.lds_per_block = (uint32_t)std::max(lds, 0);
        // the allocation-free entry point's scratch: this device's copy of the
        // array, looked up once with the device current
        if (device < 0 || device >= Context::kMaxDevices)
            return fail(PCUBE_INVALID, "device index %d beyond the library's %d", device, Context::kMaxDevices);
        void* at = nullptr;
        HIP_TRY(hipGetSymbolAddress(&at, HIP_SYMBOL(scratch)));
        if (!at) return fail(PCUBE_DEVICE, "no device address for the static scratch array");
        g.scratch_at[device] = at;
        g.device = device;
    }
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot));
    const double frac_cap = (double)tot * fraction;
    const int64_t headroom = (int64_t)fr - (int64_t)(256ull << 20) + (int64_t)g.reserved;
    int64_t c = (int64_t)std::min<double>((double)memory_cap, frac_cap);
    c = std::min<int64_t>(c, headroom);
    if (c <= 0) return fail(PCUBE_NOMEM, "no permitted device memory");
    g.cap = (uint64_t)c;
    if (cap) *cap = g.cap;
    if (free_bytes) *free_bytes = fr;
    if (total_bytes) *total_bytes = tot;
    return PCUBE_OK;
}

int pcube_memory(uint64_t* reserved, uint64_t* peak, uint64_t* cap) {
    Entry lock;
    if (reserved) *reserved = g.reserved;
    if (peak) *peak = g.peak;
    if (cap) *cap = g.cap;
    return PCUBE_OK;
}

const uint64_t* pcube_memory_peak(void) { return &g.peak; }

int pcube_device_info(char* name, size_t name_len, int* compute_units) {
    Entry lock;
    if (g.device < 0) return fail(PCUBE_INVALID, "pcube_init has not been called");
    if (name && name_len) snprintf(name, name_len, "%s", g.name);
    if (compute_units) *compute_units = g.compute_units;
    return PCUBE_OK;
}

int pcube_kernel_info(int index, char* name, size_t name_len, int* registers, int* local_bytes, int* shared_bytes) {
    Entry lock;
    const int n = (int)(sizeof kKernels / sizeof kKernels[0]);
    if (index < 0 || index >= n) return fail(PCUBE_INVALID, "kernel index out of range");
    if (name && name_len) snprintf(name, name_len, "%s", kKernels[index].name);
    hipFuncAttributes attr;
    memset(&attr, 0, sizeof attr);
    HIP_TRY(hipFuncGetAttributes(&attr, kKernels[index].fn));
    if (registers) *registers = attr.numRegs;
    if (local_bytes) *local_bytes = (int)attr.localSizeBytes;
    if (shared_bytes) *shared_bytes = (int)attr.sharedSizeBytes;
    return PCUBE_OK;
}

int pcube_alloc(uint64_t bytes, uint64_t* address) {
    Entry lock;
    void* p = nullptr;
    TRY(dev_alloc(bytes, &p));
    *address = (uint64_t)p;
    return PCUBE_OK;
}

int pcube_free(uint64_t address) {
    Entry lock;
    // stream-ordered reuse: the next user of the block queues behind this one
    return dev_free((void*)address);
}

int pcube_upload(uint64_t dst, const void* src, uint64_t bytes) {
    Entry lock;
    if (!bytes) return PCUBE_OK;
    HIP_TRY(hipMemcpyAsync((void*)dst, src, bytes, hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    return PCUBE_OK;
}

int pcube_download(void* dst, uint64_t src, uint64_t bytes) {
    Entry lock;
    if (!bytes) return PCUBE_OK;
    HIP_TRY(hipMemcpyAsync(dst, (const void*)src, bytes, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    return PCUBE_OK;
}

int pcube_pinned_alloc(uint64_t bytes, void** host) {
    Entry lock;
    if (!host || !bytes) return fail(PCUBE_INVALID, "pinned allocation needs a size and an output pointer");
    if (g.device < 0) return fail(PCUBE_INVALID, "pcube_init has not been called");
    *host = nullptr;
    if (hipHostMalloc(host, bytes, hipHostMallocDefault) != hipSuccess || !*host) {
        *host = nullptr;
        return fail(PCUBE_NOMEM, "pinned host allocation failed");
    }
    return PCUBE_OK;
}

int pcube_pinned_free(void* host) {
    Entry lock;
    if (!host) return PCUBE_OK;
    // the stream may still copy from or into the block
    HIP_TRY(hipStreamSynchronize(g.stream));
    HIP_TRY(hipHostFree(host));
    return PCUBE_OK;
}

// Host window producer: every y = c + d - w*M + base with c in C, d in D (both
// ascending, values in [0, M)), w in {0, 1} and lo <= c + d - w*M < hi, written
// in ascending order.  The walk goes up C: for each c the admissible d form
// the range [lo + w*M - c, hi + w*M - c), whose ends only move down as c
// grows, so two cursors into D walk it once per w.
int pcube_host_focus_window(const uint64_t* C, uint64_t nc, const uint64_t* D, uint64_t nd, uint64_t M, uint64_t lo,
                            uint64_t hi, uint64_t base, uint64_t* out, uint64_t cap, uint64_t* count) {
    if (!count || (nc && !C) || (nd && !D) || (cap && !out)) return fail(PCUBE_INVALID, "null host window buffer");
    *count = 0;
    if (lo > hi || hi > M || M >= (1ull << 63)) return fail(PCUBE_INVALID, "host window outside [0, M]");
    uint64_t n = 0;
    for (int w = 0; w < 2; ++w) {
        const uint64_t a = lo + (w ? M : 0), b = hi + (w ? M : 0);
        uint64_t dl = nd, dh = nd;   // first index of D at or above a - c, and b - c
        for (uint64_t i = 0; i < nc; ++i) {
            const uint64_t c = C[i];
            if (c >= b) break;
            const uint64_t la = a > c ? a - c : 0, lb = b - c;
            while (dh > 0 && D[dh - 1] >= lb) --dh;
            while (dl > 0 && D[dl - 1] >= la) --dl;
            for (uint64_t k = dl; k < dh; ++k) {
                if (n >= cap) {
                    *count = n + 1;
                    return fail(PCUBE_OVERFLOW, "host window output overflow");
                }
                out[n++] = c + D[k] - (w ? M : 0) + base;
            }
        }
    }
    std::sort(out, out + n);
    *count = n;
    return PCUBE_OK;
}

int pcube_focus_create(uint64_t modulus, const uint64_t* terms, const uint32_t* sizes, uint32_t nrings_c,
                       uint32_t nrings_d, int* focus, uint64_t* nc, uint64_t* nd) {
    Entry lock;
    if (g.device < 0) return fail(PCUBE_INVALID, "pcube_init has not been called");
    if (modulus == 0 || modulus >= (1ull << 63)) return fail(PCUBE_INVALID, "focus modulus must lie in [1, 2^63)");
    auto f = std::make_unique<Focus>();
    f->M = modulus;
    uint64_t cterms = 0;
    for (uint32_t k = 0; k < nrings_c; ++k) cterms += sizes[k];
    TRY(build_half(modulus, terms, sizes, nrings_c, f->C, &f->nc, f->work));
    TRY(build_half(modulus, terms + cterms, sizes + nrings_c, nrings_d, f->D, &f->nd, f->work));
    // bucket directory over C: about one entry per bucket
    uint32_t log_nc = 0;
    while ((2ull << log_nc) <= f->nc) ++log_nc;
    uint32_t bits = 0;
    while (bits < 64 && ((modulus - 1) >> bits) != 0) ++bits;
    f->dshift = bits > log_nc ? bits - log_nc : 0;
    f->nbk = ((modulus - 1) >> f->dshift) + 1;
    TRY(f->dir.ensure((f->nbk + 1) * 4));
    hipLaunchKernelGGL(k_directory, dim3(blocks_for(f->nbk + 1, 256)), dim3(256), 0, g.stream, f->C.as<uint64_t>(),
                       f->nc, f->dshift, f->nbk, f->dir.as<uint32_t>());
    HIP_TRY(hipStreamSynchronize(g.stream));
    *nc = f->nc;
    *nd = f->nd;
    g_focus.push_back(std::move(f));
    *focus = (int)g_focus.size() - 1;
    return PCUBE_OK;
}

int pcube_focus_arrays(int focus, uint64_t* c_address, uint64_t* d_address) {
    Entry lock;
    Focus* f = focus_at(focus);
    if (!f) return fail(PCUBE_INVALID, "unknown focus");
    *c_address = (uint64_t)f->C.p;
    *d_address = (uint64_t)f->D.p;
    return PCUBE_OK;
}

int pcube_focus_destroy(int focus) {
    Entry lock;
    if (!focus_at(focus)) return fail(PCUBE_INVALID, "unknown focus");
    HIP_TRY(hipStreamSynchronize(g.stream));
    g_focus[focus].reset();
    return PCUBE_OK;
}

int pcube_focus_window(int focus, uint64_t low, uint64_t high, uint64_t out, uint64_t cap, uint64_t* count) {
    Entry lock;
    *count = 0;
    Focus* f = focus_at(focus);
    if (!f || low > high) return fail(PCUBE_INVALID, "invalid focus window");
    if (low == high) return PCUBE_OK;
    uint64_t offsets[2] = {0, 0};
    const int rc = focus_windows(*f, 1, &low, &high, (uint64_t*)out, cap, offsets);
    *count = offsets[1];
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(g.stream));
    return PCUBE_OK;
}

int pcube_focus_wheel(int focus, const uint32_t* moduli, const uint32_t* sizes, const uint32_t* residues,
                      uint32_t nrings, uint64_t* inner_modulus, uint32_t* inner_residues) {
    Entry lock;
    Focus* f = focus_at(focus);
    if (!f) return fail(PCUBE_INVALID, "unknown focus");
    uint64_t product = 1;
    for (uint32_t k = 0; k < nrings; ++k) product *= moduli[k];
    if (product != f->M) return fail(PCUBE_INVALID, "ring moduli do not multiply to the focus modulus");
    Wheel& wh = f->wheel;
    wh.release();
    wh.ready = false;
    wh.rate = 1.0;
    uint64_t M1 = 1, n1 = 1;
    uint32_t k = 0;
    while (k < nrings && M1 * moduli[k] < kInnerModulusLimit && n1 * sizes[k] <= kMaxInnerResidues) {
        M1 *= moduli[k];
        n1 *= sizes[k];
        ++k;
    }
    if (nrings - k > (uint32_t)kMaxOuterRings) return fail(PCUBE_INVALID, "too many outer rings for the wheel");
    // entries carry r mod q_j of the outer rings as bytes above r when every
    // outer q < 256 and there are at most kPackedOuterRings of them
    wh.packed = nrings - k <= (uint32_t)kPackedOuterRings ? 1u : 0u;
    for (uint32_t j = k; j < nrings; ++j)
        if (moduli[j] >= 256) wh.packed = 0;
    const uint32_t npack = wh.packed ? nrings - k : 0;
    // inner residues in ascending order, one ring at a time: the residues mod
    // M*q are t*M + r for t in [0, q) and r in the sorted residues mod M, kept
    // when (t*M + r) mod q is allowed; visiting t, then r, yields them sorted.
    // The last ring's pass also forms the packed outer residues, each as
    // (r mod q_j + t*M mod q_j) mod q_j.
    std::vector<uint64_t> cur(1, 0), next;
    std::vector<uint32_t> bytes(1, 0), next_bytes, rq;
    std::vector<uint8_t> ok, ro[kPackedOuterRings];
    uint64_t M = 1, offset = 0;
    for (uint32_t i = 0; i < k; ++i) {
        const uint32_t q = moduli[i];
        ok.assign(q, 0);
        for (uint32_t e = 0; e < sizes[i]; ++e) {
            if (residues[offset + e] >= q) return fail(PCUBE_INVALID, "ring residue not below its modulus");
            ok[residues[offset + e]] = 1;
        }
        rq.resize(cur.size());
        for (size_t x = 0; x < cur.size(); ++x) rq[x] = (uint32_t)(cur[x] % q);
        const bool last = i + 1 == k;
        uint32_t oq[kPackedOuterRings] = {}, osh[kPackedOuterRings] = {}, ostep[kPackedOuterRings] = {};
        if (last)
            for (uint32_t j = 0; j < npack; ++j) {
                oq[j] = moduli[k + j];
                ostep[j] = (uint32_t)(M % oq[j]);
                ro[j].resize(cur.size());
                for (size_t x = 0; x < cur.size(); ++x) ro[j][x] = (uint8_t)(cur[x] % oq[j]);
            }
        const uint32_t mq = (uint32_t)(M % q);
        // distinct residues of coprime rings give exactly |R| * |R_q| survivors;
        // candidates are written unconditionally and kept by advancing n
        const size_t cap = cur.size() * sizes[i];
        next.resize(cap + 1);
        if (last) next_bytes.resize(cap + 1);
        size_t n = 0;
        uint32_t shift = 0;   // t*M mod q
        for (uint32_t t = 0; t < q; ++t) {
            const uint64_t base = (uint64_t)t * M;
            for (size_t x = 0; x < cur.size(); ++x) {
                uint32_t res = rq[x] + shift;
                res = res >= q ? res - q : res;
                const size_t at = n < cap ? n : cap;
                next[at] = base + cur[x];
                if (last) {
                    uint32_t b = 0;
                    for (uint32_t j = 0; j < npack; ++j) {
                        uint32_t m = ro[j][x] + osh[j];
                        b |= (m >= oq[j] ? m - oq[j] : m) << (8 * j);
                    }
                    next_bytes[at] = b;
                }
                n += ok[res];
            }
            shift += mq;
            shift = shift >= q ? shift - q : shift;
            for (uint32_t j = 0; j < npack; ++j) {
                osh[j] += ostep[j];
                osh[j] = osh[j] >= oq[j] ? osh[j] - oq[j] : osh[j];
            }
        }
        if (n > cap) return fail(PCUBE_INVALID, "focus rings are not pairwise coprime");
        next.resize(n);
        cur.swap(next);
        if (last) {
            next_bytes.resize(n);
            bytes.swap(next_bytes);
        }
        M *= q;
        offset += sizes[i];
    }
    if (k == 0) bytes.assign(1, 0);   // r = 0: every outer residue is 0
    wh.R1host.swap(cur);
    wh.M1 = M1;
    wh.n1 = (uint32_t)wh.R1host.size();
    wh.nouter = nrings - k;
    std::vector<uint32_t> bits;
    uint32_t bit = 0;
    for (uint32_t j = 0; j < wh.nouter; ++j) {
        const uint32_t q = moduli[k + j];
        Modulus m = make_modulus(q);
        m.b = (uint32_t)(M1 % q);
        m.bit = bit;
        bits.resize((bit + q + 31) / 32 + 1, 0);
        for (uint32_t e = 0; e < sizes[k + j]; ++e) {
            const uint32_t idx = bit + residues[offset + e];
            bits[idx / 32] |= 1u << (idx % 32);
        }
        offset += sizes[k + j];
        bit += q;
        wh.outer[j] = m;
        wh.rate *= (double)sizes[k + j] / q;
    }
    if (bits.empty()) bits.push_back(0);
    if (bits.size() > (size_t)kMaxOuterBits / 32) return fail(PCUBE_INVALID, "outer rings too large for the wheel");
    wh.allow_words = (uint32_t)bits.size();
    std::vector<uint64_t> entries(wh.n1);
    for (uint32_t i = 0; i < wh.n1; ++i)
        entries[i] = wh.R1host[i] | (wh.packed ? (uint64_t)bytes[i] << kResidueBits : 0ull);
    TRY(wh.R1.ensure(std::max<uint64_t>(8, (uint64_t)wh.n1 * 8)));
    TRY(wh.allow.ensure(bits.size() * 4));
    TRY(wh.flag.ensure(4));
    TRY(fill32(wh.flag.as<uint32_t>(), 1, 0));
    HIP_TRY(hipMemcpyAsync(wh.R1.p, entries.data(), (uint64_t)wh.n1 * 8, hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipMemcpyAsync(wh.allow.p, bits.data(), bits.size() * 4, hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    wh.ready = true;
    if (inner_modulus) *inner_modulus = M1;
    if (inner_residues) *inner_residues = wh.n1;
    return PCUBE_OK;
}

int pcube_focus_windows(int focus, uint32_t n, const uint64_t* lows, const uint64_t* highs, uint64_t out,
                        uint64_t cap, uint64_t* offsets) {
    Entry lock;
    Focus* f = focus_at(focus);
    if (!f) return fail(PCUBE_INVALID, "unknown focus");
    for (uint32_t w = 0; w < n; ++w)
        if (lows[w] > highs[w]) return fail(PCUBE_INVALID, "window %u has low > high", w);
    TRY(focus_windows(*f, n, lows, highs, (uint64_t*)out, cap, offsets));
    HIP_TRY(hipStreamSynchronize(g.stream));
    return PCUBE_OK;
}

static int match_start(int matcher, uint64_t us, uint64_t nu, uint64_t vs, uint64_t nv, uint64_t v_first,
                       uint64_t tile_width, uint32_t ntiles, uint64_t cap, int engine, const uint64_t* geo);

// A match enqueued right behind a batch's windows, before their sizes are
// known: nu and nv are estimates (the launch's sizes; the u window is padded to
// nu), geo receives the true counts on the device, and launched tells whether
// the match was enqueued.
struct Speculation {
    int matcher;
    uint64_t first, width, cap;
    int engine;
    uint64_t nu, nv;
    uint64_t* geo;
    bool launched;
};

// pcube_prepare, clearing nclear words at clear in the bounds kernel (the
// matcher's status and counters when the match follows in the same call);
// with spec, the match is enqueued behind the windows before the wait.
static int prepare_windows(int focus_u, int focus_v, uint32_t ntiles, const uint64_t* tiles, uint64_t us_out,
                           uint64_t us_cap, uint64_t vs_out, uint64_t vs_cap, uint64_t* nu, uint64_t* nv,
                           uint64_t* bounds, uint64_t* clear, uint64_t nclear, Speculation* spec = nullptr) {
    Focus* fu = focus_at(focus_u);
    Focus* fv = focus_at(focus_v);
    if (!fu || !fv || fu == fv || !fu->wheel.ready || !fv->wheel.ready || ntiles == 0 || ntiles > kMaxPrepareTiles)
        return fail(PCUBE_INVALID, "prepare needs the two focus wheels and 1 to 64 subtiles");
    const uint64_t ulo = tiles[2], uhi = tiles[4 * (ntiles - 1) + 3];
    const uint64_t vlo = tiles[0], vhi = tiles[4 * (ntiles - 1) + 1];
    if (ulo > uhi || vlo > vhi) return fail(PCUBE_INVALID, "invalid batch geometry");
    std::vector<WheelWindow> pu, pv;
    std::vector<uint64_t> su, sv;
    const uint64_t nseg_u = wheel_plan(fu->wheel, 1, &ulo, &uhi, pu, su);
    const uint64_t nseg_v = wheel_plan(fv->wheel, 1, &vlo, &vhi, pv, sv);
    TRY(wheel_check(nseg_u));
    TRY(wheel_check(nseg_v));
    // nothing is uploaded: the subtile queries ((u_min, u_end) per subtile, then
    // (v_first, v_end)) and both one-window plans travel with the kernel arguments
    const uint64_t nq = 4 * (uint64_t)ntiles, down_words = nq + 4;
    TRY(g_result_host.ensure(down_words * 8));
    PrepareQueries queries;
    for (uint32_t t = 0; t < ntiles; ++t) {
        queries.q[2 * t] = tiles[4 * t + 2];
        queries.q[2 * t + 1] = tiles[4 * t + 3];
        queries.q[2 * ntiles + 2 * t] = tiles[4 * t];
        queries.q[2 * ntiles + 2 * t + 1] = tiles[4 * t + 1];
    }
    WheelInputs in_u{}, in_v{};
    in_u.flag = fu->wheel.flag.as<uint32_t>();
    in_u.n = 1;
    in_u.nseg = nseg_u;
    in_u.win0 = pu[0];
    in_v.flag = fv->wheel.flag.as<uint32_t>();
    in_v.n = 1;
    in_v.nseg = nseg_v;
    in_v.win0 = pv[0];
    uint32_t cap_u = wheel_slot_cap(fu->wheel), cap_v = wheel_slot_cap(fv->wheel);
    // the bounds kernel writes its results straight into coherent pinned memory
    const uint64_t* r = g_result_host.as<uint64_t>();
    // the compaction answers the subtile queries itself (one window per wheel):
    // u queries to r[0, 2n), v queries to r[2n, 4n), the totals and overflow
    // flags to r[nq .. nq + 3]
    uint64_t* rw = g_result_host.as<uint64_t>();
    WheelPrepared wp;
    memset(&wp, 0, sizeof wp);
    static_assert(2 * kMaxPrepareTiles <= kMaxWindowQueries, "a wheel answers two queries per subtile");
    for (uint32_t k = 0; k < 2 * ntiles; ++k) {
        wp.qa.q[k] = queries.q[k];
        wp.qb.q[k] = queries.q[2 * ntiles + k];
    }
    wp.qa.nq = wp.qb.nq = 2 * ntiles;
    wp.qa.out = rw;
    wp.qb.out = rw + 2 * ntiles;
    wp.qa.total = rw + nq;
    wp.qb.total = rw + nq + 1;
    wp.qa.flag = rw + nq + 2;
    wp.qb.flag = rw + nq + 3;
    wp.clear = clear;
    wp.nclear = nclear;
    if (spec) {
        wp.qa.dev_total = spec->geo;
        wp.qb.dev_total = spec->geo + 1;
        wp.qa.pad_to = std::min(spec->nu, us_cap);
    }
    for (;;) {
        bool fused = false;
        if (spec && spec->launched) {
            // the windows are being made again (a slot overflowed): so is the match
            matcher_at(spec->matcher)->pending.active = false;
            matcher_at(spec->matcher)->head_zero = 0;
            spec->launched = false;
        }
        TRY(wheel_enqueue_pair(fu->wheel, in_u, (uint64_t*)us_out, us_cap, cap_u, fv->wheel, in_v, (uint64_t*)vs_out,
                               vs_cap, cap_v, &wp, &fused));
        if (!fused)
            hipLaunchKernelGGL(k_prepare_bounds, dim3(blocks_for(nq, 256 / kWave)), dim3(256), 0, g.stream,
                               (const uint64_t*)us_out, (const uint64_t*)fu->wheel.win_info.as<uint64_t>(), us_cap,
                               (const uint64_t*)vs_out, (const uint64_t*)fv->wheel.win_info.as<uint64_t>(), vs_cap,
                               queries, ntiles, rw, clear, nclear);
        // the host waits for the bounds only, not for a match queued behind them
        static hipEvent_t bounds_done = nullptr;
        if (!bounds_done) HIP_TRY(hipEventCreateWithFlags(&bounds_done, hipEventDisableTiming));
        HIP_TRY(hipEventRecord(bounds_done, g.stream));
        if (spec && fused) {
            // the compaction clears the matcher's status and counters ahead of its match
            Matcher* smt = matcher_at(spec->matcher);
            if (clear && clear == smt->status.as<uint64_t>()) smt->head_zero = nclear * 8;
            spec->launched = match_start(spec->matcher, us_out, spec->nu, vs_out, spec->nv, spec->first, spec->width,
                                         ntiles, spec->cap, spec->engine, spec->geo) == PCUBE_OK;
        }
        TRY(event_wait(bounds_done));
        const bool again_u = r[nq + 2] && cap_u < (uint32_t)kWheelWaveSpan;
        const bool again_v = r[nq + 3] && cap_v < (uint32_t)kWheelWaveSpan;
        if (!again_u && !again_v) break;
        // a slot overflowed: once more with room for every candidate
        if (again_u) { cap_u = kWheelWaveSpan; TRY(fill32(in_u.flag, 1, 0)); }
        if (again_v) { cap_v = kWheelWaveSpan; TRY(fill32(in_v.flag, 1, 0)); }
    }
    *nu = r[nq];
    *nv = r[nq + 1];
    if (*nu > us_cap || *nv > vs_cap)
        return fail(PCUBE_OVERFLOW, "focus buffer overflow; tile invalid (%llu, %llu coordinates)",
                    (unsigned long long)*nu, (unsigned long long)*nv);
    for (uint32_t t = 0; t < ntiles; ++t) {
        bounds[4 * t] = r[2 * t];
        bounds[4 * t + 1] = r[2 * t + 1];
        bounds[4 * t + 2] = r[2 * ntiles + 2 * t];
        bounds[4 * t + 3] = r[2 * ntiles + 2 * t + 1];
    }
    return PCUBE_OK;
}

int pcube_prepare(int focus_u, int focus_v, uint32_t ntiles, const uint64_t* tiles, uint64_t us_out, uint64_t us_cap,
                  uint64_t vs_out, uint64_t vs_cap, uint64_t* nu, uint64_t* nv, uint64_t* bounds) {
    Entry lock;
    return prepare_windows(focus_u, focus_v, ntiles, tiles, us_out, us_cap, vs_out, vs_cap, nu, nv, bounds, nullptr,
                           0);
}

int pcube_tile_ranges(uint64_t A, uint64_t B, uint64_t lo0, uint64_t lo1, uint64_t hi0, uint64_t hi1, uint64_t first,
                      uint64_t end, uint64_t width, uint64_t* tiles, uint32_t* ntiles) {
    *ntiles = 0;
    if (B == 0 || A >= (1ull << 63) || B >= (1ull << 63) || lo1 >= (1ull << 32) || hi1 >= (1ull << 32) ||
        first >= end || end > B || width == 0 || (end - first - 1) / width + 1 > 64)
        return fail(PCUBE_INVALID, "invalid batch geometry");
    // u_min = (lo + v A) / B + 1 for the first v of a subtile, u_end = (hi + v A) / B + 1
    // for its last v; lo, hi < 2^96 and v A < 2^126, so the sums fit 128 bits
    const unsigned __int128 lo = ((unsigned __int128)lo1 << 64) | lo0, hi = ((unsigned __int128)hi1 << 64) | hi0;
    uint32_t n = 0;
    for (uint64_t v = first;; v += width) {
        const bool last = end - v <= width;
        const uint64_t stop = last ? end : v + width;
        tiles[4 * n] = v;
        tiles[4 * n + 1] = stop;
        tiles[4 * n + 2] = (uint64_t)((lo + (unsigned __int128)v * A) / B) + 1;
        tiles[4 * n + 3] = (uint64_t)((hi + (unsigned __int128)(stop - 1) * A) / B) + 1;
        ++n;
        if (last) break;
    }
    *ntiles = n;
    return PCUBE_OK;
}

int pcube_tile_ranges_many(uint64_t A, uint64_t B, uint64_t lo0, uint64_t lo1, uint64_t hi0, uint64_t hi1, uint32_t n,
                           const uint64_t* firsts, const uint64_t* ends, const uint64_t* widths, uint64_t* tiles,
                           uint64_t tiles_cap, uint32_t* counts) {
    uint64_t at = 0;
    uint64_t one[4 * 64];
    for (uint32_t w = 0; w < n; ++w) {
        uint32_t got = 0;
        const int rc = pcube_tile_ranges(A, B, lo0, lo1, hi0, hi1, firsts[w], ends[w], widths[w], one, &got);
        if (rc) return fail(rc, "window %u: invalid batch geometry (a nonempty v range inside [0, B) of at most 64 subtiles)", w);
        if (at + got > tiles_cap)
            return fail(PCUBE_INVALID, "tile buffer of %llu subtiles too small at window %u", (unsigned long long)tiles_cap, w);
        memcpy(tiles + 4 * at, one, (size_t)got * 32);
        counts[w] = got;
        at += got;
    }
    return PCUBE_OK;
}

// pcube_prepare_range, clearing nclear words at clear as prepare_windows does.
static int prepare_range(int focus_u, int focus_v, uint64_t A, uint64_t B, uint64_t lo0, uint64_t lo1, uint64_t hi0,
                         uint64_t hi1, uint64_t first, uint64_t end, uint64_t width, uint64_t* tiles, uint32_t* ntiles,
                         uint64_t us_out, uint64_t us_cap, uint64_t vs_out, uint64_t vs_cap, uint64_t* nu, uint64_t* nv,
                         uint64_t* bounds, uint64_t* totals, uint64_t* clear, uint64_t nclear,
                         Speculation* spec = nullptr) {
    Focus* fu = focus_at(focus_u);
    Focus* fv = focus_at(focus_v);
    if (!fu || !fv) return fail(PCUBE_INVALID, "unknown focus");
    TRY(pcube_tile_ranges(A, B, lo0, lo1, hi0, hi1, first, end, width, tiles, ntiles));
    const uint32_t n = *ntiles;
    // the reference's bounded period count for both windows (pseudocube/gpu.py:100-102)
    const uint64_t ulo = tiles[2], uhi = tiles[4 * (n - 1) + 3];
    if (uhi > ulo && (uhi - 1) / fu->M - ulo / fu->M + 1 > 4096)
        return fail(PCUBE_RANGE, "focus window exceeds bounded period count");
    if ((end - 1) / fv->M - first / fv->M + 1 > 4096)
        return fail(PCUBE_RANGE, "focus window exceeds bounded period count");
    TRY(prepare_windows(focus_u, focus_v, n, tiles, us_out, us_cap, vs_out, vs_cap, nu, nv, bounds, clear, nclear,
                        spec));
    if (totals) {
        // bounds hold (u_first, u_end, v_first, v_end) index pairs < 2^32, so every product fits
        uint64_t focus = 0, opportunities = 0, largest = 0;
        for (uint32_t k = 0; k < n; ++k) {
            const uint64_t du = bounds[4 * k + 1] - bounds[4 * k], dv = bounds[4 * k + 3] - bounds[4 * k + 2];
            focus += du;
            opportunities += du * dv;
            largest = std::max(largest, du * dv);
        }
        totals[0] = focus;
        totals[1] = opportunities;
        totals[2] = largest;
    }
    return PCUBE_OK;
}

int pcube_prepare_range(int focus_u, int focus_v, uint64_t A, uint64_t B, uint64_t lo0, uint64_t lo1, uint64_t hi0,
                        uint64_t hi1, uint64_t first, uint64_t end, uint64_t width, uint64_t* tiles, uint32_t* ntiles,
                        uint64_t us_out, uint64_t us_cap, uint64_t vs_out, uint64_t vs_cap, uint64_t* nu, uint64_t* nv,
                        uint64_t* bounds, uint64_t* totals) {
    Entry lock;
    return prepare_range(focus_u, focus_v, A, B, lo0, lo1, hi0, hi1, first, end, width, tiles, ntiles, us_out, us_cap,
                         vs_out, vs_cap, nu, nv, bounds, totals, nullptr, 0);
}

int pcube_prepare_match(int focus_u, int focus_v, uint64_t A, uint64_t B, uint64_t lo0, uint64_t lo1, uint64_t hi0,
                        uint64_t hi1, uint64_t first, uint64_t end, uint64_t width, uint64_t* tiles, uint32_t* ntiles,
                        uint64_t us_out, uint64_t us_cap, uint64_t vs_out, uint64_t vs_cap, uint64_t* nu, uint64_t* nv,
                        uint64_t* bounds, uint64_t* totals, int matcher, uint64_t cap, int engine,
                        uint64_t max_opportunities, uint64_t* launch) {
    Entry lock;
    if (!totals || !launch) return fail(PCUBE_INVALID, "null totals or launch");
    *launch = 0;
    // a matcher with no batch to collect gets its status words and counters
    // cleared by this batch's bounds kernel, so its match needs no fill launch;
    // once it has seen a batch, the match is enqueued behind the windows sized
    // from the last batch's windows plus a margin, before their sizes are known
    Matcher* mt = matcher_at(matcher);
    const bool idle = mt && !mt->pending.active;
    const bool clears = idle && mt->status.bytes;
    Speculation spec{};
    Speculation* sp = nullptr;
    if (idle && engine == 0 && mt->sbits && mt->last_nu && mt->last_nv && mt->geo.ensure(16) == PCUBE_OK) {
        spec.matcher = matcher;
        spec.first = first;
        spec.width = width;
        spec.cap = cap;
        spec.engine = engine;
        spec.nu = std::min(us_cap, mt->last_nu + mt->last_nu / 32 + 256);
        spec.nv = std::min(vs_cap, mt->last_nv + mt->last_nv / 32 + 256);
        spec.geo = mt->geo.as<uint64_t>();
        sp = &spec;
    }
    const int rc = prepare_range(focus_u, focus_v, A, B, lo0, lo1, hi0, hi1, first, end, width, tiles, ntiles, us_out,
                                 us_cap, vs_out, vs_cap, nu, nv, bounds, totals,
                                 clears ? mt->status.as<uint64_t>() : nullptr, clears ? mt->status.bytes / 8 : 0, sp);
    if (rc != PCUBE_OK) {
        if (spec.launched) {   // a failed prepare leaves no batch to collect
            mt->pending.active = false;
            mt->head_zero = 0;
        }
        return rc;
    }
    if (clears && !spec.launched) mt->head_zero = mt->status.bytes;
    if (idle) {
        mt->last_nu = *nu;
        mt->last_nv = *nv;
    }
    if (spec.launched) {
        Pending& pd = mt->pending;
        if (*nu <= spec.nu && *nv <= spec.nv && totals[2] <= max_opportunities) {
            pd.nu_real = *nu;
            pd.nv_real = *nv;
            *launch = mt->launches;
            return PCUBE_OK;
        }
        // the estimate fell short, or the batch will be rejected: the enqueued
        // match runs to its end unread, having written status and counters
        pd.active = false;
        mt->head_zero = 0;
    }
    // a batch the caller will reject, or a matcher with a batch still to collect, is left alone
    if (!mt || mt->pending.active || totals[2] > max_opportunities) return PCUBE_OK;
    // a start that fails here is left for pcube_match_start to report when the batch is run
    if (match_start(matcher, us_out, *nu, vs_out, *nv, first, width, *ntiles, cap, engine, nullptr) == PCUBE_OK)
        *launch = mt->launches;
    return PCUBE_OK;
}

int pcube_match_adoptable(int matcher, uint64_t launch, int* adoptable) {
    Entry lock;
    if (!adoptable) return fail(PCUBE_INVALID, "null result");
    Matcher* mt = matcher_at(matcher);
    *adoptable = mt && launch && mt->pending.active && mt->launches == launch ? 1 : 0;
    return PCUBE_OK;
}

int pcube_searchsorted(uint64_t sorted, uint64_t n, const uint64_t* queries, uint64_t nq, uint64_t* out) {
    Entry lock;
    if (!nq) return PCUBE_OK;
    void *dq = nullptr, *dout = nullptr;
    TRY(dev_alloc(nq * 8, &dq));
    int rc = dev_alloc(nq * 8, &dout);
    if (rc) { dev_free(dq); return rc; }
    HIP_TRY(hipMemcpyAsync(dq, queries, nq * 8, hipMemcpyHostToDevice, g.stream));
    hipLaunchKernelGGL(k_searchsorted, dim3(blocks_for(nq, 256)), dim3(256), 0, g.stream, (const uint64_t*)sorted, n,
                       (const uint64_t*)dq, nq, (uint64_t*)dout);
    HIP_TRY(hipMemcpyAsync(out, dout, nq * 8, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    dev_free(dq);
    dev_free(dout);
    return PCUBE_OK;
}

int pcube_matcher_create(uint64_t A, uint64_t B, uint64_t lo0, uint64_t lo1, uint64_t hi0, uint64_t hi1,
                         const uint32_t* moduli, uint32_t nres, int* matcher) {
    Entry lock;
    if (g.device < 0) return fail(PCUBE_INVALID, "pcube_init has not been called");
    if (nres == 0 || nres > (uint32_t)kMaxResiduals) return fail(PCUBE_INVALID, "between 1 and 128 residual filters required");
    if (B == 0) return fail(PCUBE_INVALID, "B must be positive");
    auto mt = std::make_unique<Matcher>();
    mt->A = A; mt->B = B;
    const unsigned __int128 mu = ~(unsigned __int128)0 / B;
    mt->B_mu = U128{(uint64_t)mu, (uint64_t)(mu >> 64)};
    mt->lo = U128{lo0, lo1};
    mt->hi = U128{hi0, hi1};
    mt->nres = nres;
    mt->nbits = std::min<uint32_t>(nres, kMaxBitsetFilters);
    mt->ncache = std::min<uint32_t>(nres - mt->nbits, kCachedResidues);
    mt->ng = (nres + 3) / 4;
    std::vector<uint32_t> allow;
    uint32_t bit = 0;
    for (uint32_t k = 0; k < nres; ++k) {
        const uint32_t q = moduli[k];
        if (q < 2 || q >= 65536) return fail(PCUBE_INVALID, "modulus %u out of range", q);
        Modulus m;
        m.q = q;
        m.inv = (uint32_t)((1ull << 32) / q);
        m.r32 = (uint32_t)((1ull << 32) % q);
        m.r64 = (uint32_t)(((unsigned __int128)1 << 64) % q);
        m.a = (uint32_t)(A % q);
        m.b = (uint32_t)(B % q);
        m.bit = bit;
        m.pad = 0;
        if (k < mt->nbits && m.a == 0) return fail(PCUBE_INVALID, "bitset modulus %u divides A", q);
        for (uint32_t a = 0; a < q; ++a) {
            const uint32_t idx = bit + a;
            if (allow.size() <= idx / 32) allow.resize(idx / 32 + 1, 0);
            if (allowed_residue(q, a)) allow[idx / 32] |= 1u << (idx % 32);
        }
        bit += q;
        mt->mods.push_back(m);
    }
    allow.resize((allow.size() + 3) / 4 * 4, 0);   // whole 16-byte units for staged copies
    if (allow.size() > (size_t)kMaxAllowWords) return fail(PCUBE_INVALID, "allowed-residue table too large");
    mt->allow_words = (uint32_t)allow.size();
    uint32_t rb = 0;
    for (uint32_t k = 0; k < kMaxBitsetFilters; ++k) {
        mt->row_base[k] = rb;
        if (k < mt->nbits) rb += mt->mods[k].q;
    }
    mt->mask_rows = rb;
    // staged engine: the longest prefix of at most kStagedFilters filters whose
    // q rows fit a column of the chunk image and whose A residue is invertible;
    // filter k keeps P[j] = allowed_q[(-j*A) mod q] for j < q + 64, plus two
    // words of margin
    std::vector<uint32_t> pstr;
    for (uint32_t k = 0; k < std::min<uint32_t>(nres, kStagedFilters); ++k) {
        const Modulus& md = mt->mods[k];
        if (md.q > (uint32_t)kStageLines || md.a == 0 || std::gcd(md.a, md.q) != 1) break;
        const uint64_t ainv = modinv(md.a, md.q);
        mt->s_nbeta[k] = (uint32_t)((md.q - (uint64_t)md.b * ainv % md.q) % md.q);
        mt->s_pstr_off[k] = (uint32_t)pstr.size();
        const uint32_t nbits = md.q + kWave, nwords = (nbits + 31) / 32 + 2;
        const size_t o = pstr.size();
        pstr.resize(o + nwords, 0);
        for (uint32_t j = 0; j < nbits; ++j) {
            const uint32_t res = (uint32_t)((md.q - (uint64_t)(j % md.q) * md.a % md.q) % md.q);
            if (allowed_residue(md.q, res)) pstr[o + j / 32] |= 1u << (j % 32);
        }
        mt->slines = std::max(mt->slines, md.q);
        ++mt->sbits;
    }
    mt->sncache = std::min<uint32_t>(nres - mt->sbits, kCachedResidues);
    if (mt->sbits) {
        TRY(mt->d_pstr.ensure(pstr.size() * 4));
        HIP_TRY(hipMemcpyAsync(mt->d_pstr.p, pstr.data(), pstr.size() * 4, hipMemcpyHostToDevice, g.stream));
    }
    TRY(mt->d_mods.ensure(nres * sizeof(Modulus)));
    TRY(mt->d_allow.ensure(allow.size() * 4));
    TRY(mt->status.ensure(4 * 8));
    HIP_TRY(hipMemcpyAsync(mt->d_mods.p, mt->mods.data(), nres * sizeof(Modulus), hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipMemcpyAsync(mt->d_allow.p, allow.data(), allow.size() * 4, hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    g_matchers.push_back(std::move(mt));
    *matcher = (int)g_matchers.size() - 1;
    return PCUBE_OK;
}

int pcube_matcher_table_fill(int matcher, int value) {
    Entry lock;
    Matcher* mt = matcher_at(matcher);
    if (!mt) return fail(PCUBE_INVALID, "unknown matcher");
    if (value != 0 && value != 1) return fail(PCUBE_INVALID, "a table fill takes 0 or 1");
    const uint32_t word = value ? 0xffffffffu : 0u;
    for (Buffer* b : {&mt->d_allow, &mt->d_pstr})
        if (b->p && b->bytes / 4)
            hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(b->bytes / 4, 256)), dim3(256), 0, g.stream,
                               b->as<uint32_t>(), b->bytes / 4, word);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(g.stream));
    return PCUBE_OK;
}

int pcube_matcher_destroy(int matcher) {
    Entry lock;
    if (!matcher_at(matcher)) return fail(PCUBE_INVALID, "unknown matcher");
    HIP_TRY(hipStreamSynchronize(g.stream));
    g_matchers[matcher].reset();
    return PCUBE_OK;
}

// pcube_match_start; with geo, nu and nv are upper bounds and geo receives the
// true counts on the device (Pending::geo).
static int match_start(int matcher, uint64_t us, uint64_t nu, uint64_t vs, uint64_t nv, uint64_t v_first,
                       uint64_t tile_width, uint32_t ntiles, uint64_t cap, int engine, const uint64_t* geo) {
    Matcher* mt = matcher_at(matcher);
    if (!mt) return fail(PCUBE_INVALID, "unknown matcher");
    if (ntiles == 0 || ntiles > 65536 || tile_width == 0 || (engine != 0 && engine != 1))
        return fail(PCUBE_INVALID, "invalid tile geometry or engine");
    Pending& pd = mt->pending;
    pd = Pending{};
    pd.us = us; pd.nu = nu; pd.vs = vs; pd.nv = nv; pd.v_first = v_first; pd.tile_width = tile_width;
    pd.ntiles = ntiles; pd.cap = cap; pd.engine = engine; pd.geo = geo;
    const uint64_t nc = (uint64_t)ntiles * (mt->ng + 3);
    pd.list_cap = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)ntiles * cap, 1ull << 22));
    if (mt->status.bytes < (nc + 4) * 8) mt->head_zero = 0;   // a new buffer holds no zeros yet
    TRY(mt->status.ensure((nc + 4) * 8));   // status words, then the batch's counters
    TRY(mt->list.ensure(pd.list_cap * 24));
    TRY(mt->down.ensure((nc + 4) * 8));
    for (hipEvent_t& e : mt->ev)
        if (!e) HIP_TRY(hipEventCreate(&e));
    TRY(match_launch(*mt));
    pd.launch = mt->launches;
    pd.active = true;
    return PCUBE_OK;
}

int pcube_match_start(int matcher, uint64_t us, uint64_t nu, uint64_t vs, uint64_t nv, uint64_t v_first,
                      uint64_t tile_width, uint32_t ntiles, uint64_t cap, int engine) {
    Entry lock;
    return match_start(matcher, us, nu, vs, nv, v_first, tile_width, ntiles, cap, engine, nullptr);
}

int pcube_match_finish(int matcher, uint64_t* counters, uint64_t* values, uint64_t values_cap, double* seconds) {
    Entry lock;
    Matcher* mt = matcher_at(matcher);
    if (!mt) return fail(PCUBE_INVALID, "unknown matcher");
    Pending& pd = mt->pending;
    if (!pd.active) return fail(PCUBE_INVALID, "no batch started on this matcher");
    pd.active = false;
    if (mt->launches != pd.launch)
        return fail(PCUBE_INVALID, "a lease on this matcher reused the batch buffers before the batch was collected");
    const uint64_t nc = (uint64_t)pd.ntiles * (mt->ng + 3);
    uint64_t* hd = mt->down.as<uint64_t>();
    TRY(event_wait(mt->ev[5]));
    if (hd[3] > mt->cand_cap) {
        // a bitset-survivor region overflowed: size the regions from its count and run again
        if (hd[3] > kMaxWaveCandidates) {
            pd.engine = kUnstagedBitsets;
            if (pd.geo) {   // the unstaged engine takes exact sizes
                pd.nu = pd.nu_real;
                pd.nv = pd.nv_real;
                pd.geo = nullptr;
            }
        } else {
            mt->cand_cap = (hd[3] + 63) / 64 * 64;
        }
        TRY(match_launch(*mt));
        TRY(event_wait(mt->ev[5]));
    }
    const uint64_t listed = hd[0];
    const uint32_t err = (uint32_t)hd[1];
    // the counters are read where the download left them: the pinned buffer is
    // rewritten only by this matcher's next launch
    const uint64_t* host = hd + 4;
    if (seconds) {
        float a = 0, b = 0, c = 0;
        HIP_TRY(hipEventElapsedTime(&a, mt->ev[0], mt->ev[1]));
        HIP_TRY(hipEventElapsedTime(&b, mt->ev[1], mt->ev[2]));
        if (mt->kernel_span && pd.nv) HIP_TRY(hipEventElapsedTime(&c, mt->ev[1], mt->ev[2]));
        else HIP_TRY(hipEventElapsedTime(&c, mt->ev[3], mt->ev[4]));
        seconds[0] = a / 1000.0;
        seconds[1] = b / 1000.0;
        seconds[2] = c / 1000.0;
    }
    memcpy(counters, host, nc * 8);
    int rc = batch_status(*mt, host, pd.ntiles, pd.cap, err);
    int status = PCUBE_OK;
    if (!rc) rc = collect_values(*mt, pd.ntiles, host, pd.cap, pd.list_cap, values, values_cap, &status, listed);
    // status and counters are cleared before the next match: by the next
    // prepare's bounds kernel, or by the fill the next launch adds itself
    return rc ? rc : status;
}

int pcube_match(int matcher, uint64_t us, uint64_t nu, uint64_t vs, uint64_t nv, uint64_t v_first,
                uint64_t tile_width, uint32_t ntiles, uint64_t cap, int engine, uint64_t* counters, uint64_t* values,
                uint64_t values_cap, double* seconds) {
    Entry lock;
    TRY(pcube_match_start(matcher, us, nu, vs, nv, v_first, tile_width, ntiles, cap, engine));
    return pcube_match_finish(matcher, counters, values, values_cap, seconds);
}

int pcube_lease(int matcher, int focus_u, int focus_v, uint32_t nbatches, const uint32_t* batch_tiles,
                const uint64_t* tiles, uint64_t tile_width, uint64_t cap, uint64_t window_cap, uint64_t* bounds,
                uint64_t* batch_sizes, uint64_t* counters, uint64_t* values, uint64_t values_cap, double* seconds) {
    Entry lock;
    const double t_start = now_seconds();
    Matcher* mt = matcher_at(matcher);
    Focus* fu = focus_at(focus_u);
    Focus* fv = focus_at(focus_v);
    if (!mt || !fu || !fv || nbatches == 0) return fail(PCUBE_INVALID, "invalid lease arguments");
    uint64_t ntot = 0;
    for (uint32_t b = 0; b < nbatches; ++b) {
        if (batch_tiles[b] == 0 || batch_tiles[b] > 64) return fail(PCUBE_INVALID, "1..64 subtiles per batch");
        ntot += batch_tiles[b];
    }
    const uint64_t u_low = tiles[2], u_high = tiles[(ntot - 1) * 4 + 3];
    const uint64_t v_low = tiles[0], v_high = tiles[(ntot - 1) * 4 + 1];
    if (u_low > u_high || v_low >= v_high) return fail(PCUBE_INVALID, "invalid lease geometry");
    // lease-wide windows
    Buffer lus, lvs, queries, positions;
    TRY(lus.ensure(std::max<uint64_t>(8, window_cap * 8)));
    TRY(lvs.ensure(std::max<uint64_t>(8, window_cap * 8)));
    hipEvent_t ev[4];
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
    HIP_TRY(hipEventRecord(ev[0], g.stream));
    uint64_t totals[2] = {0, 0};
    // (u_min, u_end) queries into the u window, (v_first, v_end) into the v window
    const uint64_t nq = 4 * ntot;
    TRY(g_lease_up.ensure(nq * 8));
    uint64_t* q = g_lease_up.as<uint64_t>();
    for (uint64_t t = 0; t < ntot; ++t) {
        q[2 * t] = tiles[t * 4 + 2];
        q[2 * t + 1] = tiles[t * 4 + 3];
        q[2 * ntot + 2 * t] = tiles[t * 4 + 0];
        q[2 * ntot + 2 * t + 1] = tiles[t * 4 + 1];
    }
    TRY(queries.ensure(nq * 8));
    const uint64_t* up;   // subtile bounds: (u_min, u_end) positions per subtile, then (v_first, v_end)
    std::vector<uint64_t> fallback;
    if (fu->wheel.ready && fv->wheel.ready && fu != fv) {
        // both windows in one paired wheel launch, the bound searches behind them
        // reading the window sizes on the device, and one synchronisation: the
        // results arrive in coherent pinned memory
        HIP_TRY(hipMemcpyAsync(queries.p, q, nq * 8, hipMemcpyHostToDevice, g.stream));
        std::vector<WheelWindow> pu, pv;
        std::vector<uint64_t> su, sv;
        const uint64_t u_end = std::max(u_low, u_high);
        const uint64_t nseg_u = wheel_plan(fu->wheel, 1, &u_low, &u_end, pu, su);
        const uint64_t nseg_v = wheel_plan(fv->wheel, 1, &v_low, &v_high, pv, sv);
        TRY(wheel_check(nseg_u));
        TRY(wheel_check(nseg_v));
        WheelInputs in_u{}, in_v{};
        in_u.flag = fu->wheel.flag.as<uint32_t>();
        in_u.n = 1;
        in_u.nseg = nseg_u;
        in_u.win0 = pu[0];
        in_v.flag = fv->wheel.flag.as<uint32_t>();
        in_v.n = 1;
        in_v.nseg = nseg_v;
        in_v.win0 = pv[0];
        uint32_t cap_u = wheel_slot_cap(fu->wheel), cap_v = wheel_slot_cap(fv->wheel);
        TRY(g_result_host.ensure((nq + 4) * 8));
        const uint64_t* r = g_result_host.as<uint64_t>();
        for (;;) {
            TRY(wheel_enqueue_pair(fu->wheel, in_u, lus.as<uint64_t>(), window_cap, cap_u, fv->wheel, in_v,
                                   lvs.as<uint64_t>(), window_cap, cap_v));
            hipLaunchKernelGGL(k_lease_bounds, dim3(blocks_for(nq, 256 / kWave)), dim3(256), 0, g.stream,
                               (const uint64_t*)lus.as<uint64_t>(), (const uint64_t*)fu->wheel.win_info.as<uint64_t>(),
                               window_cap, (const uint64_t*)lvs.as<uint64_t>(),
                               (const uint64_t*)fv->wheel.win_info.as<uint64_t>(), window_cap,
                               (const uint64_t*)queries.as<uint64_t>(), 2 * ntot, 2 * ntot,
                               g_result_host.as<uint64_t>());
            HIP_TRY(hipGetLastError());
            TRY(stream_wait());
            const bool again_u = r[nq + 2] && cap_u < (uint32_t)kWheelWaveSpan;
            const bool again_v = r[nq + 3] && cap_v < (uint32_t)kWheelWaveSpan;
            if (!again_u && !again_v) break;
            // a slot overflowed: once more with room for every candidate
            if (again_u) { cap_u = kWheelWaveSpan; TRY(fill32(in_u.flag, 1, 0)); }
            if (again_v) { cap_v = kWheelWaveSpan; TRY(fill32(in_v.flag, 1, 0)); }
        }
        totals[0] = r[nq];
        totals[1] = r[nq + 1];
        if (totals[0] > window_cap || totals[1] > window_cap)
            return fail(PCUBE_OVERFLOW, "lease focus window overflow (%llu, %llu > %llu)",
                        (unsigned long long)totals[0], (unsigned long long)totals[1], (unsigned long long)window_cap);
        up = r;
    } else {
        uint64_t offs[2] = {0, 0};
        if (u_high > u_low) {
            TRY(lease_window(*fu, u_low, u_high, lus.as<uint64_t>(), window_cap, offs));
            totals[0] = offs[1];
        }
        TRY(lease_window(*fv, v_low, v_high, lvs.as<uint64_t>(), window_cap, offs));
        totals[1] = offs[1];
        TRY(positions.ensure(nq * 8));
        HIP_TRY(hipMemcpyAsync(queries.p, q, nq * 8, hipMemcpyHostToDevice, g.stream));
        hipLaunchKernelGGL(k_searchsorted, dim3(blocks_for(2 * ntot, 256)), dim3(256), 0, g.stream,
                           lus.as<uint64_t>(), (uint64_t)totals[0], queries.as<uint64_t>(), 2 * ntot,
                           positions.as<uint64_t>());
        hipLaunchKernelGGL(k_searchsorted, dim3(blocks_for(2 * ntot, 256)), dim3(256), 0, g.stream,
                           lvs.as<uint64_t>(), (uint64_t)totals[1], queries.as<uint64_t>() + 2 * ntot, 2 * ntot,
                           positions.as<uint64_t>() + 2 * ntot);
        fallback.resize(nq);
        HIP_TRY(hipMemcpyAsync(fallback.data(), positions.p, nq * 8, hipMemcpyDeviceToHost, g.stream));
        TRY(stream_wait());
        up = fallback.data();
    }
    const uint64_t* vp = up + 2 * ntot;
    HIP_TRY(hipEventRecord(ev[1], g.stream));
    // per batch geometry
    const uint64_t stride = mt->ng + 3;
    const uint64_t list_cap = std::max<uint64_t>(1, std::min<uint64_t>(ntot * cap, 1ull << 22));
    ++mt->launches;   // the batch buffers are reused: a started batch can no longer be collected
    TRY(mt->counters.ensure(ntot * stride * 8));
    TRY(mt->list.ensure(list_cap * 24));
    // every batch of the lease, then one download; a bitset-survivor list that
    // overflowed in any batch is sized from the largest count and the lease
    // runs again
    TRY(g_lease_down.ensure((ntot * stride + 4) * 8));
    uint64_t* host = g_lease_down.as<uint64_t>();
    uint64_t* st = host + ntot * stride;   // status words after the counters
    int geometry_rc = PCUBE_OK;
    int engine = 0;
    for (int attempt = 0; attempt < 2; ++attempt) {
        TRY(fill64(mt->counters.as<uint64_t>(), ntot * stride, 0));
        mt->head_zero = 0;
        TRY(fill64(mt->status.as<uint64_t>(), 4, 0));
        uint64_t t0 = 0;
        for (uint32_t b = 0; b < nbatches; ++b) {
            const uint64_t nt = batch_tiles[b], t1 = t0 + nt;
            const uint64_t ua = up[2 * t0], ub = up[2 * (t1 - 1) + 1];
            const uint64_t va = vp[2 * t0], vb = vp[2 * (t1 - 1) + 1];
            batch_sizes[2 * b] = ub - ua;
            batch_sizes[2 * b + 1] = vb - va;
            for (uint64_t t = t0; t < t1; ++t) {
                bounds[4 * t + 0] = up[2 * t] - ua;
                bounds[4 * t + 1] = up[2 * t + 1] - ua;
                bounds[4 * t + 2] = vp[2 * t] - va;
                bounds[4 * t + 3] = vp[2 * t + 1] - va;
                const uint64_t du = bounds[4 * t + 1] - bounds[4 * t], dv = bounds[4 * t + 3] - bounds[4 * t + 2];
                if (du * dv > 50000000ull && geometry_rc == PCUBE_OK)
                    geometry_rc = fail(PCUBE_OVERFLOW, "per-subtile candidate opportunity limit exceeded");
            }
            if (geometry_rc == PCUBE_OK) {
                TRY(match_enqueue(*mt, lus.as<uint64_t>() + ua, ub - ua, lvs.as<uint64_t>() + va, vb - va,
                                  tiles[4 * t0], tile_width, (uint32_t)nt, cap, engine,
                                  mt->counters.as<uint64_t>() + t0 * stride, t0, list_cap, nullptr, nullptr));
            }
            t0 = t1;
        }
        HIP_TRY(hipEventRecord(ev[2], g.stream));
        HIP_TRY(hipMemcpyAsync(host, mt->counters.p, ntot * stride * 8, hipMemcpyDeviceToHost, g.stream));
        HIP_TRY(hipMemcpyAsync(st, mt->status.p, 4 * 8, hipMemcpyDeviceToHost, g.stream));
        HIP_TRY(hipEventRecord(ev[3], g.stream));
        TRY(stream_wait());
        if (geometry_rc || st[3] <= mt->cand_cap) break;
        if (st[3] > kMaxWaveCandidates) engine = kUnstagedBitsets;
        else mt->cand_cap = (st[3] + 63) / 64 * 64;
    }
    const uint32_t err = (uint32_t)st[1];
    const uint64_t listed = st[0];
    memcpy(counters, host, ntot * stride * 8);
    if (seconds) {
        float a = 0, m = 0;
        HIP_TRY(hipEventElapsedTime(&a, ev[0], ev[1]));
        HIP_TRY(hipEventElapsedTime(&m, ev[1], ev[2]));
        seconds[0] = a / 1000.0;
        seconds[1] = 0.0;
        seconds[2] = m / 1000.0;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    lus.release(); lvs.release(); queries.release(); positions.release();
    if (geometry_rc) return geometry_rc;
    int rc = batch_status(*mt, counters, ntot, cap, err);
    if (rc) return rc;
    int status = PCUBE_OK;
    TRY(collect_values(*mt, ntot, counters, cap, list_cap, values, values_cap, &status, listed));
    if (seconds) seconds[3] = now_seconds() - t_start;
    return status;
}

int pcube_batch_create(const int* matchers, uint32_t ncfg, int focus_u, int focus_v, uint64_t cap,
                       uint32_t blocks_per_cu, int use_graph, int* engine) {
    Entry lock;
    if (ncfg == 0 || ncfg > (uint32_t)kMaxConfigs) return fail(PCUBE_INVALID, "1 to %d configurations per batch", kMaxConfigs);
    Focus* fu = focus_at(focus_u);
    Focus* fv = focus_at(focus_v);
    if (!fu || !fv || fu == fv || !fu->wheel.ready || !fv->wheel.ready)
        return fail(PCUBE_INVALID, "batched form needs the two distinct focus wheels");
    auto e = std::make_unique<BatchEngine>();
    memset(&e->args, 0, sizeof e->args);
    for (uint32_t c = 0; c < ncfg; ++c) {
        Matcher* mt = matcher_at(matchers[c]);
        if (!mt) return fail(PCUBE_INVALID, "unknown matcher");
        Matcher* m0 = matcher_at(matchers[0]);
        if (mt->A != m0->A || mt->B != m0->B || mt->lo.lo != m0->lo.lo || mt->lo.hi != m0->lo.hi ||
            mt->hi.lo != m0->hi.lo || mt->hi.hi != m0->hi.hi)
            return fail(PCUBE_INVALID, "batched configurations must share A, B and the interval");
        BatchConfig& bc = e->args.cfg[c];
        bc.mods = mt->d_mods.as<Modulus>();
        bc.allow = mt->d_allow.as<uint32_t>();
        bc.nres = mt->nres; bc.nbits = mt->nbits; bc.ncache = mt->ncache; bc.ng = mt->ng;
        bc.allow_words = mt->allow_words; bc.mask_rows = mt->mask_rows;
        // the staged builder's permuted strings cover the four bitset filters when its
        // staged prefix is at least as long
        bc.pstr = mt->d_pstr.as<uint32_t>();
        bc.transpose = mt->sbits >= mt->nbits && mt->d_pstr.p ? 1u : 0u;
        for (uint32_t k = 0; k < kMaxBitsetFilters; ++k) {
            bc.pstr_off[k] = k < mt->sbits ? mt->s_pstr_off[k] : 0u;
            bc.nbeta[k] = k < mt->sbits ? mt->s_nbeta[k] : 0u;
        }
        bc.chunk_base[0] = 0;
        for (uint32_t k = 0; k < kMaxBitsetFilters; ++k) {
            bc.row_base[k] = mt->row_base[k];
            bc.chunk_base[k + 1] = bc.chunk_base[k] + (k < mt->nbits ? (mt->mods[k].q + kWave - 1) / kWave : 0);
        }
        e->matchers.push_back(matchers[c]);
    }
    Matcher* m0 = matcher_at(matchers[0]);
    e->args.A = m0->A; e->args.B = m0->B; e->args.B_mu = m0->B_mu; e->args.lo = m0->lo; e->args.hi = m0->hi;
    e->args.ncfg = ncfg;
    e->focus_u = focus_u; e->focus_v = focus_v;
    e->cap = cap;
    e->args.cap = cap;
    e->args.list_cap = e->list_cap;
    e->blocks_match = std::max<uint32_t>(1, blocks_per_cu) * (uint32_t)std::max(1, g.compute_units);
    e->use_graph = use_graph != 0;
    TRY(e->cfg.ensure(sizeof(BatchConfig) * kMaxConfigs));
    HIP_TRY(hipMemcpyAsync(e->cfg.p, e->args.cfg, sizeof(BatchConfig) * ncfg, hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    e->args.dcfg = e->cfg.as<BatchConfig>();
    TRY(e->totals.ensure(sizeof(BatchTotals)));
    TRY(e->list.ensure(e->list_cap * 24));
    TRY(e->list_count.ensure(8));
    TRY(e->error.ensure(4));
    TRY(e->cursor.ensure(4));
    g_batch.push_back(std::move(e));
    *engine = (int)g_batch.size() - 1;
    return PCUBE_OK;
}

int pcube_batch_destroy(int engine) {
    Entry lock;
    if (engine < 0 || engine >= (int)g_batch.size() || !g_batch[engine]) return fail(PCUBE_INVALID, "unknown batch engine");
    HIP_TRY(hipStreamSynchronize(g.stream));
    g_batch[engine].reset();
    return PCUBE_OK;
}

int pcube_batch_run(int engine, uint32_t nwin, const uint32_t* win_cfg, const uint32_t* win_tiles,
                    const uint64_t* tiles, const uint64_t* tile_width, const uint32_t* mode, uint32_t rows_per_chunk,
                    uint64_t* bounds, uint64_t* sizes, uint64_t* counters, uint64_t* values, uint64_t values_cap,
                    double* seconds, uint64_t* graph_launches) {
    Entry lock;
    const double t_start = now_seconds();
    if (engine < 0 || engine >= (int)g_batch.size() || !g_batch[engine]) return fail(PCUBE_INVALID, "unknown batch engine");
    BatchEngine& e = *g_batch[engine];
    if (nwin == 0 || rows_per_chunk == 0) return fail(PCUBE_INVALID, "empty batch");
    Focus* fu = focus_at(e.focus_u);
    Focus* fv = focus_at(e.focus_v);
    // window geometry from the subtiles
    std::vector<uint64_t> tfirst(nwin + 1, 0);
    for (uint32_t w = 0; w < nwin; ++w) {
        if (win_cfg[w] >= e.args.ncfg || win_tiles[w] == 0 || tile_width[w] == 0 || mode[w] > 1)
            return fail(PCUBE_INVALID, "window %u: bad configuration, subtiles, width or engine", w);
        tfirst[w + 1] = tfirst[w] + win_tiles[w];
    }
    const uint64_t ntot = tfirst[nwin];
    std::vector<uint64_t> ulo(nwin), uhi(nwin), vlo(nwin), vhi(nwin);
    for (uint32_t w = 0; w < nwin; ++w) {
        const uint64_t a = tfirst[w], z = tfirst[w + 1] - 1;
        ulo[w] = tiles[4 * a + 2];
        uhi[w] = tiles[4 * z + 3];
        vlo[w] = tiles[4 * a + 0];
        vhi[w] = tiles[4 * z + 1];
        if (ulo[w] > uhi[w] || vlo[w] >= vhi[w]) return fail(PCUBE_INVALID, "window %u: empty or inverted range", w);
    }
    hipEvent_t ev[4];
    for (auto& x : ev) HIP_TRY(hipEventCreate(&x));
    HIP_TRY(hipEventRecord(ev[0], g.stream));
    // all windows of both wheels, into buffers sized by the expected counts;
    // an overflow sizes them exactly and runs again
    std::vector<uint64_t> offu(nwin + 1), offv(nwin + 1);
    auto windows_of = [&](Focus& f, const std::vector<uint64_t>& lo, const std::vector<uint64_t>& hi, Buffer& buf,
                          std::vector<uint64_t>& off) -> int {
        double span = 0;
        for (uint32_t w = 0; w < nwin; ++w) span += (double)(hi[w] - lo[w]);
        const double density = f.wheel.rate * f.wheel.n1 / (double)f.wheel.M1;
        uint64_t cap = std::max<uint64_t>((uint64_t)(span * density * 1.02) + 64ull * nwin + 4096, buf.bytes / 8);
        TRY(buf.ensure(cap * 8));
        int rc = wheel_run(f.wheel, nwin, lo.data(), hi.data(), buf.as<uint64_t>(), cap, off.data());
        if (rc == PCUBE_OVERFLOW) {
            cap = off[nwin];
            TRY(buf.ensure(cap * 8));
            rc = wheel_run(f.wheel, nwin, lo.data(), hi.data(), buf.as<uint64_t>(), cap, off.data());
        }
        return rc;
    };
    TRY(windows_of(*fu, ulo, uhi, e.lus, offu));
    TRY(windows_of(*fv, vlo, vhi, e.lvs, offv));
    // subtile bounds inside each window
    const uint64_t nq = 4 * ntot;
    std::vector<uint64_t> qb(nq), ql(nq), qk(nq), qo(nq);
    for (uint32_t w = 0; w < nwin; ++w)
        for (uint64_t t = tfirst[w]; t < tfirst[w + 1]; ++t) {
            for (int k = 0; k < 2; ++k) {
                qb[4 * t + k] = offu[w]; ql[4 * t + k] = offu[w + 1] - offu[w]; qk[4 * t + k] = tiles[4 * t + 2 + k];
                qb[4 * t + 2 + k] = offv[w]; ql[4 * t + 2 + k] = offv[w + 1] - offv[w]; qk[4 * t + 2 + k] = tiles[4 * t + k];
            }
        }
    TRY(e.q_begin.ensure(nq * 8)); TRY(e.q_len.ensure(nq * 8)); TRY(e.q_key.ensure(nq * 8)); TRY(e.q_out.ensure(nq * 8));
    // u queries search the u array, v queries the v array: two launches over interleaved halves
    std::vector<uint64_t> ub(2 * ntot), ul(2 * ntot), uk(2 * ntot), vb(2 * ntot), vl(2 * ntot), vk(2 * ntot);
    for (uint64_t t = 0; t < ntot; ++t)
        for (int k = 0; k < 2; ++k) {
            ub[2 * t + k] = qb[4 * t + k]; ul[2 * t + k] = ql[4 * t + k]; uk[2 * t + k] = qk[4 * t + k];
            vb[2 * t + k] = qb[4 * t + 2 + k]; vl[2 * t + k] = ql[4 * t + 2 + k]; vk[2 * t + k] = qk[4 * t + 2 + k];
        }
    HIP_TRY(hipMemcpyAsync(e.q_begin.p, ub.data(), 2 * ntot * 8, hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipMemcpyAsync(e.q_len.p, ul.data(), 2 * ntot * 8, hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipMemcpyAsync(e.q_key.p, uk.data(), 2 * ntot * 8, hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipMemcpyAsync(e.q_begin.as<uint64_t>() + 2 * ntot, vb.data(), 2 * ntot * 8, hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipMemcpyAsync(e.q_len.as<uint64_t>() + 2 * ntot, vl.data(), 2 * ntot * 8, hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipMemcpyAsync(e.q_key.as<uint64_t>() + 2 * ntot, vk.data(), 2 * ntot * 8, hipMemcpyHostToDevice, g.stream));
    hipLaunchKernelGGL(k_searchsorted_seg, dim3(blocks_for(2 * ntot, 256)), dim3(256), 0, g.stream,
                       (const uint64_t*)e.lus.as<uint64_t>(), (const uint64_t*)e.q_begin.as<uint64_t>(),
                       (const uint64_t*)e.q_len.as<uint64_t>(), (const uint64_t*)e.q_key.as<uint64_t>(), 2 * ntot,
                       e.q_out.as<uint64_t>());
    hipLaunchKernelGGL(k_searchsorted_seg, dim3(blocks_for(2 * ntot, 256)), dim3(256), 0, g.stream,
                       (const uint64_t*)e.lvs.as<uint64_t>(), (const uint64_t*)(e.q_begin.as<uint64_t>() + 2 * ntot),
                       (const uint64_t*)(e.q_len.as<uint64_t>() + 2 * ntot),
                       (const uint64_t*)(e.q_key.as<uint64_t>() + 2 * ntot), 2 * ntot, e.q_out.as<uint64_t>() + 2 * ntot);
    // the subtile bounds travel back with the counters; descriptors need only the window sizes
    std::vector<uint64_t> pu(2 * ntot), pv(2 * ntot);
    HIP_TRY(hipMemcpyAsync(pu.data(), e.q_out.p, 2 * ntot * 8, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipMemcpyAsync(pv.data(), e.q_out.as<uint64_t>() + 2 * ntot, 2 * ntot * 8, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipEventRecord(ev[1], g.stream));
    // descriptors
    std::vector<BatchWindow> desc(nwin);
    uint64_t rows = 0, mask_words = 0, mask_waves = 0, rho_items = 0, chunks = 0, ncounters = 0;
    int geometry_rc = PCUBE_OK;
    for (uint32_t w = 0; w < nwin; ++w) {
        const BatchConfig& c = e.args.cfg[win_cfg[w]];
        BatchWindow d;
        memset(&d, 0, sizeof d);
        const uint64_t nu = offu[w + 1] - offu[w], nv = offv[w + 1] - offv[w];
        if (nu >= (1ull << 32) - 64 || nv >= (1ull << 32)) return fail(PCUBE_INVALID, "window %u too large", w);
        sizes[2 * w] = nu;
        sizes[2 * w + 1] = nv;
        d.nu = (uint32_t)nu; d.nv = (uint32_t)nv;
        d.words = (uint32_t)((nu + 63) / 64);
        d.ntiles = win_tiles[w];
        d.cfg = win_cfg[w];
        d.mode = mode[w];
        d.v_first = vlo[w];
        d.tile_width = tile_width[w];
        d.tile_base = tfirst[w];
        d.row_begin = rows;
        d.mask_begin = mask_waves;
        d.rho_begin = rho_items;
        d.chunk_begin = chunks;
        // pointers are filled after the buffers are sized; keep offsets for now
        d.us = (const uint64_t*)(uintptr_t)(offu[w] * 8);
        d.vs = (const uint64_t*)(uintptr_t)(offv[w] * 8);
        d.counters = (uint64_t*)(uintptr_t)(ncounters * 8);
        if (d.mode == 0 && nu) {
            // mask items: (run of kMaskRun words, filter) pairs for a transposed build, else
            // (chunk, word) pairs;
            // residue items: one per u (a 16-byte row of kCachedResidues residues)
            d.masks = (uint64_t*)(uintptr_t)(mask_words * 8);
            d.rho = (uint16_t*)(uintptr_t)(rho_items * 2 * kCachedResidues);
            mask_words += (uint64_t)c.mask_rows * d.words;
            mask_waves += c.transpose ? (uint64_t)c.nbits * ((d.words + kMaskRun - 1) / kMaskRun)
                                      : (uint64_t)c.chunk_base[c.nbits] * d.words;
            rho_items += nu;
        }
        rows += nv;
        chunks += (nv + rows_per_chunk - 1) / rows_per_chunk;
        ncounters += (uint64_t)win_tiles[w] * (c.ng + 3);
        desc[w] = d;
    }
    if (chunks >= (1ull << 32)) return fail(PCUBE_INVALID, "too many row chunks");
    TRY(e.win.ensure(nwin * sizeof(BatchWindow)));
    TRY(e.rows.ensure(std::max<uint64_t>(1, rows) * sizeof(Row)));
    TRY(e.counters.ensure(std::max<uint64_t>(1, ncounters) * 8));
    TRY(e.masks.ensure(std::max<uint64_t>(8, mask_words * 8)));
    TRY(e.rho.ensure(std::max<uint64_t>(16, rho_items * 2 * kCachedResidues)));
    for (auto& d : desc) {
        d.us = (const uint64_t*)((char*)e.lus.p + (uintptr_t)d.us);
        d.vs = (const uint64_t*)((char*)e.lvs.p + (uintptr_t)d.vs);
        d.counters = (uint64_t*)((char*)e.counters.p + (uintptr_t)d.counters);
        d.masks = (uint64_t*)((char*)e.masks.p + (uintptr_t)d.masks);
        d.rho = (uint16_t*)((char*)e.rho.p + (uintptr_t)d.rho);
    }
    BatchTotals tot;
    memset(&tot, 0, sizeof tot);
    tot.rows = rows; tot.mask_waves = mask_waves; tot.rho_items = rho_items; tot.chunks = chunks;
    tot.counters = ncounters; tot.nwin = nwin; tot.rows_per_chunk = rows_per_chunk;
    HIP_TRY(hipMemcpyAsync(e.win.p, desc.data(), nwin * sizeof(BatchWindow), hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipMemcpyAsync(e.totals.p, &tot, sizeof tot, hipMemcpyHostToDevice, g.stream));
    e.args.win = e.win.as<BatchWindow>();
    e.args.totals = e.totals.as<BatchTotals>();
    e.args.rows = e.rows.as<Row>();
    e.args.list = e.list.as<uint64_t>();
    e.args.list_count = e.list_count.as<uint64_t>();
    e.args.counters = e.counters.as<uint64_t>();
    e.args.error = e.error.as<uint32_t>();
    e.args.cursor = e.cursor.as<uint32_t>();
    HIP_TRY(hipEventRecord(ev[2], g.stream));
    TRY(batch_launch(e));
    HIP_TRY(hipEventRecord(ev[3], g.stream));
    std::vector<uint64_t> host(std::max<uint64_t>(1, ncounters));
    uint32_t err = 0;
    uint64_t listed = 0;
    HIP_TRY(hipMemcpyAsync(host.data(), e.counters.p, ncounters * 8, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipMemcpyAsync(&err, e.error.p, 4, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipMemcpyAsync(&listed, e.list_count.p, 8, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    memcpy(counters, host.data(), ncounters * 8);
    for (uint64_t t = 0; t < ntot; ++t) {
        for (int k = 0; k < 2; ++k) {
            bounds[4 * t + k] = pu[2 * t + k];
            bounds[4 * t + 2 + k] = pv[2 * t + k];
        }
        const uint64_t du = bounds[4 * t + 1] - bounds[4 * t], dv = bounds[4 * t + 3] - bounds[4 * t + 2];
        if (du * dv > 50000000ull && geometry_rc == PCUBE_OK)
            geometry_rc = fail(PCUBE_OVERFLOW, "per-subtile candidate opportunity limit exceeded");
    }
    if (seconds) {
        float a = 0, b = 0;
        HIP_TRY(hipEventElapsedTime(&a, ev[0], ev[1]));
        HIP_TRY(hipEventElapsedTime(&b, ev[2], ev[3]));
        seconds[0] = a / 1000.0;
        seconds[1] = 0.0;
        seconds[2] = b / 1000.0;
    }
    for (auto& x : ev) (void)hipEventDestroy(x);
    if (graph_launches) *graph_launches = e.graph_launches;
    if (geometry_rc) return geometry_rc;
    // status and values, subtile by subtile (each window's counter stride is its configuration's)
    if (err & 1u) return fail(PCUBE_INVALID, "a v coordinate lies outside its window's subtiles");
    std::vector<uint64_t> tile_ng(ntot), tile_counter(ntot);
    {
        uint64_t off = 0;
        for (uint32_t w = 0; w < nwin; ++w) {
            const uint32_t ng = e.args.cfg[win_cfg[w]].ng;
            for (uint64_t t = tfirst[w]; t < tfirst[w + 1]; ++t) {
                tile_ng[t] = ng;
                tile_counter[t] = off;
                off += ng + 3;
            }
        }
    }
    for (uint64_t t = 0; t < ntot; ++t)
        if (host[tile_counter[t] + tile_ng[t] + 2]) return fail(PCUBE_UNDERFLOW, "unsigned underflow flagged");
    if (err & 2u) return fail(PCUBE_RANGE, "an emitted value lies outside (lo, hi]");
    std::vector<uint64_t> voff(ntot + 1, 0);
    int status = PCUBE_OK;
    for (uint64_t t = 0; t < ntot; ++t) {
        const uint64_t emitted = host[tile_counter[t] + tile_ng[t] + 1];
        if (emitted > e.cap && status == PCUBE_OK) status = fail(PCUBE_OVERFLOW, "output overflow in subtile %llu", (unsigned long long)t);
        voff[t + 1] = voff[t] + std::min(emitted, e.cap);
    }
    if (status) return status;
    if (voff[ntot] > e.list_cap) return fail(PCUBE_OVERFLOW, "survivor list overflow");
    if (listed != voff[ntot])
        return fail(PCUBE_MISMATCH, "survivor list holds %llu values, counters say %llu", (unsigned long long)listed,
                    (unsigned long long)voff[ntot]);
    if (listed > values_cap) return fail(PCUBE_OVERFLOW, "host value buffer too small");
    if (listed) {
        std::vector<uint64_t> triples(listed * 3);
        HIP_TRY(hipMemcpyAsync(triples.data(), e.list.p, listed * 24, hipMemcpyDeviceToHost, g.stream));
        HIP_TRY(hipStreamSynchronize(g.stream));
        std::vector<uint64_t> cur(voff.begin(), voff.end() - 1);
        for (uint64_t i = 0; i < listed; ++i) {
            const uint64_t t = triples[3 * i];
            if (t >= ntot || cur[t] >= voff[t + 1]) return fail(PCUBE_INVALID, "corrupt survivor list");
            values[2 * cur[t]] = triples[3 * i + 1];
            values[2 * cur[t] + 1] = triples[3 * i + 2];
            ++cur[t];
        }
    }
    if (seconds) seconds[3] = now_seconds() - t_start;
    return PCUBE_OK;
}

}  // extern "C"

namespace {

typedef unsigned __int128 u128_t;

struct TextOut {
    char* p;
    uint64_t n = 0, cap;
    bool ok = true;
    void put(const char* s, size_t len) {
        if (!ok || n + len > cap) {
            ok = false;
            return;
        }
        memcpy(p + n, s, len);
        n += len;
    }
    void put(const char* s) { put(s, strlen(s)); }
    void digits(uint64_t v) {
        char buf[24];
        int i = 24;
        do {
            buf[--i] = char('0' + v % 10);
            v /= 10;
        } while (v);
        put(buf + i, (size_t)(24 - i));
    }
    // a JSON integer, or its decimal string from 2^53 on (the encoder's rule)
    void integer(uint64_t v) {
        const bool quoted = v >= (1ull << 53);
        if (quoted) put("\"", 1);
        digits(v);
        if (quoted) put("\"", 1);
    }
};

// The primes up to p by trial division, and whether p is a prime = 1 mod 3
// (the predicate's precondition, pseudocube/oracle.py:14-16).
bool predicate_primes(uint32_t p, std::vector<uint32_t>& primes) {
    primes.clear();
    for (uint32_t q = 2; q <= p; ++q) {
        bool prime = true;
        for (uint32_t d = 2; d * d <= q; ++d)
            if (q % d == 0) { prime = false; break; }
        if (prime) primes.push_back(q);
    }
    return !primes.empty() && primes.back() == p && p % 3 == 1 && p >= 7;
}

// local3(x, p) of pseudocube/oracle.py:36-41: x > 0, x mod 9 in {1, 8}, no prime
// q <= p divides x, and x^((q-1)/3) = 1 mod q for every q = 1 mod 3.  Residues of x
// come from its 64-bit halves: x mod q = ((hi mod q)(2^64 mod q) + lo mod q) mod q,
// with 2^64 mod q in wrap[k] (wrap empty: computed here).
bool local_predicate(u128_t x, const std::vector<uint32_t>& primes, const std::vector<uint32_t>& wrap = {}) {
    const uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
    const uint64_t m9 = ((hi % 9) * 7 + lo % 9) % 9;   // 2^64 = 7 mod 9
    bool ok = x > 0 && (m9 == 1 || m9 == 8);
    for (size_t k = 0; ok && k < primes.size(); ++k) {
        const uint64_t q = primes[k];
        const uint64_t r64 = k < wrap.size() ? wrap[k] : (UINT64_MAX % q + 1) % q;
        const uint64_t r = ((hi % q) * r64 + lo % q) % q;
        if (r == 0) return false;
        if (q % 3 == 1) {
            uint64_t e = (q - 1) / 3, base = r, acc = 1;
            while (e) {
                if (e & 1) acc = acc * base % q;
                base = base * base % q;
                e >>= 1;
            }
            if (acc != 1) return false;
        }
    }
    return ok;
}

// floor(x^(1/3)) by bisection on [0, 6981463658332): 6981463658331^3 < 2^128 <=
// 6981463658332^3, so every probe's cube fits 128 bits (review/verify_cube.py:13-17).
uint64_t cube_root_floor(u128_t x) {
    uint64_t lo = 0, hi = 6981463658332ull;
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) >> 1;
        if ((u128_t)mid * mid * mid <= x) lo = mid; else hi = mid;
    }
    return lo;
}

// a^-1 mod m for gcd(a, m) = 1 (extended Euclid), or 0 when not invertible.
uint64_t inverse_mod(uint64_t a, uint64_t m) {
    int64_t t = 0, nt = 1;
    uint64_t r = m, nr = a % m;
    while (nr) {
        const uint64_t q = r / nr;
        const int64_t tt = t - (int64_t)q * nt;
        t = nt; nt = tt;
        const uint64_t rr = r - q * nr;
        r = nr; nr = rr;
    }
    if (r != 1) return 0;
    return t < 0 ? (uint64_t)(t + (int64_t)m) : (uint64_t)t;
}

void decimal_u128(u128_t x, char* buf, int& len) {
    char tmp[48];
    int i = 48;
    do {
        tmp[--i] = char('0' + (int)(x % 10));
        x /= 10;
    } while (x);
    len = 48 - i;
    memcpy(buf, tmp + i, (size_t)len);
}

}  // namespace

extern "C" {

int pcube_host_verify(const uint64_t* limbs, uint64_t n, uint32_t p, uint8_t* local_ok, uint8_t* is_cube,
                      uint64_t* cube_root) {
    // host code only: the exact predicate of pseudocube/oracle.py:36-41 and the
    // cube root of review/verify_cube.py:13-17, in unsigned __int128
    if (p < 7 || p > 100000) return fail(PCUBE_INVALID, "prime bound out of range");
    std::vector<uint32_t> primes;
    if (!predicate_primes(p, primes)) return fail(PCUBE_INVALID, "p must be a prime = 1 mod 3");
    for (uint64_t i = 0; i < n; ++i) {
        const u128_t x = ((u128_t)limbs[2 * i + 1] << 64) | limbs[2 * i];
        const uint64_t r = cube_root_floor(x);
        local_ok[i] = local_predicate(x, primes) ? 1 : 0;
        is_cube[i] = ((u128_t)r * r * r == x) ? 1 : 0;
        cube_root[i] = r;
    }
    return PCUBE_OK;
}

int pcube_batch_values(uint64_t nvalues, const uint64_t* limbs, uint32_t nsub, const uint64_t* counts,
                       const uint64_t* firsts, const uint64_t* ends, const uint32_t* primes_of, uint64_t A,
                       uint64_t B, uint64_t lo0, uint64_t lo1, uint64_t hi0, uint64_t hi1, uint8_t* cube,
                       uint64_t* roots, char* text, uint64_t text_cap, uint64_t* text_offsets, uint32_t* failure) {
    // host code only: the checks of pseudocube/batch_experiment.py:121-167 on every
    // emitted value of many subtiles at once, subtiles spread over a few threads
    *failure = 0;
    if (B == 0) return fail(PCUBE_INVALID, "B must be positive");
    std::vector<uint64_t> at(nsub + 1, 0);
    for (uint32_t s = 0; s < nsub; ++s) at[s + 1] = at[s] + counts[s];
    if (at[nsub] != nvalues) return fail(PCUBE_INVALID, "subtile counts sum to %llu, %llu values given",
                                         (unsigned long long)at[nsub], (unsigned long long)nvalues);
    const uint64_t ainv = inverse_mod(A % B, B);
    if (!ainv) return fail(PCUBE_INVALID, "A is not invertible mod B");
    const u128_t lo = ((u128_t)lo1 << 64) | lo0, hi = ((u128_t)hi1 << 64) | hi0;
    // the prime lists (and 2^64 mod q) of the bounds in use
    std::map<uint32_t, std::pair<std::vector<uint32_t>, std::vector<uint32_t>>> lists;
    for (uint32_t s = 0; s < nsub; ++s) {
        if (!counts[s] || lists.count(primes_of[s])) continue;
        std::vector<uint32_t> primes, wrap;
        if (!predicate_primes(primes_of[s], primes)) return fail(PCUBE_INVALID, "p must be a prime = 1 mod 3");
        for (uint32_t q : primes) wrap.push_back((uint32_t)((UINT64_MAX % q + 1) % q));
        lists.emplace(primes_of[s], std::make_pair(std::move(primes), std::move(wrap)));
    }
    std::vector<std::string> texts(nsub);
    std::vector<uint32_t> failed(nsub, 0);
    auto work = [&](uint32_t s) {
        const uint64_t n = counts[s], a = at[s];
        if (!n) return;
        const auto& pl = lists.at(primes_of[s]);
        std::vector<u128_t> xs(n);
        for (uint64_t i = 0; i < n; ++i) xs[i] = ((u128_t)limbs[2 * (a + i) + 1] << 64) | limbs[2 * (a + i)];
        std::sort(xs.begin(), xs.end());
        std::string& out = texts[s];
        out.reserve(n * 44 + 3);
        out.push_back('[');
        for (uint64_t i = 0; i < n; ++i) {
            const u128_t x = xs[i];
            if (i && xs[i - 1] == x) { failed[s] = 1; return; }
            if (!(lo < x && x <= hi)) { failed[s] = 2; return; }
            // canonical v = -x * A^-1 mod B must lie in the subtile (pseudocube/focus.py:116-118)
            const uint64_t xm = (uint64_t)(x % B);
            const uint64_t v = (uint64_t)((u128_t)((B - xm) % B) * ainv % B);
            if (v < firsts[s] || v >= ends[s]) { failed[s] = 4; return; }
            if (!local_predicate(x, pl.first, pl.second)) { failed[s] = 3; return; }
            const uint64_t r = cube_root_floor(x);
            cube[a + i] = ((u128_t)r * r * r == x) ? 1 : 0;
            roots[a + i] = r;
            char buf[48];
            int len = 0;
            decimal_u128(x, buf, len);
            if (i) out.push_back(',');
            out.push_back('"');
            out.append(buf, (size_t)len);
            out.push_back('"');
        }
        out.append("]\n");
    };
    const uint32_t workers = (uint32_t)std::min<uint64_t>(
        {8, std::max(1u, std::thread::hardware_concurrency()), nvalues / 2048 + 1});
    if (workers <= 1) {
        for (uint32_t s = 0; s < nsub; ++s) work(s);
    } else {
        // subtiles dealt out in value-balanced runs
        std::vector<std::thread> pool;
        std::vector<uint32_t> cut(workers + 1, nsub);
        cut[0] = 0;
        for (uint32_t w = 1, s = 0; w < workers; ++w) {
            while (s < nsub && at[s] < nvalues * w / workers) ++s;
            cut[w] = s;
        }
        for (uint32_t w = 0; w < workers; ++w)
            pool.emplace_back([&, w] {
                for (uint32_t s = cut[w]; s < cut[w + 1]; ++s) work(s);
            });
        for (auto& th : pool) th.join();
    }
    static const char* const kFailure[] = {"", "duplicate canonical output", "value outside (lo, hi]",
                                           "GPU/CPU disagreement", "output assigned to wrong subtile"};
    for (uint32_t s = 0; s < nsub; ++s)
        if (failed[s]) {
            *failure = failed[s];
            return fail(PCUBE_MISMATCH, "%s (subtile %u of the pass)", kFailure[failed[s]], s);
        }
    TextOut out{text, 0, text_cap};
    text_offsets[0] = 0;
    for (uint32_t s = 0; s < nsub; ++s) {
        out.put(texts[s].data(), texts[s].size());
        text_offsets[s + 1] = out.n;
    }
    if (!out.ok) return fail(PCUBE_OVERFLOW, "value text exceeds %llu bytes", (unsigned long long)text_cap);
    return PCUBE_OK;
}

int pcube_lds_probe(uint32_t pattern, uint32_t rounds, double* seconds, double* bytes) {
    Entry lock;
    if (pattern > 2 || !rounds) return fail(PCUBE_INVALID, "probe pattern 0..2 and at least one round");
    const uint32_t grid = (uint32_t)g.compute_units;
    void* out = nullptr;
    TRY(dev_alloc((uint64_t)grid * kProbeThreads * 4, &out));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    hipLaunchKernelGGL(k_lds_probe, dim3(grid), dim3(kProbeThreads), 0, g.stream, pattern, rounds, (uint32_t*)out);
    HIP_TRY(hipEventRecord(e0, g.stream));
    hipLaunchKernelGGL(k_lds_probe, dim3(grid), dim3(kProbeThreads), 0, g.stream, pattern, rounds, (uint32_t*)out);
    HIP_TRY(hipEventRecord(e1, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    if (seconds) *seconds = ms / 1000.0;
    if (bytes) *bytes = (double)grid * kProbeThreads * rounds * 16 * 16;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dev_free(out);
    return PCUBE_OK;
}

int pcube_valu_probe(uint32_t rounds, double* seconds, double* lane_ops) {
    Entry lock;
    if (!rounds) return fail(PCUBE_INVALID, "at least one probe round");
    const uint32_t grid = (uint32_t)g.compute_units * 8;   // 32 waves per CU
    void* out = nullptr;
    TRY(dev_alloc((uint64_t)grid * 256 * 4, &out));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    hipLaunchKernelGGL(k_valu_probe, dim3(grid), dim3(256), 0, g.stream, rounds, (uint32_t*)out);
    HIP_TRY(hipEventRecord(e0, g.stream));
    hipLaunchKernelGGL(k_valu_probe, dim3(grid), dim3(256), 0, g.stream, rounds, (uint32_t*)out);
    HIP_TRY(hipEventRecord(e1, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    if (seconds) *seconds = ms / 1000.0;
    if (lane_ops) *lane_ops = (double)grid * 256 * rounds * kValuProbeOps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dev_free(out);
    return PCUBE_OK;
}

int pcube_cube_count(uint64_t first, uint64_t last, uint32_t p, uint64_t* count, uint32_t* good, uint64_t good_words,
                     double* seconds) {
    Entry lock;
    if (first < 1 || last < first || last > UINT32_MAX || p < 2 || p > 2000)
        return fail(PCUBE_INVALID, "cube count needs 1 <= first <= last <= 2^32 - 1 and 2 <= p <= 2000");
    if (g.device < 0) return fail(PCUBE_INVALID, "pcube_init has not been called");
    // the odd primes up to p
    std::vector<uint32_t> primes;
    for (uint32_t q = 3; q <= p; q += 2) {
        bool prime = true;
        for (uint32_t d = 3; d * d <= q && prime; d += 2) prime = q % d != 0;
        if (prime) primes.push_back(q);
    }
    const uint64_t r0 = first | 1u;
    const uint64_t nodd = r0 > last ? 0 : (last - r0) / 2 + 1;
    const uint64_t need_words = (nodd + 31) / 32;
    if (good && good_words < need_words) return fail(PCUBE_INVALID, "root bitset needs %llu words",
                                                     (unsigned long long)need_words);
    void *dprimes = nullptr, *dcount = nullptr, *dgood = nullptr;
    TRY(dev_alloc(std::max<uint64_t>(4, primes.size() * 4), &dprimes));
    int rc = dev_alloc(8, &dcount);
    if (!rc && good && need_words) rc = dev_alloc(need_words * 4, &dgood);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc) {
        if (!primes.empty())
            HIP_TRY(hipMemcpyAsync(dprimes, primes.data(), primes.size() * 4, hipMemcpyHostToDevice, g.stream));
        rc = fill64((uint64_t*)dcount, 1, 0);
    }
    if (!rc && nodd) {
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        HIP_TRY(hipEventRecord(e0, g.stream));
        hipLaunchKernelGGL(k_cube_count, dim3((uint32_t)((nodd + kCubeSegment - 1) / kCubeSegment)),
                           dim3(kCubeThreads), 0, g.stream, r0, nodd, (const uint32_t*)dprimes,
                           (uint32_t)primes.size(), (unsigned long long*)dcount, (uint32_t*)dgood);
        HIP_TRY(hipEventRecord(e1, g.stream));
    }
    uint64_t total = 0;
    if (!rc) {
        HIP_TRY(hipMemcpyAsync(&total, dcount, 8, hipMemcpyDeviceToHost, g.stream));
        if (dgood) HIP_TRY(hipMemcpyAsync(good, dgood, need_words * 4, hipMemcpyDeviceToHost, g.stream));
        HIP_TRY(hipStreamSynchronize(g.stream));
        float ms = 0;
        if (e0) HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        if (seconds) *seconds = ms / 1000.0;
        if (count) *count = total;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dev_free(dprimes);
    if (dcount) dev_free(dcount);
    if (dgood) dev_free(dgood);
    return rc;
}

int pcube_cube_count_noalloc(uint64_t first, uint64_t last, uint32_t p, uint64_t* count, uint32_t* good,
                             uint64_t good_words) {
    Entry lock;
    if (first < 1 || last < first || last > UINT32_MAX || p < 2 || p > 2000)
        return fail(PCUBE_INVALID, "cube count needs 1 <= first <= last <= 2^32 - 1 and 2 <= p <= 2000");
    if (g.device < 0) return fail(PCUBE_INVALID, "pcube_init has not been called");
    uint8_t* base = (uint8_t*)g.scratch_at[g.device];
    if (!base) return fail(PCUBE_DEVICE, "the static scratch array has no address on device %d", g.device);
    // the odd primes up to p, on the host stack (no allocation)
    uint32_t primes[kNoAllocMaxPrimes];
    uint32_t nprimes = 0;
    for (uint32_t q = 3; q <= p; q += 2) {
        bool prime = true;
        for (uint32_t d = 3; d * d <= q && prime; d += 2) prime = q % d != 0;
        if (prime) primes[nprimes++] = q;   // p <= 2000 keeps nprimes <= 302
    }
    const uint64_t r0 = first | 1u;
    const uint64_t nodd = r0 > last ? 0 : (last - r0) / 2 + 1;
    const uint64_t need_words = (nodd + 31) / 32;
    if (good && good_words < need_words)
        return fail(PCUBE_INVALID, "root bitset needs %llu words", (unsigned long long)need_words);
    if (good && need_words > kNoAllocGoodWords)
        return fail(PCUBE_INVALID, "a listed range of %llu odd roots exceeds the static scratch maximum of %u",
                    (unsigned long long)nodd, kNoAllocGoodWords * 32);
    unsigned long long* dcount = (unsigned long long*)base;
    uint32_t* dprimes = (uint32_t*)(base + kNoAllocPrimes);
    uint32_t* dgood = good && need_words ? (uint32_t*)(base + kNoAllocGood) : nullptr;
    // the array keeps the previous call's contents: the counter starts at zero,
    // the prime table and every root word the kernel reports are rewritten
    if (nprimes) HIP_TRY(hipMemcpyAsync(dprimes, primes, nprimes * 4, hipMemcpyHostToDevice, g.stream));
    TRY(fill64((uint64_t*)dcount, 1, 0));
    if (nodd)
        hipLaunchKernelGGL(k_cube_count, dim3((uint32_t)((nodd + kCubeSegment - 1) / kCubeSegment)),
                           dim3(kCubeThreads), 0, g.stream, r0, nodd, (const uint32_t*)dprimes, nprimes, dcount,
                           dgood);
    HIP_TRY(hipGetLastError());
    uint64_t total = 0;
    HIP_TRY(hipMemcpyAsync(&total, dcount, 8, hipMemcpyDeviceToHost, g.stream));
    if (dgood) HIP_TRY(hipMemcpyAsync(good, dgood, need_words * 4, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    if (count) *count = total;
    return PCUBE_OK;
}

int pcube_copy_probe(uint64_t bytes, double* seconds) {
    Entry lock;
    const uint64_t n = bytes / 8;
    if (!n) return fail(PCUBE_INVALID, "probe size must be at least 8 bytes");
    void *src = nullptr, *dst = nullptr;
    TRY(dev_alloc(n * 8, &src));
    int rc = dev_alloc(n * 8, &dst);
    if (rc) { dev_free(src); return rc; }
    TRY(fill64((uint64_t*)src, n, 0x9e3779b97f4a7c15ull));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, g.stream));
    hipLaunchKernelGGL(k_copy_probe, dim3(8192), dim3(256), 0, g.stream, (const uint64_t*)src, (uint64_t*)dst, n);
    HIP_TRY(hipEventRecord(e1, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    if (seconds) *seconds = ms / 1000.0;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dev_free(src);
    dev_free(dst);
    return PCUBE_OK;
}

// Text writer for pcube_scope_records: fails once cap would be exceeded.

int pcube_scope_records(uint32_t ntiles, const uint64_t* bounds, const uint64_t* counters, uint32_t stride,
                        uint32_t ng, const uint64_t* local, const char* sha, char* out, uint64_t cap,
                        uint64_t* length) {
    if (!bounds && ntiles) return fail(PCUBE_INVALID, "null subtile bounds");
    if ((!counters || !sha) && ntiles) return fail(PCUBE_INVALID, "null counters or digests");
    if (stride < ng + 1 || !out || !length) return fail(PCUBE_INVALID, "invalid record layout");
    TextOut t{out, 0, cap};
    t.put("[", 1);
    for (uint32_t k = 0; k < ntiles; ++k) {
        const uint64_t* c = counters + (uint64_t)k * stride;
        t.put(k ? ",{\"end\":\"" : "{\"end\":\"");
        t.digits(bounds[2 * k + 1]);
        t.put("\",\"first\":\"");
        t.digits(bounds[2 * k]);
        t.put("\",\"group_survivors\":[");
        for (uint32_t g = 0; g < ng; ++g) {
            if (g) t.put(",", 1);
            t.integer(c[1 + g]);
        }
        t.put("],\"local_survivors\":");
        t.integer(local ? local[k] : 0);
        t.put(",\"output_sha256\":\"");
        t.put(sha + 64 * (uint64_t)k, 64);
        t.put("\",\"valid_focused_candidates\":");
        t.integer(c[0]);
        t.put("}", 1);
    }
    t.put("]\n", 2);
    if (!t.ok) return fail(PCUBE_OVERFLOW, "scope records exceed %llu bytes", (unsigned long long)cap);
    *length = t.n;
    return PCUBE_OK;
}

int pcube_scope_records_many(uint32_t nbatches, const uint32_t* per_batch, const uint64_t* bounds,
                             const uint64_t* counters, uint32_t stride, uint32_t ng, const uint64_t* local,
                             const char* sha, char* out, uint64_t cap, uint64_t* offsets) {
    if ((!per_batch && nbatches) || !offsets || !out) return fail(PCUBE_INVALID, "null batch layout");
    offsets[0] = 0;
    uint64_t k = 0;
    for (uint32_t b = 0; b < nbatches; ++b) {
        uint64_t length = 0;
        const uint32_t nt = per_batch[b];
        const int rc = pcube_scope_records(nt, bounds ? bounds + 2 * k : nullptr,
                                           counters ? counters + k * stride : nullptr, stride, ng,
                                           local ? local + k : nullptr, sha ? sha + 64 * k : nullptr,
                                           out + offsets[b], cap - offsets[b], &length);
        if (rc) return rc;
        offsets[b + 1] = offsets[b] + length;
        k += nt;
    }
    return PCUBE_OK;
}

int pcube_floor_div(const uint64_t* in, uint64_t* out, uint64_t count) {
    Entry lock;
    if (!count) return PCUBE_OK;
    std::vector<uint64_t> image(5 * count);
    for (uint64_t i = 0; i < count; ++i) {
        const uint64_t B = in[3 * i + 2];
        if (B == 0) return fail(PCUBE_INVALID, "division by zero in case %llu", (unsigned long long)i);
        const unsigned __int128 mu = ~(unsigned __int128)0 / B;
        image[5 * i] = in[3 * i];
        image[5 * i + 1] = in[3 * i + 1];
        image[5 * i + 2] = B;
        image[5 * i + 3] = (uint64_t)mu;
        image[5 * i + 4] = (uint64_t)(mu >> 64);
    }
    void *din = nullptr, *dout = nullptr;
    TRY(dev_alloc(count * 40, &din));
    int rc = dev_alloc(count * 16, &dout);
    if (rc) { dev_free(din); return rc; }
    HIP_TRY(hipMemcpyAsync(din, image.data(), count * 40, hipMemcpyHostToDevice, g.stream));
    hipLaunchKernelGGL(k_floor_div, dim3(blocks_for(count, 256)), dim3(256), 0, g.stream, (const uint64_t*)din,
                       (uint64_t*)dout, count);
    HIP_TRY(hipMemcpyAsync(out, dout, count * 16, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    dev_free(din);
    dev_free(dout);
    return PCUBE_OK;
}

int pcube_bit_transpose(const uint64_t* in, uint64_t* out, uint64_t count) {
    Entry lock;
    if (!count) return PCUBE_OK;
    if (count > (1ull << 24)) return fail(PCUBE_INVALID, "at most 2^24 matrices per call");
    const uint64_t bytes = count * kWave * 8;
    void *din = nullptr, *dout = nullptr;
    TRY(dev_alloc(bytes, &din));
    int rc = dev_alloc(bytes, &dout);
    if (rc) { dev_free(din); return rc; }
    HIP_TRY(hipMemcpyAsync(din, in, bytes, hipMemcpyHostToDevice, g.stream));
    hipLaunchKernelGGL(k_bit_transpose, dim3(blocks_for(count * kWave, 256)), dim3(256), 0, g.stream,
                       (const uint64_t*)din, (uint64_t*)dout, count);
    HIP_TRY(hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    dev_free(din);
    dev_free(dout);
    return PCUBE_OK;
}

int pcube_arithmetic(const uint64_t* in, uint64_t* out, uint64_t count) {
    Entry lock;
    if (!count) return PCUBE_OK;
    void *din = nullptr, *dout = nullptr;
    TRY(dev_alloc(count * 32, &din));
    int rc = dev_alloc(count * 24, &dout);
    if (rc) { dev_free(din); return rc; }
    HIP_TRY(hipMemcpyAsync(din, in, count * 32, hipMemcpyHostToDevice, g.stream));
    hipLaunchKernelGGL(k_arithmetic, dim3(blocks_for(count, 256)), dim3(256), 0, g.stream, (const uint64_t*)din,
                       (uint64_t*)dout, count);
    HIP_TRY(hipMemcpyAsync(out, dout, count * 24, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    dev_free(din);
    dev_free(dout);
    return PCUBE_OK;
}

}  // extern "C"
