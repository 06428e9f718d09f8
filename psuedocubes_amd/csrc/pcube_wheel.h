#pragma once
// CRT wheel generator for focus windows (gfx950).
//
// A focus of modulus M is a product of rings q_1 < q_2 < ... with allowed
// residue sets R_i.  Its coordinates in a window [low, high) are generated as
// y = t*M1 + r, where M1 is the product of an inner prefix of rings and r runs
// over the sorted inner residues R1 (the CRT combinations of the inner R_i);
// each candidate is then tested against the outer rings with
// (r mod q + t*M1 mod q) mod q in R_q.  Candidates are visited in (t, r) order,
// so each window's output is ascending with no sort.  The cost is proportional
// to the window width times |R1| / M1, independent of the wheel's size, so the
// inner prefix takes every ring it can (M1 < 2^40, |R1| <= 2^27 entries of 8
// bytes): each ring moved inside divides the candidates read and tested by
// q / |R_q|, about 3 for the cubic-residue rings.
//
// Work is cut into segments of kWheelSegment consecutive candidates of one t;
// a block handles one segment, each of its four waves 1024 candidates.  One
// test pass writes each wave's survivors (their inner residues r) into a slot
// of slot_cap entries and its count; a scan over segments gives each segment's
// output offset (a paired launch with few segments skips it: each compaction
// block sums its predecessors' counts); a compaction pass copies the slots to
// the dense, ascending output and writes every window's offset.  A wave with
// more survivors than slot_cap raises the overflow flag and the host runs
// again with slot_cap = 1024, which cannot overflow.
//
// Replaces the role of GPUFocus.window (pseudocube/gpu.py:19-35, 94-114):
// same coordinate set, same ascending order.
#include "pcube_kernels.h"

namespace pcube {

constexpr int kWheelThreads = 256;
constexpr int kWheelWaves = kWheelThreads / kWave;
constexpr int kWheelSteps = 16;                               // 64-candidate steps per wave
constexpr int kWheelSegment = kWheelThreads * kWheelSteps;    // candidates per block
constexpr int kWheelWaveSpan = kWheelSteps * kWave;           // candidates per wave
constexpr int kWheelScanChunk = 1024;                         // segments per segment-scan block
constexpr int kMaxOuterRings = 8;
constexpr int kMaxOuterBits = 8 * 1024;   // allowed-residue bits of the outer rings
// An R1 entry is one 64-bit word: the inner residue r in the low
// kResidueBits bits and, when the wheel is packed, r mod q_j of the first
// kPackedOuterRings outer rings as bytes above it.
constexpr int kResidueBits = 40;
constexpr uint64_t kResidueMask = (1ull << kResidueBits) - 1;
constexpr int kPackedOuterRings = (64 - kResidueBits) / 8;

struct WheelWindow {
    uint64_t low, high;  // the window [low, high)
    uint64_t t0;         // first wheel turn, low / M1
    uint32_t ib, ie;     // first candidate index at turn t0, end index at the last turn
    uint32_t nt;         // number of turns t1 - t0 + 1
    uint32_t s0, sm, sl; // segments in the first turn, each middle turn, the last turn
    uint64_t seg_first;  // first segment of this window in the launch
};

struct WheelArgs {
    const uint64_t* R1;          // sorted inner residues r, outer residues packed above them
    const WheelWindow* windows;  // nwin windows; null for one window carried in win0
    const uint64_t* seg_first;   // nwin + 1 prefix of segment counts; null with win0
    const uint32_t* allow;       // outer-ring allowed bits, rings at Modulus::bit
    uint64_t* slots;             // slot_cap inner residues per (segment, wave)
    uint32_t* wave_count;        // survivors per (segment, wave)
    uint32_t* flag;              // set to 1 when a wave overflows its slot
    uint64_t M1, nseg;
    uint32_t n1, nwin, nouter, allow_words, packed, slot_cap;
    Modulus outer[kMaxOuterRings];   // q, inv, r32 for reduction; b = M1 mod q; bit offset
    WheelWindow win0;                // the plan of a one-window launch, passed with the arguments
};

// Plan of window i: from the arguments for a one-window launch, else from memory.
__device__ __forceinline__ WheelWindow wheel_window(const WheelArgs& a, uint32_t i) {
    return a.windows ? a.windows[i] : a.win0;
}

// First segment of window v (v <= nwin); a one-window launch has 0 and nseg.
__device__ __forceinline__ uint64_t wheel_seg_first(const WheelArgs& a, uint32_t v) {
    return a.seg_first ? a.seg_first[v] : (v == 0 ? 0ull : a.nseg);
}

// Window of a segment: the last window whose first segment is <= seg.
__device__ __forceinline__ uint32_t wheel_window_of(const WheelArgs& a, uint64_t seg) {
    uint32_t l = 0, h = a.nwin;
    while (h - l > 1) {
        const uint32_t m = (l + h) >> 1;
        if (wheel_seg_first(a, m) <= seg) l = m; else h = m;
    }
    return l;
}

// Turn t and candidate range [begin, end) of segment s of window w.
__device__ __forceinline__ void wheel_segment(const WheelWindow& w, uint32_t n1, uint32_t s, uint64_t& t,
                                              uint32_t& begin, uint32_t& end) {
    if (w.nt == 1) {
        t = w.t0;
        begin = w.ib + s * kWheelSegment;
        end = min(begin + kWheelSegment, w.ie);
    } else if (s < w.s0) {
        t = w.t0;
        begin = w.ib + s * kWheelSegment;
        end = min(begin + kWheelSegment, n1);
    } else if (s - w.s0 < (w.nt - 2) * w.sm) {
        const uint32_t r = s - w.s0;
        t = w.t0 + 1 + r / w.sm;
        begin = (r % w.sm) * kWheelSegment;
        end = min(begin + kWheelSegment, n1);
    } else {
        const uint32_t r = s - w.s0 - (w.nt - 2) * w.sm;
        t = w.t0 + w.nt - 1;
        begin = r * kWheelSegment;
        end = min(begin + kWheelSegment, w.ie);
    }
}

// Test pass: one block per segment, four waves per block, each wave 16 steps
// of 64 consecutive candidates.  Loads are coalesced and a wave ballot keeps
// the survivors in candidate order within the wave's slot.
// kOuter >= 0: exactly kOuter outer rings with residues packed as bytes, checks
// unrolled; kOuter < 0: any outer rings, residues reduced in the loop.
template <int kOuter>
__device__ __forceinline__ void wheel_slots_block(const WheelArgs& a, uint64_t seg) {
    __shared__ uint32_t allow[kMaxOuterBits / 32];
    __shared__ uint32_t s_window;
    for (uint32_t t = threadIdx.x; t < a.allow_words; t += kWheelThreads) allow[t] = a.allow[t];
    if (threadIdx.x == 0) s_window = wheel_window_of(a, seg);
    __syncthreads();
    const WheelWindow w = wheel_window(a, s_window);
    uint64_t t;
    uint32_t begin, end;
    wheel_segment(w, a.n1, uint32_t(seg - w.seg_first), t, begin, end);  // segments of one window fit 32 bits
    const uint64_t ybase = t * a.M1;
    uint32_t off[kMaxOuterRings];
#pragma unroll
    for (int j = 0; j < kMaxOuterRings; ++j)
        off[j] = j < (int)a.nouter ? mod32(mod64(t, a.outer[j]) * a.outer[j].b, a.outer[j]) : 0;
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const uint32_t wbegin = begin + wave * kWheelWaveSpan;
    const uint32_t wend = min(wbegin + kWheelWaveSpan, end);
    uint64_t* slot = a.slots + (seg * kWheelWaves + wave) * a.slot_cap;
    uint32_t count = 0;
    // issue every load of this wave first: one memory latency instead of sixteen
    uint64_t e[kWheelSteps];
#pragma unroll
    for (int k = 0; k < kWheelSteps; ++k) {
        const uint32_t i = wbegin + k * kWave + lane;
        e[k] = i < wend ? a.R1[i] : 0ull;
    }
#pragma unroll
    for (int k = 0; k < kWheelSteps; ++k) {
        const uint32_t i = wbegin + k * kWave + lane;
        const uint64_t r = e[k] & kResidueMask;
        const uint64_t y = ybase + r;
        bool ok = i < wend && y >= w.low && y < w.high;
        if (kOuter >= 0) {
            const uint32_t bytes = uint32_t(e[k] >> kResidueBits);
#pragma unroll
            for (int j = 0; j < (kOuter > 0 ? kOuter : 0); ++j) {
                uint32_t m = ((bytes >> (8 * j)) & 0xffu) + off[j];
                m = m >= a.outer[j].q ? m - a.outer[j].q : m;
                ok &= allowed_bit(allow, a.outer[j].bit + m);
            }
        } else {
            for (uint32_t j = 0; j < a.nouter; ++j) {
                uint32_t m = mod64(r, a.outer[j]) + off[j];
                m = m >= a.outer[j].q ? m - a.outer[j].q : m;
                ok = ok && allowed_bit(allow, a.outer[j].bit + m);
            }
        }
        const uint64_t live = __ballot(ok);
        if (ok) {
            const uint32_t at = count + lanes_below(live);
            if (at < a.slot_cap) slot[at] = r;
        }
        count += __popcll(live);
    }
    if (lane == 0) {
        a.wave_count[seg * kWheelWaves + wave] = count;
        if (count > a.slot_cap) *a.flag = 1u;
    }
}

template <int kOuter>
__global__ __launch_bounds__(kWheelThreads) void k_wheel_slots(WheelArgs a) {
    wheel_slots_block<kOuter>(a, blockIdx.x);
}

// Two wheels' test passes in one launch (the u and v windows of a batch):
// blocks [0, a.nseg) take a's segments, the rest b's, so the two latency
// chains overlap.  Both wheels use the same instantiation.
// (one call per branch: a reference chosen at run time between the two argument
// structs would be copied to scratch memory)
template <int kOuter>
__global__ __launch_bounds__(kWheelThreads) void k_wheel_slots_pair(WheelArgs a, WheelArgs b) {
    if (blockIdx.x >= a.nseg) wheel_slots_block<kOuter>(b, blockIdx.x - a.nseg);
    else wheel_slots_block<kOuter>(a, blockIdx.x);
}

// Inclusive scan over the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wheel_wave_scan(uint32_t x, uint32_t lane) {
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, kWave);
        if (lane >= (uint32_t)d) x += y;
    }
    return x;
}

// Segment scan: block c covers segments [1024c, 1024c + 1024), four per
// thread.  seg_excl[s] = survivors of the chunk's segments before s;
// chunk_total[c] = survivors of the chunk.
__device__ __forceinline__ void wheel_seg_scan_block(const uint32_t* __restrict__ wave_count, uint64_t nseg,
                                                     uint32_t* __restrict__ seg_excl,
                                                     uint64_t* __restrict__ chunk_total, uint32_t chunk) {
    __shared__ uint32_t wave_sum[256 / kWave];
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const uint64_t s0 = (uint64_t)chunk * kWheelScanChunk + 4 * threadIdx.x;
    uint32_t c[4];
    uint32_t mine = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        c[k] = 0;
        if (s0 + k < nseg) {
            const uint4 v = *(const uint4*)(wave_count + (s0 + k) * kWheelWaves);
            c[k] = v.x + v.y + v.z + v.w;
        }
        mine += c[k];
    }
    const uint32_t inc = wheel_wave_scan(mine, lane);
    if (lane == kWave - 1) wave_sum[wave] = inc;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t k = 0; k < wave; ++k) base += wave_sum[k];
    uint32_t run = base + inc - mine;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (s0 + k < nseg) seg_excl[s0 + k] = run;
        run += c[k];
    }
    if (threadIdx.x == 255) chunk_total[chunk] = run;
}

__global__ __launch_bounds__(256) void k_wheel_seg_scan(const uint32_t* __restrict__ wave_count, uint64_t nseg,
                                                        uint32_t* __restrict__ seg_excl,
                                                        uint64_t* __restrict__ chunk_total) {
    wheel_seg_scan_block(wave_count, nseg, seg_excl, chunk_total, blockIdx.x);
}

// One wheel's scan state: its per-wave counts and segment count, the segment
// offsets and chunk totals it receives, and its number of scan chunks.
struct WheelScan {
    const uint32_t* wave_count;
    uint64_t nseg;
    uint32_t* seg_excl;
    uint64_t* chunk_total;
    uint32_t nchunks;
};

// Two wheels' segment scans in one launch: blocks [0, a.nchunks) scan a.
__global__ __launch_bounds__(256) void k_wheel_seg_scan_pair(WheelScan a, WheelScan b) {
    const bool second = blockIdx.x >= a.nchunks;
    const WheelScan& w = second ? b : a;
    wheel_seg_scan_block(w.wave_count, w.nseg, w.seg_excl, w.chunk_total,
                         second ? blockIdx.x - a.nchunks : blockIdx.x);
}

// Subtile queries of a one-window compaction (a prepared batch's bounds),
// answered by the segments themselves: result k = the number of window
// coordinates below q[k] (searchsorted, side left), written to out[k]; the
// last segment writes the window's total to *total and the overflow flag to
// *flag (out, total and flag may be host memory), the total also to
// *dev_total when set, and fills positions [total, pad_to) of the window with
// 2^64 - 1, above every coordinate, so a match sized before the total was
// known sees a sorted window of pad_to coordinates.
constexpr uint32_t kMaxWindowQueries = 128;
struct WheelQueries {
    uint64_t q[kMaxWindowQueries];
    uint32_t nq;
    uint64_t* out;
    uint64_t* total;
    uint64_t* flag;
    uint64_t* dev_total;
    uint64_t pad_to;
};

// Compaction: one block per segment.  Output offset of the segment = the
// chunk totals before its chunk + its exclusive offset within the chunk; each
// wave copies its slot to out, dropping positions at or beyond cap.  With
// direct (few segments, no scan pass), the block instead sums the survivor
// counts of every earlier segment itself.  The first segment of each window
// writes that window's offset; the last segment writes the total into every
// later entry of win_info, and win_info[nwin + 1] receives the overflow flag.
__device__ __forceinline__ void wheel_compact_block(const WheelArgs& a, const uint32_t* __restrict__ seg_excl,
                                                    const uint64_t* __restrict__ chunk_total,
                                                    uint64_t* __restrict__ out, uint64_t cap,
                                                    uint64_t* __restrict__ win_info, uint64_t seg, bool direct,
                                                    const WheelQueries* queries = nullptr) {
    __shared__ uint64_t partial[kWheelThreads / kWave];
    __shared__ uint32_t below[kWheelThreads / kWave];
    __shared__ uint32_t s_window;
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    uint64_t acc = 0;
    if (direct) {
        for (uint64_t k = threadIdx.x; k < seg; k += kWheelThreads) {
            const uint4 c = *(const uint4*)(a.wave_count + k * kWheelWaves);
            acc += (uint64_t)c.x + c.y + c.z + c.w;
        }
    } else {
        const uint64_t chunk = seg / kWheelScanChunk;
        for (uint64_t k = threadIdx.x; k < chunk; k += kWheelThreads) acc += chunk_total[k];
    }
#pragma unroll
    for (int d = kWave / 2; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, kWave);
    if (lane == 0) partial[wave] = acc;
    if (threadIdx.x == 0) s_window = wheel_window_of(a, seg);
    __syncthreads();
    const uint64_t prefix = partial[0] + partial[1] + partial[2] + partial[3] + (direct ? 0u : seg_excl[seg]);
    const uint4 counts = *(const uint4*)(a.wave_count + seg * kWheelWaves);
    const uint32_t cw[4] = {counts.x, counts.y, counts.z, counts.w};
    uint64_t base = prefix;
    for (uint32_t k = 0; k < wave; ++k) base += cw[k];
    const WheelWindow w = wheel_window(a, s_window);
    uint64_t t;
    uint32_t begin, end;
    wheel_segment(w, a.n1, uint32_t(seg - w.seg_first), t, begin, end);
    const uint64_t ybase = t * a.M1;
    const uint32_t n = min(cw[wave], a.slot_cap);
    const uint64_t* slot = a.slots + (seg * kWheelWaves + wave) * a.slot_cap;
    for (uint32_t i = lane; i < n; i += kWave) {
        const uint64_t o = base + i;
        if (o < cap) out[o] = ybase + slot[i];
    }
    if (threadIdx.x == 0) {
        for (int64_t v = s_window; v >= 0 && wheel_seg_first(a, (uint32_t)v) == seg; --v) win_info[v] = prefix;
        if (seg + 1 == a.nseg) {
            const uint64_t total = prefix + cw[0] + cw[1] + cw[2] + cw[3];
            for (uint32_t v = s_window + 1; v <= a.nwin; ++v) win_info[v] = total;
            win_info[a.nwin + 1] = *a.flag;
            if (queries) {
                *queries->total = total;
                *queries->flag = *a.flag;
                if (queries->dev_total) *queries->dev_total = total;
            }
        }
    }
    if (!queries) return;
    if (seg + 1 == a.nseg) {
        const uint64_t total = prefix + cw[0] + cw[1] + cw[2] + cw[3];
        for (uint64_t o = total + threadIdx.x; o < min(queries->pad_to, cap); o += kWheelThreads) out[o] = ~0ull;
    }
    // the segment's coordinates lie in [lo, hi): from its first candidate to the
    // next segment's first (the first segment from 0, the last to 2^64 - 1), so
    // each query falls in exactly one segment, which counts its own survivors
    // below it (every wave's slot is ascending)
    const uint64_t lo = seg == 0 ? 0ull : ybase + (a.R1[begin] & kResidueMask);
    uint64_t hi = ~0ull;
    if (seg + 1 < a.nseg) {
        uint64_t t2;
        uint32_t b2, e2;
        wheel_segment(w, a.n1, uint32_t(seg + 1 - w.seg_first), t2, b2, e2);
        hi = t2 * a.M1 + (a.R1[b2] & kResidueMask);
    }
    // 64 queries tested at a time, one per lane; every wave finds the same hits
    for (uint32_t k0 = 0; k0 < queries->nq; k0 += kWave) {
        const uint32_t k = k0 + lane;
        const uint64_t mine = k < queries->nq ? queries->q[k] : 0ull;
        uint64_t hits = __ballot(k < queries->nq && mine >= lo && mine < hi);
        while (hits) {
            const uint32_t j = (uint32_t)__builtin_ctzll(hits);
            hits &= hits - 1;
            const uint64_t q = readlane64(mine, (int)j);
            uint32_t c = 0;
            for (uint32_t i = lane; i < n; i += kWave) c += ybase + slot[i] < q ? 1u : 0u;
#pragma unroll
            for (int d = kWave / 2; d >= 1; d >>= 1) c += __shfl_xor(c, d, kWave);
            if (lane == 0) below[wave] = c;
            __syncthreads();
            if (threadIdx.x == 0) queries->out[k0 + j] = prefix + below[0] + below[1] + below[2] + below[3];
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(kWheelThreads) void k_wheel_compact(WheelArgs a, const uint32_t* __restrict__ seg_excl,
                                                                 const uint64_t* __restrict__ chunk_total,
                                                                 uint64_t* __restrict__ out, uint64_t cap,
                                                                 uint64_t* __restrict__ win_info) {
    wheel_compact_block(a, seg_excl, chunk_total, out, cap, win_info, blockIdx.x, false);
}

// One wheel's compaction outputs.
struct WheelOut {
    const uint32_t* seg_excl;
    const uint64_t* chunk_total;
    uint64_t* out;
    uint64_t cap;
    uint64_t* win_info;
    uint32_t direct;   // offsets summed by the compaction itself (no scan pass)
};

// A prepared batch's compaction: both wheels' outputs, their subtile queries,
// and words to clear on the way (a matcher's status and counters, or none).
struct WheelPrepared {
    WheelOut oa, ob;
    WheelQueries qa, qb;
    uint64_t* clear;
    uint64_t nclear;
};

// Segments per wheel up to which a paired launch skips the scan pass and each
// compaction block sums the earlier segments' counts (at most 8 loads a thread).
constexpr uint64_t kDirectPrefixSegments = 2048;

// Two wheels' compactions in one launch: blocks [0, a.nseg) compact a.
__global__ __launch_bounds__(kWheelThreads) void k_wheel_compact_pair(WheelArgs a, WheelOut oa, WheelArgs b,
                                                                      WheelOut ob) {
    if (blockIdx.x >= a.nseg)
        wheel_compact_block(b, ob.seg_excl, ob.chunk_total, ob.out, ob.cap, ob.win_info, blockIdx.x - a.nseg,
                            ob.direct != 0);
    else
        wheel_compact_block(a, oa.seg_excl, oa.chunk_total, oa.out, oa.cap, oa.win_info, blockIdx.x, oa.direct != 0);
}

// k_wheel_compact_pair with the subtile queries answered by the segments
// (one window per wheel): the separate bounds search of a prepared batch.
__global__ __launch_bounds__(kWheelThreads) void k_wheel_compact_prepared(WheelArgs a, WheelArgs b,
                                                                          const WheelPrepared p) {
    for (uint64_t i = blockIdx.x * (uint64_t)kWheelThreads + threadIdx.x; i < p.nclear;
         i += (uint64_t)gridDim.x * kWheelThreads)
        p.clear[i] = 0;
    if (blockIdx.x >= a.nseg)
        wheel_compact_block(b, p.ob.seg_excl, p.ob.chunk_total, p.ob.out, p.ob.cap, p.ob.win_info,
                            blockIdx.x - a.nseg, p.ob.direct != 0, &p.qb);
    else
        wheel_compact_block(a, p.oa.seg_excl, p.oa.chunk_total, p.oa.out, p.oa.cap, p.oa.win_info, blockIdx.x,
                            p.oa.direct != 0, &p.qa);
}

#define PCUBE_WHEEL(K)                                              \
    template __global__ void k_wheel_slots<K>(WheelArgs);           \
    template __global__ void k_wheel_slots_pair<K>(WheelArgs, WheelArgs);
PCUBE_WHEEL(-1)
PCUBE_WHEEL(0)
PCUBE_WHEEL(1)
PCUBE_WHEEL(2)
PCUBE_WHEEL(3)
#undef PCUBE_WHEEL

// Host-side selection of the instantiation for a wheel.
inline const void* wheel_kernel(uint32_t packed, uint32_t nouter) {
    if (!packed) return (const void*)k_wheel_slots<-1>;
    switch (nouter) {
        case 0: return (const void*)k_wheel_slots<0>;
        case 1: return (const void*)k_wheel_slots<1>;
        case 2: return (const void*)k_wheel_slots<2>;
        default: return (const void*)k_wheel_slots<3>;
    }
}

// The paired test pass for two wheels of the same instantiation.
inline const void* wheel_pair_kernel(uint32_t packed, uint32_t nouter) {
    if (!packed) return (const void*)k_wheel_slots_pair<-1>;
    switch (nouter) {
        case 0: return (const void*)k_wheel_slots_pair<0>;
        case 1: return (const void*)k_wheel_slots_pair<1>;
        case 2: return (const void*)k_wheel_slots_pair<2>;
        default: return (const void*)k_wheel_slots_pair<3>;
    }
}

}  // namespace pcube
