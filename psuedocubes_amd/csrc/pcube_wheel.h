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
// to the window width times |R1| / M1, independent of the wheel's size.
//
// Work is cut into segments of kWheelSegment consecutive candidates of one t;
// a block handles one segment.  Pass 1 counts each segment's survivors, a scan
// turns the counts into offsets, pass 2 writes survivors in order.
//
// Replaces the role of GPUFocus.window (pseudocube/gpu.py:19-35, 94-114):
// same coordinate set, same ascending order.
#include "pcube_kernels.h"

namespace pcube {

constexpr int kWheelThreads = 256;
constexpr int kWheelWaves = kWheelThreads / kWave;
constexpr int kWheelSteps = 16;                               // 64-candidate steps per wave
constexpr int kWheelSegment = kWheelThreads * kWheelSteps;    // candidates per block
constexpr int kMaxOuterRings = 8;
constexpr int kMaxOuterBits = 8 * 1024;   // allowed-residue bits of the outer rings

struct WheelWindow {
    uint64_t low, high;  // the window [low, high)
    uint64_t t0;         // first wheel turn, low / M1
    uint32_t ib, ie;     // first candidate index at turn t0, end index at the last turn
    uint32_t nt;         // number of turns t1 - t0 + 1
    uint32_t s0, sm, sl; // segments in the first turn, each middle turn, the last turn
    uint64_t seg_first;  // first segment of this window in the launch
};

struct WheelArgs {
    const uint2* R1;             // sorted inner residues r with their outer residues packed as bytes
    const WheelWindow* windows;  // nwin windows
    const uint64_t* seg_first;   // nwin + 1 prefix of segment counts
    const uint32_t* allow;       // outer-ring allowed bits, rings at Modulus::bit
    uint32_t* wave_count;        // pass 1 output, one per (segment, wave)
    const uint32_t* wave_offset; // pass 2 input, exclusive scan of wave_count
    uint64_t* out;               // pass 2 output
    uint64_t M1, cap;
    uint32_t n1, nwin, nouter, allow_words, packed;
    Modulus outer[kMaxOuterRings];   // q, inv, r32 for reduction; b = M1 mod q; bit offset
};

// One block per segment, four waves per block, each wave 16 steps of 64
// consecutive candidates: loads are coalesced and a wave ballot keeps the
// survivors in candidate order.
// kOuter >= 0: exactly kOuter outer rings with residues packed as bytes, checks
// unrolled; kOuter < 0: any outer rings, residues reduced in the loop.
template <bool kWrite, int kOuter>
__global__ __launch_bounds__(kWheelThreads) void k_wheel(WheelArgs a) {
    __shared__ uint32_t allow[kMaxOuterBits / 32];
    __shared__ uint32_t s_window;
    for (uint32_t t = threadIdx.x; t < a.allow_words; t += kWheelThreads) allow[t] = a.allow[t];
    const uint64_t seg = blockIdx.x;
    if (threadIdx.x == 0) {
        uint32_t l = 0, h = a.nwin;  // last window with seg_first <= seg
        while (h - l > 1) {
            const uint32_t m = (l + h) >> 1;
            if (a.seg_first[m] <= seg) l = m; else h = m;
        }
        s_window = l;
    }
    __syncthreads();
    const WheelWindow w = a.windows[s_window];
    const uint32_t s = uint32_t(seg - w.seg_first);   // segments of one window fit 32 bits (host checks)
    // turn and candidate range of this segment
    uint64_t t;
    uint32_t begin, end;
    if (w.nt == 1) {
        t = w.t0;
        begin = w.ib + s * kWheelSegment;
        end = min(begin + kWheelSegment, w.ie);
    } else if (s < w.s0) {
        t = w.t0;
        begin = w.ib + s * kWheelSegment;
        end = min(begin + kWheelSegment, a.n1);
    } else if (s - w.s0 < (w.nt - 2) * w.sm) {
        const uint32_t r = s - w.s0;
        t = w.t0 + 1 + r / w.sm;
        begin = (r % w.sm) * kWheelSegment;
        end = min(begin + kWheelSegment, a.n1);
    } else {
        const uint32_t r = s - w.s0 - (w.nt - 2) * w.sm;
        t = w.t0 + w.nt - 1;
        begin = r * kWheelSegment;
        end = min(begin + kWheelSegment, w.ie);
    }
    const uint64_t ybase = t * a.M1;
    uint32_t off[kMaxOuterRings];
#pragma unroll
    for (int j = 0; j < kMaxOuterRings; ++j)
        off[j] = j < (int)a.nouter ? mod32(mod64(t, a.outer[j]) * a.outer[j].b, a.outer[j]) : 0;
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const uint32_t wbegin = begin + wave * (kWheelSteps * kWave);
    const uint32_t wend = min(wbegin + kWheelSteps * kWave, end);
    uint64_t pos = kWrite ? (uint64_t)a.wave_offset[seg * kWheelWaves + wave] : 0;
    uint32_t count = 0;
    // issue every load of this wave first: one memory latency instead of sixteen
    uint2 e[kWheelSteps];
#pragma unroll
    for (int k = 0; k < kWheelSteps; ++k) {
        const uint32_t i = wbegin + k * kWave + lane;
        e[k] = i < wend ? a.R1[i] : make_uint2(0, 0);
    }
#pragma unroll
    for (int k = 0; k < kWheelSteps; ++k) {
        const uint32_t i = wbegin + k * kWave + lane;
        const uint64_t y = ybase + e[k].x;
        bool ok = i < wend && y >= w.low && y < w.high;
        if (kOuter >= 0) {
#pragma unroll
            for (int j = 0; j < (kOuter > 0 ? kOuter : 0); ++j) {
                uint32_t m = ((e[k].y >> (8 * j)) & 0xffu) + off[j];
                m = m >= a.outer[j].q ? m - a.outer[j].q : m;
                ok &= allowed_bit(allow, a.outer[j].bit + m);
            }
        } else {
            for (uint32_t j = 0; j < a.nouter; ++j) {
                uint32_t m = mod32(e[k].x, a.outer[j]) + off[j];
                m = m >= a.outer[j].q ? m - a.outer[j].q : m;
                ok = ok && allowed_bit(allow, a.outer[j].bit + m);
            }
        }
        const uint64_t live = __ballot(ok);
        if (kWrite && ok) {
            const uint64_t slot = pos + lanes_below(live);
            if (slot < a.cap) a.out[slot] = y;
        }
        pos += __popcll(live);
        count += __popcll(live);
    }
    if (!kWrite && lane == 0) a.wave_count[seg * kWheelWaves + wave] = count;
}

#define PCUBE_WHEEL(K)                                      \
    template __global__ void k_wheel<false, K>(WheelArgs);  \
    template __global__ void k_wheel<true, K>(WheelArgs);
PCUBE_WHEEL(-1)
PCUBE_WHEEL(0)
PCUBE_WHEEL(1)
PCUBE_WHEEL(2)
PCUBE_WHEEL(3)
PCUBE_WHEEL(4)
#undef PCUBE_WHEEL

// Host-side selection of the instantiation for a wheel.
template <bool kWrite>
inline const void* wheel_kernel(uint32_t packed, uint32_t nouter) {
    if (!packed) return (const void*)k_wheel<kWrite, -1>;
    switch (nouter) {
        case 0: return (const void*)k_wheel<kWrite, 0>;
        case 1: return (const void*)k_wheel<kWrite, 1>;
        case 2: return (const void*)k_wheel<kWrite, 2>;
        case 3: return (const void*)k_wheel<kWrite, 3>;
        default: return (const void*)k_wheel<kWrite, 4>;
    }
}

}  // namespace pcube
