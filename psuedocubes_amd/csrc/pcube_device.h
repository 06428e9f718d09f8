#pragma once
// HIP kernels for the p=619 doubly-focused pseudocube batch engine (gfx950).
//
// Reference behaviour restated here (file:line of AlexanderEastwood/psuedocubes):
//   k_partial        CRT partial sums of a ring group        pseudocube/gpu.py:82-92
//   k_window_*       focus coordinates in a window           pseudocube/gpu.py:19-35,94-114
//   k_searchsorted   subtile bounds                          pseudocube/batch_experiment.py:39-43
//   k_rows           per-v exact u bounds, 128-bit           pseudocube/batch_experiment.py:209-215
//   k_masks          first-four residue bitsets              pseudocube/batch_experiment.py:184-194
//   k_match          bitset intersection, residual chain,
//                    group counts, survivor output           pseudocube/batch_experiment.py:198-235
//   k_match_direct   per-position chain without bitsets      pseudocube/gpu.py:147-181
//   k_arithmetic     wide-limb self test                     kernels/pseudocube.cuh:6-11
//   k_floor_div      the u-interval division, self test      pseudocube/focus.py:120-121
//   k_bit_transpose  the mask builders' wave transpose, self test
#include "pcube_kernels.h"

namespace pcube {

// --------------------------------------------------------------------------
// Fills and scans

// Zero two ranges in one launch.
__global__ void k_fill2_u64(uint64_t* __restrict__ a, uint64_t na, uint64_t* __restrict__ b, uint64_t nb) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < na + nb; i += (uint64_t)gridDim.x * blockDim.x) {
        if (i < na) a[i] = 0;
        else b[i - na] = 0;
    }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t value) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    for (; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = value;
}
__global__ void k_fill_u64(uint64_t* __restrict__ p, uint64_t n, uint64_t value) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    for (; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = value;
}

constexpr int kScanThreads = 1024;
constexpr int kScanPerThread = 4;
constexpr int kScanSegment = kScanThreads * kScanPerThread;

// Exclusive scan of one value per thread across a 1024-thread block.
__device__ uint32_t block_exclusive_scan(uint32_t value, uint32_t* total) {
    __shared__ uint32_t wave_sums[kScanThreads / kWave];
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    uint32_t inclusive = value;
    for (int d = 1; d < kWave; d <<= 1) {
        uint32_t up = __shfl_up(inclusive, d, kWave);
        if (lane >= (uint32_t)d) inclusive += up;
    }
    if (lane == kWave - 1) wave_sums[wave] = inclusive;
    __syncthreads();
    if (wave == 0) {
        uint32_t s = lane < kScanThreads / kWave ? wave_sums[lane] : 0;
        uint32_t inc = s;
        for (int d = 1; d < kWave; d <<= 1) {
            uint32_t up = __shfl_up(inc, d, kWave);
            if (lane >= (uint32_t)d) inc += up;
        }
        if (lane < kScanThreads / kWave) wave_sums[lane] = inc - s;
        if (lane == kScanThreads / kWave - 1) *total = inc;
    }
    __syncthreads();
    return wave_sums[wave] + inclusive - value;
}

// out[i] = sum of in[0..i) within this block's segment; partial[block] = segment sum.
__global__ __launch_bounds__(kScanThreads) void k_scan_local(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                             uint64_t n, uint32_t* __restrict__ partial) {
    __shared__ uint32_t total;
    uint64_t base = (uint64_t)blockIdx.x * kScanSegment + (uint64_t)threadIdx.x * kScanPerThread;
    uint32_t v[kScanPerThread], sum = 0;
    for (int e = 0; e < kScanPerThread; ++e) {
        v[e] = base + e < n ? in[base + e] : 0;
        sum += v[e];
    }
    uint32_t offset = block_exclusive_scan(sum, &total);
    for (int e = 0; e < kScanPerThread; ++e) {
        if (base + e < n) out[base + e] = offset;
        offset += v[e];
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

__global__ void k_scan_add(uint32_t* __restrict__ out, uint64_t n, const uint32_t* __restrict__ block_offsets) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] += block_offsets[i / kScanSegment];
}

// --------------------------------------------------------------------------
// Focus construction: partial CRT sums and bucket sort

// out[i] = sum over rings k of terms_k[digit_k(i)] mod M, first ring least significant.
__global__ void k_partial(const uint64_t* __restrict__ terms, const uint32_t* __restrict__ sizes, uint32_t nrings,
                          uint64_t M, uint64_t n, uint64_t* __restrict__ out) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t rest = i, sum = 0;
    uint32_t offset = 0;
    for (uint32_t k = 0; k < nrings; ++k) {
        uint32_t s = sizes[k];
        uint64_t digit = rest % s;
        rest /= s;
        sum += terms[offset + digit];
        if (sum >= M) sum -= M;
        offset += s;
    }
    out[i] = sum;
}

__global__ void k_bucket_count(const uint64_t* __restrict__ values, uint64_t n, uint64_t low, uint32_t shift,
                               uint32_t* __restrict__ hist) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&hist[(values[i] - low) >> shift], 1u);
}

__global__ void k_bucket_scatter(const uint64_t* __restrict__ values, uint64_t n, uint64_t low, uint32_t shift,
                                 const uint32_t* __restrict__ offsets, uint32_t* __restrict__ cursor,
                                 uint64_t* __restrict__ out) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t y = values[i];
    uint32_t b = uint32_t((y - low) >> shift);
    out[offsets[b] + atomicAdd(&cursor[b], 1u)] = y;
}

// Insertion sort inside each bucket [offsets[b], offsets[b+1]) clamped to cap.
__global__ void k_bucket_sort(uint64_t* __restrict__ out, const uint32_t* __restrict__ offsets, uint64_t nbuckets,
                              uint64_t cap) {
    uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (b >= nbuckets) return;
    uint64_t s = offsets[b], e = offsets[b + 1];
    if (e > cap) e = cap;
    for (uint64_t i = s + 1; i < e; ++i) {
        uint64_t key = out[i], j = i;
        while (j > s && out[j - 1] > key) {
            out[j] = out[j - 1];
            --j;
        }
        out[j] = key;
    }
}

// dir[k] = first index i with C[i] >= k << shift, for k in [0, nbk]; dir[nbk] = nc.
__global__ void k_directory(const uint64_t* __restrict__ C, uint64_t nc, uint32_t shift, uint64_t nbk,
                            uint32_t* __restrict__ dir) {
    uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (k > nbk) return;
    if (k == nbk) {
        dir[k] = uint32_t(nc);
        return;
    }
    uint64_t key = k << shift, l = 0, h = nc;
    while (l < h) {
        uint64_t m = l + ((h - l) >> 1);
        if (C[m] < key) l = m + 1; else h = m;
    }
    dir[k] = uint32_t(l);
}

// --------------------------------------------------------------------------
// Focus windows
//
// All coordinates y in [low, high) with (y mod M) = (c + d) mod M for c in the
// sorted half C and d in the sorted half D.  Thread = (period, chunk of kChunk
// consecutive D entries, loaded up front).  Per wrap w in {0, 1}, c must lie
// in [a + wM - d, b + wM - d); walking the chunk's d values downwards moves that
// interval upwards, so two pointers into C only advance.  The first pointer
// position comes from the bucket directory of C.
//
// Output modes:
//   kCount  count outputs per value bucket (first pass of the exact path)
//   kEmit   scatter into bucket ranges fixed by the scanned counts (second pass)
//   kSlots  single pass: append into kSlotsPerBucket fixed slots per value
//           bucket; a full bucket raises the overflow flag and the caller
//           reruns the exact two-pass path

constexpr int kChunk = 8;
constexpr int kSlotsPerBucket = 16;
enum WindowMode { kCount = 0, kEmit = 1, kSlots = 2 };

struct WindowArgs {
    const uint64_t* C;
    const uint64_t* D;
    const uint32_t* dir;
    uint64_t nc, nd, nbk, M, low, high, per0, cap;
    uint32_t dshift, nper, nchunks, bshift;
};

__device__ __forceinline__ uint64_t lower_bound_dir(const WindowArgs& w, uint64_t x) {
    uint64_t k = x >> w.dshift;
    if (k >= w.nbk) return w.nc;
    uint64_t i = w.dir[k];
    while (i < w.nc && w.C[i] < x) ++i;
    return i;
}

template <int kMode>
__global__ __launch_bounds__(256) void k_window(WindowArgs w, uint32_t* __restrict__ hist,
                                                const uint32_t* __restrict__ offsets, uint32_t* __restrict__ cursor,
                                                uint64_t* __restrict__ out, uint32_t* __restrict__ flag) {
    uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= (uint64_t)w.nper * w.nchunks) return;
    const uint32_t p = uint32_t(t / w.nchunks), chunk = uint32_t(t % w.nchunks);
    const uint64_t base = (w.per0 + p) * w.M;
    const uint64_t a = w.low > base ? w.low - base : 0;
    const uint64_t b = (w.high - base) < w.M ? w.high - base : w.M;
    const uint64_t j0 = (uint64_t)chunk * kChunk;
    const int n = int(w.nd - j0 < (uint64_t)kChunk ? w.nd - j0 : (uint64_t)kChunk);
    uint64_t dv[kChunk];
#pragma unroll
    for (int e = 0; e < kChunk; ++e) dv[e] = e < n ? w.D[j0 + e] : 0;
    for (int wrap = 0; wrap < 2; ++wrap) {
        const uint64_t lw = a + (wrap ? w.M : 0), hw = b + (wrap ? w.M : 0);
        uint64_t pl = 0, pr = 0;
        bool started = false;
#pragma unroll
        for (int e = kChunk - 1; e >= 0; --e) {
            if (e >= n) continue;
            const uint64_t d = dv[e];
            if (hw <= d) continue;
            const uint64_t cl = lw > d ? lw - d : 0, ch = hw - d;
            if (cl >= w.M) break;
            if (!started) {
                pl = lower_bound_dir(w, cl);
                started = true;
            } else {
                while (pl < w.nc && w.C[pl] < cl) ++pl;
            }
            if (pr < pl) pr = pl;
            while (pr < w.nc && w.C[pr] < ch) ++pr;
            for (uint64_t i = pl; i < pr; ++i) {
                const uint64_t y = w.C[i] + d - (wrap ? w.M : 0) + base;
                const uint32_t bucket = uint32_t((y - w.low) >> w.bshift);
                if (kMode == kCount) {
                    atomicAdd(&hist[bucket], 1u);
                } else if (kMode == kEmit) {
                    const uint64_t slot = (uint64_t)offsets[bucket] + atomicAdd(&cursor[bucket], 1u);
                    if (slot < w.cap) out[slot] = y;
                } else {
                    const uint32_t slot = atomicAdd(&hist[bucket], 1u);
                    if (slot < (uint32_t)kSlotsPerBucket) out[(uint64_t)bucket * kSlotsPerBucket + slot] = y;
                    else atomicOr(flag, 1u);
                }
            }
        }
    }
}

// Single-pass windows, second step: bucket b holds min(count, kSlotsPerBucket)
// values in its slots; sort them and write them at the scanned offset.
__global__ void k_slots_compact(const uint64_t* __restrict__ slots, const uint32_t* __restrict__ offsets,
                                uint64_t nbuckets, uint64_t* __restrict__ out, uint64_t cap) {
    uint64_t bk = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (bk >= nbuckets) return;
    const uint32_t s = offsets[bk];
    uint32_t n = offsets[bk + 1] - s;
    if (n == 0) return;
    if (n > (uint32_t)kSlotsPerBucket) n = kSlotsPerBucket;
    uint64_t v[kSlotsPerBucket];
#pragma unroll
    for (int i = 0; i < kSlotsPerBucket; ++i) v[i] = i < (int)n ? slots[bk * kSlotsPerBucket + i] : ~0ull;
#pragma unroll
    for (int i = 1; i < kSlotsPerBucket; ++i)
#pragma unroll
        for (int j = i; j > 0; --j)
            if (v[j - 1] > v[j]) { uint64_t x = v[j]; v[j] = v[j - 1]; v[j - 1] = x; }
#pragma unroll
    for (int i = 0; i < kSlotsPerBucket; ++i)
        if (i < (int)n && (uint64_t)s + i < cap) out[(uint64_t)s + i] = v[i];
}

// np.searchsorted(sorted, queries, side='left')
__global__ void k_searchsorted(const uint64_t* __restrict__ sorted, uint64_t n, const uint64_t* __restrict__ queries,
                               uint64_t nq, uint64_t* __restrict__ out) {
    uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= nq) return;
    uint64_t key = queries[t], l = 0, h = n;
    while (l < h) {
        uint64_t m = l + ((h - l) >> 1);
        if (sorted[m] < key) l = m + 1; else h = m;
    }
    out[t] = l;
}

// --------------------------------------------------------------------------
// Matching

// Per-v row data computed once per batch.  The first 32 bytes are what the
// bitset stage reads; v serves the residual chain.
struct Row {
    uint32_t left, right;   // u index range [left, right) with lo < u*B - v*A <= hi
    uint32_t tile, flags;   // subtile index; bit0: a v outside [v_first, v_first + ntiles*width)
    uint32_t r[kStagedFilters];       // v mod q for the bitset filters (staged engine: LDS byte offset per read slot)
    uint64_t v;
    uint64_t pad;
};
static_assert(sizeof(Row) == 64, "Row is one 64-byte record");

// Staged engine: the bitset filter (column of the chunk image) that the lane
// of row j reads through offset slot i.  Slots 0-3 hold one group of four
// filters and slots 4-7 the other, both rotated by j mod 16, the lane's
// position in its 16-lane LDS cycle (pcube_staged.h).
__host__ __device__ __forceinline__ uint32_t stage_column(uint32_t i, uint32_t j) {
    const uint32_t g = j & 15;
    return ((i & 3) ^ (g & 3)) + 4 * ((i >> 2) ^ (g >> 3));
}

struct MatchArgs {
    const uint64_t* us;
    const uint64_t* vs;
    const Row* rows;
    const uint64_t* masks;     // [sum of bitset q rows][words]
    const uint16_t* rho;       // [nu][kCachedResidues] (u*B) mod q for the cached chain filters
    const Modulus* mods;       // [nres] in filter order
    const uint32_t* allow;     // allowed-residue bits, rows at Modulus::bit
    uint64_t* counters;        // [ntiles][ng + 3]
    uint64_t* list;            // emitted values as (global tile, low, high) triples
    uint64_t* list_count;      // next free triple
    uint32_t* error;           // bit0 invalid tile, bit1 value outside (lo, hi]
    uint64_t A, B, v_first, tile_width, cap, list_cap, tile_base;
    U128 B_mu;                 // floor((2^128 - 1) / B), for exact divisions by B
    U128 lo, hi;
    uint32_t nu, nv, words, nres, nbits, ncache, ng, ntiles, allow_words, rows_per_wave;
    uint32_t row_base[kStagedFilters];
    // staged engine: masks in chunk images of stage_lines lines x 8 columns x kStageWords words
    const uint32_t* pstr;      // per bitset filter, its permuted allowed string, from pstr_off[k]
    uint32_t staged, stage_lines, nstage;
    uint32_t nbeta[kStagedFilters];    // (-B * A^-1) mod q per bitset filter
    uint32_t pstr_off[kStagedFilters];
    uint32_t* cand;            // bitset survivor records, item_cap per wave of the staged match
    uint64_t* cand_max;        // largest wave record count above item_cap seen since the last clear
    uint32_t item_cap;
};

// State of one search for the first index i in [0, nu) with bound < us[i]*B
// (nu when there is none): the answer lies in [l, h], and P(i) =
// (us[i]*B <= bound) holds exactly below it; done once the answer is in l.
// P is tested as us[i] < key with key = floor(bound / B) + 1, one 64-bit
// compare per probe instead of a 64 x 64 -> 128 product.
// hval = us[h] whenever h < nu, and once done, val = us[l] whenever l < nu:
// every probe that moves h or ends the search has loaded that coordinate.
struct AboveSearch {
    AboveKey key;
    uint32_t l, h;
    bool done;
    uint64_t hval, val;
};

// With a directory (dir[k] = us[k * stride], ndir entries, in LDS): the first
// directory entry above the bound leaves the answer among the at most stride
// coordinates before it, so few probes reach global memory.
__device__ __forceinline__ void above_directory(const MatchArgs& m, AboveSearch& s, const uint64_t* dir, uint32_t ndir,
                                                uint32_t stride, uint32_t a) {
    if (a == 0) {
        s.l = 0;
        s.val = dir[0];
        s.done = true;
        return;
    }
    s.l = (a - 1) * stride + 1;
    s.h = a * stride < m.nu ? a * stride : m.nu;
    if (s.h < m.nu) s.hval = dir[a];   // a * stride < nu, so a < ndir
}

// Narrow a search with eight probes of us already loaded (x[k] = us[l + k*step]
// for the in-range k).  kTrack keeps hval and val.
template <bool kTrack>
__device__ __forceinline__ void above_round(const MatchArgs& m, AboveSearch& s, uint32_t step, const uint64_t (&x)[8]) {
    uint32_t below = 0;
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k)
        below += s.l + k * step < s.h && below_key(x[k], s.key) ? 1u : 0u;
    // x[below] = us[l + below*step] when that probe is in range (below < 8)
    uint64_t xb = 0;
    if (kTrack) {
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) xb = k == below ? x[k] : xb;
    }
    if (step == 1) {   // the last round: the window held at most eight coordinates
        if (kTrack) s.val = s.l + below < s.h ? xb : s.hval;
        s.l += below;
        s.done = true;
        return;
    }
    if (below == 0) {
        if (kTrack) s.val = x[0];
        s.done = true;
        return;
    }
    const uint32_t top = s.l + below * step;
    s.l += (below - 1) * step + 1;
    if (top < s.h) {
        s.h = top;
        if (kTrack) s.hval = xb;
    }
}

// first index with bound < us[i]*B for two bounds at once (the two ends of a
// row or of a piece).  The searches advance together, so each round's LDS or
// global probes of both are in flight at the same time: eight probes per
// search per round, a 63-coordinate window in two rounds of loads and a
// final round over the last eight.  kTrack also returns u0 = us[r0] (when
// r0 < nu) from the probes, at the cost of the registers that follow it.
template <bool kTrack>
__device__ __forceinline__ void first_above_keys(const MatchArgs& m, const AboveKey& k0, const AboveKey& k1,
                                                 const uint64_t* dir, uint32_t ndir, uint32_t stride, uint32_t& r0,
                                                 uint32_t& r1, uint64_t& u0) {
    AboveSearch s[2] = {{k0, 0, m.nu, false, 0, 0}, {k1, 0, m.nu, false, 0, 0}};
    if (dir) {
        uint32_t a[2] = {0, 0}, e[2] = {ndir, ndir};
        while (a[0] < e[0] || a[1] < e[1]) {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                if (a[i] < e[i]) {
                    const uint32_t mid = a[i] + ((e[i] - a[i]) >> 1);
                    if (!below_key(dir[mid], s[i].key)) e[i] = mid; else a[i] = mid + 1;
                }
            }
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) above_directory(m, s[i], dir, ndir, stride, a[i]);
    }
    while (!s[0].done || !s[1].done) {
        uint32_t step[2];
        uint64_t x[2][8];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            step[i] = s[i].h - s[i].l > 8 ? (s[i].h - s[i].l + 7) / 8 : 1u;
#pragma unroll
            for (uint32_t k = 0; k < 8; ++k) {
                const uint32_t at = s[i].l + k * step[i];
                x[i][k] = !s[i].done && at < s[i].h ? m.us[at] : 0ull;
            }
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
            if (!s[i].done) above_round<kTrack>(m, s[i], step[i], x[i]);
    }
    r0 = s[0].l;
    r1 = s[1].l;
    if (kTrack) u0 = s[0].val;   // us[r0] when r0 < nu
}

template <bool kTrack>
__device__ __forceinline__ void first_above_pair(const MatchArgs& m, const U128& b0, const U128& b1,
                                                 const uint64_t* dir, uint32_t ndir, uint32_t stride, uint32_t& r0,
                                                 uint32_t& r1, uint64_t& u0) {
    first_above_keys<kTrack>(m, above_key(b0, m.B, m.B_mu), above_key(b1, m.B, m.B_mu), dir, ndir, stride, r0, r1,
                             u0);
}

// The parts of row j (coordinate v) that depend on v alone: flags and
// subtile, the bitset offsets, v*A and the search keys of the row's two u
// bounds.
struct RowPrep {
    Row row;
    U128 shift;
    AboveKey lower, upper;
};

__device__ __forceinline__ RowPrep row_prep(const MatchArgs& m, uint32_t j, uint64_t v) {
    RowPrep p;
    Row& row = p.row;
    row.v = v;
    row.flags = 0;
    uint64_t tile = 0;
    if (v < m.v_first || (tile = (v - m.v_first) / m.tile_width) >= m.ntiles) {
        row.flags = 1;
        tile = 0;
        atomicOr(m.error, 1u);
    }
    row.tile = uint32_t(tile);
    p.shift = mul_wide(v, m.A);
    p.lower = above_key(add_wide(p.shift, m.lo), m.B, m.B_mu);
    p.upper = above_key(add_wide(p.shift, m.hi), m.B, m.B_mu);
    // the staged engine keeps, per offset slot, the chunk-image byte offset of
    // the mask half its lane reads first there: 256 * line + 32 * column + 16 *
    // P, with line v mod q of the slot's filter (line 0, all ones, for an
    // absent filter) and P = (g >> 2) & 1 the lane's first half; the other
    // half is the neighbouring 16 bytes (offset xor 16).  Slot k of row j reads
    // column stage_column(k, j) = k ^ x with x = (g & 3) | (g >> 3) << 2,
    // g = j mod 16: the eight entries are formed in column order (the moduli
    // are the same for every lane) and permuted by three conditional swaps.
    uint32_t res[kStagedFilters];
#pragma unroll
    for (uint32_t f = 0; f < (uint32_t)kStagedFilters; ++f) {
        const uint32_t r = f < m.nbits ? mod64(v, m.mods[f]) : 0u;
        res[f] = m.staged ? r * kStagedFilters + f : r;
    }
    if (m.staged) {
        const uint32_t g = j & 15, x = (g & 3) | ((g >> 3) << 2), half = (g >> 2) & 1;
#pragma unroll
        for (uint32_t bit = 1; bit < (uint32_t)kStagedFilters; bit <<= 1) {
            const bool swap = (x & bit) != 0;
#pragma unroll
            for (uint32_t k = 0; k < (uint32_t)kStagedFilters; ++k) {
                if (k & bit) continue;
                const uint32_t a = res[k], b = res[k | bit];
                res[k] = swap ? b : a;
                res[k | bit] = swap ? a : b;
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < (uint32_t)kStagedFilters; ++k) res[k] = 32 * res[k] + 16 * half;
    }
#pragma unroll
    for (uint32_t k = 0; k < (uint32_t)kStagedFilters; ++k) row.r[k] = res[k];
    row.pad = 0;
    return p;
}

// The rest of row j: exact u bounds by 128-bit binary search, the underflow
// check, and the record's store.  With a directory (kDir) the search also
// yields us[left]; without one, us[left] is loaded.
template <bool kDir>
__device__ __forceinline__ Row row_finish(const MatchArgs& m, uint32_t j, RowPrep& p, const uint64_t* dir,
                                          uint32_t ndir, uint32_t stride) {
    Row& row = p.row;
    // first index with u*B > bound, for the lower and the upper bound
    uint32_t left, right;
    uint64_t u_left = 0;
    first_above_keys<kDir>(m, p.lower, p.upper, dir, ndir, stride, left, right, u_left);
    row.left = left;
    row.right = right;
    if (!kDir && left < right) u_left = m.us[left];
    if (row.flags) row.right = row.left;  // an invalid row contributes nothing
    // u*B >= v*A holds for every u in range because u*B > v*A + lo; a violation
    // (impossible for sorted input) is reported as the underflow counter.
    if (row.left < row.right && less_wide(mul_wide(u_left, m.B), p.shift))
        atomicAdd(&m.counters[(uint64_t)row.tile * (m.ng + 3) + m.ng + 2], 1ull);
    ((Row*)m.rows)[j] = row;
    return row;
}

// Row j (coordinate v = vs[j]): exact u bounds, subtile, v mod q.
template <bool kDir>
__device__ __forceinline__ Row build_row(const MatchArgs& m, uint32_t j, uint64_t v, const uint64_t* dir = nullptr,
                                         uint32_t ndir = 0, uint32_t stride = 0) {
    RowPrep p = row_prep(m, j, v);
    return row_finish<kDir>(m, j, p, dir, ndir, stride);
}

__global__ void k_rows(MatchArgs m) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < m.nv) build_row<false>(m, j, m.vs[j]);
}

// Staged engine rows: as k_rows, plus each subtile's valid-position count,
// sum of right - left over its rows.  Rows are sorted by v, so a wave's rows
// span few subtiles: one atomic per (wave, subtile).
constexpr int kRowDirectory = 4096;   // u coordinates a row block samples into LDS

// Stride (a power of two >= 64) and entry count of the directory of nu u
// coordinates: every stride-th coordinate, at most kRowDirectory entries.
__host__ __device__ inline void directory_shape(uint32_t nu, uint32_t& stride, uint32_t& ndir) {
    stride = 64;
    while ((nu + stride - 1) / stride > (uint32_t)kRowDirectory) stride <<= 1;
    ndir = (nu + stride - 1) / stride;
}

// Fill dir with every stride-th u coordinate (stride a power of two >= 64 with
// at most kRowDirectory samples) with the whole 256-thread block; ends with a
// barrier.
__device__ __forceinline__ void load_directory(const MatchArgs& m, uint64_t* dir, uint32_t& ndir, uint32_t& stride) {
    directory_shape(m.nu, stride, ndir);
    for (uint32_t first = 0; first < ndir; first += 256 * 16) {
        uint64_t v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint32_t t = first + 256 * k + threadIdx.x;
            v[k] = m.us[(t < ndir ? t : ndir - 1) * stride];
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint32_t t = first + 256 * k + threadIdx.x;
            if (t < ndir) dir[t] = v[k];
        }
    }
    __syncthreads();
}

// Staged engine rows for one 256-row block: as k_rows, plus each subtile's
// valid-position count, sum of right - left over its rows.  Rows are sorted by
// v, so a wave's rows span few subtiles: one atomic per (wave, subtile).
// Rows from nv_real on (up to m.nv) are padding: an empty u range, no counts.
__device__ __forceinline__ void staged_rows_block(const MatchArgs& m, uint32_t block, uint64_t* dir,
                                                  uint32_t nv_real) {
    const uint32_t j = block * 256 + threadIdx.x;
    if (j >= nv_real && j < m.nv) {
        Row pad{};
        pad.v = ~0ull;
        ((Row*)m.rows)[j] = pad;
    }
    // v first, then the directory's loads; the parts of the row that need v
    // alone are computed while the directory is in flight
    const uint64_t v = j < nv_real ? m.vs[j] : 0ull;
    uint32_t ndir, stride;
    directory_shape(m.nu, stride, ndir);
    static_assert(kRowDirectory == 256 * 16, "one round of sixteen loads per thread fills the directory");
    uint64_t d[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const uint32_t t = 256 * k + threadIdx.x;
        d[k] = t < ndir ? m.us[t * stride] : 0ull;   // only the directory's own entries are loaded
    }
    RowPrep p;
    if (j < nv_real) p = row_prep(m, j, v);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const uint32_t t = 256 * k + threadIdx.x;
        if (t < ndir) dir[t] = d[k];
    }
    __syncthreads();
    uint32_t tile = 0xffffffffu;
    uint64_t count = 0;
    if (j < nv_real) {
        const Row row = row_finish<true>(m, j, p, ndir ? dir : nullptr, ndir, stride);
        if (!row.flags) {
            tile = row.tile;
            count = row.right - row.left;
        }
    }
    uint64_t pending = __ballot(tile != 0xffffffffu);
    while (pending) {
        const uint32_t first = __builtin_ctzll(pending);
        const uint32_t t = __shfl(tile, first, kWave);
        const bool mine = tile == t;
        uint64_t part = mine ? count : 0;
        for (int d = kWave / 2; d > 0; d >>= 1) part += __shfl_xor(part, d, kWave);
        if ((threadIdx.x & (kWave - 1)) == first && part)
            atomicAdd(&m.counters[(uint64_t)t * (m.ng + 3)], part);
        pending &= ~__ballot(mine);
    }
}

// Lane-exchange helpers of the 64 x 64 bit wave transpose.  DPP moves within
// 16-lane rows (every source lane exists for these patterns):
//   row_ror:8 -> lane ^ 8; row_half_mirror then quad_perm [3,2,1,0] -> lane ^ 4;
//   quad_perm [2,3,0,1] -> lane ^ 2; quad_perm [1,0,3,2] -> lane ^ 1.
template <int kCtrl>
__device__ __forceinline__ uint32_t dpp_move(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, kCtrl, 0xf, 0xf, true);
}

template <int kB>
__device__ __forceinline__ uint32_t lane_xor(uint32_t x) {
    if constexpr (kB == 8) return dpp_move<0x128>(x);
    else if constexpr (kB == 4) return dpp_move<0x1b>(dpp_move<0x141>(x));
    else if constexpr (kB == 2) return dpp_move<0x4e>(x);
    else return dpp_move<0xb1>(x);
}

// Transpose stage kB < 16 on one 32-bit half (its 2kB-bit groups lie inside
// the half): y is the half of lane ^ kB.  A lane with bit kB clear keeps the
// low kB bits of each group and takes the partner's low bits into the high
// ones; a lane with it set keeps the high bits and takes the partner's high
// bits into the low ones.  One rotation of y and one bit select.
template <int kB, uint32_t kHigh>
__device__ __forceinline__ uint32_t transpose_merge(uint32_t x, uint32_t lane) {
    const uint32_t y = lane_xor<kB>(x);
    const bool up = (lane & kB) == 0;
    const uint32_t s = __builtin_amdgcn_alignbit(y, y, up ? 32 - kB : kB);   // y rotated toward the kept blocks
    const uint32_t take = up ? kHigh : ~kHigh;                               // bits that come from the partner
    return (s & take) | (x & ~take);
}

// Lane t holds column t of a 64 x 64 bit matrix (bit k = element (k, t));
// returns row `lane` (bit t = element (lane, t)).  Each stage exchanges one
// bit of the row index with the same bit of the lane index: the 32-bit
// halves across lanes 32 apart (v_permlane32_swap: lanes 32-63 of the first
// operand trade with lanes 0-31 of the second), the 16-bit quarters across
// lanes 16 apart (v_permlane16_swap: odd 16-lane rows of the first operand
// trade with even rows of the second, then a byte permute), the smaller
// blocks by DPP.  No LDS traffic and no branches.
__device__ __forceinline__ uint64_t wave_transpose(uint64_t x, uint32_t lane) {
    uint32_t lo = uint32_t(x), hi = uint32_t(x >> 32);
    {
        const auto r = __builtin_amdgcn_permlane32_swap(lo, hi, false, false);
        lo = r[0];
        hi = r[1];
    }
    {
        // per half: the lane with bit 16 clear ends with its own low quarter
        // and the partner's low quarter above it, the other lane with both
        // high quarters, low first; the rotated copy supplies the latter
        const auto a = __builtin_amdgcn_permlane16_swap(lo, __builtin_amdgcn_alignbit(lo, lo, 16), false, false);
        const auto b = __builtin_amdgcn_permlane16_swap(hi, __builtin_amdgcn_alignbit(hi, hi, 16), false, false);
        lo = __builtin_amdgcn_perm(a[1], a[0], 0x05040100u);
        hi = __builtin_amdgcn_perm(b[1], b[0], 0x05040100u);
    }
    lo = transpose_merge<8, 0xff00ff00u>(lo, lane);
    hi = transpose_merge<8, 0xff00ff00u>(hi, lane);
    lo = transpose_merge<4, 0xf0f0f0f0u>(lo, lane);
    hi = transpose_merge<4, 0xf0f0f0f0u>(hi, lane);
    lo = transpose_merge<2, 0xccccccccu>(lo, lane);
    hi = transpose_merge<2, 0xccccccccu>(hi, lane);
    lo = transpose_merge<1, 0xaaaaaaaau>(lo, lane);
    hi = transpose_merge<1, 0xaaaaaaaau>(hi, lane);
    return ((uint64_t)hi << 32) | lo;
}

// One 64-bit word of the residue bitsets: wave builds 64 rows (residues
// r0..r0+63 of v mod q) of word w; lane t tests u index 64*w + t and a wave
// ballot packs the 64 tests.  Row r, bit t: allowed_q[(u*B - r*A) mod q].
// chunk c of the k-th bitset prime starts at chunk_base[k].
__device__ __forceinline__ void build_mask_word(const MatchArgs& m, const uint32_t* chunk_base,
                                                const uint32_t* row_base, uint32_t c, uint32_t w, uint32_t lane) {
    uint32_t k = 0;
    while (k + 1 < m.nbits && c >= chunk_base[k + 1]) ++k;
    const Modulus mod = m.mods[k];
    const uint32_t r0 = (c - chunk_base[k]) * kWave;
    const uint32_t i = w * kWave + lane;
    const bool valid = i < m.nu;
    // residue for row r0: (u*B - r0*A) mod q, then step by -A mod q per row
    uint32_t res = 0;
    if (valid) {
        uint32_t rho = mod32(mod64(m.us[i], mod) * mod.b, mod);
        uint32_t sub = mod32(r0 * mod.a, mod);   // r0 < 2^16, a < 2^16
        res = rho >= sub ? rho - sub : rho + mod.q - sub;
    }
    uint64_t mine = 0;
    const uint32_t rows = mod.q - r0 < (uint32_t)kWave ? mod.q - r0 : (uint32_t)kWave;
    for (uint32_t rr = 0; rr < rows; ++rr) {
        const bool ok = valid && allowed_bit(m.allow, mod.bit + res);
        const uint64_t word = __ballot(ok);
        if (lane == rr) mine = word;
        res = res >= mod.a ? res - mod.a : res + mod.q - mod.a;
    }
    if (lane < rows) ((uint64_t*)m.masks)[(uint64_t)(row_base[k] + r0 + lane) * m.words + w] = mine;
}

__global__ __launch_bounds__(256) void k_masks(MatchArgs m, uint32_t chunks_total, uint32_t chunk_base1,
                                               uint32_t chunk_base2, uint32_t chunk_base3) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / kWave;
    if (wave >= (uint64_t)chunks_total * m.words) return;
    const uint32_t base[kMaxBitsetFilters] = {0, chunk_base1, chunk_base2, chunk_base3};
    build_mask_word(m, base, m.row_base, uint32_t(wave / m.words), uint32_t(wave % m.words), lane);
}

// rho[i][c] = (u_i * B) mod q for the cached chain filters c = 0..ncache-1.
// Rows are kCachedResidues entries wide (16 bytes) whatever ncache is, so the
// chain reads one u coordinate's residues with a single vector load.  One
// thread builds u_i's whole row (the moduli are the same for every lane).
__device__ __forceinline__ void build_rho(const MatchArgs& m, uint32_t i) {
    const uint64_t u = m.us[i];
    uint32_t packed[kCachedResidues / 2] = {};
#pragma unroll
    for (uint32_t c = 0; c < (uint32_t)kCachedResidues; ++c) {
        if (c < m.ncache) {
            const Modulus mod = m.mods[m.nbits + c];
            packed[c / 2] |= mod32(mod64(u, mod) * mod.b, mod) << (16 * (c & 1));
        }
    }
    ((uint4*)m.rho)[i] = make_uint4(packed[0], packed[1], packed[2], packed[3]);
}

__global__ void k_rho(MatchArgs m) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < m.nu) build_rho(m, (uint32_t)i);
}

constexpr int kMatchThreads = 256;
constexpr int kMatchWaves = kMatchThreads / kWave;
constexpr int kQueue = 2 * kWave;
constexpr int kWordPrefetch = 5;   // 64-word steps of one row whose loads are issued together
constexpr int kGroupSlots = kMaxResiduals / 4 + 2;
constexpr int kMaxConfigs = 3;     // filter configurations resident in one batched launch

// LDS of a matching block: filter tables of kConfigs configurations, and per
// wave a survivor queue and group counters.
template <int kConfigs>
struct MatchSharedT {
    uint32_t allow[kConfigs][kMaxAllowWords];
    Modulus mods[kConfigs][kMaxResiduals];
    uint32_t queue[kMatchWaves][kQueue];
    uint32_t groups[kMatchWaves][kGroupSlots];
};

// One wave's view of the LDS for one configuration.
struct Lds {
    const uint32_t* allow;
    const Modulus* mods;
    uint32_t* queue;
    uint32_t* groups;
};

template <int kConfigs>
__device__ void load_tables(MatchSharedT<kConfigs>& sh, uint32_t cfg, const uint32_t* allow, uint32_t allow_words,
                            const Modulus* mods, uint32_t nres) {
    for (uint32_t t = threadIdx.x; t < allow_words; t += blockDim.x) sh.allow[cfg][t] = allow[t];
    for (uint32_t t = threadIdx.x; t < nres; t += blockDim.x) sh.mods[cfg][t] = mods[t];
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t x) {
    for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
    return x;
}

// Flush one wave's counts for a subtile into the global counters.
__device__ void flush_tile(const MatchArgs& m, const Lds& L, uint32_t lane, uint32_t tile, uint32_t& valid,
                           uint64_t& survivors) {
    const uint64_t total = wave_sum(valid);
    uint64_t* cnt = m.counters + (uint64_t)tile * (m.ng + 3);
    if (lane == 0) {
        if (total) atomicAdd(&cnt[0], total);
        if (survivors) atomicAdd(&cnt[1], survivors);
    }
    for (uint32_t g = 2 + lane; g <= m.ng; g += kWave) {
        const uint32_t n = L.groups[g];
        if (n) atomicAdd(&cnt[g], (uint64_t)n);
        L.groups[g] = 0;
    }
    valid = 0;
    survivors = 0;
    wave_sync();
}

// Survivor output: the subtile's slot counter decides whether the value is
// kept (slot < cap, pseudocube/batch_experiment.py:229-230); kept values go
// to one shared list tagged with their subtile.
__device__ __forceinline__ void emit(const MatchArgs& m, uint32_t tile, const U128& x) {
    uint64_t* cnt = m.counters + (uint64_t)tile * (m.ng + 3);
    const uint64_t slot = atomicAdd(&cnt[m.ng + 1], 1ull);
    if (slot < m.cap) {
        const uint64_t g = atomicAdd(m.list_count, 1ull);
        if (g < m.list_cap) {
            uint64_t* dst = m.list + g * 3;
            dst[0] = m.tile_base + tile;
            dst[1] = x.lo;
            dst[2] = x.hi;
        }
    }
}

// Run up to 64 queued bitset survivors of one v row through the rest of the
// filter chain.  Lane l takes queue entry l.  Filters nbits..nbits+ncache-1
// use the per-u residue table plus the per-row offset sigma; later filters
// build x = u*B - v*A exactly and reduce it.  Group ends count survivors
// with a wave ballot.
__device__ void run_chain(const MatchArgs& m, const Lds& L, uint32_t lane, uint32_t count, const Row& row,
                          const U128& shift, uint32_t sigma) {
    bool alive = lane < count;
    const uint32_t i = alive ? L.queue[lane] : 0;
    uint32_t k = m.nbits;
    // all cached residues of this u in one 16-byte load
    uint4 packed = make_uint4(0, 0, 0, 0);
    if (alive && m.ncache) packed = ((const uint4*)m.rho)[i];
    const uint32_t words[4] = {packed.x, packed.y, packed.z, packed.w};
#pragma unroll
    for (int c = 0; c < kCachedResidues; ++c) {
        if ((uint32_t)c >= m.ncache) break;
        const Modulus mod = L.mods[k];
        const uint32_t s = __builtin_amdgcn_readlane(sigma, c);
        if (alive) {
            uint32_t r = ((words[c >> 1] >> (16 * (c & 1))) & 0xffffu) + s;
            r = r >= mod.q ? r - mod.q : r;
            alive = allowed_bit(L.allow, mod.bit + r);
        }
        const uint64_t live = __ballot(alive);
        if ((k & 3) == 3 || k + 1 == m.nres) {
            if (lane == 0 && live) L.groups[k / 4 + 1] += __popcll(live);
        }
        if (!live) return;
        ++k;
    }
    // the filters on exact x: each surviving lane's x in turn is tested by the
    // whole wave, lane l taking filter k + l, 64 filters per round; the first
    // failing filter ends it, and each group closed before it gains one (the
    // group ends of one round are distinct counters of this wave)
    U128 x = {0, 0};
    if (alive) x = sub_wide(mul_wide(m.us[i], m.B), shift);
    for (uint64_t pending = __ballot(alive); pending; pending &= pending - 1) {
        const int src = __builtin_ctzll(pending);
        const U128 xs = {readlane64(x.lo, src), readlane64(x.hi, src)};
        bool passed = true;
        for (uint32_t k0 = k; k0 < m.nres; k0 += kWave) {
            const uint32_t f = k0 + lane;
            bool pass = true;
            if (f < m.nres) {
                const Modulus mod = L.mods[f];
                pass = allowed_bit(L.allow, mod.bit + mod128(xs, mod));
            }
            const uint64_t fails = __ballot(!pass);
            const uint32_t upto = fails ? (uint32_t)__builtin_ctzll(fails) : (uint32_t)kWave;
            if (lane < upto && f < m.nres && ((f & 3) == 3 || f + 1 == m.nres)) L.groups[f / 4 + 1] += 1;
            wave_sync();
            if (fails) {
                passed = false;
                break;
            }
        }
        if (passed && lane == 0) {
            if (!less_wide(m.lo, xs) || less_wide(m.hi, xs)) atomicOr(m.error, 2u);
            emit(m, row.tile, xs);
        }
    }
}

// Rows [j0, j1) of one window through the bitset engine, by one wave.
__device__ void match_rows_masked(const MatchArgs& m, const Lds& L, uint32_t lane, uint64_t j0, uint64_t j1) {
    uint32_t tile = 0xffffffffu, valid = 0;
    uint64_t survivors = 0;
    for (uint64_t j = j0; j < j1; ++j) {
        const Row row = m.rows[j];
        if (row.tile != tile) {
            if (tile != 0xffffffffu) flush_tile(m, L, lane, tile, valid, survivors);
            tile = row.tile;
        }
        if (row.left >= row.right) continue;
        const U128 shift = mul_wide(row.v, m.A);
        // sigma_c = (-v*A) mod q for the cached chain filters, one per lane
        uint32_t sigma = 0;
        if (lane < m.ncache) {
            const Modulus mod = L.mods[m.nbits + lane];
            const uint32_t r = mod32(mod64(row.v, mod) * mod.a, mod);
            sigma = r ? mod.q - r : 0;
        }
        const uint64_t* mrow[kMaxBitsetFilters];
        for (uint32_t k = 0; k < kMaxBitsetFilters; ++k)
            mrow[k] = m.masks + (uint64_t)(m.row_base[k] + (k < m.nbits ? row.r[k] : 0)) * m.words;
        const uint32_t wl = row.left >> 6, wr = (row.right + 63) >> 6;
        uint32_t queued = 0;
        for (uint32_t cbase = wl; cbase < wr; cbase += kWave * kWordPrefetch) {
            // all bitset words of up to kWordPrefetch steps are requested before any is used
            uint64_t chunk[kWordPrefetch];
#pragma unroll
            for (int step = 0; step < kWordPrefetch; ++step) {
                const uint32_t w = cbase + step * kWave + lane;
                uint64_t bits = 0;
                if (w < wr) {
                    bits = ~0ull;
                    const uint32_t base = w << 6;
                    if (base < row.left) bits &= ~0ull << (row.left - base);
                    if (row.right < base + 64) bits &= (1ull << (row.right - base)) - 1ull;
                    valid += __popcll(bits);
#pragma unroll
                    for (int k = 0; k < kMaxBitsetFilters; ++k)
                        if ((uint32_t)k < m.nbits) bits &= mrow[k][w];
                }
                chunk[step] = bits;
            }
#pragma unroll
            for (int step = 0; step < kWordPrefetch; ++step) {
                if (cbase + step * kWave >= wr) break;
                const uint32_t w = cbase + step * kWave + lane;
                uint64_t bits = chunk[step];
                // wave-level compaction of the surviving u indices into the queue
                for (;;) {
                    const bool has = bits != 0;
                    const uint64_t ballot = __ballot(has);
                    if (!ballot) break;
                    if (has) {
                        L.queue[queued + lanes_below(ballot)] = (w << 6) + __builtin_ctzll(bits);
                        bits &= bits - 1;
                    }
                    queued += __popcll(ballot);
                    if (queued >= kWave) {
                        wave_sync();
                        survivors += kWave;
                        run_chain(m, L, lane, kWave, row, shift, sigma);
                        wave_sync();
                        const uint32_t rest = queued - kWave;
                        uint32_t moved = 0;
                        if (lane < rest) moved = L.queue[kWave + lane];
                        wave_sync();
                        if (lane < rest) L.queue[lane] = moved;
                        queued = rest;
                        wave_sync();
                    }
                }
            }
        }
        if (queued) {
            wave_sync();
            survivors += queued;
            run_chain(m, L, lane, queued, row, shift, sigma);
            wave_sync();
        }
    }
    if (tile != 0xffffffffu) flush_tile(m, L, lane, tile, valid, survivors);
}

// Rows [j0, j1) through the direct engine: every position of every row goes
// through the whole chain with x = u*B - v*A built in 128 bits and reduced
// limb by limb (no bitsets, no residue tables).
__device__ void match_rows_direct(const MatchArgs& m, const Lds& L, uint32_t lane, uint64_t j0, uint64_t j1) {
    uint32_t tile = 0xffffffffu, valid = 0;
    uint64_t survivors = 0;
    for (uint64_t j = j0; j < j1; ++j) {
        const Row row = m.rows[j];
        if (row.tile != tile) {
            if (tile != 0xffffffffu) flush_tile(m, L, lane, tile, valid, survivors);
            tile = row.tile;
        }
        const U128 shift = mul_wide(row.v, m.A);
        for (uint32_t base = row.left; base < row.right; base += kWave) {
            const uint32_t i = base + lane;
            bool alive = i < row.right;
            U128 x = {0, 0};
            if (alive) {
                const U128 positive = mul_wide(m.us[i], m.B);
                if (less_wide(positive, shift)) {
                    atomicAdd(&m.counters[(uint64_t)row.tile * (m.ng + 3) + m.ng + 2], 1ull);
                    alive = false;
                } else {
                    x = sub_wide(positive, shift);
                    ++valid;
                }
            }
            for (uint32_t k = 0; k < m.nres; ++k) {
                const Modulus mod = L.mods[k];
                if (alive) alive = allowed_bit(L.allow, mod.bit + mod128(x, mod));
                const uint64_t live = __ballot(alive);
                if ((k & 3) == 3 || k + 1 == m.nres) {
                    const uint32_t g = k / 4 + 1;
                    if (g == 1) survivors += __popcll(live);
                    else if (lane == 0 && live) L.groups[g] += __popcll(live);
                }
                if (!live) break;
            }
            if (alive) {
                if (!less_wide(m.lo, x) || less_wide(m.hi, x)) atomicOr(m.error, 2u);
                emit(m, row.tile, x);
            }
        }
    }
    if (tile != 0xffffffffu) flush_tile(m, L, lane, tile, valid, survivors);
}

// Masked matcher for one batch.  One wave per v row at a time; each wave owns
// a contiguous run of rows, and blocks are renumbered so that each XCD
// (blocks are dealt round-robin over 8 XCDs) owns one contiguous eighth of
// the rows and so one compact band of mask columns in its L2.
__global__ __launch_bounds__(kMatchThreads) void k_match(MatchArgs m) {
    __shared__ MatchSharedT<1> sh;
    load_tables(sh, 0, m.allow, m.allow_words, m.mods, m.nres);
    for (uint32_t t = threadIdx.x; t < kMatchWaves * kGroupSlots; t += kMatchThreads) (&sh.groups[0][0])[t] = 0;
    __syncthreads();
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const Lds L{sh.allow[0], sh.mods[0], sh.queue[wave], sh.groups[wave]};
    const uint32_t per_xcd = gridDim.x / 8;
    const uint32_t logical = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    const uint64_t first_row = ((uint64_t)logical * kMatchWaves + wave) * m.rows_per_wave;
    uint64_t last_row = first_row + m.rows_per_wave;
    if (last_row > m.nv) last_row = m.nv;
    match_rows_masked(m, L, lane, first_row, last_row);
}

__global__ __launch_bounds__(kMatchThreads) void k_match_direct(MatchArgs m) {
    __shared__ MatchSharedT<1> sh;
    load_tables(sh, 0, m.allow, m.allow_words, m.mods, m.nres);
    for (uint32_t t = threadIdx.x; t < kMatchWaves * kGroupSlots; t += kMatchThreads) (&sh.groups[0][0])[t] = 0;
    __syncthreads();
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const Lds L{sh.allow[0], sh.mods[0], sh.queue[wave], sh.groups[wave]};
    const uint64_t first_row = ((uint64_t)blockIdx.x * kMatchWaves + wave) * m.rows_per_wave;
    uint64_t last_row = first_row + m.rows_per_wave;
    if (last_row > m.nv) last_row = m.nv;
    match_rows_direct(m, L, lane, first_row, last_row);
}

// --------------------------------------------------------------------------
// Batched form: many windows of uneven length and mixed configuration in one
// pass.  Every kernel has a fixed grid and reads its work totals from device
// memory, so the whole launch sequence can be captured once in a hipGraph and
// replayed with new window descriptors.

struct BatchConfig {       // one filter configuration (a libpcube matcher)
    const Modulus* mods;
    const uint32_t* allow;
    const uint32_t* pstr;      // permuted allowed strings P[j] = allowed_q[(-j*A) mod q] of the bitset filters
    uint32_t nres, nbits, ncache, ng, allow_words, mask_rows;
    uint32_t transpose;        // every bitset filter has its permuted string: masks by wave transpose
    uint32_t row_base[kMaxBitsetFilters];
    uint32_t chunk_base[kMaxBitsetFilters + 1];   // 64-row mask chunks per bitset prime, prefix
    uint32_t pstr_off[kMaxBitsetFilters], nbeta[kMaxBitsetFilters];   // string offset, (-B/A) mod q
};

// Words w0 .. w0+7 of every row of bitset filter k, row-major batch layout, by the
// method of the staged builder: for word w lane t takes the u at index 64w + t,
// and for the 64 rows r0..r0+63 of each chunk the 64-bit window P[k0 .. k0+63],
// k0 = (r0 + (u mod q) * nbeta) mod q, whose bit j is allowed_q[(u*B - (r0 + j)*A)
// mod q]; a wave transpose turns the 64 columns into the 64 row words (u indices
// past nu give zero bits, as in build_mask_word).  Lane j then stores its row's
// eight words as one 64-byte run, so every store fills whole segments.
constexpr uint32_t kMaskRun = 8;   // words one batched mask item builds per row
__device__ __forceinline__ void build_mask_filter_run(const MatchArgs& m, const BatchConfig& c, uint32_t k,
                                                      uint32_t w0, uint32_t lane) {
    const Modulus mod = m.mods[k];
    const uint32_t* P = c.pstr + c.pstr_off[k];
    uint32_t beta[kMaskRun];
    bool live[kMaskRun];
#pragma unroll
    for (uint32_t ww = 0; ww < kMaskRun; ++ww) {
        const uint32_t i = (w0 + ww) * kWave + lane;
        live[ww] = i < m.nu;
        beta[ww] = live[ww] ? mod32(mod64(m.us[i], mod) * c.nbeta[k], mod) : 0u;
    }
    const uint32_t nwords = min(kMaskRun, m.words - w0);
    for (uint32_t r0 = 0; r0 < mod.q; r0 += kWave) {
        uint64_t row[kMaskRun];
#pragma unroll
        for (uint32_t ww = 0; ww < kMaskRun; ++ww) {
            // bits k0 .. k0 + 63 of P by two funnel shifts (k0 < q: the three
            // words lie in the string and its margin)
            uint32_t k0 = beta[ww] + r0;
            k0 = k0 >= mod.q ? k0 - mod.q : k0;
            const uint32_t a = k0 >> 5, sh = k0 & 31;
            const uint32_t p0 = P[a], p1 = P[a + 1], p2 = P[a + 2];
            const uint64_t col = ((uint64_t)__builtin_amdgcn_alignbit(p2, p1, sh) << 32) |
                                 __builtin_amdgcn_alignbit(p1, p0, sh);
            row[ww] = wave_transpose(live[ww] ? col : 0ull, lane);
        }
        if (r0 + lane < mod.q) {
            // row_base from the configuration (device memory) and constant indices into
            // row: a run-time index would put the arguments or the row in scratch memory
            uint64_t* out = (uint64_t*)m.masks + (uint64_t)(c.row_base[k] + r0 + lane) * m.words + w0;
            if (nwords == kMaskRun && ((uintptr_t)out & 15) == 0) {
#pragma unroll
                for (uint32_t ww = 0; ww < kMaskRun; ww += 2)
                    *(uint4*)(out + ww) = make_uint4(uint32_t(row[ww]), uint32_t(row[ww] >> 32), uint32_t(row[ww + 1]),
                                                     uint32_t(row[ww + 1] >> 32));
            } else {
#pragma unroll
                for (uint32_t ww = 0; ww < kMaskRun; ++ww)
                    if (ww < nwords) out[ww] = row[ww];
            }
        }
    }
}

struct BatchWindow {       // one window: its coordinates, tables and counters
    const uint64_t* us;
    const uint64_t* vs;
    uint64_t* masks;
    uint16_t* rho;
    uint64_t* counters;
    uint64_t v_first, tile_width, tile_base;
    uint64_t row_begin;    // first global row
    uint64_t mask_begin;   // first global mask wave item
    uint64_t rho_begin;    // first global residue-cache item
    uint64_t chunk_begin;  // first global row chunk
    uint32_t nu, nv, words, ntiles, cfg, mode;   // mode 0 bitset engine, 1 direct engine
};

struct BatchTotals {       // device resident, rewritten before each launch
    uint64_t rows, mask_waves, rho_items, chunks, counters;
    uint32_t nwin, rows_per_chunk;
};

struct BatchArgs {
    const BatchWindow* win;
    const BatchTotals* totals;
    Row* rows;
    uint64_t* list;
    uint64_t* list_count;
    uint64_t* counters;        // all windows' counters, totals->counters words
    uint32_t* error;
    uint32_t* cursor;          // persistent work-queue cursor
    uint64_t A, B, cap, list_cap;
    U128 B_mu;                 // floor((2^128 - 1) / B)
    U128 lo, hi;
    uint32_t ncfg;
    BatchConfig cfg[kMaxConfigs];   // host copy (the kernels index dcfg: a run-time index into
                                    // kernel arguments would copy them to scratch memory)
    const BatchConfig* dcfg;        // the same configurations in device memory
};

// last window w with field(w) <= key (fields ascending over windows)
template <class F>
__device__ __forceinline__ uint32_t find_window(const BatchWindow* win, uint32_t nwin, uint64_t key, F field) {
    uint32_t l = 0, h = nwin;
    while (h - l > 1) {
        const uint32_t mid = (l + h) >> 1;
        if (field(win[mid]) <= key) l = mid; else h = mid;
    }
    return l;
}

__device__ __forceinline__ MatchArgs window_args(const BatchArgs& b, const BatchWindow& w) {
    const BatchConfig& c = b.dcfg[w.cfg];
    MatchArgs m{};   // every field the batched engine does not set is zero (staged = 0: plain row records)
    m.us = w.us; m.vs = w.vs;
    m.rows = b.rows + w.row_begin;
    m.masks = w.masks; m.rho = w.rho;
    m.mods = c.mods; m.allow = c.allow;
    m.counters = w.counters;
    m.list = b.list; m.list_count = b.list_count; m.error = b.error;
    m.A = b.A; m.B = b.B; m.B_mu = b.B_mu; m.v_first = w.v_first; m.tile_width = w.tile_width; m.cap = b.cap;
    m.list_cap = b.list_cap; m.tile_base = w.tile_base;
    m.lo = b.lo; m.hi = b.hi;
    m.nu = w.nu; m.nv = w.nv; m.words = w.words;
    m.nres = c.nres; m.nbits = c.nbits; m.ncache = c.ncache; m.ng = c.ng;
    m.ntiles = w.ntiles; m.allow_words = c.allow_words; m.rows_per_wave = 0;
    for (int k = 0; k < kMaxBitsetFilters; ++k) m.row_base[k] = c.row_base[k];
    return m;
}

__global__ void k_batch_fill(BatchArgs b) {
    const uint64_t n = b.totals->counters;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        b.counters[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *b.list_count = 0;
        *b.error = 0;
        *b.cursor = 0;
    }
}

__global__ void k_batch_rows(BatchArgs b) {
    const BatchTotals t = *b.totals;
    for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < t.rows;
         g += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t w = find_window(b.win, t.nwin, g, [](const BatchWindow& x) { return x.row_begin; });
        const BatchWindow win = b.win[w];
        const MatchArgs m = window_args(b, win);
        const uint32_t j = uint32_t(g - win.row_begin);
        build_row<false>(m, j, m.vs[j]);
    }
}

// Mask items: per window, (run of kMaskRun words, filter) pairs when its
// configuration builds by transpose (every row of the filter), else (64-row chunk, word)
// pairs.  Each wave takes one contiguous run of the items, so it finds its first
// window once and then walks the windows in order.
__global__ __launch_bounds__(256) void k_batch_masks(BatchArgs b) {
    const BatchTotals t = *b.totals;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / kWave);
    const uint64_t per = (t.mask_waves + waves - 1) / waves;
    uint64_t g = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / kWave * per;
    const uint64_t stop = min(t.mask_waves, g + per);
    if (g >= stop) return;
    uint32_t w = find_window(b.win, t.nwin, g, [](const BatchWindow& x) { return x.mask_begin; });
    while (g < stop && w < t.nwin) {
        const BatchWindow win = b.win[w];
        const uint64_t next = w + 1 < t.nwin ? b.win[w + 1].mask_begin : t.mask_waves;
        const MatchArgs m = window_args(b, win);
        const BatchConfig& c = b.dcfg[win.cfg];
        for (; g < stop && g < next; ++g) {
            const uint64_t local = g - win.mask_begin;
            if (c.transpose)
                build_mask_filter_run(m, c, uint32_t(local % m.nbits), uint32_t(local / m.nbits) * kMaskRun, lane);
            else
                build_mask_word(m, c.chunk_base, c.row_base, uint32_t(local / win.words), uint32_t(local % win.words),
                                lane);
        }
        ++w;
    }
}

// Residue cache rows: one thread per u coordinate writes its kCachedResidues
// residues (u*B) mod q as one 16-byte row.
__global__ void k_batch_rho(BatchArgs b) {
    const BatchTotals t = *b.totals;
    for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < t.rho_items;
         g += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t w = find_window(b.win, t.nwin, g, [](const BatchWindow& x) { return x.rho_begin; });
        const BatchWindow win = b.win[w];
        const BatchConfig& c = b.dcfg[win.cfg];
        const uint64_t i = g - win.rho_begin;
        const uint64_t u = win.us[i];
        uint32_t packed[kCachedResidues / 2] = {};
#pragma unroll
        for (int k = 0; k < kCachedResidues; ++k) {
            if ((uint32_t)k < c.ncache) {
                const Modulus mod = c.mods[c.nbits + k];
                packed[k / 2] |= mod32(mod64(u, mod) * mod.b, mod) << (16 * (k & 1));
            }
        }
        ((uint4*)win.rho)[i] = make_uint4(packed[0], packed[1], packed[2], packed[3]);
    }
}

// Persistent matcher: each wave takes chunks of rows_per_chunk rows of one
// window from the device queue until the queue is exhausted.
__global__ __launch_bounds__(kMatchThreads) void k_batch_match(BatchArgs b) {
    __shared__ MatchSharedT<kMaxConfigs> sh;
    for (uint32_t c = 0; c < b.ncfg; ++c)
        load_tables(sh, c, b.dcfg[c].allow, b.dcfg[c].allow_words, b.dcfg[c].mods, b.dcfg[c].nres);
    for (uint32_t t = threadIdx.x; t < kMatchWaves * kGroupSlots; t += kMatchThreads) (&sh.groups[0][0])[t] = 0;
    __syncthreads();
    const BatchTotals t = *b.totals;
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    for (;;) {
        uint32_t chunk = 0;
        if (lane == 0) chunk = atomicAdd(b.cursor, 1u);
        chunk = __builtin_amdgcn_readfirstlane(chunk);
        if (chunk >= t.chunks) break;
        const uint32_t w = find_window(b.win, t.nwin, chunk, [](const BatchWindow& x) { return x.chunk_begin; });
        const BatchWindow win = b.win[w];
        const MatchArgs m = window_args(b, win);
        const Lds L{sh.allow[win.cfg], sh.mods[win.cfg], sh.queue[wave], sh.groups[wave]};
        const uint64_t j0 = (uint64_t)(chunk - win.chunk_begin) * t.rows_per_chunk;
        const uint64_t j1 = min<uint64_t>(j0 + t.rows_per_chunk, win.nv);
        if (win.mode == 0) match_rows_masked(m, L, lane, j0, j1);
        else match_rows_direct(m, L, lane, j0, j1);
    }
}

// Plain streaming copy of n uint64 words: the calibration probe for the HBM
// byte counters (FETCH_SIZE / WRITE_SIZE read for a known traffic).
__global__ void k_copy_probe(const uint64_t* __restrict__ src, uint64_t* __restrict__ dst, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Wide-limb self test: out[3i..3i+2] = low, high, borrow of in[4i]*in[4i+1] - in[4i+2]*in[4i+3].
// in[5i..5i+4] = (x.lo, x.hi, B, mu.lo, mu.hi) with mu = floor((2^128 - 1) / B);
// out[2i..2i+1] = floor(x / B), the division the u-bound searches use.
__global__ void k_floor_div(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, uint64_t count) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= count) return;
    const U128 q = floor_div_wide(U128{in[5 * i], in[5 * i + 1]}, in[5 * i + 2], U128{in[5 * i + 3], in[5 * i + 4]});
    out[2 * i] = q.lo;
    out[2 * i + 1] = q.hi;
}

// One 64 x 64 bit matrix per wave: in[64i + t] is column t of matrix i,
// out[64i + r] its row r (the transpose the bitset mask builders apply).
__global__ void k_bit_transpose(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, uint64_t count) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i / kWave >= count) return;   // whole waves: count matrices are count * 64 lanes
    out[i] = wave_transpose(in[i], (uint32_t)(i & (kWave - 1)));
}

__global__ void k_arithmetic(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, uint64_t count) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= count) return;
    const U128 left = mul_wide(in[4 * i], in[4 * i + 1]), right = mul_wide(in[4 * i + 2], in[4 * i + 3]);
    const U128 diff = sub_wide(left, right);
    out[3 * i] = diff.lo;
    out[3 * i + 1] = diff.hi;
    out[3 * i + 2] = less_wide(left, right) ? 1 : 0;
}

}  // namespace pcube
