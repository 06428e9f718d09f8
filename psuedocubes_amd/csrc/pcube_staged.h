#pragma once
// LDS-staged residue-bitset engine for one batch (gfx950).
//
// The first kStagedFilters residual filters (193, 199, 211, 223, 229, 241,
// 271, 277 at p = 619) are folded into residue bitsets: for filter q and
// residue class r = v mod q, bit t of mask word w is allowed_q[(u*B - r*A)
// mod q] for the u at index 64w + t.  Masks are laid out in chunks of
// kStageWords words, [chunk][sum of q rows][kStageWords], so one chunk of every
// filter is one contiguous block of stage_rows * 32 bytes.
//
// Work item = (stage chunk s, piece of kPieceRows consecutive v rows).  One
// persistent block per CU walks a list of items; for each it intersects the
// chunk, staged in LDS, with every row of the piece: a thread takes whole rows
// and ANDs, for each of the chunk's words, the kStagedFilters words of its
// row's residue classes.  The next item's chunk streams into a second LDS
// buffer (global_load_lds) while the current one is matched.  Positions
// surviving all bitset filters (about 1 in 6561 at p = 619) are written to the
// item's region of a candidate list, and the block runs them through the rest
// of the filter chain after its items, exactly as in the residue-bitset engine: cached
// residues (u*B mod q) + sigma_v for the next kCachedResidues filters, then
// exact 128-bit x = u*B - v*A reduced limb by limb.
//
// Counts follow the reference's group structure (filters in groups of four,
// pseudocube/batch_experiment.py:198-235): group 1 and group 2 are popcounts
// of the bitset intersections after filters 3 and 7, later groups are counted
// in the chain; valid positions per subtile are summed per row by the row
// blocks of k_prefilter_staged.  Counts and the emitted set equal the
// residue-bitset engine's.
#include "pcube_device.h"

namespace pcube {

constexpr int kStageWords = 4;        // mask words (256 u coordinates) per stage chunk
constexpr int kPieceRows = 2048;      // v rows per work item
constexpr int kStagedThreads = 1024;  // one persistent block per CU
constexpr int kStagedWaves = kStagedThreads / kWave;
constexpr int kRowsPerLane = kPieceRows / kStagedThreads;   // rows of a piece each thread takes
constexpr int kTileSlots = 64;        // subtiles whose group counts a block keeps in LDS
constexpr int kMaxStageRows = 2048;   // bitset rows (sum of q) one chunk may hold

// A staged row is two 16-byte halves (words 0-1, words 2-3).  Rows whose bit 3
// is set store them swapped, so that rows 8 apart start in different LDS bank
// quads and a wave reading 64 rows spreads over all 16 quads.
__host__ __device__ __forceinline__ uint32_t stage_swizzle(uint32_t row) { return (row >> 3) & 1u; }

// One butterfly stage of a 64 x 64 bit transpose across a wave.
template <int kB, uint64_t kMask>
__device__ __forceinline__ uint64_t transpose_step(uint64_t x, uint32_t lane) {
    const uint64_t y = __shfl_xor(x, kB, kWave);
    return (lane & kB) ? ((x & kMask) | ((y & kMask) >> kB)) : ((x & ~kMask) | ((y & ~kMask) << kB));
}

// Lane t holds column t of a 64 x 64 bit matrix (bit k = element (k, t));
// returns row `lane` (bit t = element (lane, t)).
__device__ __forceinline__ uint64_t wave_transpose(uint64_t x, uint32_t lane) {
    x = transpose_step<32, 0xffffffff00000000ull>(x, lane);
    x = transpose_step<16, 0xffff0000ffff0000ull>(x, lane);
    x = transpose_step<8, 0xff00ff00ff00ff00ull>(x, lane);
    x = transpose_step<4, 0xf0f0f0f0f0f0f0f0ull>(x, lane);
    x = transpose_step<2, 0xccccccccccccccccull>(x, lane);
    x = transpose_step<1, 0xaaaaaaaaaaaaaaaaull>(x, lane);
    return x;
}

// Staged residue bitsets.  Wave (s, c) builds rows r0..r0+63 of one bitset
// filter for the kStageWords words of stage chunk s.  Lane t takes the u at
// index 64w + t.  With P[k] = allowed_q[(-k*A) mod q] (stored repeated, q + 64
// bits) and k0 = (r0 + (u mod q) * nbeta) mod q, bit j of the 64-bit window
// P[k0 .. k0 + 63] is allowed_q[(u*B - (r0 + j)*A) mod q]: the column of this
// u over rows r0..r0+63.  A wave transpose turns the 64 columns into the 64
// row words.  chunk_base[k] = first 64-row chunk of filter k.
__device__ __forceinline__ void staged_mask_wave(const MatchArgs& m, uint64_t g, uint32_t chunks_total,
                                                 const uint32_t* base) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    if (g >= (uint64_t)chunks_total * m.nstage) return;
    const uint32_t s = uint32_t(g / chunks_total), c = uint32_t(g % chunks_total);
    uint32_t k = 0;
    while (k + 1 < m.nbits && c >= base[k + 1]) ++k;
    const Modulus mod = m.mods[k];
    const uint32_t r0 = (c - base[k]) * kWave;
    const uint32_t rows = mod.q - r0 < (uint32_t)kWave ? mod.q - r0 : (uint32_t)kWave;
    const uint32_t* P = m.pstr + m.pstr_off[k];
    uint64_t out[kStageWords];
#pragma unroll
    for (int ww = 0; ww < kStageWords; ++ww) {
        const uint32_t i = (s * kStageWords + ww) * kWave + lane;
        uint64_t col = 0;
        if (i < m.nu) {
            uint32_t k0 = mod32(mod64(m.us[i], mod) * m.nbeta[k], mod) + r0;
            k0 = k0 >= mod.q ? k0 - mod.q : k0;
            const uint32_t w0 = k0 >> 5, sh = k0 & 31;
            const uint64_t low = ((uint64_t)P[w0 + 1] << 32) | P[w0];
            col = (low >> sh) | (sh ? (uint64_t)P[w0 + 2] << (64 - sh) : 0ull);
        }
        out[ww] = wave_transpose(col, lane);
    }
    if (lane < rows) {
        const uint32_t row = m.row_base[k] + r0 + lane;
        const uint32_t swap = stage_swizzle(row);
        uint4* dst = (uint4*)((uint64_t*)m.masks + ((uint64_t)s * m.stage_rows + row) * kStageWords);
        dst[swap] = make_uint4(uint32_t(out[0]), uint32_t(out[0] >> 32), uint32_t(out[1]), uint32_t(out[1] >> 32));
        dst[swap ^ 1] = make_uint4(uint32_t(out[2]), uint32_t(out[2] >> 32), uint32_t(out[3]), uint32_t(out[3] >> 32));
    }
}

// Copy n 16-byte units from global memory to LDS with the whole block, up to
// kBatch loads per thread in flight before any store.
// Loads are unconditional (indices past n read element n - 1) so that the
// batch stays in registers and all its loads are in flight together.
template <int kThreads, int kBatch>
__device__ __forceinline__ void stage_copy(uint4* __restrict__ dst, const uint4* __restrict__ src, uint32_t n) {
    for (uint32_t first = 0; first < n; first += kThreads * kBatch) {
        uint4 v[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
            const uint32_t t = first + k * kThreads + threadIdx.x;
            v[k] = src[t < n ? t : n - 1];
        }
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
            const uint32_t t = first + k * kThreads + threadIdx.x;
            if (t < n) dst[t] = v[k];
        }
    }
}

// LDS of a staged block, carved from dynamic shared memory.
struct Staged {
    const uint32_t* allow;
    const Modulus* mods;
    uint32_t* tilecnt;       // [kTileSlots][ng]: groups 1..ng of subtiles t0..t0+63
    uint32_t* queue;         // this wave's survivors, (row offset << 8) | u offset within the item
};


// Add n to group g (1-based) of a subtile: LDS when the subtile is within the
// block's window, else the global counters.
__device__ __forceinline__ void staged_add(const MatchArgs& m, const Staged& S, uint32_t t0, uint32_t tile, uint32_t g,
                                           uint32_t n) {
    const uint32_t slot = tile - t0;
    if (slot < (uint32_t)kTileSlots) atomicAdd(&S.tilecnt[slot * m.ng + g - 1], n);
    else atomicAdd(&m.counters[(uint64_t)tile * (m.ng + 3) + g], (uint64_t)n);
}

// Run up to 64 bitset survivors (lane l takes row j, u index i, live when
// l < count) through the filters after the bitsets.
__device__ __forceinline__ void staged_chain(const MatchArgs& m, const Staged& S, uint32_t lane, uint32_t count,
                                             uint32_t j, uint32_t i) {
    bool alive = lane < count;
    if (!alive) j = i = 0;
    const Row row = m.rows[j];
    uint32_t k = m.nbits;
    uint4 packed = make_uint4(0, 0, 0, 0);
    if (alive && m.ncache) packed = ((const uint4*)m.rho)[i];
    const uint32_t words[4] = {packed.x, packed.y, packed.z, packed.w};
#pragma unroll
    for (int c = 0; c < kCachedResidues; ++c) {
        if ((uint32_t)c >= m.ncache) break;
        const Modulus mod = S.mods[k];
        if (alive) {
            uint32_t r = ((words[c >> 1] >> (16 * (c & 1))) & 0xffffu) + row.sigma[c];
            r = r >= mod.q ? r - mod.q : r;
            alive = allowed_bit(S.allow, mod.bit + r);
        }
        if (((k & 3) == 3 || k + 1 == m.nres) && alive) staged_add(m, S, 0, row.tile, k / 4 + 1, 1);
        if (!__ballot(alive)) return;
        ++k;
    }
    U128 x = {0, 0};
    if (alive) x = sub_wide(mul_wide(m.us[i], m.B), mul_wide(row.v, m.A));
    for (; k < m.nres; ++k) {
        const Modulus mod = S.mods[k];
        if (alive) alive = allowed_bit(S.allow, mod.bit + mod128(x, mod));
        if (((k & 3) == 3 || k + 1 == m.nres) && alive) staged_add(m, S, 0, row.tile, k / 4 + 1, 1);
        if (!__ballot(alive)) return;
    }
    if (alive) {
        if (!less_wide(m.lo, x) || less_wide(m.hi, x)) atomicOr(m.error, 2u);
        emit(m, row.tile, x);
    }
}

// Append count (<= 64) queued survivors to the work item's region of the
// candidate list; the item's fill counter lives in LDS.
__device__ __forceinline__ void cand_flush(const Staged& S, uint32_t lane, uint32_t count, uint32_t* region,
                                           uint32_t cap, uint32_t* fill) {
    uint32_t at = 0;
    if (lane == 0) at = atomicAdd(fill, count);
    at = __shfl(at, 0, kWave);
    if (lane < count && at + lane < cap) region[at + lane] = S.queue[lane];
}

// Queue the set bits of one mask word as (row offset << 8) | (u offset of the word + bit).
__device__ __forceinline__ void staged_enqueue(const Staged& S, uint32_t lane, uint64_t b, uint32_t tag,
                                               uint32_t& queued, uint32_t* region, uint32_t cap, uint32_t* fill) {
    for (;;) {
        const bool has = b != 0;
        const uint64_t ballot = __ballot(has);
        if (!ballot) return;
        if (has) {
            S.queue[queued + lanes_below(ballot)] = tag + (uint32_t)__builtin_ctzll(b);
            b &= b - 1;
        }
        queued += __popcll(ballot);
        if (queued >= kWave) {
            wave_sync();
            cand_flush(S, lane, kWave, region, cap, fill);
            wave_sync();
            const uint32_t rest = queued - kWave;
            uint32_t moved = 0;
            if (lane < rest) moved = S.queue[kWave + lane];
            wave_sync();
            if (lane < rest) S.queue[lane] = moved;
            queued = rest;
            wave_sync();
        }
    }
}

// Chunks [c_lo, c_hi] that the rows of piece p reach (none when c_lo > c_hi):
// rows are sorted by v, so from the u bounds of the piece's first and last
// rows, computed here exactly as the row builder does.
__device__ __forceinline__ void piece_chunks(const MatchArgs& m, uint32_t p, const uint64_t* dir, uint32_t ndir,
                                             uint32_t stride, uint32_t& c_lo, uint32_t& c_hi) {
    const uint32_t j0 = p * kPieceRows, j1 = min(m.nv, j0 + kPieceRows);
    const U128 lower = add_wide(mul_wide(m.vs[j0], m.A), m.lo);
    const U128 upper = add_wide(mul_wide(m.vs[j1 - 1], m.A), m.hi);
    const uint32_t left0 = first_above(m, lower, dir, ndir, stride);
    const uint32_t right1 = first_above(m, upper, dir, ndir, stride);
    c_lo = 1;
    c_hi = 0;
    if (right1 > left0) {
        c_lo = (left0 >> 6) / kStageWords;
        c_hi = min(((right1 - 1) >> 6) / kStageWords, m.nstage - 1);
    }
}

// Work list of the staged match: every (piece, chunk) pair whose rows reach
// the chunk, piece-major, at most items_cap of them; nitems receives the full
// count.  One 256-thread block: thread t takes a run of consecutive pieces, a
// block scan turns the runs' counts into offsets.
__device__ __forceinline__ void staged_plan_block(const MatchArgs& m, uint32_t npieces, uint2* items,
                                                  uint32_t items_cap, uint32_t* nitems, uint64_t* dir) {
    __shared__ uint32_t wave_total[256 / kWave];
    uint32_t ndir, stride;
    load_directory(m, dir, ndir, stride);
    const uint64_t* d = ndir ? dir : nullptr;
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const uint32_t per = (npieces + 255) / 256;
    const uint32_t p0 = min(npieces, threadIdx.x * per), p1 = min(npieces, p0 + per);
    uint32_t count = 0, lo1 = 1, hi1 = 0;   // the range of a one-piece run is kept
    for (uint32_t p = p0; p < p1; ++p) {
        uint32_t lo, hi;
        piece_chunks(m, p, d, ndir, stride, lo, hi);
        if (lo <= hi) count += hi - lo + 1;
        lo1 = lo;
        hi1 = hi;
    }
    uint32_t inc = count;
#pragma unroll
    for (int dd = 1; dd < kWave; dd <<= 1) {
        const uint32_t y = __shfl_up(inc, dd, kWave);
        if (lane >= (uint32_t)dd) inc += y;
    }
    if (lane == kWave - 1) wave_total[wave] = inc;
    __syncthreads();
    uint32_t at = inc - count;
    for (uint32_t w = 0; w < wave; ++w) at += wave_total[w];
    for (uint32_t p = p0; p < p1; ++p) {
        uint32_t lo = lo1, hi = hi1;
        if (p1 - p0 > 1) piece_chunks(m, p, d, ndir, stride, lo, hi);
        for (uint32_t c = lo; c <= hi && lo <= hi; ++c, ++at)
            if (at < items_cap) items[at] = make_uint2(p, c);
    }
    if (threadIdx.x == 255) *nitems = at;
}

// Geometry of the fused prefilter launch.
struct PrefilterArgs {
    uint32_t rows_blocks, mask_blocks, rho_blocks, chunks_total, npieces, items_cap;
    uint32_t chunk_base[kStagedFilters];
    uint2* items;
    uint32_t* nitems;
};

// Everything the staged match needs, in one launch of independent 256-thread
// blocks: block 0 plans the work items; then the row blocks (exact u bounds,
// mask row offsets, sigmas, valid counts), the mask waves and the residue
// cache.
__global__ __launch_bounds__(256) void k_prefilter_staged(MatchArgs m, PrefilterArgs pa) {
    __shared__ uint64_t dir[kRowDirectory];
    uint32_t b = blockIdx.x;
    if (b == 0) {
        staged_plan_block(m, pa.npieces, pa.items, pa.items_cap, pa.nitems, dir);
        return;
    }
    b -= 1;
    if (b < pa.rows_blocks) {
        staged_rows_block(m, b, dir);
        return;
    }
    b -= pa.rows_blocks;
    if (b < pa.mask_blocks) {
        staged_mask_wave(m, (uint64_t)b * (256 / kWave) + threadIdx.x / kWave, pa.chunks_total, pa.chunk_base);
        return;
    }
    b -= pa.mask_blocks;
    const uint64_t t = (uint64_t)b * 256 + threadIdx.x;
    if (t < (uint64_t)m.nu * kCachedResidues) build_rho(m, t);
}

// Row headers of one work item's rows for this thread: (left, right, tile,
// flags) and the LDS byte offsets of the filters' mask rows.
struct ItemRows {
    uint4 h0[kRowsPerLane], h1[kRowsPerLane];
};

__device__ __forceinline__ void load_item_rows(const MatchArgs& m, uint32_t p, ItemRows& R) {
    const uint32_t j0 = p * kPieceRows, j1 = min(m.nv, j0 + kPieceRows);
#pragma unroll
    for (int t = 0; t < kRowsPerLane; ++t) {
        const uint32_t j = j0 + threadIdx.x + kStagedThreads * t;
        const uint4* rp = (const uint4*)&m.rows[j < j1 ? j : j1 - 1];
        R.h0[t] = rp[0];
        R.h1[t] = rp[1];
    }
}

// Start the copy of stage chunk s into an LDS buffer: each wave moves 64
// sixteen-byte units per instruction straight from global memory to LDS.
// Complete once the wave's vector memory counter drains.
__device__ __forceinline__ void stage_async(const MatchArgs& m, uint32_t s, uint8_t* buffer, uint32_t wave,
                                            uint32_t lane) {
    const uint4* src = (const uint4*)(m.masks + (uint64_t)s * m.stage_rows * kStageWords);
    const uint32_t n = m.stage_rows * kStageWords / 2;
    for (uint32_t q = wave; q * kWave < n; q += kStagedWaves) {
        const uint32_t t = q * kWave + lane;
        if (t < n) __builtin_amdgcn_global_load_lds((const void*)(src + t), (void*)(buffer + q * kWave * 16), 16, 0, 0);
    }
}

// Static LDS of the staged match (one block per CU): two mask buffers, so the
// next item's chunk can stream in while the current one is read, plus the
// filter tables, per-subtile counters, survivor queues and the item's fill
// counter.  Distinct arrays let the compiler see that reads of one buffer do
// not wait for the direct-to-LDS copies into the other.
constexpr int kStageBufferUnits = kMaxStageRows * kStageWords * 8 / 16;
constexpr uint32_t kStagedLdsBytes = 2 * kStageBufferUnits * 16 + kMaxAllowWords * 4 + kMaxResiduals * 32 +
                                     kTileSlots * (kMaxResiduals / 4) * 4 + kStagedWaves * kQueue * 4 + 16;

struct ItemState {
    uint32_t item, total;
    uint2 it;
    ItemRows rows;
};

// Match one work item reading chunk buffer `cur` while the next item's chunk
// streams into `nxt`.  Returns false when the block has no further item.
template <int kF, bool kLast>
__device__ __forceinline__ bool staged_item(const MatchArgs& m, const uint2* items, const Staged& S, ItemState& st,
                                            const uint4 (&cur)[kStageBufferUnits], uint4 (&nxt)[kStageBufferUnits],
                                            uint32_t* fill, uint32_t lane, uint32_t wave) {
    const uint32_t p = st.it.x, s = st.it.y;
    const uint32_t j0 = p * kPieceRows, j1 = min(m.nv, j0 + kPieceRows);
    const uint32_t t0 = m.rows[j0].tile;
    __builtin_amdgcn_s_waitcnt(0);   // this wave's part of the chunk and its row headers have arrived
    if (threadIdx.x == 0) *fill = 0;
    uint32_t* region = m.cand + (uint64_t)st.item * m.item_cap;
    __syncthreads();
    // the next item's chunk and rows are in flight while this one is matched
    const uint32_t next = st.item + gridDim.x;
    const ItemRows cur_rows = st.rows;
    uint2 nit = st.it;
    if (next < st.total) {
        nit = items[next];
        load_item_rows(m, nit.x, st.rows);
        stage_async(m, nit.y, (uint8_t*)nxt, wave, lane);
    }
    const uint8_t* masks = (const uint8_t*)cur;
    const uint32_t base0 = (s * kStageWords) << 6;   // first u index of the chunk
    uint32_t tile_cur = 0xffffffffu, acc1 = 0, acc2 = 0;
    uint32_t queued = 0;
#pragma unroll
    for (int t = 0; t < kRowsPerLane; ++t) {
        const uint32_t j = j0 + threadIdx.x + kStagedThreads * t;
        if (j0 + (threadIdx.x - lane) + kStagedThreads * t < j1) {   // the wave has rows in the piece
            const bool live = j < j1;
            const uint4 h0 = cur_rows.h0[t], h1 = cur_rows.h1[t];
            const uint32_t tile = h0.z;
            if (live && tile != tile_cur) {
                if (tile_cur != 0xffffffffu) {
                    if (acc1) staged_add(m, S, t0, tile_cur, 1, acc1);
                    if (acc2) staged_add(m, S, t0, tile_cur, 2, acc2);
                }
                tile_cur = tile;
                acc1 = acc2 = 0;
            }
            // range of the row's u indices within the chunk, clamped to the chunk's words
            const int dl = live ? (int)(h0.x - base0) : 0, dr = live ? (int)(h0.y - base0) : 0;
            uint64_t b[kStageWords];
#pragma unroll
            for (int ww = 0; ww < kStageWords; ++ww) {
                const int lo = min(max(dl - 64 * ww, 0), 64), hi = min(max(dr - 64 * ww, 0), 64);
                b[ww] = hi > lo ? (~0ull << lo) & (~0ull >> (64 - hi)) : 0ull;
            }
            // each half (two words) of every filter's row is read before any is used
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                uint4 mw[kF];
#pragma unroll
                for (int k = 0; k < kF; ++k) {
                    const uint32_t pair = k < 2 ? h1.x : k < 4 ? h1.y : k < 6 ? h1.z : h1.w;
                    const uint32_t off = ((pair >> (16 * (k & 1))) & 0xffffu) ^ (16u * h);
                    mw[k] = *(const uint4*)(masks + off);
                }
#pragma unroll
                for (int k = 0; k < kF; ++k) {
                    b[2 * h] &= ((uint64_t)mw[k].y << 32) | mw[k].x;
                    b[2 * h + 1] &= ((uint64_t)mw[k].w << 32) | mw[k].z;
                    if ((k & 3) == 3 || (kLast && k == kF - 1)) {
                        const uint32_t n = __popcll(b[2 * h]) + __popcll(b[2 * h + 1]);
                        if (k < 4) acc1 += n;
                        else acc2 += n;
                    }
                }
            }
            const uint32_t tag = ((j - j0) << 8);
#pragma unroll
            for (int ww = 0; ww < kStageWords; ++ww)
                staged_enqueue(S, lane, b[ww], tag + 64 * ww, queued, region, m.item_cap, fill);
        }
    }
    if (queued) {
        wave_sync();
        cand_flush(S, lane, queued, region, m.item_cap, fill);
    }
    if (tile_cur != 0xffffffffu) {
        if (acc1) staged_add(m, S, t0, tile_cur, 1, acc1);
        if (acc2) staged_add(m, S, t0, tile_cur, 2, acc2);
    }
    __syncthreads();
    if (threadIdx.x == 0) m.item_count[st.item] = *fill;
    for (uint32_t t = threadIdx.x; t < (uint32_t)kTileSlots * m.ng; t += kStagedThreads) {
        const uint32_t n = S.tilecnt[t];
        if (n) {
            atomicAdd(&m.counters[(uint64_t)(t0 + t / m.ng) * (m.ng + 3) + t % m.ng + 1], (uint64_t)n);
            S.tilecnt[t] = 0;
        }
    }
    if (next >= st.total) return false;
    st.item = next;
    st.it = nit;
    return true;
}

// Persistent staged match: block b takes work items b, b + grid, ...  For an
// item (piece p, chunk s) thread l intersects the staged chunk with rows
// j0 + l + 1024 t of the piece; the next item's chunk and row headers are in
// flight meanwhile.  Group counts of the piece's subtiles collect in LDS and
// go to the global counters at the end of each item.  After its last item the
// block runs its items' survivors through the residual chain.  kF = bitset filters;
// kLast: the last residual filter is the last bitset filter (it closes a
// group).
template <int kF, bool kLast>
__global__ __launch_bounds__(kStagedThreads) void k_match_staged(MatchArgs m, const uint2* items, uint32_t items_cap,
                                                                 const uint32_t* nitems) {
    __shared__ uint4 buffer0[kStageBufferUnits];
    __shared__ uint4 buffer1[kStageBufferUnits];
    __shared__ uint32_t allow[kMaxAllowWords];
    __shared__ Modulus mods[kMaxResiduals];
    __shared__ uint32_t tilecnt[kTileSlots * (kMaxResiduals / 4)];
    __shared__ uint32_t queues[kStagedWaves * kQueue];
    __shared__ uint32_t fill;   // survivors of the current item
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    ItemState st;
    st.total = min(*nitems, items_cap);
    st.item = blockIdx.x;
    if (st.item >= st.total) return;
    // allow and the Modulus table are whole 16-byte units (both buffers are allocated so)
    stage_copy<kStagedThreads, 1>((uint4*)allow, (const uint4*)m.allow, (m.allow_words + 3) / 4);
    stage_copy<kStagedThreads, 1>((uint4*)mods, (const uint4*)m.mods, m.nres * (uint32_t)sizeof(Modulus) / 16);
    for (uint32_t t = threadIdx.x; t < (uint32_t)kTileSlots * m.ng; t += kStagedThreads) tilecnt[t] = 0;
    const Staged S{allow, mods, tilecnt, queues + wave * kQueue};
    st.it = items[st.item];
    load_item_rows(m, st.it.x, st.rows);
    stage_async(m, st.it.y, (uint8_t*)buffer0, wave, lane);
    for (;;) {
        if (!staged_item<kF, kLast>(m, items, S, st, buffer0, buffer1, &fill, lane, wave)) break;
        if (!staged_item<kF, kLast>(m, items, S, st, buffer1, buffer0, &fill, lane, wave)) break;
    }
    // the residual chain over this block's items' survivors, 64 at a time; the
    // block's waves take the 64-survivor slices of its items in turn
    __syncthreads();   // every item's region and count is written
    uint32_t slice = 0;
    for (uint32_t item = blockIdx.x; item < st.total; item += gridDim.x) {
        uint32_t n = m.item_count[item];
        if (n > m.item_cap) {
            if (threadIdx.x == 0) atomicMax((unsigned long long*)m.cand_max, (unsigned long long)n);
            n = m.item_cap;
        }
        const uint2 it = items[item];
        const uint32_t j0 = it.x * kPieceRows, u0 = (it.y * kStageWords) << 6;
        const uint32_t* region = m.cand + (uint64_t)item * m.item_cap;
        for (uint32_t first = 0; first < n; first += kWave, ++slice) {
            if (slice % kStagedWaves != wave) continue;
            const uint32_t count = min(n - first, (uint32_t)kWave);
            const uint32_t e = lane < count ? region[first + lane] : 0u;
            staged_chain(m, S, lane, count, j0 + (e >> 8), u0 + (e & 0xffu));
        }
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < (uint32_t)kTileSlots * m.ng; t += kStagedThreads) {
        const uint32_t c = tilecnt[t];
        if (c) atomicAdd(&m.counters[(uint64_t)(t / m.ng) * (m.ng + 3) + t % m.ng + 1], (uint64_t)c);
    }
}

#define PCUBE_STAGED(F)                                                                               \
    template __global__ void k_match_staged<F, false>(MatchArgs, const uint2*, uint32_t, const uint32_t*); \
    template __global__ void k_match_staged<F, true>(MatchArgs, const uint2*, uint32_t, const uint32_t*);
PCUBE_STAGED(1)
PCUBE_STAGED(2)
PCUBE_STAGED(3)
PCUBE_STAGED(4)
PCUBE_STAGED(5)
PCUBE_STAGED(6)
PCUBE_STAGED(7)
PCUBE_STAGED(8)
#undef PCUBE_STAGED

template <bool kLast>
inline const void* staged_kernel_for(uint32_t nbits) {
    switch (nbits) {
        case 1: return (const void*)k_match_staged<1, kLast>;
        case 2: return (const void*)k_match_staged<2, kLast>;
        case 3: return (const void*)k_match_staged<3, kLast>;
        case 4: return (const void*)k_match_staged<4, kLast>;
        case 5: return (const void*)k_match_staged<5, kLast>;
        case 6: return (const void*)k_match_staged<6, kLast>;
        case 7: return (const void*)k_match_staged<7, kLast>;
        default: return (const void*)k_match_staged<8, kLast>;
    }
}
// The instantiation for a matcher: nbits bitset filters out of nres.
inline const void* staged_kernel(uint32_t nbits, uint32_t nres) {
    return nbits == nres ? staged_kernel_for<true>(nbits) : staged_kernel_for<false>(nbits);
}

}  // namespace pcube
