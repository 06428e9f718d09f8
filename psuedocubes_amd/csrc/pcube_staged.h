#pragma once
// LDS-staged residue-bitset engine for one batch (gfx950).
//
// The first kStagedFilters residual filters (193, 199, 211, 223, 229, 241,
// 271, 277 at p = 619) are folded into residue bitsets: for filter q and
// residue class r = v mod q, bit t of mask word w is allowed_q[(u*B - r*A)
// mod q] for the u at index 64w + t.  Masks are laid out in chunks of
// kStageWords words.  One chunk of every filter is one image of stage_lines
// 256-byte lines: line r holds, in its 32-byte column f, filter f's row r
// (words 0-1 in the first 16 bytes, words 2-3 in the second).
//
// Bank-conflict-free reads.  A ds_read_b128 serves 16 lanes per cycle, and
// a line's 16 (column, half) pairs cover all 64 banks once.  Lane g = lane mod
// 16 of every wave reads, at step s = 0..15, the pair s ^ g: filter
// (s ^ g) & 3 of group (s ^ g) >> 3 (filters 0-3 or 4-7), half
// ((s ^ g) >> 2) & 1.  The 16 lanes of a cycle then hit 16 distinct bank
// groups whatever lines their rows use.  The row builder stores a row's eight
// column offsets in the order its lane reads them (the row's lane is fixed by
// its index).
//
// Work item = (stage chunk s, piece of kPieceRows consecutive v rows), planned
// chunk-major; work unit = 64 rows of an item's piece (one row per lane).  Two
// persistent blocks per CU each take a contiguous run of units.  A block stages
// the chunk image of its run's current segment (the units of one chunk) in its
// LDS, and its waves take the segment's units in turn: a lane takes one row and
// ANDs, for each of the chunk's words, the kStagedFilters words of its row's
// residue classes.  Waves run independently within a segment; the block
// synchronises only when the chunk changes (then the next image streams in with
// global_load_lds while the other block of the CU keeps its SIMDs busy).
// Positions surviving all bitset filters (about 1 in 6561 at p = 619) go to the
// wave's own survivor list, and the block runs them through the rest of the
// filter chain after its run, exactly as in the residue-bitset engine: cached
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
constexpr int kPieceRows = 1024;      // v rows per work item
// Work units: unit x is rows 64 (x mod 16) .. 64 (x mod 16) + 63 of the piece of
// item x / 16 against that item's chunk, so a block's contiguous run of units
// follows the chunk-major item list.
constexpr uint32_t kUnitsPerItem = kPieceRows / 64;
constexpr int kStagedThreads = 512;   // threads of a persistent block
constexpr int kStagedBlocksPerCU = 2; // persistent blocks per CU (4 waves per SIMD, 128 VGPRs)
constexpr uint32_t kFirstRoundWeight = 80;   // a first-round block's share of units, in 1/64 of an average run (measured)
constexpr uint32_t kPlanLdsPieces = 1024;    // pieces whose reach the planner keeps in LDS
constexpr int kStagedWaves = kStagedThreads / kWave;
constexpr int kTileSlots = 64;        // subtiles whose group counts a block keeps in LDS
constexpr int kItemGroups = 2;        // groups counted by the intersection itself (popcounts)
constexpr int kStageLines = 280;      // lines of a chunk image: the largest staged q
constexpr int kStageLineBytes = kStagedFilters * kStageWords * 8;   // 8 columns of 32 bytes
static_assert(kStagedFilters == 8 && kStageLineBytes == 256, "the read rotation assumes 8 columns of 32 bytes");

static_assert((kStageLines + 64 + 31) / 32 + 2 <= kWave, "a permuted string takes one load per lane");

// Staged residue bitsets.  Wave (s, k) builds every row of bitset filter k
// (its column of the chunk image) for the kStageWords words of stage chunk s,
// 64 rows at a time.  Lane t takes the u at index 64w + t and reduces it once:
// with P[j] = allowed_q[(-j*A) mod q] (stored repeated, q + 64 bits) and
// k0 = (r0 + (u mod q) * nbeta) mod q, bit j of the 64-bit window
// P[k0 .. k0 + 63] is allowed_q[(u*B - (r0 + j)*A) mod q]: the column of this
// u over rows r0..r0+63.  A wave transpose turns the 64 columns into the 64
// row words.
// Words of a staged filter's permuted string: q + 64 bits with q <= kStageLines,
// plus the two words of margin the matcher stores after each.
constexpr int kStrWords = (kStageLines + 64 + 31) / 32 + 2;

__device__ __forceinline__ void staged_mask_wave(const MatchArgs& m, uint64_t g) {
    // the filter's permuted string, read as 64-bit windows, per wave in LDS
    __shared__ uint32_t strings[256 / kWave][kStrWords];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    if (g >= (uint64_t)m.nbits * m.nstage) return;
    const uint32_t s = uint32_t(g / m.nbits), k = uint32_t(g % m.nbits);
    const Modulus mod = m.mods[k];
    uint32_t* P = strings[threadIdx.x / kWave];
    if (lane < (mod.q + 64 + 31) / 32 + 2) P[lane] = m.pstr[m.pstr_off[k] + lane];
    wave_sync();
    // (u mod q) * nbeta mod q of this lane's u in each word, once for all rows
    uint32_t base[kStageWords];
    bool valid[kStageWords];
#pragma unroll
    for (int ww = 0; ww < kStageWords; ++ww) {
        const uint32_t i = (s * kStageWords + ww) * kWave + lane;
        valid[ww] = i < m.nu;
        base[ww] = valid[ww] ? mod32(mod64(m.us[i], mod) * m.nbeta[k], mod) : 0u;
    }
    uint64_t* image = (uint64_t*)m.masks + (uint64_t)s * m.stage_lines * (kStageLineBytes / 8);
    for (uint32_t r0 = 0; r0 < mod.q; r0 += kWave) {
        const uint32_t rows = mod.q - r0 < (uint32_t)kWave ? mod.q - r0 : (uint32_t)kWave;
        uint64_t out[kStageWords];
#pragma unroll
        for (int ww = 0; ww < kStageWords; ++ww) {
            // bits k0 .. k0 + 63 of P by two funnel shifts (k0 < q, so the
            // three words lie in the loaded string)
            uint32_t k0 = base[ww] + r0;
            k0 = k0 >= mod.q ? k0 - mod.q : k0;
            const uint32_t w0 = k0 >> 5, sh = k0 & 31;
            const uint32_t p0 = P[w0], p1 = P[w0 + 1], p2 = P[w0 + 2];
            const uint64_t col = ((uint64_t)__builtin_amdgcn_alignbit(p2, p1, sh) << 32) |
                                 __builtin_amdgcn_alignbit(p1, p0, sh);
            out[ww] = wave_transpose(valid[ww] ? col : 0ull, lane);
        }
        if (lane < rows) {
            uint4* dst = (uint4*)(image + (r0 + lane) * (kStageLineBytes / 8) + k * kStageWords);
            dst[0] = make_uint4(uint32_t(out[0]), uint32_t(out[0] >> 32), uint32_t(out[1]), uint32_t(out[1] >> 32));
            dst[1] = make_uint4(uint32_t(out[2]), uint32_t(out[2] >> 32), uint32_t(out[3]), uint32_t(out[3] >> 32));
        }
    }
    // the columns of absent filters (fewer than kStagedFilters staged) keep one
    // all-ones line 0, which their rows read
    if (k == 0 && lane < kStagedFilters - m.nbits) {
        uint4* dst = (uint4*)(image + (m.nbits + lane) * kStageWords);
        dst[0] = dst[1] = make_uint4(~0u, ~0u, ~0u, ~0u);
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

// The LDS a staged block's helpers use (static arrays of k_match_staged).
// The filter tables are set only for the residual chain after the units.
struct Staged {
    const uint32_t* allow;
    const Modulus* mods;
    uint32_t* tilecnt;       // [kTileSlots][stride]: groups 1..stride of subtiles 0..kTileSlots-1 (units only)
    uint32_t stride;         // groups per subtile kept in tilecnt
    uint2* queue;            // this wave's survivors: (row index j, u index i)
};

// Add n to group g (1-based) of a subtile: LDS for the first kTileSlots
// subtiles of the batch (all of them for a batch of at most 64), else the
// global counters.
__device__ __forceinline__ void staged_add(const MatchArgs& m, const Staged& S, uint32_t tile, uint32_t g,
                                           uint32_t n) {
    if (tile < (uint32_t)kTileSlots) atomicAdd(&S.tilecnt[tile * S.stride + g - 1], n);
    else atomicAdd(&m.counters[(uint64_t)tile * (m.ng + 3) + g], (uint64_t)n);
}

// Add 1 to group g (1-based) of its row's subtile for each lane where pred
// holds, in the global counters: one atomic for the wave when those lanes share
// a subtile (the chain's counts are few: a few thousand per batch).
__device__ __forceinline__ void chain_count(const MatchArgs& m, bool pred, uint32_t tile, uint32_t g, uint32_t lane) {
    const uint64_t active = __ballot(pred);
    if (!active) return;
    const uint32_t lead = __builtin_amdgcn_readlane(tile, __builtin_ctzll(active));
    if (!__ballot(pred && tile != lead)) {
        if (lane == 0) atomicAdd(&m.counters[(uint64_t)lead * (m.ng + 3) + g], (uint64_t)__popcll(active));
    } else if (pred) {
        atomicAdd(&m.counters[(uint64_t)tile * (m.ng + 3) + g], 1ull);
    }
}

// Run up to 64 bitset survivors (lane l takes row j, u index i, live when
// l < count) through the filters after the bitsets; the filter tables are
// S.allow and S.mods, the later groups' counts go to the global counters.
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
            // (u*B - v*A) mod q = rho + sigma, sigma = (-v*A) mod q, formed here for
            // the lanes still alive (the row records do not carry it)
            const uint32_t va = mod32(mod64(row.v, mod) * mod.a, mod);
            uint32_t r = ((words[c >> 1] >> (16 * (c & 1))) & 0xffffu) + (va ? mod.q - va : 0u);
            r = r >= mod.q ? r - mod.q : r;
            alive = allowed_bit(S.allow, mod.bit + r);
        }
        if ((k & 3) == 3 || k + 1 == m.nres) chain_count(m, alive, row.tile, k / 4 + 1, lane);
        if (!__ballot(alive)) return;
        ++k;
    }
    // the filters on exact x = u*B - v*A: few lanes get here (about 1 in 81 of
    // the bitset survivors at p = 619), so each survivor in turn is tested by the
    // whole wave, lane l taking filter k + l, 64 filters per round; the first
    // failing filter ends it, and the groups closed before it are counted
    U128 x = {0, 0};
    if (alive) x = sub_wide(mul_wide(m.us[i], m.B), mul_wide(row.v, m.A));
    for (uint64_t pending = __ballot(alive); pending; pending &= pending - 1) {
        const int src = __builtin_ctzll(pending);
        const U128 xs = {readlane64(x.lo, src), readlane64(x.hi, src)};
        const uint32_t tile = __builtin_amdgcn_readlane(row.tile, src);
        bool passed = true;
        for (uint32_t k0 = k; k0 < m.nres; k0 += kWave) {
            const uint32_t f = k0 + lane;
            bool pass = true;
            if (f < m.nres) {
                const Modulus mod = S.mods[f];
                pass = allowed_bit(S.allow, mod.bit + mod128(xs, mod));
            }
            const uint64_t fails = __ballot(!pass);
            const uint32_t upto = fails ? (uint32_t)__builtin_ctzll(fails) : (uint32_t)kWave;
            if (lane < upto && f < m.nres && ((f & 3) == 3 || f + 1 == m.nres))
                atomicAdd(&m.counters[(uint64_t)tile * (m.ng + 3) + f / 4 + 1], 1ull);
            if (fails) {
                passed = false;
                break;
            }
        }
        if (passed && lane == 0) {
            if (!less_wide(m.lo, xs) || less_wide(m.hi, xs)) atomicOr(m.error, 2u);
            emit(m, tile, xs);
        }
    }
}

// A survivor record: the bitset survivors of one row in one chunk, 48 bytes:
// (row j, u index of word 0, u index of word 2, 0), then the four 64-bit
// survivor words (words 0-1 cover u from the first index on, words 2-3 from
// the second).
constexpr int kRecordUnits = 3;   // 16-byte units of a record

// Write the record of each lane with survivors to the wave's region (fill
// records written so far; those past cap are counted, not written).
__device__ __forceinline__ void write_records(const uint64_t (&w)[4], uint32_t j, uint32_t up, uint32_t uq,
                                              uint4* region, uint32_t cap, uint32_t& fill) {
    const bool has = (w[0] | w[1] | w[2] | w[3]) != 0;
    const uint64_t ballot = __ballot(has);
    if (!ballot) return;
    if (has) {
        const uint32_t at = fill + lanes_below(ballot);
        if (at < cap) {
            uint4* rec = region + (uint64_t)at * kRecordUnits;
            rec[0] = make_uint4(j, up, uq, 0);
            rec[1] = make_uint4(uint32_t(w[0]), uint32_t(w[0] >> 32), uint32_t(w[1]), uint32_t(w[1] >> 32));
            rec[2] = make_uint4(uint32_t(w[2]), uint32_t(w[2] >> 32), uint32_t(w[3]), uint32_t(w[3] >> 32));
        }
    }
    fill += (uint32_t)__popcll(ballot);
}

// Run a wave's n survivor records through the residual chain: lane l expands
// record l of each group of 64 into (row, u index) pairs, one bit per lane per
// round, into the wave's queue; each 64 queued pairs take one chain call.
__device__ __forceinline__ void chain_records(const MatchArgs& m, const Staged& C, uint32_t lane,
                                              const uint4* region, uint32_t n) {
    uint32_t queued = 0;
    for (uint32_t r0 = 0; r0 < n; r0 += kWave) {
        uint64_t b[4] = {0, 0, 0, 0};
        uint32_t j = 0, up = 0, uq = 0;
        if (r0 + lane < n) {
            const uint4* rec = region + (uint64_t)(r0 + lane) * kRecordUnits;
            const uint4 h = rec[0], x = rec[1], y = rec[2];
            j = h.x;
            up = h.y;
            uq = h.z;
            b[0] = ((uint64_t)x.y << 32) | x.x;
            b[1] = ((uint64_t)x.w << 32) | x.z;
            b[2] = ((uint64_t)y.y << 32) | y.x;
            b[3] = ((uint64_t)y.w << 32) | y.z;
        }
        for (;;) {
            const bool has = (b[0] | b[1] | b[2] | b[3]) != 0;
            const uint64_t ballot = __ballot(has);
            if (!ballot) break;
            if (has) {
                const int w = b[0] ? 0 : b[1] ? 1 : b[2] ? 2 : 3;
                const uint64_t x = w == 0 ? b[0] : w == 1 ? b[1] : w == 2 ? b[2] : b[3];
                const uint32_t u0 = w < 2 ? up + 64 * w : uq + 64 * (w - 2);
                C.queue[queued + lanes_below(ballot)] = make_uint2(j, u0 + (uint32_t)__builtin_ctzll(x));
                const uint64_t rest = x & (x - 1);
                b[0] = w == 0 ? rest : b[0];
                b[1] = w == 1 ? rest : b[1];
                b[2] = w == 2 ? rest : b[2];
                b[3] = w == 3 ? rest : b[3];
            }
            queued += (uint32_t)__popcll(ballot);
            if (queued >= kWave) {
                wave_sync();
                const uint2 e = C.queue[lane];
                staged_chain(m, C, lane, kWave, e.x, e.y);
                wave_sync();
                const uint32_t left = queued - kWave;
                uint2 moved = make_uint2(0, 0);
                if (lane < left) moved = C.queue[kWave + lane];
                wave_sync();
                if (lane < left) C.queue[lane] = moved;
                queued = left;
                wave_sync();
            }
        }
    }
    if (queued) {
        wave_sync();
        const uint2 e = lane < queued ? C.queue[lane] : make_uint2(0, 0);
        staged_chain(m, C, lane, queued, e.x, e.y);
    }
}

// End i of the pieces' reach (i = 2p: the lower end of piece p's first row;
// i = 2p + 1: the upper end of its last row): the row's v coordinate, the search
// key of its bound, and the u index of the end.  Rows are sorted by v, so
// [left0, right1) = [end 2p, end 2p + 1) holds every u index the piece's rows
// reach.  With a directory (dir[k] = us[k * stride] in LDS) an end comes from it
// alone, widened to the directory interval that holds the exact end (no global
// probe); without one, it is exact.  Both ends are nondecreasing over the
// pieces, and a piece's items may include a chunk beyond each end that none of
// its rows touches (the match skips such units).
__device__ __forceinline__ uint64_t piece_end(const MatchArgs& m, uint32_t i, uint32_t nv) {
    const uint32_t j0 = (i / 2) * kPieceRows;
    return j0 < nv ? m.vs[i & 1 ? min(nv, j0 + kPieceRows) - 1 : j0] : 0ull;
}

__device__ __forceinline__ AboveKey piece_end_key(const MatchArgs& m, uint32_t i, uint64_t v) {
    return above_key(add_wide(mul_wide(v, m.A), i & 1 ? m.hi : m.lo), m.B, m.B_mu);
}

__device__ __forceinline__ uint32_t piece_end_index(const MatchArgs& m, uint32_t i, const AboveKey& key,
                                                    const uint64_t* dir, uint32_t ndir, uint32_t stride) {
    if (!dir) {
        uint32_t r0, r1;
        uint64_t unused;
        first_above_keys<false>(m, key, key, dir, ndir, stride, r0, r1, unused);
        return r0;
    }
    // the first directory entry a above the bound: the exact end lies in
    // ((a - 1) * stride, a * stride] (0 when a = 0)
    uint32_t a = 0, e = ndir;
    while (a < e) {
        const uint32_t mid = a + ((e - a) >> 1);
        if (!below_key(dir[mid], key)) e = mid; else a = mid + 1;
    }
    if (i & 1) return (uint32_t)min((uint64_t)a * stride, (uint64_t)m.nu);
    return a == 0 ? 0u : (a - 1) * stride + 1;
}

// Share of the work before run r, in 1/64 of an average run, when the first-round
// blocks (block id < grid / 2, the first to reach each CU) take first64 / 64 of
// an average run and the others (128 - first64) / 64.  Runs are grouped by XCD
// (run = (b mod 8) * grid / 8 + b / 8), so within each XCD's runs the first half
// belongs to first-round blocks.  The two blocks of a CU then end together
// although the older one's waves win the CU's issue arbitration.
__device__ __forceinline__ uint64_t run_weight_prefix(uint32_t r, uint32_t grid, uint32_t first64) {
    const uint32_t per = grid / 8;
    if (grid % 16 != 0 || first64 == 64) return (uint64_t)r * 64;
    const uint32_t x = r / per, k = r % per, half = per / 2;
    const uint64_t within = k < half ? (uint64_t)k * first64 : (uint64_t)half * first64 + (uint64_t)(k - half) * (128 - first64);
    return (uint64_t)x * per * 64 + within;
}

// Work of the staged match, chunk-major: for every stage chunk c (u indices
// [256c, 256c + 256)) the pieces whose rows reach it, in piece order; item k
// of this list is one (piece, chunk) pair of 16 work units.  With piece p
// reaching [left0_p, right1_p), both nondecreasing in p, the pieces reaching
// chunk c are the run [a_c, b_c): a_c the first with right1 > 256c, b_c the
// first with left0 >= 256c + 256.  (A piece whose rows reach nothing, right1 <=
// left0, can lie inside a run: its item has no work.)  The list is kept as its
// chunk table: chunks[c] = (a_c, first item of chunk c) for c <= nstage, so item
// k of chunk c is piece a_c + k - chunks[c].y.  Chunk-major order lets a block
// walking consecutive items keep one chunk image in LDS while only the row
// headers change.
// Geometry of the fused prefilter launch and of the match that follows it.
struct PrefilterArgs {
    uint32_t rows_blocks, mask_blocks, rho_blocks, npieces;
    uint32_t reach_lds;  // the pieces' reach fits the dynamic LDS after the directory
    uint32_t grid;       // blocks of the staged match
    uint32_t first64;    // a first-round match block's share of the work (1/64 of an average run)
    uint2* chunks;       // [nstage + 1]: first piece and first item of each chunk
    uint32_t* reach;     // [npieces][2]: the u indices each piece's rows reach (when not in LDS)
    uint2* runs;         // [grid + 1]: first work unit of each match block's run and that unit's chunk
    // a launch sized before the windows were known: geo[0], geo[1] = the true u
    // and v counts (at most nu and nv); rows from geo[1] on are padding, and
    // us holds 2^64 - 1 from geo[0] on.  Null: nu and nv are exact.
    const uint64_t* geo;
};

// Rows of the batch that are real (the rest, up to nv, are padding).
__device__ __forceinline__ uint32_t real_rows(const MatchArgs& m, const PrefilterArgs& pa) {
    return pa.geo ? (uint32_t)min(pa.geo[1], (uint64_t)m.nv) : m.nv;
}

// Cost of a chunk change inside a match block's run, in work units: the block
// waits for the next chunk image about as long as its waves take for this
// many units (measured).
constexpr uint32_t kChunkChangeUnits = 16;

// The planner (one 256-thread block, at raised wave priority: the match waits
// for it).  One thread per end of each piece's reach: the row coordinate loads
// with the u directory, the search key is formed while the directory arrives,
// and the end goes to `reach` (two words per piece; in LDS after the directory
// for kLds, else global).  Then thread t takes a run of consecutive chunks,
// finds each chunk's piece run [a_c, b_c) and keeps it in LDS over the
// directory, no longer needed, when that holds nstage entries (else the runs
// are searched again where needed); a block scan of the runs' item counts
// gives the chunk table, and each match block's run starts where the cost
// (units plus kChunkChangeUnits per nonempty chunk) reaches its share.
template <bool kLds>
__device__ __forceinline__ void staged_plan_block(const MatchArgs& m, const PrefilterArgs& pa, uint64_t* dir) {
    __shared__ uint32_t wave_items[256 / kWave], wave_cost[256 / kWave];
    __shared__ uint32_t items_before[257], cost_before[257];   // per thread's chunk run; [256]: the totals
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const uint32_t npieces = pa.npieces, nstage = m.nstage;
    // the other blocks of this CU (rows, bitsets) yield issue slots to the planner
    __builtin_amdgcn_s_setprio(3);
    uint32_t ndir, stride;
    directory_shape(m.nu, stride, ndir);
    // thread i < 2 npieces takes one end of piece i / 2 (even i: the lower end of
    // its first row, odd i: the upper end of its last): the coordinate loads
    // with the directory's, the search key is formed before the directory is
    // stored, and a binary search over the directory gives the end
    static_assert(kRowDirectory == 256 * 16, "one round of sixteen loads per thread fills the directory");
    const uint32_t ends = 2 * npieces;
    // pieces past the real rows (padding) reach [nu, nu): nothing, and the ends
    // stay nondecreasing
    const uint32_t nv_real = real_rows(m, pa);
    uint64_t v_end = 0;
    if (threadIdx.x < ends) v_end = piece_end(m, threadIdx.x, nv_real);
    uint64_t d[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const uint32_t t = 256 * k + threadIdx.x;
        d[k] = t < ndir ? m.us[t * stride] : 0ull;
    }
    AboveKey key0{};
    if (threadIdx.x < ends) key0 = piece_end_key(m, threadIdx.x, v_end);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const uint32_t t = 256 * k + threadIdx.x;
        if (t < ndir) dir[t] = d[k];
    }
    __syncthreads();
    {
        uint32_t* reach = kLds ? (uint32_t*)(dir + (ndir ? ndir : 1)) : pa.reach;
        for (uint32_t i = threadIdx.x; i < ends; i += 256) {
            if ((i / 2) * kPieceRows >= nv_real) {
                reach[i] = m.nu;
                continue;
            }
            const AboveKey key = i == threadIdx.x ? key0 : piece_end_key(m, i, piece_end(m, i, nv_real));
            reach[i] = piece_end_index(m, i, key, ndir ? dir : nullptr, ndir, stride);
        }
    }
    __threadfence_block();
    __syncthreads();
    const uint32_t* reach = kLds ? (const uint32_t*)(dir + (ndir ? ndir : 1)) : pa.reach;
    constexpr uint32_t kChunkU = kStageWords * 64;
    const uint32_t per = (nstage + 255) / 256;
    const uint32_t c0 = min(nstage, threadIdx.x * per), c1 = min(nstage, c0 + per);
    // each chunk's piece run [a_c, b_c): a_c the first piece with right1 > 256c,
    // b_c the first with left0 >= 256c + 256 (two interleaved searches); kept in
    // LDS over the directory, no longer read, when it holds nstage entries
    uint2* span = (uint2*)dir;
    const bool tabled = nstage <= ndir;
    auto search_span = [&](uint32_t c) -> uint2 {
        uint32_t al = 0, ah = npieces, bl = 0, bh = npieces;
        const uint32_t ka = c * kChunkU + 1, kb = (c + 1) * kChunkU;
        while (al < ah || bl < bh) {
            if (al < ah) {
                const uint32_t mid = (al + ah) >> 1;
                if (reach[2 * mid + 1] < ka) al = mid + 1; else ah = mid;
            }
            if (bl < bh) {
                const uint32_t mid = (bl + bh) >> 1;
                if (reach[2 * mid] < kb) bl = mid + 1; else bh = mid;
            }
        }
        return make_uint2(al, bl);
    };
    if (tabled) {
        for (uint32_t c = c0; c < c1; ++c) span[c] = search_span(c);
        __syncthreads();
    }
    auto chunk_span = [&](uint32_t c) -> uint2 { return tabled ? span[c] : search_span(c); };
    // items of this thread's chunks, and their cost: each nonempty chunk's units
    // plus one chunk change
    uint32_t count = 0, cost = 0;
    for (uint32_t c = c0; c < c1; ++c) {
        const uint2 ab = chunk_span(c);
        if (ab.y > ab.x) {
            count += ab.y - ab.x;
            cost += kChunkChangeUnits + kUnitsPerItem * (ab.y - ab.x);
        }
    }
    uint32_t inc = count, inc_cost = cost;
#pragma unroll
    for (int dd = 1; dd < kWave; dd <<= 1) {
        const uint32_t y = __shfl_up(inc, dd, kWave), z = __shfl_up(inc_cost, dd, kWave);
        if (lane >= (uint32_t)dd) {
            inc += y;
            inc_cost += z;
        }
    }
    if (lane == kWave - 1) {
        wave_items[wave] = inc;
        wave_cost[wave] = inc_cost;
    }
    __syncthreads();
    uint32_t at = inc - count, kat = inc_cost - cost;
    for (uint32_t w = 0; w < wave; ++w) {
        at += wave_items[w];
        kat += wave_cost[w];
    }
    items_before[threadIdx.x] = at;
    cost_before[threadIdx.x] = kat;
    if (threadIdx.x == 255) {
        items_before[256] = at + count;
        cost_before[256] = kat + cost;
        pa.chunks[nstage] = make_uint2(npieces, at + count);
    }
    for (uint32_t c = c0; c < c1; ++c) {
        const uint2 ab = chunk_span(c);
        pa.chunks[c] = make_uint2(ab.x, at);
        if (ab.y > ab.x) at += ab.y - ab.x;
    }
    __syncthreads();
    // the match blocks' runs: run r starts at the unit where the cost reaches its
    // share (run_weight_prefix) of the total; a chunk change costs
    // kChunkChangeUnits before the chunk's first unit
    const uint32_t total_units = items_before[256] * kUnitsPerItem;
    const uint64_t total_cost = cost_before[256];
    for (uint32_t r = threadIdx.x; r <= pa.grid; r += 256) {
        uint32_t u = total_units, chunk = 0;
        if (r < pa.grid && total_cost) {
            const uint64_t target = total_cost * run_weight_prefix(r, pa.grid, pa.first64) / ((uint64_t)pa.grid * 64);
            // the last thread whose chunks' cost starts at or below the target
            uint32_t lo = 0, hi = 256;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (cost_before[mid] <= target) lo = mid; else hi = mid - 1;
            }
            const uint32_t t0 = min(nstage, lo * per), t1 = min(nstage, t0 + per);
            uint64_t k = cost_before[lo];
            uint32_t unit = items_before[lo] * kUnitsPerItem;
            for (uint32_t c = t0; c < t1; ++c) {
                const uint2 ab = chunk_span(c);
                if (ab.y <= ab.x) continue;
                const uint32_t units = kUnitsPerItem * (ab.y - ab.x);
                if (target < k + kChunkChangeUnits + units) {
                    const uint64_t off = target - k;
                    u = unit + (off < kChunkChangeUnits ? 0u : (uint32_t)(off - kChunkChangeUnits));
                    chunk = c;
                    break;
                }
                k += kChunkChangeUnits + units;
                unit += units;
            }
            u = min(u, total_units);
        }
        pa.runs[r] = make_uint2(u, chunk);
    }
}


// Everything the staged match needs, in one launch of independent 256-thread
// blocks: block 0 plans the work items; then the row blocks (exact u bounds,
// mask row offsets, sigmas, valid counts), the mask waves and the residue
// cache.
// The u directory lives in dynamic LDS sized to the batch (directory_shape of
// nu, 8 bytes an entry, then the planner's piece reach when it fits), so the
// blocks that never read it stay small.
__global__ __launch_bounds__(256) void k_prefilter_staged(MatchArgs m, PrefilterArgs pa) {
    extern __shared__ uint64_t prefilter_dir[];
    uint64_t* dir = prefilter_dir;
    uint32_t b = blockIdx.x;
    if (b == 0) {
        if (pa.reach_lds) staged_plan_block<true>(m, pa, dir);
        else staged_plan_block<false>(m, pa, dir);
        return;
    }
    b -= 1;
    if (b < pa.rows_blocks) {
        staged_rows_block(m, b, dir, real_rows(m, pa));
        return;
    }
    b -= pa.rows_blocks;
    if (b < pa.mask_blocks) {
        staged_mask_wave(m, (uint64_t)b * (256 / kWave) + threadIdx.x / kWave);
        return;
    }
    b -= pa.mask_blocks;
    const uint64_t i = (uint64_t)b * 256 + threadIdx.x;
    if (i < m.nu) build_rho(m, (uint32_t)i);
}

// Row header of one row of a work unit: (left, right, tile) and the LDS byte
// offsets of the mask halves it reads first (h1: slots 0-3, h2: slots 4-7).
struct UnitRow {
    uint3 h0;
    uint4 h1, h2;
};

// Makes the compiler wait for a unit's row header here (an empty statement that
// reads its registers), so that the next unit's headers, loaded after this, may
// stay in flight while this unit is matched: the vector memory counter waits
// in order, so without it the first use of this header would also wait for
// those loads.
__device__ __forceinline__ void unit_row_ready(const UnitRow& R) {
    asm volatile("" ::"v"(R.h0.x), "v"(R.h0.y), "v"(R.h0.z), "v"(R.h1.x), "v"(R.h1.y), "v"(R.h1.z), "v"(R.h1.w),
                 "v"(R.h2.x), "v"(R.h2.y), "v"(R.h2.z), "v"(R.h2.w));
}

// Unit s of piece p is rows j0 + 64 s .. j0 + 64 s + 63 of the piece (j0 =
// 1024 p), lane l taking row j0 + 64 s + l; lanes past the piece's end read
// its last row (their results are discarded).
__device__ __forceinline__ void load_unit_row(const MatchArgs& m, uint32_t p, uint32_t s, uint32_t lane, UnitRow& R) {
    const uint32_t j0 = p * kPieceRows, j1 = min(m.nv, j0 + kPieceRows);
    const uint32_t j = j0 + s * kWave + lane;
    const uint4* rp = (const uint4*)&m.rows[j < j1 ? j : j1 - 1];
    R.h0 = *(const uint3*)rp;
    R.h1 = rp[1];
    R.h2 = rp[2];
}

// Start the copy of stage chunk s into an LDS buffer: each wave moves 64
// sixteen-byte units per instruction straight from global memory to LDS.
// Complete once the wave's vector memory counter drains.
__device__ __forceinline__ void stage_async(const MatchArgs& m, uint32_t s, uint8_t* buffer, uint32_t wave,
                                            uint32_t lane) {
    const uint4* src = (const uint4*)(m.masks + (uint64_t)s * m.stage_lines * (kStageLineBytes / 8));
    const uint32_t n = m.stage_lines * (kStageLineBytes / 16);
    for (uint32_t q = wave; q * kWave < n; q += kStagedWaves) {
        const uint32_t t = q * kWave + lane;
        if (t < n) __builtin_amdgcn_global_load_lds((const void*)(src + t), (void*)(buffer + q * kWave * 16), 16, 0, 0);
    }
}

// Static LDS of the staged match (two blocks per CU): one chunk image, the
// per-subtile counts of groups 1 and 2 and the waves' survivor queues.  The
// chain after the units keeps the filter tables and the per-subtile counts of
// every group in the image buffer.
constexpr int kStageBufferUnits = kStageLines * kStageLineBytes / 16;
constexpr uint32_t kStagedLdsBytes =
    kStageBufferUnits * 16 + kTileSlots * kItemGroups * 4 + kStagedWaves * kQueue * (uint32_t)sizeof(uint2);
static_assert(kStagedBlocksPerCU * kStagedLdsBytes <= 160 * 1024, "two staged blocks share one CU's LDS");
static_assert(kMaxAllowWords * 4 + kMaxResiduals * sizeof(Modulus) + kTileSlots * (kMaxResiduals / 4) * 4 <=
                  kStageBufferUnits * 16,
              "the chain's filter tables and group counts fit the mask buffer");


// One segment of a match block's run: the items of one chunk, [item0,
// item_end) of the chunk-major list, item k being piece piece0 + k - item0.
struct Segment {
    uint32_t chunk, piece0, item0, item_end;
};

__device__ __forceinline__ Segment chunk_segment(const uint2* chunks, uint32_t c) {
    const uint2 d = chunks[c], e = chunks[c + 1];
    return Segment{c, d.x, d.y, e.y};
}

// Piece of work unit x (x in the segment).
__device__ __forceinline__ uint32_t unit_piece(const Segment& sg, uint32_t x) {
    return sg.piece0 + x / kUnitsPerItem - sg.item0;
}

// The two 64-bit words of a 16-byte mask half.
__device__ __forceinline__ uint64_t half_word(const uint4& h, int w) {
    return w == 0 ? ((uint64_t)h.y << 32) | h.x : ((uint64_t)h.w << 32) | h.z;
}

__device__ __forceinline__ uint4 and3(const uint4& a, const uint4& b, const uint4& c) {
    return make_uint4(a.x & b.x & c.x, a.y & b.y & c.y, a.z & b.z & c.z, a.w & b.w & c.w);
}

// Bits lo..hi-1 of a mask word for a row whose chunk-relative u range is
// [dl, dr), the word starting at chunk offset `first`.
__device__ __forceinline__ uint64_t word_range(int dl, int dr, int first) {
    const int lo = min(max(dl - first, 0), 64), hi = min(max(dr - first, 0), 64);
    return hi > lo ? (~0ull << lo) & (~0ull >> (64 - hi)) : 0ull;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
#pragma unroll
    for (int d = kWave / 2; d >= 1; d >>= 1) x += __shfl_xor(x, d, kWave);
    return x;
}

// Group 1 and 2 counts of a unit's rows for their subtiles: one add per group
// when the wave's counting rows share a subtile (the two counts, each at most
// 256 per row, summed packed in one 32-bit word), else one per row.
__device__ __forceinline__ void flush_groups(const MatchArgs& m, const Staged& S, bool has, uint32_t tile,
                                             uint32_t c1, uint32_t c2, uint32_t lane) {
    const uint64_t active = __ballot(has && (c1 | c2));
    if (!active) return;
    const uint32_t lead = __shfl(tile, __builtin_ctzll(active), kWave);
    if (!__ballot(has && (c1 | c2) && tile != lead)) {
        const uint32_t s = wave_sum(has ? c1 | (c2 << 16) : 0u);
        if (lane == 0) {
            if (s & 0xffffu) staged_add(m, S, lead, 1, s & 0xffffu);
            if (s >> 16) staged_add(m, S, lead, 2, s >> 16);
        }
    } else if (has) {
        if (c1) staged_add(m, S, tile, 1, c1);
        if (c2) staged_add(m, S, tile, 2, c2);
    }
}

// One row's intersection with the staged chunk, reading its 16 mask halves in the
// rotated order of its lane g (see the header): slots 0-3 (group X: filters 0-3
// for g < 8, 4-7 otherwise) then slots 4-7 (group Y, the other four), first the
// half P = (g >> 2) & 1 of every slot, then the other half Q.  c1 counts the
// positions passing group A (X for g < 8, Y otherwise) and c2 those passing both
// groups; words receives the four words of survivors.  kRange: apply the row's
// chunk-relative u range [dl, dr) (some row of the wave covers the chunk only in
// part); without it every position of the chunk is in range.
template <int kF, bool kLast, bool kRange>
__device__ __forceinline__ void staged_row(const uint8_t* image, int first_p, int first_q, uint32_t g, const UnitRow& R,
                                           int dl, int dr, uint32_t& c1, uint32_t& c2, uint64_t (&words)[4]) {
    // group 1 closes after filter 3 (or after the last one), group 2 after filter 7
    constexpr bool kCount1 = kF >= 4 || kLast;
    constexpr bool kCount2 = kF == 8 || (kLast && kF > 4);
    // byte offsets of half P of each slot; half Q is the other 16 bytes of the column
    const uint32_t offs[kStagedFilters] = {R.h1.x, R.h1.y, R.h1.z, R.h1.w, R.h2.x, R.h2.y, R.h2.z, R.h2.w};
    uint64_t both[2][2];
    uint32_t acc_x = 0, acc_y = 0, acc2 = 0;
    // half P of all eight slots, then half Q: the eight reads of a half are in
    // flight together, and a half is reduced to its counts and survivors at once
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int first = h == 0 ? first_p : first_q;
        uint4 v[kStagedFilters];
#pragma unroll
        for (int i = 0; i < kStagedFilters; ++i) v[i] = *(const uint4*)(image + (h == 0 ? offs[i] : offs[i] ^ 16u));
        const uint4 x01 = and3(v[0], v[1], v[2]), y01 = and3(v[4], v[5], v[6]);
        const uint4 xh = make_uint4(x01.x & v[3].x, x01.y & v[3].y, x01.z & v[3].z, x01.w & v[3].w);
        const uint4 yh = make_uint4(y01.x & v[7].x, y01.y & v[7].y, y01.z & v[7].z, y01.w & v[7].w);
#pragma unroll
        for (int w = 0; w < 2; ++w) {
            uint64_t x = half_word(xh, w), y = half_word(yh, w);
            if (kRange) {
                const uint64_t range = word_range(dl, dr, first + 64 * w);
                x &= range;
                y &= range;
            }
            both[h][w] = x & y;
            if (kCount1) {
                acc_x += __popcll(x);
                acc_y += __popcll(y);
            }
            if (kCount2) acc2 += __popcll(both[h][w]);
        }
        // eight reads in flight at a time keep the registers within four waves per SIMD
        __builtin_amdgcn_sched_barrier(0);
    }
    c1 = g < 8 ? acc_x : acc_y;
    c2 = acc2;
    words[0] = both[0][0];
    words[1] = both[0][1];
    words[2] = both[1][0];
    words[3] = both[1][1];
}

// One work unit: the rows of unit s of piece p (row j0 + 64 s + lane for this
// lane) against the staged image of chunk `chunk`, through staged_row (skipped
// when no row of the wave reaches the chunk; without range masks when every row
// covers it whole).  Survivor records go to the wave's region; the lane's
// group 1 and 2 counts and its row's subtile are returned (zero counts when
// the wave had nothing to do).
template <int kF, bool kLast>
__device__ __forceinline__ void staged_unit(const MatchArgs& m, uint32_t p, uint32_t s, uint32_t chunk,
                                            const UnitRow& R, const uint4* image, uint32_t lane, uint4* region,
                                            uint32_t& fill, uint32_t& c1, uint32_t& c2) {
    c1 = c2 = 0;
    const uint32_t j0 = p * kPieceRows, j1 = min(m.nv, j0 + kPieceRows);
    const uint32_t jr = j0 + s * kWave;
    if (jr >= j1) return;   // the piece ends before this unit
    const uint32_t j = jr + lane;
    const bool live = j < j1;
    constexpr int kChunkU = kStageWords * 64;
    const uint32_t base0 = chunk * kChunkU;   // first u index of the chunk
    // range of the row's u indices within the chunk
    const int dl = live ? (int)(R.h0.x - base0) : 0, dr = live ? (int)(R.h0.y - base0) : 0;
    const bool touches = live && dr > 0 && dl < kChunkU;
    const bool full = touches && dl <= 0 && dr >= kChunkU;
    if (!__ballot(touches)) return;
    const uint32_t g = lane & 15, hp = (g >> 2) & 1;                  // hp: the lane's half P of a column
    const int first_p = 128 * (int)hp, first_q = 128 - first_p;        // chunk offset of each half's first word
    uint64_t words[4];
    if (!__ballot(!full))
        staged_row<kF, kLast, false>((const uint8_t*)image, first_p, first_q, g, R, dl, dr, c1, c2, words);
    else
        staged_row<kF, kLast, true>((const uint8_t*)image, first_p, first_q, g, R, dl, dr, c1, c2, words);
    write_records(words, j, base0 + first_p, base0 + first_q, region, m.item_cap, fill);
}

// Group 1 and 2 counts of a wave's consecutive units, kept per lane while the
// counting rows of its units lie in one subtile and added to the block's
// counts when the subtile changes; a unit whose counting rows span two
// subtiles adds its own counts row by row.
struct GroupAcc {
    uint32_t tile;   // subtile of the held counts (wave-uniform), ~0u when none
    uint32_t c1, c2;
};

__device__ __forceinline__ void acc_flush(const MatchArgs& m, const Staged& S, GroupAcc& A, uint32_t lane) {
    if (A.tile == ~0u) return;
    const uint32_t s1 = wave_sum(A.c1), s2 = wave_sum(A.c2);
    if (lane == 0) {
        if (s1) staged_add(m, S, A.tile, 1, s1);
        if (s2) staged_add(m, S, A.tile, 2, s2);
    }
    A.tile = ~0u;
    A.c1 = A.c2 = 0;
}

__device__ __forceinline__ void acc_add(const MatchArgs& m, const Staged& S, GroupAcc& A, uint32_t tile, uint32_t c1,
                                        uint32_t c2, uint32_t lane) {
    const bool counted = (c1 | c2) != 0;
    const uint64_t active = __ballot(counted);
    if (!active) return;
    const uint32_t t0 = __builtin_amdgcn_readlane(tile, __builtin_ctzll(active));
    if (__ballot(counted && tile != t0)) {
        flush_groups(m, S, true, tile, c1, c2, lane);
        return;
    }
    if (t0 != A.tile) {
        acc_flush(m, S, A, lane);
        A.tile = t0;
    }
    A.c1 += c1;
    A.c2 += c2;
}

// Persistent staged match, kStagedBlocksPerCU blocks per CU.  The work units
// (16 per item of the chunk-major list) are split into one contiguous run per
// block (planned by cost, run_weight_prefix); the runs of one XCD's blocks are
// adjacent, so an L2 holds the chunk images its blocks stage.  A run is walked
// in segments of one chunk: the block stages the chunk's image once and each
// wave takes a contiguous share of the segment's units on its own, with the
// next unit's row headers in flight; the block synchronises only when the
// chunk changes.  Group 1 and 2 counts collect in LDS for the whole run; each
// wave writes its survivor records to its own region (item_cap records) and,
// when its units are done, runs them through the residual chain (filter tables
// from global memory, later groups' counts to the global counters) while the
// other waves may still be matching.  kF = bitset filters; kLast:
// the last residual filter is the last bitset filter (it closes a group).  The
// waves-per-EU bound keeps the kernel within 128 VGPRs, so both blocks of a CU
// stay resident.
template <int kF, bool kLast>
__global__ __launch_bounds__(kStagedThreads) __attribute__((amdgpu_waves_per_eu(4)))
void k_match_staged(MatchArgs m, const uint2* chunks, const uint2* runs) {
    __shared__ uint4 image[kStageBufferUnits];
    __shared__ uint32_t tilecnt[kTileSlots * kItemGroups];
    __shared__ uint2 queues[kStagedWaves * kQueue];
    // the wave index is the same in every lane: held in a scalar register, the
    // unit loop's control is scalar
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    // consecutive block ids go round robin over the 8 XCDs: group the runs by XCD
    const uint32_t grid = gridDim.x;
    const uint32_t run = grid % 8 == 0 ? (blockIdx.x % 8) * (grid / 8) + blockIdx.x / 8 : blockIdx.x;
    const uint2 mine = runs[run], next = runs[run + 1];
    const uint32_t u_first = mine.x, u_last = next.x;
    if (u_first >= u_last) return;
    // the first chunk image streams in while its items' descriptor loads
    stage_async(m, mine.y, (uint8_t*)image, wave, lane);
    Segment sg = chunk_segment(chunks, mine.y);
    for (uint32_t t = threadIdx.x; t < (uint32_t)(kTileSlots * kItemGroups); t += kStagedThreads) tilecnt[t] = 0;
    const Staged S{nullptr, nullptr, tilecnt, kItemGroups, queues + wave * kQueue};
    uint4* region = (uint4*)m.cand + ((uint64_t)blockIdx.x * kStagedWaves + wave) * m.item_cap * kRecordUnits;
    uint32_t fill = 0;
    GroupAcc A{~0u, 0, 0};
    // a segment [u, seg_end) of n units gives wave w the contiguous units
    // [u + n w / 8, u + n (w + 1) / 8), so its consecutive units mostly share a subtile
    uint32_t u = u_first;
    uint32_t seg_end = min(u_last, sg.item_end * kUnitsPerItem);
    uint32_t x = u + (seg_end - u) * wave / kStagedWaves;
    uint32_t x_end = u + (seg_end - u) * (wave + 1) / kStagedWaves;
    UnitRow R;
    if (x < x_end) load_unit_row(m, unit_piece(sg, x), x % kUnitsPerItem, lane, R);
    __builtin_amdgcn_s_waitcnt(0);   // this wave's part of the image has arrived
    __syncthreads();
    for (;;) {
        // unrolled by two: rows of unit x in R while those of x + 1 load into R2,
        // then the other way round
        UnitRow R2;
        uint32_t c1, c2;
        while (x < x_end) {
            const uint32_t p = unit_piece(sg, x), s = x % kUnitsPerItem;
            ++x;
            unit_row_ready(R);
            if (x < x_end) load_unit_row(m, unit_piece(sg, x), x % kUnitsPerItem, lane, R2);
            staged_unit<kF, kLast>(m, p, s, sg.chunk, R, image, lane, region, fill, c1, c2);
            acc_add(m, S, A, R.h0.z, c1, c2, lane);
            if (x >= x_end) break;
            const uint32_t p2 = unit_piece(sg, x), s2 = x % kUnitsPerItem;
            ++x;
            unit_row_ready(R2);
            if (x < x_end) load_unit_row(m, unit_piece(sg, x), x % kUnitsPerItem, lane, R);
            staged_unit<kF, kLast>(m, p2, s2, sg.chunk, R2, image, lane, region, fill, c1, c2);
            acc_add(m, S, A, R2.h0.z, c1, c2, lane);
        }
        if (seg_end >= u_last) break;
        // the next segment: the next chunk with items (the last entry of the
        // table holds the item count, above every item of the run); its first
        // unit's rows are in flight while the block waits for every wave to
        // finish with the image
        u = seg_end;
        uint32_t c = sg.chunk + 1;
        while (chunks[c + 1].y <= u / kUnitsPerItem) ++c;
        sg = chunk_segment(chunks, c);
        seg_end = min(u_last, sg.item_end * kUnitsPerItem);
        x = u + (seg_end - u) * wave / kStagedWaves;
        x_end = u + (seg_end - u) * (wave + 1) / kStagedWaves;
        if (x < x_end) load_unit_row(m, unit_piece(sg, x), x % kUnitsPerItem, lane, R);
        __syncthreads();
        stage_async(m, sg.chunk, (uint8_t*)image, wave, lane);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    }
    acc_flush(m, S, A, lane);
    // records past the region's capacity: the host sizes the regions from the
    // largest count and runs the batch again
    if (fill > m.item_cap && lane == 0) atomicMax((unsigned long long*)m.cand_max, (unsigned long long)fill);
    const uint32_t listed = min(fill, m.item_cap);
    // the residual chain over this wave's survivors, 64 at a time, as soon as the
    // wave's units are done (the other waves may still be matching): the filter
    // tables read from global memory, where they stay in the caches
    const Staged C{m.allow, m.mods, nullptr, m.ng, S.queue};
    chain_records(m, C, lane, region, listed);
    __syncthreads();   // every wave is done with its unit counts
    for (uint32_t t = threadIdx.x; t < (uint32_t)(kTileSlots * kItemGroups); t += kStagedThreads) {
        const uint32_t n = tilecnt[t];
        if (n) atomicAdd(&m.counters[(uint64_t)(t / kItemGroups) * (m.ng + 3) + t % kItemGroups + 1], (uint64_t)n);
    }
}

// LDS read probe, the calibration of the staged engine's LDS roofline: each
// thread keeps 16 ds_read_b128 in flight per round over a 64 KB buffer.
// Pattern 0: consecutive 16-byte units (no bank conflicts, the peak);
// 1: the staged engine's reads (random lines, each step's lane-rotated column
// and half); 2: every lane reads a random 16-byte unit (an unrotated layout).
// Every address stays below 64 KB (lines < 256).  A round costs two 3-input
// ORs per read, so the VALU keeps pace with the LDS.
constexpr int kProbeUnits = 4096;
constexpr int kProbeThreads = 1024;   // one block of 16 waves per CU
__global__ __launch_bounds__(kProbeThreads) void k_lds_probe(uint32_t pattern, uint32_t rounds, uint32_t* out) {
    __shared__ uint4 buf[kProbeUnits];
    for (uint32_t t = threadIdx.x; t < (uint32_t)kProbeUnits; t += kProbeThreads)
        buf[t] = make_uint4(t, 3 * t, 5 * t, 7 * t);
    __syncthreads();
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave, g = lane & 15;
    uint32_t a[16];
#pragma unroll
    for (uint32_t i = 0; i < 16; ++i) {
        uint32_t h = (lane * 977u + wave * 131u + i * 7919u + blockIdx.x * 104729u) * 0x9e3779b1u;
        h ^= h >> 15;
        if (pattern == 0) {
            a[i] = (((wave * 16 + i) * kWave + lane) * 16) & 0xfff0u;
        } else if (pattern == 1) {
            const uint32_t c = i ^ g;
            a[i] = (h >> 8) % 256u * kStageLineBytes + ((c & 3) + 4 * (c >> 3)) * 32 + ((c >> 2) & 1) * 16;
        } else {
            a[i] = (h >> 8) % (uint32_t)kProbeUnits * 16;
        }
    }
    uint32_t acc0 = 0, acc1 = 0;
    const uint8_t* base = (const uint8_t*)buf;
    for (uint32_t r = 0; r < rounds; ++r) {
        asm volatile("" ::: "memory");   // the reads of every round are issued again
        uint4 v[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = *(const uint4*)(base + a[i]);
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            acc0 = acc0 | v[i].x | v[i].y;
            acc1 = acc1 | v[i].z | v[i].w;
        }
    }
    out[blockIdx.x * kProbeThreads + threadIdx.x] = acc0 ^ acc1;
}

// Integer-VALU issue probe, the calibration of the VALU peak: eight independent
// add/xor chains per thread, kValuProbeOps 32-bit VALU instructions per round
// (the count of this build's loop body), 4 blocks of 256 threads per CU.
constexpr uint32_t kValuProbeOps = 16;
__global__ __launch_bounds__(256) void k_valu_probe(uint32_t rounds, uint32_t* out) {
    uint32_t a[8], b[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        a[k] = threadIdx.x + 977u * k + blockIdx.x;
        b[k] = 0x9e3779b9u * (k + 1);
    }
    for (uint32_t r = 0; r < rounds; ++r) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            a[k] += b[k];
            b[k] ^= a[k];
        }
    }
    uint32_t x = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) x ^= a[k] ^ b[k];
    out[blockIdx.x * 256 + threadIdx.x] = x;
}

#define PCUBE_STAGED(F)                                                                               \
    template __global__ void k_match_staged<F, false>(MatchArgs, const uint2*, const uint2*); \
    template __global__ void k_match_staged<F, true>(MatchArgs, const uint2*, const uint2*);
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
