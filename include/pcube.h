/*
 * libpcube: MI355X (gfx950) engine for the p=619 doubly-focused pseudocube
 * sieve of AlexanderEastwood/psuedocubes, behind a plain C ABI.
 *
 * Conventions
 *   - Device buffers are passed as uint64_t addresses obtained from
 *     pcube_alloc (or offsets into them); host buffers as ordinary pointers.
 *   - Every call returns a status code; pcube_last_error() describes the most
 *     recent failure of the calling thread.  The Python layer maps the codes to the reference's
 *     exceptions: OVERFLOW -> OverflowError, INVALID -> ValueError,
 *     UNDERFLOW/RANGE/MISMATCH -> ArithmeticError, NOMEM -> MemoryError.
 *   - All work runs on one library-owned HIP stream of default priority; each
 *     entry point returns after its results are on the host unless it says
 *     otherwise.  Calls may come from any host thread: one library lock
 *     serialises them, and each makes the library's device current on its
 *     thread before it allocates or launches.
 *   - 128-bit quantities are split into (low, high) uint64 limbs.
 */
#ifndef PCUBE_H
#define PCUBE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PCUBE_OK = 0,
    PCUBE_OVERFLOW = 1,   /* output or focus buffer overflow                 */
    PCUBE_INVALID = 2,    /* invalid arguments or geometry                    */
    PCUBE_UNDERFLOW = 3,  /* unsigned underflow flagged while building x      */
    PCUBE_NOMEM = 4,      /* device memory cap exceeded                        */
    PCUBE_DEVICE = 5,     /* HIP runtime failure                               */
    PCUBE_RANGE = 6,      /* an emitted x fell outside (lo, hi]               */
    PCUBE_MISMATCH = 7    /* the survivor list disagrees with the counters    */
};

const char *pcube_last_error(void);

/* --- runtime: replaces pseudocube/gpu.py:38-67 (Runtime) ----------------- */

/* Number of visible devices (replaces runtime.cp.cuda.runtime.getDeviceCount(),
 * tools/reproduce.py:36). */
int pcube_device_count(int *count);

/* Select the device, create the stream, and fix the memory cap as
 * min(memory_cap, total*fraction, free - 256 MiB) (pseudocube/gpu.py:42-45).
 * cap/free/total may be NULL. */
int pcube_init(int device, uint64_t memory_cap, double fraction, uint64_t *cap, uint64_t *free_bytes,
               uint64_t *total_bytes);

/* Reserved and peak device bytes of the capped pool (pseudocube/gpu.py:49-52). */
int pcube_memory(uint64_t *reserved, uint64_t *peak, uint64_t *cap);

/* Address of the pool's peak device bytes, for a caller that checks the cap
 * before every step without a call (pseudocube/gpu.py:49-52 reads the pool's
 * total each time); the word is updated as the pool grows. */
const uint64_t *pcube_memory_peak(void);

/* Device name and compute-unit count. */
int pcube_device_info(char *name, size_t name_len, int *compute_units);

/* Kernel attributes for the compile record (pseudocube/gpu.py:54-61).
 * index in [0, count); returns INVALID past the end. */
int pcube_kernel_info(int index, char *name, size_t name_len, int *registers, int *local_bytes,
                      int *shared_bytes);

/* --- device buffers (the capped pool behind every array) ------------------ */
/* Every library kernel and copy runs on the library's one stream, so
 * pcube_free returns a block to the pool at once, without waiting: work that
 * later reuses the block is enqueued behind every earlier use of it.  Blocks
 * go back to the driver only after that stream has drained. */
int pcube_alloc(uint64_t bytes, uint64_t *address);
int pcube_free(uint64_t address);
int pcube_upload(uint64_t dst, const void *src, uint64_t bytes);
int pcube_download(void *dst, uint64_t src, uint64_t bytes);

/* --- producer/consumer staging: replaces the slots and the CPU producer of
 * tools/batch_window_experiment.py:52-82, 256-336 -------------------------- */

/* Page-locked host memory for the staging slots between a window producer and
 * a matcher consumer (the reference's shared-memory slots of 2 x 1e6 uint64,
 * tools/batch_window_experiment.py:314-318); copies to and from it run as DMA.
 * pcube_pinned_free waits for the library stream before it frees. */
int pcube_pinned_alloc(uint64_t bytes, void **host);
int pcube_pinned_free(void *host);

/* Host window producer with the arguments and codes of the reference's CPU
 * producer focus_window (review/focus_window_cpu.cpp:4-6, bound at
 * tools/batch_window_experiment.py:58-64): the focus coordinates in
 * [lo, hi) of one period (0 <= lo <= hi <= M < 2^63), offset by base, from the
 * ascending half arrays C and D, written in ascending order (the reference
 * sorts after the call, tools/batch_window_experiment.py:82).  Returns 0, 1
 * with *count = cap + 1 when more than cap exist (as the reference: written + 1),
 * or 2. */
int pcube_host_focus_window(const uint64_t *C, uint64_t nc, const uint64_t *D, uint64_t nd, uint64_t M,
                            uint64_t lo, uint64_t hi, uint64_t base, uint64_t *out, uint64_t cap,
                            uint64_t *count);

/* --- focus wheels: replaces pseudocube/gpu.py:70-114 (GPUFocus) ----------- */

/* Build the two sorted half arrays C and D of one focus on the device
 * (partial CRT sums, pseudocube/gpu.py:82-92) plus a bucket directory of C.
 * terms: concatenated CRT terms of the C rings then the D rings;
 * sizes: ring sizes, nrings_c for C followed by nrings_d for D. */
int pcube_focus_create(uint64_t modulus, const uint64_t *terms, const uint32_t *sizes, uint32_t nrings_c,
                       uint32_t nrings_d, int *focus, uint64_t *nc, uint64_t *nd);
int pcube_focus_arrays(int focus, uint64_t *c_address, uint64_t *d_address);
int pcube_focus_destroy(int focus);

/* Sorted focus coordinates in [low, high) written to the device buffer out
 * (capacity cap), count in *count.  Same codes as the reference's CPU
 * producer (review/focus_window_cpu.cpp:4-6): 0 ok, 1 overflow (*count is the
 * full size, > cap), 2 invalid.  Replaces GPUFocus.window (pseudocube/gpu.py:94-114). */
int pcube_focus_window(int focus, uint64_t low, uint64_t high, uint64_t out, uint64_t cap, uint64_t *count);

/* Attach the CRT wheel generator to a focus: ring moduli in ascending order,
 * their allowed residue counts and the concatenated (coefficient-adjusted)
 * residues, exactly Focus.rings (pseudocube/focus.py:35-47, 108-114).  The
 * inner ring prefix takes every ring that keeps its modulus M1 < 2^40 and at
 * most 2^27 residues (8 bytes each on the device); later windows of this
 * focus are generated as t*M1 + r and need no sort.
 * inner_modulus/inner_residues may be NULL. */
int pcube_focus_wheel(int focus, const uint32_t *moduli, const uint32_t *sizes, const uint32_t *residues,
                      uint32_t nrings, uint64_t *inner_modulus, uint32_t *inner_residues);

/* n windows [lows[w], highs[w]) of one focus in one pass, concatenated in
 * window order into out (capacity cap), each ascending; offsets[0..n] on the
 * host delimit them.  OVERFLOW (nothing written, offsets[n] = total) when the
 * total exceeds cap.  Needs the wheel.  The batched generalisation of
 * GPUFocus.window (pseudocube/gpu.py:94-114). */
int pcube_focus_windows(int focus, uint32_t n, const uint64_t *lows, const uint64_t *highs, uint64_t out,
                        uint64_t cap, uint64_t *offsets);

/* prepare() of one compute batch in one call (pseudocube/batch_experiment.py:51-57):
 * the u window [u_min of the first subtile, u_end of the last) into us_out,
 * the v window [v_first of the first, v_end of the last) into vs_out, and
 * the subtile bounds (a, b, c, d) = searchsorted of (u_min, u_end) in us and
 * (v_first, v_end) in vs.  tiles: ntiles x (v_first, v_end, u_min, u_end),
 * 1 <= ntiles <= 64 (INVALID otherwise).
 * OVERFLOW (nu and nv set, window contents unspecified) when a window exceeds its cap. */
int pcube_prepare(int focus_u, int focus_v, uint32_t ntiles, const uint64_t *tiles, uint64_t us_out, uint64_t us_cap,
                  uint64_t vs_out, uint64_t vs_cap, uint64_t *nu, uint64_t *nv, uint64_t *bounds);

/* tile_ranges(first, end, width) of the reference (pseudocube/batch_experiment.py:27-32)
 * in exact 128-bit host arithmetic: up to 64 x (v_first, v_end, u_min, u_end)
 * into tiles, *ntiles of them, with u_min = (lo + v A) / B + 1 at a subtile's
 * first v and u_end = (hi + v A) / B + 1 at its last.  lo = lo1:lo0, hi =
 * hi1:hi0.  INVALID for an empty range, end > B, more than 64 subtiles, A or B
 * >= 2^63, or lo or hi >= 2^96.  Host only. */
int pcube_tile_ranges(uint64_t A, uint64_t B, uint64_t lo0, uint64_t lo1, uint64_t hi0, uint64_t hi1, uint64_t first,
                      uint64_t end, uint64_t width, uint64_t *tiles, uint32_t *ntiles);

/* pcube_tile_ranges of n windows [firsts[w], ends[w]) with subtile widths
 * widths[w], concatenated in window order into tiles (room for tiles_cap
 * subtiles, 4 uint64 each), counts[w] subtiles for window w.  INVALID (naming
 * the window) as pcube_tile_ranges, or when the tiles do not fit.  Host only:
 * the geometry of the batched form's windows in one call. */
int pcube_tile_ranges_many(uint64_t A, uint64_t B, uint64_t lo0, uint64_t lo1, uint64_t hi0, uint64_t hi1, uint32_t n,
                           const uint64_t *firsts, const uint64_t *ends, const uint64_t *widths, uint64_t *tiles,
                           uint64_t tiles_cap, uint32_t *counts);

/* prepare() of one compute batch from its canonical v range [first, end) and
 * subtile width (pseudocube/batch_experiment.py:27-32, 51-57): the geometry of
 * pcube_tile_ranges into tiles, then pcube_prepare on it.  totals (host, 3
 * uint64, may be NULL) receives the batch's focus_u = sum of (u_end - u_first),
 * its cartesian opportunities = sum of (u_end - u_first) * (v_end - v_first)
 * over the subtile bounds, and the largest single subtile's opportunities
 * (pseudocube/batch_experiment.py:53-57, 150-166).  Errors as
 * pcube_tile_ranges, and RANGE when a window spans more than 4096 focus
 * periods (nothing is generated). */
int pcube_prepare_range(int focus_u, int focus_v, uint64_t A, uint64_t B, uint64_t lo0, uint64_t lo1, uint64_t hi0,
                        uint64_t hi1, uint64_t first, uint64_t end, uint64_t width, uint64_t *tiles, uint32_t *ntiles,
                        uint64_t us_out, uint64_t us_cap, uint64_t vs_out, uint64_t vs_cap, uint64_t *nu, uint64_t *nv,
                        uint64_t *bounds, uint64_t *totals);

/* np.searchsorted(sorted, queries, 'left') on the device; host queries/out
 * (pseudocube/batch_experiment.py:41-42). */
int pcube_searchsorted(uint64_t sorted, uint64_t n, const uint64_t *queries, uint64_t nq, uint64_t *out);

/* --- matcher: replaces pseudocube/batch_experiment.py:79-254 -------------- */

/* Residual moduli in filter order (cubic ascending, then the rest,
 * pseudocube/gpu.py:134); at most 128. */
int pcube_matcher_create(uint64_t A, uint64_t B, uint64_t lo0, uint64_t lo1, uint64_t hi0, uint64_t hi1,
                         const uint32_t *moduli, uint32_t nres, int *matcher);
int pcube_matcher_destroy(int matcher);

/* Overwrite the matcher's device allowed-residue tables (the allowed bits of
 * every filter and the bitset builders' permuted copies) with all residues
 * allowed (value 1) or none (value 0).  Every later batch of the matcher, in
 * any engine, reads the mutated tables.  For tests that a corrupted device
 * table cannot reach a receipt: the reference fills its device byte table
 * with 1s and requires host verification to raise ArithmeticError
 * (REL!tools/gpu_worker.py:96-103, `bad.table.fill(1)`). */
int pcube_matcher_table_fill(int matcher, int value);

/* Match one batch of up to 64 subtiles (or one subtile with ntiles = 1).
 *   us/vs        device addresses of sorted u and v coordinates
 *   v_first      tile_id = (v - v_first) / tile_width must be < ntiles
 *   counters     host, ntiles * (ng + 3) uint64: [valid, group_1..group_ng,
 *                emitted, underflow] per subtile, ng = ceil(nres / 4)
 *                (pseudocube/batch_experiment.py:207-234)
 *   values       host, min(emitted, cap) (low, high) limb pairs per subtile,
 *                concatenated in subtile order; values_cap pairs available
 *   engine       0 = residue-bitset engine, 1 = direct per-position engine
 *   seconds      host, 3 doubles (may be NULL): prefilter build (rows, bitsets,
 *                residue cache), matching (every kernel after the prefilter),
 *                and of that the staged match kernel alone (engine 0)
 * Returns OK, OVERFLOW (a subtile emitted more than cap; counters valid),
 * UNDERFLOW, RANGE, MISMATCH (the survivor list disagrees with the emitted
 * counts; counters valid, values not), INVALID. */
int pcube_match(int matcher, uint64_t us, uint64_t nu, uint64_t vs, uint64_t nv, uint64_t v_first,
                uint64_t tile_width, uint32_t ntiles, uint64_t cap, int engine, uint64_t *counters,
                uint64_t *values, uint64_t values_cap, double *seconds);

/* pcube_match in two halves, so the caller can prepare the host side of the
 * result row while the device works: pcube_match_start validates and enqueues
 * the batch (same arguments and meaning as pcube_match) and returns at once;
 * pcube_match_finish waits for it and fills counters, values and seconds as
 * pcube_match does, with the same return codes.  One batch per matcher may be
 * pending; other calls may run in between, except a pcube_lease on the same
 * matcher, which reuses its batch buffers: finish then returns INVALID.  The
 * split of the reference's MaskedMatcher.run
 * (pseudocube/batch_experiment.py:104-167). */
int pcube_match_start(int matcher, uint64_t us, uint64_t nu, uint64_t vs, uint64_t nv, uint64_t v_first,
                      uint64_t tile_width, uint32_t ntiles, uint64_t cap, int engine);
int pcube_match_finish(int matcher, uint64_t *counters, uint64_t *values, uint64_t values_cap, double *seconds);

/* prepare() and the start of the batch's match in one call, so the match is
 * enqueued the moment the subtile bounds are known, with no return to the
 * caller in between (the reference's prepare then MaskedMatcher.run,
 * pseudocube/batch_experiment.py:51-57, 104-167).  Arguments up to totals and
 * the return codes are those of pcube_prepare_range (totals is required here).
 * Then, when the matcher has no batch pending and the largest subtile's
 * opportunities are at most max_opportunities, it runs
 * pcube_match_start(matcher, us_out, *nu, vs_out, *nv, first, width, *ntiles,
 * cap, engine); a start that fails is not reported (pcube_match_start reports
 * it when the caller starts the batch itself).  launch receives a nonzero
 * token when the match was started, else 0. */
int pcube_prepare_match(int focus_u, int focus_v, uint64_t A, uint64_t B, uint64_t lo0, uint64_t lo1, uint64_t hi0,
                        uint64_t hi1, uint64_t first, uint64_t end, uint64_t width, uint64_t *tiles, uint32_t *ntiles,
                        uint64_t us_out, uint64_t us_cap, uint64_t vs_out, uint64_t vs_cap, uint64_t *nu, uint64_t *nv,
                        uint64_t *bounds, uint64_t *totals, int matcher, uint64_t cap, int engine,
                        uint64_t max_opportunities, uint64_t *launch);

/* adoptable = 1 when the batch that pcube_prepare_match started with token
 * launch is still the matcher's pending batch (nothing was started on the
 * matcher and no lease used it since), so pcube_match_finish collects it;
 * else 0 and the caller starts the batch itself.  Host only. */
int pcube_match_adoptable(int matcher, uint64_t launch, int *adoptable);

/* --- lease pipeline: replaces pseudocube/optimized.py:81-90 ---------------- */

/* Run consecutive compute batches (a lease, or one pass of it) with one window
 * per focus over all of them and no host synchronisation between batches.
 *   tiles        host, ntiles_total * 4 uint64: (v_first, v_end, u_min, u_end)
 *                per subtile, exactly tile_ranges() of each batch, batches
 *                consecutive; batch_tiles[b] subtiles belong to batch b
 *   bounds       host out, ntiles_total * 4 uint64: (a, b, c, d) subtile
 *                bounds into the batch's u and v windows
 *   batch_sizes  host out, nbatches * 2 uint64: (nu, nv) per batch
 *   counters     host out, ntiles_total * (ng + 3)
 *   values       host out as in pcube_match, all batches concatenated
 *   seconds      host out, 4 doubles: windows, prefilter, match, total
 * Returns OK or the first failing code (OVERFLOW for any window or subtile
 * overflow, after which counters describe what happened). */
int pcube_lease(int matcher, int focus_u, int focus_v, uint32_t nbatches, const uint32_t *batch_tiles,
                const uint64_t *tiles, uint64_t tile_width, uint64_t cap, uint64_t window_cap, uint64_t *bounds,
                uint64_t *batch_sizes, uint64_t *counters, uint64_t *values, uint64_t values_cap,
                double *seconds);

/* --- batched form: many uneven windows, mixed prime bounds, one device pass --- */

/* A batched engine over up to 3 matchers (configurations sharing A, B and
 * (lo, hi], e.g. p = 313, 433, 619) and the two focus wheels.  cap is the
 * per-subtile output capacity; the persistent match kernel runs
 * blocks_per_cu blocks of 256 threads per compute unit; use_graph captures
 * the fixed-grid launch sequence in a hipGraph on first use and replays it.
 * No reference counterpart as one call: it runs what
 * tools/run_production.py:51-83 / pseudocube/optimized.py:81-90 run window
 * by window, for many windows at once. */
int pcube_batch_create(const int *matchers, uint32_t ncfg, int focus_u, int focus_v, uint64_t cap,
                       uint32_t blocks_per_cu, int use_graph, int *engine);
int pcube_batch_destroy(int engine);

/* Run nwin windows.  Window w uses configuration win_cfg[w], owns
 * win_tiles[w] consecutive subtiles of `tiles` ((v_first, v_end, u_min,
 * u_end) each, as tile_ranges() gives), has subtile width tile_width[w] and
 * engine mode[w] (0 residue bitsets, 1 direct).  Rows are handed to the
 * persistent kernel rows_per_chunk at a time.  Outputs as pcube_lease:
 * bounds (4 per subtile), sizes (nu, nv per window), counters (per window
 * win_tiles[w] x (ng + 3) of its configuration, concatenated), values
 * (emitted pairs in subtile order), seconds (windows, 0, match, total) and
 * the engine's total graph launches. */
int pcube_batch_run(int engine, uint32_t nwin, const uint32_t *win_cfg, const uint32_t *win_tiles,
                    const uint64_t *tiles, const uint64_t *tile_width, const uint32_t *mode, uint32_t rows_per_chunk,
                    uint64_t *bounds, uint64_t *sizes, uint64_t *counters, uint64_t *values, uint64_t values_cap,
                    double *seconds, uint64_t *graph_launches);

/* --- host verification (CPU): the exact local predicate of every emitted value
 * For each (low, high) value: local_ok = local3(n, p) (pseudocube/oracle.py:36-41),
 * is_cube = n is a perfect cube, cube_root = floor(n^(1/3)) (review/verify_cube.py:13-17). */
int pcube_host_verify(const uint64_t *limbs, uint64_t n, uint32_t p, uint8_t *local_ok, uint8_t *is_cube,
                      uint64_t *cube_root);

/* --- host verification of a device pass's emitted values (CPU) ------------
 * The checks of the reference's host collection (pseudocube/batch_experiment.py:121-167)
 * on the values of nsub subtiles at once.  counts[s] values of subtile s follow those
 * of subtile s - 1 in limbs ((low, high) per value).  Each subtile's values are sorted
 * ascending.  A duplicate, a value outside (lo, hi], a canonical v = -x A^-1 mod B
 * outside [firsts[s], ends[s]) (pseudocube/focus.py:116-118), or a value failing
 * local3(x, primes_of[s]) (pseudocube/oracle.py:36-41) returns MISMATCH with *failure
 * 1, 2, 4 or 3.  Otherwise, per value in sorted order, cube[i] = x is a perfect cube
 * and roots[i] = floor(x^(1/3)) (review/verify_cube.py:13-17); text holds each
 * subtile's sorted values as the canonical JSON list of decimal strings with the
 * encoder's final newline (pseudocube/accounting.py:17-31), subtile s at
 * [text_offsets[s], text_offsets[s + 1]), empty when it has no values.  OVERFLOW when
 * the text exceeds text_cap; INVALID when the counts do not sum to nvalues, a prime
 * bound is not a prime = 1 mod 3, or A is not invertible mod B.  Host only. */
int pcube_batch_values(uint64_t nvalues, const uint64_t *limbs, uint32_t nsub, const uint64_t *counts,
                       const uint64_t *firsts, const uint64_t *ends, const uint32_t *primes_of, uint64_t A,
                       uint64_t B, uint64_t lo0, uint64_t lo1, uint64_t hi0, uint64_t hi1, uint8_t *cube,
                       uint64_t *roots, char *text, uint64_t text_cap, uint64_t *text_offsets, uint32_t *failure);

/* --- self test: replaces the `arithmetic` kernel (kernels/pseudocube.cuh:6-11) */
/* out[3i..3i+2] = (low, high, borrow) of in[4i]*in[4i+1] - in[4i+2]*in[4i+3]; host arrays. */
int pcube_arithmetic(const uint64_t *in, uint64_t *out, uint64_t count);

/* --- receipts: the lease transcript's bytes (pseudocube/optimized.py:26-78) --- */

/* The canonical encoding (pseudocube/accounting.py:17-31: sorted keys, no
 * spaces, integers >= 2^53 as decimal strings, a final newline) of one batch's
 * subtile_scopes list, whose records hold end and first as decimal strings,
 * group_survivors, local_survivors, output_sha256 and valid_focused_candidates
 * (pseudocube/batch_experiment.py:149-158).  Subtile k: bounds[2k..2k+1] =
 * (first, end); counters + k*stride = [valid, group_1 .. group_ng, ...];
 * local[k] its local survivors (local NULL: all 0); sha + 64k its 64 hex
 * digits.  Writes at most cap bytes to out, the length to *length; OVERFLOW
 * when they do not fit.  Host only. */
int pcube_scope_records(uint32_t ntiles, const uint64_t *bounds, const uint64_t *counters, uint32_t stride,
                        uint32_t ng, const uint64_t *local, const char *sha, char *out, uint64_t cap,
                        uint64_t *length);

/* pcube_scope_records for the consecutive batches of one device pass: batch b
 * holds per_batch[b] subtiles, and its encoding is out[offsets[b] ..
 * offsets[b+1]) (offsets has nbatches + 1 entries, offsets[0] = 0).  bounds,
 * counters, local and sha as above, over every subtile of the pass in order.
 * OVERFLOW when the encodings do not fit in cap bytes.  Host only. */
int pcube_scope_records_many(uint32_t nbatches, const uint32_t *per_batch, const uint64_t *bounds,
                             const uint64_t *counters, uint32_t stride, uint32_t ng, const uint64_t *local,
                             const char *sha, char *out, uint64_t cap, uint64_t *offsets);

/* --- self test: the division behind u_interval (pseudocube/focus.py:120-121) --- */
/* The row builders find a row's u bounds as first indices with u >= floor((v*A + lo)/B) + 1
 * and u >= floor((v*A + hi)/B) + 1, dividing 128-bit values by B exactly (Barrett, with
 * floor((2^128 - 1)/B) precomputed).  out[2i..2i+1] = (low, high) of floor(x / B) for
 * x = in[3i] + 2^64 in[3i+1] and B = in[3i+2]; host arrays.  INVALID when a B is 0. */
int pcube_floor_div(const uint64_t *in, uint64_t *out, uint64_t count);

/* --- self test: the transpose behind the residue bitsets (pseudocube/batch_experiment.py:184-194) --- */
/* The mask builders form each 64-row block of a filter's bitset as 64 columns (one u
 * coordinate per lane, its bits over 64 residues of v) and transpose it across the
 * wave.  out[64i + r] = row r of matrix i, whose column t is in[64i + t] (bit k of
 * column t = bit t of row k); host arrays of 64 * count words.  INVALID above 2^24
 * matrices. */
int pcube_bit_transpose(const uint64_t *in, uint64_t *out, uint64_t count);

/* --- full-domain control: replaces review/cube_count.c (main 12-38) -------- */
/* Number of roots r in [first, last] coprime to every prime q <= p, counted
 * on the device (the reference's control: 38,322,775 for [1000000001,
 * 1442249570], p = 619, tools/check_cpu.py:24-29).  1 <= first <= last <=
 * 2^32 - 1 and 2 <= p <= 2000, as the reference requires, else INVALID.
 * good (host, may be NULL): one bit per odd root from first | 1, set when the
 * root is coprime, ceil(odds / 32) words; *seconds the kernel's duration. */
int pcube_cube_count(uint64_t first, uint64_t last, uint32_t p, uint64_t *count, uint32_t *good, uint64_t good_words,
                     double *seconds);

/* pcube_cube_count without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops and harnesses that forbid hipMalloc).  Same
 * arguments, limits and results as pcube_cube_count, without the timing.  Its
 * working memory (prime table, counter, root bitset) is one static device
 * array of 264,448 bytes, which each device holds until the process ends;
 * nothing is allocated or freed.  Limit: with good given, the range may hold
 * at most 2^21 odd roots (65,536 bitset words), else INVALID; a count without
 * good has no limit beyond pcube_cube_count's.  Every call on a device shares
 * that device's array and keeps the last call's contents: within a process
 * the library lock and the call's final synchronisation order the calls;
 * calls that overlapped on one device would race on it. */
int pcube_cube_count_noalloc(uint64_t first, uint64_t last, uint32_t p, uint64_t *count, uint32_t *good,
                             uint64_t good_words);

/* --- measurement: calibration probe for the HBM byte counters ------------ */
/* Copies `bytes` between two fresh device buffers with a plain streaming
 * kernel; *seconds is its event-timed duration.  No reference counterpart. */
int pcube_copy_probe(uint64_t bytes, double *seconds);

/* --- measurement: LDS read bandwidth of the staged engine's access pattern - */
/* One 1024-thread block per CU, each thread keeping 16 ds_read_b128 in flight
 * for `rounds` rounds over 64 KB of LDS.  pattern 0: consecutive 16-byte units
 * (the peak); 1: the staged engine's lane-rotated column reads at random lines;
 * 2: random 16-byte units per lane.  *seconds is the event-timed duration of
 * the second of two launches, *bytes the LDS bytes one launch reads.  No
 * reference counterpart. */
int pcube_lds_probe(uint32_t pattern, uint32_t rounds, double *seconds, double *bytes);

/* --- measurement: integer-VALU issue rate --------------------------------- */
/* 32 waves per CU, each lane running eight independent add/xor chains: 16
 * 32-bit VALU instructions per round for `rounds` rounds.  *seconds is the
 * event-timed duration of the second of two launches, *lane_ops the VALU lane
 * operations one launch executes.  No reference counterpart. */
int pcube_valu_probe(uint32_t rounds, double *seconds, double *lane_ops);

#ifdef __cplusplus
}
#endif
#endif
