/*
 * CPU oracle for the p=619 doubly-focused pseudocube batch engine.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product package links, loads or
 * calls this file; only tests/, the smoke check and the CPU-baseline leg of
 * the benchmark use it, and only as the checker.
 *
 * It restates, in plain C with unsigned __int128, the reference algorithm:
 *   oracle_partial  - CRT partial sums of one ring group
 *                     (pseudocube/gpu.py:82-92 `partial`, pseudocube/focus.py:49-56)
 *   oracle_window   - sorted focus coordinates in [low, high), period by period,
 *                     two wraps, two monotone pointers over sorted C and D
 *                     (pseudocube/gpu.py:94-114, review/focus_window_cpu.cpp:4-25,
 *                     docs/mathematics.md:32-34)
 *   oracle_match    - the residue-bitset batch matcher: bounds by exact 128-bit
 *                     comparison, first-four-filter bitsets, x = uB - vA built
 *                     as unsigned __int128, later filters by x mod q, group
 *                     counting every fourth filter, per-subtile counters and
 *                     output slots (pseudocube/batch_experiment.py:184-235).
 *
 * Residues in the filter chain are taken from the full 128-bit x, not from a
 * decomposition, so the GPU kernels' per-coordinate residue tables are checked
 * against an independent computation.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

enum { ORACLE_OK = 0, ORACLE_OVERFLOW = 1, ORACLE_INVALID = 2 };

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

/* out[i] = sum over rings k of terms_k[digit_k(i)] mod M, first ring least
 * significant in the mixed radix.  Returns 2 if n is not the ring product. */
int oracle_partial(const uint64_t *terms, const uint32_t *sizes, uint32_t nrings,
                   uint64_t M, uint64_t *out, uint64_t n) {
    uint64_t expect = 1;
    for (uint32_t k = 0; k < nrings; ++k) expect *= sizes[k];
    if (expect != n || M == 0 || M >= (1ULL << 63)) return ORACLE_INVALID;
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t rest = i, sum = 0, offset = 0;
        for (uint32_t k = 0; k < nrings; ++k) {
            sum += terms[offset + rest % sizes[k]];
            if (sum >= M) sum -= M;
            rest /= sizes[k];
            offset += sizes[k];
        }
        out[i] = sum;
    }
    return ORACLE_OK;
}

/* Every coordinate y in [low, high) with y mod M = (c + d) mod M for some
 * c in C, d in D (both sorted ascending, values in [0, M)), written in
 * ascending order.  Returns 1 with *count = cap + 1 when more than cap exist. */
int oracle_window(const uint64_t *C, uint64_t nc, const uint64_t *D, uint64_t nd, uint64_t M,
                  uint64_t low, uint64_t high, uint64_t *out, uint64_t cap, uint64_t *count) {
    *count = 0;
    if (low > high || M == 0 || M >= (1ULL << 63)) return ORACLE_INVALID;
    if (low == high) return ORACLE_OK;
    uint64_t written = 0;
    for (uint64_t period = low / M; period <= (high - 1) / M; ++period) {
        uint64_t base = period * M;
        uint64_t a = low > base ? low - base : 0;
        uint64_t b = high - base < M ? high - base : M;
        for (int wrap = 0; wrap < 2; ++wrap) {
            /* c + d - wrap*M in [a, b)  <=>  c in [a + wrap*M - d, b + wrap*M - d).
             * Walking d downwards moves that interval upwards, so both
             * pointers into C only advance. */
            uint64_t lw = a + (wrap ? M : 0), hw = b + (wrap ? M : 0);
            uint64_t pl = 0, pr = 0;
            for (uint64_t j = nd; j-- > 0;) {
                uint64_t d = D[j];
                if (hw <= d) continue;
                uint64_t cl = lw > d ? lw - d : 0, ch = hw - d;
                if (cl >= M) break;
                while (pl < nc && C[pl] < cl) ++pl;
                if (pr < pl) pr = pl;
                while (pr < nc && C[pr] < ch) ++pr;
                for (uint64_t i = pl; i < pr; ++i) {
                    if (written >= cap) {
                        *count = written + 1;
                        return ORACLE_OVERFLOW;
                    }
                    out[written++] = C[i] + d - (wrap ? M : 0) + base;
                }
            }
        }
    }
    qsort(out, written, sizeof(uint64_t), cmp_u64);
    *count = written;
    return ORACLE_OK;
}

/* ---------------------------------------------------------------------- */
/* Batch matcher                                                           */

static int is_allowed(uint32_t q, uint32_t a) {
    if (q == 9) return a == 1 || a == 8;
    if (a == 0) return 0;
    if (q % 3 != 1) return 1;
    /* cubic residue by exhaustive search over cubes */
    for (uint32_t t = 1; t < q; ++t)
        if ((uint64_t)t * t % q * t % q == a) return 1;
    return 0;
}

typedef struct {
    const uint64_t *us, *vs;
    uint64_t nu, nv, A, B, v_first, tile_width, cap;
    u128 lo, hi;
    const uint32_t *mod;
    uint32_t nres, ng, nbits, ntiles, words;
    uint8_t **allow;          /* allow[k][m] = 1 if residue m passes filter k */
    uint64_t **mask;          /* mask[k][r * words + w] for the first nbits filters */
    uint64_t *counters, *values;
    volatile int error;
    uint64_t next;            /* work cursor (rows or mask rows) */
} job_t;

static u128 mul64(uint64_t a, uint64_t b) { return (u128)a * b; }

/* first i in [0, n) with us[i]*B > bound, or n */
static uint64_t first_above(const uint64_t *us, uint64_t n, uint64_t B, u128 bound) {
    uint64_t l = 0, h = n;
    while (l < h) {
        uint64_t m = l + (h - l) / 2;
        if (mul64(us[m], B) > bound) h = m; else l = m + 1;
    }
    return l;
}

static void *build_masks(void *arg) {
    job_t *J = (job_t *)arg;
    uint32_t total = 0;
    for (uint32_t k = 0; k < J->nbits; ++k) total += J->mod[k];
    for (;;) {
        uint64_t row = __atomic_fetch_add(&J->next, 1, __ATOMIC_RELAXED);
        if (row >= total) break;
        uint32_t k = 0, r = (uint32_t)row;
        while (r >= J->mod[k]) r -= J->mod[k++];
        uint32_t q = J->mod[k];
        uint32_t sub = (uint32_t)((uint64_t)r * (J->A % q) % q);
        uint32_t bq = (uint32_t)(J->B % q);
        uint64_t *dst = J->mask[k] + (uint64_t)r * J->words;
        memset(dst, 0, J->words * sizeof(uint64_t));
        for (uint64_t i = 0; i < J->nu; ++i) {
            uint32_t m = (uint32_t)(((J->us[i] % q) * bq + q - sub) % q);
            if (J->allow[k][m]) dst[i >> 6] |= 1ULL << (i & 63);
        }
    }
    return NULL;
}

static void *match_rows(void *arg) {
    job_t *J = (job_t *)arg;
    const uint32_t ng = J->ng, stride = ng + 3;
    uint64_t *pop = (uint64_t *)calloc(ng + 1, sizeof(uint64_t));
    for (;;) {
        uint64_t j0 = __atomic_fetch_add(&J->next, 64, __ATOMIC_RELAXED);
        if (j0 >= J->nv || J->error) break;
        uint64_t j1 = j0 + 64 < J->nv ? j0 + 64 : J->nv;
        for (uint64_t j = j0; j < j1; ++j) {
            uint64_t v = J->vs[j];
            if (v < J->v_first || (v - J->v_first) / J->tile_width >= J->ntiles) { J->error = 1; break; }
            uint64_t tile = (v - J->v_first) / J->tile_width;
            uint64_t *cnt = J->counters + tile * stride;
            memset(pop, 0, (ng + 1) * sizeof(uint64_t));
            u128 shift = mul64(v, J->A);
            uint64_t left = first_above(J->us, J->nu, J->B, shift + J->lo);
            uint64_t right = first_above(J->us, J->nu, J->B, shift + J->hi);
            const uint64_t *rows[4];
            for (uint32_t k = 0; k < J->nbits; ++k)
                rows[k] = J->mask[k] + (v % J->mod[k]) * J->words;
            uint64_t underflow = 0;
            for (uint64_t w = left >> 6; w < ((right + 63) >> 6); ++w) {
                uint64_t base = w << 6, bits = ~0ULL;
                if (base < left) bits &= ~0ULL << (left - base);
                if (right < base + 64) bits &= (1ULL << (right - base)) - 1;
                pop[0] += (uint64_t)__builtin_popcountll(bits);
                for (uint32_t k = 0; k < J->nbits; ++k) bits &= rows[k][w];
                while (bits) {
                    uint64_t i = base + (uint64_t)__builtin_ctzll(bits);
                    bits &= bits - 1;
                    u128 positive = mul64(J->us[i], J->B);
                    if (positive < shift) { ++underflow; continue; }
                    u128 x = positive - shift;
                    ++pop[1];
                    int pass = 1;
                    for (uint32_t k = 4; k < J->nres; ++k) {
                        if (!J->allow[k][(uint32_t)(x % J->mod[k])]) { pass = 0; break; }
                        if (k % 4 == 3 || k + 1 == J->nres) ++pop[k / 4 + 1];
                    }
                    if (!pass) continue;
                    uint64_t slot = __atomic_fetch_add(&cnt[ng + 1], 1, __ATOMIC_RELAXED);
                    if (slot < J->cap) {
                        uint64_t *dst = J->values + (tile * J->cap + slot) * 2;
                        dst[0] = (uint64_t)x;
                        dst[1] = (uint64_t)(x >> 64);
                    }
                }
            }
            for (uint32_t g = 0; g <= ng; ++g)
                if (pop[g]) __atomic_fetch_add(&cnt[g], pop[g], __ATOMIC_RELAXED);
            if (underflow) __atomic_fetch_add(&cnt[ng + 2], underflow, __ATOMIC_RELAXED);
        }
    }
    free(pop);
    return NULL;
}

static void run_threads(void *(*fn)(void *), job_t *J, int threads) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t tid[256];
    J->next = 0;
    for (int t = 1; t < threads; ++t) pthread_create(&tid[t], NULL, fn, J);
    fn(J);
    for (int t = 1; t < threads; ++t) pthread_join(tid[t], NULL);
}

/* counters: ntiles x (ng + 3) = [valid, group_1..group_ng, emitted, underflow],
 * zeroed here.  values: ntiles x cap x 2 limbs (lo, hi).  Returns 0, or 2 for
 * invalid arguments (a v outside the tiles, or bad moduli). */
int oracle_match(const uint64_t *us, uint64_t nu, const uint64_t *vs, uint64_t nv,
                 uint64_t A, uint64_t B, uint64_t lo0, uint64_t lo1, uint64_t hi0, uint64_t hi1,
                 const uint32_t *moduli, uint32_t nres, uint64_t v_first, uint64_t tile_width,
                 uint32_t ntiles, uint64_t cap, uint64_t *counters, uint64_t *values, int threads) {
    if (nres == 0 || tile_width == 0 || ntiles == 0) return ORACLE_INVALID;
    job_t J;
    memset(&J, 0, sizeof J);
    J.us = us; J.vs = vs; J.nu = nu; J.nv = nv; J.A = A; J.B = B;
    J.v_first = v_first; J.tile_width = tile_width; J.cap = cap; J.ntiles = ntiles;
    J.lo = ((u128)lo1 << 64) | lo0;
    J.hi = ((u128)hi1 << 64) | hi0;
    J.mod = moduli; J.nres = nres; J.ng = (nres + 3) / 4; J.nbits = nres < 4 ? nres : 4;
    J.words = (uint32_t)((nu + 63) / 64);
    J.counters = counters; J.values = values;
    memset(counters, 0, (size_t)ntiles * (J.ng + 3) * sizeof(uint64_t));
    J.allow = (uint8_t **)calloc(nres, sizeof(uint8_t *));
    for (uint32_t k = 0; k < nres; ++k) {
        uint32_t q = moduli[k];
        if (q < 2 || q > 65535) return ORACLE_INVALID;
        J.allow[k] = (uint8_t *)malloc(q);
        for (uint32_t a = 0; a < q; ++a) J.allow[k][a] = (uint8_t)is_allowed(q, a);
    }
    J.mask = (uint64_t **)calloc(4, sizeof(uint64_t *));
    for (uint32_t k = 0; k < J.nbits; ++k)
        J.mask[k] = (uint64_t *)malloc((size_t)moduli[k] * (J.words ? J.words : 1) * sizeof(uint64_t));
    run_threads(build_masks, &J, threads);
    run_threads(match_rows, &J, threads);
    for (uint32_t k = 0; k < J.nbits; ++k) free(J.mask[k]);
    for (uint32_t k = 0; k < nres; ++k) free(J.allow[k]);
    free(J.mask);
    free(J.allow);
    return J.error ? ORACLE_INVALID : ORACLE_OK;
}

/* Wide-limb known-answer helper: (a*b - c*d) mod 2^128 and the borrow,
 * the contract of the reference's CPU limb oracle (review/wide.c:3-7). */
int oracle_products(uint64_t a, uint64_t b, uint64_t c, uint64_t d, uint64_t *low, uint64_t *high) {
    u128 left = mul64(a, b), right = mul64(c, d), diff = left - right;
    *low = (uint64_t)diff;
    *high = (uint64_t)(diff >> 64);
    return left < right;
}


/* ---- Full-domain control: roots coprime to every prime <= p ------------------
 * Restates review/cube_count.c:12-38: the number of r in [first, last] with no
 * prime factor q <= p.  Odd roots only (2 <= p), sieved in byte segments of
 * 2^20 odd roots.  When roots is given, the first roots_cap coprime roots are
 * stored and *nroots receives how many there are in all. */
uint64_t oracle_cube_count(uint64_t first, uint64_t last, uint32_t p, uint64_t *roots, uint64_t roots_cap,
                           uint64_t *nroots) {
    uint32_t primes[400], np = 0;
    for (uint32_t q = 3; q <= p; q += 2) {
        int prime = 1;
        for (uint32_t d = 3; d * d <= q; d += 2)
            if (q % d == 0) { prime = 0; break; }
        if (prime) primes[np++] = q;
    }
    const uint64_t seg = 1u << 20;
    unsigned char *flag = malloc(seg);
    uint64_t count = 0;
    for (uint64_t start = first | 1u; start <= last; start += 2 * seg) {
        uint64_t n = (last - start) / 2 + 1;
        if (n > seg) n = seg;
        memset(flag, 1, n);
        for (uint32_t k = 0; k < np; ++k) {
            const uint64_t q = primes[k];
            uint64_t m = (start + q - 1) / q * q;
            if (!(m & 1)) m += q;
            for (uint64_t i = (m - start) / 2; i < n; i += q) flag[i] = 0;
        }
        for (uint64_t i = 0; i < n; ++i)
            if (flag[i]) {
                if (roots && count < roots_cap) roots[count] = start + 2 * i;
                ++count;
            }
    }
    free(flag);
    if (nroots) *nroots = count;
    return count;
}
