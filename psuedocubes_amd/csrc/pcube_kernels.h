// Device-side building blocks shared by the pseudocube kernels (gfx950, wave64).
//
// Exact integer arithmetic only: 64x64->128 products, 128-bit add/sub/compare,
// and Barrett reduction by small moduli (q < 2^16) with integer reciprocals.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcube {

constexpr int kWave = 64;
constexpr int kMaxResiduals = 128;   // residual filters per configuration
constexpr int kMaxBitsetFilters = 4; // filters folded into the residue bitsets
constexpr int kStagedFilters = 8;    // filters folded into the LDS-staged residue bitsets
constexpr int kCachedResidues = 8;   // chain residues precomputed per u coordinate
constexpr int kMaxAllowWords = 2048; // 32-bit words of allowed-residue bits

struct U128 {
    uint64_t lo, hi;
};

__device__ __forceinline__ U128 mul_wide(uint64_t a, uint64_t b) {
    return U128{a * b, __umul64hi(a, b)};
}
__device__ __forceinline__ U128 add_wide(U128 a, U128 b) {
    uint64_t lo = a.lo + b.lo;
    return U128{lo, a.hi + b.hi + (lo < a.lo ? 1ull : 0ull)};
}
__device__ __forceinline__ U128 sub_wide(U128 a, U128 b) {
    return U128{a.lo - b.lo, a.hi - b.hi - (a.lo < b.lo ? 1ull : 0ull)};
}
__device__ __forceinline__ bool less_wide(U128 a, U128 b) {
    return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo);
}

// floor(x / B) for any x < 2^128, with mu = floor((2^128 - 1) / B) (Barrett).
// t = floor(x * mu / 2^128), formed exactly from the four 64 x 64 products, is
// floor(x / B) or one less, so one comparison of the remainder settles it.
__device__ __forceinline__ U128 floor_div_wide(U128 x, uint64_t B, U128 mu) {
    const U128 p0 = mul_wide(x.lo, mu.lo), a = mul_wide(x.lo, mu.hi), b = mul_wide(x.hi, mu.lo);
    const U128 c = mul_wide(x.hi, mu.hi);
    // the 2^64 column: p0.hi + a.lo + b.lo and its carry into the 2^128 column
    uint64_t s = p0.hi + a.lo;
    uint64_t carry = s < a.lo ? 1ull : 0ull;
    s += b.lo;
    carry += s < b.lo ? 1ull : 0ull;
    U128 t = add_wide(add_wide(c, U128{a.hi, 0}), add_wide(U128{b.hi, 0}, U128{carry, 0}));
    // r = x - t*B (mod 2^128) lies in [0, 2B)
    U128 tb = mul_wide(t.lo, B);
    tb.hi += t.hi * B;
    const U128 r = sub_wide(x, tb);
    if (r.hi != 0 || r.lo >= B) t = add_wide(t, U128{1, 0});
    return t;
}

// The least u with u*B > bound, as a 64-bit key (u >= key exactly when
// u*B > bound); beyond every 64-bit u when inf.
struct AboveKey {
    uint64_t key;
    bool inf;
};

__device__ __forceinline__ AboveKey above_key(U128 bound, uint64_t B, U128 mu) {
    const U128 q = floor_div_wide(bound, B, mu);
    return AboveKey{q.lo + 1, q.hi != 0 || q.lo == ~0ull};
}

// u*B <= bound, for the bound the key was made from
__device__ __forceinline__ bool below_key(uint64_t u, const AboveKey& k) {
    return k.inf || u < k.key;
}

// Per-modulus constants for exact reduction.
//   q     modulus (2 <= q < 2^16)
//   inv   floor(2^32 / q): Barrett factor, x - q*mulhi(x, inv) lies in [0, 2q)
//   r32   2^32 mod q
//   r64   2^64 mod q
//   a, b  A mod q, B mod q
//   bit   offset of this modulus's row in the allowed-residue bit table
struct Modulus {
    uint32_t q, inv, r32, r64, a, b, bit, pad;
};

__device__ __forceinline__ uint32_t mod32(uint32_t x, const Modulus& m) {
    uint32_t r = x - m.q * __umulhi(x, m.inv);
    return r >= m.q ? r - m.q : r;
}
__device__ __forceinline__ uint32_t mod64(uint64_t x, const Modulus& m) {
    uint32_t h = mod32(uint32_t(x >> 32), m), l = mod32(uint32_t(x), m);
    return mod32(h * m.r32 + l, m);  // < q^2 + q < 2^32 for q < 2^16
}
__device__ __forceinline__ uint32_t mod128(U128 x, const Modulus& m) {
    return mod32(mod64(x.hi, m) * m.r64 + mod64(x.lo, m), m);
}
__device__ __forceinline__ bool allowed_bit(const uint32_t* bits, uint32_t index) {
    return (bits[index >> 5] >> (index & 31)) & 1u;
}

// Full wave execution barrier with LDS ordering between lanes of one wave.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
// A 64-bit value of lane src, in every lane.
__device__ __forceinline__ uint64_t readlane64(uint64_t x, int src) {
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(x >> 32), src) << 32) |
           (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, src);
}

// Number of set bits of `mask` at lanes below this one.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

}  // namespace pcube
