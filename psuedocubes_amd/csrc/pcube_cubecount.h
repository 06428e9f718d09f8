#pragma once
// Full-domain control on the GPU: how many roots r in [first, last] are
// coprime to every prime q <= p, i.e. how many genuine cubes r^3 the sieve's
// domain holds (the reference counts 38,322,775 for r in [1000000001,
// 1442249570], p = 619, with review/cube_count.c and
// review/cube_count_independent.c, run by tools/check_cpu.py:24-29).
//
// Odd roots only (2 <= p divides every even one).  Block b sieves odd indices
// [b*kCubeSegment, (b+1)*kCubeSegment), odd index i standing for root
// r0 + 2i, as a bitset in LDS: for each odd prime q <= p the threads of the
// block mark q's odd multiples in the segment, stride q, with LDS atomics;
// then every thread popcounts its words of the unmarked set into one global
// counter.  When `good` is given (short ranges), the unmarked words are also
// written out for the host to list the roots.
#include "pcube_kernels.h"

namespace pcube {

constexpr uint32_t kCubeSegment = 1u << 18;   // odd indices per block: 32 KB of bits
constexpr uint32_t kCubeThreads = 1024;

__global__ __launch_bounds__(kCubeThreads) void k_cube_count(uint64_t r0, uint64_t nodd, const uint32_t* primes,
                                                             uint32_t nprimes, unsigned long long* count,
                                                             uint32_t* good) {
    __shared__ uint32_t bits[kCubeSegment / 32];
    const uint64_t i0 = (uint64_t)blockIdx.x * kCubeSegment;
    const uint32_t n = (uint32_t)min((uint64_t)kCubeSegment, nodd - i0);
    for (uint32_t w = threadIdx.x; w < kCubeSegment / 32; w += kCubeThreads) bits[w] = 0;
    __syncthreads();
    const uint64_t seg_first = r0 + 2 * i0;   // odd root of index 0 of the segment
    for (uint32_t k = 0; k < nprimes; ++k) {
        const uint64_t q = primes[k];
        // first odd multiple of q at or after seg_first
        uint64_t m = (seg_first + q - 1) / q * q;
        if (!(m & 1)) m += q;
        const uint64_t first = (m - seg_first) / 2;   // its odd index in the segment
        for (uint64_t i = first + (uint64_t)threadIdx.x * q; i < n; i += (uint64_t)kCubeThreads * q)
            atomicOr(&bits[i >> 5], 1u << (i & 31));
    }
    __syncthreads();
    uint32_t mine = 0;
    for (uint32_t w = threadIdx.x; w * 32 < n; w += kCubeThreads) {
        uint32_t g = ~bits[w];
        if (w * 32 + 32 > n) g &= (1u << (n - w * 32)) - 1u;   // n - 32w < 32 here
        mine += __popc(g);
        if (good) good[(i0 >> 5) + w] = g;
    }
#pragma unroll
    for (int d = kWave / 2; d >= 1; d >>= 1) mine += __shfl_xor(mine, d, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0 && mine) atomicAdd(count, (unsigned long long)mine);
}

}  // namespace pcube
