// rng_device.hpp -- counter-based random numbers for the device-side generators.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Every draw of a link l is
// addressed by a counter {lo32(l), hi16(l) | block << 16, lo32(n), hi32(n)}
// with key = seed: n selects the refresh / sweep, block the use:
//   0..3        momenta (one Box-Muller pair each)
//   16..39      hot start (4 attempts x 6 blocks)
//   62          accept word of a batched chain's trajectory (link 0)
//   63          Metropolis accept word
//   64..32767   Metropolis proposals, 8 blocks per update (at most 4000 updates)
//   32768..     heat bath, 256 blocks per SU(2) subgroup
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lqcd {

// first Philox block of the heat-bath draws
constexpr uint32_t kHeatBathBlock = 0x8000u;

__device__ __forceinline__ uint4 philox4x32(uint4 c, uint2 k)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
        const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
    return c;
}

__device__ __forceinline__ uint4 philox_link(uint64_t link, uint32_t block, uint64_t n, uint64_t seed)
{
    const uint4 ctr = make_uint4((uint32_t)link, (uint32_t)((link >> 32) & 0xFFFFu) | (block << 16),
                                 (uint32_t)n, (uint32_t)(n >> 32));
    return philox4x32(ctr, make_uint2((uint32_t)seed, (uint32_t)(seed >> 32)));
}

// 53-bit uniform in [0, 1) from two words
__device__ __forceinline__ double u53(uint32_t a, uint32_t b)
{
    return (double)(((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6)) * 0x1.0p-53;
}

} // namespace lqcd
