// geometry.hpp -- periodic lattice geometry shared by host and device code.
//
// Sites are numbered as the reference's LatticePoint::to_index
// (src/lattice.rs:909-915): s = sum_i x_i * prod_{k<i} L_k, x_0 fastest.  The
// device stores each field as a structure of arrays over sites (see
// fields.hpp), so a wavefront of 64 consecutive sites reads 64 consecutive
// values of the same matrix element: coalesced.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lqcd {

constexpr int kMaxDim = 4;
// the field planes of one direction (9 link elements) span less than 2^32
// bytes: the device addresses them with 32-bit offsets
constexpr uint64_t kMaxPlaneBytes = 0xFFFFFFFFull;

struct Geom {
    int dim;
    int extent[kMaxDim];
    int64_t stride[kMaxDim];
    int64_t nsite;   // number of sites
    int64_t npad;    // row length of every SoA plane (>= nsite, multiple of 256)
};

// Offsets to the +dir and -dir neighbours of site s (periodic wrap).  Site
// indices of one lattice fit in 31 bits (the device field access uses 32-bit
// byte offsets), so the offsets are 32-bit.
template <int D>
struct Neighbours {
    int32_t up[D];   // s + up[d]   = site + d_hat
    int32_t dn[D];   // s + dn[d]   = site - d_hat
};

template <int D>
__device__ __forceinline__ Neighbours<D> neighbours(const Geom& g, int64_t s)
{
    Neighbours<D> n;
    uint32_t r = (uint32_t)s;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const uint32_t L = (uint32_t)g.extent[d];
        const uint32_t x = (d == D - 1) ? r : r % L;
        r /= L;
        const int32_t st = (int32_t)g.stride[d];
        n.up[d] = (x == L - 1) ? -(int32_t)(L - 1) * st : st;
        n.dn[d] = (x == 0) ? (int32_t)(L - 1) * st : -st;
    }
    return n;
}

// Workgroups are dispatched round robin over the 8 XCDs (workgroup b runs on
// XCD b mod 8).  Give XCD k the contiguous logical range
// [k q + min(k, r), ...) of n = 8 q + r blocks.
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t n)
{
    const int64_t k = b & 7, i = b >> 3, q = n >> 3, r = n & 7;
    return k * q + (k < r ? k : r) + i;
}

// Site block of logical block j for the slab order: P blocks per slice of the
// slowest direction, G slices per XCD slab; within a slab the slice index runs
// fastest.
__device__ __forceinline__ int64_t slab_block(int64_t j, int64_t P, int64_t G)
{
    const int64_t q = j / (G * P), r = j % (G * P);
    return (q * G + r % G) * P + r / G;
}

template <int D>
__device__ __forceinline__ int site_parity(const Geom& g, int64_t s)
{
    uint32_t r = (uint32_t)s;
    uint32_t p = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const uint32_t L = (uint32_t)g.extent[d];
        const uint32_t x = (d == D - 1) ? r : r % L;
        r /= L;
        p += x;
    }
    return (int)(p & 1);
}

} // namespace lqcd
