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

struct Geom {
    int dim;
    int extent[kMaxDim];
    int64_t stride[kMaxDim];
    int64_t nsite;   // number of sites
    int64_t npad;    // row length of every SoA plane (>= nsite, multiple of 256)
};

// Offsets to the +dir and -dir neighbours of site s (periodic wrap).
template <int D>
struct Neighbours {
    int64_t up[D];   // s + up[d]   = site + d_hat
    int64_t dn[D];   // s + dn[d]   = site - d_hat
};

template <int D>
__device__ __forceinline__ Neighbours<D> neighbours(const Geom& g, int64_t s)
{
    Neighbours<D> n;
    int64_t r = s;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int L = g.extent[d];
        const int x = (d == D - 1) ? (int)r : (int)(r % L);
        r /= L;
        const int64_t st = g.stride[d];
        n.up[d] = (x == L - 1) ? -(int64_t)(L - 1) * st : st;
        n.dn[d] = (x == 0) ? (int64_t)(L - 1) * st : -st;
    }
    return n;
}

template <int D>
__device__ __forceinline__ int site_parity(const Geom& g, int64_t s)
{
    int64_t r = s;
    int p = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int L = g.extent[d];
        const int x = (d == D - 1) ? (int)r : (int)(r % L);
        r /= L;
        p += x;
    }
    return p & 1;
}

} // namespace lqcd
