// generators.hpp -- per-link random draws on the device (momenta, hot start),
// shared by the single-lattice kernels and the batched chains so that a chain
// and a lattice with the same seed draw the same numbers.
#pragma once
#include "rng_device.hpp"
#include "su3_device.hpp"

namespace lqcd {

// Gaussian momenta, e^a(l) ~ N(0, sigma) for link l: block j (0..3) of the
// Philox stream (counter n, key seed) is one Box-Muller pair.
__device__ __forceinline__ void momenta_link(uint64_t l, uint64_t seed, uint64_t n, double sigma,
                                             double (&e)[8])
{
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
        const uint4 o = philox_link(l, j, n, seed);
        const double u = u53(o.x, o.y);
        const double w = u53(o.z, o.w);
        const double rad = sqrt(-2.0 * log(1.0 - u)) * sigma;
        double sn, cs;
        sincospi(2.0 * w, &sn, &cs);
        e[2 * j] = rad * cs;
        e[2 * j + 1] = rad * sn;
    }
}

// Hot start: random_su3 (src/su3.rs:322-347) -- two complex 3-vectors with
// re, im ~ U[-1, 1), redrawn while |v1| <= eps or |v1 . v2| <= eps
// (non-conjugated dot), then the column Gram-Schmidt of orthonormalize_matrix.
// The 12 uniforms of attempt r come from Philox blocks 16 + 6 r .. 21 + 6 r of
// link l with counter 0; at most 4 attempts.
__device__ __forceinline__ void random_su3_link(uint64_t l, uint64_t seed, mat3<double>& m)
{
    const double eps = 2.220446049250313e-16;
    for (int attempt = 0; attempt < 4; ++attempt) {
        double v[12];
#pragma unroll
        for (uint32_t j = 0; j < 6; ++j) {
            const uint4 o = philox_link(l, 16u + 6u * (uint32_t)attempt + j, 0, seed);
            v[2 * j] = 2.0 * u53(o.x, o.y) - 1.0;
            v[2 * j + 1] = 2.0 * u53(o.z, o.w) - 1.0;
        }
        // v1 = (v0 + i v1, v2 + i v3, v4 + i v5), v2 = (v6 + i v7, ...)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            m.e[r * 3 + 0] = {v[2 * r], v[2 * r + 1]};
            m.e[r * 3 + 1] = {v[6 + 2 * r], v[6 + 2 * r + 1]};
            m.e[r * 3 + 2] = {0.0, 0.0};
        }
        double n2 = 0.0, dre = 0.0, dim_ = 0.0;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const cplx<double> a = m.e[r * 3 + 0], b = m.e[r * 3 + 1];
            n2 += a.re * a.re + a.im * a.im;
            dre += a.re * b.re - a.im * b.im;
            dim_ += a.re * b.im + a.im * b.re;
        }
        if (sqrt(n2) > eps && sqrt(dre * dre + dim_ * dim_) > eps) break;
    }
    orthonormalize(m);
}

} // namespace lqcd
