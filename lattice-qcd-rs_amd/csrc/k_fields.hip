// k_fields.hip -- layout conversion, cold start, device momenta, reunitarisation.
#include "dispatch.hpp"
#include "fields.hpp"
#include "rng_device.hpp"

namespace lqcd {

// AoS (the reference's Vec<Matrix3<Complex<f64>>>: link = site*D + mu,
// column-major 3x3, re/im interleaved) -> SoA row-major planes.
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_links_from_aos(const double* __restrict__ aos,
                                                           vec2_t<S>* __restrict__ U, Geom g)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= g.nsite) return;
    const double* p = aos + (s * D + mu) * 18;
    mat3<double> m;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) m.e[r * 3 + c] = {p[(c * 3 + r) * 2], p[(c * 3 + r) * 2 + 1]};
    store_link<double, S>(U, g.npad, mu, s, m);
}

template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_links_to_aos(const vec2_t<S>* __restrict__ U,
                                                         double* __restrict__ aos, Geom g)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= g.nsite) return;
    const mat3<double> m = load_link<double, S>(U, g.npad, mu, s);
    double* p = aos + (s * D + mu) * 18;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            p[(c * 3 + r) * 2] = m.e[r * 3 + c].re;
            p[(c * 3 + r) * 2 + 1] = m.e[r * 3 + c].im;
        }
}

// E: AoS [site][mu][8] f64 <-> SoA planes (mu*8 + a)
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_efield_from_aos(const double* __restrict__ aos,
                                                            S* __restrict__ E, Geom g)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= g.nsite) return;
    double e[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) e[a] = aos[(s * D + mu) * 8 + a];
    store_e<double, S>(E, g.npad, mu, s, e);
}

template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_efield_to_aos(const S* __restrict__ E,
                                                          double* __restrict__ aos, Geom g)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= g.nsite) return;
    double e[8];
    load_e<double, S>(E, g.npad, mu, s, e);
#pragma unroll
    for (int a = 0; a < 8; ++a) aos[(s * D + mu) * 8 + a] = e[a];
}

template <typename S>
__global__ __launch_bounds__(kBlock) void k_links_cold(vec2_t<S>* __restrict__ U, Geom g)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= g.nsite) return;
    mat3<double> m;
    set_identity(m);
    store_link<double, S>(U, g.npad, mu, s, m);
}

template <typename S>
__global__ __launch_bounds__(kBlock) void k_efield_zero(S* __restrict__ E, Geom g)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= g.nsite) return;
    double e[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    store_e<double, S>(E, g.npad, mu, s, e);
}

// Gaussian momenta, e^a_mu(x) ~ N(0, sigma): link l = site*D + mu, block j
// (0..3) uses counter {lo32(l), hi16(l) | j << 16, lo32(n), hi32(n)} and key =
// seed; each block is one Box-Muller pair.
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_device_momenta(S* __restrict__ E, Geom g, uint64_t seed,
                                                           uint64_t counter, double sigma)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= g.nsite) return;
    const uint64_t l = (uint64_t)s * D + mu;
    const uint2 key = make_uint2((uint32_t)seed, (uint32_t)(seed >> 32));
    double e[8];
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
        const uint4 ctr = make_uint4((uint32_t)l, (uint32_t)((l >> 32) & 0xFFFFu) | (j << 16),
                                     (uint32_t)counter, (uint32_t)(counter >> 32));
        const uint4 o = philox4x32(ctr, key);
        const double u = u53(o.x, o.y);
        const double w = u53(o.z, o.w);
        const double rad = sqrt(-2.0 * log(1.0 - u)) * sigma;
        double sn, cs;
        sincospi(2.0 * w, &sn, &cs);
        e[2 * j] = rad * cs;
        e[2 * j + 1] = rad * sn;
    }
    store_e<double, S>(E, g.npad, mu, s, e);
}

// Hot start on the device: random_su3 (src/su3.rs:322-347) -- two complex
// 3-vectors with re, im ~ U[-1, 1), redrawn while |v1| <= eps or
// |v1 . v2| <= eps (non-conjugated dot), then the column Gram-Schmidt of
// orthonormalize_matrix.  The 12 uniforms of attempt r come from the Philox
// blocks j = 16 + 6 r .. 21 + 6 r of link l (counter as for the momenta,
// counter word 2..3 = 0); at most 4 attempts.
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_links_random(vec2_t<S>* __restrict__ U, Geom g, uint64_t seed)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= g.nsite) return;
    const uint64_t l = (uint64_t)s * D + mu;
    const uint2 key = make_uint2((uint32_t)seed, (uint32_t)(seed >> 32));
    const double eps = 2.220446049250313e-16;
    mat3<double> m;
    for (int attempt = 0; attempt < 4; ++attempt) {
        double v[12];
#pragma unroll
        for (uint32_t j = 0; j < 6; ++j) {
            const uint32_t blk = 16u + 6u * (uint32_t)attempt + j;
            const uint4 ctr = make_uint4((uint32_t)l, (uint32_t)((l >> 32) & 0xFFFFu) | (blk << 16), 0u, 0u);
            const uint4 o = philox4x32(ctr, key);
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
    store_link<double, S>(U, g.npad, mu, s, m);
}

// LinkMatrix::normalize (src/field.rs:897-901) in double arithmetic
template <typename S>
__global__ __launch_bounds__(kBlock) void k_reunitarize(vec2_t<S>* __restrict__ U, Geom g)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= g.nsite) return;
    mat3<double> m = load_link<double, S>(U, g.npad, mu, s);
    orthonormalize(m);
    store_link<double, S>(U, g.npad, mu, s, m);
}

// Streaming copy of n 16-byte words: the reference point for achievable HBM
// bandwidth and for calibrating byte counters.
__global__ __launch_bounds__(kBlock) void k_stream_copy(const uint4* __restrict__ a, uint4* __restrict__ b,
                                                         int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

hipError_t launch_stream_copy(hipStream_t stream, const void* src, void* dst, size_t bytes)
{
    const int64_t n = (int64_t)(bytes / 16);
    k_stream_copy<<<8192, kBlock, 0, stream>>>((const uint4*)src, (uint4*)dst, n);
    return hipGetLastError();
}

static dim3 link_grid(const lqcd_lattice* L)
{
    return dim3(grid_for(L->g.nsite), (unsigned)L->g.dim);
}

hipError_t launch_links_from_aos(const lqcd_lattice* L, const double* d_aos, void* U)
{
    LQCD_DISPATCH_DS(L, k_links_from_aos<D, S><<<link_grid(L), kBlock, 0, L->ctx->stream>>>(
                            d_aos, (vec2_t<S>*)U, L->g));
    return hipGetLastError();
}

hipError_t launch_links_to_aos(const lqcd_lattice* L, const void* U, double* d_aos)
{
    LQCD_DISPATCH_DS(L, k_links_to_aos<D, S><<<link_grid(L), kBlock, 0, L->ctx->stream>>>(
                            (const vec2_t<S>*)U, d_aos, L->g));
    return hipGetLastError();
}

hipError_t launch_efield_from_aos(const lqcd_lattice* L, const double* d_aos, void* E)
{
    LQCD_DISPATCH_DS(L, k_efield_from_aos<D, S><<<link_grid(L), kBlock, 0, L->ctx->stream>>>(
                            d_aos, (S*)E, L->g));
    return hipGetLastError();
}

hipError_t launch_efield_to_aos(const lqcd_lattice* L, const void* E, double* d_aos)
{
    LQCD_DISPATCH_DS(L, k_efield_to_aos<D, S><<<link_grid(L), kBlock, 0, L->ctx->stream>>>(
                            (const S*)E, d_aos, L->g));
    return hipGetLastError();
}

hipError_t launch_links_cold(const lqcd_lattice* L, void* U)
{
    LQCD_DISPATCH_S(L, k_links_cold<S><<<link_grid(L), kBlock, 0, L->ctx->stream>>>(
                           (vec2_t<S>*)U, L->g));
    return hipGetLastError();
}

hipError_t launch_efield_zero(const lqcd_lattice* L, void* E)
{
    LQCD_DISPATCH_S(L, k_efield_zero<S><<<link_grid(L), kBlock, 0, L->ctx->stream>>>((S*)E, L->g));
    return hipGetLastError();
}

hipError_t launch_device_momenta(const lqcd_lattice* L, void* E, uint64_t seed, uint64_t counter,
                                 double sigma)
{
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_MOMENTA, &ev);
    LQCD_DISPATCH_DS(L, k_device_momenta<D, S><<<link_grid(L), kBlock, 0, L->ctx->stream>>>(
                            (S*)E, L->g, seed, counter, sigma));
    prof_end(L->ctx, &ev);
    return hipGetLastError();
}

hipError_t launch_links_random(const lqcd_lattice* L, void* U, uint64_t seed)
{
    LQCD_DISPATCH_DS(L, k_links_random<D, S><<<link_grid(L), kBlock, 0, L->ctx->stream>>>(
                            (vec2_t<S>*)U, L->g, seed));
    return hipGetLastError();
}

hipError_t launch_reunitarize(const lqcd_lattice* L, void* U)
{
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_REUNIT, &ev);
    LQCD_DISPATCH_S(L, k_reunitarize<S><<<link_grid(L), kBlock, 0, L->ctx->stream>>>(
                           (vec2_t<S>*)U, L->g));
    prof_end(L->ctx, &ev);
    return hipGetLastError();
}

} // namespace lqcd
