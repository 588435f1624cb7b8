// k_fields.hip -- layout conversion, cold start, device momenta, reunitarisation.
#include "dispatch.hpp"
#include "fields.hpp"
#include "generators.hpp"

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

// Gaussian momenta e^a_mu(x) ~ N(0, sigma) (momenta_link), link l = site*D + mu.
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_device_momenta(S* __restrict__ E, Geom g, uint64_t seed,
                                                           uint64_t counter, double sigma)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= g.nsite) return;
    double e[8];
    momenta_link((uint64_t)s * D + mu, seed, counter, sigma, e);
    store_e<double, S>(E, g.npad, mu, s, e);
}

// Hot start on the device (random_su3_link), link l = site * D + mu.
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_links_random(vec2_t<S>* __restrict__ U, Geom g, uint64_t seed)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= g.nsite) return;
    mat3<double> m;
    random_su3_link((uint64_t)s * D + mu, seed, m);
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

// Row 2 of every link rebuilt from rows 0 and 1 (conj(row0 x row1), in the
// storage precision, the arithmetic of load_link_r12): after an integration
// whose kernels stored only rows 0, 1 of the SU(3) links they drifted.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_complete_row2(vec2_t<S>* __restrict__ U, Geom g)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= g.nsite) return;
    store_link_row2<S, S>(U, g.npad, mu, s, load_link_r12<S, S>(U, g.npad, mu, s));
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

// Streaming read of n 16-byte words, four loads in flight per thread: the
// reference point for kernels that only read.  Each thread folds its words
// with xor and stores the fold (into its own slot of sink, grid x kBlock
// words) only when it equals `never`, so the loads stay and nothing is
// written in practice.
__global__ __launch_bounds__(kBlock) void k_stream_read(const uint4* __restrict__ a, int64_t n, uint32_t never,
                                                         uint32_t* __restrict__ sink)
{
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t acc = 0;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const uint4 v0 = a[i], v1 = a[i + stride], v2 = a[i + 2 * stride], v3 = a[i + 3 * stride];
        acc ^= (v0.x ^ v0.y ^ v0.z ^ v0.w) ^ (v1.x ^ v1.y ^ v1.z ^ v1.w) ^ (v2.x ^ v2.y ^ v2.z ^ v2.w) ^
               (v3.x ^ v3.y ^ v3.z ^ v3.w);
    }
    for (; i < n; i += stride) {
        const uint4 v = a[i];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    if (acc == never) sink[(int64_t)blockIdx.x * kBlock + threadIdx.x] = acc;
}

hipError_t launch_stream_read(hipStream_t stream, const void* src, size_t bytes, uint32_t* sink)
{
    const int64_t n = (int64_t)(bytes / 16);
    k_stream_read<<<kStreamReadBlocks, kBlock, 0, stream>>>((const uint4*)src, n, 0x9e3779b9u, sink);
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

hipError_t launch_complete_row2(const lqcd_lattice* L, void* U)
{
    LQCD_DISPATCH_S(L, k_complete_row2<S><<<link_grid(L), kBlock, 0, L->ctx->stream>>>((vec2_t<S>*)U, L->g));
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
