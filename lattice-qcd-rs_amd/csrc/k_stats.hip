// k_stats.hip -- sums behind the statistics entry points
// (src/statistics/mod.rs): mean, variance in its two-pass and one-pass forms,
// covariance.  Values are n doubles, or n complex numbers as (re, im) pairs,
// resident in device memory.  Each launch is a fixed grid of partial sums
// (four independent accumulators per thread, grid-stride), wave64 butterflies,
// then one block summing the partials in block order: results are the same
// run to run for a given n and buffer alignment.  Buffers aligned to 16 bytes
// are read 16 bytes per lane (two reals or one complex value), which keeps
// twice the bytes in flight per load; others a double at a time.
#include <hip/hip_runtime.h>

#include "dispatch.hpp"
#include "internal.hpp"
#include "observables_device.hpp"

namespace lqcd {

namespace {

constexpr int kStatsUnroll = 4;

struct cval {
    double re, im;
};

// element i; complex values are read as two doubles, so the caller's pointer
// needs only the alignment of a double
template <bool CPLX>
__device__ __forceinline__ cval value_at(const double* __restrict__ x, int64_t i)
{
    if constexpr (CPLX)
        return {x[2 * i], x[2 * i + 1]};
    else
        return {x[i], 0.0};
}

// num-complex product (a + bi)(c + di) = (ac - bd) + (ad + bc)i
__device__ __forceinline__ cval cmul(cval a, cval b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// NK sums per element, each one double (real) or a (re, im) pair (complex):
//   STATS_MOMENTS : x, x x
//   STATS_CENTERED: (x - m)(x - m), m = (center[0] + i center[1]) / n
//   STATS_COV     : x, y, x y
template <int OP>
struct StatsK;
template <>
struct StatsK<STATS_MOMENTS> {
    static constexpr int N = 2;
};
template <>
struct StatsK<STATS_CENTERED> {
    static constexpr int N = 1;
};
template <>
struct StatsK<STATS_COV> {
    static constexpr int N = 3;
};

// the sums of one element a (and b = y's element for STATS_COV) added to acc
template <bool CPLX, int OP>
__device__ __forceinline__ void add_terms(cval a, cval b, cval m, double* acc)
{
    cval t[StatsK<OP>::N];
    if constexpr (OP == STATS_MOMENTS) {
        t[0] = a;
        t[1] = cmul(a, a);
    } else if constexpr (OP == STATS_CENTERED) {
        const cval d = {a.re - m.re, a.im - m.im};
        t[0] = cmul(d, d);
    } else {
        t[0] = a;
        t[1] = b;
        t[2] = cmul(a, b);
    }
#pragma unroll
    for (int k = 0; k < StatsK<OP>::N; ++k) {
        acc[CPLX ? 2 * k : k] += t[k].re;
        if constexpr (CPLX) acc[2 * k + 1] += t[k].im;
    }
}

// unit u of the grid-stride walk: with VEC one 16-byte word (two reals, or
// one complex value), else one element
template <bool CPLX, int OP, bool VEC>
__device__ __forceinline__ void add_unit(const double* __restrict__ x, const double* __restrict__ y, int64_t u,
                                         cval m, double* acc)
{
    if constexpr (VEC) {
        const double2 p = reinterpret_cast<const double2*>(x)[u];
        double2 q = make_double2(0.0, 0.0);
        if constexpr (OP == STATS_COV) q = reinterpret_cast<const double2*>(y)[u];
        if constexpr (CPLX) {
            add_terms<CPLX, OP>({p.x, p.y}, {q.x, q.y}, m, acc);
        } else {
            add_terms<CPLX, OP>({p.x, 0.0}, {q.x, 0.0}, m, acc);
            add_terms<CPLX, OP>({p.y, 0.0}, {q.y, 0.0}, m, acc);
        }
    } else {
        cval b = {0.0, 0.0};
        if constexpr (OP == STATS_COV) b = value_at<CPLX>(y, u);
        add_terms<CPLX, OP>(value_at<CPLX>(x, u), b, m, acc);
    }
}

template <bool CPLX, int OP, bool VEC>
__global__ __launch_bounds__(kBlock) void k_stats_partials(const double* __restrict__ x,
                                                           const double* __restrict__ y, int64_t n,
                                                           const double* __restrict__ center,
                                                           double* __restrict__ partials)
{
    constexpr int NK = StatsK<OP>::N;
    constexpr int K = CPLX ? 2 * NK : NK;
    cval m = {0.0, 0.0};
    if constexpr (OP == STATS_CENTERED) {
        // mean = sum / len, as statistics::mean computes it
        m.re = center[0] / (double)n;
        if constexpr (CPLX) m.im = center[1] / (double)n;
    }
    double acc[kStatsUnroll][K];
#pragma unroll
    for (int u = 0; u < kStatsUnroll; ++u)
#pragma unroll
        for (int k = 0; k < K; ++k) acc[u][k] = 0.0;
    // units of the walk: pairs of reals when VEC reads real values two at a
    // time (an odd last element is thread 0's), else elements
    const int64_t units = (VEC && !CPLX) ? n / 2 : n;
    const int64_t step = (int64_t)gridDim.x * kBlock;
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    // kStatsUnroll independent loads in flight per thread
    for (; i + (kStatsUnroll - 1) * step < units; i += kStatsUnroll * step) {
#pragma unroll
        for (int u = 0; u < kStatsUnroll; ++u) add_unit<CPLX, OP, VEC>(x, y, i + u * step, m, acc[u]);
    }
    for (; i < units; i += step) add_unit<CPLX, OP, VEC>(x, y, i, m, acc[0]);
    if constexpr (VEC && !CPLX)
        if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) add_unit<CPLX, OP, false>(x, y, n - 1, m, acc[0]);
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = (acc[0][k] + acc[1][k]) + (acc[2][k] + acc[3][k]);
    block_reduce_to<K>(v, partials + (int64_t)blockIdx.x * K);
}

template <int K>
__global__ __launch_bounds__(kBlock) void k_stats_final(const double* __restrict__ partials, int nblocks,
                                                        double* __restrict__ out)
{
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kBlock)
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] += partials[(int64_t)b * K + k];
    block_reduce_to<K>(acc, out);
}

template <bool CPLX, int OP>
hipError_t launch_stats_t(lqcd_ctx* ctx, const double* x, const double* y, int64_t n, const double* center,
                          double* partials, int nblocks, double* out)
{
    constexpr int K = CPLX ? 2 * StatsK<OP>::N : StatsK<OP>::N;
    lqcd_prof_event ev;
    prof_begin(ctx, PROF_STATS, &ev);
    const bool vec = ((uintptr_t)x & 15) == 0 && (OP != STATS_COV || ((uintptr_t)y & 15) == 0);
    if (vec)
        k_stats_partials<CPLX, OP, true><<<nblocks, kBlock, 0, ctx->stream>>>(x, y, n, center, partials);
    else
        k_stats_partials<CPLX, OP, false><<<nblocks, kBlock, 0, ctx->stream>>>(x, y, n, center, partials);
    prof_end(ctx, &ev);
    k_stats_final<K><<<1, kBlock, 0, ctx->stream>>>(partials, nblocks, out);
    return hipGetLastError();
}

} // namespace

int stats_blocks(int64_t n)
{
    int64_t b = (n + (int64_t)kBlock * kStatsUnroll - 1) / ((int64_t)kBlock * kStatsUnroll);
    if (b > kMaxReduceBlocks) b = kMaxReduceBlocks;
    if (b < 1) b = 1;
    return (int)b;
}

hipError_t launch_stats(lqcd_ctx* ctx, int op, bool cplx, const double* x, const double* y, int64_t n,
                        const double* center, double* partials, int nblocks, double* out)
{
    switch (op) {
    case STATS_MOMENTS:
        return cplx ? launch_stats_t<true, STATS_MOMENTS>(ctx, x, y, n, center, partials, nblocks, out)
                    : launch_stats_t<false, STATS_MOMENTS>(ctx, x, y, n, center, partials, nblocks, out);
    case STATS_CENTERED:
        return cplx ? launch_stats_t<true, STATS_CENTERED>(ctx, x, y, n, center, partials, nblocks, out)
                    : launch_stats_t<false, STATS_CENTERED>(ctx, x, y, n, center, partials, nblocks, out);
    case STATS_COV:
        return cplx ? launch_stats_t<true, STATS_COV>(ctx, x, y, n, center, partials, nblocks, out)
                    : launch_stats_t<false, STATS_COV>(ctx, x, y, n, center, partials, nblocks, out);
    default:
        return hipErrorInvalidValue;
    }
}

} // namespace lqcd
