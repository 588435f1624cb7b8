// k_observables.hip -- plaquette / Wilson action / E-field energy reductions.
//
// All reductions run in double arithmetic (also for f32-stored fields) and in
// a fixed order: a grid of at most kMaxReduceBlocks blocks, each thread
// walking a fixed stride, wave64 butterfly shuffles, then one block that sums
// the per-block partials in block order.  The result is bitwise reproducible
// run to run for a given lattice shape.
#include "dispatch.hpp"
#include "observables_device.hpp"

namespace lqcd {

// plaquette_site summed over the lattice
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_plaquette_partials(const vec2_t<S>* __restrict__ U, Geom g,
                                                               double* __restrict__ partials)
{
    double acc[3] = {0.0, 0.0, 0.0};
    const int64_t step = (int64_t)gridDim.x * kBlock;
    for (int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x; s < g.nsite; s += step)
        plaquette_site<D, S>(U, g, s, acc);
    block_reduce_to<3>(acc, partials + (int64_t)blockIdx.x * kReduceSlots);
}

// both terms of H in one pass over the lattice (hamiltonian_total,
// src/simulation/state.rs:229-231): the plaquette sums of plaquette_site and
// the E energy sum of the site's D links (acc[3]); the E loads stream while
// the plaquette's link loads wait
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_hamiltonian_partials(const vec2_t<S>* __restrict__ U,
                                                                 const S* __restrict__ E, Geom g,
                                                                 double* __restrict__ partials)
{
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const int64_t step = (int64_t)gridDim.x * kBlock;
    for (int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x; s < g.nsite; s += step) {
#pragma unroll
        for (int mu = 0; mu < D; ++mu) acc[3] += efield_energy_link<S>(E, g, mu, s);
        plaquette_site<D, S>(U, g, s, acc);
    }
    block_reduce_to<4>(acc, partials + (int64_t)blockIdx.x * kReduceSlots);
}

// sum over links of Su3Adjoint::trace_squared = sum_a e_a^2 / 2
// (src/field.rs:164-167, hamiltonian_efield src/simulation/state.rs:1370-1385)
template <typename S>
__global__ __launch_bounds__(kBlock) void k_efield_partials(const S* __restrict__ E, Geom g,
                                                            double* __restrict__ partials)
{
    double acc[1] = {0.0};
    const int64_t step = (int64_t)gridDim.x * kBlock;
    for (int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x; s < g.nsite; s += step)
        for (int mu = 0; mu < g.dim; ++mu) acc[0] += efield_energy_link<S>(E, g, mu, s);
    block_reduce_to<1>(acc, partials + (int64_t)blockIdx.x * kReduceSlots);
}

__global__ __launch_bounds__(kBlock) void k_mh_partials(const double* __restrict__ p,
                                                        const uint32_t* __restrict__ acc_flags,
                                                        int64_t n, double* __restrict__ partials)
{
    double acc[2] = {0.0, 0.0};
    const int64_t step = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += step) {
        acc[0] += p[i];
        acc[1] += (double)acc_flags[i];
    }
    block_reduce_to<2>(acc, partials + (int64_t)blockIdx.x * kReduceSlots);
}

template <int K>
__global__ __launch_bounds__(kBlock) void k_final_reduce(const double* __restrict__ partials,
                                                         int nblocks, double* __restrict__ out)
{
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kBlock)
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] += partials[(int64_t)b * kReduceSlots + k];
    block_reduce_to<K>(acc, out);
}

// max over lanes of a wave, then over the waves of the block (thread 0 gets it)
__device__ __forceinline__ double block_max(double v)
{
    v = wave_max(v);
    __shared__ double sh[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double m = sh[0];
#pragma unroll
    for (int w = 1; w < kBlock / 64; ++w) m = fmax(m, sh[w]);
    return m;
}

// largest su3_distance_link over the lattice: per-block maxima, then one
// block over them
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_su3_distance_partials(const vec2_t<S>* __restrict__ U, Geom g,
                                                                  double* __restrict__ partials)
{
    double dev = 0.0;
    const int64_t step = (int64_t)gridDim.x * kBlock;
    for (int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x; s < g.nsite; s += step)
#pragma unroll
        for (int mu = 0; mu < D; ++mu) dev = fmax(dev, su3_distance_link(load_link<double, S>(U, g.npad, mu, s)));
    const double m = block_max(dev);
    if (threadIdx.x == 0) partials[(int64_t)blockIdx.x * kReduceSlots] = m;
}

__global__ __launch_bounds__(kBlock) void k_final_max(const double* __restrict__ partials, int nblocks,
                                                       double* __restrict__ out)
{
    double m = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kBlock) m = fmax(m, partials[(int64_t)b * kReduceSlots]);
    m = block_max(m);
    if (threadIdx.x == 0) out[0] = m;
}

static int reduce_blocks(int64_t n)
{
    int64_t b = (n + kBlock - 1) / kBlock;
    if (b > kMaxReduceBlocks) b = kMaxReduceBlocks;
    if (b < 1) b = 1;
    return (int)b;
}

hipError_t launch_plaquette_sum(const lqcd_lattice* L, const void* U, double* d_out3)
{
    const int nb = reduce_blocks(L->g.nsite);
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_PLAQUETTE, &ev);
    LQCD_DISPATCH_DS(L, k_plaquette_partials<D, S><<<nb, kBlock, 0, L->ctx->stream>>>(
                            (const vec2_t<S>*)U, L->g, L->partials));
    prof_end(L->ctx, &ev);
    k_final_reduce<3><<<1, kBlock, 0, L->ctx->stream>>>(L->partials, nb, d_out3);
    return hipGetLastError();
}

hipError_t launch_hamiltonian_sum(const lqcd_lattice* L, const void* U, const void* E, double* d_out4)
{
    const int nb = reduce_blocks(L->g.nsite);
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_PLAQUETTE, &ev);
    LQCD_DISPATCH_DS(L, k_hamiltonian_partials<D, S><<<nb, kBlock, 0, L->ctx->stream>>>(
                            (const vec2_t<S>*)U, (const S*)E, L->g, L->partials));
    prof_end(L->ctx, &ev);
    k_final_reduce<4><<<1, kBlock, 0, L->ctx->stream>>>(L->partials, nb, d_out4);
    return hipGetLastError();
}

hipError_t launch_su3_distance(const lqcd_lattice* L, const void* U, double* d_out1)
{
    const int nb = reduce_blocks(L->g.nsite);
    LQCD_DISPATCH_DS(L, k_su3_distance_partials<D, S><<<nb, kBlock, 0, L->ctx->stream>>>(
                            (const vec2_t<S>*)U, L->g, L->partials));
    k_final_max<<<1, kBlock, 0, L->ctx->stream>>>(L->partials, nb, d_out1);
    return hipGetLastError();
}

hipError_t launch_efield_energy(const lqcd_lattice* L, const void* E, double* d_out1)
{
    const int nb = reduce_blocks(L->g.nsite);
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_EFIELD_ENERGY, &ev);
    LQCD_DISPATCH_S(L, k_efield_partials<S><<<nb, kBlock, 0, L->ctx->stream>>>((const S*)E, L->g,
                                                                                L->partials));
    prof_end(L->ctx, &ev);
    k_final_reduce<1><<<1, kBlock, 0, L->ctx->stream>>>(L->partials, nb, d_out1);
    return hipGetLastError();
}

hipError_t launch_mh_stats(const lqcd_lattice* L, double* d_out2)
{
    const int64_t n = (int64_t)L->nlink;
    const int nb = reduce_blocks(n);
    k_mh_partials<<<nb, kBlock, 0, L->ctx->stream>>>(L->mh_p, L->mh_acc, n, L->partials);
    k_final_reduce<2><<<1, kBlock, 0, L->ctx->stream>>>(L->partials, nb, d_out2);
    return hipGetLastError();
}

hipError_t launch_final_reduce_at(const lqcd_lattice* L, const double* partials, int nblocks, double* d_out2)
{
    k_final_reduce<2><<<1, kBlock, 0, L->ctx->stream>>>(partials, nblocks, d_out2);
    return hipGetLastError();
}

hipError_t launch_final_reduce1(const lqcd_lattice* L, int nblocks, double* d_out1)
{
    k_final_reduce<1><<<1, kBlock, 0, L->ctx->stream>>>(L->partials, nblocks, d_out1);
    return hipGetLastError();
}

} // namespace lqcd
