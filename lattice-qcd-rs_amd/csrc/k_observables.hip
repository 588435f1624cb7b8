// k_observables.hip -- plaquette / Wilson action / E-field energy reductions.
//
// All reductions run in double arithmetic (also for f32-stored fields) and in
// a fixed order: a grid of at most kMaxReduceBlocks blocks, each thread
// walking a fixed stride, wave64 butterfly shuffles, then one block that sums
// the per-block partials in block order.  The result is bitwise reproducible
// run to run for a given lattice shape.
#include "dispatch.hpp"
#include "fields.hpp"

namespace lqcd {

template <int K>
__device__ __forceinline__ void block_reduce_to(double (&v)[K], double* __restrict__ dst)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] += __shfl_xor(v[k], off, 64);
    __shared__ double sh[kBlock / 64][K];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) sh[wave][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double t = 0.0;
#pragma unroll
            for (int w = 0; w < kBlock / 64; ++w) t += sh[w][k];
            dst[k] = t;
        }
}

// Per site: sum_{i<j} Tr P_ij(x) and sum_{i<j} (1 - Re Tr P_ij(x) / 3), with
// P_ij(x) = U_i(x) U_j(x+i) U_i(x+j)^+ U_j(x)^+ (LinkMatrix::pij,
// src/field.rs:761-771), as Tr[(U_i(x) U_j(x+i)) (U_j(x) U_i(x+j))^+].
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_plaquette_partials(const vec2_t<S>* __restrict__ U, Geom g,
                                                               double* __restrict__ partials)
{
    double acc[3] = {0.0, 0.0, 0.0};
    const int64_t step = (int64_t)gridDim.x * kBlock;
    for (int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x; s < g.nsite; s += step) {
        const Neighbours<D> nb = neighbours<D>(g, s);
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const mat3<double> ui = load_link<double, S>(U, g.npad, i, s);
#pragma unroll
            for (int j = i + 1; j < D; ++j) {
                const mat3<double> uj = load_link<double, S>(U, g.npad, j, s);
                const mat3<double> uj_pi = load_link<double, S>(U, g.npad, j, s + nb.up[i]);
                const mat3<double> ui_pj = load_link<double, S>(U, g.npad, i, s + nb.up[j]);
                mat3<double> a, b;
                mul<false, false>(ui, uj_pi, a);
                mul<false, false>(uj, ui_pj, b);
                const cplx<double> tr = trace_mul<false, true>(a, b);
                acc[0] += tr.re;
                acc[1] += tr.im;
                acc[2] += 1.0 - tr.re / 3.0;
            }
        }
    }
    block_reduce_to<3>(acc, partials + (int64_t)blockIdx.x * kReduceSlots);
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
        for (int mu = 0; mu < g.dim; ++mu) {
            double e[8];
            load_e<double, S>(E, g.npad, mu, s, e);
            double t = 0.0;
#pragma unroll
            for (int a = 0; a < 8; ++a) t += e[a] * e[a];
            acc[0] += t / 2.0;
        }
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

} // namespace lqcd
