// k_batch.hip -- kernels of the batched chains: one launch covers every
// chain.  Each wavefront takes one tile (64 sites of one chain, or those
// sites' links of one direction), looks the chain up and runs the same
// per-site / per-link arithmetic as the single-lattice kernels
// (md_device.hpp, observables_device.hpp, generators.hpp), so a chain of the
// batch evolves exactly as a lattice of its own would.
//
// Each chain is an independent HybridMonteCarloDiagnostic
// (src/simulation/monte_carlo/hybrid_monte_carlo.rs:465-471, 584-613): momenta
// refresh, H, leapfrog, H, accept with probability clamp(exp(H_old - H_new)),
// and on reject the old links stay.  The accept decision is taken on the
// device so that a whole trajectory needs no host round trip.
#include "../../include/lqcd.h"
#include "batch.hpp"
#include "dispatch.hpp"
#include "generators.hpp"
#include "md_device.hpp"
#include "observables_device.hpp"

namespace lqcd {

// Philox block of the per-chain accept word (link 0, counter = trajectory)
constexpr uint32_t kAcceptBlock = 62;

// MD over all chains.  Work units are tiles (SITE) or (tile, mu) pairs, mu
// fastest; workgroup wg (XCD-remapped) owns the contiguous unit range
// [wg * chunk, (wg + 1) * chunk) and its waves take every W-th unit of it.
template <int D, typename S, int DR, bool SITE, bool R12>
__global__ __launch_bounds__(kBlock) void k_batch_md(const vec2_t<S>* __restrict__ U_in,
                                                     vec2_t<S>* __restrict__ U_out, S* __restrict__ E,
                                                     const ChainDesc* __restrict__ cd,
                                                     const int32_t* __restrict__ tile_chain,
                                                     int64_t nunit, int64_t chunk, double kick, double dt)
{
    const int W = (int)(blockDim.x >> 6);
    // the same on every lane: readfirstlane keeps the chain lookup scalar
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & 63);
    const int64_t wg = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t begin = wg * chunk;
    const int64_t end = begin + chunk < nunit ? begin + chunk : nunit;
    for (int64_t u = begin + wave; u < end; u += W) {
        const int64_t tile = SITE ? u : u / D;
        const int mu = SITE ? 0 : (int)(u % D);
        const ChainDesc& d = cd[tile_chain[tile]];
        const int64_t s = (tile - d.tile0) * kTile + lane;
        if (s >= d.g.nsite) continue;
        const S fs = (S)md_force_scale(d.size);
        const S df = (S)md_drift_scale(dt, d.size);
        const vec2_t<S>* ui = U_in + d.u_off;
        vec2_t<S>* uo = DR >= 0 ? U_out + d.u_off : nullptr;
        S* e = E + d.e_off;
        // two-row kernels store the drifted links as rows 0, 1 (the
        // trajectory rebuilds row 2 of its final links, k_batch_links_map<2>)
        constexpr bool R01 = R12 && DR == DRIFT_EXP;
        if (SITE)
            md_site<D, S, S, DR, R12, false, R01>(ui, uo, e, d.g, s, fs, (S)kick, df);
        else
            md_link<D, S, S, DR, R12, R01>(ui, uo, e, d.g, mu, s, fs, (S)kick, df);
    }
}

// one wave per tile: wave w of block b takes tile 4 b + w
__device__ __forceinline__ int64_t tile_of_wave()
{
    return (int64_t)blockIdx.x * (kBlock / 64) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_batch_momenta(S* __restrict__ E, const ChainDesc* __restrict__ cd,
                                                          const int32_t* __restrict__ tile_chain,
                                                          int64_t ntile, const uint64_t* __restrict__ counter)
{
    const int64_t tile = tile_of_wave();
    if (tile >= ntile) return;
    const ChainDesc& d = cd[tile_chain[tile]];
    const int64_t s = (tile - d.tile0) * kTile + (threadIdx.x & 63);
    if (s >= d.g.nsite) return;
    const uint64_t n = *counter;
#pragma unroll
    for (int mu = 0; mu < D; ++mu) {
        double e[8];
        momenta_link((uint64_t)s * D + mu, d.seed, n, d.sigma, e);
        store_e<double, S>(E + d.e_off, d.g.npad, mu, s, e);
    }
}

template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_batch_hot_start(vec2_t<S>* __restrict__ U,
                                                            const ChainDesc* __restrict__ cd,
                                                            const int32_t* __restrict__ tile_chain, int64_t ntile)
{
    const int64_t tile = tile_of_wave();
    if (tile >= ntile) return;
    const ChainDesc& d = cd[tile_chain[tile]];
    const int64_t s = (tile - d.tile0) * kTile + (threadIdx.x & 63);
    if (s >= d.g.nsite) return;
    for (int mu = 0; mu < D; ++mu) {
        mat3<double> m;
        random_su3_link((uint64_t)s * D + mu, d.seed, m);
        store_link<double, S>(U + d.u_off, d.g.npad, mu, s, m);
    }
}

// MODE 0: identity links; 1: reunitarise in place (LinkMatrix::normalize,
// src/field.rs:897-901, in double arithmetic); 2: rebuild row 2 of SU(3)
// links from rows 0, 1 in the storage precision (after two-row MD kernels)
template <int D, typename S, int MODE>
__global__ __launch_bounds__(kBlock) void k_batch_links_map(vec2_t<S>* __restrict__ U,
                                                            const ChainDesc* __restrict__ cd,
                                                            const int32_t* __restrict__ tile_chain, int64_t ntile)
{
    const int64_t tile = tile_of_wave();
    if (tile >= ntile) return;
    const ChainDesc& d = cd[tile_chain[tile]];
    const int64_t s = (tile - d.tile0) * kTile + (threadIdx.x & 63);
    if (s >= d.g.nsite) return;
#pragma unroll
    for (int mu = 0; mu < D; ++mu) {
        if constexpr (MODE == 2) {
            store_link_row2<S, S>(U + d.u_off, d.g.npad, mu, s, load_link_r12<S, S>(U + d.u_off, d.g.npad, mu, s));
            continue;
        }
        mat3<double> m;
        if (MODE == 0) {
            set_identity(m);
        } else {
            m = load_link<double, S>(U + d.u_off, d.g.npad, mu, s);
            orthonormalize(m);
        }
        store_link<double, S>(U + d.u_off, d.g.npad, mu, s, m);
    }
}

// per-tile sums [Re Tr P, Im Tr P, 1 - Re Tr P / 3, e^2 / 2] into partials[tile][4]
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_batch_tile_sums(const vec2_t<S>* __restrict__ U,
                                                            const S* __restrict__ E,
                                                            const ChainDesc* __restrict__ cd,
                                                            const int32_t* __restrict__ tile_chain,
                                                            int64_t ntile, double* __restrict__ partials)
{
    const int64_t tile = tile_of_wave();
    if (tile >= ntile) return;
    const ChainDesc& d = cd[tile_chain[tile]];
    const int64_t s = (tile - d.tile0) * kTile + (threadIdx.x & 63);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (s < d.g.nsite) {
        plaquette_site<D, S>(U + d.u_off, d.g, s, acc);
        if (E)
#pragma unroll
            for (int mu = 0; mu < D; ++mu) acc[3] += efield_energy_link<S>(E + d.e_off, d.g, mu, s);
    }
    wave_sum<4>(acc);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k) partials[tile * 4 + k] = acc[k];
}

// largest su3_distance_link of each tile's links, then of all tiles
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_batch_su3_distance(const vec2_t<S>* __restrict__ U,
                                                               const ChainDesc* __restrict__ cd,
                                                               const int32_t* __restrict__ tile_chain,
                                                               int64_t ntile, double* __restrict__ partials)
{
    const int64_t tile = tile_of_wave();
    if (tile >= ntile) return;
    const ChainDesc& d = cd[tile_chain[tile]];
    const int64_t s = (tile - d.tile0) * kTile + (threadIdx.x & 63);
    double dev = 0.0;
    if (s < d.g.nsite)
#pragma unroll
        for (int mu = 0; mu < D; ++mu)
            dev = fmax(dev, su3_distance_link(load_link<double, S>(U + d.u_off, d.g.npad, mu, s)));
    dev = wave_max(dev);
    if ((threadIdx.x & 63) == 0) partials[tile * 4] = dev;
}

__global__ __launch_bounds__(kBlock) void k_batch_max_tiles(const double* __restrict__ partials, int64_t ntile,
                                                            double* __restrict__ out)
{
    double m = 0.0;
    for (int64_t t = threadIdx.x; t < ntile; t += kBlock) m = fmax(m, partials[t * 4]);
    m = wave_max(m);
    __shared__ double sh[kBlock / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = sh[0];
        for (int w = 1; w < kBlock / 64; ++w) r = fmax(r, sh[w]);
        out[0] = r;
    }
}

// block c sums the tiles of chain c in a fixed order
__global__ __launch_bounds__(kBlock) void k_batch_chain_sums(const double* __restrict__ partials,
                                                             const ChainDesc* __restrict__ cd,
                                                             double* __restrict__ out)
{
    const ChainDesc& d = cd[blockIdx.x];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t t = threadIdx.x; t < d.ntile; t += kBlock)
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[k] += partials[(d.tile0 + t) * 4 + k];
    block_reduce_to<4>(acc, out + (int64_t)blockIdx.x * 4);
}

// one thread per chain: H = beta sum(1 - Re Tr P / 3) + beta sum e^2 / 2
// (hamiltonian_links + hamiltonian_efield, src/simulation/state.rs:821-849,
// 1370-1385), p = clamp(exp(H_old - H_new), 0, 1) with NaN -> 1
// (hybrid_monte_carlo.rs:584-589), accepted iff p == 1 or
// word < floor(p 2^64) (rand 0.8 Bernoulli), word from Philox block 62.
__global__ void k_batch_accept(const ChainDesc* __restrict__ cd, int64_t nchain, int dim,
                               const double* __restrict__ h_old, const double* __restrict__ h_new,
                               const uint64_t* __restrict__ counter, uint32_t* __restrict__ accept,
                               lqcd_chain_stats* __restrict__ stats)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchain) return;
    const ChainDesc& d = cd[c];
    const double* o = h_old + c * 4;
    const double* w = h_new + c * 4;
    const double ho = __dadd_rn(__dmul_rn(d.beta, o[2]), __dmul_rn(d.beta, o[3]));
    const double hn = __dadd_rn(__dmul_rn(d.beta, w[2]), __dmul_rn(d.beta, w[3]));
    const double p = fmax(fmin(exp(ho - hn), 1.0), 0.0);
    bool acc = true;
    if (!(p >= 1.0)) {
        const uint4 r = philox_link(0, kAcceptBlock, *counter, d.seed);
        const uint64_t word = ((uint64_t)r.x << 32) | (uint64_t)r.y;
        acc = word < (uint64_t)(p * 18446744073709551616.0);
    }
    accept[c] = acc ? 1u : 0u;
    const double nplaq = (double)d.g.nsite * (double)(dim * (dim - 1) / 2);
    lqcd_chain_stats st;
    st.h_old = ho;
    st.h_new = hn;
    st.delta_h = hn - ho;
    st.probability = p;
    st.plaquette = (acc ? w[0] : o[0]) / nplaq / 3.0;
    st.accepted = acc ? 1 : 0;
    st.reserved = 0;
    stats[c] = st;
}

// accepted chains take the links of U_final
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_batch_accept_copy(const vec2_t<S>* __restrict__ U_final,
                                                              vec2_t<S>* __restrict__ U,
                                                              const ChainDesc* __restrict__ cd,
                                                              const int32_t* __restrict__ tile_chain,
                                                              const uint32_t* __restrict__ accept, int64_t ntile)
{
    const int64_t tile = tile_of_wave();
    if (tile >= ntile) return;
    const int32_t c = tile_chain[tile];
    if (!accept[c]) return;
    const ChainDesc& d = cd[c];
    const int64_t s = (tile - d.tile0) * kTile + (threadIdx.x & 63);
    if (s >= d.g.nsite) return;
    const vec2_t<S>* src = U_final + d.u_off + s;
    vec2_t<S>* dst = U + d.u_off + s;
#pragma unroll
    for (int k = 0; k < D * 9; ++k) dst[(int64_t)k * d.g.npad] = src[(int64_t)k * d.g.npad];
}

__global__ void k_batch_counter_advance(uint64_t* __restrict__ counter)
{
    if (threadIdx.x == 0) counter[0] += 1;
}

// ------------------------------------------------------------------ launchers

#define BATCH_DISPATCH(b, ...)                                                      \
    do {                                                                            \
        switch ((b)->dim) {                                                         \
        case 2: {                                                                   \
            constexpr int D = 2;                                                    \
            if ((b)->precision == 1) { using S = float; __VA_ARGS__; }              \
            else { using S = double; __VA_ARGS__; }                                 \
        } break;                                                                    \
        case 3: {                                                                   \
            constexpr int D = 3;                                                    \
            if ((b)->precision == 1) { using S = float; __VA_ARGS__; }              \
            else { using S = double; __VA_ARGS__; }                                 \
        } break;                                                                    \
        case 4: {                                                                   \
            constexpr int D = 4;                                                    \
            if ((b)->precision == 1) { using S = float; __VA_ARGS__; }              \
            else { using S = double; __VA_ARGS__; }                                 \
        } break;                                                                    \
        default:                                                                    \
            return hipErrorInvalidValue;                                            \
        }                                                                           \
    } while (0)

static unsigned tile_grid(const lqcd_batch* b) { return grid_for(b->ntile, kBlock / 64); }

// workgroups of kernel k resident on the whole device at once (block threads each)
template <typename K>
static int64_t resident_blocks(K kernel, int block)
{
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess) return 0;
    return (int64_t)cus * per_cu;
}

// launch == false only settles the grid (the occupancy query), so that a
// later launch under stream capture makes no device query
template <int D, typename S, int DR, bool SITE, bool R12>
static hipError_t batch_md(const lqcd_batch* b, int block, bool persistent, const void* U_in, void* U_out,
                           double kick, double dt, bool launch)
{
    const int W = block / 64;
    const int64_t nunit = SITE ? b->ntile : b->ntile * D;
    int64_t groups = (nunit + W - 1) / W;
    if (persistent) {
        static int64_t resident[5] = {-1, -1, -1, -1, -1};  // by block / 64
        int64_t& res = resident[W];
        if (res < 0) res = resident_blocks(k_batch_md<D, S, DR, SITE, R12>, block);
        if (res > 0 && res < groups) groups = res;
    }
    if (!launch) return hipSuccess;
    const int64_t chunk = (nunit + groups - 1) / groups;
    k_batch_md<D, S, DR, SITE, R12><<<(unsigned)groups, block, 0, b->ctx->stream>>>(
        (const vec2_t<S>*)U_in, (vec2_t<S>*)U_out, (S*)b->E, b->d_chains, b->d_tile_chain, nunit, chunk,
        kick, dt);
    return hipGetLastError();
}

template <int D, typename S, int DR>
static hipError_t batch_md_shape(const lqcd_batch* b, int config, const void* U_in, void* U_out, double kick,
                                 double dt, bool launch)
{
    const BatchShape sh = batch_shape(config);
    // links read as two rows: exp-drift integrations of SU(3) links
    if constexpr (DR != DRIFT_ADDITIVE) {
        if (b->md_r12) {
            if (sh.site)
                return batch_md<D, S, DR, true, true>(b, sh.block, sh.persistent, U_in, U_out, kick, dt, launch);
            return batch_md<D, S, DR, false, true>(b, sh.block, sh.persistent, U_in, U_out, kick, dt, launch);
        }
    }
    if (sh.site) return batch_md<D, S, DR, true, false>(b, sh.block, sh.persistent, U_in, U_out, kick, dt, launch);
    return batch_md<D, S, DR, false, false>(b, sh.block, sh.persistent, U_in, U_out, kick, dt, launch);
}

static hipError_t batch_md_op(const lqcd_batch* b, int config, const void* U_in, void* U_out, double kick,
                              double dt, int op, bool launch)
{
    if (config < 0 || config >= kBatchConfigs) return hipErrorInvalidValue;
    if (op == BATCH_KICK) {
        BATCH_DISPATCH(b, return (batch_md_shape<D, S, -1>(b, config, U_in, U_out, kick, dt, launch)));
    } else if (op == DRIFT_EXP) {
        BATCH_DISPATCH(b, return (batch_md_shape<D, S, DRIFT_EXP>(b, config, U_in, U_out, kick, dt, launch)));
    } else {
        BATCH_DISPATCH(b, return (batch_md_shape<D, S, DRIFT_ADDITIVE>(b, config, U_in, U_out, kick, dt, launch)));
    }
    return hipErrorInvalidValue;
}

hipError_t launch_batch_md(const lqcd_batch* b, int config, const void* U_in, void* U_out, double kick,
                           double dt, int op)
{
    lqcd_prof_event ev;
    prof_begin(b->ctx, op == BATCH_KICK ? PROF_KICK : PROF_KICK_DRIFT, &ev);
    const hipError_t e = batch_md_op(b, config, U_in, U_out, kick, dt, op, true);
    prof_end(b->ctx, &ev);
    return e;
}

hipError_t prepare_batch_md(const lqcd_batch* b, int config, int op)
{
    return batch_md_op(b, config, nullptr, nullptr, 0.0, 0.0, op, false);
}

hipError_t launch_batch_momenta(const lqcd_batch* b)
{
    lqcd_prof_event ev;
    prof_begin(b->ctx, PROF_MOMENTA, &ev);
    BATCH_DISPATCH(b, k_batch_momenta<D, S><<<tile_grid(b), kBlock, 0, b->ctx->stream>>>(
                          (S*)b->E, b->d_chains, b->d_tile_chain, b->ntile, b->d_counter));
    prof_end(b->ctx, &ev);
    return hipGetLastError();
}

hipError_t launch_batch_hot_start(const lqcd_batch* b)
{
    BATCH_DISPATCH(b, k_batch_hot_start<D, S><<<tile_grid(b), kBlock, 0, b->ctx->stream>>>(
                          (vec2_t<S>*)b->U[0], b->d_chains, b->d_tile_chain, b->ntile));
    return hipGetLastError();
}

hipError_t launch_batch_cold(const lqcd_batch* b)
{
    BATCH_DISPATCH(b, k_batch_links_map<D, S, 0><<<tile_grid(b), kBlock, 0, b->ctx->stream>>>(
                          (vec2_t<S>*)b->U[0], b->d_chains, b->d_tile_chain, b->ntile));
    return hipGetLastError();
}

hipError_t launch_batch_reunitarize(const lqcd_batch* b)
{
    lqcd_prof_event ev;
    prof_begin(b->ctx, PROF_REUNIT, &ev);
    BATCH_DISPATCH(b, k_batch_links_map<D, S, 1><<<tile_grid(b), kBlock, 0, b->ctx->stream>>>(
                          (vec2_t<S>*)b->U[0], b->d_chains, b->d_tile_chain, b->ntile));
    prof_end(b->ctx, &ev);
    return hipGetLastError();
}

hipError_t launch_batch_complete_row2(const lqcd_batch* b, void* U)
{
    BATCH_DISPATCH(b, k_batch_links_map<D, S, 2><<<tile_grid(b), kBlock, 0, b->ctx->stream>>>(
                          (vec2_t<S>*)U, b->d_chains, b->d_tile_chain, b->ntile));
    return hipGetLastError();
}

hipError_t launch_batch_sums(const lqcd_batch* b, const void* U, const void* E, double* d_out)
{
    lqcd_prof_event ev;
    prof_begin(b->ctx, PROF_PLAQUETTE, &ev);
    BATCH_DISPATCH(b, k_batch_tile_sums<D, S><<<tile_grid(b), kBlock, 0, b->ctx->stream>>>(
                          (const vec2_t<S>*)U, (const S*)E, b->d_chains, b->d_tile_chain, b->ntile,
                          b->d_partials));
    prof_end(b->ctx, &ev);
    k_batch_chain_sums<<<(unsigned)b->chains.size(), kBlock, 0, b->ctx->stream>>>(b->d_partials, b->d_chains,
                                                                                 d_out);
    return hipGetLastError();
}

hipError_t launch_batch_su3_distance(const lqcd_batch* b, const void* U, double* d_out1)
{
    BATCH_DISPATCH(b, k_batch_su3_distance<D, S><<<tile_grid(b), kBlock, 0, b->ctx->stream>>>(
                          (const vec2_t<S>*)U, b->d_chains, b->d_tile_chain, b->ntile, b->d_partials));
    k_batch_max_tiles<<<1, kBlock, 0, b->ctx->stream>>>(b->d_partials, b->ntile, d_out1);
    return hipGetLastError();
}

hipError_t launch_batch_accept(const lqcd_batch* b)
{
    const int64_t n = (int64_t)b->chains.size();
    k_batch_accept<<<grid_for(n, 64), 64, 0, b->ctx->stream>>>(b->d_chains, n, b->dim, b->d_h_old, b->d_h_new,
                                                                 b->d_counter, b->d_accept, b->d_stats);
    return hipGetLastError();
}

hipError_t launch_batch_accept_copy(const lqcd_batch* b, const void* U_final)
{
    BATCH_DISPATCH(b, k_batch_accept_copy<D, S><<<tile_grid(b), kBlock, 0, b->ctx->stream>>>(
                          (const vec2_t<S>*)U_final, (vec2_t<S>*)b->U[0], b->d_chains, b->d_tile_chain,
                          b->d_accept, b->ntile));
    return hipGetLastError();
}

hipError_t launch_batch_counter_advance(const lqcd_batch* b)
{
    k_batch_counter_advance<<<1, 64, 0, b->ctx->stream>>>(b->d_counter);
    return hipGetLastError();
}

} // namespace lqcd
