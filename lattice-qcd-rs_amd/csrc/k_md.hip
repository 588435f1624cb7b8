// k_md.hip -- gauge force, momentum kick and link drift.
//
// The force is derivative_e of src/simulation/state.rs:1420-1448,
//   de^a_mu(x) = -sqrt(2/3)/a * Im Tr(T^a U_mu(x) Sigma_mu(x)),
//   Sigma_mu(x) = sum_{nu != mu} [ U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
//                               + U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu) ],
// which is the sum of S_{mu,+-nu}(x)^+ over the signed directions of the
// reference (LinkMatrix::sij, src/field.rs:743-758).  The eight generator
// projections are taken in closed form from the nine entries of U Sigma.
//
// The kick+drift kernels read U_in and write U_out (double buffered): one
// launch performs e += kick * F(U_in) and then the drift of its own link with
// the updated e -- one pass over HBM per leapfrog step.
//
// Work decompositions (chosen per lattice by the autotuner in capi.cpp):
//   VAR_LINK      one thread per link, grid (site blocks, D): the plain layout.
//   VAR_LINK_XCD  one thread per link, 1-D grid where the D links of a site
//                 block are consecutive blocks and consecutive blocks are
//                 remapped onto the same XCD, so neighbouring site blocks share
//                 that XCD's L2 instead of being spread over all eight.
//   VAR_SITE_XCD  one thread per site for all D links: the D staple sums are
//                 built together, which shares the loads (11 matrices per
//                 link instead of 19) and the product
//                 W = U_nu(x+mu) U_mu(x+nu)^+ between the upper staples of
//                 (mu, nu) and (nu, mu); same XCD remap.
//   VAR_SITE_SLAB per site as VAR_SITE_XCD, but each XCD gets a slab of the
//                 slowest direction and walks its site blocks with that
//                 direction fastest, so the blocks holding the +-1 neighbours
//                 in the slowest direction run next to each other in time and
//                 share the XCD's L2 (needs 256 | sites per slice and 8 | L_slow;
//                 otherwise it falls back to VAR_SITE_XCD's order).
#include "dispatch.hpp"
#include "fields.hpp"

namespace lqcd {

// Workgroups are dispatched round robin over the 8 XCDs (workgroup b runs on
// XCD b mod 8).  Give XCD k the contiguous logical range
// [k q + min(k, r), ...) of n = 8 q + r blocks.
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t n)
{
    const int64_t k = b & 7, i = b >> 3, q = n >> 3, r = n & 7;
    return k * q + (k < r ? k : r) + i;
}

// Site block of logical block j for VAR_SITE_SLAB: P blocks per slice of the
// slowest direction, G slices per XCD slab; within a slab the slice index runs
// fastest.
__device__ __forceinline__ int64_t slab_block(int64_t j, int64_t P, int64_t G)
{
    const int64_t q = j / (G * P), r = j % (G * P);
    return (q * G + r % G) * P + r / G;
}

template <int D, typename S, typename T>
__device__ __forceinline__ void staple_sum(const vec2_t<S>* __restrict__ U, const Geom& g, int mu,
                                           int64_t s, const Neighbours<D>& nb, mat3<T>& sigma)
{
    set_zero(sigma);
    const int64_t s_pmu = s + nb.up[mu];
#pragma unroll
    for (int nu = 0; nu < D; ++nu) {
        if (nu == mu) continue;
        // upper staple: U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
        {
            const mat3<T> a = load_link<T, S>(U, g.npad, nu, s_pmu);
            const mat3<T> b = load_link<T, S>(U, g.npad, mu, s + nb.up[nu]);
            mat3<T> ab;
            mul<false, true>(a, b, ab);
            const mat3<T> c = load_link<T, S>(U, g.npad, nu, s);
            mul_add<false, true>(ab, c, sigma);
        }
        // lower staple: U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)
        {
            const int64_t s_mnu = s + nb.dn[nu];
            const mat3<T> a = load_link<T, S>(U, g.npad, nu, s_pmu + nb.dn[nu]);
            const mat3<T> b = load_link<T, S>(U, g.npad, mu, s_mnu);
            mat3<T> ab;
            mul<true, true>(a, b, ab);
            const mat3<T> c = load_link<T, S>(U, g.npad, nu, s_mnu);
            mul_add<false, false>(ab, c, sigma);
        }
    }
}

// All D staple sums of site s.  Upper staples pairwise through
// W = U_nu(x+mu) U_mu(x+nu)^+:  Sigma_mu += W U_nu(x)^+,
// Sigma_nu += W^+ U_mu(x)^+.
template <int D, typename S, typename T>
__device__ __forceinline__ void staple_sums_site(const vec2_t<S>* __restrict__ U, const Geom& g,
                                                 int64_t s, const Neighbours<D>& nb,
                                                 mat3<T> (&sig)[D])
{
#pragma unroll
    for (int mu = 0; mu < D; ++mu) set_zero(sig[mu]);
#pragma unroll
    for (int mu = 0; mu < D; ++mu)
#pragma unroll
        for (int nu = mu + 1; nu < D; ++nu) {
            mat3<T> w;
            {
                const mat3<T> a = load_link<T, S>(U, g.npad, nu, s + nb.up[mu]);
                const mat3<T> b = load_link<T, S>(U, g.npad, mu, s + nb.up[nu]);
                mul<false, true>(a, b, w);
            }
            {
                const mat3<T> c = load_link<T, S>(U, g.npad, nu, s);
                mul_add<false, true>(w, c, sig[mu]);
            }
            {
                const mat3<T> c = load_link<T, S>(U, g.npad, mu, s);
                mul_add<true, true>(w, c, sig[nu]);
            }
        }
#pragma unroll
    for (int nu = 0; nu < D; ++nu) {
        const int64_t s_mnu = s + nb.dn[nu];
        const mat3<T> c = load_link<T, S>(U, g.npad, nu, s_mnu);
#pragma unroll
        for (int mu = 0; mu < D; ++mu) {
            if (mu == nu) continue;
            const mat3<T> a = load_link<T, S>(U, g.npad, nu, s_mnu + nb.up[mu]);
            const mat3<T> b = load_link<T, S>(U, g.npad, mu, s_mnu);
            mat3<T> ab;
            mul<true, true>(a, b, ab);
            mul_add<false, false>(ab, c, sig[mu]);
        }
    }
}

// drift of one link with momentum e:
//   additive (integrate_link, src/integrator/mod.rs:216-233):
//       U' = U + dt * (E U) * i sqrt(6) / a
//   exp:  U' = exp(i dt sqrt(6)/a E) U  (su3_exp_i, src/su3.rs:832-855)
template <typename T, int MODE>
__device__ __forceinline__ void drift_link(const mat3<T>& u, const T* e, T f, mat3<T>& out)
{
    if (MODE == DRIFT_ADDITIVE) {
        mat3<T> em, eu;
        adjoint_to_matrix(e, em);
        mul<false, false>(em, u, eu);
#pragma unroll
        for (int k = 0; k < 9; ++k)
            out.e[k] = {fma(-f, eu.e[k].im, u.e[k].re), fma(f, eu.e[k].re, u.e[k].im)};
    } else {
        T v[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) v[a] = f * e[a];
        mat3<T> r;
        exp_i(v, r);
        mul<false, false>(r, u, out);
    }
}

// what happens with the force of one link (DR: -1 kick only, else drift mode)
template <int DR, typename S, typename T>
__device__ __forceinline__ void finish_link(const mat3<T>& u, const mat3<T>& sigma,
                                            vec2_t<S>* __restrict__ U_out, S* __restrict__ E,
                                            const Geom& g, int mu, int64_t s, T fscale, T kick,
                                            T dfac)
{
    mat3<T> m;
    mul<false, false>(u, sigma, m);
    T f[8], e[8];
    im_trace_generators(m, f);
    load_e<T, S>(E, g.npad, mu, s, e);
#pragma unroll
    for (int a = 0; a < 8; ++a) e[a] = fma(kick, fscale * f[a], e[a]);
    store_e<T, S>(E, g.npad, mu, s, e);
    if (DR >= 0) {
        mat3<T> un;
        drift_link<T, DR>(u, e, dfac, un);
        store_link<T, S>(U_out, g.npad, mu, s, un);
    }
}

template <int D, typename S, typename T>
__global__ __launch_bounds__(kBlock) void k_force_store(const vec2_t<S>* __restrict__ U, Geom g,
                                                        T fscale, double* __restrict__ out, double oscale)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= g.nsite) return;
    const Neighbours<D> nb = neighbours<D>(g, s);
    mat3<T> sigma, m;
    staple_sum<D, S, T>(U, g, mu, s, nb, sigma);
    const mat3<T> u = load_link<T, S>(U, g.npad, mu, s);
    mul<false, false>(u, sigma, m);
    T f[8];
    im_trace_generators(m, f);
    double* p = out + (s * D + mu) * 8;
#pragma unroll
    for (int a = 0; a < 8; ++a) p[a] = oscale * (double)(fscale * f[a]);
}

// VAR_LINK: grid (site blocks, D)
template <int D, typename S, typename T, int DR>
__global__ __launch_bounds__(kBlock) void k_md_link(const vec2_t<S>* __restrict__ U_in,
                                                    vec2_t<S>* __restrict__ U_out,
                                                    S* __restrict__ E, Geom g, T fscale, T kick,
                                                    T dfac)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= g.nsite) return;
    const Neighbours<D> nb = neighbours<D>(g, s);
    mat3<T> sigma;
    staple_sum<D, S, T>(U_in, g, mu, s, nb, sigma);
    const mat3<T> u = load_link<T, S>(U_in, g.npad, mu, s);
    finish_link<DR, S, T>(u, sigma, U_out, E, g, mu, s, fscale, kick, dfac);
}

// VAR_LINK_XCD: 1-D grid of (site block, mu) pairs, mu fastest, XCD remapped
template <int D, typename S, typename T, int DR>
__global__ __launch_bounds__(kBlock) void k_md_link_xcd(const vec2_t<S>* __restrict__ U_in,
                                                        vec2_t<S>* __restrict__ U_out,
                                                        S* __restrict__ E, Geom g, T fscale, T kick,
                                                        T dfac)
{
    const int64_t logical = xcd_remap(blockIdx.x, gridDim.x);
    const int mu = (int)(logical % D);
    const int64_t s = (logical / D) * kBlock + threadIdx.x;
    if (s >= g.nsite) return;
    const Neighbours<D> nb = neighbours<D>(g, s);
    mat3<T> sigma;
    staple_sum<D, S, T>(U_in, g, mu, s, nb, sigma);
    const mat3<T> u = load_link<T, S>(U_in, g.npad, mu, s);
    finish_link<DR, S, T>(u, sigma, U_out, E, g, mu, s, fscale, kick, dfac);
}

template <int D, typename S, typename T, int DR>
__device__ __forceinline__ void drift_site(const vec2_t<S>* __restrict__ U_in, vec2_t<S>* __restrict__ U_out,
                                           const Geom& g, int64_t s, const T (&e)[D][8], T dfac)
{
#pragma unroll
    for (int mu = 0; mu < D; ++mu) {
        const mat3<T> u = load_link<T, S>(U_in, g.npad, mu, s);
        mat3<T> un;
        drift_link<T, DR>(u, e[mu], dfac, un);
        store_link<T, S>(U_out, g.npad, mu, s, un);
    }
}

// VAR_SITE_XCD / VAR_SITE_SLAB (slab_p > 0): one thread per site, all D links
template <int D, typename S, typename T, int DR>
__global__ __launch_bounds__(kBlock) void k_md_site_xcd(const vec2_t<S>* __restrict__ U_in,
                                                        vec2_t<S>* __restrict__ U_out,
                                                        S* __restrict__ E, Geom g, T fscale, T kick,
                                                        T dfac, int64_t slab_p, int64_t slab_g)
{
    int64_t logical = xcd_remap(blockIdx.x, gridDim.x);
    if (slab_p > 0) logical = slab_block(logical, slab_p, slab_g);
    const int64_t s = logical * kBlock + threadIdx.x;
    if (s >= g.nsite) return;
    const Neighbours<D> nb = neighbours<D>(g, s);
    // phase 1: the D forces (the staple sums die here)
    T f[D][8];
    {
        mat3<T> sig[D];
        staple_sums_site<D, S, T>(U_in, g, s, nb, sig);
#pragma unroll
        for (int mu = 0; mu < D; ++mu) {
            const mat3<T> u = load_link<T, S>(U_in, g.npad, mu, s);
            mat3<T> m;
            mul<false, false>(u, sig[mu], m);
            im_trace_generators(m, f[mu]);
        }
    }
    // phase 2: momentum kicks of the D links
    T e[D][8];
#pragma unroll
    for (int mu = 0; mu < D; ++mu) {
        load_e<T, S>(E, g.npad, mu, s, e[mu]);
#pragma unroll
        for (int a = 0; a < 8; ++a) e[mu][a] = fma(kick, fscale * f[mu][a], e[mu][a]);
        store_e<T, S>(E, g.npad, mu, s, e[mu]);
    }
    // phase 3: drifts (U_mu(x) is read again, from cache)
    if (DR >= 0) drift_site<D, S, T, DR>(U_in, U_out, g, s, e, dfac);
}

template <int D, typename S, typename T, int MODE>
__global__ __launch_bounds__(kBlock) void k_drift(const vec2_t<S>* __restrict__ U_in,
                                                  vec2_t<S>* __restrict__ U_out,
                                                  const S* __restrict__ E, Geom g, T dfac)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= g.nsite) return;
    T e[8];
    const mat3<T> u = load_link<T, S>(U_in, g.npad, mu, s);
    load_e<T, S>(E, g.npad, mu, s, e);
    mat3<T> un;
    drift_link<T, MODE>(u, e, dfac, un);
    store_link<T, S>(U_out, g.npad, mu, s, un);
}

static dim3 link_grid(const lqcd_lattice* L)
{
    return dim3(grid_for(L->g.nsite), (unsigned)L->g.dim);
}

// -sqrt(2 / CA) / a with CA = 3 (derivative_e, src/simulation/state.rs:1426)
static double force_scale(const lqcd_lattice* L) { return -sqrt(2.0 / 3.0) / L->size; }
// dt * sqrt(2 CA) / a (derivative_u, src/simulation/state.rs:1413-1416)
static double drift_scale(const lqcd_lattice* L, double dt) { return dt * sqrt(6.0) / L->size; }

template <int D, typename S, int DR>
static void launch_md_variant(const lqcd_lattice* L, int variant, const void* U_in, void* U_out, void* E,
                              double kick, double dfac)
{
    const S fs = (S)force_scale(L), k = (S)kick, df = (S)dfac;
    const auto* ui = (const vec2_t<S>*)U_in;
    auto* uo = (vec2_t<S>*)U_out;
    auto* e = (S*)E;
    hipStream_t st = L->ctx->stream;
    const unsigned nsb = grid_for(L->g.nsite);
    switch (variant) {
    case VAR_SITE_XCD:
    case VAR_SITE_SLAB: {
        int64_t slab_p = 0, slab_g = 0;
        const int64_t Ls = L->g.extent[D - 1];
        const int64_t slice = L->g.nsite / Ls;
        if (variant == VAR_SITE_SLAB && slice % kBlock == 0 && Ls % 8 == 0) {
            slab_p = slice / kBlock;
            slab_g = Ls / 8;
        }
        k_md_site_xcd<D, S, S, DR><<<nsb, kBlock, 0, st>>>(ui, uo, e, L->g, fs, k, df, slab_p, slab_g);
        break;
    }
    case VAR_LINK_XCD:
        k_md_link_xcd<D, S, S, DR><<<nsb * (unsigned)D, kBlock, 0, st>>>(ui, uo, e, L->g, fs, k, df);
        break;
    default:
        k_md_link<D, S, S, DR><<<link_grid(L), kBlock, 0, st>>>(ui, uo, e, L->g, fs, k, df);
        break;
    }
}

static int chosen_variant(const lqcd_lattice* L)
{
    return (L->md_variant >= 0 && L->md_variant < VAR_COUNT) ? L->md_variant : VAR_LINK_XCD;
}

hipError_t launch_force_store(const lqcd_lattice* L, const void* U, double* d_out, double scale)
{
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_FORCE_STORE, &ev);
    LQCD_DISPATCH_DS(L, k_force_store<D, S, S><<<link_grid(L), kBlock, 0, L->ctx->stream>>>(
                            (const vec2_t<S>*)U, L->g, (S)force_scale(L), d_out, scale));
    prof_end(L->ctx, &ev);
    return hipGetLastError();
}

hipError_t launch_kick_variant(const lqcd_lattice* L, int variant, const void* U, void* E, double kick)
{
    LQCD_DISPATCH_DS(L, launch_md_variant<D, S, -1>(L, variant, U, nullptr, E, kick, 0.0));
    return hipGetLastError();
}

hipError_t launch_kick_drift_variant(const lqcd_lattice* L, int variant, const void* U_in, void* U_out,
                                     void* E, double kick, double dt, int drift_mode)
{
    const double df = drift_scale(L, dt);
    if (drift_mode == DRIFT_EXP) {
        LQCD_DISPATCH_DS(L, launch_md_variant<D, S, DRIFT_EXP>(L, variant, U_in, U_out, E, kick, df));
    } else {
        LQCD_DISPATCH_DS(L, launch_md_variant<D, S, DRIFT_ADDITIVE>(L, variant, U_in, U_out, E, kick, df));
    }
    return hipGetLastError();
}

hipError_t launch_kick(const lqcd_lattice* L, const void* U, void* E, double kick)
{
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_KICK, &ev);
    const hipError_t e = launch_kick_variant(L, chosen_variant(L), U, E, kick);
    prof_end(L->ctx, &ev);
    return e;
}

hipError_t launch_kick_drift(const lqcd_lattice* L, const void* U_in, void* U_out, void* E, double kick,
                             double dt, int drift_mode)
{
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_KICK_DRIFT, &ev);
    const hipError_t e =
        launch_kick_drift_variant(L, chosen_variant(L), U_in, U_out, E, kick, dt, drift_mode);
    prof_end(L->ctx, &ev);
    return e;
}

hipError_t launch_drift(const lqcd_lattice* L, const void* U_in, void* U_out, const void* E, double dt,
                        int drift_mode)
{
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_DRIFT, &ev);
    if (drift_mode == DRIFT_EXP) {
        LQCD_DISPATCH_DS(L, k_drift<D, S, S, DRIFT_EXP><<<link_grid(L), kBlock, 0, L->ctx->stream>>>(
                                (const vec2_t<S>*)U_in, (vec2_t<S>*)U_out, (const S*)E, L->g,
                                (S)drift_scale(L, dt)));
    } else {
        LQCD_DISPATCH_DS(L, k_drift<D, S, S, DRIFT_ADDITIVE><<<link_grid(L), kBlock, 0, L->ctx->stream>>>(
                                (const vec2_t<S>*)U_in, (vec2_t<S>*)U_out, (const S*)E, L->g,
                                (S)drift_scale(L, dt)));
    }
    prof_end(L->ctx, &ev);
    return hipGetLastError();
}

} // namespace lqcd
