// k_mc.hip -- checkerboard Metropolis sub-sweep.
//
// MetropolisHastingsSweep (src/simulation/monte_carlo/metropolis_hastings_sweep.rs:146-174)
// updates links one by one in storage order.  Links U_mu(x) of one class
// (mu fixed, x of fixed parity) never enter each other's staples, so a class
// is updated in one launch; a sweep is the 2*D classes for mu = 0..D-1,
// parity 0 then 1.  Per link, in double arithmetic:
//   A        = staple (src/simulation/monte_carlo/mod.rs:339-362)
//   U'       = R_k ... R_1 U with the given proposals
//   dS       = -(beta/3) Re Tr((U' - U) A)       (mod.rs:324-334)
//   p        = clamp(exp(-dS), 0, 1), NaN -> 1 (Rust min/max semantics)
//   accepted = p == 1 or u < floor(p * 2^64)     (rand 0.8 Bernoulli)
#include "dispatch.hpp"
#include "fields.hpp"

namespace lqcd {

template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_metropolis(vec2_t<S>* __restrict__ U, Geom g, int mu,
                                                       int parity, double beta, int n_update,
                                                       const double* __restrict__ proposals,
                                                       const uint64_t* __restrict__ uniforms,
                                                       double* __restrict__ p_out,
                                                       uint32_t* __restrict__ acc_out)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t half = g.nsite / 2;
    if (k >= half) return;
    // extent[0] is even: sites 2k and 2k+1 differ only in x_0 (even -> odd),
    // so exactly one of them has the requested parity.
    const int64_t s0 = 2 * k;
    const int par0 = site_parity<D>(g, s0);
    const int64_t s = s0 + ((par0 + parity) & 1);
    const Neighbours<D> nb = neighbours<D>(g, s);

    mat3<double> a;
    set_zero(a);
    const int64_t s_pmu = s + nb.up[mu];
#pragma unroll
    for (int nu = 0; nu < D; ++nu) {
        if (nu == mu) continue;
        {
            const mat3<double> x = load_link<double, S>(U, g.npad, nu, s_pmu);
            const mat3<double> y = load_link<double, S>(U, g.npad, mu, s + nb.up[nu]);
            mat3<double> xy;
            mul<false, true>(x, y, xy);
            const mat3<double> z = load_link<double, S>(U, g.npad, nu, s);
            mul_add<false, true>(xy, z, a);
        }
        {
            const int64_t s_mnu = s + nb.dn[nu];
            const mat3<double> x = load_link<double, S>(U, g.npad, nu, s_pmu + nb.dn[nu]);
            const mat3<double> y = load_link<double, S>(U, g.npad, mu, s_mnu);
            mat3<double> xy;
            mul<true, true>(x, y, xy);
            const mat3<double> z = load_link<double, S>(U, g.npad, nu, s_mnu);
            mul_add<false, false>(xy, z, a);
        }
    }
    const mat3<double> u = load_link<double, S>(U, g.npad, mu, s);
    const int64_t li = s * D + mu;
    const int64_t nlink = g.nsite * D;
    mat3<double> nw = u;
    for (int j = 0; j < n_update; ++j) {
        const double* pr = proposals + ((int64_t)j * nlink + li) * 18;
        mat3<double> r, t;
#pragma unroll
        for (int rr = 0; rr < 3; ++rr)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
                r.e[rr * 3 + cc] = {pr[(cc * 3 + rr) * 2], pr[(cc * 3 + rr) * 2 + 1]};
        mul<false, false>(r, nw, t);
        nw = t;
    }
    mat3<double> diff;
#pragma unroll
    for (int e = 0; e < 9; ++e) diff.e[e] = {nw.e[e].re - u.e[e].re, nw.e[e].im - u.e[e].im};
    const double ds = -trace_mul<false, false>(diff, a).re * beta / 3.0;
    const double p = fmax(fmin(exp(-ds), 1.0), 0.0);
    bool accept;
    if (p >= 1.0) {
        accept = true;
    } else {
        const uint64_t p_int = (uint64_t)(p * 18446744073709551616.0);
        accept = uniforms[li] < p_int;
    }
    p_out[li] = p;
    acc_out[li] = accept ? 1u : 0u;
    if (accept) store_link<double, S>(U, g.npad, mu, s, nw);
}

hipError_t launch_metropolis_class(const lqcd_lattice* L, void* U, int mu, int parity, double beta,
                                   size_t n_update, const double* d_proposals,
                                   const uint64_t* d_uniforms)
{
    const unsigned grid = grid_for(L->g.nsite / 2);
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_METROPOLIS, &ev);
    LQCD_DISPATCH_DS(L, k_metropolis<D, S><<<grid, kBlock, 0, L->ctx->stream>>>(
                            (vec2_t<S>*)U, L->g, mu, parity, beta, (int)n_update, d_proposals,
                            d_uniforms, L->mh_p, L->mh_acc));
    prof_end(L->ctx, &ev);
    return hipGetLastError();
}

} // namespace lqcd
