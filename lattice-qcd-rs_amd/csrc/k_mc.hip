// k_mc.hip -- checkerboard Metropolis sub-sweep.
//
// MetropolisHastingsSweep (src/simulation/monte_carlo/metropolis_hastings_sweep.rs:146-174)
// updates links one by one in storage order.  Links U_mu(x) of one class
// (mu fixed, x of fixed parity) never enter each other's staples, so a class
// is updated in one launch; a sweep is the 2*D classes for mu = 0..D-1,
// parity 0 then 1.  Per link, in double arithmetic:
//   A        = staple (src/simulation/monte_carlo/mod.rs:339-362)
//   U'       = R_k ... R_1 U with R_j given or drawn on the device
//   dS       = -(beta/3) Re Tr((U' - U) A)       (mod.rs:324-334)
//   p        = clamp(exp(-dS), 0, 1), NaN -> 1 (Rust min/max semantics)
//   accepted = p == 1 or u < floor(p * 2^64)     (rand 0.8 Bernoulli)
#include "dispatch.hpp"
#include "fields.hpp"
#include "rng_device.hpp"

namespace lqcd {

// su2::random_su2_close_to_unity (src/su2.rs:80-100) from four uniforms in
// [0, 1): r = 2 w - 1 (Uniform(-1, 1)), x = spread * r / |r| (r kept when
// |r| <= eps), x0 = +-sqrt(1 - |x|^2) (+ when w3 < 0.5, i.e. the Bernoulli(0.5)
// draw is true), then complex_matrix_from_vec (src/su2.rs:134-140):
// x0 + i sum x_k sigma_k with the reference's sigma_3 = identity
// (src/su2.rs:39-45).  m = [m00, m01, m10, m11].
__device__ __forceinline__ void su2_close_to_unity(const double* w, double spread, cplx<double>* m)
{
    const double r0 = 2.0 * w[0] - 1.0, r1 = 2.0 * w[1] - 1.0, r2 = 2.0 * w[2] - 1.0;
    const double n = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
    const bool keep = n <= 2.220446049250313e-16;
    const double x1 = (keep ? r0 : r0 / n) * spread;
    const double x2 = (keep ? r1 : r1 / n) * spread;
    const double x3 = (keep ? r2 : r2 / n) * spread;
    const double x0u = sqrt(1.0 - (x1 * x1 + x2 * x2 + x3 * x3));
    const double x0 = w[3] < 0.5 ? x0u : -x0u;
    m[0] = {x0, x3};
    m[1] = {x2, x1};
    m[2] = {-x2, x1};
    m[3] = {x0, x3};
}

// orthonormalize(random_su3_close_to_unity(spread)) (src/su3.rs:385-398,
// metropolis_hastings_sweep.rs:126-143): R = get_r(a) get_s(b) get_t(c), its
// adjoint when the last Bernoulli(0.5) is true, then the column Gram-Schmidt.
// Draws of update j of link l: Philox blocks 64 + 8 j + (0..6), 14 uniforms
// w0..w13 (a: w0..w3, b: w4..w7, c: w8..w11, adjoint: w12).
__device__ __forceinline__ void device_proposal(uint64_t l, int j, uint64_t seed, uint64_t counter,
                                                double spread, mat3<double>& out)
{
    double w[14];
#pragma unroll
    for (int b = 0; b < 7; ++b) {
        const uint4 o = philox_link(l, 64u + 8u * (uint32_t)j + (uint32_t)b, counter, seed);
        w[2 * b] = u53(o.x, o.y);
        w[2 * b + 1] = u53(o.z, o.w);
    }
    cplx<double> a[4], bm[4], c[4];
    su2_close_to_unity(w + 0, spread, a);
    su2_close_to_unity(w + 4, spread, bm);
    su2_close_to_unity(w + 8, spread, c);
    mat3<double> R, S, T, RS, X;
    set_zero(R);
    set_zero(S);
    set_zero(T);
    R.e[0] = a[0]; R.e[1] = a[1]; R.e[3] = a[2]; R.e[4] = a[3]; R.e[8] = {1.0, 0.0};
    S.e[0] = bm[0]; S.e[2] = bm[1]; S.e[6] = bm[2]; S.e[8] = bm[3]; S.e[4] = {1.0, 0.0};
    T.e[0] = {1.0, 0.0}; T.e[4] = c[0]; T.e[5] = c[1]; T.e[7] = c[2]; T.e[8] = c[3];
    mul<false, false>(R, S, RS);
    mul<false, false>(RS, T, X);
    if (w[12] < 0.5) {
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int k = 0; k < 3; ++k) out.e[r * 3 + k] = {X.e[k * 3 + r].re, -X.e[k * 3 + r].im};
    } else {
        out = X;
    }
    orthonormalize(out);
}

// the accept word of link l: Philox block 63, (x << 32) | y
__device__ __forceinline__ uint64_t device_accept_word(uint64_t l, uint64_t seed, uint64_t counter)
{
    const uint4 o = philox_link(l, 63u, counter, seed);
    return ((uint64_t)o.x << 32) | (uint64_t)o.y;
}

template <int D, typename S, bool RNG>
__global__ __launch_bounds__(kBlock) void k_metropolis(vec2_t<S>* __restrict__ U, Geom g, int mu,
                                                       int parity, double beta, int n_update,
                                                       const double* __restrict__ proposals,
                                                       const uint64_t* __restrict__ uniforms,
                                                       double spread, uint64_t seed, uint64_t counter,
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
        mat3<double> r, t;
        if (RNG) {
            device_proposal((uint64_t)li, j, seed, counter, spread, r);
        } else {
            const double* pr = proposals + ((int64_t)j * nlink + li) * 18;
#pragma unroll
            for (int rr = 0; rr < 3; ++rr)
#pragma unroll
                for (int cc = 0; cc < 3; ++cc)
                    r.e[rr * 3 + cc] = {pr[(cc * 3 + rr) * 2], pr[(cc * 3 + rr) * 2 + 1]};
        }
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
        const uint64_t word = RNG ? device_accept_word((uint64_t)li, seed, counter) : uniforms[li];
        accept = word < p_int;
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
    LQCD_DISPATCH_DS(L, k_metropolis<D, S, false><<<grid, kBlock, 0, L->ctx->stream>>>(
                            (vec2_t<S>*)U, L->g, mu, parity, beta, (int)n_update, d_proposals,
                            d_uniforms, 0.0, 0, 0, L->mh_p, L->mh_acc));
    prof_end(L->ctx, &ev);
    return hipGetLastError();
}

hipError_t launch_metropolis_class_rng(const lqcd_lattice* L, void* U, int mu, int parity, double beta,
                                       size_t n_update, double spread, uint64_t seed, uint64_t counter)
{
    const unsigned grid = grid_for(L->g.nsite / 2);
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_METROPOLIS, &ev);
    LQCD_DISPATCH_DS(L, k_metropolis<D, S, true><<<grid, kBlock, 0, L->ctx->stream>>>(
                            (vec2_t<S>*)U, L->g, mu, parity, beta, (int)n_update, nullptr, nullptr,
                            spread, seed, counter, L->mh_p, L->mh_acc));
    prof_end(L->ctx, &ev);
    return hipGetLastError();
}

} // namespace lqcd
