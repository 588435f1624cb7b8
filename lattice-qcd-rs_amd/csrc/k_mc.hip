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

// staple A_mu(x) of monte_carlo/mod.rs:339-362 in double arithmetic:
// sum_{nu != mu} U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ + U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)
// R12: links read as two rows (load_link_r12), for links in SU(3)
template <int D, typename S, bool R12 = false>
__device__ __forceinline__ void mc_staple(const vec2_t<S>* __restrict__ U, const Geom& g, int mu, int64_t s,
                                          mat3<double>& a)
{
    const Neighbours<D> nb = neighbours<D>(g, s);
    set_zero(a);
    const int64_t s_pmu = s + nb.up[mu];
#pragma unroll
    for (int nu = 0; nu < D; ++nu) {
        if (nu == mu) continue;
        {
            const mat3<double> x = load_link_sel<R12, double, S>(U, g.npad, nu, s_pmu);
            const mat3<double> y = load_link_sel<R12, double, S>(U, g.npad, mu, s + nb.up[nu]);
            mat3<double> xy;
            mul<false, true>(x, y, xy);
            const mat3<double> z = load_link_sel<R12, double, S>(U, g.npad, nu, s);
            mul_add<false, true>(xy, z, a);
        }
        {
            const int64_t s_mnu = s + nb.dn[nu];
            const mat3<double> x = load_link_sel<R12, double, S>(U, g.npad, nu, s_pmu + nb.dn[nu]);
            const mat3<double> y = load_link_sel<R12, double, S>(U, g.npad, mu, s_mnu);
            mat3<double> xy;
            mul<true, true>(x, y, xy);
            const mat3<double> z = load_link_sel<R12, double, S>(U, g.npad, nu, s_mnu);
            mul_add<false, false>(xy, z, a);
        }
    }
}

// site of class (parity) handled by thread k: extent[0] is even, so sites 2k
// and 2k+1 differ only in x_0 and exactly one of them has the parity
template <int D>
__device__ __forceinline__ int64_t class_site(const Geom& g, int64_t k, int parity)
{
    const int64_t s0 = 2 * k;
    return s0 + ((site_parity<D>(g, s0) + parity) & 1);
}

// Order of a class launch's workgroups: each covers 2 kBlock consecutive
// sites (kBlock class links); XCD k gets a contiguous range of them
// (xcd_remap), walked slab by slab of the slowest direction with that
// direction fastest (slab_block, when slab_p > 0), so the workgroups whose
// staples share links run on one XCD close together in time and find them in
// its L2.  The class index k of a thread does not depend on the order, so the
// results do not either.
struct ClassOrder {
    int64_t slab_p = 0;   // workgroups per slice of the slowest direction (0: XCD remap only)
    int64_t slab_g = 0;   // slices per XCD slab
};

__device__ __forceinline__ int64_t class_index(const ClassOrder& co)
{
    int64_t logical = xcd_remap(blockIdx.x, gridDim.x);
    if (co.slab_p > 0) logical = slab_block(logical, co.slab_p, co.slab_g);
    return logical * kBlock + threadIdx.x;
}

template <int D, typename S, bool RNG, bool R12 = false>
__global__ __launch_bounds__(kBlock) void k_metropolis(vec2_t<S>* __restrict__ U, Geom g, int mu,
                                                       int parity, double beta, int n_update,
                                                       const double* __restrict__ proposals,
                                                       const uint64_t* __restrict__ uniforms,
                                                       double spread, uint64_t seed, uint64_t counter,
                                                       double* __restrict__ p_out,
                                                       uint32_t* __restrict__ acc_out, ClassOrder co)
{
    const int64_t k = class_index(co);
    const int64_t half = g.nsite / 2;
    if (k >= half) return;
    const int64_t s = class_site<D>(g, k, parity);
    mat3<double> a;
    mc_staple<D, S, R12>(U, g, mu, s, a);
    const mat3<double> u = load_link_sel<R12, double, S>(U, g.npad, mu, s);
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
    // per-link outputs in class order (contiguous for the launch); only their
    // sums are read
    const int64_t o = (int64_t)(mu * 2 + parity) * half + k;
    p_out[o] = p;
    acc_out[o] = accept ? 1u : 0u;
    if (accept) store_link<double, S>(U, g.npad, mu, s, nw);
}

// MetropolisHastingsDeltaDiagnostic::next_element_default
// (metropolis_hastings.rs:394-419, potential_modif :374-391): one link drawn
// uniformly, new link = orthonormalize(random_su3_close_to_unity(spread)) U,
// dS of delta_s_old_new_cmp (monte_carlo/mod.rs:324-334), accepted with
// p = clamp(exp(-dS), 0, 1).  Draws of step `counter`: the link from Philox
// block 62 of link 0 (site = (x y) * nsite >> 64 over the 64-bit word of
// x, y; direction = z * D >> 32), then the proposal and accept word of that
// link as in the device sweep.  One lane works; out = p, accepted, link.
template <int D, typename S>
__global__ __launch_bounds__(64) void k_metropolis_link(vec2_t<S>* __restrict__ U, Geom g, double beta,
                                                        double spread, uint64_t seed, uint64_t counter,
                                                        double* __restrict__ out)
{
    if (threadIdx.x != 0) return;
    const uint4 o = philox_link(0, 62u, counter, seed);
    const uint64_t w = ((uint64_t)o.x << 32) | (uint64_t)o.y;
    const int64_t s = (int64_t)__umul64hi(w, (uint64_t)g.nsite);
    const int mu = (int)(((uint64_t)o.z * (uint64_t)D) >> 32);
    mat3<double> a, r, nw;
    mc_staple<D, S>(U, g, mu, s, a);
    const mat3<double> u = load_link<double, S>(U, g.npad, mu, s);
    const int64_t li = s * D + mu;
    device_proposal((uint64_t)li, 0, seed, counter, spread, r);
    mul<false, false>(r, u, nw);
    mat3<double> diff;
#pragma unroll
    for (int e = 0; e < 9; ++e) diff.e[e] = {nw.e[e].re - u.e[e].re, nw.e[e].im - u.e[e].im};
    const double ds = -trace_mul<false, false>(diff, a).re * beta / 3.0;
    const double p = fmax(fmin(exp(-ds), 1.0), 0.0);
    bool accept = true;
    if (p < 1.0) accept = device_accept_word((uint64_t)li, seed, counter) < (uint64_t)(p * 18446744073709551616.0);
    out[0] = p;
    out[1] = accept ? 1.0 : 0.0;
    out[2] = (double)li;
    if (accept) store_link<double, S>(U, g.npad, mu, s, nw);
}

// ---------------------------------------------------------------- heat bath
// HeatBathSweep (simulation/monte_carlo/heat_bath.rs:73-123) with draws from
// the device stream: subgroup g (0 r, 1 s, 2 t) of link l in sweep n uses
// Philox blocks kHeatBathBlock + 256 g + k (key seed, counter n):
//   k = 2t, 2t+1 (t < kHeatBathAttempts) attempt t of the norm sampler:
//       w0 = r in [0,1), r_i = 1 - w_{i+1} in (0,1] for ModifiedNormal
//   k = 128 + 2q, +1 (q < 2)  direction draw q: x_i = 2 w_i - 1
//   k = 132 + 2q, +1 (q < 2)  random_su2 draw q: (re0, im0, re1, im1) = 2 w - 1
constexpr int kHeatBathAttempts = 64;

__device__ __forceinline__ void hb_draws(uint64_t l, uint32_t blk, uint64_t n, uint64_t seed, double (&w)[4])
{
#pragma unroll
    for (uint32_t b = 0; b < 2; ++b) {
        const uint4 o = philox_link(l, blk + b, n, seed);
        w[2 * b] = u53(o.x, o.y);
        w[2 * b + 1] = u53(o.z, o.w);
    }
}

__device__ __forceinline__ cplx<double> conj2(cplx<double> a) { return {a.re, -a.im}; }

// heat_bath_su2 (heat_bath.rs:73-87): k = sqrt(Re det M); if k is normal,
// X v with v = M^+ / k and X from HeatBathDistribution(beta k)
// (distribution.rs:193-214: x0 from HeatBathDistributionNorm, a random
// direction scaled to sqrt(1 - x0^2), complex_matrix_from_vec with the
// reference's sigma_3 = identity); otherwise su2::random_su2.  The norm
// sampler gives up after kHeatBathAttempts tries: the factor is then the
// identity and `exhausted` counts it.  2x2 matrices row-major.
__device__ __forceinline__ void hb_su2(const cplx<double> (&st)[4], double beta, uint64_t l, int grp, uint64_t n,
                                       uint64_t seed, cplx<double> (&out)[4], uint32_t& exhausted)
{
    const uint32_t base = kHeatBathBlock + 256u * (uint32_t)grp;
    const cplx<double> det = {st[0].re * st[3].re - st[0].im * st[3].im - (st[1].re * st[2].re - st[1].im * st[2].im),
                              st[0].re * st[3].im + st[0].im * st[3].re - (st[1].re * st[2].im + st[1].im * st[2].re)};
    const double k = sqrt(det.re);
    if (isfinite(k) && k >= 2.2250738585072014e-308) {
        const double p = beta * k;
        double x0 = 0.0;
        bool ok = false;
        for (int t = 0; t < kHeatBathAttempts && !ok; ++t) {
            double w[4];
            hb_draws(l, base + 2u * (uint32_t)t, n, seed, w);
            const double r = w[0];
            const double c = cos(2.0 * 3.14159265358979323846 * (1.0 - w[2]));
            const double lambda = sqrt(-(log(1.0 - w[1]) + c * c * log(1.0 - w[3])) / (2.0 * p));
            if (r * r <= 1.0 - lambda * lambda) {
                x0 = 1.0 - 2.0 * lambda * lambda;
                ok = true;
            }
        }
        if (!ok) {
            exhausted += 1;
            out[0] = {1.0, 0.0};
            out[1] = {0.0, 0.0};
            out[2] = {0.0, 0.0};
            out[3] = {1.0, 0.0};
            return;
        }
        double x[3] = {0.0, 0.0, 0.0}, nrm = 0.0;
        for (int q = 0; q < 2; ++q) {
            double w[4];
            hb_draws(l, base + 128u + 2u * (uint32_t)q, n, seed, w);
            for (int i = 0; i < 3; ++i) x[i] = 2.0 * w[i] - 1.0;
            nrm = sqrt((x[0] * x[0] + x[1] * x[1]) + x[2] * x[2]);
            if (nrm > 2.220446049250313e-16) break;
        }
        const double sc = sqrt(1.0 - x0 * x0) / nrm;
#pragma unroll
        for (int i = 0; i < 3; ++i) x[i] *= sc;
        const cplx<double> X[4] = {{x0, x[2]}, {x[1], x[0]}, {-x[1], x[0]}, {x0, x[2]}};
        const cplx<double> v[4] = {{st[0].re / k, -st[0].im / k}, {st[2].re / k, -st[2].im / k},
                                   {st[1].re / k, -st[1].im / k}, {st[3].re / k, -st[3].im / k}};
        out[0] = {X[0].re * v[0].re - X[0].im * v[0].im + X[1].re * v[2].re - X[1].im * v[2].im,
                  X[0].re * v[0].im + X[0].im * v[0].re + X[1].re * v[2].im + X[1].im * v[2].re};
        out[1] = {X[0].re * v[1].re - X[0].im * v[1].im + X[1].re * v[3].re - X[1].im * v[3].im,
                  X[0].re * v[1].im + X[0].im * v[1].re + X[1].re * v[3].im + X[1].im * v[3].re};
        out[2] = {X[2].re * v[0].re - X[2].im * v[0].im + X[3].re * v[2].re - X[3].im * v[2].im,
                  X[2].re * v[0].im + X[2].im * v[0].re + X[3].re * v[2].im + X[3].im * v[2].re};
        out[3] = {X[2].re * v[1].re - X[2].im * v[1].im + X[3].re * v[3].re - X[3].im * v[3].im,
                  X[2].re * v[1].im + X[2].im * v[1].re + X[3].re * v[3].im + X[3].im * v[3].re};
    } else {
        double w[4] = {0.0, 0.0, 0.0, 0.0};
        double nrm = 0.0;
        for (int q = 0; q < 2; ++q) {
            hb_draws(l, base + 132u + 2u * (uint32_t)q, n, seed, w);
            for (int i = 0; i < 4; ++i) w[i] = 2.0 * w[i] - 1.0;
            nrm = sqrt(((w[0] * w[0] + w[1] * w[1]) + w[2] * w[2]) + w[3] * w[3]);
            if (isfinite(nrm) && nrm >= 2.2250738585072014e-308) break;
        }
        const cplx<double> a0 = {w[0] / nrm, w[1] / nrm}, b0 = {w[2] / nrm, w[3] / nrm};
        out[0] = a0;
        out[1] = b0;
        out[2] = {-b0.re, b0.im};
        out[3] = conj2(a0);
    }
}

// sub-blocks of get_sub_block_r/s/t (su3.rs:559-640) and their embeddings
// get_r/get_s/get_t (su3.rs:428-530): rows/cols (0,1), (0,2), (1,2)
__device__ __forceinline__ void sub_block_rc(int grp, int& i, int& j)
{
    i = grp == 2 ? 1 : 0;
    j = grp == 0 ? 1 : 2;
}

template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_heat_bath(vec2_t<S>* __restrict__ U, Geom g, int mu, int parity,
                                                      double beta, uint64_t seed, uint64_t counter,
                                                      double* __restrict__ p_out, uint32_t* __restrict__ exhausted,
                                                      ClassOrder co)
{
    const int64_t k = class_index(co);
    if (k >= g.nsite / 2) return;
    const int64_t s = class_site<D>(g, k, parity);
    mat3<double> a;
    mc_staple<D, S>(U, g, mu, s, a);
    const mat3<double> u = load_link<double, S>(U, g.npad, mu, s);
    const int64_t li = s * D + mu;
    mat3<double> ua, acc;
    mul<false, false>(u, a, ua);
    set_identity(acc);
    uint32_t ex = 0;
#pragma unroll
    for (int grp = 0; grp < 3; ++grp) {
        // project_to_su2_unorm of the sub-block of (acc U) A (su2.rs:155-157)
        mat3<double> w;
        mul<false, false>(acc, ua, w);
        int i, j;
        sub_block_rc(grp, i, j);
        const cplx<double> m00 = w.e[i * 3 + i], m01 = w.e[i * 3 + j], m10 = w.e[j * 3 + i], m11 = w.e[j * 3 + j];
        const cplx<double> tr = {m00.re + m11.re, -(m00.im + m11.im)};
        const cplx<double> proj[4] = {{m00.re - m00.re + tr.re, m00.im + m00.im + tr.im},
                                      {m01.re - m10.re, m01.im + m10.im},
                                      {m10.re - m01.re, m10.im + m01.im},
                                      {m11.re - m11.re + tr.re, m11.im + m11.im + tr.im}};
        cplx<double> h[4];
        hb_su2(proj, beta, (uint64_t)li, grp, counter, seed, h, ex);
        mat3<double> f, t;
        set_identity(f);
        f.e[i * 3 + i] = h[0];
        f.e[i * 3 + j] = h[1];
        f.e[j * 3 + i] = h[2];
        f.e[j * 3 + j] = h[3];
        mul<false, false>(f, acc, t);
        acc = t;
    }
    mat3<double> nw;
    mul<false, false>(acc, u, nw);
    store_link<double, S>(U, g.npad, mu, s, nw);
    const int64_t o = (int64_t)(mu * 2 + parity) * (g.nsite / 2) + k;
    p_out[o] = 0.0;
    exhausted[o] = ex;
}

// --------------------------------------------------------- overrelaxation
// inverse of a 3x3 complex matrix by cofactors; returns det
__device__ __forceinline__ cplx<double> inverse3(const mat3<double>& a, mat3<double>& inv)
{
    cplx<double> c[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            const cplx<double> p = cmul(a.e[i1 * 3 + j1], a.e[i2 * 3 + j2]);
            const cplx<double> q = cmul(a.e[i1 * 3 + j2], a.e[i2 * 3 + j1]);
            c[i * 3 + j] = {p.re - q.re, p.im - q.im};
        }
    cplx<double> det = {0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const cplx<double> t = cmul(a.e[j], c[j]);
        det = {det.re + t.re, det.im + t.im};
    }
    const double dd = det.re * det.re + det.im * det.im;
    const cplx<double> idet = {det.re / dd, -det.im / dd};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) inv.e[i * 3 + j] = cmul(c[j * 3 + i], idet);
    return det;
}

// Unitary factor W of A = W P (W = U V^+ of the SVD A = U S V^+): scaled
// Newton iteration X <- (g X + (g X)^{-+}) / 2 with g = |det X|^{-1/3}.
__device__ __forceinline__ void polar_unitary(const mat3<double>& a, mat3<double>& x)
{
    x = a;
    for (int it = 0; it < 100; ++it) {
        mat3<double> inv;
        const cplx<double> det = inverse3(x, inv);
        const double gm = pow(sqrt(det.re * det.re + det.im * det.im), -1.0 / 3.0);
        double diff = 0.0, nrm = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const cplx<double> t = {inv.e[j * 3 + i].re, -inv.e[j * 3 + i].im};
                const cplx<double> v = {0.5 * (gm * x.e[i * 3 + j].re + t.re / gm),
                                        0.5 * (gm * x.e[i * 3 + j].im + t.im / gm)};
                diff += sqrt((v.re - x.e[i * 3 + j].re) * (v.re - x.e[i * 3 + j].re) +
                             (v.im - x.e[i * 3 + j].im) * (v.im - x.e[i * 3 + j].im));
                nrm += sqrt(v.re * v.re + v.im * v.im);
                x.e[i * 3 + j] = v;
            }
        if (diff <= 1e-15 * nrm) break;
    }
}

// Singular vectors of A = U S V^+ by one-sided Jacobi: plane rotations of
// column pairs (p, q) applied to W = A V until every pair is orthogonal; then
// the columns of W are U S.  Each rotation removes the Hermitian inner product
// w_p^+ w_q of the pair: with w_p^+ w_q = g e^{i phi}, the 2x2 Gram matrix
// [[|w_p|^2, g], [g, |w_q|^2]] is diagonalised by the real Jacobi angle t and
// the phase e^{i phi} moved onto column q.
__device__ __forceinline__ void svd3_jacobi(const mat3<double>& a, mat3<double>& u, mat3<double>& v)
{
    mat3<double> w = a;
    set_identity(v);
    for (int sweep = 0; sweep < 30; ++sweep) {
        double off = 0.0, diag = 0.0;
#pragma unroll
        for (int pq = 0; pq < 3; ++pq) {
            const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
            double al = 0.0, be = 0.0;
            cplx<double> ga = {0.0, 0.0};
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const cplx<double> x = w.e[i * 3 + p], y = w.e[i * 3 + q];
                al += x.re * x.re + x.im * x.im;
                be += y.re * y.re + y.im * y.im;
                ga.re += x.re * y.re + x.im * y.im;
                ga.im += x.re * y.im - x.im * y.re;
            }
            const double gn = hypot(ga.re, ga.im);
            off += gn;
            diag += sqrt(al * be);
            if (gn <= 1e-300) continue;
            const cplx<double> ph = {ga.re / gn, ga.im / gn};
            const double zeta = (be - al) / (2.0 * gn);
            const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const cplx<double> x = w.e[i * 3 + p], y = w.e[i * 3 + q];
                const cplx<double> yp = {y.re * ph.re + y.im * ph.im, y.im * ph.re - y.re * ph.im};
                const cplx<double> xq = {x.re * ph.re - x.im * ph.im, x.re * ph.im + x.im * ph.re};
                w.e[i * 3 + p] = {c * x.re - sn * yp.re, c * x.im - sn * yp.im};
                w.e[i * 3 + q] = {sn * xq.re + c * y.re, sn * xq.im + c * y.im};
                const cplx<double> vx = v.e[i * 3 + p], vy = v.e[i * 3 + q];
                const cplx<double> vyp = {vy.re * ph.re + vy.im * ph.im, vy.im * ph.re - vy.re * ph.im};
                const cplx<double> vxq = {vx.re * ph.re - vx.im * ph.im, vx.re * ph.im + vx.im * ph.re};
                v.e[i * 3 + p] = {c * vx.re - sn * vyp.re, c * vx.im - sn * vyp.im};
                v.e[i * 3 + q] = {sn * vxq.re + c * vy.re, sn * vxq.im + c * vy.im};
            }
        }
        if (off <= 1e-16 * diag) break;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double n = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) n += w.e[i * 3 + k].re * w.e[i * 3 + k].re + w.e[i * 3 + k].im * w.e[i * 3 + k].im;
        n = sqrt(n);
        const double r = n > 0.0 ? 1.0 / n : 1.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) u.e[i * 3 + k] = {w.e[i * 3 + k].re * r, w.e[i * 3 + k].im * r};
    }
}

// OverrelaxationSweepRotation::get_modif (overrelaxation.rs:66-78):
//   rot = polar factor of staple^+, new link = rot U^+ rot;
// OverrelaxationSweepReverse::get_modif (overrelaxation.rs:136-150):
//   staple^+ = W S V^+, new link = W reverse(W^+ U V) V^+, where su3::reverse
//   (su3.rs:705-714) negates the off-diagonal entries.  The result does not
//   depend on the phases or the order of the singular vectors.
template <int D, typename S, bool REVERSE>
__global__ __launch_bounds__(kBlock) void k_overrelax(vec2_t<S>* __restrict__ U, Geom g, int mu, int parity,
                                                      ClassOrder co)
{
    const int64_t k = class_index(co);
    if (k >= g.nsite / 2) return;
    const int64_t s = class_site<D>(g, k, parity);
    mat3<double> a, ad, t, nw;
    mc_staple<D, S>(U, g, mu, s, a);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) ad.e[i * 3 + j] = {a.e[j * 3 + i].re, -a.e[j * 3 + i].im};
    const mat3<double> u = load_link<double, S>(U, g.npad, mu, s);
    if (REVERSE) {
        mat3<double> w, v, m;
        svd3_jacobi(ad, w, v);
        mul<true, false>(w, u, t);
        mul<false, false>(t, v, m);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                if (i != j) m.e[i * 3 + j] = {-m.e[i * 3 + j].re, -m.e[i * 3 + j].im};
        mul<false, false>(w, m, t);
        mul<false, true>(t, v, nw);
    } else {
        mat3<double> rot;
        polar_unitary(ad, rot);
        mul<false, true>(rot, u, t);
        mul<false, false>(t, rot, nw);
    }
    store_link<double, S>(U, g.npad, mu, s, nw);
}

// the slab order of a class launch when the lattice allows it: whole
// workgroups per slice of the slowest direction, 8 | its extent
static ClassOrder class_order(const lqcd_lattice* L)
{
    ClassOrder co;
    const int64_t Ls = L->g.extent[L->g.dim - 1];
    const int64_t slice = L->g.nsite / Ls;
    if (slice % (2 * kBlock) == 0 && Ls % 8 == 0) {
        co.slab_p = slice / (2 * kBlock);
        co.slab_g = Ls / 8;
    }
    return co;
}

hipError_t launch_heat_bath_class(const lqcd_lattice* L, void* U, int mu, int parity, double beta, uint64_t seed,
                                  uint64_t counter)
{
    const unsigned grid = grid_for(L->g.nsite / 2);
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_HEAT_BATH, &ev);
    LQCD_DISPATCH_DS(L, k_heat_bath<D, S><<<grid, kBlock, 0, L->ctx->stream>>>(
                            (vec2_t<S>*)U, L->g, mu, parity, beta, seed, counter, L->mh_p, L->mh_acc,
                            class_order(L)));
    prof_end(L->ctx, &ev);
    return hipGetLastError();
}

hipError_t launch_overrelax_class(const lqcd_lattice* L, void* U, int mu, int parity, bool reverse)
{
    const unsigned grid = grid_for(L->g.nsite / 2);
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_OVERRELAX, &ev);
    if (reverse)
        LQCD_DISPATCH_DS(L, k_overrelax<D, S, true><<<grid, kBlock, 0, L->ctx->stream>>>((vec2_t<S>*)U, L->g,
                                                                                         mu, parity, class_order(L)));
    else
        LQCD_DISPATCH_DS(L, k_overrelax<D, S, false><<<grid, kBlock, 0, L->ctx->stream>>>((vec2_t<S>*)U, L->g,
                                                                                          mu, parity, class_order(L)));
    prof_end(L->ctx, &ev);
    return hipGetLastError();
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
                            d_uniforms, 0.0, 0, 0, L->mh_p, L->mh_acc, class_order(L)));
    prof_end(L->ctx, &ev);
    return hipGetLastError();
}

hipError_t launch_metropolis_link(const lqcd_lattice* L, void* U, double beta, double spread, uint64_t seed,
                                  uint64_t counter, double* d_out3)
{
    LQCD_DISPATCH_DS(L, k_metropolis_link<D, S><<<1, 64, 0, L->ctx->stream>>>(
                            (vec2_t<S>*)U, L->g, beta, spread, seed, counter, d_out3));
    return hipGetLastError();
}

hipError_t launch_metropolis_class_rng(const lqcd_lattice* L, void* U, int mu, int parity, double beta,
                                       size_t n_update, double spread, uint64_t seed, uint64_t counter)
{
    const unsigned grid = grid_for(L->g.nsite / 2);
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_METROPOLIS, &ev);
    if (L->md_r12)
        LQCD_DISPATCH_DS(L, k_metropolis<D, S, true, true><<<grid, kBlock, 0, L->ctx->stream>>>(
                                (vec2_t<S>*)U, L->g, mu, parity, beta, (int)n_update, nullptr, nullptr,
                                spread, seed, counter, L->mh_p, L->mh_acc, class_order(L)));
    else
        LQCD_DISPATCH_DS(L, k_metropolis<D, S, true, false><<<grid, kBlock, 0, L->ctx->stream>>>(
                                (vec2_t<S>*)U, L->g, mu, parity, beta, (int)n_update, nullptr, nullptr,
                                spread, seed, counter, L->mh_p, L->mh_acc, class_order(L)));
    prof_end(L->ctx, &ev);
    return hipGetLastError();
}

} // namespace lqcd
