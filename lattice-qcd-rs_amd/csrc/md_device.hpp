// md_device.hpp -- gauge force, momentum kick and link drift of one site or
// one link, shared by the single-lattice kernels (k_md.hip) and the batched
// chains (k_batch.hip) so both run the same arithmetic.
//
// The force is derivative_e of src/simulation/state.rs:1420-1448,
//   de^a_mu(x) = -sqrt(2/3)/a * Im Tr(T^a U_mu(x) Sigma_mu(x)),
//   Sigma_mu(x) = sum_{nu != mu} [ U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
//                               + U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu) ],
// which is the sum of S_{mu,+-nu}(x)^+ over the signed directions of the
// reference (LinkMatrix::sij, src/field.rs:743-758).  The eight generator
// projections are taken in closed form from the nine entries of U Sigma.
#pragma once
#include "fields.hpp"
#include "internal.hpp"

namespace lqcd {

// Workgroups are dispatched round robin over the 8 XCDs (workgroup b runs on
// XCD b mod 8).  Give XCD k the contiguous logical range
// [k q + min(k, r), ...) of n = 8 q + r blocks.
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t n)
{
    const int64_t k = b & 7, i = b >> 3, q = n >> 3, r = n & 7;
    return k * q + (k < r ? k : r) + i;
}

// Site block of logical block j for the slab order: P blocks per slice of the
// slowest direction, G slices per XCD slab; within a slab the slice index runs
// fastest.
__device__ __forceinline__ int64_t slab_block(int64_t j, int64_t P, int64_t G)
{
    const int64_t q = j / (G * P), r = j % (G * P);
    return (q * G + r % G) * P + r / G;
}

template <int D, typename S, typename T, bool R12 = false>
__device__ __forceinline__ void staple_sum(const vec2_t<S>* __restrict__ U, const Geom& g, int mu,
                                           int64_t s, const Neighbours<D>& nb, mat3<T>& sigma)
{
    set_zero(sigma);
    const int64_t s_pmu = s + nb.up[mu];
    if constexpr (R12) {
        // links and staples in SU(3): products as rows 0, 1 (mul_rows01),
        // the staple's row 2 rebuilt into Sigma (add_su3)
#pragma unroll
        for (int nu = 0; nu < D; ++nu) {
            if (nu == mu) continue;
            {
                const mat3<T> a = load_link_rows01<T, S>(U, g.npad, nu, s_pmu);
                const mat3<T> b = load_link_r12<T, S>(U, g.npad, mu, s + nb.up[nu]);
                mat3<T> ab, st;
                mul_rows01<false, true>(a, b, ab);
                const mat3<T> c = load_link_r12<T, S>(U, g.npad, nu, s);
                mul_rows01<false, true>(ab, c, st);
                add_su3(st, sigma);
            }
            {
                const int64_t s_mnu = s + nb.dn[nu];
                const mat3<T> a = load_link_r12<T, S>(U, g.npad, nu, s_pmu + nb.dn[nu]);
                const mat3<T> b = load_link_r12<T, S>(U, g.npad, mu, s_mnu);
                mat3<T> ab, st;
                mul_rows01<true, true>(a, b, ab);
                const mat3<T> c = load_link_r12<T, S>(U, g.npad, nu, s_mnu);
                mul_rows01<false, false>(ab, c, st);
                add_su3(st, sigma);
            }
        }
        return;
    }
#pragma unroll
    for (int nu = 0; nu < D; ++nu) {
        if (nu == mu) continue;
        // upper staple: U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
        {
            const mat3<T> a = load_link_sel<R12, T, S>(U, g.npad, nu, s_pmu);
            const mat3<T> b = load_link_sel<R12, T, S>(U, g.npad, mu, s + nb.up[nu]);
            mat3<T> ab;
            mul<false, true>(a, b, ab);
            const mat3<T> c = load_link_sel<R12, T, S>(U, g.npad, nu, s);
            mul_add<false, true>(ab, c, sigma);
        }
        // lower staple: U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)
        {
            const int64_t s_mnu = s + nb.dn[nu];
            const mat3<T> a = load_link_sel<R12, T, S>(U, g.npad, nu, s_pmu + nb.dn[nu]);
            const mat3<T> b = load_link_sel<R12, T, S>(U, g.npad, mu, s_mnu);
            mat3<T> ab;
            mul<true, true>(a, b, ab);
            const mat3<T> c = load_link_sel<R12, T, S>(U, g.npad, nu, s_mnu);
            mul_add<false, false>(ab, c, sigma);
        }
    }
}

// The site's own D links, completed to nine entries once per launch and kept
// in LDS slots private to the thread ([mu * 9 + k][kBlock], consecutive lanes
// in consecutive words): each is used by D - 1 upper staples, the force and
// the drift, and re-reading it from LDS is cheaper than reloading and
// completing it.  No barrier: a thread reads only its own slots.
template <typename T>
__device__ __forceinline__ mat3<T> own_get(const vec2_t<T>* own, int mu)
{
    mat3<T> m;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const vec2_t<T> v = own[(mu * 9 + k) * kBlock + threadIdx.x];
        m.e[k] = {v.x, v.y};
    }
    return m;
}
template <typename T>
__device__ __forceinline__ void own_put(vec2_t<T>* own, int mu, const mat3<T>& m)
{
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        vec2_t<T> v;
        v.x = m.e[k].re;
        v.y = m.e[k].im;
        own[(mu * 9 + k) * kBlock + threadIdx.x] = v;
    }
}

// All D staple sums of site s.  Upper staples pairwise through
// W = U_nu(x+mu) U_mu(x+nu)^+:  Sigma_mu += W U_nu(x)^+,
// Sigma_nu += W^+ U_mu(x)^+.
template <int D, typename S, typename T, bool R12 = false, bool OWN = false>
__device__ __forceinline__ void staple_sums_site(const vec2_t<S>* __restrict__ U, const Geom& g,
                                                 int64_t s, const Neighbours<D>& nb,
                                                 mat3<T> (&sig)[D], const vec2_t<T>* own = nullptr)
{
#pragma unroll
    for (int mu = 0; mu < D; ++mu) set_zero(sig[mu]);
    if constexpr (R12) {
        // links, W and staples in SU(3): products as rows 0, 1 (mul_rows01);
        // W gets its row 2 rebuilt (it is used as W^+), the staples theirs
        // inside the accumulation (add_su3)
#pragma unroll
        for (int mu = 0; mu < D; ++mu)
#pragma unroll
            for (int nu = mu + 1; nu < D; ++nu) {
                mat3<T> w, st;
                {
                    const mat3<T> a = load_link_rows01<T, S>(U, g.npad, nu, s + nb.up[mu]);
                    const mat3<T> b = load_link_r12<T, S>(U, g.npad, mu, s + nb.up[nu]);
                    mul_rows01<false, true>(a, b, w);
                    complete_su3(w);
                }
                {
                    const mat3<T> c = OWN ? own_get<T>(own, nu) : load_link_r12<T, S>(U, g.npad, nu, s);
                    mul_rows01<false, true>(w, c, st);
                    add_su3(st, sig[mu]);
                }
                {
                    const mat3<T> c = OWN ? own_get<T>(own, mu) : load_link_r12<T, S>(U, g.npad, mu, s);
                    mul_rows01<true, true>(w, c, st);
                    add_su3(st, sig[nu]);
                }
            }
#pragma unroll
        for (int nu = 0; nu < D; ++nu) {
            const int64_t s_mnu = s + nb.dn[nu];
            const mat3<T> c = load_link_r12<T, S>(U, g.npad, nu, s_mnu);
#pragma unroll
            for (int mu = 0; mu < D; ++mu) {
                if (mu == nu) continue;
                const mat3<T> a = load_link_r12<T, S>(U, g.npad, nu, s_mnu + nb.up[mu]);
                const mat3<T> b = load_link_r12<T, S>(U, g.npad, mu, s_mnu);
                mat3<T> ab, st;
                mul_rows01<true, true>(a, b, ab);
                mul_rows01<false, false>(ab, c, st);
                add_su3(st, sig[mu]);
            }
        }
        return;
    }
#pragma unroll
    for (int mu = 0; mu < D; ++mu)
#pragma unroll
        for (int nu = mu + 1; nu < D; ++nu) {
            mat3<T> w;
            {
                const mat3<T> a = load_link_sel<R12, T, S>(U, g.npad, nu, s + nb.up[mu]);
                const mat3<T> b = load_link_sel<R12, T, S>(U, g.npad, mu, s + nb.up[nu]);
                mul<false, true>(a, b, w);
            }
            {
                const mat3<T> c = load_link_sel<R12, T, S>(U, g.npad, nu, s);
                mul_add<false, true>(w, c, sig[mu]);
            }
            {
                const mat3<T> c = load_link_sel<R12, T, S>(U, g.npad, mu, s);
                mul_add<true, true>(w, c, sig[nu]);
            }
        }
#pragma unroll
    for (int nu = 0; nu < D; ++nu) {
        const int64_t s_mnu = s + nb.dn[nu];
        const mat3<T> c = load_link_sel<R12, T, S>(U, g.npad, nu, s_mnu);
#pragma unroll
        for (int mu = 0; mu < D; ++mu) {
            if (mu == nu) continue;
            const mat3<T> a = load_link_sel<R12, T, S>(U, g.npad, nu, s_mnu + nb.up[mu]);
            const mat3<T> b = load_link_sel<R12, T, S>(U, g.npad, mu, s_mnu);
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
    load_e<T, S, kCacheStream>(E, g.npad, mu, s, e);
#pragma unroll
    for (int a = 0; a < 8; ++a) e[a] = fma(kick, fscale * f[a], e[a]);
    store_e<T, S, kCacheStream>(E, g.npad, mu, s, e);
    if (DR >= 0) {
        mat3<T> un;
        drift_link<T, DR>(u, e, dfac, un);
        store_link<T, S, kCacheStream>(U_out, g.npad, mu, s, un);
    }
}

// one link (mu, s): e += kick * F(U_in); U_out = drift(U_in, e) unless DR < 0.
// R12: links read as two rows (load_link_r12), for links in SU(3).
template <int D, typename S, typename T, int DR, bool R12 = false>
__device__ __forceinline__ void md_link(const vec2_t<S>* __restrict__ U_in, vec2_t<S>* __restrict__ U_out,
                                        S* __restrict__ E, const Geom& g, int mu, int64_t s, T fscale,
                                        T kick, T dfac)
{
    const Neighbours<D> nb = neighbours<D>(g, s);
    mat3<T> sigma;
    staple_sum<D, S, T, R12>(U_in, g, mu, s, nb, sigma);
    const mat3<T> u = load_link_sel<R12, T, S>(U_in, g.npad, mu, s);
    finish_link<DR, S, T>(u, sigma, U_out, E, g, mu, s, fscale, kick, dfac);
}

// all D links of site s, sharing the staple loads
template <int D, typename S, typename T, int DR, bool R12 = false, bool OWN = false>
__device__ __forceinline__ void md_site(const vec2_t<S>* __restrict__ U_in, vec2_t<S>* __restrict__ U_out,
                                        S* __restrict__ E, const Geom& g, int64_t s, T fscale, T kick,
                                        T dfac, vec2_t<T>* own = nullptr)
{
    const Neighbours<D> nb = neighbours<D>(g, s);
    // OWN (two-row reads, f32): the own links into the thread's LDS slots
    if constexpr (OWN) {
#pragma unroll
        for (int mu = 0; mu < D; ++mu) own_put<T>(own, mu, load_link_r12<T, S>(U_in, g.npad, mu, s));
    }
    // phase 1: the D forces (the staple sums die here)
    T f[D][8];
    {
        mat3<T> sig[D];
        staple_sums_site<D, S, T, R12, OWN>(U_in, g, s, nb, sig, own);
#pragma unroll
        for (int mu = 0; mu < D; ++mu) {
            const mat3<T> u = OWN ? own_get<T>(own, mu) : load_link_sel<R12, T, S>(U_in, g.npad, mu, s);
            mat3<T> m;
            mul<false, false>(u, sig[mu], m);
            im_trace_generators(m, f[mu]);
        }
    }
    // phase 2: momentum kicks of the D links
    T e[D][8];
#pragma unroll
    for (int mu = 0; mu < D; ++mu) {
        load_e<T, S, kCacheStream>(E, g.npad, mu, s, e[mu]);
#pragma unroll
        for (int a = 0; a < 8; ++a) e[mu][a] = fma(kick, fscale * f[mu][a], e[mu][a]);
        store_e<T, S, kCacheStream>(E, g.npad, mu, s, e[mu]);
    }
    // phase 3: drifts (U_mu(x) is read again, from cache)
    if (DR >= 0) {
#pragma unroll
        for (int mu = 0; mu < D; ++mu) {
            const mat3<T> u = OWN ? own_get<T>(own, mu) : load_link_sel<R12, T, S>(U_in, g.npad, mu, s);
            mat3<T> un;
            drift_link<T, DR>(u, e[mu], dfac, un);
            store_link<T, S, kCacheStream>(U_out, g.npad, mu, s, un);
        }
    }
}

// -sqrt(2 / CA) / a with CA = 3 (derivative_e, src/simulation/state.rs:1426)
__host__ __device__ inline double md_force_scale(double a) { return -sqrt(2.0 / 3.0) / a; }
// dt * sqrt(2 CA) / a (derivative_u, src/simulation/state.rs:1413-1416)
__host__ __device__ inline double md_drift_scale(double dt, double a) { return dt * sqrt(6.0) / a; }

} // namespace lqcd
