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

// ---- LDS tile of two-row links (VAR_SITE_TILE) ----
// A workgroup's kBlock sites are kTileRows = kBlock / L0 whole rows of x_0
// (L0 = extent[0]) at consecutive x_1 and fixed x_2.., and the tile holds the
// rows 0, 1 of all D links of those rows and of the rows x_1 - 1 and
// x_1 + kTileRows (periodic) next to them.  Slot of the site at tile row
// r in [-1, rows] and x_0 = c: (r + 1) L0 + c; layout [nu][entry 0..5][slot]
// (consecutive lanes read consecutive words).
struct TileShape {
    int L0;      // extent[0], divides kBlock
    int rows;    // kBlock / L0
    int slots;   // (rows + 2) * L0
};

template <typename T>
__device__ __forceinline__ mat3<T> tile_rows01(const vec2_t<T>* tile, const TileShape& ts, int nu, int slot)
{
    mat3<T> m;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const vec2_t<T> v = tile[(nu * 6 + k) * ts.slots + slot];
        m.e[k] = {v.x, v.y};
    }
    return m;
}

template <typename T>
__device__ __forceinline__ mat3<T> tile_r12(const vec2_t<T>* tile, const TileShape& ts, int nu, int slot)
{
    mat3<T> m = tile_rows01<T>(tile, ts, nu, slot);
    complete_su3(m);
    return m;
}

// slot of the neighbour (d0, d1) of the site at tile row r, x_0 = c
__device__ __forceinline__ int tile_slot(const TileShape& ts, int r, int c, int d0, int d1)
{
    int cc = c + d0;
    cc = cc < 0 ? cc + ts.L0 : (cc >= ts.L0 ? cc - ts.L0 : cc);
    return (r + d1 + 1) * ts.L0 + cc;
}

// staple_sums_site (two-row links, own links from the tile) with every link
// whose site differs from x in x_0 and x_1 only read from the tile and the
// others (a step in x_2.. ) from U.
template <int D, typename S, typename T>
__device__ __forceinline__ void staple_sums_site_tile(const vec2_t<S>* __restrict__ U, const Geom& g, int64_t s,
                                                      const Neighbours<D>& nb, const vec2_t<T>* tile,
                                                      const TileShape& ts, int r, int c, mat3<T> (&sig)[D])
{
    // offsets (d0, d1) of a unit step in direction d, and whether it stays in the tile
    auto in_tile = [](int d) { return d < 2; };
    auto dx0 = [](int d, int sign) { return d == 0 ? sign : 0; };
    auto dx1 = [](int d, int sign) { return d == 1 ? sign : 0; };
    const int own = tile_slot(ts, r, c, 0, 0);
#pragma unroll
    for (int mu = 0; mu < D; ++mu) set_zero(sig[mu]);
#pragma unroll
    for (int mu = 0; mu < D; ++mu)
#pragma unroll
        for (int nu = mu + 1; nu < D; ++nu) {
            mat3<T> w, st;
            {
                const mat3<T> a = in_tile(mu)
                                      ? tile_rows01<T>(tile, ts, nu, tile_slot(ts, r, c, dx0(mu, 1), dx1(mu, 1)))
                                      : load_link_rows01<T, S>(U, g.npad, nu, s + nb.up[mu]);
                const mat3<T> b = in_tile(nu)
                                      ? tile_r12<T>(tile, ts, mu, tile_slot(ts, r, c, dx0(nu, 1), dx1(nu, 1)))
                                      : load_link_r12<T, S>(U, g.npad, mu, s + nb.up[nu]);
                mul_rows01<false, true>(a, b, w);
                complete_su3(w);
            }
            {
                const mat3<T> cm = tile_r12<T>(tile, ts, nu, own);
                mul_rows01<false, true>(w, cm, st);
                add_su3(st, sig[mu]);
            }
            {
                const mat3<T> cm = tile_r12<T>(tile, ts, mu, own);
                mul_rows01<true, true>(w, cm, st);
                add_su3(st, sig[nu]);
            }
        }
#pragma unroll
    for (int nu = 0; nu < D; ++nu) {
        const int64_t s_mnu = s + nb.dn[nu];
        const int mslot = tile_slot(ts, r, c, dx0(nu, -1), dx1(nu, -1));
        const mat3<T> cm = in_tile(nu) ? tile_r12<T>(tile, ts, nu, mslot) : load_link_r12<T, S>(U, g.npad, nu, s_mnu);
#pragma unroll
        for (int mu = 0; mu < D; ++mu) {
            if (mu == nu) continue;
            const mat3<T> a =
                (in_tile(nu) && in_tile(mu))
                    ? tile_r12<T>(tile, ts, nu, tile_slot(ts, r, c, dx0(nu, -1) + dx0(mu, 1), dx1(nu, -1) + dx1(mu, 1)))
                    : load_link_r12<T, S>(U, g.npad, nu, s_mnu + nb.up[mu]);
            const mat3<T> b = in_tile(nu) ? tile_r12<T>(tile, ts, mu, mslot) : load_link_r12<T, S>(U, g.npad, mu, s_mnu);
            mat3<T> ab, st;
            mul_rows01<true, true>(a, b, ab);
            mul_rows01<false, false>(ab, cm, st);
            add_su3(st, sig[mu]);
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

// the exp drift of one link as rows 0 and 1 only (the update stays in SU(3)):
// exp(iX) U with the product formed for rows 0, 1
template <typename T>
__device__ __forceinline__ void drift_link_rows01(const mat3<T>& u, const T* e, T f, mat3<T>& out)
{
    T v[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) v[a] = f * e[a];
    mat3<T> r;
    exp_i(v, r);
    mul_rows01<false, false>(r, u, out);
}

// drift of link (mu, s) into U_out: all nine entries, or with R01 (exp
// drift of SU(3) links) rows 0, 1 only
template <int DR, bool R01, typename S, typename T>
__device__ __forceinline__ void drift_store(const mat3<T>& u, const T* e, T dfac, vec2_t<S>* __restrict__ U_out,
                                            const Geom& g, int mu, int64_t s)
{
    mat3<T> un;
    if constexpr (R01) {
        static_assert(DR == DRIFT_EXP, "rows-only stores need the SU(3)-preserving exp drift");
        drift_link_rows01<T>(u, e, dfac, un);
        store_link_rows01<T, S, kCacheStream>(U_out, g.npad, mu, s, un);
    } else {
        drift_link<T, DR>(u, e, dfac, un);
        store_link<T, S, kCacheStream>(U_out, g.npad, mu, s, un);
    }
}

// what happens with the force of one link (DR: -1 kick only, else drift mode)
template <int DR, typename S, typename T, bool R01 = false>
__device__ __forceinline__ void finish_link(const mat3<T>& u, const mat3<T>& sigma,
                                            vec2_t<S>* __restrict__ U_out, S* __restrict__ E,
                                            const Geom& g, int mu, int64_t s, T fscale, T kick,
                                            T dfac)
{
    mat3<T> m;
    mul<false, false>(u, sigma, m);
    T f[8], e[8];
    im_trace_combinations(m, f);
    load_e<T, S, kCacheStream>(E, g.npad, mu, s, e);
    kick_generators(f, kick * fscale, e);
    store_e<T, S, kCacheStream>(E, g.npad, mu, s, e);
    if constexpr (DR >= 0) drift_store<DR, R01, S, T>(u, e, dfac, U_out, g, mu, s);
}

// one link (mu, s): e += kick * F(U_in); U_out = drift(U_in, e) unless DR < 0.
// R12: links read as two rows (load_link_r12), for links in SU(3); R01: the
// drifted link is stored as rows 0, 1 only.
template <int D, typename S, typename T, int DR, bool R12 = false, bool R01 = false>
__device__ __forceinline__ void md_link(const vec2_t<S>* __restrict__ U_in, vec2_t<S>* __restrict__ U_out,
                                        S* __restrict__ E, const Geom& g, int mu, int64_t s, T fscale,
                                        T kick, T dfac)
{
    const Neighbours<D> nb = neighbours<D>(g, s);
    mat3<T> sigma;
    staple_sum<D, S, T, R12>(U_in, g, mu, s, nb, sigma);
    const mat3<T> u = load_link_sel<R12, T, S>(U_in, g.npad, mu, s);
    finish_link<DR, S, T, R01>(u, sigma, U_out, E, g, mu, s, fscale, kick, dfac);
}

// The Hamiltonian sums a fused MD launch can produce on the way (SUMS): the
// trace sum sum_mu Re Tr(U_mu(x) Sigma_mu(x)) of the links it reads (every
// plaquette enters it through each of its 4 links, so it is 4 sum_P Re Tr P)
// and the momentum sum sum_mu,a e^a_mu(x)^2 of E before the kick
// (SUMS_BEFORE_KICK: the first launch of a trajectory, H_old) or after it
// (SUMS_AFTER_KICK: the final kick, H_new); accumulated per thread in double.
enum MdSums { SUMS_NONE = 0, SUMS_BEFORE_KICK = 1, SUMS_AFTER_KICK = 2 };
template <typename T>
struct MdAcc {
    T tr = T(0);
    T e2 = T(0);
};

template <typename T>
__device__ __forceinline__ T re_trace(const mat3<T>& m)
{
    return m.e[0].re + m.e[4].re + m.e[8].re;
}

template <typename T>
__device__ __forceinline__ T e_squares8(const T* e)
{
    T q = T(0);
#pragma unroll
    for (int a = 0; a < 8; ++a) q = fma(e[a], e[a], q);
    return q;
}

// all D links of site s, sharing the staple loads.  SUMS: add to acc (see
// MdSums).  row2 (two-row reads only): row 2 of the site's links, completed
// here anyway, is written to row2 (the buffer the links were read from, so a
// final kick leaves complete links behind).
template <int D, typename S, typename T, int DR, bool R12 = false, bool OWN = false, bool R01 = false,
          int SUMS = SUMS_NONE>
__device__ __forceinline__ void md_site(const vec2_t<S>* __restrict__ U_in, vec2_t<S>* __restrict__ U_out,
                                        S* __restrict__ E, const Geom& g, int64_t s, T fscale, T kick,
                                        T dfac, vec2_t<T>* own = nullptr, MdAcc<T>* acc = nullptr,
                                        vec2_t<S>* row2 = nullptr)
{
    const Neighbours<D> nb = neighbours<D>(g, s);
    // OWN (two-row reads, f32): the own links into the thread's LDS slots
    if constexpr (OWN) {
#pragma unroll
        for (int mu = 0; mu < D; ++mu) own_put<T>(own, mu, load_link_r12<T, S>(U_in, g.npad, mu, s));
    }
    // phase 1: the D forces (the staple sums die here), as the unscaled
    // generator combinations: the kick folds in their factors
    T f[D][8];
    T e[D][8];
    {
        mat3<T> sig[D];
        staple_sums_site<D, S, T, R12, OWN>(U_in, g, s, nb, sig, own);
        // the momenta of the D links: their loads run under the D products U Sigma
#pragma unroll
        for (int mu = 0; mu < D; ++mu) load_e<T, S, kCacheStream>(E, g.npad, mu, s, e[mu]);
#pragma unroll
        for (int mu = 0; mu < D; ++mu) {
            const mat3<T> u = OWN ? own_get<T>(own, mu) : load_link_sel<R12, T, S>(U_in, g.npad, mu, s);

            mat3<T> m;
            mul<false, false>(u, sig[mu], m);
            im_trace_combinations(m, f[mu]);
            if constexpr (SUMS != SUMS_NONE) acc->tr += re_trace(m);
        }
    }
    // phase 2: momentum kicks of the D links
#pragma unroll
    for (int mu = 0; mu < D; ++mu) {
        if constexpr (SUMS == SUMS_BEFORE_KICK) acc->e2 += e_squares8<T>(e[mu]);
        kick_generators(f[mu], kick * fscale, e[mu]);
        if constexpr (SUMS == SUMS_AFTER_KICK) acc->e2 += e_squares8<T>(e[mu]);
        store_e<T, S, kCacheStream>(E, g.npad, mu, s, e[mu]);
    }
    if constexpr (R12) {
        if (row2) {
#pragma unroll
            for (int mu = 0; mu < D; ++mu) {
                const mat3<T> u = OWN ? own_get<T>(own, mu) : load_link_r12<T, S>(U_in, g.npad, mu, s);
                store_link_row2<T, S>(row2, g.npad, mu, s, u);
            }
        }
    }
    // phase 3: drifts (U_mu(x) is read again, from cache)
    if constexpr (DR >= 0) {
#pragma unroll
        for (int mu = 0; mu < D; ++mu) {
            const mat3<T> u = OWN ? own_get<T>(own, mu) : load_link_sel<R12, T, S>(U_in, g.npad, mu, s);
            drift_store<DR, R01, S, T>(u, e[mu], dfac, U_out, g, mu, s);
        }
    }
}

// -sqrt(2 / CA) / a with CA = 3 (derivative_e, src/simulation/state.rs:1426)
__host__ __device__ inline double md_force_scale(double a) { return -sqrt(2.0 / 3.0) / a; }
// dt * sqrt(2 CA) / a (derivative_u, src/simulation/state.rs:1413-1416)
__host__ __device__ inline double md_drift_scale(double dt, double a) { return dt * sqrt(6.0) / a; }

} // namespace lqcd
