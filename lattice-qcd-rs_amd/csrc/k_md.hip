// k_md.hip -- gauge force, momentum kick and link drift.
//
// One thread per link (x, mu): blockIdx.y = mu, 256 consecutive sites per
// block.  The force is derivative_e of src/simulation/state.rs:1420-1448,
//   de^a_mu(x) = -sqrt(2/3)/a * Im Tr(T^a U_mu(x) Sigma_mu(x)),
//   Sigma_mu(x) = sum_{nu != mu} [ U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
//                               + U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu) ],
// which is the sum of S_{mu,+-nu}(x)^+ over the signed directions of the
// reference (LinkMatrix::sij, src/field.rs:743-758).  The eight generator
// projections are taken in closed form from the nine entries of U Sigma.
//
// The fused kick+drift kernel reads U_in and writes U_out (double buffered),
// so one launch performs e += kick * F(U_in) followed by the drift of its own
// link with the updated e: one pass over HBM per leapfrog step.
#include "dispatch.hpp"
#include "fields.hpp"

namespace lqcd {

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

// F^a = fscale * Im Tr(T^a U Sigma); returns U in u for the drift.
template <int D, typename S, typename T>
__device__ __forceinline__ void force_link(const vec2_t<S>* __restrict__ U, const Geom& g, int mu,
                                           int64_t s, T fscale, mat3<T>& u, T* f)
{
    const Neighbours<D> nb = neighbours<D>(g, s);
    mat3<T> sigma, m;
    staple_sum<D, S, T>(U, g, mu, s, nb, sigma);
    u = load_link<T, S>(U, g.npad, mu, s);
    mul<false, false>(u, sigma, m);
    im_trace_generators(m, f);
#pragma unroll
    for (int a = 0; a < 8; ++a) f[a] *= fscale;
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
        for (int k = 0; k < 9; ++k) out.e[k] = {fma(-f, eu.e[k].im, u.e[k].re), fma(f, eu.e[k].re, u.e[k].im)};
    } else {
        T v[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) v[a] = f * e[a];
        mat3<T> r;
        exp_i(v, r);
        mul<false, false>(r, u, out);
    }
}

template <int D, typename S, typename T>
__global__ __launch_bounds__(kBlock) void k_force_store(const vec2_t<S>* __restrict__ U, Geom g,
                                                        T fscale, double* __restrict__ out, double oscale)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= g.nsite) return;
    mat3<T> u;
    T f[8];
    force_link<D, S, T>(U, g, mu, s, fscale, u, f);
    double* p = out + (s * D + mu) * 8;
#pragma unroll
    for (int a = 0; a < 8; ++a) p[a] = oscale * (double)f[a];
}

template <int D, typename S, typename T>
__global__ __launch_bounds__(kBlock) void k_kick(const vec2_t<S>* __restrict__ U, S* __restrict__ E,
                                                 Geom g, T fscale, T kick)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= g.nsite) return;
    mat3<T> u;
    T f[8], e[8];
    force_link<D, S, T>(U, g, mu, s, fscale, u, f);
    load_e<T, S>(E, g.npad, mu, s, e);
#pragma unroll
    for (int a = 0; a < 8; ++a) e[a] = fma(kick, f[a], e[a]);
    store_e<T, S>(E, g.npad, mu, s, e);
}

template <int D, typename S, typename T, int MODE>
__global__ __launch_bounds__(kBlock) void k_kick_drift(const vec2_t<S>* __restrict__ U_in,
                                                       vec2_t<S>* __restrict__ U_out,
                                                       S* __restrict__ E, Geom g, T fscale, T kick,
                                                       T dfac)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= g.nsite) return;
    mat3<T> u, un;
    T f[8], e[8];
    force_link<D, S, T>(U_in, g, mu, s, fscale, u, f);
    load_e<T, S>(E, g.npad, mu, s, e);
#pragma unroll
    for (int a = 0; a < 8; ++a) e[a] = fma(kick, f[a], e[a]);
    store_e<T, S>(E, g.npad, mu, s, e);
    drift_link<T, MODE>(u, e, dfac, un);
    store_link<T, S>(U_out, g.npad, mu, s, un);
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

hipError_t launch_force_store(const lqcd_lattice* L, const void* U, double* d_out, double scale)
{
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_FORCE_STORE, &ev);
    LQCD_DISPATCH_DS(L, k_force_store<D, S, S><<<link_grid(L), kBlock, 0, L->ctx->stream>>>(
                            (const vec2_t<S>*)U, L->g, (S)force_scale(L), d_out, scale));
    prof_end(L->ctx, &ev);
    return hipGetLastError();
}

hipError_t launch_kick(const lqcd_lattice* L, const void* U, void* E, double kick)
{
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_KICK, &ev);
    LQCD_DISPATCH_DS(L, k_kick<D, S, S><<<link_grid(L), kBlock, 0, L->ctx->stream>>>(
                            (const vec2_t<S>*)U, (S*)E, L->g, (S)force_scale(L), (S)kick));
    prof_end(L->ctx, &ev);
    return hipGetLastError();
}

hipError_t launch_kick_drift(const lqcd_lattice* L, const void* U_in, void* U_out, void* E, double kick,
                             double dt, int drift_mode)
{
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_KICK_DRIFT, &ev);
    if (drift_mode == DRIFT_EXP) {
        LQCD_DISPATCH_DS(L, k_kick_drift<D, S, S, DRIFT_EXP><<<link_grid(L), kBlock, 0, L->ctx->stream>>>(
                                (const vec2_t<S>*)U_in, (vec2_t<S>*)U_out, (S*)E, L->g,
                                (S)force_scale(L), (S)kick, (S)drift_scale(L, dt)));
    } else {
        LQCD_DISPATCH_DS(L, k_kick_drift<D, S, S, DRIFT_ADDITIVE>
                         <<<link_grid(L), kBlock, 0, L->ctx->stream>>>(
                             (const vec2_t<S>*)U_in, (vec2_t<S>*)U_out, (S*)E, L->g,
                             (S)force_scale(L), (S)kick, (S)drift_scale(L, dt)));
    }
    prof_end(L->ctx, &ev);
    return hipGetLastError();
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
