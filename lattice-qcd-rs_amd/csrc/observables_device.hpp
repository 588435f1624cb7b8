// observables_device.hpp -- per-site plaquette and E-field energy terms and
// the wave64 reduction, shared by the single-lattice reductions
// (k_observables.hip) and the batched chains (k_batch.hip).
#pragma once
#include "fields.hpp"
#include "internal.hpp"

namespace lqcd {

// Site s adds sum_{i<j} Re Tr P_ij(x), Im Tr P_ij(x) and (1 - Re Tr P_ij(x) / 3)
// to acc[0..2], with P_ij(x) = U_i(x) U_j(x+i) U_i(x+j)^+ U_j(x)^+
// (LinkMatrix::pij, src/field.rs:761-771) taken as
// Tr[(U_i(x) U_j(x+i)) (U_j(x) U_i(x+j))^+], in double arithmetic.
template <int D, typename S>
__device__ __forceinline__ void plaquette_site(const vec2_t<S>* __restrict__ U, const Geom& g, int64_t s,
                                               double* acc)
{
    const Neighbours<D> nb = neighbours<D>(g, s);
#pragma unroll
    for (int i = 0; i < D; ++i) {
#pragma unroll
        for (int j = i + 1; j < D; ++j) {
            const mat3<double> ui = load_link<double, S>(U, g.npad, i, s);
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
            // f32 storage: one plaquette at a time.  Without this fence the
            // scheduler issues the loads of all plaquettes up front and the
            // kernel drops to one wave per SIMD (measured 429 us against 174 us
            // per 32^4 launch); with f64 storage the fence costs more than the
            // occupancy gains (235 us against 331 us).
            if constexpr (sizeof(S) == 4) __builtin_amdgcn_sched_barrier(0);
        }
    }
}

// Su3Adjoint::trace_squared = sum_a e_a^2 / 2 of link (mu, s)
// (src/field.rs:164-167)
template <typename S>
__device__ __forceinline__ double efield_energy_link(const S* __restrict__ E, const Geom& g, int mu, int64_t s)
{
    double e[8];
    load_e<double, S>(E, g.npad, mu, s, e);
    double t = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) t += e[a] * e[a];
    return t / 2.0;
}

// distance of one link from SU(3): max(max_ij |(U U^+ - 1)_ij|, |det U - 1|)
// in double arithmetic; NaN entries give +inf
__device__ __forceinline__ double su3_distance_link(const mat3<double>& u)
{
    mat3<double> p;
    mul<false, true>(u, u, p);
    double dev = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double re = p.e[i * 3 + j].re - (i == j ? 1.0 : 0.0), im = p.e[i * 3 + j].im;
            const double d = sqrt(re * re + im * im);
            dev = d == d ? fmax(dev, d) : INFINITY;
        }
    const cplx<double> dt = det(u);
    const double d = sqrt((dt.re - 1.0) * (dt.re - 1.0) + dt.im * dt.im);
    return d == d ? fmax(dev, d) : INFINITY;
}

// butterfly max over the 64 lanes of a wave
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// butterfly sum over the 64 lanes of a wave; every lane gets the total
template <int K>
__device__ __forceinline__ void wave_sum(double (&v)[K])
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] += __shfl_xor(v[k], off, 64);
}

// sum of v over a kBlock-thread block, written to dst[0..K) by thread 0:
// wave sums first, then the waves in order
template <int K>
__device__ __forceinline__ void block_reduce_to(double (&v)[K], double* __restrict__ dst)
{
    wave_sum<K>(v);
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

} // namespace lqcd
