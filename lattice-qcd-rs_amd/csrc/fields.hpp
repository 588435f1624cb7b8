// fields.hpp -- structure-of-arrays layout of the link and E fields in HBM.
//
// Links: plane (mu * 9 + e) holds element e = row * 3 + col of U_mu(x) for
// every site x, as a 2-vector (re, im) of the storage type S (8 B for f32,
// 16 B for f64).  E: plane (mu * 8 + a) holds e^a_mu(x) as a scalar S.  Every
// plane has g.npad entries, so a wavefront over 64 consecutive sites issues
// one fully coalesced 512 B (f32) / 1 KiB (f64) request per element.
#pragma once
#include "geometry.hpp"
#include "su3_device.hpp"

namespace lqcd {

template <typename S>
struct vec2;
template <>
struct vec2<float> {
    using type = float2;
};
template <>
struct vec2<double> {
    using type = double2;
};
template <typename S>
using vec2_t = typename vec2<S>::type;

template <typename T, typename S>
__device__ __forceinline__ mat3<T> load_link(const vec2_t<S>* __restrict__ U, int64_t npad, int mu,
                                             int64_t s)
{
    const vec2_t<S>* p = U + (int64_t)mu * 9 * npad + s;
    mat3<T> m;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const vec2_t<S> v = p[(int64_t)k * npad];
        m.e[k] = {T(v.x), T(v.y)};
    }
    return m;
}

template <typename T, typename S>
__device__ __forceinline__ void store_link(vec2_t<S>* __restrict__ U, int64_t npad, int mu, int64_t s,
                                           const mat3<T>& m)
{
    vec2_t<S>* p = U + (int64_t)mu * 9 * npad + s;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        vec2_t<S> v;
        v.x = S(m.e[k].re);
        v.y = S(m.e[k].im);
        p[(int64_t)k * npad] = v;
    }
}

template <typename T, typename S>
__device__ __forceinline__ void load_e(const S* __restrict__ E, int64_t npad, int mu, int64_t s, T* e)
{
    const S* p = E + (int64_t)mu * 8 * npad + s;
#pragma unroll
    for (int a = 0; a < 8; ++a) e[a] = T(p[(int64_t)a * npad]);
}

template <typename T, typename S>
__device__ __forceinline__ void store_e(S* __restrict__ E, int64_t npad, int mu, int64_t s, const T* e)
{
    S* p = E + (int64_t)mu * 8 * npad + s;
#pragma unroll
    for (int a = 0; a < 8; ++a) p[(int64_t)a * npad] = S(e[a]);
}

} // namespace lqcd
