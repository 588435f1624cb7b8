// fields.hpp -- structure-of-arrays layout of the link and E fields in HBM.
//
// Links: plane (mu * 9 + e) holds element e = row * 3 + col of U_mu(x) for
// every site x, as a 2-vector (re, im) of the storage type S (8 B for f32,
// 16 B for f64).  E: plane (mu * 8 + a) holds e^a_mu(x) as a scalar S.  Every
// plane has g.npad entries, so a wavefront over 64 consecutive sites issues
// one fully coalesced 512 B (f32) / 1 KiB (f64) request per element.
//
// Accesses are raw buffer loads / stores: the resource covers the planes of
// one direction (uniform, so it lives in scalar registers), the site enters
// as a 32-bit byte offset and the element plane as the scalar offset.  No
// per-lane 64-bit address arithmetic is left, and an access outside the
// direction's planes reads zero instead of faulting.  The planes of one
// direction must therefore span less than 4 GiB (kMaxPlaneBytes).
#pragma once
#include "geometry.hpp"
#include "su3_device.hpp"

namespace lqcd {

// descriptor word 3 of a raw buffer resource on the gfx9 family
constexpr int kBufferFlags = 0x00020000;

// cache-policy operand of a buffer access on gfx950: 0 is the default; bit 1
// (nt) marks the line non-temporal, so the L2 evicts it first.  Fields that a
// kernel reads or writes exactly once (E, the new links) use kCacheStream and
// leave the L2 to the neighbour links that other sites read again.
constexpr int kCacheDefault = 0;
constexpr int kCacheStream = 2;

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

// resource over `planes` planes of npad elements of type E starting at p
template <typename E>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t planes_rsrc(const E* p, int64_t npad, int planes)
{
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)((uint64_t)planes * npad * sizeof(E)),
                                             kBufferFlags);
}

template <typename E, int CP = kCacheDefault>
__device__ __forceinline__ E buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff)
{
    if constexpr (sizeof(E) == 4) {
        return __builtin_bit_cast(E, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, (int)soff, CP));
    } else if constexpr (sizeof(E) == 8) {
        return __builtin_bit_cast(E, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, (int)soff, CP));
    } else {
        static_assert(sizeof(E) == 16, "element of 4, 8 or 16 bytes");
        return __builtin_bit_cast(E, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, (int)soff, CP));
    }
}

template <typename E, int CP = kCacheDefault>
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff, E v)
{
    if constexpr (sizeof(E) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)off, (int)soff, CP);
    } else if constexpr (sizeof(E) == 8) {
        using V = __attribute__((ext_vector_type(2))) uint32_t;
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, v), r, (int)off, (int)soff, CP);
    } else {
        static_assert(sizeof(E) == 16, "element of 4, 8 or 16 bytes");
        using V = __attribute__((ext_vector_type(4))) uint32_t;
        // soffset 0, the plane offset in the VGPR offset: the compiler inserts
        // the wait state that a VALU write of the data VGPRs right after a
        // store of more than 64 bits needs only when soffset is not a
        // register, and without it such writes reach memory in place of the
        // stored data
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V, v), r, (int)(off + soff), 0, CP);
    }
}

template <typename T, typename S, int CP = kCacheDefault>
__device__ __forceinline__ mat3<T> load_link(const vec2_t<S>* __restrict__ U, int64_t npad, int mu,
                                             int64_t s)
{
    using E = vec2_t<S>;
    const __amdgpu_buffer_rsrc_t r = planes_rsrc(U + (int64_t)mu * 9 * npad, npad, 9);
    const uint32_t off = (uint32_t)s * (uint32_t)sizeof(E);
    const uint32_t plane = (uint32_t)npad * (uint32_t)sizeof(E);
    mat3<T> m;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const E v = buf_load<E, CP>(r, off, (uint32_t)k * plane);
        m.e[k] = {T(v.x), T(v.y)};
    }
    return m;
}

// The first two rows of U_mu(x) (planes 0-5); row 2 is not written.
template <typename T, typename S>
__device__ __forceinline__ mat3<T> load_link_rows01(const vec2_t<S>* __restrict__ U, int64_t npad, int mu,
                                                    int64_t s)
{
    using E = vec2_t<S>;
    const __amdgpu_buffer_rsrc_t r = planes_rsrc(U + (int64_t)mu * 9 * npad, npad, 9);
    const uint32_t off = (uint32_t)s * (uint32_t)sizeof(E);
    const uint32_t plane = (uint32_t)npad * (uint32_t)sizeof(E);
    mat3<T> m;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const E v = buf_load<E>(r, off, (uint32_t)k * plane);
        m.e[k] = {T(v.x), T(v.y)};
    }
    return m;
}

// U_mu(x) from its first two rows: for a matrix of SU(3) the third row is
// the complex conjugate of the cross product of the first two,
//   u_2c = conj(u_0,c+1 u_1,c+2 - u_0,c+2 u_1,c+1),
// so six of the nine planes are read.  Exact for SU(3); callers use it only on
// links checked to be in SU(3) to rounding.
template <typename T, typename S>
__device__ __forceinline__ mat3<T> load_link_r12(const vec2_t<S>* __restrict__ U, int64_t npad, int mu,
                                                 int64_t s)
{
    mat3<T> m = load_link_rows01<T, S>(U, npad, mu, s);
    complete_su3(m);
    return m;
}

// load_link, or load_link_r12 when R12
template <bool R12, typename T, typename S>
__device__ __forceinline__ mat3<T> load_link_sel(const vec2_t<S>* __restrict__ U, int64_t npad, int mu, int64_t s)
{
    if constexpr (R12)
        return load_link_r12<T, S>(U, npad, mu, s);
    else
        return load_link<T, S>(U, npad, mu, s);
}

template <typename T, typename S, int CP = kCacheDefault>
__device__ __forceinline__ void store_link(vec2_t<S>* __restrict__ U, int64_t npad, int mu, int64_t s,
                                           const mat3<T>& m)
{
    using E = vec2_t<S>;
    const __amdgpu_buffer_rsrc_t r = planes_rsrc(U + (int64_t)mu * 9 * npad, npad, 9);
    const uint32_t off = (uint32_t)s * (uint32_t)sizeof(E);
    const uint32_t plane = (uint32_t)npad * (uint32_t)sizeof(E);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        E v;
        v.x = S(m.e[k].re);
        v.y = S(m.e[k].im);
        buf_store<E, CP>(r, off, (uint32_t)k * plane, v);
    }
}

// Rows 0 and 1 of U_mu(x) only (planes 0-5): for links in SU(3) whose
// readers rebuild row 2 (load_link_r12) or which get it back from
// complete_row2_link before any other reader.
template <typename T, typename S, int CP = kCacheDefault>
__device__ __forceinline__ void store_link_rows01(vec2_t<S>* __restrict__ U, int64_t npad, int mu, int64_t s,
                                                  const mat3<T>& m)
{
    using E = vec2_t<S>;
    const __amdgpu_buffer_rsrc_t r = planes_rsrc(U + (int64_t)mu * 9 * npad, npad, 9);
    const uint32_t off = (uint32_t)s * (uint32_t)sizeof(E);
    const uint32_t plane = (uint32_t)npad * (uint32_t)sizeof(E);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        E v;
        v.x = S(m.e[k].re);
        v.y = S(m.e[k].im);
        buf_store<E, CP>(r, off, (uint32_t)k * plane, v);
    }
}

// Row 2 of U_mu(x) only (planes 6-8).
template <typename T, typename S, int CP = kCacheDefault>
__device__ __forceinline__ void store_link_row2(vec2_t<S>* __restrict__ U, int64_t npad, int mu, int64_t s,
                                                const mat3<T>& m)
{
    using E = vec2_t<S>;
    const __amdgpu_buffer_rsrc_t r = planes_rsrc(U + (int64_t)mu * 9 * npad, npad, 9);
    const uint32_t off = (uint32_t)s * (uint32_t)sizeof(E);
    const uint32_t plane = (uint32_t)npad * (uint32_t)sizeof(E);
#pragma unroll
    for (int k = 6; k < 9; ++k) {
        E v;
        v.x = S(m.e[k].re);
        v.y = S(m.e[k].im);
        buf_store<E, CP>(r, off, (uint32_t)k * plane, v);
    }
}

template <typename T, typename S, int CP = kCacheDefault>
__device__ __forceinline__ void load_e(const S* __restrict__ E, int64_t npad, int mu, int64_t s, T* e)
{
    const __amdgpu_buffer_rsrc_t r = planes_rsrc(E + (int64_t)mu * 8 * npad, npad, 8);
    const uint32_t off = (uint32_t)s * (uint32_t)sizeof(S);
    const uint32_t plane = (uint32_t)npad * (uint32_t)sizeof(S);
#pragma unroll
    for (int a = 0; a < 8; ++a) e[a] = T(buf_load<S, CP>(r, off, (uint32_t)a * plane));
}

template <typename T, typename S, int CP = kCacheDefault>
__device__ __forceinline__ void store_e(S* __restrict__ E, int64_t npad, int mu, int64_t s, const T* e)
{
    const __amdgpu_buffer_rsrc_t r = planes_rsrc(E + (int64_t)mu * 8 * npad, npad, 8);
    const uint32_t off = (uint32_t)s * (uint32_t)sizeof(S);
    const uint32_t plane = (uint32_t)npad * (uint32_t)sizeof(S);
#pragma unroll
    for (int a = 0; a < 8; ++a) buf_store<S, CP>(r, off, (uint32_t)a * plane, S(e[a]));
}

} // namespace lqcd
