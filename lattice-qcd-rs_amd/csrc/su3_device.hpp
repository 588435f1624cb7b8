// su3_device.hpp -- SU(3)/su(3) arithmetic for gfx950 kernels.
//
// Matrices are row-major 3x3 complex held in registers (e[r*3+c]).  Products
// take compile-time adjoint flags so a^dagger * b^dagger etc. cost the same as
// a * b: the conjugations fold into the sign modifiers of v_fma.  Everything
// is templated on the arithmetic type T (float or double).
#pragma once
#include <hip/hip_runtime.h>

namespace lqcd {

template <typename T>
struct cplx {
    T re, im;
};

template <typename T>
struct mat3 {
    cplx<T> e[9];
    __device__ __forceinline__ cplx<T>& operator()(int r, int c) { return e[r * 3 + c]; }
    __device__ __forceinline__ const cplx<T>& operator()(int r, int c) const { return e[r * 3 + c]; }
};

template <typename T>
__device__ __forceinline__ void set_zero(mat3<T>& a)
{
#pragma unroll
    for (int k = 0; k < 9; ++k) a.e[k] = {T(0), T(0)};
}

template <typename T>
__device__ __forceinline__ void set_identity(mat3<T>& a)
{
    set_zero(a);
    a.e[0].re = T(1);
    a.e[4].re = T(1);
    a.e[8].re = T(1);
}

// element (i, k) of a or of a^dagger
template <bool DAG, typename T>
__device__ __forceinline__ cplx<T> elem(const mat3<T>& a, int i, int k)
{
    if (DAG) {
        const cplx<T> z = a.e[k * 3 + i];
        return {z.re, -z.im};
    }
    return a.e[i * 3 + k];
}

// c = op(a) * op(b); c may not alias a or b.
// Accumulation order: the terms k != i first, then k == i with its real
// product last.  For near-identity factors the diagonal entry is then one
// rounding of 1 - e1 - e2 + e1 e2 + (small terms); rounding the leading
// product on its own would drop the positive e1 e2 every time (the rest is
// exactly representable), a systematic shrinking of f32 links under repeated
// link updates.
template <bool DA, bool DB, typename T>
__device__ __forceinline__ void mul(const mat3<T>& a, const mat3<T>& b, mat3<T>& c)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            T re = T(0), im = T(0);
#pragma unroll
            for (int kk = 1; kk < 3; ++kk) {
                const int k = (i + kk) % 3;
                const cplx<T> x = elem<DA>(a, i, k);
                const cplx<T> y = elem<DB>(b, k, j);
                re = fma(x.re, y.re, re);
                re = fma(-x.im, y.im, re);
                im = fma(x.re, y.im, im);
                im = fma(x.im, y.re, im);
            }
            const cplx<T> x = elem<DA>(a, i, i);
            const cplx<T> y = elem<DB>(b, i, j);
            re = fma(-x.im, y.im, re);
            im = fma(x.re, y.im, im);
            im = fma(x.im, y.re, im);
            re = fma(x.re, y.re, re);
            c.e[i * 3 + j] = {re, im};
        }
}

// rows 0 and 1 of c = op(a) * op(b), in the accumulation order of mul; row 2
// of c is not written.  Needs rows 0, 1 of op(a) and all of op(b).
template <bool DA, bool DB, typename T>
__device__ __forceinline__ void mul_rows01(const mat3<T>& a, const mat3<T>& b, mat3<T>& c)
{
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            T re = T(0), im = T(0);
#pragma unroll
            for (int kk = 1; kk < 3; ++kk) {
                const int k = (i + kk) % 3;
                const cplx<T> x = elem<DA>(a, i, k);
                const cplx<T> y = elem<DB>(b, k, j);
                re = fma(x.re, y.re, re);
                re = fma(-x.im, y.im, re);
                im = fma(x.re, y.im, im);
                im = fma(x.im, y.re, im);
            }
            const cplx<T> x = elem<DA>(a, i, i);
            const cplx<T> y = elem<DB>(b, i, j);
            re = fma(-x.im, y.im, re);
            im = fma(x.re, y.im, im);
            im = fma(x.im, y.re, im);
            re = fma(x.re, y.re, re);
            c.e[i * 3 + j] = {re, im};
        }
}

// Entry (2, c) of a matrix of SU(3) from its rows 0 and 1:
// conj(m_0,c+1 m_1,c+2 - m_0,c+2 m_1,c+1), added to acc, with one rounding
// per part and the leading product added last.
template <typename T>
__device__ __forceinline__ cplx<T> su3_row2_add(const mat3<T>& m, int c, cplx<T> acc)
{
    const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
    const cplx<T> a = m.e[c1], b = m.e[3 + c2], d = m.e[c2], f = m.e[3 + c1];
    const T re = fma(a.re, b.re, fma(-a.im, b.im, fma(-d.re, f.re, fma(d.im, f.im, acc.re))));
    const T nim = fma(d.re, f.im, fma(d.im, f.re, fma(-a.re, b.im, fma(-a.im, b.re, acc.im))));
    return {re, nim};
}

// row 2 of m in SU(3) from its rows 0 and 1
template <typename T>
__device__ __forceinline__ void complete_su3(mat3<T>& m)
{
    const cplx<T> z = {T(0), T(0)};
    const cplx<T> r0 = su3_row2_add(m, 0, z), r1 = su3_row2_add(m, 1, z), r2 = su3_row2_add(m, 2, z);
    m.e[6] = r0;
    m.e[7] = r1;
    m.e[8] = r2;
}

// acc += m for m in SU(3) given by its rows 0 and 1
template <typename T>
__device__ __forceinline__ void add_su3(const mat3<T>& m, mat3<T>& acc)
{
#pragma unroll
    for (int k = 0; k < 6; ++k) acc.e[k] = {acc.e[k].re + m.e[k].re, acc.e[k].im + m.e[k].im};
#pragma unroll
    for (int c = 0; c < 3; ++c) acc.e[6 + c] = su3_row2_add(m, c, acc.e[6 + c]);
}

// c += op(a) * op(b)
template <bool DA, bool DB, typename T>
__device__ __forceinline__ void mul_add(const mat3<T>& a, const mat3<T>& b, mat3<T>& c)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            T re = c.e[i * 3 + j].re, im = c.e[i * 3 + j].im;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const cplx<T> x = elem<DA>(a, i, k);
                const cplx<T> y = elem<DB>(b, k, j);
                re = fma(x.re, y.re, re);
                re = fma(-x.im, y.im, re);
                im = fma(x.re, y.im, im);
                im = fma(x.im, y.re, im);
            }
            c.e[i * 3 + j] = {re, im};
        }
}

// Re Tr(op(a) * op(b)) without forming the product
template <bool DA, bool DB, typename T>
__device__ __forceinline__ cplx<T> trace_mul(const mat3<T>& a, const mat3<T>& b)
{
    T re = T(0), im = T(0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const cplx<T> x = elem<DA>(a, i, k);
            const cplx<T> y = elem<DB>(b, k, i);
            re = fma(x.re, y.re, re);
            re = fma(-x.im, y.im, re);
            im = fma(x.re, y.im, im);
            im = fma(x.im, y.re, im);
        }
    return {re, im};
}

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b)
{
    return {fma(a.re, b.re, -a.im * b.im), fma(a.re, b.im, a.im * b.re)};
}

template <typename T>
__device__ __forceinline__ cplx<T> det(const mat3<T>& m)
{
    const cplx<T> a = cmul(m.e[4], m.e[8]), b = cmul(m.e[7], m.e[5]);
    const cplx<T> c = cmul(m.e[3], m.e[8]), d = cmul(m.e[6], m.e[5]);
    const cplx<T> f = cmul(m.e[3], m.e[7]), g = cmul(m.e[6], m.e[4]);
    const cplx<T> m0 = {a.re - b.re, a.im - b.im};
    const cplx<T> m1 = {c.re - d.re, c.im - d.im};
    const cplx<T> m2 = {f.re - g.re, f.im - g.im};
    const cplx<T> t0 = cmul(m.e[0], m0), t1 = cmul(m.e[1], m1), t2 = cmul(m.e[2], m2);
    return {t0.re - t1.re + t2.re, t0.im - t1.im + t2.im};
}

// su(3) element sum_a v_a T^a (T^a = Gell-Mann / 2) as a hermitian matrix
template <typename T>
__device__ __forceinline__ void adjoint_to_matrix(const T* v, mat3<T>& m)
{
    const T h = T(0.5);
    const T r3 = T(0.28867513459481288225); // 1 / (2 sqrt 3)
    m.e[0] = {h * v[2] + r3 * v[7], T(0)};
    m.e[4] = {-h * v[2] + r3 * v[7], T(0)};
    m.e[8] = {T(-2) * r3 * v[7], T(0)};
    m.e[1] = {h * v[0], -h * v[1]};
    m.e[3] = {h * v[0], h * v[1]};
    m.e[2] = {h * v[3], -h * v[4]};
    m.e[6] = {h * v[3], h * v[4]};
    m.e[5] = {h * v[5], -h * v[6]};
    m.e[7] = {h * v[5], h * v[6]};
}

// out_a = Im Tr(T^a M), closed form of the eight generator projections
template <typename T>
__device__ __forceinline__ void im_trace_generators(const mat3<T>& M, T* out)
{
    const T h = T(0.5);
    const T r3 = T(0.28867513459481288225);
    out[0] = h * (M.e[1].im + M.e[3].im);
    out[1] = h * (M.e[1].re - M.e[3].re);
    out[2] = h * (M.e[0].im - M.e[4].im);
    out[3] = h * (M.e[2].im + M.e[6].im);
    out[4] = h * (M.e[2].re - M.e[6].re);
    out[5] = h * (M.e[5].im + M.e[7].im);
    out[6] = h * (M.e[5].re - M.e[7].re);
    out[7] = r3 * (M.e[0].im + M.e[4].im - T(2) * M.e[8].im);
}

// The combinations whose multiples are the eight Im Tr(T^a M): out_a / 2 for
// a < 7 and out_7 / (2 sqrt 3) are im_trace_generators(M); callers that scale
// the projections anyway (the momentum kick) fold those factors into their
// scale (kick_generators).
template <typename T>
__device__ __forceinline__ void im_trace_combinations(const mat3<T>& M, T* out)
{
    out[0] = M.e[1].im + M.e[3].im;
    out[1] = M.e[1].re - M.e[3].re;
    out[2] = M.e[0].im - M.e[4].im;
    out[3] = M.e[2].im + M.e[6].im;
    out[4] = M.e[2].re - M.e[6].re;
    out[5] = M.e[5].im + M.e[7].im;
    out[6] = M.e[5].re - M.e[7].re;
    out[7] = M.e[0].im + M.e[4].im - T(2) * M.e[8].im;
}

// e_a += k Im Tr(T^a M) from the combinations c of im_trace_combinations
template <typename T>
__device__ __forceinline__ void kick_generators(const T* c, T k, T* e)
{
    const T kh = k * T(0.5), kr = k * T(0.28867513459481288225);
#pragma unroll
    for (int a = 0; a < 7; ++a) e[a] = fma(kh, c[a], e[a]);
    e[7] = fma(kr, c[7], e[7]);
}

// 1/k! for k = 0..25
__device__ __constant__ const double kInvFactorial[26] = {
    1.0,
    1.0,
    0.5,
    0.16666666666666666,
    0.041666666666666664,
    0.008333333333333333,
    0.001388888888888889,
    0.0001984126984126984,
    2.48015873015873e-05,
    2.7557319223985893e-06,
    2.755731922398589e-07,
    2.505210838544172e-08,
    2.08767569878681e-09,
    1.6059043836821613e-10,
    1.1470745597729725e-11,
    7.647163731819816e-13,
    4.779477332387385e-14,
    2.8114572543455206e-15,
    1.5619206968586225e-16,
    8.22063524662433e-18,
    4.110317623312165e-19,
    1.9572941063391263e-20,
    8.896791392450574e-22,
    3.8681701706306835e-23,
    1.6117375710961184e-24,
    6.446950284384474e-26,
};

// X^2 of a hermitian X (real diagonal a, b, c; X_01 = u, X_02 = v,
// X_12 = w): hermitian, so three real diagonal and three complex upper
// entries are computed and the lower ones conjugated.
template <typename T>
__device__ __forceinline__ void hermitian_square(const mat3<T>& X, mat3<T>& Y)
{
    const T a = X.e[0].re, b = X.e[4].re, c = X.e[8].re;
    const cplx<T> u = X.e[1], v = X.e[2], w = X.e[5];
    const T uu = fma(u.re, u.re, u.im * u.im);
    const T vv = fma(v.re, v.re, v.im * v.im);
    const T ww = fma(w.re, w.re, w.im * w.im);
    Y.e[0] = {fma(a, a, uu + vv), T(0)};
    Y.e[4] = {fma(b, b, uu + ww), T(0)};
    Y.e[8] = {fma(c, c, vv + ww), T(0)};
    const T ab = a + b, ac = a + c, bc = b + c;
    // Y01 = (a + b) u + v conj(w)
    Y.e[1] = {fma(ab, u.re, fma(v.re, w.re, v.im * w.im)), fma(ab, u.im, fma(v.im, w.re, -v.re * w.im))};
    // Y02 = (a + c) v + u w
    Y.e[2] = {fma(ac, v.re, fma(u.re, w.re, -u.im * w.im)), fma(ac, v.im, fma(u.re, w.im, u.im * w.re))};
    // Y12 = (b + c) w + conj(u) v
    Y.e[5] = {fma(bc, w.re, fma(u.re, v.re, u.im * v.im)), fma(bc, w.im, fma(u.re, v.im, -u.im * v.re))};
    Y.e[3] = {Y.e[1].re, -Y.e[1].im};
    Y.e[6] = {Y.e[2].re, -Y.e[2].im};
    Y.e[7] = {Y.e[5].re, -Y.e[5].im};
}

// det X of a hermitian X (real):
// abc - a|w|^2 - b|v|^2 - c|u|^2 + 2 Re(u w conj(v))
template <typename T>
__device__ __forceinline__ T hermitian_det(const mat3<T>& X)
{
    const T a = X.e[0].re, b = X.e[4].re, c = X.e[8].re;
    const cplx<T> u = X.e[1], v = X.e[2], w = X.e[5];
    const T uw_re = fma(u.re, w.re, -u.im * w.im);
    const T uw_im = fma(u.re, w.im, u.im * w.re);
    const T cyc = fma(uw_re, v.re, uw_im * v.im);
    T r = a * fma(b, c, -fma(w.re, w.re, w.im * w.im));
    r = fma(-b, fma(v.re, v.re, v.im * v.im), r);
    r = fma(-c, fma(u.re, u.re, u.im * u.im), r);
    return fma(T(2), cyc, r);
}

// Cayley-Hamilton recursion of su3_exp_i (src/su3.rs:832-855) started at
// order K: q0 = 1/K!, then q0 <- 1/k! + d q2, q1 <- i (q0 - t q2),
// q2 <- i q1 for k = K-1 .. 0 (the Taylor series to order K reduced by
// X^3 = -t X - i d).  For hermitian X, d = i det X is imaginary: d = i delta.
// Fully unrolled.
template <int K, typename T>
__device__ __forceinline__ void ch_recursion(T delta, T t, cplx<T>& q0, cplx<T>& q1, cplx<T>& q2)
{
    q0 = {T(kInvFactorial[K]), T(0)};
    q1 = {T(0), T(0)};
    q2 = {T(0), T(0)};
#pragma unroll
    for (int k = K - 1; k >= 0; --k) {
        const cplx<T> dq2 = {-delta * q2.im, delta * q2.re};
        const cplx<T> q0n = {T(kInvFactorial[k]) + dq2.re, dq2.im};
        const cplx<T> w = {fma(-t, q2.re, q0.re), fma(-t, q2.im, q0.im)};
        const cplx<T> q1n = {-w.im, w.re};
        const cplx<T> q2n = {-q1.im, q1.re};
        q0 = q0n;
        q1 = q1n;
        q2 = q2n;
    }
}

// exp(i X) for hermitian traceless X given by its adjoint coordinates v
// (X = sum v_a T^a): with d = i det X and t = -Tr(X^2)/2,
// exp(iX) = q0 + q1 X + q2 X^2 (the recursion of su3_exp_i,
// src/su3.rs:832-855).  The reference always sums to order 25.  Here the
// series stops at K = 12 (f64) / 7 (f32), which is within the working
// precision while r = |X|_F (a bound of the spectral radius) is at most
// 0.25 / 0.35: the remainder r^(K+1)/(K+1)! e^r is below 1e-17 / 1e-8.
// Larger X are scaled by 2^-m into that range and the result squared m
// times (m = 0 for molecular-dynamics step sizes).
template <typename T>
__device__ __forceinline__ void exp_i(const T* v_in, mat3<T>& out)
{
    constexpr bool f32 = sizeof(T) == 4;
    constexpr int K = f32 ? 7 : 12;
    const T max_norm2 = f32 ? T(0.35 * 0.35) : T(0.25 * 0.25);
    T ss = T(0);
#pragma unroll
    for (int a = 0; a < 8; ++a) ss = fma(v_in[a], v_in[a], ss);
    int m = 0;
    T scale = T(1);
    while (T(0.5) * ss * scale * scale > max_norm2 && m < 40) {
        scale *= T(0.5);
        ++m;
    }
    T v[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) v[a] = v_in[a];
    // scaling is exact (a power of two) and rare at molecular-dynamics step
    // sizes: lanes without it skip the multiplications
    if (m > 0) {
#pragma unroll
        for (int a = 0; a < 8; ++a) v[a] *= scale;
        ss *= scale * scale;
    }
    mat3<T> X, X2;
    adjoint_to_matrix(v, X);
    hermitian_square(X, X2);
    const T t = T(-0.25) * ss;
    cplx<T> q0, q1, q2;
    ch_recursion<K>(hermitian_det(X), t, q0, q1, q2);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const cplx<T> x = X.e[k], y = X2.e[k];
        T re = fma(q1.re, x.re, -q1.im * x.im);
        T im = fma(q1.re, x.im, q1.im * x.re);
        re = fma(q2.re, y.re, re);
        re = fma(-q2.im, y.im, re);
        im = fma(q2.re, y.im, im);
        im = fma(q2.im, y.re, im);
        out.e[k] = {re, im};
    }
    out.e[0].re += q0.re; out.e[0].im += q0.im;
    out.e[4].re += q0.re; out.e[4].im += q0.im;
    out.e[8].re += q0.re; out.e[8].im += q0.im;
    for (int i = 0; i < m; ++i) {
        mat3<T> sq;
        mul<false, false>(out, out, sq);
        out = sq;
    }
}

// column Gram-Schmidt of orthonormalize_matrix, src/su3.rs:279-303:
// c0 /= |c0| (unless |c0| <= eps), c1 -= c0 (c0^dagger c1), c1 /= |c1|
// (same rule), c2 = conj(c0) x conj(c1).  Double arithmetic only.
__device__ __forceinline__ void orthonormalize(mat3<double>& m)
{
    const double eps = 2.220446049250313e-16;
    cplx<double> a[3], b[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        a[i] = m.e[i * 3 + 0];
        b[i] = m.e[i * 3 + 1];
    }
    double n0 = sqrt(a[0].re * a[0].re + a[0].im * a[0].im + a[1].re * a[1].re +
                     a[1].im * a[1].im + a[2].re * a[2].re + a[2].im * a[2].im);
    if (n0 > eps) {
#pragma unroll
        for (int i = 0; i < 3; ++i) a[i] = {a[i].re / n0, a[i].im / n0};
    }
    // s = sum conj(a_i) b_i
    cplx<double> s = {0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        s.re += a[i].re * b[i].re + a[i].im * b[i].im;
        s.im += a[i].re * b[i].im - a[i].im * b[i].re;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const cplx<double> p = cmul(a[i], s);
        b[i] = {b[i].re - p.re, b[i].im - p.im};
    }
    double n1 = sqrt(b[0].re * b[0].re + b[0].im * b[0].im + b[1].re * b[1].re +
                     b[1].im * b[1].im + b[2].re * b[2].re + b[2].im * b[2].im);
    if (n1 > eps) {
#pragma unroll
        for (int i = 0; i < 3; ++i) b[i] = {b[i].re / n1, b[i].im / n1};
    }
    const cplx<double> ca[3] = {{a[0].re, -a[0].im}, {a[1].re, -a[1].im}, {a[2].re, -a[2].im}};
    const cplx<double> cb[3] = {{b[0].re, -b[0].im}, {b[1].re, -b[1].im}, {b[2].re, -b[2].im}};
    cplx<double> c[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int j = (i + 1) % 3, k = (i + 2) % 3;
        const cplx<double> p = cmul(ca[j], cb[k]);
        const cplx<double> q = cmul(ca[k], cb[j]);
        c[i] = {p.re - q.re, p.im - q.im};
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        m.e[i * 3 + 0] = a[i];
        m.e[i * 3 + 1] = b[i];
        m.e[i * 3 + 2] = c[i];
    }
}

template <typename T, typename U>
__device__ __forceinline__ mat3<T> convert(const mat3<U>& m)
{
    mat3<T> r;
#pragma unroll
    for (int k = 0; k < 9; ++k) r.e[k] = {T(m.e[k].re), T(m.e[k].im)};
    return r;
}

} // namespace lqcd
