// k_gauss.hip -- Gauss law of the E field (EField::gauss, gauss_sum_div,
// project_to_gauss_step; src/field.rs:1172-1337).
//
//   G(x) = sum_i [ E_i(x) - U_i(x-i)^+ E_i(x-i) U_i(x-i) ]
//          (U_{-i}(x) = U_i(x-i)^+, LinkMatrix::matrix src/field.rs:726-740)
//   div  = sum_x | Tr( (sum_a T^a) G(x) ) |
//   step e'^a_i(x) = 2 Re Tr[ T^a ( (U_i(x) G(x) U_i(x)^+ G(x+i) - G(x)) K + T^a e^a_i(x) ) ]
//                  = e^a_i(x) + 2 Re Tr[ T^a (U G U^+ G(x+i) - G) K ],  K = 0.12
//
// The iteration works on double-precision copies of E (for f32 storage as
// well: the reference's stop rule, div <= N * 320 * eps, is far below what
// an f32 field can represent).  EField::gauss keeps G as double planes
// (k_gauss_tiles, or k_gauss for lattices without the streamed tiling); the
// projection's steps compute it where they need it (k_gauss_stream, or
// k_gauss_fused), each pair walking the sites in the same order so that the
// deviation is summed the same way in both.
#include <cstdlib>

#include "dispatch.hpp"
#include "observables_device.hpp"

namespace lqcd {

// Re Tr(T^a M), closed form of the eight generator projections
__device__ __forceinline__ void re_trace_generators(const mat3<double>& M, double* out)
{
    const double h = 0.5;
    const double r3 = 0.28867513459481288225; // 1 / (2 sqrt 3)
    out[0] = h * (M.e[1].re + M.e[3].re);
    out[1] = h * (M.e[3].im - M.e[1].im);
    out[2] = h * (M.e[0].re - M.e[4].re);
    out[3] = h * (M.e[2].re + M.e[6].re);
    out[4] = h * (M.e[6].im - M.e[2].im);
    out[5] = h * (M.e[5].re + M.e[7].re);
    out[6] = h * (M.e[7].im - M.e[5].im);
    out[7] = r3 * (M.e[0].re + M.e[4].re - 2.0 * M.e[8].re);
}

// | Tr( (sum_a T^a) G ) |; sum_a T^a = [[1/2 + r, (1-i)/2, (1-i)/2],
// [(1+i)/2, -1/2 + r, (1-i)/2], [(1+i)/2, (1+i)/2, -2r]], r = 1/(2 sqrt 3)
__device__ __forceinline__ double gsum_trace_abs(const mat3<double>& G)
{
    const double r = 0.28867513459481288225;
    const cplx<double> S[9] = {{0.5 + r, 0.0}, {0.5, -0.5}, {0.5, -0.5},
                               {0.5, 0.5},     {-0.5 + r, 0.0}, {0.5, -0.5},
                               {0.5, 0.5},     {0.5, 0.5},  {-2.0 * r, 0.0}};
    double re = 0.0, im = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const cplx<double> a = S[i * 3 + j], b = G.e[j * 3 + i];
            re += a.re * b.re - a.im * b.im;
            im += a.re * b.im + a.im * b.re;
        }
    return sqrt(re * re + im * im);
}

// ---- Hermitian 3x3 matrices ----
// E(x) (adjoint_to_matrix), U^+ E U and G(x) are Hermitian; they are carried
// as the real diagonal and the upper triangle (0,1), (0,2), (1,2), and their
// products skip the imaginary parts of the diagonal that are zero.
struct herm3 {
    double d[3];
    cplx<double> o[3];
};

__device__ __forceinline__ constexpr int hidx(int j, int k) { return j + k - 1; }

// element (j, k) of a Hermitian matrix, j != k
__device__ __forceinline__ cplx<double> hoff(const herm3& h, int j, int k)
{
    if (j < k) return h.o[hidx(j, k)];
    const cplx<double> v = h.o[hidx(k, j)];
    return {v.re, -v.im};
}

__device__ __forceinline__ mat3<double> herm_full(const herm3& h)
{
    mat3<double> m;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) m.e[j * 3 + k] = j == k ? cplx<double>{h.d[j], 0.0} : hoff(h, j, k);
    return m;
}

// the Hermitian matrix of an adjoint vector (adjoint_to_matrix)
__device__ __forceinline__ herm3 herm_from_adjoint(const double* v)
{
    const double h = 0.5;
    const double r3 = 0.28867513459481288225; // 1 / (2 sqrt 3)
    herm3 m;
    m.d[0] = h * v[2] + r3 * v[7];
    m.d[1] = -h * v[2] + r3 * v[7];
    m.d[2] = -2.0 * r3 * v[7];
    m.o[0] = {h * v[0], -h * v[1]};
    m.o[1] = {h * v[3], -h * v[4]};
    m.o[2] = {h * v[5], -h * v[6]};
    return m;
}

// p = op(a) H for a general a (op: identity or adjoint) and Hermitian H
template <bool DA>
__device__ __forceinline__ void mul_gen_herm(const mat3<double>& a, const herm3& h, mat3<double>& p)
{
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            double re = 0.0, im = 0.0;
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const cplx<double> x = elem<DA>(a, j, m);
                if (m == l) {
                    re = fma(x.re, h.d[l], re);
                    im = fma(x.im, h.d[l], im);
                } else {
                    const cplx<double> y = hoff(h, m, l);
                    re = fma(x.re, y.re, re);
                    re = fma(-x.im, y.im, re);
                    im = fma(x.re, y.im, im);
                    im = fma(x.im, y.re, im);
                }
            }
            p.e[j * 3 + l] = {re, im};
        }
}

// H = a op(u) when the product is known to be Hermitian: the diagonal's real
// part and the upper triangle
template <bool DU>
__device__ __forceinline__ herm3 mul_to_herm(const mat3<double>& a, const mat3<double>& u)
{
    herm3 r;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        double re = 0.0;
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            const cplx<double> x = a.e[j * 3 + l], y = elem<DU>(u, l, j);
            re = fma(x.re, y.re, re);
            re = fma(-x.im, y.im, re);
        }
        r.d[j] = re;
    }
#pragma unroll
    for (int jk = 0; jk < 3; ++jk) {
        const int j = jk == 2 ? 1 : 0, k = jk == 0 ? 1 : 2;
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            const cplx<double> x = a.e[j * 3 + l], y = elem<DU>(u, l, k);
            re = fma(x.re, y.re, re);
            re = fma(-x.im, y.im, re);
            im = fma(x.re, y.im, im);
            im = fma(x.im, y.re, im);
        }
        r.o[jk] = {re, im};
    }
    return r;
}

// element (j, k) of the product of two Hermitian matrices A B (only the real
// part when j == k)
__device__ __forceinline__ cplx<double> herm_mul_elem(const herm3& A, const herm3& B, int j, int k)
{
    double re = 0.0, im = 0.0;
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        if (l == j && l == k) {
            re = fma(A.d[l], B.d[l], re);
        } else if (l == j) {
            const cplx<double> y = hoff(B, l, k);
            re = fma(A.d[l], y.re, re);
            im = fma(A.d[l], y.im, im);
        } else if (l == k) {
            const cplx<double> x = hoff(A, j, l);
            re = fma(x.re, B.d[l], re);
            im = fma(x.im, B.d[l], im);
        } else {
            const cplx<double> x = hoff(A, j, l), y = hoff(B, l, k);
            re = fma(x.re, y.re, re);
            re = fma(-x.im, y.im, re);
            if (j != k) {
                im = fma(x.re, y.im, im);
                im = fma(x.im, y.re, im);
            }
        }
    }
    return {re, j == k ? 0.0 : im};
}

// G(x) = sum_i [ E_i(x) - U_i(x-i)^+ E_i(x-i) U_i(x-i) ] (EField::gauss) with
// the site offsets dn[i] of x - i
// KEEP: E_i(x) of directions i < 3 as loaded into keep, for a caller that
// steps those links next
template <int D, typename S, bool KEEP = false>
__device__ __forceinline__ herm3 gauss_h(const vec2_t<S>* __restrict__ U, const double* __restrict__ Ed,
                                         const Geom& g, int64_t s, const int32_t (&dn)[D], double (&keep)[3][8])
{
    herm3 G;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        G.d[j] = 0.0;
        G.o[j] = {0.0, 0.0};
    }
#pragma unroll
    for (int i = 0; i < D; ++i) {
        double e[8], em[8];
        load_e<double, double>(Ed, g.npad, i, s, e);
        const int64_t sm = s + dn[i];
        load_e<double, double>(Ed, g.npad, i, sm, em);
        const mat3<double> u = load_link<double, S>(U, g.npad, i, sm);
        if constexpr (KEEP)
            if (i < 3)
#pragma unroll
                for (int a = 0; a < 8; ++a) keep[i][a] = e[a];
        const herm3 Ei = herm_from_adjoint(e), Em = herm_from_adjoint(em);
        mat3<double> t;
        mul_gen_herm<true>(u, Em, t);
        const herm3 t2 = mul_to_herm<false>(t, u);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            G.d[j] = G.d[j] + (Ei.d[j] - t2.d[j]);
            G.o[j] = {G.o[j].re + (Ei.o[j].re - t2.o[j].re), G.o[j].im + (Ei.o[j].im - t2.o[j].im)};
        }
    }
    return G;
}

template <int D, typename S>
__device__ __forceinline__ herm3 gauss_h(const vec2_t<S>* __restrict__ U, const double* __restrict__ Ed,
                                         const Geom& g, int64_t s, const int32_t (&dn)[D])
{
    double unused[3][8];
    return gauss_h<D, S, false>(U, Ed, g, s, dn, unused);
}

template <int D, typename S>
__device__ __forceinline__ void gauss_at(const vec2_t<S>* __restrict__ U, const double* __restrict__ Ed,
                                         const Geom& g, int64_t s, mat3<double>& G)
{
    const Neighbours<D> nb = neighbours<D>(g, s);
    int32_t dn[D];
#pragma unroll
    for (int i = 0; i < D; ++i) dn[i] = nb.dn[i];
    G = herm_full(gauss_h<D, S>(U, Ed, g, s, dn));
}

// one project_to_gauss_step of the E of link (x, i):
// e += 2 Re Tr[ T^a (U G(x) U^+ G(x+i) - G(x)) K ] for the eight a
__device__ __forceinline__ void gauss_step_link(const mat3<double>& u, const herm3& G, const herm3& Gn, double (&e)[8])
{
    mat3<double> a;
    mul_gen_herm<false>(u, G, a);
    const herm3 b = mul_to_herm<true>(a, u);
    const double K = 0.12;
    mat3<double> c;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const cplx<double> v = herm_mul_elem(b, Gn, j, k);
            const cplx<double> gv = j == k ? cplx<double>{G.d[j], 0.0} : hoff(G, j, k);
            c.e[j * 3 + k] = {(v.re - gv.re) * K, (v.im - gv.im) * K};
        }
    double p[8];
    re_trace_generators(c, p);
#pragma unroll
    for (int a8 = 0; a8 < 8; ++a8) e[a8] = fma(2.0, p[a8], e[a8]);
}

// G(x) into double planes (9 x npad double2); per-block partials of div.
// Grid-stride walk over the sites (the walk of k_gauss_fused).
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_gauss(const vec2_t<S>* __restrict__ U, const double* __restrict__ Ed,
                                                  Geom g, double2* __restrict__ Gp,
                                                  double* __restrict__ partials)
{
    double acc[1] = {0.0};
    const int64_t step = (int64_t)gridDim.x * kBlock;
    for (int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x; s < g.nsite; s += step) {
        mat3<double> G;
        gauss_at<D, S>(U, Ed, g, s, G);
        store_link<double, double>(Gp, g.npad, 0, s, G);
        acc[0] += gsum_trace_abs(G);
    }
    block_reduce_to<1>(acc, partials + (int64_t)blockIdx.x * kReduceSlots);
}

// LDS image of a Hermitian matrix: 9 doubles at stride n
__device__ __forceinline__ void herm_put(double* base, int n, int i, const herm3& h)
{
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        base[j * n + i] = h.d[j];
        base[(3 + 2 * j) * n + i] = h.o[j].re;
        base[(4 + 2 * j) * n + i] = h.o[j].im;
    }
}

__device__ __forceinline__ herm3 herm_get(const double* base, int n, int i)
{
    herm3 h;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        h.d[j] = base[j * n + i];
        h.o[j] = {base[(3 + 2 * j) * n + i], base[(4 + 2 * j) * n + i]};
    }
    return h;
}

// The projection steps store the new E non-temporally: the next step, a
// launch of its own, reads it, and the L2 is better left to the E and U
// that neighbouring sites read again within the step.

// One project_to_gauss_step of the whole field with G computed on the fly,
// for any lattice: Ed_out = step(Ed_in), and the block partials of
// div(Ed_in).  The kernel walks the lattice in tiles of kBlock consecutive
// sites (tile t = blockIdx.x + k gridDim.x, the sites of k_gauss's
// grid-stride loop, so the div partials are summed in the same order as
// there).  G(x) of the tile goes to LDS; the step of link (x, i) takes
// G(x + i) from there when x + i lies in the tile and otherwise computes it
// again from E and U (gauss_h, the same arithmetic, so the value is the same
// either way).  G is never written to HBM.
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_gauss_fused(const vec2_t<S>* __restrict__ U,
                                                        const double* __restrict__ Ed_in, Geom g,
                                                        double* __restrict__ Ed_out,
                                                        double* __restrict__ partials)
{
    __shared__ double gl[9 * kBlock];
    double acc[1] = {0.0};
    const int64_t ntile = (g.nsite + kBlock - 1) / kBlock;
    for (int64_t t = blockIdx.x; t < ntile; t += gridDim.x) {
        const int64_t t0 = t * kBlock;
        const int64_t s = t0 + threadIdx.x;
        const bool live = s < g.nsite;
        const Neighbours<D> nb = neighbours<D>(g, live ? s : 0);
        herm3 G;
        if (live) {
            int32_t dn[D];
#pragma unroll
            for (int i = 0; i < D; ++i) dn[i] = nb.dn[i];
            G = gauss_h<D, S>(U, Ed_in, g, s, dn);
            acc[0] += gsum_trace_abs(herm_full(G));
            herm_put(gl, kBlock, threadIdx.x, G);
        }
        __syncthreads();
        if (live) {
#pragma unroll
            for (int i = 0; i < D; ++i) {
                const int64_t sn = s + nb.up[i];
                herm3 Gn;
                if (sn >= t0 && sn < t0 + kBlock && sn < g.nsite) {
                    Gn = herm_get(gl, kBlock, (int)(sn - t0));
                } else {
                    const Neighbours<D> nn = neighbours<D>(g, sn);
                    int32_t dn[D];
#pragma unroll
                    for (int k = 0; k < D; ++k) dn[k] = nn.dn[k];
                    Gn = gauss_h<D, S>(U, Ed_in, g, sn, dn);
                }
                const mat3<double> u = load_link<double, S>(U, g.npad, i, s);
                double e[8];
                load_e<double, double>(Ed_in, g.npad, i, s, e);
                gauss_step_link(u, G, Gn, e);
                store_e<double, double, kCacheStream>(Ed_out, g.npad, i, s, e);
            }
        }
        // the next tile overwrites gl
        __syncthreads();
    }
    block_reduce_to<1>(acc, partials + (int64_t)blockIdx.x * kReduceSlots);
}

// ---- streamed tiles (D = 4) ----
// A workgroup owns a tile of L0 x a x b sites (x_0 fastest, kBlock = L0 a b
// threads, one site each; a wavefront holds whole x_0 rows of one x_2 layer)
// and walks C consecutive slices x_3 = c C .. c C + C - 1 of it.  Workgroup
// w (after the XCD remap) takes tile w mod (n1 n2) of a slice and chunk
// w / (n1 n2), so each XCD streams a slab of x_3.  Per slice the LDS holds
// G of the tile at x_3 and x_3 + 1 (a ring of two: the x_3 + 1 tile is the
// next slice's own) and G of the tile's halo in +x_1 (L0 b sites) and +x_2
// (L0 a sites); every G(x + i) of a step comes from LDS, and per site the
// kernel evaluates 1 + (a + b) / (a b) + 1 / C G's instead of one per
// neighbour outside a tile of consecutive sites.
struct GaussTiles {
    int a = 0, b = 0, C = 0;   // 0: no streamed tiling for this lattice
    int n1 = 0, n2 = 0;        // tiles along x_1, x_2
    int nblocks = 0;           // n1 n2 (L3 / C)
};

constexpr int kGaussHaloMax = kBlock;

struct TileSite {
    int x[4];
};

__device__ __forceinline__ int64_t tile_site_index(const Geom& g, const TileSite& t)
{
    return t.x[0] + g.stride[1] * t.x[1] + g.stride[2] * t.x[2] + g.stride[3] * t.x[3];
}

// offsets of x - i (periodic) from the coordinates
__device__ __forceinline__ void tile_dn(const Geom& g, const TileSite& t, int32_t (&dn)[4])
{
#pragma unroll
    for (int i = 0; i < 4; ++i)
        dn[i] = t.x[i] == 0 ? (int32_t)((g.extent[i] - 1) * g.stride[i]) : -(int32_t)g.stride[i];
}

template <typename S, bool KEEP = false>
__device__ __forceinline__ herm3 gauss_tile_site(const vec2_t<S>* __restrict__ U, const double* __restrict__ Ed,
                                                 const Geom& g, const TileSite& t, double (&keep)[3][8])
{
    int32_t dn[4];
    tile_dn(g, t, dn);
    return gauss_h<4, S, KEEP>(U, Ed, g, tile_site_index(g, t), dn, keep);
}

template <typename S>
__device__ __forceinline__ herm3 gauss_tile_site(const vec2_t<S>* __restrict__ U, const double* __restrict__ Ed,
                                                 const Geom& g, const TileSite& t)
{
    double unused[3][8];
    return gauss_tile_site<S, false>(U, Ed, g, t, unused);
}

// this thread's site of the workgroup's tile at x_3 = 0, and the first slice
__device__ __forceinline__ TileSite tile_own(const Geom& g, const GaussTiles& gt, int& t0)
{
    const int64_t w = xcd_remap(blockIdx.x, gridDim.x);
    const int tps = gt.n1 * gt.n2;
    const int tt = (int)(w % tps), ch = (int)(w / tps);
    const int L0 = g.extent[0];
    const int c0 = (int)threadIdx.x % L0, r = (int)threadIdx.x / L0;
    TileSite o;
    o.x[0] = c0;
    o.x[1] = (tt % gt.n1) * gt.a + r % gt.a;
    o.x[2] = (tt / gt.n1) * gt.b + r / gt.a;
    o.x[3] = 0;
    t0 = ch * gt.C;
    return o;
}

// G(x) into double planes and the block partials of div, in the walk of
// k_gauss_stream (the same sites per thread in the same order)
template <typename S>
__global__ __launch_bounds__(kBlock) void k_gauss_tiles(const vec2_t<S>* __restrict__ U,
                                                        const double* __restrict__ Ed, Geom g,
                                                        double2* __restrict__ Gp, double* __restrict__ partials,
                                                        GaussTiles gt)
{
    int t0;
    TileSite o = tile_own(g, gt, t0);
    double acc[1] = {0.0};
    for (int k = 0; k < gt.C; ++k) {
        o.x[3] = t0 + k;
        const mat3<double> G = herm_full(gauss_tile_site<S>(U, Ed, g, o));
        store_link<double, double>(Gp, g.npad, 0, tile_site_index(g, o), G);
        acc[0] += gsum_trace_abs(G);
    }
    block_reduce_to<1>(acc, partials + (int64_t)blockIdx.x * kReduceSlots);
}

// the steps of the links (x, 0..2) of one slice: G(x) and G(x + i) from the
// LDS tile rc or the halo, E_i(x) as the G evaluation of x loaded it
template <typename S>
__device__ __forceinline__ void gauss_step_slice(const vec2_t<S>* __restrict__ U, double* __restrict__ Ed_out,
                                                 const Geom& g, const double* rc, const double* halo, int nh,
                                                 int64_t s, int tid, int j0, int j1, int j2, bool in1, bool in2,
                                                 const double (&ek)[3][8])
{
    const herm3 G = herm_get(rc, kBlock, tid);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        herm3 Gn;
        if (i == 0) Gn = herm_get(rc, kBlock, j0);
        else if (i == 1) Gn = in1 ? herm_get(rc, kBlock, j1) : herm_get(halo, nh, j1);
        else Gn = in2 ? herm_get(rc, kBlock, j2) : herm_get(halo, nh, j2);
        // the last use of this link in the launch: loaded evict-first
        const mat3<double> u = load_link<double, S, kCacheStream>(U, g.npad, i, s);
        double e[8];
#pragma unroll
        for (int a8 = 0; a8 < 8; ++a8) e[a8] = ek[i][a8];
        gauss_step_link(u, G, Gn, e);
        store_e<double, double, kCacheStream>(Ed_out, g.npad, i, s, e);
    }
}

template <typename S>
__global__ __launch_bounds__(kBlock) void k_gauss_stream(const vec2_t<S>* __restrict__ U,
                                                         const double* __restrict__ Ed_in, Geom g,
                                                         double* __restrict__ Ed_out,
                                                         double* __restrict__ partials, GaussTiles gt)
{
    // LDS: the ring of two tiles, then the halo (9 doubles per site each)
    extern __shared__ double lds[];
    int t0;
    TileSite o = tile_own(g, gt, t0);
    const int L0 = g.extent[0];
    const int tid = (int)threadIdx.x;
    const int r1 = (tid / L0) % gt.a, r2 = (tid / L0) / gt.a, c0 = tid % L0;
    // halo site of this thread: +x_1 rows (h < L0 b), then +x_2 rows
    const int nh1 = L0 * gt.b, nh = L0 * (gt.a + gt.b);
    double* const halo = lds + 18 * kBlock;
    TileSite hs = o;
    if (tid < nh1) {
        hs.x[1] = (o.x[1] - r1 + gt.a) % g.extent[1];
        hs.x[2] = o.x[2] - r2 + tid / L0;
    } else if (tid < nh) {
        hs.x[1] = o.x[1] - r1 + (tid - nh1) / L0;
        hs.x[2] = (o.x[2] - r2 + gt.b) % g.extent[2];
    }
    hs.x[0] = c0;
    // LDS index of G(x + i) for i = 0, 1, 2: in the tile, or in the halo
    const int j0 = tid - c0 + (c0 + 1) % L0;
    const bool in1 = r1 + 1 < gt.a, in2 = r2 + 1 < gt.b;
    const int j1 = in1 ? tid + L0 : c0 + L0 * r2;
    const int j2 = in2 ? tid + L0 * gt.a : nh1 + c0 + L0 * r1;

    // Slice t0: G of the tile and its halo, then the steps of its links in
    // x_0, x_1, x_2.  Iteration k (slice t = t0 + k): G of the tile at t + 1
    // (the ring's other half), the halo of t + 1 and the steps of its links
    // in x_0..x_2 (while t + 1 is in the chunk), and the steps of the links
    // in x_3 of slice t, which need G(t) and G(t + 1).  Each link is stepped
    // in the iteration whose G evaluations have just loaded its E and U.
    double acc[1] = {0.0};
    double ek[3][8];
    o.x[3] = t0;
    {
        const herm3 G = gauss_tile_site<S, true>(U, Ed_in, g, o, ek);
        acc[0] += gsum_trace_abs(herm_full(G));
        herm_put(lds, kBlock, tid, G);
        if (tid < nh) {
            hs.x[3] = t0;
            herm_put(halo, nh, tid, gauss_tile_site<S>(U, Ed_in, g, hs));
        }
    }
    __syncthreads();
    gauss_step_slice<S>(U, Ed_out, g, lds, halo, nh, tile_site_index(g, o), tid, j0, j1, j2, in1, in2, ek);
    __syncthreads();
    for (int k = 0; k < gt.C; ++k) {
        const int cur = k & 1;
        double* const rc = lds + 9 * kBlock * cur;
        double* const rn = lds + 9 * kBlock * (cur ^ 1);
        const int t = t0 + k;
        const bool more = k + 1 < gt.C;
        TileSite on = o;
        on.x[3] = (t + 1) % g.extent[3];
        {
            const herm3 Gn = gauss_tile_site<S, true>(U, Ed_in, g, on, ek);
            if (more) acc[0] += gsum_trace_abs(herm_full(Gn));
            herm_put(rn, kBlock, tid, Gn);
        }
        if (more && tid < nh) {
            hs.x[3] = t + 1;
            herm_put(halo, nh, tid, gauss_tile_site<S>(U, Ed_in, g, hs));
        }
        __syncthreads();
        {
            o.x[3] = t;
            const int64_t s = tile_site_index(g, o);
            const herm3 G = herm_get(rc, kBlock, tid), Gn = herm_get(rn, kBlock, tid);
            const mat3<double> u = load_link<double, S>(U, g.npad, 3, s);
            double e[8];
            load_e<double, double>(Ed_in, g.npad, 3, s, e);
            gauss_step_link(u, G, Gn, e);
            store_e<double, double, kCacheStream>(Ed_out, g.npad, 3, s, e);
        }
        if (more)
            gauss_step_slice<S>(U, Ed_out, g, rn, halo, nh, tile_site_index(g, on), tid, j0, j1, j2, in1, in2, ek);
        // the next iteration overwrites rc and the halo
        __syncthreads();
    }
    block_reduce_to<1>(acc, partials + (int64_t)blockIdx.x * kReduceSlots);
}

// E (storage type) <-> double copy
template <typename Sin, typename Sout>
__global__ __launch_bounds__(kBlock) void k_efield_convert(const Sin* __restrict__ in, Sout* __restrict__ out,
                                                           Geom g)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= g.nsite) return;
    double e[8];
    load_e<double, Sin>(in, g.npad, mu, s, e);
    store_e<double, Sout>(out, g.npad, mu, s, e);
}

// G planes -> [site][18] column-major re/im (the reference's CMatrix3 layout)
__global__ __launch_bounds__(kBlock) void k_gauss_to_aos(const double2* __restrict__ Gp, Geom g,
                                                         double* __restrict__ out)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= g.nsite) return;
    const mat3<double> G = load_link<double, double>(Gp, g.npad, 0, s);
    double* p = out + s * 18;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            p[(c * 3 + r) * 2] = G.e[r * 3 + c].re;
            p[(c * 3 + r) * 2 + 1] = G.e[r * 3 + c].im;
        }
}

hipError_t launch_gauss_to_aos(const lqcd_lattice* L, double* d_out)
{
    k_gauss_to_aos<<<grid_for(L->g.nsite), kBlock, 0, L->ctx->stream>>>((const double2*)L->gauss_g, L->g, d_out);
    return hipGetLastError();
}

static int gauss_blocks(int64_t n)
{
    int64_t b = (n + kBlock - 1) / kBlock;
    if (b > kMaxReduceBlocks) b = kMaxReduceBlocks;
    return (int)(b < 1 ? 1 : b);
}

// The streamed tiling of a D = 4 lattice, or a = 0 when it does not apply
// (or LQCD_GAUSS_TILES=0 pins the site walk): L0 divides 64 (whole x_0 rows
// per wavefront), a tile L0 x a x b of kBlock sites with the rows of a
// wavefront in one x_2 layer, a | L1, b | L2 and the smallest halo L0 (a + b);
// C | L3 the largest of 32, 16, .., 1 that leaves min(512, half the tiles of the
// lattice) workgroups or more (512: one round of two workgroups per CU) and at
// most kMaxReduceBlocks.
static GaussTiles gauss_tiles(const Geom& g)
{
    GaussTiles gt;
    if (g.dim != 4) return gt;
    const char* env = std::getenv("LQCD_GAUSS_TILES");
    if (env && env[0] == '0' && env[1] == 0) return gt;
    const int L0 = g.extent[0], L1 = g.extent[1], L2 = g.extent[2], L3 = g.extent[3];
    if (L0 > 64 || 64 % L0 != 0) return gt;
    const int m = kBlock / L0, rpw = 64 / L0;
    int best_a = 0, best_b = 0;
    for (int a = rpw; a <= m; a += rpw) {
        if (m % a != 0) continue;
        const int b = m / a;
        if (L1 % a != 0 || L2 % b != 0) continue;
        if (L0 * (a + b) > kGaussHaloMax) continue;
        if (best_a == 0 || a + b < best_a + best_b) {
            best_a = a;
            best_b = b;
        }
    }
    if (best_a == 0) return gt;
    const int64_t tiles = (int64_t)(L1 / best_a) * (L2 / best_b);
    const int64_t all = tiles * L3;
    const int64_t target = all / 2 < 512 ? all / 2 : 512;
    int C = 0;
    for (int c : {32, 16, 8, 4, 2, 1})
        if (L3 % c == 0 && all / c >= target && all / c <= kMaxReduceBlocks) {
            C = c;
            break;
        }
    if (C == 0) return gt;
    gt.a = best_a;
    gt.b = best_b;
    gt.C = C;
    gt.n1 = L1 / best_a;
    gt.n2 = L2 / best_b;
    gt.nblocks = (int)(all / C);
    return gt;
}

hipError_t launch_gauss(const lqcd_lattice* L, const double* Ed, double* d_div)
{
    const GaussTiles gt = gauss_tiles(L->g);
    int nb;
    if (gt.a > 0) {
        nb = gt.nblocks;
        LQCD_DISPATCH_S(L, k_gauss_tiles<S><<<nb, kBlock, 0, L->ctx->stream>>>(
                               (const vec2_t<S>*)L->U[L->cur], Ed, L->g, (double2*)L->gauss_g, L->partials, gt));
    } else {
        nb = gauss_blocks(L->g.nsite);
        LQCD_DISPATCH_DS(L, k_gauss<D, S><<<nb, kBlock, 0, L->ctx->stream>>>(
                                (const vec2_t<S>*)L->U[L->cur], Ed, L->g, (double2*)L->gauss_g, L->partials));
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_final_reduce1(L, nb, d_div);
}

hipError_t launch_gauss_fused(const lqcd_lattice* L, const double* Ed_in, double* Ed_out, double* d_div)
{
    const GaussTiles gt = gauss_tiles(L->g);
    int nb;
    if (gt.a > 0) {
        nb = gt.nblocks;
        const size_t lds = sizeof(double) * 9 * (2 * kBlock + (size_t)L->g.extent[0] * (gt.a + gt.b));
        LQCD_DISPATCH_S(L, k_gauss_stream<S><<<nb, kBlock, lds, L->ctx->stream>>>(
                               (const vec2_t<S>*)L->U[L->cur], Ed_in, L->g, Ed_out, L->partials, gt));
    } else {
        nb = gauss_blocks(L->g.nsite);
        LQCD_DISPATCH_DS(L, k_gauss_fused<D, S><<<nb, kBlock, 0, L->ctx->stream>>>(
                                (const vec2_t<S>*)L->U[L->cur], Ed_in, L->g, Ed_out, L->partials));
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_final_reduce1(L, nb, d_div);
}

hipError_t launch_efield_to_double(const lqcd_lattice* L, const void* E, double* Ed)
{
    const dim3 grid(grid_for(L->g.nsite), (unsigned)L->g.dim);
    LQCD_DISPATCH_S(L, k_efield_convert<S, double><<<grid, kBlock, 0, L->ctx->stream>>>((const S*)E, Ed, L->g));
    return hipGetLastError();
}

hipError_t launch_efield_from_double(const lqcd_lattice* L, const double* Ed, void* E)
{
    const dim3 grid(grid_for(L->g.nsite), (unsigned)L->g.dim);
    LQCD_DISPATCH_S(L, k_efield_convert<double, S><<<grid, kBlock, 0, L->ctx->stream>>>(Ed, (S*)E, L->g));
    return hipGetLastError();
}

} // namespace lqcd
