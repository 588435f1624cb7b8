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
// (k_gauss); the projection's steps compute it where they need it
// (k_gauss_fused).
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

template <int D, typename S>
__device__ __forceinline__ void gauss_at(const vec2_t<S>* __restrict__ U, const double* __restrict__ Ed,
                                         const Geom& g, int64_t s, mat3<double>& G)
{
    const Neighbours<D> nb = neighbours<D>(g, s);
    set_zero(G);
#pragma unroll
    for (int i = 0; i < D; ++i) {
        double e[8], em[8];
        load_e<double, double>(Ed, g.npad, i, s, e);
        const int64_t sm = s + nb.dn[i];
        load_e<double, double>(Ed, g.npad, i, sm, em);
        const mat3<double> u = load_link<double, S>(U, g.npad, i, sm);
        mat3<double> Ei, Em, t, t2;
        adjoint_to_matrix(e, Ei);
        adjoint_to_matrix(em, Em);
        mul<true, false>(u, Em, t);
        mul<false, false>(t, u, t2);
#pragma unroll
        for (int k = 0; k < 9; ++k)
            G.e[k] = {G.e[k].re + (Ei.e[k].re - t2.e[k].re), G.e[k].im + (Ei.e[k].im - t2.e[k].im)};
    }
}

// G(x) into double planes (9 x npad double2); per-block partials of div
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

// One project_to_gauss_step of the whole field with G computed on the fly:
// Ed_out = step(Ed_in), and the block partials of div(Ed_in).  The kernel
// walks the lattice in tiles of kBlock consecutive sites (tile t = blockIdx.x
// + k gridDim.x, the sites of k_gauss's grid-stride loop, so the div partials
// are summed in the same order as there).  G(x) of the tile goes to LDS; the
// step of link (x, i) takes G(x + i) from there when x + i lies in the tile
// (x + 0 always for extent[0] dividing kBlock, x + 1 for all but the last
// row) and otherwise computes it again from E and U (gauss_at, the same
// arithmetic, so the value is the same either way).  G is never written to
// HBM.
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_gauss_fused(const vec2_t<S>* __restrict__ U,
                                                        const double* __restrict__ Ed_in, Geom g,
                                                        double* __restrict__ Ed_out,
                                                        double* __restrict__ partials)
{
    __shared__ double gl[18][kBlock];
    double acc[1] = {0.0};
    const int64_t ntile = (g.nsite + kBlock - 1) / kBlock;
    for (int64_t t = blockIdx.x; t < ntile; t += gridDim.x) {
        const int64_t t0 = t * kBlock;
        const int64_t s = t0 + threadIdx.x;
        const bool live = s < g.nsite;
        mat3<double> G;
        if (live) {
            gauss_at<D, S>(U, Ed_in, g, s, G);
            acc[0] += gsum_trace_abs(G);
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                gl[2 * k][threadIdx.x] = G.e[k].re;
                gl[2 * k + 1][threadIdx.x] = G.e[k].im;
            }
        }
        __syncthreads();
        if (live) {
            const Neighbours<D> nb = neighbours<D>(g, s);
#pragma unroll
            for (int i = 0; i < D; ++i) {
                const int64_t sn = s + nb.up[i];
                mat3<double> Gn;
                if (sn >= t0 && sn < t0 + kBlock && sn < g.nsite) {
                    const int j = (int)(sn - t0);
#pragma unroll
                    for (int k = 0; k < 9; ++k) Gn.e[k] = {gl[2 * k][j], gl[2 * k + 1][j]};
                } else {
                    gauss_at<D, S>(U, Ed_in, g, sn, Gn);
                }
                const mat3<double> u = load_link<double, S>(U, g.npad, i, s);
                mat3<double> a, b, c;
                mul<false, false>(u, G, a);
                mul<false, true>(a, u, b);
                mul<false, false>(b, Gn, c);
                const double K = 0.12;
#pragma unroll
                for (int k = 0; k < 9; ++k)
                    c.e[k] = {(c.e[k].re - G.e[k].re) * K, (c.e[k].im - G.e[k].im) * K};
                double p[8], e[8];
                re_trace_generators(c, p);
                load_e<double, double>(Ed_in, g.npad, i, s, e);
#pragma unroll
                for (int a8 = 0; a8 < 8; ++a8) e[a8] = fma(2.0, p[a8], e[a8]);
                store_e<double, double>(Ed_out, g.npad, i, s, e);
            }
        }
        // the next tile overwrites gl
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

hipError_t launch_gauss(const lqcd_lattice* L, const double* Ed, double* d_div)
{
    const int nb = gauss_blocks(L->g.nsite);
    LQCD_DISPATCH_DS(L, k_gauss<D, S><<<nb, kBlock, 0, L->ctx->stream>>>(
                            (const vec2_t<S>*)L->U[L->cur], Ed, L->g, (double2*)L->gauss_g, L->partials));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_final_reduce1(L, nb, d_div);
}

hipError_t launch_gauss_fused(const lqcd_lattice* L, const double* Ed_in, double* Ed_out, double* d_div)
{
    const int nb = gauss_blocks(L->g.nsite);
    LQCD_DISPATCH_DS(L, k_gauss_fused<D, S><<<nb, kBlock, 0, L->ctx->stream>>>(
                            (const vec2_t<S>*)L->U[L->cur], Ed_in, L->g, Ed_out, L->partials));
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
