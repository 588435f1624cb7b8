// k_md.hip -- gauge force, momentum kick and link drift on one lattice.
//
// The per-link and per-site arithmetic lives in md_device.hpp.  The
// kick+drift kernels read U_in and write U_out (double buffered): one launch
// performs e += kick * F(U_in) and then the drift of its own link with the
// updated e -- one pass over HBM per leapfrog step.
//
// Work decompositions (chosen per lattice by the autotuner in capi.cpp):
//   VAR_LINK      one thread per link, grid (site blocks, D): the plain layout.
//   VAR_LINK_XCD  one thread per link, 1-D grid where the D links of a site
//                 block are consecutive blocks and consecutive blocks are
//                 remapped onto the same XCD, so neighbouring site blocks share
//                 that XCD's L2 instead of being spread over all eight.
//   VAR_SITE_XCD  one thread per site for all D links: the D staple sums are
//                 built together, which shares the loads (11 matrices per
//                 link instead of 19) and the product
//                 W = U_nu(x+mu) U_mu(x+nu)^+ between the upper staples of
//                 (mu, nu) and (nu, mu); same XCD remap.
//   VAR_SITE_SLAB per site as VAR_SITE_XCD, but each XCD gets a slab of the
//                 slowest direction and walks its site blocks with that
//                 direction fastest, so the blocks holding the +-1 neighbours
//                 in the slowest direction run next to each other in time and
//                 share the XCD's L2 (needs 256 | sites per slice and 8 | L_slow;
//                 otherwise it falls back to VAR_SITE_XCD's order).
//   VAR_SITE_TILE per site in VAR_SITE_SLAB's order with an LDS tile of
//                 two-row f32 links: each workgroup stages its whole x_0 rows
//                 and the x_1 rows on either side, and reads the links one or
//                 two steps away in x_0 / x_1 from LDS (20 of the 44 link
//                 reads of a D = 4 site); the others come from U_in.
#include "dispatch.hpp"
#include "md_device.hpp"
#include "observables_device.hpp"

namespace lqcd {

// derivative_e (src/simulation/state.rs:1420-1448) of every link, times
// oscale, in the host layout [site][mu][8]: one thread per site, the D staple
// sums sharing their products (staple_sums_site).  A block's output is one
// contiguous run of kForceBlock * D * 8 doubles; the threads put their
// values into an LDS image of it (rows of D * 8 + 1 doubles, so that the
// lanes' 8-byte writes fall into different banks) and the block stores the
// image as whole 16-byte words, lane after lane.
constexpr int kForceBlock = 128;

template <int D, typename S, typename T>
__global__ __launch_bounds__(kForceBlock) void k_force_store(const vec2_t<S>* __restrict__ U, Geom g,
                                                             T fscale, double* __restrict__ out,
                                                             double oscale)
{
    constexpr int kRow = D * 8 + 1;
    __shared__ double img[kForceBlock * kRow];
    const int64_t s0 = (int64_t)blockIdx.x * kForceBlock;
    const int64_t s = s0 + threadIdx.x;
    if (s < g.nsite) {
        const Neighbours<D> nb = neighbours<D>(g, s);
        mat3<T> sig[D];
        staple_sums_site<D, S, T>(U, g, s, nb, sig);
#pragma unroll
        for (int mu = 0; mu < D; ++mu) {
            const mat3<T> u = load_link<T, S>(U, g.npad, mu, s);
            mat3<T> m;
            mul<false, false>(u, sig[mu], m);
            T f[8];
            im_trace_generators(m, f);
#pragma unroll
            for (int a = 0; a < 8; ++a) img[threadIdx.x * kRow + mu * 8 + a] = oscale * (double)(fscale * f[a]);
        }
    }
    __syncthreads();
    // word w of the block's run: site w / (D * 4), doubles 2 (w % (D * 4)) + 0, 1
    const int n = (int)min<int64_t>(kForceBlock, g.nsite - s0);
    double2* o = reinterpret_cast<double2*>(out + s0 * D * 8);
    for (int w = threadIdx.x; w < n * D * 4; w += kForceBlock) {
        const int site = w / (D * 4), k = 2 * (w % (D * 4));
        o[w] = make_double2(img[site * kRow + k], img[site * kRow + k + 1]);
    }
}

// derivative_u (src/simulation/state.rs:1407-1417) of every link, double
// arithmetic: ((E U) i sqrt(2 CA)) / a, written in the host layout
// [site][mu][column-major 3x3][re, im]
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_derivative_u(const vec2_t<S>* __restrict__ U,
                                                         const S* __restrict__ E, Geom g, double inv_a,
                                                         double* __restrict__ out)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= g.nsite) return;
    double e[8];
    load_e<double, S>(E, g.npad, mu, s, e);
    mat3<double> em, eu;
    adjoint_to_matrix(e, em);
    mul<false, false>(em, load_link<double, S>(U, g.npad, mu, s), eu);
    const double c = sqrt(6.0);
    double* p = out + (s * D + mu) * 18;
#pragma unroll
    for (int col = 0; col < 3; ++col)
#pragma unroll
        for (int row = 0; row < 3; ++row) {
            // (x + iy) (0 + i c) = -c y + i c x, then times (1/a + 0i)
            const cplx<double> z = eu.e[row * 3 + col];
            p[(col * 3 + row) * 2] = (-c * z.im) * inv_a;
            p[(col * 3 + row) * 2 + 1] = (c * z.re) * inv_a;
        }
}

// VAR_LINK: grid (site blocks, D)
template <int D, typename S, typename T, int DR, bool R12>
__global__ __launch_bounds__(kBlock) void k_md_link(const vec2_t<S>* __restrict__ U_in,
                                                    vec2_t<S>* __restrict__ U_out,
                                                    S* __restrict__ E, Geom g, T fscale, T kick,
                                                    T dfac)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int mu = blockIdx.y;
    if (s >= g.nsite) return;
    md_link<D, S, T, DR, R12, R12 && DR == DRIFT_EXP>(U_in, U_out, E, g, mu, s, fscale, kick, dfac);
}

// VAR_LINK_XCD: 1-D grid of (site block, mu) pairs, mu fastest, XCD remapped
template <int D, typename S, typename T, int DR, bool R12>
__global__ __launch_bounds__(kBlock) void k_md_link_xcd(const vec2_t<S>* __restrict__ U_in,
                                                        vec2_t<S>* __restrict__ U_out,
                                                        S* __restrict__ E, Geom g, T fscale, T kick,
                                                        T dfac)
{
    const int64_t logical = xcd_remap(blockIdx.x, gridDim.x);
    const int mu = (int)(logical % D);
    const int64_t s = (logical / D) * kBlock + threadIdx.x;
    if (s >= g.nsite) return;
    md_link<D, S, T, DR, R12, R12 && DR == DRIFT_EXP>(U_in, U_out, E, g, mu, s, fscale, kick, dfac);
}

// VAR_SITE_XCD / VAR_SITE_SLAB (slab_p > 0): one thread per site, all D links.
// SUMS: the workgroup's Hamiltonian sums (MdSums) go to partials[blockIdx.x];
// row2: see md_site.
template <int D, typename S, typename T, int DR, bool R12, int SUMS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 2 : 1)))
void k_md_site_xcd(const vec2_t<S>* __restrict__ U_in,
                                                        vec2_t<S>* __restrict__ U_out,
                                                        S* __restrict__ E, Geom g, T fscale, T kick,
                                                        T dfac, int64_t slab_p, int64_t slab_g,
                                                        double* __restrict__ partials, vec2_t<S>* row2)
{
    int64_t logical = xcd_remap(blockIdx.x, gridDim.x);
    if (slab_p > 0) logical = slab_block(logical, slab_p, slab_g);
    const int64_t s = logical * kBlock + threadIdx.x;
    // f32 two-row reads keep the own links in LDS (72 KiB at D = 4: two
    // workgroups per CU, the occupancy the registers allow anyway)
    constexpr bool OWN = R12 && sizeof(T) == 4;
    __shared__ vec2_t<T> own_lds[OWN ? D * 9 * kBlock : 1];
    constexpr bool R01 = R12 && DR == DRIFT_EXP;
    if constexpr (SUMS == SUMS_NONE) {
        if (s >= g.nsite) return;
        md_site<D, S, T, DR, R12, OWN, R01>(U_in, U_out, E, g, s, fscale, kick, dfac, own_lds, nullptr, row2);
    } else {
        MdAcc<T> acc;
        if (s < g.nsite)
            md_site<D, S, T, DR, R12, OWN, R01, SUMS>(U_in, U_out, E, g, s, fscale, kick, dfac, own_lds, &acc,
                                                      row2);
        double v[2] = {(double)acc.tr, (double)acc.e2};
        block_reduce_to<2>(v, partials + (int64_t)blockIdx.x * kReduceSlots);
    }
}

// VAR_SITE_TILE: per site as VAR_SITE_SLAB (same block order), two-row
// links only: the workgroup first stages rows 0, 1 of the links of its tile
// and of the two neighbouring x_1 rows in LDS (TileShape), then reads every
// link reachable by steps in x_0 and x_1 from there and the rest from U_in.
template <int D, typename S, typename T, int DR, int SUMS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2)))
void k_md_site_tile(const vec2_t<S>* __restrict__ U_in,
                                                         vec2_t<S>* __restrict__ U_out,
                                                         S* __restrict__ E, Geom g, T fscale, T kick,
                                                         T dfac, int64_t slab_p, int64_t slab_g,
                                                         double* __restrict__ partials, vec2_t<S>* row2)
{
    extern __shared__ unsigned char tile_raw[];
    auto* tile = reinterpret_cast<vec2_t<T>*>(tile_raw);
    int64_t logical = xcd_remap(blockIdx.x, gridDim.x);
    if (slab_p > 0) logical = slab_block(logical, slab_p, slab_g);
    TileShape ts;
    ts.L0 = g.extent[0];
    ts.rows = kBlock / ts.L0;
    ts.slots = (ts.rows + 2) * ts.L0;
    const int64_t t0 = logical * kBlock;
    const int r = threadIdx.x / ts.L0, c = threadIdx.x % ts.L0;
    const int64_t s = t0 + threadIdx.x;
    // the tile's own rows, then the rows x_1 - 1 and x_1 + rows
#pragma unroll
    for (int nu = 0; nu < D; ++nu) {
        const mat3<T> m = load_link_rows01<T, S>(U_in, g.npad, nu, s);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            vec2_t<T> v;
            v.x = m.e[k].re;
            v.y = m.e[k].im;
            tile[(nu * 6 + k) * ts.slots + (r + 1) * ts.L0 + c] = v;
        }
    }
    {
        const int L1 = g.extent[1];
        const int64_t x1_0 = (t0 / ts.L0) % L1;
        const int64_t base = t0 - x1_0 * ts.L0;
        for (int j = threadIdx.x; j < 2 * ts.L0 * D; j += kBlock) {
            const int h = j % (2 * ts.L0), nu = j / (2 * ts.L0);
            const int hr = h / ts.L0, hc = h % ts.L0;
            const int64_t y1 = hr == 0 ? (x1_0 + L1 - 1) % L1 : (x1_0 + ts.rows) % L1;
            const mat3<T> m = load_link_rows01<T, S>(U_in, g.npad, nu, base + y1 * ts.L0 + hc);
            const int slot = (hr == 0 ? 0 : ts.rows + 1) * ts.L0 + hc;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                vec2_t<T> v;
                v.x = m.e[k].re;
                v.y = m.e[k].im;
                tile[(nu * 6 + k) * ts.slots + slot] = v;
            }
        }
    }
    __syncthreads();
    const Neighbours<D> nb = neighbours<D>(g, s);
    const int own = tile_slot(ts, r, c, 0, 0);
    MdAcc<T> acc;
    T f[D][8];
    {
        mat3<T> sig[D];
        staple_sums_site_tile<D, S, T>(U_in, g, s, nb, tile, ts, r, c, sig);
#pragma unroll
        for (int mu = 0; mu < D; ++mu) {
            const mat3<T> u = tile_r12<T>(tile, ts, mu, own);
            mat3<T> m;
            mul<false, false>(u, sig[mu], m);
            im_trace_combinations(m, f[mu]);
            if constexpr (SUMS != SUMS_NONE) acc.tr += re_trace(m);
        }
    }
    T e[D][8];
#pragma unroll
    for (int mu = 0; mu < D; ++mu) {
        load_e<T, S, kCacheStream>(E, g.npad, mu, s, e[mu]);
        if constexpr (SUMS == SUMS_BEFORE_KICK) acc.e2 += e_squares8<T>(e[mu]);
        kick_generators(f[mu], kick * fscale, e[mu]);
        if constexpr (SUMS == SUMS_AFTER_KICK) acc.e2 += e_squares8<T>(e[mu]);
        store_e<T, S, kCacheStream>(E, g.npad, mu, s, e[mu]);
    }
    if (row2) {
#pragma unroll
        for (int mu = 0; mu < D; ++mu) store_link_row2<T, S>(row2, g.npad, mu, s, tile_r12<T>(tile, ts, mu, own));
    }
    if constexpr (DR >= 0) {
#pragma unroll
        for (int mu = 0; mu < D; ++mu) {
            const mat3<T> u = tile_r12<T>(tile, ts, mu, own);
            drift_store<DR, DR == DRIFT_EXP, S, T>(u, e[mu], dfac, U_out, g, mu, s);
        }
    }
    if constexpr (SUMS != SUMS_NONE) {
        double v[2] = {(double)acc.tr, (double)acc.e2};
        block_reduce_to<2>(v, partials + (int64_t)blockIdx.x * kReduceSlots);
    }
}

// whether VAR_SITE_TILE can run on this lattice: two-row f32 links, whole
// x_0 rows per workgroup that never straddle an x_2 boundary, and an LDS tile
// of at most 64 KiB
bool md_tile_fits(const lqcd_lattice* L)
{
    const Geom& g = L->g;
    if (L->precision != 1 || g.dim < 2 || g.nsite % kBlock != 0) return false;
    const int L0 = g.extent[0];
    if (L0 < 2 || kBlock % L0 != 0) return false;
    const int rows = kBlock / L0;
    if (g.extent[1] % rows != 0 || g.extent[1] < rows) return false;
    return md_tile_bytes(L) <= 65536;
}

size_t md_tile_bytes(const lqcd_lattice* L)
{
    const int L0 = L->g.extent[0];
    const size_t slots = (size_t)(kBlock / L0 + 2) * (size_t)L0;
    return (size_t)L->g.dim * 6 * slots * 2 * sizeof(float);
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

// the per-site kernels with the Hamiltonian sums of the launch (x->sums: 0,
// or SUMS_BEFORE_KICK for a kick+drift, SUMS_AFTER_KICK for a kick) and the
// row-2 write (x->row2)
template <int D, typename S, int DR, bool R12, int SUMS>
static void launch_site(const lqcd_lattice* L, bool tile, unsigned nsb, const vec2_t<S>* ui, vec2_t<S>* uo, S* e,
                        S fs, S k, S df, int64_t slab_p, int64_t slab_g, const MdExtra* x)
{
    hipStream_t st = L->ctx->stream;
    double* partials = x ? x->partials : nullptr;
    auto* row2 = (vec2_t<S>*)(x ? x->row2 : nullptr);
    if constexpr (R12 && sizeof(S) == 4) {
        if (tile) {
            k_md_site_tile<D, S, S, DR, SUMS><<<nsb, kBlock, md_tile_bytes(L), st>>>(ui, uo, e, L->g, fs, k, df, slab_p,
                                                                                    slab_g, partials, row2);
            return;
        }
    }
    k_md_site_xcd<D, S, S, DR, R12, SUMS><<<nsb, kBlock, 0, st>>>(ui, uo, e, L->g, fs, k, df, slab_p, slab_g,
                                                                  partials, row2);
}

template <int D, typename S, int DR, bool R12>
static void launch_md_variant_r(const lqcd_lattice* L, int variant, const void* U_in, void* U_out, void* E,
                                double kick, double dfac, const MdExtra* x)
{
    const S fs = (S)md_force_scale(L->size), k = (S)kick, df = (S)dfac;
    const auto* ui = (const vec2_t<S>*)U_in;
    auto* uo = (vec2_t<S>*)U_out;
    auto* e = (S*)E;
    hipStream_t st = L->ctx->stream;
    const unsigned nsb = grid_for(L->g.nsite);
    const int sums = x ? x->sums : SUMS_NONE;
    switch (variant) {
    case VAR_SITE_TILE:
    case VAR_SITE_XCD:
    case VAR_SITE_SLAB: {
        const bool tile = variant == VAR_SITE_TILE && R12 && sizeof(S) == 4 && md_tile_fits(L);
        int64_t slab_p = 0, slab_g = 0;
        const int64_t Ls = L->g.extent[D - 1];
        const int64_t slice = L->g.nsite / Ls;
        if (variant != VAR_SITE_XCD && slice % kBlock == 0 && Ls % 8 == 0) {
            slab_p = slice / kBlock;
            slab_g = Ls / 8;
        }
        constexpr int SUMS_ON = DR >= 0 ? SUMS_BEFORE_KICK : SUMS_AFTER_KICK;
        if (sums == SUMS_NONE)
            launch_site<D, S, DR, R12, SUMS_NONE>(L, tile, nsb, ui, uo, e, fs, k, df, slab_p, slab_g, x);
        else
            launch_site<D, S, DR, R12, SUMS_ON>(L, tile, nsb, ui, uo, e, fs, k, df, slab_p, slab_g, x);
        break;
    }
    case VAR_LINK_XCD:
        k_md_link_xcd<D, S, S, DR, R12><<<nsb * (unsigned)D, kBlock, 0, st>>>(ui, uo, e, L->g, fs, k, df);
        break;
    default:
        k_md_link<D, S, S, DR, R12><<<link_grid(L), kBlock, 0, st>>>(ui, uo, e, L->g, fs, k, df);
        break;
    }
}

// links read as two rows (R12) when the integration uses the exp drift and
// has checked them to be in SU(3) (lqcd_lattice::md_r12)
template <int D, typename S, int DR>
static void launch_md_variant(const lqcd_lattice* L, int variant, const void* U_in, void* U_out, void* E,
                              double kick, double dfac, const MdExtra* x)
{
    if constexpr (DR != DRIFT_ADDITIVE) {
        if (L->md_r12) {
            launch_md_variant_r<D, S, DR, true>(L, variant, U_in, U_out, E, kick, dfac, x);
            return;
        }
    }
    launch_md_variant_r<D, S, DR, false>(L, variant, U_in, U_out, E, kick, dfac, x);
}

static int chosen_variant(const lqcd_lattice* L)
{
    return (L->md_variant >= 0 && L->md_variant < VAR_COUNT) ? L->md_variant : VAR_LINK_XCD;
}

hipError_t launch_derivative_u(const lqcd_lattice* L, double* d_out)
{
    LQCD_DISPATCH_DS(L, k_derivative_u<D, S><<<link_grid(L), kBlock, 0, L->ctx->stream>>>(
                            (const vec2_t<S>*)L->U[L->cur], (const S*)L->E, L->g, 1.0 / L->size, d_out));
    return hipGetLastError();
}

hipError_t launch_force_store(const lqcd_lattice* L, const void* U, double* d_out, double scale)
{
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_FORCE_STORE, &ev);
    LQCD_DISPATCH_DS(L, k_force_store<D, S, S><<<grid_for(L->g.nsite, kForceBlock), kForceBlock, 0,
                                                L->ctx->stream>>>((const vec2_t<S>*)U, L->g,
                                                                  (S)md_force_scale(L->size), d_out, scale));
    prof_end(L->ctx, &ev);
    return hipGetLastError();
}

hipError_t launch_kick_variant(const lqcd_lattice* L, int variant, const void* U, void* E, double kick,
                               const MdExtra* x)
{
    LQCD_DISPATCH_DS(L, launch_md_variant<D, S, -1>(L, variant, U, nullptr, E, kick, 0.0, x));
    return hipGetLastError();
}

hipError_t launch_kick_drift_variant(const lqcd_lattice* L, int variant, const void* U_in, void* U_out,
                                     void* E, double kick, double dt, int drift_mode, const MdExtra* x)
{
    const double df = md_drift_scale(dt, L->size);
    if (drift_mode == DRIFT_EXP) {
        LQCD_DISPATCH_DS(L, launch_md_variant<D, S, DRIFT_EXP>(L, variant, U_in, U_out, E, kick, df, x));
    } else {
        LQCD_DISPATCH_DS(L, launch_md_variant<D, S, DRIFT_ADDITIVE>(L, variant, U_in, U_out, E, kick, df, x));
    }
    return hipGetLastError();
}

hipError_t launch_kick(const lqcd_lattice* L, const void* U, void* E, double kick, const MdExtra* x)
{
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_KICK, &ev);
    const hipError_t e = launch_kick_variant(L, chosen_variant(L), U, E, kick, x);
    prof_end(L->ctx, &ev);
    return e;
}

hipError_t launch_kick_drift(const lqcd_lattice* L, const void* U_in, void* U_out, void* E, double kick,
                             double dt, int drift_mode, const MdExtra* x)
{
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_KICK_DRIFT, &ev);
    const hipError_t e =
        launch_kick_drift_variant(L, chosen_variant(L), U_in, U_out, E, kick, dt, drift_mode, x);
    prof_end(L->ctx, &ev);
    return e;
}

bool md_variant_has_sums(int variant)
{
    return variant == VAR_SITE_XCD || variant == VAR_SITE_SLAB || variant == VAR_SITE_TILE;
}

hipError_t launch_drift(const lqcd_lattice* L, const void* U_in, void* U_out, const void* E, double dt,
                        int drift_mode)
{
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_DRIFT, &ev);
    const double df = md_drift_scale(dt, L->size);
    if (drift_mode == DRIFT_EXP) {
        LQCD_DISPATCH_DS(L, k_drift<D, S, S, DRIFT_EXP><<<link_grid(L), kBlock, 0, L->ctx->stream>>>(
                                (const vec2_t<S>*)U_in, (vec2_t<S>*)U_out, (const S*)E, L->g, (S)df));
    } else {
        LQCD_DISPATCH_DS(L, k_drift<D, S, S, DRIFT_ADDITIVE><<<link_grid(L), kBlock, 0, L->ctx->stream>>>(
                                (const vec2_t<S>*)U_in, (vec2_t<S>*)U_out, (const S*)E, L->g, (S)df));
    }
    prof_end(L->ctx, &ev);
    return hipGetLastError();
}

} // namespace lqcd
