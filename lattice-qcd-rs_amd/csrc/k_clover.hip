// k_clover.hip -- clover, field-strength tensor and chromomagnetic field of
// every site (LinkMatrix::clover / f_mu_nu / magnetic_field_vec,
// src/field.rs:806-868), in double arithmetic on either storage type.
//
// The output is in the host layout of the reference (18 doubles per matrix,
// column-major 3x3 complex): [site][i][j] for the clover and F (stores
// staged through LDS), [site][dir] for B.  F takes one thread per (site,
// pair i < j) (k_f_mu_nu_blocks); the clover and B one thread per site.
#include "dispatch.hpp"
#include "fields.hpp"

namespace lqcd {

// neighbour of s one step along +d or -d (periodic)
__device__ __forceinline__ int64_t site_step(const Geom& g, int64_t s, int d, bool pos)
{
    const int64_t L = g.extent[d], st = g.stride[d];
    const int64_t x = (s / st) % L;
    if (pos) return x == L - 1 ? s - (L - 1) * st : s + st;
    return x == 0 ? s + (L - 1) * st : s - st;
}

__device__ __forceinline__ void adjoint(const mat3<double>& a, mat3<double>& b)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) b.e[i * 3 + j] = {a.e[j * 3 + i].re, -a.e[j * 3 + i].im};
}

// LinkMatrix::matrix (field.rs:726-740): U_d(y) for +d, U_d(y - d)^+ for -d
template <typename S>
__device__ __forceinline__ mat3<double> link_signed(const vec2_t<S>* __restrict__ U, const Geom& g, int64_t y,
                                                   int d, bool pos)
{
    if (pos) return load_link<double, S>(U, g.npad, d, y);
    const mat3<double> m = load_link<double, S>(U, g.npad, d, site_step(g, y, d, false));
    mat3<double> r;
    adjoint(m, r);
    return r;
}

// LinkMatrix::pij (field.rs:743-771) for signed directions a, b:
//   S_ab(x) = U_b(x) U_a(x+b) U_b(x+a)^+,  P_ab(x) = U_a(x) S_ab(x)^+
template <typename S>
__device__ __forceinline__ void pij_signed(const vec2_t<S>* __restrict__ U, const Geom& g, int64_t s, int a,
                                           bool pa, int b, bool pb, mat3<double>& p)
{
    const int64_t xb = site_step(g, s, b, pb), xa = site_step(g, s, a, pa);
    mat3<double> t, sab;
    {
        const mat3<double> ub = link_signed<S>(U, g, s, b, pb);
        const mat3<double> ua_b = link_signed<S>(U, g, xb, a, pa);
        mul<false, false>(ub, ua_b, t);
    }
    {
        const mat3<double> ub_a = link_signed<S>(U, g, xa, b, pb);
        mul<false, true>(t, ub_a, sab);
    }
    const mat3<double> ua = link_signed<S>(U, g, s, a, pa);
    mul<false, true>(ua, sab, p);
}

// LinkMatrix::clover (field.rs:806-820):
//   P(i, j) + P(j, -i) + P(-i, -j) + P(-j, i)
template <typename S>
__device__ __forceinline__ void clover_at(const vec2_t<S>* __restrict__ U, const Geom& g, int64_t s, int i,
                                          int j, mat3<double>& c)
{
    mat3<double> p;
    pij_signed<S>(U, g, s, i, true, j, true, c);
    pij_signed<S>(U, g, s, j, true, i, false, p);
#pragma unroll
    for (int k = 0; k < 9; ++k) c.e[k] = {c.e[k].re + p.e[k].re, c.e[k].im + p.e[k].im};
    pij_signed<S>(U, g, s, i, false, j, false, p);
#pragma unroll
    for (int k = 0; k < 9; ++k) c.e[k] = {c.e[k].re + p.e[k].re, c.e[k].im + p.e[k].im};
    pij_signed<S>(U, g, s, j, false, i, true, p);
#pragma unroll
    for (int k = 0; k < 9; ++k) c.e[k] = {c.e[k].re + p.e[k].re, c.e[k].im + p.e[k].im};
}

// Link loaders: U_d(y) through a resource over the 9 planes of direction d
// (uniform d), or through one resource over all 9 D planes (d may differ
// between the lanes of a wave; the planes must span less than 4 GiB).
template <typename S>
struct DirLinks {
    const vec2_t<S>* __restrict__ U;
    int64_t npad;
    __device__ __forceinline__ mat3<double> operator()(int d, int64_t y) const
    {
        return load_link<double, S>(U, npad, d, y);
    }
};

template <typename S, int D>
struct AllLinks {
    const vec2_t<S>* __restrict__ U;
    int64_t npad;
    __device__ __forceinline__ mat3<double> operator()(int d, int64_t y) const
    {
        using E = vec2_t<S>;
        const __amdgpu_buffer_rsrc_t r = planes_rsrc(U, npad, 9 * D);
        const uint32_t plane = (uint32_t)npad * (uint32_t)sizeof(E);
        const uint32_t off = ((uint32_t)d * 9u * (uint32_t)npad + (uint32_t)y) * (uint32_t)sizeof(E);
        mat3<double> m;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const E v = buf_load<E>(r, off, (uint32_t)k * plane);
            m.e[k] = {double(v.x), double(v.y)};
        }
        return m;
    }
};

// One leaf of the clover in the form of pij: S = op(ub) op(uab) op(uba)^+,
// P = op(ua) S^+, where op(.) is the adjoint when the template flag is set
// and each matrix is the link U_d(site) named by its arguments.
template <typename LD, bool DUB, bool DUAB, bool DUBA, bool DUA>
__device__ __forceinline__ void leaf(const LD& ld, int db, int64_t sb, int dab, int64_t sab, int dba, int64_t sba,
                                     int da, int64_t sa, mat3<double>& p)
{
    mat3<double> t, st;
    {
        const mat3<double> ub = ld(db, sb);
        const mat3<double> uab = ld(dab, sab);
        mul<DUB, DUAB>(ub, uab, t);
    }
    {
        const mat3<double> uba = ld(dba, sba);
        mul<false, !DUBA>(t, uba, st);
    }
    const mat3<double> ua = ld(da, sa);
    mul<DUA, true>(ua, st, p);
}

// clover for i != j with the site's neighbour offsets: an offset along one
// direction depends only on that coordinate, so offsets along different
// directions add.  The four leaves are pij(i, j), pij(j, -i), pij(-i, -j),
// pij(-j, i) with U_{-d}(y) = U_d(y - d)^+ (field.rs:726-771):
//   P(i,j):   ub = U_j(x),      uab = U_i(x+j),      uba = U_j(x+i),      ua = U_i(x)
//   P(j,-i):  ub = U_i(x-i)^+,  uab = U_j(x-i),      uba = U_i(x+j-i)^+,  ua = U_j(x)
//   P(-i,-j): ub = U_j(x-j)^+,  uab = U_i(x-j-i)^+,  uba = U_j(x-i-j)^+,  ua = U_i(x-i)^+
//   P(-j,i):  ub = U_i(x),      uab = U_j(x+i-j)^+,  uba = U_i(x-j),      ua = U_j(x-j)^+
template <int D, typename LD>
__device__ __forceinline__ void clover_offdiag(const LD& ld, int64_t s, const Neighbours<D>& nb, int i, int j,
                                               mat3<double>& c)
{
    // the offsets for directions i, j by selects (i, j may differ between
    // lanes: an indexed read of nb would go through scratch memory)
    int32_t ui = 0, uj = 0, di = 0, dj = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        ui = d == i ? nb.up[d] : ui;
        uj = d == j ? nb.up[d] : uj;
        di = d == i ? nb.dn[d] : di;
        dj = d == j ? nb.dn[d] : dj;
    }
    const int64_t pi = s + ui, pj = s + uj, mi = s + di, mj = s + dj;
    const int64_t pj_mi = pj + di, mi_mj = mi + dj, pi_mj = pi + dj;
    // one leaf at a time: without the scheduling fences the loads of all
    // four leaves are issued up front and the kernel needs 250 VGPRs
    mat3<double> p;
    leaf<LD, false, false, false, false>(ld, j, s, i, pj, j, pi, i, s, c);
    __builtin_amdgcn_sched_barrier(0);
    leaf<LD, true, false, true, false>(ld, i, mi, j, mi, i, pj_mi, j, s, p);
#pragma unroll
    for (int k = 0; k < 9; ++k) c.e[k] = {c.e[k].re + p.e[k].re, c.e[k].im + p.e[k].im};
    __builtin_amdgcn_sched_barrier(0);
    leaf<LD, true, true, true, true>(ld, j, mj, i, mi_mj, j, mi_mj, i, mi, p);
#pragma unroll
    for (int k = 0; k < 9; ++k) c.e[k] = {c.e[k].re + p.e[k].re, c.e[k].im + p.e[k].im};
    __builtin_amdgcn_sched_barrier(0);
    leaf<LD, false, true, false, true>(ld, i, s, j, pi_mj, i, mj, j, mj, p);
#pragma unroll
    for (int k = 0; k < 9; ++k) c.e[k] = {c.e[k].re + p.e[k].re, c.e[k].im + p.e[k].im};
}

// 18 doubles, column-major, at out
__device__ __forceinline__ void store_host_matrix(double* __restrict__ out, const mat3<double>& m)
{
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r)
            *reinterpret_cast<double2*>(out + 2 * (c * 3 + r)) = make_double2(m.e[r * 3 + c].re, m.e[r * 3 + c].im);
}

__device__ __forceinline__ void add_host_matrix(double* __restrict__ out, const mat3<double>& m)
{
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            double2* q = reinterpret_cast<double2*>(out + 2 * (c * 3 + r));
            const double2 v = *q;
            *q = make_double2(v.x + m.e[r * 3 + c].re, v.y + m.e[r * 3 + c].im);
        }
}

// levi_civita(d, i, j) for i < j and d outside {i, j} (src/utils.rs:306-318):
// sign(i - d) * sign(j - d)
__device__ __forceinline__ double levi_dij(int d, int i, int j)
{
    return (i > d) == (j > d) ? 1.0 : -1.0;
}

// Matrix m of every lane of a wave to out[(site * nslot + slot) * 18], the
// wave's sites being s0 .. s0 + 63 (sites >= nsite are skipped).  The matrices
// pass through LDS so that each store instruction writes runs of whole
// 144-byte matrices of consecutive sites (about seven per instruction) instead
// of 64 pieces 144 * nslot bytes apart.  Every thread of the block calls it.
__device__ __forceinline__ void store_slot_staged(double2* __restrict__ stage, double* __restrict__ out, int64_t s0,
                                                  int64_t nsite, int nslot, int slot, const mat3<double>& m)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r)
            stage[lane * 9 + c * 3 + r] = make_double2(m.e[r * 3 + c].re, m.e[r * 3 + c].im);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int ch = k * 64 + lane;
        const int sl = ch / 9, piece = ch - sl * 9;
        const int64_t site = s0 + sl;
        if (site < nsite) reinterpret_cast<double2*>(out + (site * nslot + slot) * 18)[piece] = stage[ch];
    }
    __syncthreads();
}

// magnetic_field_vec of every site (thread per site).  inv = 1 / (8 a^2).
// Pairs i < j are computed once: clover(j, i) = clover(i, j)^+ (each leaf of
// one is the adjoint of a leaf of the other), so F_ij = (C - C^+) / (8 a^2),
// and B_d = sum_{i,j} eps(d,i,j) F_ij / (2i), which takes each unordered
// pair twice with the same sign, is the sum of eps(d,i,j) F_ij / i.
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_magnetic_field(const vec2_t<S>* __restrict__ U, Geom g,
                                                           double* __restrict__ out, double inv)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= g.nsite) return;
    const mat3<double> zero = {};
    const Neighbours<D> nb = neighbours<D>(g, s);
    for (int d = 0; d < D; ++d) store_host_matrix(out + 18 * (s * D + d), zero);
    for (int i = 0; i < D; ++i)
        for (int j = i + 1; j < D; ++j) {
            mat3<double> c, ca, f;
            clover_offdiag<D>(DirLinks<S>{U, g.npad}, s, nb, i, j, c);
            adjoint(c, ca);
#pragma unroll
            for (int k = 0; k < 9; ++k) f.e[k] = {(c.e[k].re - ca.e[k].re) * inv, (c.e[k].im - ca.e[k].im) * inv};
            for (int d = 0; d < D; ++d) {
                if (d == i || d == j) continue;
                const double sg = levi_dij(d, i, j);
                mat3<double> b;   // sg * F / i = sg * (-i F)
#pragma unroll
                for (int k = 0; k < 9; ++k) b.e[k] = {sg * f.e[k].im, -sg * f.e[k].re};
                add_host_matrix(out + 18 * (s * D + d), b);
            }
        }
}

// clover (MODE 0) or f_mu_nu (MODE 1) of every site, one thread per site,
// inv = 1 / (8 a^2) (the D (D + 1) / 2 clovers of a site share their link loads in L1 and L2;
// a thread per site and pair is 30 % slower).  clover(j, i) = clover(i, j)^+,
// so each pair i < j is computed once and F_ji = -F_ij with
// F_ij = (C - C^+) / (8 a^2); F_ii = 0 as clover(i, i) - clover(i, i) is in
// the reference.
template <int D, typename S, int MODE>
__global__ __launch_bounds__(kBlock) void k_clover_fields(const vec2_t<S>* __restrict__ U, Geom g,
                                                         double* __restrict__ out, double inv)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    // every thread reaches the staged stores; threads past the lattice
    // compute nothing and store nothing
    __shared__ double2 stage_all[kBlock * 9];
    double2* stage = stage_all + (threadIdx.x >> 6) * 64 * 9;
    const int64_t s0 = s - (threadIdx.x & 63);
    const bool active = s < g.nsite;
    const Neighbours<D> nb = neighbours<D>(g, active ? s : 0);
    constexpr int NS = D * D;
    const mat3<double> zero = {};
    for (int i = 0; i < D; ++i) {
        mat3<double> dg = zero;
        if (MODE == 0 && active) clover_at<S>(U, g, s, i, i, dg);
        store_slot_staged(stage, out, s0, g.nsite, NS, i * D + i, dg);
        for (int j = i + 1; j < D; ++j) {
            mat3<double> c = zero, ca;
            if (active) clover_offdiag<D>(DirLinks<S>{U, g.npad}, s, nb, i, j, c);
            adjoint(c, ca);
            if (MODE == 1) {
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    const cplx<double> f = {(c.e[k].re - ca.e[k].re) * inv, (c.e[k].im - ca.e[k].im) * inv};
                    c.e[k] = f;
                    ca.e[k] = {-f.re, -f.im};
                }
            }
            store_slot_staged(stage, out, s0, g.nsite, NS, i * D + j, c);
            store_slot_staged(stage, out, s0, g.nsite, NS, j * D + i, ca);
        }
    }
}

// F_mu_nu of every site (LinkMatrix::f_mu_nu, field.rs:822-834) with the
// output written in whole 128-byte lines.  A workgroup owns 32 consecutive
// sites and has one thread per (site, pair i < j): 32 D (D - 1) / 2 threads,
// the pair fixed per half wavefront (so a load instruction reads 32
// consecutive sites; the link direction differs between the halves, hence the
// resource over all planes).  Each thread forms F_ij = (C_ij - C_ij^+) / (8 a^2)
// and writes F_ij and F_ji = -F_ij into an LDS image of the sites' output
// (site-major, the host layout); the diagonal slots are zero.  The image is
// written out in two halves of the D^2 slots (8 of 16 slots, 1,152 B = nine
// lines per site for D = 4), each a run of whole lines per site.  (One row of
// x_0 per workgroup in block order: handing XCDs contiguous block ranges was
// slower here.)
// the output is written once and read by the host or a later kernel: its
// stores are marked non-temporal, so that the L2 keeps the links
using d2v = double __attribute__((ext_vector_type(2)));

template <int D>
struct FPairs {
    static constexpr int n = D * (D - 1) / 2;
    static constexpr int threads = 32 * n;
    static constexpr int slots = D * D;
    static constexpr int half0 = slots / 2;                 // slots [0, half0), then [half0, slots)
    static constexpr int stage_slots = slots - half0;      // the larger half
    static constexpr int site_doubles = stage_slots * 18 + 2;  // padded: lanes of one half wave spread over banks
};

template <int D>
__device__ __forceinline__ void pair_of(int p, int& i, int& j)
{
    i = 0;
    int k = p;
    while (k >= D - 1 - i) {
        k -= D - 1 - i;
        ++i;
    }
    j = i + 1 + k;
}

template <int D, typename S>
__global__ __launch_bounds__(FPairs<D>::threads) void k_f_mu_nu_blocks(const vec2_t<S>* __restrict__ U, Geom g,
                                                                      double* __restrict__ out, double inv)
{
    using P = FPairs<D>;
    __shared__ double stage[32 * P::site_doubles];
    const int t = threadIdx.x, pr = t >> 5, sl = t & 31;
    const int64_t s0 = (int64_t)blockIdx.x * 32;
    const int64_t s = s0 + sl;
    const int nsl = g.nsite - s0 < 32 ? (int)(g.nsite - s0) : 32;
    int i, j;
    pair_of<D>(pr, i, j);
    mat3<double> f = {};
    if (sl < nsl) {
        const Neighbours<D> nb = neighbours<D>(g, s);
        mat3<double> c;
        clover_offdiag<D>(AllLinks<S, D>{U, g.npad}, s, nb, i, j, c);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const cplx<double> a = c.e[r * 3 + q], b = c.e[q * 3 + r];
                f.e[r * 3 + q] = {(a.re - b.re) * inv, (a.im + b.im) * inv};
            }
    }
    double* const mine = stage + sl * P::site_doubles;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int lo = h == 0 ? 0 : P::half0, hi = h == 0 ? P::half0 : P::slots;
        // F_ij, F_ji = -F_ij and the diagonal zeros of this half
        auto put = [&](int slot, double sign) __attribute__((always_inline)) {
            if (slot < lo || slot >= hi) return;
            double* m = mine + (slot - lo) * 18;
#pragma unroll
            for (int q = 0; q < 3; ++q)
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    m[2 * (q * 3 + r)] = sign * f.e[r * 3 + q].re;
                    m[2 * (q * 3 + r) + 1] = sign * f.e[r * 3 + q].im;
                }
        };
        put(i * D + j, 1.0);
        put(j * D + i, -1.0);
        for (int d = pr; d < D; d += P::n) {
            const int slot = d * D + d;
            if (slot >= lo && slot < hi)
#pragma unroll
                for (int k = 0; k < 18; ++k) mine[(slot - lo) * 18 + k] = 0.0;
        }
        __syncthreads();
        // the sites' runs of (hi - lo) matrices, 9 double2 each
        const int per_site = (hi - lo) * 9;
        const int total = nsl * per_site;
        for (int q = t; q < total; q += P::threads) {
            const int site = q / per_site, r = q - site * per_site;
            const double2 v = *reinterpret_cast<const double2*>(stage + site * P::site_doubles + 2 * r);
            __builtin_nontemporal_store(d2v{v.x, v.y}, reinterpret_cast<d2v*>(out + ((s0 + site) * P::slots + lo) * 18) + r);
        }
        __syncthreads();
    }
}

// LinkMatrix::sij (WHICH 0) or pij (WHICH 1) (field.rs:743-771) at every site
// for the signed directions (a, pa), (b, pb), in the host layout
template <int D, typename S, int WHICH>
__global__ __launch_bounds__(kBlock) void k_plaquette_field(const vec2_t<S>* __restrict__ U, Geom g, int a,
                                                            bool pa, int b, bool pb, double* __restrict__ out)
{
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= g.nsite) return;
    const int64_t xb = site_step(g, s, b, pb), xa = site_step(g, s, a, pa);
    mat3<double> t, sab;
    mul<false, false>(link_signed<S>(U, g, s, b, pb), link_signed<S>(U, g, xb, a, pa), t);
    mul<false, true>(t, link_signed<S>(U, g, xa, b, pb), sab);
    if (WHICH == 0) {
        store_host_matrix(out + 18 * s, sab);
    } else {
        mat3<double> p;
        mul<false, true>(link_signed<S>(U, g, s, a, pa), sab, p);
        store_host_matrix(out + 18 * s, p);
    }
}

hipError_t launch_plaquette_field(const lqcd_lattice* L, int which, int a, bool pa, int b, bool pb, double* d_out)
{
    const unsigned grid = grid_for(L->g.nsite);
    if (which == 0)
        LQCD_DISPATCH_DS(L, k_plaquette_field<D, S, 0><<<grid, kBlock, 0, L->ctx->stream>>>(
                                (const vec2_t<S>*)L->U[L->cur], L->g, a, pa, b, pb, d_out));
    else
        LQCD_DISPATCH_DS(L, k_plaquette_field<D, S, 1><<<grid, kBlock, 0, L->ctx->stream>>>(
                                (const vec2_t<S>*)L->U[L->cur], L->g, a, pa, b, pb, d_out));
    return hipGetLastError();
}

hipError_t launch_clover_fields(const lqcd_lattice* L, int mode, double* d_out)
{
    const unsigned grid = grid_for(L->g.nsite);
    const double inv = 1.0 / (8.0 * L->size * L->size);
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_FIELDS, &ev);
    switch (mode) {
    case 0:
        LQCD_DISPATCH_DS(L, k_clover_fields<D, S, 0><<<grid, kBlock, 0, L->ctx->stream>>>(
                                (const vec2_t<S>*)L->U[L->cur], L->g, d_out, inv));
        break;
    case 1:
        if ((uint64_t)L->g.npad * 9 * L->g.dim * L->elem <= kMaxPlaneBytes) {
            const unsigned nb = (unsigned)((L->g.nsite + 31) / 32);
            LQCD_DISPATCH_DS(L, k_f_mu_nu_blocks<D, S><<<nb, FPairs<D>::threads, 0, L->ctx->stream>>>(
                                    (const vec2_t<S>*)L->U[L->cur], L->g, d_out, inv));
        } else {
            LQCD_DISPATCH_DS(L, k_clover_fields<D, S, 1><<<grid, kBlock, 0, L->ctx->stream>>>(
                                    (const vec2_t<S>*)L->U[L->cur], L->g, d_out, inv));
        }
        break;
    default:
        LQCD_DISPATCH_DS(L, k_magnetic_field<D, S><<<grid, kBlock, 0, L->ctx->stream>>>(
                                (const vec2_t<S>*)L->U[L->cur], L->g, d_out, inv));
        break;
    }
    prof_end(L->ctx, &ev);
    return hipGetLastError();
}

} // namespace lqcd
