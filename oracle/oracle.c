/*
 * oracle.c -- CPU restatement (C, f64) of lattice-qcd-rs's gauge-update path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  The product
 * (the HIP library behind include/lqcd.h) never links or calls it.
 *
 * The reference (Rust, nalgebra 0.31 / num-complex / rand 0.8.5) cannot be
 * built here (no Rust toolchain, no crates), so this file restates its
 * algorithms line by line.  Floating-point operation order follows the Rust
 * source where it is visible (no FMA contraction: built with
 * -ffp-contract=off); nalgebra internals (matrix product, determinant, norm,
 * cross, dot) are restated from that crate's published generic code.  The
 * summation order of the reference's parallel reductions (rayon par_bridge) is
 * nondeterministic; here sums run in storage order.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   geometry          reference doctests (lattice.rs:62-97,256-323), iterator
 *                     counts (test/mod.rs:21-41), exhaustive bijection/involution
 *   su3 exp           closed forms (test/mod.rs:96-146), scipy expm on random
 *                     su(3) (test/mod.rs:148-190), generator algebra (:231-250)
 *   orthonormalize    exact cases (test/mod.rs:455-567), SU(3) property
 *   plaquette / H_U   clover known answers (field.rs:1579-1711), U(1)-embedded
 *                     closed form, np.roll brute force, cold state = 3
 *   force             central finite differences of H_U, staple identity,
 *                     cold state force = 0
 *   derivative_u      closed form E*U*i*sqrt(6)/a
 *   MD / HMC          dt^2 scaling of |dH|, reversibility, cold invariance
 *                     (test/mod.rs:429-453), |dH| bound of test/mod.rs:678-698
 *   Metropolis        dS == dH brute force (metropolis_hastings.rs:479-514),
 *                     checkerboard order independence
 *   Bernoulli rule    rand 0.8.5 integer threshold, known values
 *   Gauss projection  gauss_sum_div below threshold after projection
 *   statistics        the reference's own known answers (statistics/mod.rs
 *                     doctests and mean_var test), numpy mean / var(ddof=1) /
 *                     population covariance, one-pass against two-pass forms
 * Parity unpinned (no reference execution possible): the rounding-level
 * behaviour of nalgebra/num-complex internals, and every random stream
 * (StdRng/ChaCha12 and rand_distr's ziggurat are replaced by xoshiro256++ and
 * Box-Muller; inputs are fed identically to oracle and device instead).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CA 3.0
#define F64_EPS DBL_EPSILON

static int g_threads = 0;
void or_set_threads(int n) { g_threads = n; }
int or_get_threads(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}
#define OMP_THREADS num_threads(or_get_threads())

/* ------------------------------------------------------------------------ */
/* complex arithmetic: num-complex's formulas                                */
/* ------------------------------------------------------------------------ */
static inline or_cplx cx(double re, double im) { or_cplx c = {re, im}; return c; }
static inline or_cplx cadd(or_cplx a, or_cplx b) { return cx(a.re + b.re, a.im + b.im); }
static inline or_cplx csub(or_cplx a, or_cplx b) { return cx(a.re - b.re, a.im - b.im); }
/* num-complex Mul: (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re) */
static inline or_cplx cmul(or_cplx a, or_cplx b)
{
    return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline or_cplx cconj(or_cplx a) { return cx(a.re, -a.im); }
static inline or_cplx cscale_div(or_cplx a, double d) { return cx(a.re / d, a.im / d); }
static const or_cplx C_ZERO = {0.0, 0.0};
static const or_cplx C_ONE = {1.0, 0.0};
static const or_cplx C_I = {0.0, 1.0};

#define M(mat, row, col) ((mat)->m[(col) * 3 + (row)])

static inline void mzero(or_mat3* a) { memset(a, 0, sizeof(*a)); }
static inline void mident(or_mat3* a)
{
    mzero(a);
    M(a, 0, 0) = C_ONE; M(a, 1, 1) = C_ONE; M(a, 2, 2) = C_ONE;
}
static inline void madd(const or_mat3* a, const or_mat3* b, or_mat3* c)
{
    for (int k = 0; k < 9; ++k) c->m[k] = cadd(a->m[k], b->m[k]);
}
static inline void msub(const or_mat3* a, const or_mat3* b, or_mat3* c)
{
    for (int k = 0; k < 9; ++k) c->m[k] = csub(a->m[k], b->m[k]);
}
/* Matrix * Complex scalar: each entry times s (entry on the left) */
static inline void mscal(const or_mat3* a, or_cplx s, or_mat3* c)
{
    for (int k = 0; k < 9; ++k) c->m[k] = cmul(a->m[k], s);
}

/* nalgebra generic gemm: column j of C = A[:,0]*B[0,j], then += A[:,1]*B[1,j],
 * then += A[:,2]*B[2,j] (beta = 0 then 1, alpha = 1). */
void or_matmul(const or_mat3* a, const or_mat3* b, or_mat3* c)
{
    or_mat3 r;
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            or_cplx acc = cmul(M(a, i, 0), M(b, 0, j));
            acc = cadd(cmul(M(a, i, 1), M(b, 1, j)), acc);
            acc = cadd(cmul(M(a, i, 2), M(b, 2, j)), acc);
            M(&r, i, j) = acc;
        }
    *c = r;
}
void or_adj(const or_mat3* a, or_mat3* c)
{
    or_mat3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M(&r, i, j) = cconj(M(a, j, i));
    *c = r;
}
static inline or_cplx mtrace(const or_mat3* a)
{
    or_cplx t = C_ZERO;
    t = cadd(t, M(a, 0, 0));
    t = cadd(t, M(a, 1, 1));
    t = cadd(t, M(a, 2, 2));
    return t;
}
/* nalgebra determinant, 3x3 branch (cofactor expansion along row 0) */
or_cplx or_det(const or_mat3* m)
{
    or_cplx m11 = M(m, 0, 0), m12 = M(m, 0, 1), m13 = M(m, 0, 2);
    or_cplx m21 = M(m, 1, 0), m22 = M(m, 1, 1), m23 = M(m, 1, 2);
    or_cplx m31 = M(m, 2, 0), m32 = M(m, 2, 1), m33 = M(m, 2, 2);
    or_cplx minor_m12_m23 = csub(cmul(m22, m33), cmul(m32, m23));
    or_cplx minor_m11_m23 = csub(cmul(m21, m33), cmul(m31, m23));
    or_cplx minor_m11_m22 = csub(cmul(m21, m32), cmul(m31, m22));
    return cadd(csub(cmul(m11, minor_m12_m23), cmul(m12, minor_m11_m23)),
                cmul(m13, minor_m11_m22));
}

static inline void load_mat(const double* p, or_mat3* m) { memcpy(m, p, sizeof(*m)); }
static inline void store_mat(double* p, const or_mat3* m) { memcpy(p, m, sizeof(*m)); }

/* ------------------------------------------------------------------------ */
/* geometry: src/lattice.rs                                                  */
/* ------------------------------------------------------------------------ */

/* LatticeCyclic::new, src/lattice.rs:190-201 (one extent per direction here) */
int or_lattice_init(or_lattice* l, int dim, const size_t* extent, double size)
{
    memset(l, 0, sizeof(*l));
    if (dim == 0) return -1;                                   /* ZeroDimension */
    if (dim < 0 || dim > OR_MAX_DIM) return -4;
    if (size <= 0.0 || isnan(size) || isinf(size)) return -2; /* NonPositiveSize */
    for (int i = 0; i < dim; ++i)
        if (extent[i] < 2) return -3;                           /* DimTooSmall */
    l->dim = dim;
    l->size = size;
    size_t s = 1;
    for (int i = 0; i < dim; ++i) {
        l->extent[i] = extent[i];
        l->stride[i] = s;
        s *= extent[i];
    }
    l->nsite = s;
    return 0;
}

/* LatticePoint::to_index, src/lattice.rs:909-915: sum (x_i mod L) * L^i */
size_t or_site_index(const or_lattice* l, const size_t* x)
{
    size_t idx = 0;
    for (int i = 0; i < l->dim; ++i) idx += (x[i] % l->extent[i]) * l->stride[i];
    return idx;
}

void or_site_coords(const or_lattice* l, size_t idx, size_t* x)
{
    for (int i = 0; i < l->dim; ++i) {
        x[i] = idx % l->extent[i];
        idx /= l->extent[i];
    }
}

/* LatticeCyclic::add_point_direction(_n) with n = 1, src/lattice.rs:273-323 */
size_t or_neighbor(const or_lattice* l, size_t site, int dir, int positive)
{
    size_t x[OR_MAX_DIM];
    or_site_coords(l, site, x);
    size_t L = l->extent[dir];
    const size_t shift = 1 % L;
    if (positive) {
        x[dir] = (x[dir] + shift) % L;
    } else if (x[dir] < shift) {
        x[dir] = L - (shift - x[dir]);
    } else {
        x[dir] = (x[dir] - shift) % L;
    }
    return or_site_index(l, x);
}

int or_site_parity(const or_lattice* l, size_t site)
{
    size_t x[OR_MAX_DIM];
    or_site_coords(l, site, x);
    size_t s = 0;
    for (int i = 0; i < l->dim; ++i) s += x[i];
    return (int)(s & 1u);
}

/* LatticeCyclic::link_canonical + LatticeLinkCanonical::to_index,
 * src/lattice.rs:84-97 and :923-927 */
size_t or_link_canonical(const or_lattice* l, size_t site, int dir, int positive)
{
    size_t s = positive ? site : or_neighbor(l, site, dir, 0);
    return s * (size_t)l->dim + (size_t)dir;
}

/* IteratorLatticeLinkCanonical::next, src/lattice.rs:381-419: direction
 * fastest, then the point iterator (x_0 fastest, src/lattice.rs:685-719). */
void or_link_order(const or_lattice* l, size_t* out_site, int* out_dir)
{
    size_t k = 0;
    size_t x[OR_MAX_DIM] = {0};
    for (;;) {
        for (int d = 0; d < l->dim; ++d) {
            out_site[k] = or_site_index(l, x);
            out_dir[k] = d;
            ++k;
        }
        /* IteratorLatticePoint::next */
        x[0] += 1;
        int done = 0;
        for (int i = 0; i < l->dim; ++i) {
            while (x[i] >= l->extent[i]) {
                if (i < l->dim - 1) {
                    x[i + 1] += 1;
                } else {
                    done = 1;
                    break;
                }
                x[i] -= l->extent[i];
            }
            if (done) break;
        }
        if (done) break;
    }
}

/* ------------------------------------------------------------------------ */
/* su(3): src/su3.rs, src/field.rs:29-205                                   */
/* ------------------------------------------------------------------------ */

/* GENERATORS, src/su3.rs:24-194 (Gell-Mann / 2; constants as in :179-181) */
static const double MINUS_ONE_OVER_SQRT_3 = -0.5773502691896258;
static const double ONE_OVER_2_SQRT_3 = 0.2886751345948129;

void or_generator(int a, or_mat3* g)
{
    mzero(g);
    switch (a) {
    case 0: M(g, 0, 1) = cx(0.5, 0); M(g, 1, 0) = cx(0.5, 0); break;
    case 1: M(g, 0, 1) = cx(0, -0.5); M(g, 1, 0) = cx(0, 0.5); break;
    case 2: M(g, 0, 0) = cx(0.5, 0); M(g, 1, 1) = cx(-0.5, 0); break;
    case 3: M(g, 0, 2) = cx(0.5, 0); M(g, 2, 0) = cx(0.5, 0); break;
    case 4: M(g, 0, 2) = cx(0, -0.5); M(g, 2, 0) = cx(0, 0.5); break;
    case 5: M(g, 1, 2) = cx(0.5, 0); M(g, 2, 1) = cx(0.5, 0); break;
    case 6: M(g, 1, 2) = cx(0, -0.5); M(g, 2, 1) = cx(0, 0.5); break;
    case 7:
        M(g, 0, 0) = cx(ONE_OVER_2_SQRT_3, 0);
        M(g, 1, 1) = cx(ONE_OVER_2_SQRT_3, 0);
        M(g, 2, 2) = cx(MINUS_ONE_OVER_SQRT_3, 0);
        break;
    default: break;
    }
}

static or_mat3 G_TAB[8];
static or_mat3 G_SUM; /* GENERATORS.iter().sum(), used by gauss_sum_div */
static pthread_once_t g_tables_once = PTHREAD_ONCE_INIT;
static void fill_tables(void)
{
    or_mat3 s;
    mzero(&s);
    for (int a = 0; a < 8; ++a) {
        or_generator(a, &G_TAB[a]);
        madd(&s, &G_TAB[a], &s);
    }
    G_SUM = s;
}
/* called from inside OpenMP loops: initialisation must be thread-safe */
static void init_tables(void) { pthread_once(&g_tables_once, fill_tables); }

/* Su3Adjoint::to_matrix, src/field.rs:106-112: sum over a of G_a * (v_a + 0i),
 * accumulated from the zero matrix in generator order. */
void or_adjoint_to_matrix(const double* v, or_mat3* out)
{
    init_tables();
    or_mat3 acc, t;
    mzero(&acc);
    for (int a = 0; a < 8; ++a) {
        mscal(&G_TAB[a], cx(v[a], 0.0), &t);
        madd(&acc, &t, &acc);
    }
    *out = acc;
}

/* Su3Adjoint::trace_squared, src/field.rs:164-167 */
double or_trace_squared(const double* v)
{
    double s = 0.0;
    for (int a = 0; a < 8; ++a) s += v[a] * v[a];
    return s / 2.0;
}

/* Su3Adjoint::d, src/field.rs:203-205: det(to_matrix) * i */
or_cplx or_adjoint_d(const double* v)
{
    or_mat3 m;
    or_adjoint_to_matrix(v, &m);
    return cmul(or_det(&m), C_I);
}

/* factorial_size_for_exp, src/su3.rs:720-732 */
unsigned long long or_factorial_size_for_exp(void)
{
    unsigned long long n = 7;
    unsigned __int128 fact = 1;
    while (1.0 / (double)fact >= F64_EPS) {
        n += 1;
        fact *= (unsigned __int128)(n - 7);
    }
    return n;
}

/* 1/k! with k! held as u128 and converted to f64 (src/su3.rs:735-817) */
static double inv_factorial(int k)
{
    unsigned __int128 f = 1;
    for (int i = 1; i <= k; ++i) f *= (unsigned __int128)i;
    return 1.0 / (double)f;
}

#define EXP_N 26 /* src/su3.rs:820 */

/* su3_exp_i, src/su3.rs:832-855: exp(i v^a T^a) by the Cayley-Hamilton
 * recursion over 1/k!, k = N-2 .. 0. */
void or_su3_exp_i(const double* v, or_mat3* out)
{
    const int n_loop = EXP_N - 1;
    or_cplx q0 = cx(inv_factorial(n_loop), 0.0);
    or_cplx q1 = C_ZERO, q2 = C_ZERO;
    or_cplx d = or_adjoint_d(v);
    or_cplx t = cx(-0.5 * or_trace_squared(v), 0.0); /* Su3Adjoint::t :185-187 */
    for (int i = n_loop - 1; i >= 0; --i) {
        or_cplx q0_n = cadd(cx(inv_factorial(i), 0.0), cmul(d, q2));
        or_cplx q1_n = cmul(C_I, csub(q0, cmul(t, q2)));
        or_cplx q2_n = cmul(C_I, q1);
        q0 = q0_n; q1 = q1_n; q2 = q2_n;
    }
    or_mat3 m, mm, a, b, r;
    or_adjoint_to_matrix(v, &m);
    /* CMatrix3::from_diagonal_element(q0) + m * q1 + m * m * q2 */
    mzero(&r);
    M(&r, 0, 0) = q0; M(&r, 1, 1) = q0; M(&r, 2, 2) = q0;
    mscal(&m, q1, &a);
    madd(&r, &a, &r);
    or_matmul(&m, &m, &mm);
    mscal(&mm, q2, &b);
    madd(&r, &b, out);
}

/* su3_exp_r, src/su3.rs:916-938: exp(v^a T^a) */
void or_su3_exp_r(const double* v, or_mat3* out)
{
    const int n_loop = EXP_N - 1;
    or_cplx q0 = cx(inv_factorial(n_loop), 0.0);
    or_cplx q1 = C_ZERO, q2 = C_ZERO;
    or_cplx d = or_adjoint_d(v);
    or_cplx t = cx(-0.5 * or_trace_squared(v), 0.0);
    for (int i = n_loop - 1; i >= 0; --i) {
        or_cplx q0_n = csub(cx(inv_factorial(i), 0.0), cmul(cmul(C_I, d), q2));
        or_cplx q1_n = csub(q0, cmul(t, q2));
        or_cplx q2_n = q1;
        q0 = q0_n; q1 = q1_n; q2 = q2_n;
    }
    or_mat3 m, mm, a, b, r;
    or_adjoint_to_matrix(v, &m);
    mzero(&r);
    M(&r, 0, 0) = q0; M(&r, 1, 1) = q0; M(&r, 2, 2) = q0;
    mscal(&m, q1, &a);
    madd(&r, &a, &r);
    or_matmul(&m, &m, &mm);
    mscal(&mm, q2, &b);
    madd(&r, &b, out);
}

/* ---- 3-vector helpers (nalgebra) ---- */
typedef struct { or_cplx v[3]; } vec3;

/* unrolled 3-element dot: (a0*b0 + a1*b1) + a2*b2; conj applied to a if asked */
static or_cplx vdot(const vec3* a, const vec3* b, int conj_a)
{
    or_cplx x0 = conj_a ? cconj(a->v[0]) : a->v[0];
    or_cplx x1 = conj_a ? cconj(a->v[1]) : a->v[1];
    or_cplx x2 = conj_a ? cconj(a->v[2]) : a->v[2];
    return cadd(cadd(cmul(x0, b->v[0]), cmul(x1, b->v[1])), cmul(x2, b->v[2]));
}
static double vnorm(const vec3* a) { return sqrt(vdot(a, a, 1).re); }
/* try_normalize(eps).unwrap_or(v): divide by the norm unless norm <= eps */
static vec3 vtry_normalize(const vec3* a)
{
    double n = vnorm(a);
    if (n <= F64_EPS) return *a;
    vec3 r;
    for (int i = 0; i < 3; ++i) r.v[i] = cscale_div(a->v[i], n);
    return r;
}
static vec3 vconj(const vec3* a)
{
    vec3 r;
    for (int i = 0; i < 3; ++i) r.v[i] = cconj(a->v[i]);
    return r;
}
/* nalgebra cross product */
static vec3 vcross(const vec3* a, const vec3* b)
{
    vec3 r;
    r.v[0] = csub(cmul(a->v[1], b->v[2]), cmul(a->v[2], b->v[1]));
    r.v[1] = csub(cmul(a->v[2], b->v[0]), cmul(a->v[0], b->v[2]));
    r.v[2] = csub(cmul(a->v[0], b->v[1]), cmul(a->v[1], b->v[0]));
    return r;
}

/* create_matrix_from_2_vector, src/su3.rs:279-287: columns v1, v2,
 * conj(v1) x conj(v2) */
static void create_matrix_from_2_vector(const vec3* v1, const vec3* v2, or_mat3* out)
{
    vec3 c1 = vconj(v1), c2 = vconj(v2);
    vec3 cr = vcross(&c1, &c2);
    for (int i = 0; i < 3; ++i) {
        M(out, i, 0) = v1->v[i];
        M(out, i, 1) = v2->v[i];
        M(out, i, 2) = cr.v[i];
    }
}

/* ortho_matrix_from_2_vector, src/su3.rs:290-295 */
static void ortho_matrix_from_2_vector(const vec3* v1, const vec3* v2, or_mat3* out)
{
    vec3 v1n = vtry_normalize(v1);
    vec3 c1 = vconj(&v1n);
    or_cplx s = vdot(&c1, v2, 0);
    vec3 v2t;
    for (int i = 0; i < 3; ++i) v2t.v[i] = csub(v2->v[i], cmul(v1n.v[i], s));
    vec3 v2n = vtry_normalize(&v2t);
    create_matrix_from_2_vector(&v1n, &v2n, out);
}

/* orthonormalize_matrix, src/su3.rs:299-303 (columns 0 and 1) */
void or_orthonormalize(const or_mat3* in, or_mat3* out)
{
    vec3 v1, v2;
    for (int i = 0; i < 3; ++i) {
        v1.v[i] = M(in, i, 0);
        v2.v[i] = M(in, i, 1);
    }
    or_mat3 r;
    ortho_matrix_from_2_vector(&v1, &v2, &r);
    *out = r;
}

/* ------------------------------------------------------------------------ */
/* random inputs                                                             */
/* ------------------------------------------------------------------------ */

/* xoshiro256++ seeded through SplitMix64 (Blackman & Vigna) */
static uint64_t splitmix64(uint64_t* x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
void or_rng_seed(or_rng* r, uint64_t seed)
{
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) r->s[i] = splitmix64(&x);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
uint64_t or_rng_next(or_rng* r)
{
    uint64_t* s = r->s;
    const uint64_t result = rotl(s[0] + s[3], 23) + s[0];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

/* rand 0.8 UniformFloat<f64>::sample: 52 random mantissa bits -> [1,2),
 * minus 1, times scale, plus low (scale = high - low for [-1,1) etc.) */
double or_rng_uniform(or_rng* r, double low, double high)
{
    uint64_t bits = (or_rng_next(r) >> 12) | 0x3FF0000000000000ull;
    double v12;
    memcpy(&v12, &bits, sizeof(v12));
    double v01 = v12 - 1.0;
    return v01 * (high - low) + low;
}

/* rand 0.8 Bernoulli: p == 1 always true, otherwise u64 < floor(p * 2^64) */
int or_bernoulli_accept(double p, uint64_t u)
{
    if (p >= 1.0) return 1;
    if (!(p > 0.0)) return 0;
    uint64_t p_int = (uint64_t)(p * 18446744073709551616.0);
    return u < p_int;
}
int or_rng_bernoulli(or_rng* r, double p)
{
    if (p >= 1.0) return 1; /* no draw, as rand 0.8 */
    return or_bernoulli_accept(p, or_rng_next(r));
}

/* f64 (x).exp().min(1).max(0): Rust min/max return the non-NaN operand */
double or_clamp_probability(double x) { return fmax(fmin(x, 1.0), 0.0); }

/* Box-Muller on two uniforms (stands in for rand_distr::Normal's ziggurat) */
double or_rng_normal(or_rng* r, double mean, double sigma)
{
    double u1 = 1.0 - or_rng_uniform(r, 0.0, 1.0); /* (0, 1] */
    double u2 = or_rng_uniform(r, 0.0, 1.0);
    return mean + sigma * sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* random_vec_3, src/su3.rs:350-355: re then im of each entry */
static vec3 random_vec_3(or_rng* r)
{
    vec3 v;
    for (int i = 0; i < 3; ++i) {
        double re = or_rng_uniform(r, -1.0, 1.0);
        double im = or_rng_uniform(r, -1.0, 1.0);
        v.v[i] = cx(re, im);
    }
    return v;
}

/* random_su3 / rand_su3_with_dis, src/su3.rs:322-347 */
void or_random_su3(or_rng* r, or_mat3* out)
{
    vec3 v1 = random_vec_3(r);
    while (vnorm(&v1) <= F64_EPS) v1 = random_vec_3(r);
    vec3 v2 = random_vec_3(r);
    for (;;) {
        or_cplx d = vdot(&v1, &v2, 0);
        if (hypot(d.re, d.im) > F64_EPS) break;
        v2 = random_vec_3(r);
    }
    ortho_matrix_from_2_vector(&v1, &v2, out);
}

/* su2::random_su2_close_to_unity, src/su2.rs:80-100, with
 * complex_matrix_from_vec :134-140 and PAULI_* :14-49.  PAULI_3 is the
 * identity in the reference (src/su2.rs:39-45); that is restated as is.
 * out = row-major 2x2 [m00, m01, m10, m11]. */
static void su2_close_to_unity_from(const double v[3], int positive, double spread, or_cplx out[4])
{
    double n = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    double x[3];
    for (int i = 0; i < 3; ++i) x[i] = (n <= F64_EPS ? v[i] : v[i] / n) * spread;
    double nsq = (x[0] * x[0] + x[1] * x[1]) + x[2] * x[2];
    double x0u = sqrt(1.0 - nsq);
    double x0 = positive ? x0u : -x0u;
    /* identity * x0 + sum_i PAULI_i * (0 + i x_i) */
    or_cplx m[4] = {cx(x0, 0), C_ZERO, C_ZERO, cx(x0, 0)};
    /* PAULI_1 = [[0,1],[1,0]] */
    or_cplx s1 = cx(0, x[0]);
    /* PAULI_2 = [[0,-i],[i,0]] */
    or_cplx s2 = cx(0, x[1]);
    /* PAULI_3 (reference quirk) = [[1,0],[0,1]] */
    or_cplx s3 = cx(0, x[2]);
    or_cplx p[4] = {C_ZERO, C_ZERO, C_ZERO, C_ZERO};
    p[1] = cadd(p[1], cmul(C_ONE, s1));
    p[2] = cadd(p[2], cmul(C_ONE, s1));
    p[1] = cadd(p[1], cmul(cx(0, -1), s2));
    p[2] = cadd(p[2], cmul(C_I, s2));
    p[0] = cadd(p[0], cmul(C_ONE, s3));
    p[3] = cadd(p[3], cmul(C_ONE, s3));
    for (int k = 0; k < 4; ++k) out[k] = cadd(m[k], p[k]);
}

void or_random_su2_close_to_unity(or_rng* r, double spread, or_cplx out[4])
{
    double v[3];
    for (int i = 0; i < 3; ++i) v[i] = or_rng_uniform(r, -1.0, 1.0);
    int positive = or_rng_bernoulli(r, 0.5);
    su2_close_to_unity_from(v, positive, spread, out);
}

/* get_r(a) get_s(b) get_t(c), adjoint when adj (src/su3.rs:385-398, :428-530) */
static void su3_from_su2(const or_cplx a[4], const or_cplx b[4], const or_cplx c[4], int adj,
                         or_mat3* out)
{
    or_mat3 R, S, T, RS, X;
    mzero(&R); mzero(&S); mzero(&T);
    M(&R, 0, 0) = a[0]; M(&R, 0, 1) = a[1]; M(&R, 1, 0) = a[2]; M(&R, 1, 1) = a[3];
    M(&R, 2, 2) = C_ONE;
    M(&S, 0, 0) = b[0]; M(&S, 0, 2) = b[1]; M(&S, 2, 0) = b[2]; M(&S, 2, 2) = b[3];
    M(&S, 1, 1) = C_ONE;
    M(&T, 0, 0) = C_ONE;
    M(&T, 1, 1) = c[0]; M(&T, 1, 2) = c[1]; M(&T, 2, 1) = c[2]; M(&T, 2, 2) = c[3];
    or_matmul(&R, &S, &RS);
    or_matmul(&RS, &T, &X);
    if (adj) or_adj(&X, &X);
    *out = X;
}

/* random_su3_close_to_unity, src/su3.rs:385-398, with get_r/s/t :428-530 */
void or_random_su3_close_to_unity(or_rng* r, double spread, or_mat3* out)
{
    or_cplx a[4], b[4], c[4];
    or_random_su2_close_to_unity(r, spread, a);
    or_random_su2_close_to_unity(r, spread, b);
    or_random_su2_close_to_unity(r, spread, c);
    su3_from_su2(a, b, c, or_rng_bernoulli(r, 0.5), out);
}

/* LinkMatrix::new_determinist, src/field.rs:646-659 (storage order) */
void or_hot_start(const or_lattice* l, uint64_t seed, double* links)
{
    or_rng r;
    or_rng_seed(&r, seed);
    size_t n = l->nsite * (size_t)l->dim;
    for (size_t k = 0; k < n; ++k) {
        or_mat3 m;
        or_random_su3(&r, &m);
        store_mat(links + 18 * k, &m);
    }
}

/* LinkMatrix::new_cold, src/field.rs:716-722 */
void or_cold_start(const or_lattice* l, double* links)
{
    size_t n = l->nsite * (size_t)l->dim;
    or_mat3 id;
    mident(&id);
    for (size_t k = 0; k < n; ++k) store_mat(links + 18 * k, &id);
}

/* EField::new_determinist, src/field.rs:1086-1099: per site, per direction,
 * 8 draws of N(0, sigma) */
void or_random_efield(const or_lattice* l, uint64_t seed, double sigma, double* e)
{
    or_rng r;
    or_rng_seed(&r, seed);
    size_t n = l->nsite * (size_t)l->dim * 8;
    for (size_t k = 0; k < n; ++k) e[k] = or_rng_normal(&r, 0.0, sigma);
}

/* MetropolisHastingsSweep::potential_modif, metropolis_hastings_sweep.rs:126-143:
 * orthonormalize(random_su3_close_to_unity(spread)); drawn here in storage
 * order, proposal k of link l stored at [k][l][18]. */
void or_mh_proposals(const or_lattice* l, uint64_t seed, double spread, size_t n_update,
                     double* proposals)
{
    or_rng r;
    or_rng_seed(&r, seed);
    size_t nl = l->nsite * (size_t)l->dim;
    for (size_t li = 0; li < nl; ++li)
        for (size_t k = 0; k < n_update; ++k) {
            or_mat3 m, o;
            or_random_su3_close_to_unity(&r, spread, &m);
            or_orthonormalize(&m, &o);
            store_mat(proposals + 18 * (k * nl + li), &o);
        }
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) */
void or_philox4x32(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c[4] = {ctr_in[0], ctr_in[1], ctr_in[2], ctr_in[3]};
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

/* Device momentum stream: link l, block j (0..3) -> counter
 * {lo32(l), hi16(l) | j << 16, lo32(counter), hi32(counter)}, key = seed.
 * Each block gives two 53-bit uniforms u, w -> Box-Muller pair
 * r = sqrt(-2 log(1-u)) * sigma, (r cos 2 pi w, r sin 2 pi w). */
static inline double u53(uint32_t a, uint32_t b)
{
    return (double)(((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6)) * 0x1.0p-53;
}
void or_device_momenta(const or_lattice* l, uint64_t seed, uint64_t counter, double sigma,
                       double* e)
{
    size_t nl = l->nsite * (size_t)l->dim;
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (size_t li = 0; li < nl; ++li)
        for (uint32_t j = 0; j < 4; ++j) {
            uint32_t ctr[4] = {(uint32_t)li, (uint32_t)((li >> 32) & 0xFFFFu) | (j << 16),
                               (uint32_t)counter, (uint32_t)(counter >> 32)};
            uint32_t o[4];
            or_philox4x32(ctr, key, o);
            double u = u53(o[0], o[1]);
            double w = u53(o[2], o[3]);
            double rad = sqrt(-2.0 * log(1.0 - u)) * sigma;
            double th = 2.0 * M_PI * w;
            e[li * 8 + 2 * j] = rad * cos(th);
            e[li * 8 + 2 * j + 1] = rad * sin(th);
        }
}

/* Device hot start stream: link l, attempt r (0..3), block j (0..5) ->
 * Philox counter {lo32(l), hi16(l) | (16 + 6r + j) << 16, 0, 0}, key = seed;
 * each block gives two uniforms 2*u53 - 1 in [-1, 1).  v1 = (x0 + i x1,
 * x2 + i x3, x4 + i x5), v2 = (x6 + i x7, ...); accepted as random_su3
 * (src/su3.rs:334-347) does, then ortho_matrix_from_2_vector. */
void or_device_hot_start(const or_lattice* l, uint64_t seed, double* links)
{
    size_t nl = l->nsite * (size_t)l->dim;
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (size_t li = 0; li < nl; ++li) {
        vec3 v1, v2;
        for (uint32_t attempt = 0; attempt < 4; ++attempt) {
            double x[12];
            for (uint32_t j = 0; j < 6; ++j) {
                uint32_t blk = 16u + 6u * attempt + j;
                uint32_t ctr[4] = {(uint32_t)li, (uint32_t)((li >> 32) & 0xFFFFu) | (blk << 16), 0u, 0u};
                uint32_t o[4];
                or_philox4x32(ctr, key, o);
                x[2 * j] = 2.0 * u53(o[0], o[1]) - 1.0;
                x[2 * j + 1] = 2.0 * u53(o[2], o[3]) - 1.0;
            }
            for (int r = 0; r < 3; ++r) {
                v1.v[r] = cx(x[2 * r], x[2 * r + 1]);
                v2.v[r] = cx(x[6 + 2 * r], x[6 + 2 * r + 1]);
            }
            or_cplx d = vdot(&v1, &v2, 0);
            if (vnorm(&v1) > F64_EPS && hypot(d.re, d.im) > F64_EPS) break;
        }
        or_mat3 m;
        ortho_matrix_from_2_vector(&v1, &v2, &m);
        store_mat(links + 18 * li, &m);
    }
}

/* Device Metropolis stream (the rng of MetropolisHastingsSweep::next_element,
 * metropolis_hastings_sweep.rs:146-174, drawn per link): Philox counter
 * {lo32(l), hi16(l) | blk << 16, lo32(counter), hi32(counter)}, key = seed.
 * Update j of link l uses blocks 64 + 8j + b (b = 0..6): 14 uniforms
 * w0..w13 = u53 pairs.  su2 number t (t = 0, 1, 2) takes v = 2 w_{4t+k} - 1
 * (k = 0..2) as the Uniform(-1, 1) draws and w_{4t+3} < 0.5 as its
 * Bernoulli(0.5); w12 < 0.5 is the adjoint Bernoulli; the proposal is
 * orthonormalize(get_r get_s get_t) as potential_modif (:126-143) forms it.
 * The accept word of link l is (x << 32) | y of block 63.
 * proposals [n_update][link][18] column-major, uniforms [link]. */
void or_device_mh_inputs(const or_lattice* l, uint64_t seed, uint64_t counter, double spread,
                         size_t n_update, double* proposals, uint64_t* uniforms)
{
    size_t nl = l->nsite * (size_t)l->dim;
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (size_t li = 0; li < nl; ++li) {
        uint32_t hi = (uint32_t)((li >> 32) & 0xFFFFu);
        for (size_t j = 0; j < n_update; ++j) {
            double w[14];
            for (uint32_t b = 0; b < 7; ++b) {
                uint32_t blk = 64u + 8u * (uint32_t)j + b;
                uint32_t ctr[4] = {(uint32_t)li, hi | (blk << 16), (uint32_t)counter,
                                   (uint32_t)(counter >> 32)};
                uint32_t o[4];
                or_philox4x32(ctr, key, o);
                w[2 * b] = u53(o[0], o[1]);
                w[2 * b + 1] = u53(o[2], o[3]);
            }
            or_cplx su2[3][4];
            for (int t = 0; t < 3; ++t) {
                double v[3] = {2.0 * w[4 * t] - 1.0, 2.0 * w[4 * t + 1] - 1.0, 2.0 * w[4 * t + 2] - 1.0};
                su2_close_to_unity_from(v, w[4 * t + 3] < 0.5, spread, su2[t]);
            }
            or_mat3 m, o;
            su3_from_su2(su2[0], su2[1], su2[2], w[12] < 0.5, &m);
            or_orthonormalize(&m, &o);
            store_mat(proposals + 18 * (j * nl + li), &o);
        }
        uint32_t ctr[4] = {(uint32_t)li, hi | (63u << 16), (uint32_t)counter, (uint32_t)(counter >> 32)};
        uint32_t o[4];
        or_philox4x32(ctr, key, o);
        uniforms[li] = ((uint64_t)o[0] << 32) | (uint64_t)o[1];
    }
}

/* ------------------------------------------------------------------------ */
/* stencils: src/field.rs:726-804, src/simulation/state.rs                  */
/* ------------------------------------------------------------------------ */

/* LinkMatrix::matrix, src/field.rs:726-740 */
void or_link_matrix(const or_lattice* l, const double* links, size_t site, int dir,
                    int positive, or_mat3* out)
{
    size_t idx = or_link_canonical(l, site, dir, positive);
    or_mat3 m;
    load_mat(links + 18 * idx, &m);
    if (positive)
        *out = m;
    else
        or_adj(&m, out);
}

/* LinkMatrix::sij, src/field.rs:743-758 (signed directions) */
void or_sij(const or_lattice* l, const double* links, size_t site, int di, int pi, int dj,
            int pj, or_mat3* out)
{
    or_mat3 u_j, u_i_p_j, u_j_pi, u_j_pi_d, t;
    or_link_matrix(l, links, site, dj, pj, &u_j);
    size_t p_j = or_neighbor(l, site, dj, pj);
    or_link_matrix(l, links, p_j, di, pi, &u_i_p_j);
    size_t p_i = or_neighbor(l, site, di, pi);
    or_link_matrix(l, links, p_i, dj, pj, &u_j_pi);
    or_adj(&u_j_pi, &u_j_pi_d);
    or_matmul(&u_j, &u_i_p_j, &t);
    or_matmul(&t, &u_j_pi_d, out);
}

/* LinkMatrix::pij, src/field.rs:761-771: U_i(x) * S_ij(x)^dagger */
void or_pij(const or_lattice* l, const double* links, size_t site, int di, int pi, int dj,
            int pj, or_mat3* out)
{
    or_mat3 s, sd, u_i;
    or_sij(l, links, site, di, pi, dj, pj, &s);
    or_link_matrix(l, links, site, di, pi, &u_i);
    or_adj(&s, &sd);
    or_matmul(&u_i, &sd, out);
}

/* LinkMatrix::clover, src/field.rs:806-820 (positive i, j): used by the
 * known-answer tests of src/field.rs:1579-1711 to pin signed-direction pij */
void or_clover(const or_lattice* l, const double* links, size_t site, int di, int dj,
               or_mat3* out)
{
    or_mat3 a, b, c, d, s;
    or_pij(l, links, site, di, 1, dj, 1, &a);
    or_pij(l, links, site, dj, 1, di, 0, &b);
    or_pij(l, links, site, di, 0, dj, 0, &c);
    or_pij(l, links, site, dj, 0, di, 1, &d);
    madd(&a, &b, &s);
    madd(&s, &c, &s);
    madd(&s, &d, out);
}

/* num-complex division (a + bi) / (c + di) = ((ac + bd) + (bc - ad) i) / (c^2 + d^2),
 * as used for matrix / Complex in src/field.rs:833,867 */
static or_cplx cdiv_nc(or_cplx x, or_cplx y)
{
    const double n = y.re * y.re + y.im * y.im;
    return cx((x.re * y.re + x.im * y.im) / n, (x.im * y.re - x.re * y.im) / n);
}

/* LinkMatrix::f_mu_nu, src/field.rs:822-834:
 * (clover(i, j) - clover(j, i)) / (8 a^2) */
void or_f_mu_nu(const or_lattice* l, const double* links, size_t site, int di, int dj, or_mat3* out)
{
    or_mat3 a, b, m;
    or_clover(l, links, site, di, dj, &a);
    or_clover(l, links, site, dj, di, &b);
    msub(&a, &b, &m);
    const or_cplx den = cx(8.0 * l->size * l->size, 0.0);
    for (int k = 0; k < 9; ++k) out->m[k] = cdiv_nc(m.m[k], den);
}

/* levi_civita, src/utils.rs:306-318: product over pairs j < i of
 * sign(index[i] - index[j]) */
int or_levi_civita(const int* index, int n)
{
    int prod = 1;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) prod *= index[i] == index[j] ? 0 : (index[i] > index[j] ? 1 : -1);
    return prod;
}

/* LinkMatrix::magnetic_field, src/field.rs:848-868:
 * sum_{i, j} f_mu_nu(i, j) * levi_civita(dir, i, j), divided by 2i */
void or_magnetic_field(const or_lattice* l, const double* links, size_t site, int dir, or_mat3* out)
{
    or_mat3 sum;
    mzero(&sum);
    for (int i = 0; i < l->dim; ++i) {
        or_mat3 inner;
        mzero(&inner);
        for (int j = 0; j < l->dim; ++j) {
            or_mat3 f;
            or_f_mu_nu(l, links, site, i, j, &f);
            const int idx[3] = {dir, i, j};
            const double lc = (double)or_levi_civita(idx, 3);
            for (int k = 0; k < 9; ++k) f.m[k] = cmul(f.m[k], cx(lc, 0.0));
            madd(&inner, &f, &inner);
        }
        madd(&sum, &inner, &sum);
    }
    for (int k = 0; k < 9; ++k) out->m[k] = cdiv_nc(sum.m[k], cx(0.0, 2.0));
}

/* the three observables over the whole lattice, in the host layouts
 * [site][i][j] (clover, f_mu_nu) and [site][dir] (magnetic_field_vec,
 * src/field.rs:836-845), 18 doubles per matrix */
void or_clover_field(const or_lattice* l, const double* links, double* out)
{
    const size_t D = (size_t)l->dim;
#pragma omp parallel for OMP_THREADS schedule(static)
    for (size_t s = 0; s < l->nsite; ++s)
        for (int i = 0; i < l->dim; ++i)
            for (int j = 0; j < l->dim; ++j) {
                or_mat3 m;
                or_clover(l, links, s, i, j, &m);
                store_mat(out + 18 * ((s * D + (size_t)i) * D + (size_t)j), &m);
            }
}

void or_f_mu_nu_field(const or_lattice* l, const double* links, double* out)
{
    const size_t D = (size_t)l->dim;
#pragma omp parallel for OMP_THREADS schedule(static)
    for (size_t s = 0; s < l->nsite; ++s)
        for (int i = 0; i < l->dim; ++i)
            for (int j = 0; j < l->dim; ++j) {
                or_mat3 m;
                or_f_mu_nu(l, links, s, i, j, &m);
                store_mat(out + 18 * ((s * D + (size_t)i) * D + (size_t)j), &m);
            }
}

void or_magnetic_field_vec(const or_lattice* l, const double* links, double* out)
{
    const size_t D = (size_t)l->dim;
#pragma omp parallel for OMP_THREADS schedule(static)
    for (size_t s = 0; s < l->nsite; ++s)
        for (int d = 0; d < l->dim; ++d) {
            or_mat3 m;
            or_magnetic_field(l, links, s, d, &m);
            store_mat(out + 18 * (s * D + (size_t)d), &m);
        }
}

/* per-site sum of Tr P_ij over i < j, in the nested order of
 * LinkMatrix::average_trace_plaquette, src/field.rs:788-801 */
static or_cplx site_trace_plaquette(const or_lattice* l, const double* links, size_t site)
{
    or_cplx outer = C_ZERO;
    for (int i = 0; i < l->dim; ++i) {
        or_cplx inner = C_ZERO;
        for (int j = i + 1; j < l->dim; ++j) {
            or_mat3 p;
            or_pij(l, links, site, i, 1, j, 1, &p);
            inner = cadd(inner, mtrace(&p));
        }
        outer = cadd(outer, inner);
    }
    return outer;
}

/* LinkMatrix::average_trace_plaquette, src/field.rs:775-804 */
or_cplx or_average_trace_plaquette(const or_lattice* l, const double* links)
{
    or_cplx* part = (or_cplx*)malloc(sizeof(or_cplx) * l->nsite);
#pragma omp parallel for OMP_THREADS schedule(static)
    for (size_t s = 0; s < l->nsite; ++s) part[s] = site_trace_plaquette(l, links, s);
    or_cplx sum = C_ZERO;
    for (size_t s = 0; s < l->nsite; ++s) sum = cadd(sum, part[s]);
    free(part);
    double nplaq = (double)(l->nsite * (size_t)(l->dim * (l->dim - 1) / 2));
    return cx(sum.re / nplaq, sum.im / nplaq);
}

/* LatticeStateDefault::hamiltonian_links, src/simulation/state.rs:821-849 */
double or_hamiltonian_links(const or_lattice* l, const double* links, double beta)
{
    double* part = (double*)malloc(sizeof(double) * l->nsite);
#pragma omp parallel for OMP_THREADS schedule(static)
    for (size_t s = 0; s < l->nsite; ++s) {
        double outer = 0.0;
        for (int i = 0; i < l->dim; ++i) {
            double inner = 0.0;
            for (int j = i + 1; j < l->dim; ++j) {
                or_mat3 p;
                or_pij(l, links, s, i, 1, j, 1, &p);
                inner += 1.0 - mtrace(&p).re / CA;
            }
            outer += inner;
        }
        part[s] = outer;
    }
    double sum = 0.0;
    for (size_t s = 0; s < l->nsite; ++s) sum += part[s];
    free(part);
    return sum * beta;
}

/* hamiltonian_efield, src/simulation/state.rs:1370-1385 */
double or_hamiltonian_efield(const or_lattice* l, const double* e, double beta)
{
    double sum = 0.0;
    for (size_t s = 0; s < l->nsite; ++s) {
        double site_sum = 0.0;
        for (int i = 0; i < l->dim; ++i)
            site_sum += or_trace_squared(e + ((s * (size_t)l->dim + (size_t)i) * 8));
        sum += site_sum;
    }
    return sum * beta;
}

/* derivative_e, src/simulation/state.rs:1420-1448.  Sum over the signed
 * directions in the order +0,-0,+1,-1,... (procedural_macro/src/lib.rs:56-60)
 * skipping +-dir; projection (G_a * U) * sum_s as in :1439-1441. */
void or_derivative_e(const or_lattice* l, const double* links, size_t site, double* out)
{
    init_tables();
    const double c = -sqrt(2.0 / CA);
    for (int dir = 0; dir < l->dim; ++dir) {
        or_mat3 u_i, sum_s, s, sd;
        or_link_matrix(l, links, site, dir, 1, &u_i);
        mzero(&sum_s);
        for (int d2 = 0; d2 < l->dim; ++d2) {
            if (d2 == dir) continue;
            for (int sign = 1; sign >= 0; --sign) {
                or_sij(l, links, site, dir, 1, d2, sign, &s);
                or_adj(&s, &sd);
                madd(&sum_s, &sd, &sum_s);
            }
        }
        for (int a = 0; a < 8; ++a) {
            or_mat3 gu, gus;
            or_matmul(&G_TAB[a], &u_i, &gu);
            or_matmul(&gu, &sum_s, &gus);
            out[dir * 8 + a] = c * mtrace(&gus).im / l->size;
        }
    }
}

void or_derivative_e_all(const or_lattice* l, const double* links, double* out)
{
#pragma omp parallel for OMP_THREADS schedule(static)
    for (size_t s = 0; s < l->nsite; ++s)
        or_derivative_e(l, links, s, out + s * (size_t)l->dim * 8);
}

/* derivative_u, src/simulation/state.rs:1407-1417:
 * ((E * U) * (0 + i sqrt(2 CA))) * (1/a + 0i) */
void or_derivative_u(const or_lattice* l, const double* links, const double* e, size_t site,
                     int dir, or_mat3* out)
{
    or_mat3 u, em, eu, t;
    or_link_matrix(l, links, site, dir, 1, &u);
    or_adjoint_to_matrix(e + ((site * (size_t)l->dim + (size_t)dir) * 8), &em);
    or_matmul(&em, &u, &eu);
    mscal(&eu, cx(0.0, sqrt(2.0 * CA)), &t);
    mscal(&t, cx(1.0 / l->size, 0.0), out);
}

/* Gradient of the Wilson action S = beta * sum (1 - Re Tr P / 3) with respect
 * to U -> exp(i w^a T^a) U: dS/dw^a = -(beta * a / sqrt(6)) * derivative_e^a,
 * the identity of derivative_e above with the action of
 * src/simulation/state.rs:821-849.  Layout [site][mu][8]. */
void or_wilson_backward(const or_lattice* l, const double* links, double beta, double* grad)
{
    or_derivative_e_all(l, links, grad);
    const double f = -(beta * l->size) / sqrt(2.0 * CA);
    size_t n = l->nsite * (size_t)l->dim * 8;
    for (size_t k = 0; k < n; ++k) grad[k] *= f;
}

/* ------------------------------------------------------------------------ */
/* integrators: src/integrator/mod.rs:216-254, symplectic_euler_rayon.rs     */
/* ------------------------------------------------------------------------ */

/* integrate_efield, src/integrator/mod.rs:240-254: e + deriv * dt */
void or_integrate_efield(const or_lattice* l, const double* links, const double* e_in,
                         double dt, double* e_out)
{
    const size_t D = (size_t)l->dim;
#pragma omp parallel for OMP_THREADS schedule(static)
    for (size_t s = 0; s < l->nsite; ++s) {
        double deriv[OR_MAX_DIM * 8];
        or_derivative_e(l, links, s, deriv);
        for (size_t k = 0; k < D * 8; ++k)
            e_out[s * D * 8 + k] = e_in[s * D * 8 + k] + deriv[k] * dt;
    }
}

/* integrate_link, src/integrator/mod.rs:216-233 (additive, reference mode):
 * U + derivative_u * (dt + 0i).  Exp mode (not in the reference integrator):
 * U' = su3_exp_i(e * (dt * sqrt(2 CA) / a)) * U, with su3_exp_i of
 * src/su3.rs:832-855. */
void or_integrate_link(const or_lattice* l, const double* links, const double* e, double dt,
                       int mode, double* links_out)
{
    const size_t D = (size_t)l->dim;
    const size_t nl = l->nsite * D;
    const double scale = dt * sqrt(2.0 * CA) / l->size;
#pragma omp parallel for OMP_THREADS schedule(static)
    for (size_t k = 0; k < nl; ++k) {
        size_t s = k / D;
        int dir = (int)(k % D);
        or_mat3 u, r;
        load_mat(links + 18 * k, &u);
        if (mode == OR_DRIFT_EXP) {
            double v[8];
            for (int a = 0; a < 8; ++a) v[a] = e[k * 8 + a] * scale;
            or_mat3 x;
            or_su3_exp_i(v, &x);
            or_matmul(&x, &u, &r);
        } else {
            or_mat3 du, t;
            or_derivative_u(l, links, e, s, dir, &du);
            mscal(&du, cx(dt, 0.0), &t);
            madd(&u, &t, &r);
        }
        store_mat(links_out + 18 * k, &r);
    }
}

/* SymplecticEulerRayon::integrate_symplectic, symplectic_euler_rayon.rs:222-252:
 * e_half = e + dt/2 F(U); U' = U + dt D(U, e_half); e' = e_half + dt/2 F(U') */
void or_integrate_symplectic(const or_lattice* l, double* links, double* e, double dt,
                             int mode)
{
    const size_t nl = l->nsite * (size_t)l->dim;
    double* e_half = (double*)malloc(sizeof(double) * nl * 8);
    double* u_new = (double*)malloc(sizeof(double) * nl * 18);
    or_integrate_efield(l, links, e, dt / 2.0, e_half);
    or_integrate_link(l, links, e_half, dt, mode, u_new);
    or_integrate_efield(l, u_new, e_half, dt / 2.0, e);
    memcpy(links, u_new, sizeof(double) * nl * 18);
    free(e_half);
    free(u_new);
}

/* simulate_symplectic_n, src/simulation/state.rs:470-492 (reference scheme:
 * n calls of integrate_symplectic).  The leapfrog scheme merges the two
 * half kicks between consecutive steps into one kick of dt (equal in exact
 * arithmetic).  Returns -1 (ZeroIntegration) for n == 0. */
int or_simulate_symplectic_n(const or_lattice* l, double* links, double* e, double dt,
                             size_t n, int mode, int scheme)
{
    if (n == 0) return -1;
    if (scheme == OR_SCHEME_REFERENCE) {
        for (size_t k = 0; k < n; ++k) or_integrate_symplectic(l, links, e, dt, mode);
        return 0;
    }
    const size_t nl = l->nsite * (size_t)l->dim;
    double* e2 = (double*)malloc(sizeof(double) * nl * 8);
    double* u2 = (double*)malloc(sizeof(double) * nl * 18);
    or_integrate_efield(l, links, e, dt / 2.0, e2);
    memcpy(e, e2, sizeof(double) * nl * 8);
    for (size_t k = 0; k < n; ++k) {
        or_integrate_link(l, links, e, dt, mode, u2);
        memcpy(links, u2, sizeof(double) * nl * 18);
        or_integrate_efield(l, links, e, (k + 1 < n) ? dt : dt / 2.0, e2);
        memcpy(e, e2, sizeof(double) * nl * 8);
    }
    free(e2);
    free(u2);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Monte Carlo: src/simulation/monte_carlo/                                  */
/* ------------------------------------------------------------------------ */

/* HybridMonteCarloDiagnostic, hybrid_monte_carlo.rs:584-613, with the momenta
 * given (the refresh of state.rs:1093-1108 is an input here) and one u64 for
 * the Bernoulli draw. */
int or_hmc_trajectory(const or_lattice* l, double* links, const double* momenta,
                      double beta, double dt, size_t n, int mode, int scheme,
                      uint64_t uniform, or_hmc_stats* st)
{
    if (n == 0) return -1;
    const size_t nl = l->nsite * (size_t)l->dim;
    double* u = (double*)malloc(sizeof(double) * nl * 18);
    double* e = (double*)malloc(sizeof(double) * nl * 8);
    memcpy(u, links, sizeof(double) * nl * 18);
    memcpy(e, momenta, sizeof(double) * nl * 8);
    double h_old = or_hamiltonian_links(l, links, beta) + or_hamiltonian_efield(l, momenta, beta);
    or_simulate_symplectic_n(l, u, e, dt, n, mode, scheme);
    double h_new = or_hamiltonian_links(l, u, beta) + or_hamiltonian_efield(l, e, beta);
    double p = or_clamp_probability(exp(h_old - h_new));
    p = or_clamp_probability(p);
    int acc = or_bernoulli_accept(p, uniform);
    if (acc) memcpy(links, u, sizeof(double) * nl * 18);
    if (st) {
        st->h_old = h_old;
        st->h_new = h_new;
        st->delta_h = h_new - h_old;
        st->probability = p;
        st->accepted = acc;
    }
    free(u);
    free(e);
    return 0;
}

/* staple, src/simulation/monte_carlo/mod.rs:339-362 */
void or_staple(const or_lattice* l, const double* links, size_t site, int dir, or_mat3* out)
{
    or_mat3 acc;
    mzero(&acc);
    for (int di = 0; di < l->dim; ++di) {
        if (di == dir) continue;
        or_mat3 s, el_1, u1, u2, u2a, u3, t, t2, term;
        or_sij(l, links, site, dir, 1, di, 1, &s);
        or_adj(&s, &el_1);
        size_t p_dir = or_neighbor(l, site, dir, 1);
        or_link_matrix(l, links, p_dir, di, 0, &u1);
        size_t m_di = or_neighbor(l, site, di, 0);
        or_link_matrix(l, links, m_di, dir, 1, &u2a);
        or_adj(&u2a, &u2);
        or_link_matrix(l, links, m_di, di, 1, &u3);
        or_matmul(&u1, &u2, &t);
        or_matmul(&t, &u3, &t2);
        madd(&el_1, &t2, &term);
        madd(&acc, &term, &acc);
    }
    *out = acc;
}

/* delta_s_old_new_cmp, src/simulation/monte_carlo/mod.rs:324-334 */
double or_delta_s(const or_lattice* l, const double* links, size_t site, int dir,
                  const or_mat3* new_link, double beta)
{
    or_mat3 a, old, diff, prod;
    or_staple(l, links, site, dir, &a);
    or_link_matrix(l, links, site, dir, 1, &old);
    msub(new_link, &old, &diff);
    or_matmul(&diff, &a, &prod);
    return -mtrace(&prod).re * beta / CA;
}

static void mh_update_link(const or_lattice* l, double* links, size_t site, int dir,
                           double beta, size_t n_update, const double* proposals,
                           const uint64_t* uniforms, double* psum, size_t* nacc)
{
    const size_t nl = l->nsite * (size_t)l->dim;
    const size_t li = site * (size_t)l->dim + (size_t)dir;
    or_mat3 nw, r;
    load_mat(links + 18 * li, &nw);
    for (size_t k = 0; k < n_update; ++k) {
        load_mat(proposals + 18 * (k * nl + li), &r);
        or_matmul(&r, &nw, &nw);
    }
    double p = or_clamp_probability(exp(-or_delta_s(l, links, site, dir, &nw, beta)));
    *psum += p;
    if (or_bernoulli_accept(p, uniforms[li])) {
        *nacc += 1;
        store_mat(links + 18 * li, &nw);
    }
}

/* MetropolisHastingsSweep::next_element_default,
 * metropolis_hastings_sweep.rs:146-174.  Lexicographic order is the
 * reference's storage order; checkerboard order visits, for mu = 0..D-1 and
 * parity 0 then 1, every site of that parity in storage order (links of one
 * (mu, parity) class do not enter each other's staples).  Proposals are
 * given ([k][link][18], already orthonormalized) and one u64 per link. */
int or_metropolis_sweep(const or_lattice* l, double* links, double beta, size_t n_update,
                        const double* proposals, const uint64_t* uniforms, int order,
                        or_mh_stats* st)
{
    if (n_update == 0) return -1;
    double psum = 0.0;
    size_t nacc = 0;
    if (order == OR_ORDER_LEXICOGRAPHIC) {
        for (size_t s = 0; s < l->nsite; ++s)
            for (int d = 0; d < l->dim; ++d)
                mh_update_link(l, links, s, d, beta, n_update, proposals, uniforms, &psum,
                               &nacc);
    } else {
        for (int i = 0; i < l->dim; ++i)
            if (l->extent[i] % 2) return -2;
        for (int d = 0; d < l->dim; ++d)
            for (int par = 0; par < 2; ++par)
                for (size_t s = 0; s < l->nsite; ++s)
                    if (or_site_parity(l, s) == par)
                        mh_update_link(l, links, s, d, beta, n_update, proposals, uniforms,
                                       &psum, &nacc);
    }
    if (st) {
        st->prob_replace_mean = psum / (double)(l->nsite * (size_t)l->dim);
        st->number_replace_last = nacc;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* heat bath: simulation/monte_carlo/heat_bath.rs:73-123,                    */
/* statistics/distribution.rs:147-352, su2.rs:155-157,200-217, su3.rs:428-660 */
/* ------------------------------------------------------------------------ */

/* 2x2 matrices row-major [m00, m01, m10, m11] */
static void m2_mul(const or_cplx a[4], const or_cplx b[4], or_cplx c[4])
{
    or_cplx r[4];
    r[0] = cadd(cmul(a[0], b[0]), cmul(a[1], b[2]));
    r[1] = cadd(cmul(a[0], b[1]), cmul(a[1], b[3]));
    r[2] = cadd(cmul(a[2], b[0]), cmul(a[3], b[2]));
    r[3] = cadd(cmul(a[2], b[1]), cmul(a[3], b[3]));
    memcpy(c, r, sizeof(r));
}

/* su2::project_to_su2_unorm, src/su2.rs:155-157: m - m^+ + I conj(Tr m) */
void or_su2_project_unorm(const or_cplx m[4], or_cplx out[4])
{
    or_cplx tr = cconj(cadd(m[0], m[3]));
    or_cplx r[4];
    r[0] = cadd(csub(m[0], cconj(m[0])), tr);
    r[1] = csub(m[1], cconj(m[2]));
    r[2] = csub(m[2], cconj(m[1]));
    r[3] = cadd(csub(m[3], cconj(m[3])), tr);
    memcpy(out, r, sizeof(r));
}

/* get_sub_block_r/s/t (src/su3.rs:559-640): block 0 = rows/cols (0,1),
 * 1 = (0,2), 2 = (1,2) */
static const int SUB_I[3] = {0, 0, 1};
static const int SUB_J[3] = {1, 2, 2};
static void sub_block(const or_mat3* m, int b, or_cplx out[4])
{
    int i = SUB_I[b], j = SUB_J[b];
    out[0] = M(m, i, i); out[1] = M(m, i, j); out[2] = M(m, j, i); out[3] = M(m, j, j);
}
/* get_r/get_s/get_t (src/su3.rs:428-530): embed at the block, 1 elsewhere */
static void embed_block(const or_cplx m[4], int b, or_mat3* out)
{
    int i = SUB_I[b], j = SUB_J[b];
    mident(out);
    M(out, i, i) = m[0]; M(out, i, j) = m[1]; M(out, j, i) = m[2]; M(out, j, j) = m[3];
}

/* Device heat-bath stream of link l in sweep n (key seed): subgroup g
 * (0 r, 1 s, 2 t) uses Philox blocks HB + 256 g + k, HB = 0x8000:
 *   k = 2t, 2t+1 (t < 64)  attempt t of the norm sampler: w0 = r in [0,1),
 *                          r_i = 1 - w_{i+1} in (0,1] for ModifiedNormal
 *   k = 128 + 2q, +1 (q < 2) direction draw q: x_i = 2 w_i - 1 (i < 3)
 *   k = 132 + 2q, +1 (q < 2) random_su2 draw q: (re0, im0, re1, im1) = 2 w - 1 */
#define OR_HB_BASE 0x8000u
#define OR_HB_ATTEMPTS 64
static void hb_draws(uint64_t l, uint32_t blk, uint64_t n, uint64_t seed, double w[4])
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int b = 0; b < 2; ++b) {
        uint32_t ctr[4] = {(uint32_t)l, (uint32_t)((l >> 32) & 0xFFFFu) | ((blk + (uint32_t)b) << 16),
                           (uint32_t)n, (uint32_t)(n >> 32)};
        uint32_t o[4];
        or_philox4x32(ctr, key, o);
        w[2 * b] = u53(o[0], o[1]);
        w[2 * b + 1] = u53(o[2], o[3]);
    }
}

/* HeatBathSweep::heat_bath_su2 (heat_bath.rs:73-87) with the draws above.
 * HeatBathDistributionNorm (distribution.rs:336-350): loop { r; lambda =
 * ModifiedNormal (distribution.rs:~66-76); if r^2 <= 1 - lambda^2 return
 * 1 - 2 lambda^2 }, here at most OR_HB_ATTEMPTS tries (then the subgroup
 * update is the identity and *exhausted is incremented).  Returns the 2x2
 * factor X v. */
static void hb_su2(const or_cplx st[4], double beta, uint64_t l, int g, uint64_t n, uint64_t seed,
                   or_cplx out[4], size_t* exhausted)
{
    const uint32_t base = OR_HB_BASE + 256u * (uint32_t)g;
    /* staple.determinant().real().sqrt() */
    or_cplx det = csub(cmul(st[0], st[3]), cmul(st[1], st[2]));
    double k = sqrt(det.re);
    if (isnormal(k)) {
        const double p = beta * k;
        double x0 = 0.0;
        int ok = 0;
        for (int t = 0; t < OR_HB_ATTEMPTS && !ok; ++t) {
            double w[4];
            hb_draws(l, base + 2u * (uint32_t)t, n, seed, w);
            double r = w[0];
            double r0 = 1.0 - w[1], r1 = 1.0 - w[2], r2 = 1.0 - w[3];
            double c = cos(2.0 * M_PI * r1);
            double lambda = sqrt(-(log(r0) + c * c * log(r2)) / (2.0 * p));
            if (r * r <= 1.0 - lambda * lambda) {
                x0 = 1.0 - 2.0 * lambda * lambda;
                ok = 1;
            }
        }
        if (!ok) {
            *exhausted += 1;
            out[0] = C_ONE; out[1] = C_ZERO; out[2] = C_ZERO; out[3] = C_ONE;
            return;
        }
        double x[3] = {0.0, 0.0, 0.0}, nrm = 0.0;
        for (int q = 0; q < 2; ++q) {
            double w[4];
            hb_draws(l, base + 128u + 2u * (uint32_t)q, n, seed, w);
            for (int i = 0; i < 3; ++i) x[i] = 2.0 * w[i] - 1.0;
            nrm = sqrt((x[0] * x[0] + x[1] * x[1]) + x[2] * x[2]);
            if (nrm > F64_EPS) break;
        }
        double sc = sqrt(1.0 - x0 * x0) / nrm;
        for (int i = 0; i < 3; ++i) x[i] *= sc;
        /* complex_matrix_from_vec (su2.rs:134-140), PAULI_3 = identity quirk */
        or_cplx X[4] = {cx(x0, x[2]), cx(x[1], x[0]), cx(-x[1], x[0]), cx(x0, x[2])};
        /* v = staple^+ / k */
        or_cplx v[4] = {cx(st[0].re / k, -st[0].im / k), cx(st[2].re / k, -st[2].im / k),
                        cx(st[1].re / k, -st[1].im / k), cx(st[3].re / k, -st[3].im / k)};
        m2_mul(X, v, out);
    } else {
        /* su2::random_su2 (su2.rs:200-217) */
        double w[4];
        double nrm = 0.0;
        for (int q = 0; q < 2; ++q) {
            hb_draws(l, base + 132u + 2u * (uint32_t)q, n, seed, w);
            for (int i = 0; i < 4; ++i) w[i] = 2.0 * w[i] - 1.0;
            nrm = sqrt(((w[0] * w[0] + w[1] * w[1]) + w[2] * w[2]) + w[3] * w[3]);
            if (isnormal(nrm)) break;
        }
        or_cplx a = cx(w[0] / nrm, w[1] / nrm), b = cx(w[2] / nrm, w[3] / nrm);
        out[0] = a; out[1] = b; out[2] = cx(-b.re, b.im); out[3] = cconj(a);
    }
}

/* HeatBathSweep::get_modif (heat_bath.rs:90-110): three subgroup heat baths
 * on U A, r U A, s r U A; the new link is t s r U. */
static void hb_update_link(const or_lattice* l, double* links, size_t site, int dir, double beta,
                           uint64_t seed, uint64_t n, size_t* exhausted)
{
    const size_t li = site * (size_t)l->dim + (size_t)dir;
    or_mat3 u, a, w, f[3], acc, tmp;
    load_mat(links + 18 * li, &u);
    or_staple(l, links, site, dir, &a);
    mident(&acc); /* product of the subgroup factors applied so far */
    for (int g = 0; g < 3; ++g) {
        or_mat3 left;
        or_matmul(&acc, &u, &left);
        or_matmul(&left, &a, &w);
        or_cplx blk[4], proj[4], hb[4];
        sub_block(&w, g, blk);
        or_su2_project_unorm(blk, proj);
        hb_su2(proj, beta, (uint64_t)li, g, n, seed, hb, exhausted);
        embed_block(hb, g, &f[g]);
        or_matmul(&f[g], &acc, &tmp);
        acc = tmp;
    }
    or_matmul(&acc, &u, &tmp);
    store_mat(links + 18 * li, &tmp);
}

/* HeatBathSweep::next_element_default (heat_bath.rs:112-123) with the device
 * stream, checkerboard (or lexicographic) order as or_metropolis_sweep.
 * Returns -2 for odd extents in checkerboard order. */
int or_heat_bath_sweep(const or_lattice* l, double* links, double beta, uint64_t seed,
                       uint64_t counter, int order, size_t* exhausted)
{
    size_t ex = 0;
    if (order == OR_ORDER_LEXICOGRAPHIC) {
        for (size_t s = 0; s < l->nsite; ++s)
            for (int d = 0; d < l->dim; ++d) hb_update_link(l, links, s, d, beta, seed, counter, &ex);
    } else {
        for (int i = 0; i < l->dim; ++i)
            if (l->extent[i] % 2) return -2;
        for (int d = 0; d < l->dim; ++d)
            for (int par = 0; par < 2; ++par)
                for (size_t s = 0; s < l->nsite; ++s)
                    if (or_site_parity(l, s) == par) hb_update_link(l, links, s, d, beta, seed, counter, &ex);
    }
    if (exhausted) *exhausted = ex;
    return 0;
}

/* one draw of the heat-bath norm sampler (for the distribution tests):
 * x0 of attempt sequence (l, g = 0, n, seed) with parameter p, or NaN when
 * the attempts run out */
double or_heat_bath_x0(double p, uint64_t l, uint64_t n, uint64_t seed)
{
    for (int t = 0; t < OR_HB_ATTEMPTS; ++t) {
        double w[4];
        hb_draws(l, OR_HB_BASE + 2u * (uint32_t)t, n, seed, w);
        double r = w[0];
        double r0 = 1.0 - w[1], r1 = 1.0 - w[2], r2 = 1.0 - w[3];
        double c = cos(2.0 * M_PI * r1);
        double lambda = sqrt(-(log(r0) + c * c * log(r2)) / (2.0 * p));
        if (r * r <= 1.0 - lambda * lambda) return 1.0 - 2.0 * lambda * lambda;
    }
    return NAN;
}

/* ------------------------------------------------------------------------ */
/* overrelaxation: simulation/monte_carlo/overrelaxation.rs:66-91             */
/* ------------------------------------------------------------------------ */

/* inverse of a 3x3 complex matrix by cofactors; returns det */
static or_cplx m3_inverse(const or_mat3* a, or_mat3* inv)
{
    or_cplx c[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            c[i][j] = csub(cmul(M(a, i1, j1), M(a, i2, j2)), cmul(M(a, i1, j2), M(a, i2, j1)));
        }
    or_cplx det = cadd(cadd(cmul(M(a, 0, 0), c[0][0]), cmul(M(a, 0, 1), c[0][1])), cmul(M(a, 0, 2), c[0][2]));
    double dd = det.re * det.re + det.im * det.im;
    or_cplx idet = cx(det.re / dd, -det.im / dd);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M(inv, i, j) = cmul(c[j][i], idet);
    return det;
}

/* Unitary factor W of the polar decomposition A = W P (W = U V^+ of the SVD
 * A = U S V^+, unique for invertible A), by the scaled Newton iteration
 * X <- (g X + (g X)^{-+}) / 2, g = |det X|^{-1/3}. */
void or_polar_unitary(const or_mat3* a, or_mat3* w)
{
    or_mat3 x = *a;
    for (int it = 0; it < 100; ++it) {
        or_mat3 inv;
        or_cplx det = m3_inverse(&x, &inv);
        double g = pow(hypot(det.re, det.im), -1.0 / 3.0);
        or_mat3 nx;
        double diff = 0.0, nrm = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                /* (inv^+)_ij = conj(inv_ji) */
                or_cplx t = cconj(M(&inv, j, i));
                or_cplx v = cx(0.5 * (g * M(&x, i, j).re + t.re / g), 0.5 * (g * M(&x, i, j).im + t.im / g));
                diff += hypot(v.re - M(&x, i, j).re, v.im - M(&x, i, j).im);
                nrm += hypot(v.re, v.im);
                M(&nx, i, j) = v;
            }
        x = nx;
        if (diff <= 1e-15 * nrm) break;
    }
    *w = x;
}

/* OverrelaxationSweepRotation::get_modif (overrelaxation.rs:66-78):
 * A = staple^+, rot = U V^+ of its SVD, new link = rot U^+ rot */
static void or_update_link(const or_lattice* l, double* links, size_t site, int dir)
{
    const size_t li = site * (size_t)l->dim + (size_t)dir;
    or_mat3 u, st, a, rot, ud, t, nw;
    load_mat(links + 18 * li, &u);
    or_staple(l, links, site, dir, &st);
    or_adj(&st, &a);
    or_polar_unitary(&a, &rot);
    or_adj(&u, &ud);
    or_matmul(&rot, &ud, &t);
    or_matmul(&t, &rot, &nw);
    store_mat(links + 18 * li, &nw);
}

/* SVD A = U S V^+ of a 3x3 complex matrix by one-sided (Hestenes) Jacobi:
 * rotate column pairs of A (and of V = I) until they are orthogonal; then
 * S_k = |a_k| and U_k = a_k / S_k.  Columns with S_k = 0 are left as they
 * are (the overrelaxation input is nonsingular). */
void or_svd3(const or_mat3* a, or_mat3* u, double sv[3], or_mat3* v)
{
    or_mat3 w = *a;
    mident(v);
    for (int sweep = 0; sweep < 30; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                double al = 0.0, be = 0.0;
                or_cplx ga = C_ZERO; /* a_p^+ a_q */
                for (int i = 0; i < 3; ++i) {
                    or_cplx x = M(&w, i, p), y = M(&w, i, q);
                    al += x.re * x.re + x.im * x.im;
                    be += y.re * y.re + y.im * y.im;
                    ga = cadd(ga, cmul(cconj(x), y));
                }
                double g = hypot(ga.re, ga.im);
                off += g;
                diag += sqrt(al * be);
                if (g <= 1e-300) continue;
                /* phase e^{i phi} = ga / g; real rotation of [[al, g], [g, be]] */
                or_cplx ph = cx(ga.re / g, ga.im / g);
                double zeta = (be - al) / (2.0 * g);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
                /* a_p' = c a_p - s conj(ph) a_q ... in the form that zeroes a_p'^+ a_q' */
                for (int i = 0; i < 3; ++i) {
                    or_cplx x = M(&w, i, p), y = M(&w, i, q);
                    or_cplx yp = cmul(y, cconj(ph));
                    M(&w, i, p) = cx(c * x.re - sn * yp.re, c * x.im - sn * yp.im);
                    or_cplx xq = cmul(x, ph);
                    M(&w, i, q) = cx(sn * xq.re + c * y.re, sn * xq.im + c * y.im);
                    or_cplx vx = M(v, i, p), vy = M(v, i, q);
                    or_cplx vyp = cmul(vy, cconj(ph));
                    M(v, i, p) = cx(c * vx.re - sn * vyp.re, c * vx.im - sn * vyp.im);
                    or_cplx vxq = cmul(vx, ph);
                    M(v, i, q) = cx(sn * vxq.re + c * vy.re, sn * vxq.im + c * vy.im);
                }
            }
        if (off <= 1e-16 * diag) break;
    }
    for (int k = 0; k < 3; ++k) {
        double n = 0.0;
        for (int i = 0; i < 3; ++i) n += M(&w, i, k).re * M(&w, i, k).re + M(&w, i, k).im * M(&w, i, k).im;
        n = sqrt(n);
        sv[k] = n;
        for (int i = 0; i < 3; ++i) M(u, i, k) = n > 0 ? cx(M(&w, i, k).re / n, M(&w, i, k).im / n) : M(&w, i, k);
    }
}

/* OverrelaxationSweepReverse::get_modif (overrelaxation.rs:136-150):
 * A = staple^+ = U S V^+, new link = U reverse(U^+ L V) V^+ with
 * su3::reverse (su3.rs:705-714) negating the off-diagonal entries.  The
 * result does not depend on the SVD's column phases or order. */
static void or_reverse_update_link(const or_lattice* l, double* links, size_t site, int dir)
{
    const size_t li = site * (size_t)l->dim + (size_t)dir;
    or_mat3 lk, st, a, u, v, ud, t1, m, t2, vd, nw;
    double sv[3];
    load_mat(links + 18 * li, &lk);
    or_staple(l, links, site, dir, &st);
    or_adj(&st, &a);
    or_svd3(&a, &u, sv, &v);
    or_adj(&u, &ud);
    or_matmul(&ud, &lk, &t1);
    or_matmul(&t1, &v, &m);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (i != j) M(&m, i, j) = cx(-M(&m, i, j).re, -M(&m, i, j).im);
    or_matmul(&u, &m, &t2);
    or_adj(&v, &vd);
    or_matmul(&t2, &vd, &nw);
    store_mat(links + 18 * li, &nw);
}

int or_overrelaxation_reverse_sweep(const or_lattice* l, double* links, int order)
{
    if (order == OR_ORDER_LEXICOGRAPHIC) {
        for (size_t s = 0; s < l->nsite; ++s)
            for (int d = 0; d < l->dim; ++d) or_reverse_update_link(l, links, s, d);
    } else {
        for (int i = 0; i < l->dim; ++i)
            if (l->extent[i] % 2) return -2;
        for (int d = 0; d < l->dim; ++d)
            for (int par = 0; par < 2; ++par)
                for (size_t s = 0; s < l->nsite; ++s)
                    if (or_site_parity(l, s) == par) or_reverse_update_link(l, links, s, d);
    }
    return 0;
}

int or_overrelaxation_sweep(const or_lattice* l, double* links, int order)
{
    if (order == OR_ORDER_LEXICOGRAPHIC) {
        for (size_t s = 0; s < l->nsite; ++s)
            for (int d = 0; d < l->dim; ++d) or_update_link(l, links, s, d);
    } else {
        for (int i = 0; i < l->dim; ++i)
            if (l->extent[i] % 2) return -2;
        for (int d = 0; d < l->dim; ++d)
            for (int par = 0; par < 2; ++par)
                for (size_t s = 0; s < l->nsite; ++s)
                    if (or_site_parity(l, s) == par) or_update_link(l, links, s, d);
    }
    return 0;
}

/* LatticeStateDefault::normalize_link_matrices, state.rs:754-756 ->
 * LinkMatrix::normalize, src/field.rs:897-901 */
void or_normalize_link_matrices(const or_lattice* l, double* links)
{
    const size_t nl = l->nsite * (size_t)l->dim;
#pragma omp parallel for OMP_THREADS schedule(static)
    for (size_t k = 0; k < nl; ++k) {
        or_mat3 m;
        load_mat(links + 18 * k, &m);
        or_orthonormalize(&m, &m);
        store_mat(links + 18 * k, &m);
    }
}

/* ------------------------------------------------------------------------ */
/* Gauss law: src/field.rs:1172-1337                                         */
/* ------------------------------------------------------------------------ */

/* EField::gauss, src/field.rs:1174-1195 */
void or_gauss(const or_lattice* l, const double* links, const double* e, size_t site,
              or_mat3* out)
{
    const size_t D = (size_t)l->dim;
    or_mat3 acc;
    mzero(&acc);
    for (int dir = 0; dir < l->dim; ++dir) {
        or_mat3 e_i, u_mi, u_mi_d, e_m_i, t, t2, term;
        or_adjoint_to_matrix(e + ((site * D + (size_t)dir) * 8), &e_i);
        or_link_matrix(l, links, site, dir, 0, &u_mi);
        size_t p_mi = or_neighbor(l, site, dir, 0);
        or_adjoint_to_matrix(e + ((p_mi * D + (size_t)dir) * 8), &e_m_i);
        or_matmul(&u_mi, &e_m_i, &t);
        or_adj(&u_mi, &u_mi_d);
        or_matmul(&t, &u_mi_d, &t2);
        msub(&e_i, &t2, &term);
        madd(&acc, &term, &acc);
    }
    *out = acc;
}

/* EField::gauss_sum_div, src/field.rs:1199-1220 */
double or_gauss_sum_div(const or_lattice* l, const double* links, const double* e)
{
    init_tables();
    double sum = 0.0;
    for (size_t s = 0; s < l->nsite; ++s) {
        or_mat3 g, p;
        or_gauss(l, links, e, s, &g);
        or_matmul(&G_SUM, &g, &p);
        or_cplx t = mtrace(&p);
        sum += hypot(t.re, t.im);
    }
    return sum;
}

/* EField::project_to_gauss_step, src/field.rs:1301-1337, K = 0.12 */
static void project_step(const or_lattice* l, const double* links, const double* e,
                         double* e_out)
{
    init_tables();
    const size_t D = (size_t)l->dim;
    const or_cplx K = {0.12, 0.0};
#pragma omp parallel for OMP_THREADS schedule(static)
    for (size_t s = 0; s < l->nsite; ++s) {
        for (int dir = 0; dir < l->dim; ++dir) {
            or_mat3 u, ud, g, gp, t1, t2, t3, t4, t5;
            or_link_matrix(l, links, s, dir, 1, &u);
            or_gauss(l, links, e, s, &g);
            size_t sp = or_neighbor(l, s, dir, 1);
            or_gauss(l, links, e, sp, &gp);
            or_matmul(&u, &g, &t1);
            or_adj(&u, &ud);
            or_matmul(&t1, &ud, &t2);
            or_matmul(&t2, &gp, &t3);
            msub(&t3, &g, &t4);
            mscal(&t4, K, &t5);
            for (int a = 0; a < 8; ++a) {
                or_mat3 ge, inner, prod;
                mscal(&G_TAB[a], cx(e[(s * D + (size_t)dir) * 8 + (size_t)a], 0.0), &ge);
                madd(&t5, &ge, &inner);
                or_matmul(&G_TAB[a], &inner, &prod);
                e_out[(s * D + (size_t)dir) * 8 + (size_t)a] = 2.0 * mtrace(&prod).re;
            }
        }
    }
}

/* EField::project_to_gauss, src/field.rs:1265-1294 (the stop rule uses
 * N * 4 * 8 * 10 * f64::EPSILON whatever D is).  max_iter bounds the
 * reference's unbounded loop; returns 0 on convergence, -1 on NaN,
 * -2 when max_iter outer iterations did not converge. */
int or_project_to_gauss(const or_lattice* l, const double* links, const double* e_in,
                        double* e_out, size_t max_iter, size_t* iterations)
{
    const size_t n = l->nsite * (size_t)l->dim * 8;
    double* tmp = (double*)malloc(sizeof(double) * n);
    project_step(l, links, e_in, e_out);
    size_t it = 0;
    int rc = -2;
    for (; it < max_iter; ++it) {
        double val = or_gauss_sum_div(l, links, e_out);
        if (isnan(val)) { rc = -1; break; }
        if (val <= F64_EPS * (double)(l->nsite * 4 * 8 * 10)) { rc = 0; break; }
        for (int k = 0; k < 4; ++k) {
            project_step(l, links, e_out, tmp);
            memcpy(e_out, tmp, sizeof(double) * n);
        }
    }
    if (iterations) *iterations = it;
    free(tmp);
    return rc;
}

/* ---- statistics (src/statistics/mod.rs) ------------------------------------
 * values: n reals, or n complex numbers as (re, im) pairs when cplx != 0.
 * Complex arithmetic is num-complex's: (a+bi)(c+di) = (ac-bd) + (ad+bc)i, no
 * conjugation anywhere (the complex "variance" is the mean of (x-m)^2).
 * Sums run left to right from zero, as Iterator::sum does. */

static or_cplx st_at(const double* x, size_t i, int cplx)
{
    or_cplx v = {cplx ? x[2 * i] : x[i], cplx ? x[2 * i + 1] : 0.0};
    return v;
}
static or_cplx st_add(or_cplx a, or_cplx b) { or_cplx r = {a.re + b.re, a.im + b.im}; return r; }
static or_cplx st_sub(or_cplx a, or_cplx b) { or_cplx r = {a.re - b.re, a.im - b.im}; return r; }
static or_cplx st_mul(or_cplx a, or_cplx b)
{
    or_cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static or_cplx st_div(or_cplx a, double d) { or_cplx r = {a.re / d, a.im / d}; return r; }
static void st_out(or_cplx v, int cplx, double* out)
{
    out[0] = v.re;
    if (cplx) out[1] = v.im;
}
static or_cplx st_sum(const double* x, size_t n, int cplx)
{
    or_cplx s = {0.0, 0.0};
    for (size_t i = 0; i < n; ++i) s = st_add(s, st_at(x, i, cplx));
    return s;
}

/* mean, mod.rs:321-335: sum / len (NaN for an empty input) */
void or_stats_mean(const double* x, size_t n, int cplx, double* out)
{
    st_out(st_div(st_sum(x, n, cplx), (double)n), cplx, out);
}

/* mean_and_variance, mod.rs:375-399: two passes, variance =
 * sum (x - mean)(x - mean) / (len - 1).  out = mean, variance (1 or 2 doubles
 * each).  Returns -1 for an empty input ((len - 1) underflows in the
 * reference). */
int or_stats_mean_and_variance(const double* x, size_t n, int cplx, double* out)
{
    if (n == 0) return -1;
    const or_cplx m = st_div(st_sum(x, n, cplx), (double)n);
    or_cplx v = {0.0, 0.0};
    for (size_t i = 0; i < n; ++i) {
        const or_cplx d = st_sub(st_at(x, i, cplx), m);
        v = st_add(v, st_mul(d, d));
    }
    st_out(m, cplx, out);
    st_out(st_div(v, (double)(n - 1)), cplx, out + (cplx ? 2 : 1));
    return 0;
}

/* mean_and_variance_par_iter_val, mod.rs:163-181: one pass over (x, x*x),
 * variance = (sum x^2 - (sum x)(sum x) / len) / (len - 1). */
int or_stats_mean_and_variance_par(const double* x, size_t n, int cplx, double* out)
{
    if (n == 0) return -1;
    or_cplx s = {0.0, 0.0}, s2 = {0.0, 0.0};
    for (size_t i = 0; i < n; ++i) {
        const or_cplx v = st_at(x, i, cplx);
        s = st_add(s, v);
        s2 = st_add(s2, st_mul(v, v));
    }
    const or_cplx var = st_div(st_sub(s2, st_div(st_mul(s, s), (double)n)), (double)(n - 1));
    st_out(st_div(s, (double)n), cplx, out);
    st_out(var, cplx, out + (cplx ? 2 : 1));
    return 0;
}

/* covariance, mod.rs:427-458: mean(x y) - mean(x) mean(y).  Returns 1 and
 * writes out when the lengths agree, 0 (the reference's None) otherwise. */
int or_stats_covariance(const double* x, size_t nx, const double* y, size_t ny, int cplx, double* out)
{
    if (nx != ny) return 0;
    or_cplx sp = {0.0, 0.0};
    for (size_t i = 0; i < nx; ++i) sp = st_add(sp, st_mul(st_at(x, i, cplx), st_at(y, i, cplx)));
    const or_cplx mp = st_div(sp, (double)nx);
    const or_cplx mx = st_div(st_sum(x, nx, cplx), (double)nx);
    const or_cplx my = st_div(st_sum(y, ny, cplx), (double)ny);
    st_out(st_sub(mp, st_mul(mx, my)), cplx, out);
    return 1;
}

/* covariance_par_iter_val, mod.rs:279-306: one pass over (x, y, x y),
 * (sum x y - (sum x)(sum y) / len) / len.  Same return convention. */
int or_stats_covariance_par(const double* x, size_t nx, const double* y, size_t ny, int cplx, double* out)
{
    if (nx != ny) return 0;
    or_cplx sx = {0.0, 0.0}, sy = {0.0, 0.0}, sp = {0.0, 0.0};
    for (size_t i = 0; i < nx; ++i) {
        const or_cplx a = st_at(x, i, cplx), b = st_at(y, i, cplx);
        sx = st_add(sx, a);
        sy = st_add(sy, b);
        sp = st_add(sp, st_mul(a, b));
    }
    const double len = (double)nx;
    st_out(st_div(st_sub(sp, st_div(st_mul(sx, sy), len)), len), cplx, out);
    return 1;
}
