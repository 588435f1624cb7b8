/*
 * oracle.h -- CPU restatement (C, f64) of lattice-qcd-rs's gauge-update path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (the HIP library behind include/lqcd.h) never links or calls it.
 *
 * Every function names the reference file:line it restates.  Layouts are the
 * reference's in-memory layouts:
 *   links  : [site][mu] 3x3 complex, COLUMN-MAJOR (nalgebra Matrix3), re/im
 *            interleaved -> 18 doubles per link, link index = site*D + mu
 *            (src/lattice.rs:909-929, src/field.rs:584-586)
 *   efield : [site][mu][8] doubles (src/field.rs:1025-1027)
 *   site   : sum_i x_i * prod_{k<i} L_k, x_0 fastest (src/lattice.rs:909-915)
 * The reference is hypercubic (one extent); the oracle accepts one extent per
 * direction, which reduces to the reference when all extents are equal.
 *
 * Parity status of each part is stated in the header comment of oracle.c.
 */
#ifndef LQCD_ORACLE_H
#define LQCD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_DIM 8

typedef struct { double re, im; } or_cplx;
/* column-major: m[col*3 + row] */
typedef struct { or_cplx m[9]; } or_mat3;

typedef struct {
    int dim;                  /* D */
    size_t extent[OR_MAX_DIM];
    size_t stride[OR_MAX_DIM];
    size_t nsite;
    double size;              /* lattice spacing a */
} or_lattice;

/* ---- lattice geometry (src/lattice.rs) ---- */
int    or_lattice_init(or_lattice* l, int dim, const size_t* extent, double size);
size_t or_site_index(const or_lattice* l, const size_t* x);
void   or_site_coords(const or_lattice* l, size_t idx, size_t* x);
size_t or_neighbor(const or_lattice* l, size_t site, int dir, int positive);
int    or_site_parity(const or_lattice* l, size_t site);
size_t or_link_canonical(const or_lattice* l, size_t site, int dir, int positive);
void   or_link_order(const or_lattice* l, size_t* out_site, int* out_dir);

/* ---- su(3)/SU(3) algebra (src/su3.rs, src/field.rs:106-205) ---- */
void    or_generator(int a, or_mat3* out);
void    or_adjoint_to_matrix(const double* v, or_mat3* out);
double  or_trace_squared(const double* v);
or_cplx or_adjoint_d(const double* v);
void    or_su3_exp_i(const double* v, or_mat3* out);
void    or_su3_exp_r(const double* v, or_mat3* out);
void    or_orthonormalize(const or_mat3* in, or_mat3* out);
or_cplx or_det(const or_mat3* m);
void    or_matmul(const or_mat3* a, const or_mat3* b, or_mat3* c);
void    or_adj(const or_mat3* a, or_mat3* c);
unsigned long long or_factorial_size_for_exp(void);

/* ---- random inputs: xoshiro256++ stream + the reference's samplers ---- */
typedef struct { uint64_t s[4]; } or_rng;
void     or_rng_seed(or_rng* r, uint64_t seed);
uint64_t or_rng_next(or_rng* r);
double   or_rng_uniform(or_rng* r, double low, double high);
int      or_rng_bernoulli(or_rng* r, double p);
double   or_rng_normal(or_rng* r, double mean, double sigma);
void     or_random_su3(or_rng* r, or_mat3* out);
void     or_random_su2_close_to_unity(or_rng* r, double spread, or_cplx out[4]);
void     or_random_su3_close_to_unity(or_rng* r, double spread, or_mat3* out);
void     or_hot_start(const or_lattice* l, uint64_t seed, double* links);
void     or_cold_start(const or_lattice* l, double* links);
void     or_random_efield(const or_lattice* l, uint64_t seed, double sigma, double* e);
void     or_mh_proposals(const or_lattice* l, uint64_t seed, double spread,
                         size_t n_update, double* proposals);

/* counter-based stream used by the device-side generators */
void     or_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void     or_device_hot_start(const or_lattice* l, uint64_t seed, double* links);
void     or_device_mh_inputs(const or_lattice* l, uint64_t seed, uint64_t counter, double spread,
                             size_t n_update, double* proposals, uint64_t* uniforms);
void     or_device_momenta(const or_lattice* l, uint64_t seed, uint64_t counter,
                           double sigma, double* e);

/* accept rule of rand 0.8 Bernoulli fed one u64 word */
int      or_bernoulli_accept(double p, uint64_t u);
double   or_clamp_probability(double x);

/* ---- stencils and observables (src/field.rs, src/simulation/state.rs) ---- */
void    or_link_matrix(const or_lattice* l, const double* links, size_t site, int dir,
                       int positive, or_mat3* out);
void    or_sij(const or_lattice* l, const double* links, size_t site, int di, int pi,
               int dj, int pj, or_mat3* out);
void    or_pij(const or_lattice* l, const double* links, size_t site, int di, int pi,
               int dj, int pj, or_mat3* out);
void    or_clover(const or_lattice* l, const double* links, size_t site, int di, int dj,
                  or_mat3* out);
void    or_f_mu_nu(const or_lattice* l, const double* links, size_t site, int di, int dj, or_mat3* out);
int     or_levi_civita(const int* index, int n);
void    or_magnetic_field(const or_lattice* l, const double* links, size_t site, int dir, or_mat3* out);
void    or_clover_field(const or_lattice* l, const double* links, double* out);
void    or_f_mu_nu_field(const or_lattice* l, const double* links, double* out);
void    or_magnetic_field_vec(const or_lattice* l, const double* links, double* out);
or_cplx or_average_trace_plaquette(const or_lattice* l, const double* links);
double  or_hamiltonian_links(const or_lattice* l, const double* links, double beta);
double  or_hamiltonian_efield(const or_lattice* l, const double* e, double beta);
void    or_derivative_e(const or_lattice* l, const double* links, size_t site, double* out);
void    or_derivative_e_all(const or_lattice* l, const double* links, double* out);
void    or_derivative_u(const or_lattice* l, const double* links, const double* e,
                        size_t site, int dir, or_mat3* out);
void    or_wilson_backward(const or_lattice* l, const double* links, double beta,
                           double* grad);

/* ---- integrators (src/integrator/) ---- */
enum { OR_DRIFT_ADDITIVE = 0, OR_DRIFT_EXP = 1 };
enum { OR_SCHEME_REFERENCE = 0, OR_SCHEME_LEAPFROG = 1 };
void or_integrate_efield(const or_lattice* l, const double* links, const double* e_in,
                         double dt, double* e_out);
void or_integrate_link(const or_lattice* l, const double* links, const double* e,
                       double dt, int mode, double* links_out);
void or_integrate_symplectic(const or_lattice* l, double* links, double* e, double dt,
                             int mode);
int  or_simulate_symplectic_n(const or_lattice* l, double* links, double* e, double dt,
                              size_t n, int mode, int scheme);

/* ---- Monte Carlo (src/simulation/monte_carlo/) ---- */
typedef struct {
    double delta_h;       /* H_new - H_old */
    double h_old, h_new;
    double probability;   /* prob_replace_last */
    int    accepted;      /* has_replace_last */
} or_hmc_stats;
int  or_hmc_trajectory(const or_lattice* l, double* links, const double* momenta,
                       double beta, double dt, size_t n, int mode, int scheme,
                       uint64_t uniform, or_hmc_stats* st);

void   or_staple(const or_lattice* l, const double* links, size_t site, int dir, or_mat3* out);
double or_delta_s(const or_lattice* l, const double* links, size_t site, int dir,
                  const or_mat3* new_link, double beta);

enum { OR_ORDER_LEXICOGRAPHIC = 0, OR_ORDER_CHECKERBOARD = 1 };
typedef struct {
    double prob_replace_mean;
    size_t number_replace_last;
} or_mh_stats;
int or_metropolis_sweep(const or_lattice* l, double* links, double beta, size_t n_update,
                        const double* proposals, const uint64_t* uniforms, int order,
                        or_mh_stats* st);
void or_su2_project_unorm(const or_cplx m[4], or_cplx out[4]);
int or_heat_bath_sweep(const or_lattice* l, double* links, double beta, uint64_t seed,
                       uint64_t counter, int order, size_t* exhausted);
double or_heat_bath_x0(double p, uint64_t l, uint64_t n, uint64_t seed);
void or_polar_unitary(const or_mat3* a, or_mat3* w);
int or_overrelaxation_sweep(const or_lattice* l, double* links, int order);
void or_svd3(const or_mat3* a, or_mat3* u, double sv[3], or_mat3* v);
int or_overrelaxation_reverse_sweep(const or_lattice* l, double* links, int order);

void or_normalize_link_matrices(const or_lattice* l, double* links);

/* ---- Gauss law (src/field.rs:1172-1337) ---- */
void   or_gauss(const or_lattice* l, const double* links, const double* e, size_t site,
                or_mat3* out);
double or_gauss_sum_div(const or_lattice* l, const double* links, const double* e);
int    or_project_to_gauss(const or_lattice* l, const double* links, const double* e_in,
                           double* e_out, size_t max_iter, size_t* iterations);

/* ---- statistics (src/statistics/mod.rs) ----
 * values: n reals, or n complex numbers as (re, im) pairs when cplx != 0 */
void or_stats_mean(const double* x, size_t n, int cplx, double* out);
int  or_stats_mean_and_variance(const double* x, size_t n, int cplx, double* out);
int  or_stats_mean_and_variance_par(const double* x, size_t n, int cplx, double* out);
int  or_stats_covariance(const double* x, size_t nx, const double* y, size_t ny, int cplx, double* out);
int  or_stats_covariance_par(const double* x, size_t nx, const double* y, size_t ny, int cplx,
                             double* out);

/* threads used by the parallel loops (0 = OpenMP default) */
void or_set_threads(int n);
int  or_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
