/*
 * lqcd.h -- C ABI of the MI355X (gfx950) gauge-update library.
 *
 * This is the drop-in boundary under the reference's trait API
 * (lattice-qcd-rs: LatticeState / LatticeStateWithEField / SymplecticIntegrator
 * / MonteCarlo).  A Rust FFI crate binds these functions one to one; each
 * declaration cites the reference interface it replaces.  Plain C types only:
 * handles, pointers, sizes, doubles.
 *
 * Conventions
 *  - Every function returns an lqcd_status; lqcd_last_error() gives the
 *    message of the last failure on the calling thread.
 *  - Host buffers are owned by the caller and use the reference's in-memory
 *    layouts:
 *      links  [site][mu] 3x3 complex, COLUMN-major (nalgebra Matrix3), re/im
 *             interleaved: 18 doubles per link, link index = site*D + mu
 *             (src/field.rs:584-586, src/lattice.rs:923-927);
 *      efield [site][mu][8] doubles (src/field.rs:1025-1027);
 *      site   = sum_i x_i * prod_{k<i} L_k (src/lattice.rs:909-915).
 *  - Device buffers are owned by the library; fields stay resident in HBM
 *    between calls.  Calls are synchronous unless stated otherwise; the
 *    ones that only enqueue device work (lqcd_efield_random,
 *    lqcd_normalize_link_matrices) return once it is enqueued on the
 *    context's stream: every later call on the context sees their result,
 *    and lqcd_ctx_synchronize waits for them.
 *  - A context owns one HIP stream and is not thread-safe: use one context
 *    per host thread.
 */
#ifndef LQCD_H
#define LQCD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    LQCD_OK = 0,
    /* LatticeInitializationError (src/error.rs:219-241) */
    LQCD_ERR_ZERO_DIMENSION = 1,
    LQCD_ERR_NON_POSITIVE_SIZE = 2,
    LQCD_ERR_DIM_TOO_SMALL = 3,
    /* StateInitializationError::IncompatibleSize (src/error.rs:124-171) */
    LQCD_ERR_INCOMPATIBLE_SIZE = 4,
    /* MultiIntegrationError::ZeroIntegration (src/error.rs:93-98) */
    LQCD_ERR_ZERO_INTEGRATION = 5,
    LQCD_ERR_INVALID_ARGUMENT = 6,
    LQCD_ERR_UNSUPPORTED = 7,
    LQCD_ERR_DEVICE = 8,
    LQCD_ERR_OUT_OF_MEMORY = 9,
    /* StateInitializationError::GaussProjectionError (src/error.rs:131-132) */
    LQCD_ERR_GAUSS_PROJECTION = 10,
    LQCD_ERR_IO = 11
} lqcd_status;

typedef enum { LQCD_F64 = 0, LQCD_F32 = 1 } lqcd_precision;

/* link drift: ADDITIVE = the reference's U + dt * derivative_u
 * (src/integrator/mod.rs:216-233); EXP = exp(i dt sqrt(6)/a E) U with the
 * reference's su3_exp_i (src/su3.rs:832-855). */
typedef enum { LQCD_DRIFT_ADDITIVE = 0, LQCD_DRIFT_EXP = 1 } lqcd_drift;

/* REFERENCE = n calls of integrate_symplectic, kick(dt/2) drift(dt)
 * kick(dt/2) each (symplectic_euler_rayon.rs:222-252); LEAPFROG merges the
 * two half kicks between consecutive steps (n + 1 force evaluations), which
 * is simulate_using_leapfrog_n (src/simulation/state.rs:321-358). */
typedef enum { LQCD_SCHEME_REFERENCE = 0, LQCD_SCHEME_LEAPFROG = 1 } lqcd_scheme;

typedef struct lqcd_ctx lqcd_ctx;
typedef struct lqcd_lattice lqcd_lattice;

/* ---------------------------------------------------------------- misc */
const char* lqcd_last_error(void);
const char* lqcd_version(void);
int lqcd_device_count(int* count);

/* ------------------------------------------------------------- context */
/* One HIP stream on `device`.  The reference's rayon pool
 * (src/thread.rs:684-699) is the counterpart. */
int lqcd_ctx_create(int device, lqcd_ctx** out);
int lqcd_ctx_destroy(lqcd_ctx* ctx);
int lqcd_ctx_synchronize(lqcd_ctx* ctx);

/* ------------------------------------------------------------- lattice */
/* LatticeCyclic::new(size, dim) (src/lattice.rs:190-201), with one extent per
 * direction (all equal = the reference's hypercubic lattice).  dim in 2..4.
 * Allocates links (cold start) and a zero E field on the device. */
int lqcd_lattice_create(lqcd_ctx* ctx, int dim, const size_t* extent, double size,
                        int precision, lqcd_lattice** out);
int lqcd_lattice_destroy(lqcd_lattice* lat);
/* number_of_points / number_of_canonical_links_space (src/lattice.rs:216-245) */
int lqcd_lattice_info(const lqcd_lattice* lat, size_t* nsite, size_t* nlink);

/* LatticeState::set_link_matrix / link_matrix (src/simulation/state.rs:60-75):
 * n_links must equal nsite * dim (else LQCD_ERR_INCOMPATIBLE_SIZE). */
int lqcd_links_upload(lqcd_lattice* lat, const double* links, size_t n_links);
int lqcd_links_download(lqcd_lattice* lat, double* links, size_t n_links);
/* LinkMatrix::new_cold (src/field.rs:716-722) */
int lqcd_links_cold(lqcd_lattice* lat);

/* LinkMatrix::new_determinist hot start (src/field.rs:646-659 with random_su3,
 * src/su3.rs:322-347) drawn on the device from a Philox4x32-10 stream keyed
 * by seed (one independent sub-stream per link). */
int lqcd_links_random(lqcd_lattice* lat, uint64_t seed);

/* LatticeStateWithEField::set_e_field / e_field (src/simulation/state.rs:190-200) */
int lqcd_efield_upload(lqcd_lattice* lat, const double* efield, size_t n_links);
int lqcd_efield_download(lqcd_lattice* lat, double* efield, size_t n_links);
/* EField::new_cold (src/field.rs:1131-1141) */
int lqcd_efield_cold(lqcd_lattice* lat);
/* Momentum refresh drawn on the device (replaces EField::new_determinist with
 * Normal(0, sigma), src/field.rs:1086-1099; the reference uses
 * sigma = 0.5 / beta, src/simulation/state.rs:1097).  Counter-based stream:
 * Philox4x32-10 keyed by seed, counter selects the draw.  Enqueued only. */
int lqcd_efield_random(lqcd_lattice* lat, uint64_t seed, uint64_t counter, double sigma);

/* EField::gauss (src/field.rs:1174-1195) at every site:
 * G(x) = sum_i E_i(x) - U_{-i}(x) E_i(x - i) U_{-i}(x)^+, out [site][18]
 * column-major 3x3 complex.  Computed in double arithmetic. */
int lqcd_efield_gauss(lqcd_lattice* lat, double* out, size_t n_sites);
/* EField::gauss_sum_div (src/field.rs:1199-1220): sum_x |Tr((sum_a T^a) G(x))| */
int lqcd_efield_gauss_sum_div(lqcd_lattice* lat, double* out);
/* EField::project_to_gauss (src/field.rs:1265-1337) applied to the lattice's
 * E field: one project_to_gauss_step, then four more while gauss_sum_div >
 * N * 4 * 8 * 10 * f64::EPSILON.  The iteration runs on double copies of E
 * (also for f32 storage, where it is rounded once at the end).  The
 * reference loops without bound; here at most max_iter checks are made
 * (LQCD_ERR_GAUSS_PROJECTION if not converged or NaN, E unchanged then).
 * iterations (may be NULL) gets the number of checks that failed. */
int lqcd_efield_project_gauss(lqcd_lattice* lat, size_t max_iter, size_t* iterations);

/* ---------------------------------------------------------- observables */
/* LatticeState::average_trace_plaquette (src/simulation/state.rs:115-117,
 * src/field.rs:775-804): out[0] + i out[1] = sum Tr P / (N D (D-1)/2). */
int lqcd_average_trace_plaquette(lqcd_lattice* lat, double out[2]);
/* LatticeState::hamiltonian_links (src/simulation/state.rs:821-849) */
int lqcd_hamiltonian_links(lqcd_lattice* lat, double beta, double* out);
/* LatticeStateWithEField::hamiltonian_efield (src/simulation/state.rs:1370-1385) */
int lqcd_hamiltonian_efield(lqcd_lattice* lat, double beta, double* out);

/* --------------------------------------------------- molecular dynamics */
/* LatticeStateWithEField::derivative_e for every site
 * (src/simulation/state.rs:1420-1448); out is [site][mu][8]. */
int lqcd_derivative_e(lqcd_lattice* lat, double* out, size_t n_links);
/* LatticeStateWithEField::derivative_u (src/simulation/state.rs:1407-1417)
 * of every link: ((E U) i sqrt(2 CA)) / a with the lattice's E, out in the
 * links' host layout [site][mu][18]. */
int lqcd_derivative_u(lqcd_lattice* lat, double* out, size_t n_links);
/* integrate_efield over all sites (src/integrator/mod.rs:240-254):
 * e += dt * derivative_e(U). */
int lqcd_integrate_efield(lqcd_lattice* lat, double dt);
/* integrate_link over all links (src/integrator/mod.rs:216-233): U <- drift(U, e, dt). */
int lqcd_integrate_link(lqcd_lattice* lat, double dt, int drift);
/* SimulationStateSynchronous::simulate_symplectic_n (src/simulation/state.rs:470-492)
 * with SymplecticIntegrator::integrate_symplectic
 * (src/integrator/symplectic_euler_rayon.rs:222-252).  n == 0 ->
 * LQCD_ERR_ZERO_INTEGRATION. */
int lqcd_simulate_symplectic_n(lqcd_lattice* lat, double dt, size_t n, int drift, int scheme);

/* The other transitions of SymplecticIntegrator (src/integrator/mod.rs:93-208)
 * as SymplecticEulerRayon implements them.  The handle does not know whether
 * its E is at a synchronous or a half step (SimulationStateLeap); the caller
 * keeps track, as the reference's types do.
 *   sync_sync (symplectic_euler_rayon.rs:127-144): U' = drift(U, E, dt) and
 *       E' = E + dt F(U), both from the state before the call;
 *   sync_leap (:173-193): E' = E + dt/2 F(U), U unchanged;
 *   leap_leap (:146-171): U' = drift(U, E, dt), then E' = E + dt F(U');
 *   leap_sync (:195-220): U' = drift(U, E, dt), then E' = E + dt/2 F(U').
 * simulate_sync_n / simulate_leap_n repeat sync_sync / leap_leap n times
 * (src/simulation/state.rs:397-418, 624-648; n == 0 ->
 * LQCD_ERR_ZERO_INTEGRATION).  sync_leap, n - 1 leap_leap and leap_sync are
 * simulate_using_leapfrog_n (state.rs:321-358), which is also
 * lqcd_simulate_symplectic_n with LQCD_SCHEME_LEAPFROG. */
int lqcd_integrate_sync_sync(lqcd_lattice* lat, double dt, int drift);
int lqcd_integrate_sync_leap(lqcd_lattice* lat, double dt);
int lqcd_integrate_leap_leap(lqcd_lattice* lat, double dt, int drift);
int lqcd_integrate_leap_sync(lqcd_lattice* lat, double dt, int drift);
int lqcd_simulate_sync_n(lqcd_lattice* lat, double dt, size_t n, int drift);
int lqcd_simulate_leap_n(lqcd_lattice* lat, double dt, size_t n, int drift);

/* ------------------------------------------------------------------ HMC */
typedef struct {
    double delta_h;      /* H_new - H_old */
    double h_old;
    double h_new;
    double probability;  /* HybridMonteCarloDiagnostic::prob_replace_last */
    int accepted;        /* HybridMonteCarloDiagnostic::has_replace_last */
} lqcd_hmc_stats;

/* One trajectory of HybridMonteCarloDiagnostic::next_element
 * (src/simulation/monte_carlo/hybrid_monte_carlo.rs:465-471,584-613) with the
 * momenta already in the lattice's E field (upload or lqcd_efield_random):
 * H_old, n symplectic steps, H_new, p = clamp(exp(H_old - H_new)) (NaN -> 1),
 * accept iff p == 1 or uniform < floor(p * 2^64) (rand 0.8 Bernoulli).  On
 * reject the links are those before the call.  E holds the final momenta. */
int lqcd_hmc_trajectory(lqcd_lattice* lat, double beta, double dt, size_t n, int drift,
                        int scheme, uint64_t uniform, lqcd_hmc_stats* stats);

/* ------------------------------------------------------ Wilson adjoint */
/* Forward: S_W = beta * sum_{x, i<j} (1 - Re Tr P_ij(x) / 3), the action of
 * src/simulation/state.rs:821-849. */
int lqcd_wilson_forward(lqcd_lattice* lat, double beta, double* action);
/* Backward: grad[site][mu][a] = dS_W / dw^a for U_mu(x) -> exp(i w^a T^a) U_mu(x);
 * equals -(beta a / sqrt 6) * derivative_e (src/simulation/state.rs:1420-1448).
 * Staples are recomputed, nothing is stored by the forward pass. */
int lqcd_wilson_backward(lqcd_lattice* lat, double beta, double* grad, size_t n_links);
/* The same gradient into a caller-owned device buffer (d_grad: device memory
 * of the lattice's GPU, n_links x 8 doubles), for callers that keep it on the
 * device: no host copy.  No counterpart in the reference. */
int lqcd_wilson_backward_device(lqcd_lattice* lat, double beta, double* d_grad, size_t n_links);

/* ----------------------------------------------------------- Metropolis */
typedef struct {
    double prob_replace_mean;    /* MetropolisHastingsSweep::prob_replace_mean */
    size_t number_replace_last;  /* MetropolisHastingsSweep::number_replace_last */
} lqcd_mh_stats;

/* One MetropolisHastingsSweep::next_element
 * (src/simulation/monte_carlo/metropolis_hastings_sweep.rs:146-174) in
 * checkerboard order: for mu = 0..D-1, parity 0 then 1, all links of that
 * class at once.  proposals: [n_update][link][18] column-major 3x3 complex
 * (already orthonormalized, applied as R_k...R_1 U); uniforms: one u64 per
 * link.  Needs even extents (else LQCD_ERR_UNSUPPORTED). */
int lqcd_metropolis_sweep(lqcd_lattice* lat, double beta, size_t n_update,
                          const double* proposals, const uint64_t* uniforms,
                          lqcd_mh_stats* stats);
/* The same sweep with the randomness drawn on the device, replacing the
 * rng argument of MetropolisHastingsSweep::next_element
 * (metropolis_hastings_sweep.rs:146-174) and potential_modif (:126-143).
 * Proposal j of link l is orthonormalize(random_su3_close_to_unity(spread))
 * built from 14 uniforms of Philox4x32-10 blocks 64 + 8j + (0..6), the
 * accept word from block 63; counter selects the sweep, key = seed.  The
 * stream differs from the reference's xoshiro draws but every draw has the
 * reference's distribution.  n_update <= 4000, 0 < spread < 1. */
int lqcd_metropolis_sweep_random(lqcd_lattice* lat, double beta, size_t n_update,
                                 double spread, uint64_t seed, uint64_t counter,
                                 lqcd_mh_stats* stats);

/* One MetropolisHastingsDeltaDiagnostic::next_element
 * (src/simulation/monte_carlo/metropolis_hastings.rs:374-419): a link drawn
 * uniformly, its proposal orthonormalize(random_su3_close_to_unity(spread))
 * U_old, dS from the staple (monte_carlo/mod.rs:324-362), accepted with
 * p = clamp(exp(-dS), 0, 1) by the rand 0.8 Bernoulli rule.  The draws come
 * from the device Philox stream (seed, counter = the step): the link from
 * block 62 of link 0, the proposal and accept word of that link as in
 * lqcd_metropolis_sweep_random.  Outputs (each may be NULL): the probability
 * (prob_replace_last), whether it was accepted (has_replace_last) and the
 * link index site * D + mu. */
int lqcd_metropolis_link_random(lqcd_lattice* lat, double beta, double spread, uint64_t seed, uint64_t counter,
                                double* probability, int* accepted, size_t* link);
/* HeatBathSweep::next_element (src/simulation/monte_carlo/heat_bath.rs:73-123)
 * in checkerboard order (as lqcd_metropolis_sweep), double arithmetic: per
 * link, three SU(2) subgroup heat baths on the r, s, t sub-blocks of U A
 * (su3.rs:428-660), each drawn from HeatBathDistribution(beta k)
 * (statistics/distribution.rs:147-352, including the reference's
 * sigma_3 = identity in complex_matrix_from_vec, so links leave SU(3) as in
 * the reference) or su2::random_su2 when k is not normal.  Draws come from
 * Philox4x32-10 (key seed, counter = sweep) instead of the rng argument;
 * the norm sampler's rejection loop is cut after 64 tries per subgroup
 * (factor = identity then), *exhausted (may be NULL) counts those.
 * beta > 0, even extents. */
int lqcd_heat_bath_sweep(lqcd_lattice* lat, double beta, uint64_t seed, uint64_t counter,
                         size_t* exhausted);
/* OverrelaxationSweepRotation::next_element
 * (src/simulation/monte_carlo/overrelaxation.rs:66-91) in checkerboard order:
 * new link = rot U^+ rot with rot = U V^+ of the SVD of staple^+, computed
 * as the unitary polar factor (scaled Newton iteration, double arithmetic).
 * Leaves the Wilson action unchanged.  Even extents. */
int lqcd_overrelaxation_sweep(lqcd_lattice* lat);
/* OverrelaxationSweepReverse::next_element
 * (src/simulation/monte_carlo/overrelaxation.rs:136-183) in checkerboard
 * order: with staple^+ = W S V^+ (one-sided Jacobi SVD, double arithmetic),
 * new link = W reverse(W^+ U V) V^+, su3::reverse (src/su3.rs:705-714)
 * negating the off-diagonal entries.  Leaves the Wilson action unchanged.
 * Even extents. */
int lqcd_overrelaxation_reverse_sweep(lqcd_lattice* lat);

/* ------------------------------------------------------- field strength */
/* LinkMatrix::clover (src/field.rs:806-820) at every site for every ordered
 * pair (i, j) of positive directions, i == j included: out[(site*D + i)*D + j]
 * is clover(x, i, j), 18 doubles (column-major 3x3 complex, the layout of one
 * link).  n = nsite * D * D * 18.  Double arithmetic on either storage type. */
int lqcd_clover_field(lqcd_lattice* lat, double* out, size_t n);
/* LinkMatrix::sij (which 0) or pij (which 1) (src/field.rs:743-771) at every
 * site for the signed directions (dir_i, positive_i), (dir_j, positive_j):
 * S_ij(x) = U_j(x) U_i(x+j) U_j(x+i)^+, P_ij(x) = U_i(x) S_ij(x)^+ with
 * U_{-d}(y) = U_d(y-d)^+; out[site] 18 doubles (column-major 3x3 complex),
 * n = nsite * 18. */
int lqcd_plaquette_field(lqcd_lattice* lat, int which, int dir_i, int positive_i, int dir_j, int positive_j,
                         double* out, size_t n);
/* LinkMatrix::f_mu_nu (src/field.rs:822-834), (clover(i,j) - clover(j,i)) /
 * (8 a^2), same layout and size as lqcd_clover_field. */
int lqcd_f_mu_nu_field(lqcd_lattice* lat, double* out, size_t n);
/* LinkMatrix::magnetic_field_vec (src/field.rs:836-868) at every site:
 * out[site*D + dir] = sum_{i,j} levi_civita(dir, i, j) F_ij / (2i), 18 doubles
 * each, n = nsite * D * 18 (zero for D = 2). */
int lqcd_magnetic_field_vec(lqcd_lattice* lat, double* out, size_t n);
/* The same three fields into a caller-owned device buffer (d_out: device
 * memory of the lattice's GPU, n doubles as above), for callers that keep
 * the result on the device.  No counterpart in the reference. */
enum { LQCD_FIELD_CLOVER = 0, LQCD_FIELD_F_MU_NU = 1, LQCD_FIELD_MAGNETIC = 2 };
int lqcd_field_strength_device(lqcd_lattice* lat, int what, double* d_out, size_t n);

/* ------------------------------------------------------ state interchange */
/* The serde data model of LatticeStateDefault<D> (src/simulation/state.rs:653-659)
 *   { lattice: { size, dim }, beta, link_matrix: { data: [matrix; n] } }
 * and of LatticeStateEFSyncDefault<LatticeStateDefault<D>, D> (state.rs:1046-1062)
 *   { e_field: { data: [[ { data: [8 reals] }; D ]; nsite] }, t, lattice_state },
 * a matrix being the tuple of its 9 entries in column-major order and an entry
 * the tuple (re, im) (nalgebra 0.31 / num-complex 0.4 serde), in one of two
 * serde formats: bincode 1.x `serialize` (little-endian fixed-width integers,
 * u64 sequence lengths) or serde_json.  D is not serialized: readers take it
 * from the caller and check the lengths against dim^D.  Only hypercubic
 * lattices (the reference's LatticeCyclic) can be written. */
enum { LQCD_FORMAT_BINCODE = 0, LQCD_FORMAT_JSON = 1 };
int lqcd_state_write(lqcd_lattice* lat, double beta, const char* path, int format);
int lqcd_state_ef_write(lqcd_lattice* lat, double beta, uint64_t t, const char* path, int format);
/* checks a whole file (with_efield: the EFSync layout; bincode: every length
 * prefix against the file size, which is the whole check for fixed-width
 * fields; JSON: a full parse) and returns its lattice spacing, points per
 * direction, beta and t (t may be NULL); needs no device */
int lqcd_state_read_header(const char* path, int format, int with_efield, int dim, double* size,
                           size_t* points, double* beta, uint64_t* t);
/* the same parse into host buffers: links n_links x 18 (link index site*D+mu,
 * as lqcd_links_download), efield n_links x 8 when with_efield; no device */
int lqcd_state_read_host(const char* path, int format, int with_efield, int dim, double* links,
                         size_t n_links, double* efield);
/* loads the links (and E) of a file into a lattice of the same shape and
 * spacing; beta and t (may be NULL) receive the stored values.  The whole
 * file is read before the lattice is overwritten: a file that is refused
 * leaves the lattice as it was */
int lqcd_state_read(lqcd_lattice* lat, const char* path, int format, int with_efield, double* beta,
                    uint64_t* t);

/* LatticeStateDefault::normalize_link_matrices (src/simulation/state.rs:754-756).
 * Enqueued only. */
int lqcd_normalize_link_matrices(lqcd_lattice* lat);

/* ------------------------------------------------------------- tuning */
/* Work decomposition of the MD kernels.  No counterpart in the reference
 * (the rayon pool schedules per-element closures, src/thread.rs:684-699).
 *   0  one thread per link, grid (site blocks, D)
 *   1  one thread per link, XCD-grouped 1-D grid (default before tuning)
 *   2  one thread per site for all D links (shared staple loads), XCD-grouped
 *   3  as 2, each XCD walking a slab of the slowest direction with that
 *      direction fastest (neighbour slices share L2)
 * Variants differ only in rounding (summation order).  The first integration
 * of a lattice autotunes unless a variant was set or LQCD_MD_VARIANT is set
 * in the environment. */
int lqcd_md_autotune(lqcd_lattice* lat, int drift, int* variant);
int lqcd_md_variant_set(lqcd_lattice* lat, int variant);
int lqcd_md_variant_get(const lqcd_lattice* lat, int* variant);
/* Two-row link reads in exp-drift integrations.  No counterpart in the
 * reference.  With mode 1 (default) an integration with LQCD_DRIFT_EXP first
 * measures the distance of the links from SU(3) (lqcd_su3_distance); when it
 * is at most 64 ulps of 1 in the storage type (3.8e-6 for f32, 7.1e-15 for
 * f64) the MD kernels read only the first two rows of every link and rebuild
 * the third as conj(row0 x row1), exact on SU(3); products of SU(3) factors
 * are formed the same way.  Mode 0 reads all nine entries always.
 * Device-proposal Metropolis sweeps (lqcd_metropolis_sweep_random) read the
 * staple links the same way: their proposals are in SU(3).  The library
 * keeps track of what is known (hot / cold start and reunitarisation give
 * SU(3); the additive drift, heat bath and overrelaxation leave it; uploads
 * and host proposals make it unknown) and measures again after 200
 * SU(3)-preserving updates per link.  *last_used tells whether the last
 * integration or device Metropolis sweep read two rows. */
int lqcd_md_reconstruct_set(lqcd_lattice* lat, int mode);
/* 1 (default): the MD launches of an integration are replayed from a HIP
 * graph captured on first use for the starting buffer and parameters (the
 * reference scheme only while profiling is off); 0: launched one by one.
 * Same results; the graph removes the ~10 us the device leaves between eager
 * launches.  No counterpart in the reference. */
int lqcd_md_graph_set(lqcd_lattice* lat, int enable);
int lqcd_md_reconstruct_get(const lqcd_lattice* lat, int* mode, int* last_used);
/* max over links of max(max_ij |(U U^+ - 1)_ij|, |det U - 1|), double arithmetic */
int lqcd_su3_distance(lqcd_lattice* lat, double* out);

/* ------------------------------------------------------ batched chains */
/* Many independent chains, each a LatticeStateDefault of its own extents,
 * beta and seed advanced by its own HybridMonteCarloDiagnostic
 * (src/simulation/monte_carlo/hybrid_monte_carlo.rs:465-471, 584-613), kept
 * in one set of device buffers and advanced together: one launch per MD
 * step covers every chain.  The reference runs such chains one after the
 * other; there is no batch type in it.  All chains share D and precision. */
typedef struct lqcd_batch lqcd_batch;

typedef struct {
    size_t extent[4];  /* the first D entries are used, each >= 2 */
    double beta;
    double size;       /* lattice spacing a > 0 */
    uint64_t seed;     /* hot start, momenta and accept streams of the chain */
} lqcd_chain_spec;

typedef struct {
    double h_old;        /* H before the trajectory (hamiltonian_total) */
    double h_new;        /* H at its end */
    double delta_h;      /* h_new - h_old */
    double probability;  /* probability_of_replacement */
    double plaquette;    /* Re average_trace_plaquette / 3 of the state kept */
    int32_t accepted;
    int32_t reserved;
} lqcd_chain_stats;

/* LatticeCyclic::new / LatticeStateDefault::new_cold per chain (error codes
 * as lqcd_lattice_create); every chain starts cold. */
int lqcd_batch_create(lqcd_ctx* ctx, int dim, int precision, const lqcd_chain_spec* chains,
                      size_t nchain, lqcd_batch** out);
int lqcd_batch_destroy(lqcd_batch* batch);
int lqcd_batch_info(const lqcd_batch* batch, size_t* nchain, size_t* nsite_total);
int lqcd_batch_chain_info(const lqcd_batch* batch, size_t chain, size_t* nsite, size_t* nlink);
/* LinkMatrix::new_cold (src/field.rs:716-722) for every chain */
int lqcd_batch_cold_start(lqcd_batch* batch);
/* every chain drawn as lqcd_links_random(seed of the chain) */
int lqcd_batch_hot_start(lqcd_batch* batch);
/* LatticeState::link_matrix / set_link_matrix of one chain */
int lqcd_batch_links_upload(lqcd_batch* batch, size_t chain, const double* links, size_t n_links);
int lqcd_batch_links_download(lqcd_batch* batch, size_t chain, double* links, size_t n_links);
/* Re average_trace_plaquette / 3 (src/field.rs:775-804) of every chain */
int lqcd_batch_plaquette(lqcd_batch* batch, double* plaquette);
/* The trajectory counter n of the random streams: trajectory n of chain c
 * draws its momenta as lqcd_efield_random(seed_c, n, 0.5 / beta_c) and its
 * accept word as (x << 32) | y of Philox4x32-10 counter {0, 62 << 16,
 * lo32(n), hi32(n)}, key seed_c.  Starts at 0, +1 per trajectory. */
int lqcd_batch_counter_get(const lqcd_batch* batch, uint64_t* counter);
int lqcd_batch_counter_set(lqcd_batch* batch, uint64_t counter);
/* n_traj HMC trajectories of every chain: momenta refresh, H, n_steps of
 * the scheme with step dt, H, accept on the device, rejected chains keep
 * their links; reunitarize != 0 then applies normalize_link_matrices
 * (src/simulation/state.rs:754-756).  stats: [n_traj][nchain] or NULL.
 * One trajectory is captured once as a HIP graph and replayed (see
 * lqcd_batch_graph_enable). */
int lqcd_batch_hmc(lqcd_batch* batch, double dt, size_t n_steps, int drift, int scheme,
                   size_t n_traj, int reunitarize, lqcd_chain_stats* stats);
/* Launch shape of the batched MD kernel, config in 0..11:
 *   config % 3        workgroup of 64 / 128 / 256 threads
 *   (config / 3) % 2  1: one thread per site for all D links; 0: per link
 *   config / 6        1: persistent grid (resident workgroups walk tile
 *                     ranges); 0: one workgroup per 1-4 tiles
 * Shapes differ only in rounding (staple summation order).  The first
 * lqcd_batch_hmc autotunes unless a config was set or LQCD_BATCH_CONFIG is
 * set in the environment. */
int lqcd_batch_autotune(lqcd_batch* batch, int drift, int* config);
int lqcd_batch_config_set(lqcd_batch* batch, int config);
int lqcd_batch_config_get(const lqcd_batch* batch, int* config);
/* 1 (default): replay a captured graph per trajectory; 0: launch eagerly */
int lqcd_batch_graph_enable(lqcd_batch* batch, int enable);
/* as lqcd_md_reconstruct_set / _get for the chains: lqcd_batch_hmc with
 * LQCD_DRIFT_EXP and reunitarize set reads the links as two rows when all
 * chains' links are within 64 ulps of SU(3).  No counterpart in the
 * reference. */
int lqcd_batch_reconstruct_set(lqcd_batch* batch, int mode);
int lqcd_batch_reconstruct_get(const lqcd_batch* batch, int* mode, int* last_used);

/* ------------------------------------------------------------ profiling */
/* Kernel kinds for lqcd_profile_read (HIP events around each launch). */
typedef enum {
    LQCD_KERNEL_KICK_DRIFT = 0,
    LQCD_KERNEL_KICK = 1,
    LQCD_KERNEL_DRIFT = 2,
    LQCD_KERNEL_PLAQUETTE = 3,
    LQCD_KERNEL_EFIELD_ENERGY = 4,
    LQCD_KERNEL_REUNIT = 5,
    LQCD_KERNEL_METROPOLIS = 6,
    LQCD_KERNEL_MOMENTA = 7,
    LQCD_KERNEL_FORCE = 8,
    LQCD_KERNEL_FIELDS = 9,
    LQCD_KERNEL_STATS = 10,
    LQCD_KERNEL_HEAT_BATH = 11,
    LQCD_KERNEL_OVERRELAX = 12
} lqcd_kernel_kind;

/* ---------------------------------------------------------- statistics */
/* The statistics of src/statistics/mod.rs over values resident in device
 * memory (a device pointer from hipMalloc, or lqcd_field_strength_device's
 * output): n doubles, or, when complex_values != 0, n complex numbers stored
 * as (re, im) pairs, in which case every result is a (re, im) pair too.
 * Complex arithmetic is num-complex's, with no conjugation (the complex
 * variance is the mean of (x - m)(x - m), as in the reference).  The sums
 * run on the context's stream in a fixed order: results are the same run to
 * run, and agree with the reference's sequential sums to rounding.  The calls
 * synchronize the context's stream. */
typedef enum {
    /* the sequential functions: two-pass variance sum (x - mean)^2 / (len - 1)
     * (mean_and_variance, mod.rs:375-399), covariance
     * mean(x y) - mean(x) mean(y) (mod.rs:427-458) */
    LQCD_STATS_SEQUENTIAL = 0,
    /* the *_par_iter functions: one-pass (sum x^2 - (sum x)^2 / len) / (len - 1)
     * (mod.rs:163-181), covariance (sum x y - sum x sum y / len) / len
     * (mod.rs:279-306) */
    LQCD_STATS_PARALLEL = 1
} lqcd_stats_form;

/* statistics::mean / mean_par_iter_val (mod.rs:321-335, 57-64): sum / len;
 * NaN for n = 0.  mean: 1 double, 2 when complex. */
int lqcd_stats_mean(lqcd_ctx* ctx, const double* d_values, size_t n, int complex_values, double* mean);
/* mean_and_variance (mod.rs:375-399) or mean_and_variance_par_iter_val
 * (mod.rs:163-181) by `form`: out = mean, variance (2 doubles, 4 when
 * complex).  n = 0 is LQCD_ERR_INVALID_ARGUMENT (the reference's len - 1
 * underflows); n = 1 gives a NaN variance, as in the reference. */
int lqcd_stats_mean_and_variance(lqcd_ctx* ctx, const double* d_values, size_t n, int complex_values, int form,
                                 double* out);
/* lqcd_stats_mean_and_variance with no device allocation: no counterpart in
 * the reference.  For callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc); the
 * plain entry point allocates its scratch on a context's first statistics
 * call.  Same arguments, results and errors.  The block partials and sums
 * live in one fixed device array per GPU (192 KiB + 128 B, held until the
 * process ends), whose address is resolved when a context is created on that
 * GPU.  Every value the kernels read from it is written earlier in the same
 * call, so nothing needs clearing; calls that overlap on one GPU (two host
 * threads, or two contexts' streams) race on it.  n above 2^62 is
 * LQCD_ERR_INVALID_ARGUMENT. */
int lqcd_stats_mean_and_variance_noalloc(lqcd_ctx* ctx, const double* d_values, size_t n, int complex_values,
                                         int form, double* out);
/* mean_with_error / mean_with_error_par_iter_val (mod.rs:401-405, 201-207):
 * out = mean, sqrt(variance / len), real values only. */
int lqcd_stats_mean_with_error(lqcd_ctx* ctx, const double* d_values, size_t n, int form, double out[2]);
/* covariance (mod.rs:427-458) or covariance_par_iter_val (mod.rs:279-306) by
 * `form`.  *has_value = 0 when nx != ny (the reference's None; out is left
 * alone), else 1 and out = the covariance (1 double, 2 when complex). */
int lqcd_stats_covariance(lqcd_ctx* ctx, const double* d_x, size_t nx, const double* d_y, size_t ny,
                          int complex_values, int form, double* out, int* has_value);

/* Copies `bytes` (a multiple of 16) between two device buffers `iters` times
 * with a streaming kernel and returns the read+write rate in GB/s: the
 * achievable HBM bandwidth of the device, and a reference of known size for
 * byte counters.  No counterpart in the reference. */
int lqcd_stream_copy(lqcd_ctx* ctx, size_t bytes, int iters, double* gb_per_s);
/* Reads `bytes` (a multiple of 16) of one device buffer `iters` times with a
 * streaming kernel and returns the read rate in GB/s: the reference point for
 * kernels that only read (a read-only stream can pass the copy rate without
 * any cache serving it).  No counterpart in the reference. */
int lqcd_stream_read(lqcd_ctx* ctx, size_t bytes, int iters, double* gb_per_s);

/* Record HIP events around every launch of the kinds in `mask` (bit = kind)
 * on the context's stream; mask 0 stops recording.  Clears past records. */
int lqcd_profile_enable(lqcd_ctx* ctx, unsigned mask);
/* Synchronizes, then returns the number of recorded launches of `kind` and
 * their summed duration in milliseconds. */
int lqcd_profile_read(lqcd_ctx* ctx, int kind, size_t* count, double* total_ms);

#ifdef __cplusplus
}
#endif
#endif /* LQCD_H */
