// internal.hpp -- host-side objects behind the C ABI of include/lqcd.h and the
// kernel launchers shared between translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "geometry.hpp"

struct lqcd_prof_event {
    int kind;
    hipEvent_t start, stop;
    int launches = 1;   // kernels between start and stop (a graph replay holds several)
};

struct lqcd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool profiling = false;
    std::vector<lqcd_prof_event> events;     // recorded launches of the profiled kinds
    std::vector<lqcd_prof_event> event_pool; // reusable event pairs
    unsigned profile_mask = 0;
    double* stats_scratch = nullptr;        // statistics partials and sums (device)
};

struct lqcd_lattice {
    lqcd_ctx* ctx = nullptr;
    lqcd::Geom g{};
    double size = 1.0;       // lattice spacing a
    int precision = 0;       // LQCD_F64 / LQCD_F32
    size_t nlink = 0;
    size_t elem = 8;         // bytes of one stored real
    void* U[3] = {nullptr, nullptr, nullptr};
    int cur = 0;             // buffer holding the current links
    void* E = nullptr;
    double* partials = nullptr;  // per-block reduction scratch
    double* result = nullptr;    // device reduction results
    double* host_result = nullptr; // pinned mirror of result
    double* mh_p = nullptr;        // per-link Metropolis probabilities (class order)
    uint32_t* mh_acc = nullptr;    // per-link Metropolis accept flags (class order)
    int nblocks_reduce = 0;
    int md_variant = -1;           // work decomposition of the MD kernels (-1: not tuned)
    int md_r12_mode = 1;           // 1: exp-drift integrations read SU(3) links as two rows
    bool md_r12 = false;           // set for the duration of one integration that does
    bool md_r12_last = false;      // whether the last integration did
    // leapfrog MD launches replayed from captured graphs, one per starting
    // buffer; the other fields are the launch parameters the graphs bake in
    bool md_graphs = true;
    hipGraphExec_t md_graph[3] = {nullptr, nullptr, nullptr};
    size_t mg_steps = 0;
    double mg_dt = 0.0;
    int mg_drift = -1, mg_variant = -1, mg_scheme = -1, mg_extras = -1;
    bool mg_r12 = false;
    // what is known of the distance of U[cur] from SU(3): >= 0 within
    // r12_distance, counting the SU(3)-preserving updates since that was
    // established; -1 unknown; -2 known to be off SU(3)
    int64_t su3_known = -1;
    void* gauss_g = nullptr;       // Gauss law G(x): 9 planes of double2
    double* gauss_e[3] = {nullptr, nullptr, nullptr};  // double copies of E for the projection
    double* md_partials = nullptr;  // the HMC's fused H sums: two sets of per-workgroup partials
};

namespace lqcd {

constexpr int kBlock = 256;
constexpr int kReduceSlots = 4;   // doubles per block partial
constexpr int kMaxReduceBlocks = 4096;
// device Metropolis proposals use Philox blocks 64 + 8 j + (0..6), below the
// heat-bath blocks (32768..)
constexpr size_t kMaxMhUpdates = 4000;
// Links whose distance from SU(3) (launch_su3_distance) is at most this are
// read as two rows by exp-drift integrations: 64 ulps of 1 in the storage
// type, i.e. the rounding of a stored SU(3) matrix with room for the growth
// of a trajectory.
constexpr double kR12DistanceF32 = 64.0 * 5.9604644775390625e-8;
constexpr double kR12DistanceF64 = 64.0 * 1.1102230246251565e-16;
inline double r12_distance(int precision) { return precision == 1 ? kR12DistanceF32 : kR12DistanceF64; }
// SU(3)-preserving updates (MD steps, Metropolis updates per link) after
// which the distance is measured again rather than assumed
constexpr int64_t kR12Recheck = 200;

enum ProfileKind {
    PROF_KICK_DRIFT = 0,
    PROF_KICK = 1,
    PROF_DRIFT = 2,
    PROF_PLAQUETTE = 3,
    PROF_EFIELD_ENERGY = 4,
    PROF_REUNIT = 5,
    PROF_METROPOLIS = 6,
    PROF_MOMENTA = 7,
    PROF_FORCE_STORE = 8,
    PROF_FIELDS = 9,
    PROF_STATS = 10,
    PROF_HEAT_BATH = 11,
    PROF_OVERRELAX = 12,
    PROF_NKINDS = 13
};

// ---- fields (k_fields.hip) ----
hipError_t launch_links_from_aos(const lqcd_lattice* L, const double* d_aos, void* U);
hipError_t launch_links_to_aos(const lqcd_lattice* L, const void* U, double* d_aos);
hipError_t launch_efield_from_aos(const lqcd_lattice* L, const double* d_aos, void* E);
hipError_t launch_efield_to_aos(const lqcd_lattice* L, const void* E, double* d_aos);
hipError_t launch_links_cold(const lqcd_lattice* L, void* U);
hipError_t launch_efield_zero(const lqcd_lattice* L, void* E);
hipError_t launch_device_momenta(const lqcd_lattice* L, void* E, uint64_t seed, uint64_t counter,
                                 double sigma);
hipError_t launch_reunitarize(const lqcd_lattice* L, void* U);
// row 2 of every link of U rebuilt from rows 0, 1 (links in SU(3))
hipError_t launch_complete_row2(const lqcd_lattice* L, void* U);
hipError_t launch_links_random(const lqcd_lattice* L, void* U, uint64_t seed);
hipError_t launch_stream_copy(hipStream_t stream, const void* src, void* dst, size_t bytes);
// read-only streaming probe: grid of kStreamReadBlocks x kBlock threads, sink
// holds one uint32 per thread
constexpr int kStreamReadBlocks = 8192;
hipError_t launch_stream_read(hipStream_t stream, const void* src, size_t bytes, uint32_t* sink);

// ---- observables (k_observables.hip) ----
// out[0..1] = sum Tr P (complex), out[2] = sum (1 - Re Tr P / 3)
hipError_t launch_plaquette_sum(const lqcd_lattice* L, const void* U, double* d_out3);
// the plaquette sums (d_out4[0..2], as launch_plaquette_sum) and the E energy
// sum (d_out4[3]) in one kernel; recorded as a plaquette launch
hipError_t launch_hamiltonian_sum(const lqcd_lattice* L, const void* U, const void* E, double* d_out4);
// out[0] = sum_a e_a^2 / 2 over all links
hipError_t launch_efield_energy(const lqcd_lattice* L, const void* E, double* d_out1);
// out[0] = sum p, out[1] = number accepted
hipError_t launch_mh_stats(const lqcd_lattice* L, double* d_out2);

// ---- molecular dynamics (k_md.hip) ----
enum DriftMode { DRIFT_ADDITIVE = 0, DRIFT_EXP = 1 };
enum MdVariant {
    VAR_LINK = 0,
    VAR_LINK_XCD = 1,
    VAR_SITE_XCD = 2,
    VAR_SITE_SLAB = 3,
    VAR_SITE_TILE = 4,   // falls back to VAR_SITE_SLAB where md_tile_fits is false
    VAR_COUNT = 5
};
// VAR_SITE_TILE's LDS tile: whether it applies to the lattice, and its bytes
bool md_tile_fits(const lqcd_lattice* L);
size_t md_tile_bytes(const lqcd_lattice* L);
// What an MD launch of the per-site kernels (variants 2-4) does besides its
// kick (+ drift): sums (MdSums in md_device.hpp: 0 none, 1 before the kick,
// 2 after it) of Re Tr(U Sigma) and e^2 into partials[block][0..1] (a
// workgroup per 256 sites), and with two-row reads row 2 of the links read
// written to row2 (null: not).
struct MdExtra {
    int sums = 0;
    double* partials = nullptr;
    void* row2 = nullptr;
};
// whether an MD work decomposition takes MdExtra (the per-site kernels)
bool md_variant_has_sums(int variant);
hipError_t launch_kick_variant(const lqcd_lattice* L, int variant, const void* U, void* E, double kick,
                               const MdExtra* x = nullptr);
hipError_t launch_kick_drift_variant(const lqcd_lattice* L, int variant, const void* U_in, void* U_out,
                                     void* E, double kick, double dt, int drift_mode, const MdExtra* x = nullptr);
// out[(s*D+mu)*8+a] = scale * derivative_e^a
hipError_t launch_force_store(const lqcd_lattice* L, const void* U, double* d_out, double scale);
// derivative_u of every link of U[cur] with E, host layout (18 doubles per link)
hipError_t launch_derivative_u(const lqcd_lattice* L, double* d_out);
// E += kick * derivative_e(U)
hipError_t launch_kick(const lqcd_lattice* L, const void* U, void* E, double kick, const MdExtra* x = nullptr);
// E += kick * derivative_e(U_in); U_out = drift(U_in, E, dt)
hipError_t launch_kick_drift(const lqcd_lattice* L, const void* U_in, void* U_out, void* E, double kick,
                             double dt, int drift_mode, const MdExtra* x = nullptr);
// U_out = drift(U_in, E, dt)
hipError_t launch_drift(const lqcd_lattice* L, const void* U_in, void* U_out, const void* E, double dt,
                        int drift_mode);

// ---- Metropolis (k_mc.hip) ----
hipError_t launch_metropolis_class(const lqcd_lattice* L, void* U, int mu, int parity, double beta,
                                   size_t n_update, const double* d_proposals,
                                   const uint64_t* d_uniforms);
// the same class update with proposals and accept words drawn on the device
hipError_t launch_metropolis_class_rng(const lqcd_lattice* L, void* U, int mu, int parity, double beta,
                                       size_t n_update, double spread, uint64_t seed, uint64_t counter);

// heat bath / overrelaxation of one (mu, parity) class (k_mc.hip); the heat
// bath writes the per-link count of exhausted samplers to mh_acc (mh_p = 0)
hipError_t launch_heat_bath_class(const lqcd_lattice* L, void* U, int mu, int parity, double beta, uint64_t seed,
                                  uint64_t counter);
// (reverse: OverrelaxationSweepReverse, else OverrelaxationSweepRotation)
hipError_t launch_overrelax_class(const lqcd_lattice* L, void* U, int mu, int parity, bool reverse);
// one MetropolisHastingsDeltaDiagnostic step on the device stream (seed,
// counter); d_out3 = probability, accepted, link index
hipError_t launch_metropolis_link(const lqcd_lattice* L, void* U, double beta, double spread, uint64_t seed,
                                  uint64_t counter, double* d_out3);

// clover (mode 0), f_mu_nu (1) or magnetic_field_vec (2) of U[cur] at every
// site, written to d_out in the host layout (k_clover.hip)
hipError_t launch_clover_fields(const lqcd_lattice* L, int mode, double* d_out);
// sij (which 0) or pij (which 1) of U[cur] at every site for the signed
// directions (a, pa), (b, pb), host layout (k_clover.hip)
hipError_t launch_plaquette_field(const lqcd_lattice* L, int which, int a, bool pa, int b, bool pb, double* d_out);

// ---- Gauss law (k_gauss.hip); links are U[cur], G goes to gauss_g ----
// G(x) of the double E copy Ed, and div = sum_x |Tr(sum_a T^a G(x))| into d_div
hipError_t launch_gauss(const lqcd_lattice* L, const double* Ed, double* d_div);
hipError_t launch_gauss_to_aos(const lqcd_lattice* L, double* d_out);
// one project_to_gauss_step Ed_in -> Ed_out with G computed in the kernel
// (gauss_g untouched), and div of Ed_in into d_div
hipError_t launch_gauss_fused(const lqcd_lattice* L, const double* Ed_in, double* Ed_out, double* d_div);
hipError_t launch_efield_to_double(const lqcd_lattice* L, const void* E, double* Ed);
hipError_t launch_efield_from_double(const lqcd_lattice* L, const double* Ed, void* E);
// sum of nblocks partials (slot 0) into d_out1
hipError_t launch_final_reduce1(const lqcd_lattice* L, int nblocks, double* d_out1);
// slots 0, 1 of nblocks partials at `partials`, each summed in block order, into d_out2[0..1]
hipError_t launch_final_reduce_at(const lqcd_lattice* L, const double* partials, int nblocks, double* d_out2);
// max over links of max(|U U^+ - 1|_ij, |det U - 1|) of buffer U into d_out1
hipError_t launch_su3_distance(const lqcd_lattice* L, const void* U, double* d_out1);

// ---- statistics (k_stats.hip) ----
// sums over n values (complex: (re, im) pairs) of x and x x (STATS_MOMENTS),
// of (x - m)(x - m) with m = (center[0] + i center[1]) / n (STATS_CENTERED),
// or of x, y and x y (STATS_COV), written to d_out in that order, one double
// (real) or a (re, im) pair (complex) each.  partials holds nblocks * 6
// doubles.
enum { STATS_MOMENTS = 0, STATS_CENTERED = 1, STATS_COV = 2 };
int stats_blocks(int64_t n);
hipError_t launch_stats(lqcd_ctx* ctx, int op, bool cplx, const double* x, const double* y, int64_t n,
                        const double* center, double* partials, int nblocks, double* d_out);

// resolves `device`'s address of the fixed statistics scratch (capi_stats.cpp);
// the device must be current
hipError_t stats_fixed_scratch_init(int device);

// profiling hooks (capi.cpp)
void prof_begin(lqcd_ctx* ctx, int kind, lqcd_prof_event* ev);
void prof_end(lqcd_ctx* ctx, lqcd_prof_event* ev);

} // namespace lqcd
