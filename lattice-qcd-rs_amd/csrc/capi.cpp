// capi.cpp -- the extern "C" boundary of include/lqcd.h.
//
// Owns device memory and the per-context HIP stream, validates arguments with
// the reference's error semantics, and sequences the kernels of one
// integration step / trajectory / sweep.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>

#include "../../include/lqcd.h"
#include "capi_util.hpp"
#include "dispatch.hpp"
#include "internal.hpp"

using namespace lqcd;

std::string& lqcd::last_error_slot()
{
    thread_local std::string msg;
    return msg;
}

namespace {

int other_buffers(int cur, int* a, int* b)
{
    *a = (cur + 1) % 3;
    *b = (cur + 2) % 3;
    return 0;
}

int read_results(lqcd_lattice* L, int n)
{
    HIP_TRY(hipMemcpyAsync(L->host_result, L->result, sizeof(double) * n, hipMemcpyDeviceToHost,
                           L->ctx->stream));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return LQCD_OK;
}

// H_U (beta * sum (1 - Re Tr P / 3)) of buffer U into host_result[2] and the
// plaquette trace sum into host_result[0..1]; E energy into host_result[3].
int hamiltonian_terms(lqcd_lattice* L, const void* U, bool with_e)
{
    if (with_e)
        HIP_TRY(launch_hamiltonian_sum(L, U, L->E, L->result));
    else
        HIP_TRY(launch_plaquette_sum(L, U, L->result));
    return read_results(L, with_e ? 4 : 3);
}

// Times every MD work decomposition on this lattice (fused kick+drift, kick
// 0 so E keeps its values; E is restored anyway) and keeps the fastest.
int autotune_md(lqcd_lattice* L, int drift, int* chosen)
{
    int w1, w2;
    other_buffers(L->cur, &w1, &w2);
    const size_t ebytes = (size_t)L->g.dim * 8 * (size_t)L->g.npad * L->elem;
    DevBuf ebak;
    HIP_TRY(hipMalloc(&ebak.p, ebytes));
    HIP_TRY(hipMemcpyAsync(ebak.p, L->E, ebytes, hipMemcpyDeviceToDevice, L->ctx->stream));
    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    int best = VAR_LINK_XCD;
    float best_ms = 1e30f;
    int rc = LQCD_OK;
    // two rounds over the variants, 6 launches each, keeping each variant's
    // better round: one noisy round does not decide
    float vms[VAR_COUNT];
    for (int v = 0; v < VAR_COUNT; ++v) vms[v] = 1e30f;
    for (int round = 0; round < 2 && rc == LQCD_OK; ++round)
        for (int v = 0; v < VAR_COUNT && rc == LQCD_OK; ++v) {
            const int reps = 6;
            hipError_t e = launch_kick_drift_variant(L, v, L->U[L->cur], L->U[w1], L->E, 0.0, 1e-3, drift);
            if (e == hipSuccess) e = hipEventRecord(t0, L->ctx->stream);
            for (int r = 0; r < reps && e == hipSuccess; ++r)
                e = launch_kick_drift_variant(L, v, L->U[L->cur], L->U[w1], L->E, 0.0, 1e-3, drift);
            if (e == hipSuccess) e = hipEventRecord(t1, L->ctx->stream);
            if (e == hipSuccess) e = hipEventSynchronize(t1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
            if (e != hipSuccess) {
                rc = fail(LQCD_ERR_DEVICE, std::string("autotune: ") + hipGetErrorString(e));
                break;
            }
            if (ms < vms[v]) vms[v] = ms;
        }
    for (int v = 0; v < VAR_COUNT; ++v)
        if (vms[v] < best_ms) {
            best_ms = vms[v];
            best = v;
        }
    // within 1.5 % of the fastest, the decomposition that moves the fewest
    // bytes past the L2 wins (slab order, then the LDS tile): the timing
    // rounds do not separate them reliably, and the slab order fetches about
    // half of what the plain XCD order does (32^4: 0.61 against 1.06 GB per
    // f32 launch; 8 % faster in f64)
    for (int v : {(int)VAR_SITE_SLAB, (int)VAR_SITE_TILE})
        if (vms[v] <= best_ms * 1.015f) {
            best = v;
            break;
        }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (rc != LQCD_OK) return rc;
    HIP_TRY(hipMemcpyAsync(L->E, ebak.p, ebytes, hipMemcpyDeviceToDevice, L->ctx->stream));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    L->md_variant = best;
    if (chosen) *chosen = best;
    return LQCD_OK;
}

// The MD work decomposition: LQCD_MD_VARIANT pins it, otherwise the first
// integration of a lattice autotunes it.
int ensure_md_variant(lqcd_lattice* L, int drift)
{
    if (L->md_variant >= 0) return LQCD_OK;
    const char* env = std::getenv("LQCD_MD_VARIANT");
    if (env && *env) {
        const int v = std::atoi(env);
        if (v >= 0 && v < VAR_COUNT) {
            L->md_variant = v;
            return LQCD_OK;
        }
    }
    return autotune_md(L, drift, nullptr);
}

struct R12Scope {
    lqcd_lattice* L;
    ~R12Scope() { L->md_r12 = false; }
};

// Whether U[cur] is within r12_distance of SU(3) for an operation of
// `updates` SU(3)-preserving updates per link.  A measurement
// (launch_su3_distance, one host sync) is made when nothing is known or when
// the updates since the last one would pass kR12Recheck; links known to be
// off SU(3) are not measured again until they are replaced.
int links_su3(lqcd_lattice* L, int64_t updates, bool* ok)
{
    *ok = false;
    if (L->su3_known == -2) return LQCD_OK;
    if (L->su3_known >= 0 && L->su3_known + updates <= kR12Recheck) {
        L->su3_known += updates;
        *ok = true;
        return LQCD_OK;
    }
    HIP_TRY(launch_su3_distance(L, L->U[L->cur], L->result + 8));
    HIP_TRY(hipMemcpyAsync(L->host_result + 8, L->result + 8, sizeof(double), hipMemcpyDeviceToHost,
                           L->ctx->stream));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    *ok = L->host_result[8] <= r12_distance(L->precision);
    L->su3_known = *ok ? updates : -2;
    return LQCD_OK;
}

// The MD launches of an integration from U[cur], buffers alternating w1, w2,
// replayed from a graph captured on first use.  Graph replays start each
// kernel as the last one ends; eager launches leave about 10 us between them.
//   LEAPFROG: the n fused kick+drift launches (kick dt/2 then dt, drift dt);
//             the final half kick stays outside.  With profiling on, one
//             event pair spans the replay and counts n fused launches.
//   REFERENCE: per step a fused kick(dt/2)+drift(dt) and a kick(dt/2); used
//             only with profiling off (the replay mixes two kernel kinds).
// `first` goes with the first fused launch and, for REFERENCE, `last` with
// the final kick (MdExtra; null: none).  One graph per starting buffer; all
// are captured again when a parameter they bake in changes.
static int md_graph_replay(lqcd_lattice* L, double dt, size_t n, int drift, int scheme, const MdExtra* first,
                           const MdExtra* last)
{
    const int extras = (first ? 1 : 0) | (last && last->row2 ? 2 : 0) | (last && last->sums ? 4 : 0);
    if (L->mg_steps != n || L->mg_dt != dt || L->mg_drift != drift || L->mg_variant != L->md_variant ||
        L->mg_r12 != L->md_r12 || L->mg_scheme != scheme || L->mg_extras != extras) {
        for (auto& ge : L->md_graph)
            if (ge) {
                (void)hipGraphExecDestroy(ge);
                ge = nullptr;
            }
        L->mg_steps = n;
        L->mg_dt = dt;
        L->mg_drift = drift;
        L->mg_variant = L->md_variant;
        L->mg_r12 = L->md_r12;
        L->mg_scheme = scheme;
        L->mg_extras = extras;
    }
    hipGraphExec_t& exec = L->md_graph[L->cur];
    hipStream_t st = L->ctx->stream;
    if (!exec) {
        int w1, w2;
        other_buffers(L->cur, &w1, &w2);
        // timing events stay out of the graph: it holds kernel nodes only
        const bool profiling = L->ctx->profiling;
        L->ctx->profiling = false;
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t le = hipSuccess;
        int src = L->cur, dst = w1;
        for (size_t k = 0; k < n && le == hipSuccess; ++k) {
            const MdExtra* x = k == 0 ? first : nullptr;
            if (scheme == LQCD_SCHEME_REFERENCE) {
                le = launch_kick_drift(L, L->U[src], L->U[dst], L->E, dt / 2.0, dt, drift, x);
                if (le == hipSuccess) le = launch_kick(L, L->U[dst], L->E, dt / 2.0, k + 1 == n ? last : nullptr);
            } else {
                le = launch_kick_drift(L, L->U[src], L->U[dst], L->E, k == 0 ? dt / 2.0 : dt, dt, drift, x);
            }
            src = dst;
            dst = (dst == w1) ? w2 : w1;
        }
        hipGraph_t graph = nullptr;
        const hipError_t ce = hipStreamEndCapture(st, &graph);
        L->ctx->profiling = profiling;
        if (le != hipSuccess || ce != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipGetLastError();
            return fail(LQCD_ERR_DEVICE, std::string("MD capture: ") +
                                             hipGetErrorString(le != hipSuccess ? le : ce));
        }
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) {
            exec = nullptr;
            return fail(LQCD_ERR_DEVICE, std::string("MD graph instantiate: ") + hipGetErrorString(ie));
        }
    }
    lqcd_prof_event ev;
    prof_begin(L->ctx, PROF_KICK_DRIFT, &ev);
    ev.launches = (int)n;
    HIP_TRY(hipGraphLaunch(exec, st));
    prof_end(L->ctx, &ev);
    return LQCD_OK;
}

// the buffer holding the links after n drifts from U[cur]
static int buffer_after(const lqcd_lattice* L, size_t n)
{
    int w1, w2;
    other_buffers(L->cur, &w1, &w2);
    return n == 0 ? L->cur : ((n % 2) ? w1 : w2);
}

// n steps of the chosen scheme starting from U[cur]; *last gets the buffer
// holding the final links.  U[cur] is left untouched.  An integration with
// the exp drift keeps the links in SU(3) (up to rounding), so when the
// starting links are in SU(3) its kernels read only their first two rows and
// store only the first two rows of the links they drift; the final kick of
// the per-site kernels writes row 2 of the final links back (otherwise
// k_complete_row2 does).  sums (non-null): the terms of H before and after
// the integration go to result.  With a per-site work decomposition the
// first launch and the final kick sum them on the way (MdExtra): result[0..3]
// = sum Re Tr(U Sigma) and sum e^2 over all links, before, then after, and
// *sums = true.  Otherwise the plaquette kernel runs before and after the MD
// (launch_hamiltonian_sum into result[0..3] and result[4..7]), *sums = false.
int integrate_steps(lqcd_lattice* L, double dt, size_t n, int drift, int scheme, int* last, bool* sums = nullptr)
{
    R12Scope scope{L};
    L->md_r12 = false;
    if (drift == LQCD_DRIFT_EXP && L->md_r12_mode) {
        bool ok;
        if (int rc = links_su3(L, (int64_t)n, &ok)) return rc;
        L->md_r12 = ok;
    }
    // the additive drift takes the links off SU(3)
    if (drift == LQCD_DRIFT_ADDITIVE) L->su3_known = -2;
    L->md_r12_last = L->md_r12;
    if (int rc = ensure_md_variant(L, drift)) return rc;
    const bool site = md_variant_has_sums(L->md_variant);
    const bool with_sums = sums && site && n > 0;
    const int64_t nb = (int64_t)grid_for(L->g.nsite);
    if (with_sums) {
        if (!L->md_partials)
            HIP_TRY(hipMalloc((void**)&L->md_partials, sizeof(double) * 2 * (size_t)nb * kReduceSlots));
    } else if (sums) {
        HIP_TRY(launch_hamiltonian_sum(L, L->U[L->cur], L->E, L->result));
    }
    const int final_buf = buffer_after(L, n);
    MdExtra first, fin;
    if (with_sums) {
        first.sums = 1;
        first.partials = L->md_partials;
        fin.sums = 2;
        fin.partials = L->md_partials + nb * kReduceSlots;
    }
    // the final kick reads the final links as two rows and completes them
    // anyway: it writes their row 2
    if (site && L->md_r12 && n > 0) fin.row2 = L->U[final_buf];
    const MdExtra* px = with_sums ? &first : nullptr;
    const MdExtra* pf = (with_sums || fin.row2) ? &fin : nullptr;
    int w1, w2;
    other_buffers(L->cur, &w1, &w2);
    int src = L->cur, dst = w1;
    if (scheme == LQCD_SCHEME_REFERENCE) {
        if (L->md_graphs && !L->ctx->profiling) {
            if (int rc = md_graph_replay(L, dt, n, drift, scheme, px, pf)) return rc;
            src = final_buf;
        } else {
            for (size_t k = 0; k < n; ++k) {
                HIP_TRY(launch_kick_drift(L, L->U[src], L->U[dst], L->E, dt / 2.0, dt, drift, k == 0 ? px : nullptr));
                HIP_TRY(launch_kick(L, L->U[dst], L->E, dt / 2.0, k + 1 == n ? pf : nullptr));
                src = dst;
                dst = (dst == w1) ? w2 : w1;
            }
        }
    } else {
        if (L->md_graphs) {
            if (int rc = md_graph_replay(L, dt, n, drift, scheme, px, nullptr)) return rc;
            src = final_buf;
        } else {
            for (size_t k = 0; k < n; ++k) {
                HIP_TRY(launch_kick_drift(L, L->U[src], L->U[dst], L->E, k == 0 ? dt / 2.0 : dt, dt, drift,
                                          k == 0 ? px : nullptr));
                src = dst;
                dst = (dst == w1) ? w2 : w1;
            }
        }
        HIP_TRY(launch_kick(L, L->U[src], L->E, dt / 2.0, pf));
    }
    // two-row kernels store the drifted links as rows 0, 1: the final links
    // get their row 2 back before anything else reads them
    if (L->md_r12 && n > 0 && !fin.row2) HIP_TRY(launch_complete_row2(L, L->U[src]));
    if (with_sums) {
        HIP_TRY(launch_final_reduce_at(L, first.partials, (int)nb, L->result));
        HIP_TRY(launch_final_reduce_at(L, fin.partials, (int)nb, L->result + 2));
    } else if (sums) {
        HIP_TRY(launch_hamiltonian_sum(L, L->U[src], L->E, L->result + 4));
    }
    if (sums) *sums = with_sums;
    *last = src;
    return LQCD_OK;
}

int check_links(const lqcd_lattice* L, size_t n_links)
{
    if (n_links != L->nlink)
        return fail(LQCD_ERR_INCOMPATIBLE_SIZE, "size of lattice and data are incompatible: expected " +
                                                    std::to_string(L->nlink) + " links, got " +
                                                    std::to_string(n_links));
    return LQCD_OK;
}

} // namespace

namespace lqcd {

void prof_begin(lqcd_ctx* ctx, int kind, lqcd_prof_event* ev)
{
    ev->kind = -1;
    ev->launches = 1;
    if (!ctx->profiling || !(ctx->profile_mask & (1u << kind))) return;
    lqcd_prof_event e;
    if (!ctx->event_pool.empty()) {
        e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
    } else {
        if (hipEventCreate(&e.start) != hipSuccess) return;
        if (hipEventCreate(&e.stop) != hipSuccess) {
            (void)hipEventDestroy(e.start);
            return;
        }
    }
    e.kind = kind;
    // a pooled pair keeps the count of its last use (a graph replay holds n
    // launches); every new record starts at one launch
    e.launches = 1;
    (void)hipEventRecord(e.start, ctx->stream);
    *ev = e;
}

void prof_end(lqcd_ctx* ctx, lqcd_prof_event* ev)
{
    if (ev->kind < 0) return;
    (void)hipEventRecord(ev->stop, ctx->stream);
    ctx->events.push_back(*ev);
}

} // namespace lqcd

extern "C" {

const char* lqcd_last_error(void) { return last_error_slot().c_str(); }

const char* lqcd_version(void) { return "lqcd-mi355x 0.1.0 (gfx950)"; }

int lqcd_device_count(int* count)
{
    CHECK_ARG(count, "count is null");
    HIP_TRY(hipGetDeviceCount(count));
    return LQCD_OK;
}

int lqcd_ctx_create(int device, lqcd_ctx** out)
{
    CHECK_ARG(out, "out is null");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    CHECK_ARG(device >= 0 && device < n, "device index out of range");
    HIP_TRY(hipSetDevice(device));
    lqcd_ctx* ctx = new lqcd_ctx;
    ctx->device = device;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return fail(LQCD_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    // the allocation-free statistics entry point retries a failure here
    (void)stats_fixed_scratch_init(device);
    *out = ctx;
    return LQCD_OK;
}

int lqcd_ctx_destroy(lqcd_ctx* ctx)
{
    if (!ctx) return LQCD_OK;
    DeviceGuard dg(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& e : ctx->events) {
        (void)hipEventDestroy(e.start);
        (void)hipEventDestroy(e.stop);
    }
    for (auto& e : ctx->event_pool) {
        (void)hipEventDestroy(e.start);
        (void)hipEventDestroy(e.stop);
    }
    if (ctx->stats_scratch) (void)hipFree(ctx->stats_scratch);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return LQCD_OK;
}

int lqcd_ctx_synchronize(lqcd_ctx* ctx)
{
    CHECK_ARG(ctx, "ctx is null");
    DeviceGuard dg(ctx->device);
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return LQCD_OK;
}

int lqcd_lattice_create(lqcd_ctx* ctx, int dim, const size_t* extent, double size, int precision,
                        lqcd_lattice** out)
{
    CHECK_ARG(ctx && out, "null argument");
    // LatticeCyclic::new, src/lattice.rs:190-201
    if (dim == 0) return fail(LQCD_ERR_ZERO_DIMENSION, "the lattice has zero dimension");
    if (size <= 0.0 || std::isnan(size) || std::isinf(size))
        return fail(LQCD_ERR_NON_POSITIVE_SIZE, "the lattice spacing must be positive and finite");
    CHECK_ARG(extent, "extent is null");
    if (dim < 2 || dim > kMaxDim)
        return fail(LQCD_ERR_UNSUPPORTED, "dimension must be 2, 3 or 4");
    for (int i = 0; i < dim; ++i)
        if (extent[i] < 2) return fail(LQCD_ERR_DIM_TOO_SMALL, "every extent must be at least 2");
    CHECK_ARG(precision == LQCD_F64 || precision == LQCD_F32, "unknown precision");
    DeviceGuard dg(ctx->device);

    lqcd_lattice* L = new lqcd_lattice;
    L->ctx = ctx;
    L->size = size;
    L->precision = precision;
    L->elem = precision == LQCD_F32 ? 4 : 8;
    L->g.dim = dim;
    int64_t n = 1;
    for (int i = 0; i < kMaxDim; ++i) {
        L->g.extent[i] = 1;
        L->g.stride[i] = 0;
    }
    for (int i = 0; i < dim; ++i) {
        if (extent[i] > (size_t)(1 << 30)) {
            delete L;
            return fail(LQCD_ERR_INVALID_ARGUMENT, "extent too large");
        }
        L->g.extent[i] = (int)extent[i];
        L->g.stride[i] = n;
        n *= (int64_t)extent[i];
        if (n > ((int64_t)1 << 40)) {
            delete L;
            return fail(LQCD_ERR_INVALID_ARGUMENT, "lattice too large");
        }
    }
    L->g.nsite = n;
    L->g.npad = (n + 255) / 256 * 256;
    // the device addresses the 9 planes of one direction with 32-bit offsets
    if ((uint64_t)L->g.npad * 9 * 2 * L->elem > kMaxPlaneBytes) {
        delete L;
        return fail(LQCD_ERR_UNSUPPORTED, "lattice too large: the link planes of one direction exceed 4 GiB");
    }
    L->nlink = (size_t)n * (size_t)dim;
    const size_t ubytes = (size_t)dim * 9 * (size_t)L->g.npad * 2 * L->elem;
    const size_t ebytes = (size_t)dim * 8 * (size_t)L->g.npad * L->elem;
    hipError_t e = hipSuccess;
    for (int b = 0; b < 3 && e == hipSuccess; ++b) e = hipMalloc(&L->U[b], ubytes);
    if (e == hipSuccess) e = hipMalloc(&L->E, ebytes);
    if (e == hipSuccess)
        e = hipMalloc((void**)&L->partials, sizeof(double) * kMaxReduceBlocks * kReduceSlots);
    if (e == hipSuccess) e = hipMalloc((void**)&L->result, sizeof(double) * 16);
    if (e == hipSuccess) e = hipHostMalloc((void**)&L->host_result, sizeof(double) * 16, 0);
    if (e == hipSuccess) e = hipMemsetAsync(L->U[0], 0, ubytes, ctx->stream);
    if (e == hipSuccess) e = launch_links_cold(L, L->U[0]);
    if (e == hipSuccess) e = launch_efield_zero(L, L->E);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        lqcd_lattice_destroy(L);
        return fail(e == hipErrorOutOfMemory ? LQCD_ERR_OUT_OF_MEMORY : LQCD_ERR_DEVICE,
                    std::string("lattice allocation: ") + hipGetErrorString(e));
    }
    *out = L;
    return LQCD_OK;
}

int lqcd_lattice_destroy(lqcd_lattice* L)
{
    if (!L) return LQCD_OK;
    DeviceGuard dg(L->ctx->device);
    (void)hipStreamSynchronize(L->ctx->stream);
    for (int b = 0; b < 3; ++b)
        if (L->U[b]) (void)hipFree(L->U[b]);
    if (L->E) (void)hipFree(L->E);
    if (L->partials) (void)hipFree(L->partials);
    if (L->result) (void)hipFree(L->result);
    if (L->host_result) (void)hipHostFree(L->host_result);
    if (L->mh_p) (void)hipFree(L->mh_p);
    if (L->mh_acc) (void)hipFree(L->mh_acc);
    if (L->gauss_g) (void)hipFree(L->gauss_g);
    if (L->md_partials) (void)hipFree(L->md_partials);
    for (int k = 0; k < 3; ++k)
        if (L->gauss_e[k]) (void)hipFree(L->gauss_e[k]);
    for (auto& ge : L->md_graph)
        if (ge) (void)hipGraphExecDestroy(ge);
    delete L;
    return LQCD_OK;
}

int lqcd_lattice_info(const lqcd_lattice* L, size_t* nsite, size_t* nlink)
{
    CHECK_ARG(L, "lattice is null");
    if (nsite) *nsite = (size_t)L->g.nsite;
    if (nlink) *nlink = L->nlink;
    return LQCD_OK;
}

int lqcd_links_upload(lqcd_lattice* L, const double* links, size_t n_links)
{
    CHECK_ARG(L && links, "null argument");
    if (int rc = check_links(L, n_links)) return rc;
    DeviceGuard dg(L->ctx->device);
    DevBuf st;
    HIP_TRY(hipMalloc(&st.p, sizeof(double) * 18 * L->nlink));
    HIP_TRY(hipMemcpyAsync(st.p, links, sizeof(double) * 18 * L->nlink, hipMemcpyHostToDevice,
                           L->ctx->stream));
    HIP_TRY(launch_links_from_aos(L, (const double*)st.p, L->U[L->cur]));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    L->su3_known = -1;
    return LQCD_OK;
}

int lqcd_links_download(lqcd_lattice* L, double* links, size_t n_links)
{
    CHECK_ARG(L && links, "null argument");
    if (int rc = check_links(L, n_links)) return rc;
    DeviceGuard dg(L->ctx->device);
    DevBuf st;
    HIP_TRY(hipMalloc(&st.p, sizeof(double) * 18 * L->nlink));
    HIP_TRY(launch_links_to_aos(L, L->U[L->cur], (double*)st.p));
    HIP_TRY(hipMemcpyAsync(links, st.p, sizeof(double) * 18 * L->nlink, hipMemcpyDeviceToHost,
                           L->ctx->stream));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return LQCD_OK;
}

int lqcd_links_cold(lqcd_lattice* L)
{
    CHECK_ARG(L, "lattice is null");
    DeviceGuard dg(L->ctx->device);
    HIP_TRY(launch_links_cold(L, L->U[L->cur]));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    L->su3_known = 0;
    return LQCD_OK;
}

int lqcd_links_random(lqcd_lattice* L, uint64_t seed)
{
    CHECK_ARG(L, "lattice is null");
    DeviceGuard dg(L->ctx->device);
    HIP_TRY(launch_links_random(L, L->U[L->cur], seed));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    L->su3_known = 0;
    return LQCD_OK;
}

int lqcd_efield_upload(lqcd_lattice* L, const double* efield, size_t n_links)
{
    CHECK_ARG(L && efield, "null argument");
    if (int rc = check_links(L, n_links)) return rc;
    DeviceGuard dg(L->ctx->device);
    DevBuf st;
    HIP_TRY(hipMalloc(&st.p, sizeof(double) * 8 * L->nlink));
    HIP_TRY(hipMemcpyAsync(st.p, efield, sizeof(double) * 8 * L->nlink, hipMemcpyHostToDevice,
                           L->ctx->stream));
    HIP_TRY(launch_efield_from_aos(L, (const double*)st.p, L->E));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return LQCD_OK;
}

int lqcd_efield_download(lqcd_lattice* L, double* efield, size_t n_links)
{
    CHECK_ARG(L && efield, "null argument");
    if (int rc = check_links(L, n_links)) return rc;
    DeviceGuard dg(L->ctx->device);
    DevBuf st;
    HIP_TRY(hipMalloc(&st.p, sizeof(double) * 8 * L->nlink));
    HIP_TRY(launch_efield_to_aos(L, L->E, (double*)st.p));
    HIP_TRY(hipMemcpyAsync(efield, st.p, sizeof(double) * 8 * L->nlink, hipMemcpyDeviceToHost,
                           L->ctx->stream));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return LQCD_OK;
}

int lqcd_efield_cold(lqcd_lattice* L)
{
    CHECK_ARG(L, "lattice is null");
    DeviceGuard dg(L->ctx->device);
    HIP_TRY(launch_efield_zero(L, L->E));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return LQCD_OK;
}

int lqcd_efield_random(lqcd_lattice* L, uint64_t seed, uint64_t counter, double sigma)
{
    CHECK_ARG(L, "lattice is null");
    CHECK_ARG(sigma >= 0.0 && std::isfinite(sigma), "sigma must be finite and >= 0");
    DeviceGuard dg(L->ctx->device);
    // enqueued only: later calls on the context are stream-ordered after it
    HIP_TRY(launch_device_momenta(L, L->E, seed, counter, sigma));
    return LQCD_OK;
}

static int ensure_gauss_buffers(lqcd_lattice* L)
{
    const size_t npad = (size_t)L->g.npad;
    if (!L->gauss_g) HIP_TRY(hipMalloc(&L->gauss_g, sizeof(double) * 2 * 9 * npad));
    for (int k = 0; k < 3; ++k)
        if (!L->gauss_e[k]) HIP_TRY(hipMalloc((void**)&L->gauss_e[k], sizeof(double) * 8 * npad * L->g.dim));
    return LQCD_OK;
}

int lqcd_efield_gauss(lqcd_lattice* L, double* out, size_t n_sites)
{
    CHECK_ARG(L && out, "null argument");
    if (n_sites != (size_t)L->g.nsite)
        return fail(LQCD_ERR_INCOMPATIBLE_SIZE, "size of lattice and data are incompatible: expected " +
                                                    std::to_string(L->g.nsite) + " sites, got " +
                                                    std::to_string(n_sites));
    DeviceGuard dg(L->ctx->device);
    if (int rc = ensure_gauss_buffers(L)) return rc;
    DevBuf st;
    HIP_TRY(hipMalloc(&st.p, sizeof(double) * 18 * n_sites));
    HIP_TRY(launch_efield_to_double(L, L->E, L->gauss_e[0]));
    HIP_TRY(launch_gauss(L, L->gauss_e[0], L->result));
    HIP_TRY(launch_gauss_to_aos(L, (double*)st.p));
    HIP_TRY(hipMemcpyAsync(out, st.p, sizeof(double) * 18 * n_sites, hipMemcpyDeviceToHost, L->ctx->stream));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return LQCD_OK;
}

static int clover_fields(lqcd_lattice* L, int mode, double* out, size_t n)
{
    CHECK_ARG(L && out, "null argument");
    const size_t D = (size_t)L->g.dim;
    const size_t want = (size_t)L->g.nsite * (mode == 2 ? D : D * D) * 18;
    if (n != want)
        return fail(LQCD_ERR_INCOMPATIBLE_SIZE, "size of lattice and data are incompatible: expected " +
                                                    std::to_string(want) + " doubles, got " + std::to_string(n));
    DeviceGuard dg(L->ctx->device);
    DevBuf buf;
    HIP_TRY(hipMalloc(&buf.p, sizeof(double) * want));
    HIP_TRY(launch_clover_fields(L, mode, (double*)buf.p));
    HIP_TRY(hipMemcpyAsync(out, buf.p, sizeof(double) * want, hipMemcpyDeviceToHost, L->ctx->stream));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return LQCD_OK;
}

int lqcd_clover_field(lqcd_lattice* L, double* out, size_t n) { return clover_fields(L, 0, out, n); }

int lqcd_plaquette_field(lqcd_lattice* L, int which, int dir_i, int positive_i, int dir_j, int positive_j,
                         double* out, size_t n)
{
    CHECK_ARG(L && out, "null argument");
    CHECK_ARG(which == 0 || which == 1, "which must be 0 (sij) or 1 (pij)");
    CHECK_ARG(dir_i >= 0 && dir_i < L->g.dim && dir_j >= 0 && dir_j < L->g.dim, "direction out of range");
    const size_t want = (size_t)L->g.nsite * 18;
    if (n != want)
        return fail(LQCD_ERR_INCOMPATIBLE_SIZE, "size of lattice and data are incompatible: expected " +
                                                    std::to_string(want) + " doubles, got " + std::to_string(n));
    DeviceGuard dg(L->ctx->device);
    DevBuf buf;
    HIP_TRY(hipMalloc(&buf.p, sizeof(double) * want));
    HIP_TRY(launch_plaquette_field(L, which, dir_i, positive_i != 0, dir_j, positive_j != 0, (double*)buf.p));
    HIP_TRY(hipMemcpyAsync(out, buf.p, sizeof(double) * want, hipMemcpyDeviceToHost, L->ctx->stream));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return LQCD_OK;
}

int lqcd_field_strength_device(lqcd_lattice* L, int what, double* d_out, size_t n)
{
    CHECK_ARG(L && d_out, "null argument");
    CHECK_ARG(what == LQCD_FIELD_CLOVER || what == LQCD_FIELD_F_MU_NU || what == LQCD_FIELD_MAGNETIC,
              "unknown field kind");
    const size_t D = (size_t)L->g.dim;
    const size_t want = (size_t)L->g.nsite * (what == LQCD_FIELD_MAGNETIC ? D : D * D) * 18;
    if (n != want)
        return fail(LQCD_ERR_INCOMPATIBLE_SIZE, "size of lattice and data are incompatible: expected " +
                                                    std::to_string(want) + " doubles, got " + std::to_string(n));
    DeviceGuard dg(L->ctx->device);
    HIP_TRY(launch_clover_fields(L, what, d_out));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return LQCD_OK;
}

int lqcd_f_mu_nu_field(lqcd_lattice* L, double* out, size_t n) { return clover_fields(L, 1, out, n); }

int lqcd_magnetic_field_vec(lqcd_lattice* L, double* out, size_t n) { return clover_fields(L, 2, out, n); }

int lqcd_efield_gauss_sum_div(lqcd_lattice* L, double* out)
{
    CHECK_ARG(L && out, "null argument");
    DeviceGuard dg(L->ctx->device);
    if (int rc = ensure_gauss_buffers(L)) return rc;
    HIP_TRY(launch_efield_to_double(L, L->E, L->gauss_e[0]));
    HIP_TRY(launch_gauss(L, L->gauss_e[0], L->result));
    if (int rc = read_results(L, 1)) return rc;
    *out = L->host_result[0];
    return LQCD_OK;
}

int lqcd_efield_project_gauss(lqcd_lattice* L, size_t max_iter, size_t* iterations)
{
    CHECK_ARG(L, "lattice is null");
    DeviceGuard dg(L->ctx->device);
    if (int rc = ensure_gauss_buffers(L)) return rc;
    // project_to_gauss, src/field.rs:1265-1294: one step, then while the
    // deviation is above N * 4 * 8 * 10 * eps, four more steps.  Each step is
    // one k_gauss_fused launch, which also sums the deviation of its input: a
    // check is the first launch of its group of four.  The host reads that
    // deviation while the other three steps run; they write the two buffers
    // other than the checked field, which stays the result if the check passes.
    double* const buf[3] = {L->gauss_e[0], L->gauss_e[1], L->gauss_e[2]};
    int cur = 1;
    HIP_TRY(launch_efield_to_double(L, L->E, buf[0]));
    HIP_TRY(launch_gauss_fused(L, buf[0], buf[1], L->result + 1));
    const double threshold = 2.220446049250313e-16 * (double)(L->g.nsite * 4 * 8 * 10);
    hipEvent_t checked;
    HIP_TRY(hipEventCreateWithFlags(&checked, hipEventDisableTiming));
    struct EventGuard {
        hipEvent_t e;
        ~EventGuard() { (void)hipEventDestroy(e); }
    } eg{checked};
    size_t it = 0;
    bool converged = false;
    double div = 0.0;
    for (; it < max_iter; ++it) {
        const int n1 = (cur + 1) % 3, n2 = (cur + 2) % 3;
        HIP_TRY(launch_gauss_fused(L, buf[cur], buf[n1], L->result));
        HIP_TRY(hipMemcpyAsync(L->host_result, L->result, sizeof(double), hipMemcpyDeviceToHost, L->ctx->stream));
        HIP_TRY(hipEventRecord(checked, L->ctx->stream));
        HIP_TRY(launch_gauss_fused(L, buf[n1], buf[n2], L->result + 1));
        HIP_TRY(launch_gauss_fused(L, buf[n2], buf[n1], L->result + 1));
        HIP_TRY(launch_gauss_fused(L, buf[n1], buf[n2], L->result + 1));
        HIP_TRY(hipEventSynchronize(checked));
        div = L->host_result[0];
        if (std::isnan(div)) {
            HIP_TRY(hipStreamSynchronize(L->ctx->stream));
            if (iterations) *iterations = it;
            return fail(LQCD_ERR_GAUSS_PROJECTION, "gauss projection could not finish: deviation is NaN");
        }
        if (div <= threshold) {
            converged = true;
            break;
        }
        cur = n2;
    }
    if (iterations) *iterations = it;
    if (!converged) {
        HIP_TRY(hipStreamSynchronize(L->ctx->stream));
        char msg[160];
        std::snprintf(msg, sizeof(msg),
                      "gauss projection could not finish within %zu iterations (deviation %.6e, threshold %.6e)",
                      max_iter, div, threshold);
        return fail(LQCD_ERR_GAUSS_PROJECTION, msg);
    }
    HIP_TRY(launch_efield_from_double(L, buf[cur], L->E));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return LQCD_OK;
}

int lqcd_average_trace_plaquette(lqcd_lattice* L, double out[2])
{
    CHECK_ARG(L && out, "null argument");
    DeviceGuard dg(L->ctx->device);
    if (int rc = hamiltonian_terms(L, L->U[L->cur], false)) return rc;
    const double nplaq = (double)L->g.nsite * (double)(L->g.dim * (L->g.dim - 1) / 2);
    out[0] = L->host_result[0] / nplaq;
    out[1] = L->host_result[1] / nplaq;
    return LQCD_OK;
}

int lqcd_hamiltonian_links(lqcd_lattice* L, double beta, double* out)
{
    CHECK_ARG(L && out, "null argument");
    DeviceGuard dg(L->ctx->device);
    if (int rc = hamiltonian_terms(L, L->U[L->cur], false)) return rc;
    *out = L->host_result[2] * beta;
    return LQCD_OK;
}

int lqcd_hamiltonian_efield(lqcd_lattice* L, double beta, double* out)
{
    CHECK_ARG(L && out, "null argument");
    DeviceGuard dg(L->ctx->device);
    HIP_TRY(launch_efield_energy(L, L->E, L->result + 3));
    HIP_TRY(hipMemcpyAsync(L->host_result + 3, L->result + 3, sizeof(double), hipMemcpyDeviceToHost,
                           L->ctx->stream));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    *out = L->host_result[3] * beta;
    return LQCD_OK;
}

static int force_to_host(lqcd_lattice* L, double* out, size_t n_links, double scale)
{
    CHECK_ARG(L && out, "null argument");
    if (int rc = check_links(L, n_links)) return rc;
    DeviceGuard dg(L->ctx->device);
    DevBuf st;
    HIP_TRY(hipMalloc(&st.p, sizeof(double) * 8 * L->nlink));
    HIP_TRY(launch_force_store(L, L->U[L->cur], (double*)st.p, scale));
    HIP_TRY(hipMemcpyAsync(out, st.p, sizeof(double) * 8 * L->nlink, hipMemcpyDeviceToHost,
                           L->ctx->stream));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return LQCD_OK;
}

int lqcd_derivative_e(lqcd_lattice* L, double* out, size_t n_links)
{
    return force_to_host(L, out, n_links, 1.0);
}

int lqcd_derivative_u(lqcd_lattice* L, double* out, size_t n_links)
{
    CHECK_ARG(L && out, "null argument");
    if (int rc = check_links(L, n_links)) return rc;
    DeviceGuard dg(L->ctx->device);
    DevBuf st;
    HIP_TRY(hipMalloc(&st.p, sizeof(double) * 18 * L->nlink));
    HIP_TRY(launch_derivative_u(L, (double*)st.p));
    HIP_TRY(hipMemcpyAsync(out, st.p, sizeof(double) * 18 * L->nlink, hipMemcpyDeviceToHost, L->ctx->stream));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return LQCD_OK;
}

int lqcd_integrate_efield(lqcd_lattice* L, double dt)
{
    CHECK_ARG(L, "lattice is null");
    CHECK_ARG(std::isfinite(dt), "dt must be finite");
    DeviceGuard dg(L->ctx->device);
    if (int rc = ensure_md_variant(L, LQCD_DRIFT_EXP)) return rc;
    HIP_TRY(launch_kick(L, L->U[L->cur], L->E, dt));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return LQCD_OK;
}

int lqcd_integrate_link(lqcd_lattice* L, double dt, int drift)
{
    CHECK_ARG(L, "lattice is null");
    CHECK_ARG(std::isfinite(dt), "dt must be finite");
    CHECK_ARG(drift == LQCD_DRIFT_ADDITIVE || drift == LQCD_DRIFT_EXP, "unknown drift mode");
    DeviceGuard dg(L->ctx->device);
    int w1, w2;
    other_buffers(L->cur, &w1, &w2);
    HIP_TRY(launch_drift(L, L->U[L->cur], L->U[w1], L->E, dt, drift));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    L->cur = w1;
    if (drift == LQCD_DRIFT_ADDITIVE)
        L->su3_known = -2;
    else if (L->su3_known >= 0)
        ++L->su3_known;
    return LQCD_OK;
}

// the links after an update of every link by the chosen drift
static void note_drift(lqcd_lattice* L, int drift)
{
    if (drift == LQCD_DRIFT_ADDITIVE)
        L->su3_known = -2;
    else if (L->su3_known >= 0)
        ++L->su3_known;
}

// n steps of integrate_sync_sync (symplectic_euler_rayon.rs:127-144): both
// halves from the old state, U' = drift(U, E), E' = E + dt F(U).  The drift
// writes a second buffer, so the kick still reads the old links.
static int sync_sync_steps(lqcd_lattice* L, double dt, size_t n, int drift)
{
    if (int rc = ensure_md_variant(L, drift)) return rc;
    for (size_t k = 0; k < n; ++k) {
        int w1, w2;
        other_buffers(L->cur, &w1, &w2);
        HIP_TRY(launch_drift(L, L->U[L->cur], L->U[w1], L->E, dt, drift));
        HIP_TRY(launch_kick(L, L->U[L->cur], L->E, dt));
        L->cur = w1;
        note_drift(L, drift);
    }
    return LQCD_OK;
}

// n steps of integrate_leap_leap (symplectic_euler_rayon.rs:146-171), drift
// then kick with the new links; the last step's kick is `last_kick` * dt
// (0.5 for integrate_leap_sync, :195-220).  A kick followed by the next
// step's drift is one fused launch.
static int leap_steps(lqcd_lattice* L, double dt, size_t n, int drift, double last_kick)
{
    if (int rc = ensure_md_variant(L, drift)) return rc;
    int w1, w2;
    other_buffers(L->cur, &w1, &w2);
    int src = L->cur, dst = w1;
    HIP_TRY(launch_drift(L, L->U[src], L->U[dst], L->E, dt, drift));
    for (size_t k = 1; k < n; ++k) {
        src = dst;
        dst = (dst == w1) ? w2 : w1;
        HIP_TRY(launch_kick_drift(L, L->U[src], L->U[dst], L->E, dt, dt, drift));
    }
    HIP_TRY(launch_kick(L, L->U[dst], L->E, last_kick * dt));
    L->cur = dst;
    for (size_t k = 0; k < n; ++k) note_drift(L, drift);
    return LQCD_OK;
}

static int check_integration(lqcd_lattice* L, double dt, int drift)
{
    CHECK_ARG(L, "lattice is null");
    CHECK_ARG(std::isfinite(dt), "dt must be finite");
    CHECK_ARG(drift == LQCD_DRIFT_ADDITIVE || drift == LQCD_DRIFT_EXP, "unknown drift mode");
    return LQCD_OK;
}

int lqcd_integrate_sync_sync(lqcd_lattice* L, double dt, int drift)
{
    if (int rc = check_integration(L, dt, drift)) return rc;
    DeviceGuard dg(L->ctx->device);
    if (int rc = sync_sync_steps(L, dt, 1, drift)) return rc;
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return LQCD_OK;
}

int lqcd_simulate_sync_n(lqcd_lattice* L, double dt, size_t n, int drift)
{
    if (int rc = check_integration(L, dt, drift)) return rc;
    if (n == 0) return fail(LQCD_ERR_ZERO_INTEGRATION, "no integration steps");
    DeviceGuard dg(L->ctx->device);
    if (int rc = sync_sync_steps(L, dt, n, drift)) return rc;
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return LQCD_OK;
}

int lqcd_integrate_sync_leap(lqcd_lattice* L, double dt)
{
    if (int rc = check_integration(L, dt, LQCD_DRIFT_EXP)) return rc;
    DeviceGuard dg(L->ctx->device);
    if (int rc = ensure_md_variant(L, LQCD_DRIFT_EXP)) return rc;
    HIP_TRY(launch_kick(L, L->U[L->cur], L->E, dt / 2.0));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return LQCD_OK;
}

int lqcd_integrate_leap_leap(lqcd_lattice* L, double dt, int drift)
{
    if (int rc = check_integration(L, dt, drift)) return rc;
    DeviceGuard dg(L->ctx->device);
    if (int rc = leap_steps(L, dt, 1, drift, 1.0)) return rc;
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return LQCD_OK;
}

int lqcd_simulate_leap_n(lqcd_lattice* L, double dt, size_t n, int drift)
{
    if (int rc = check_integration(L, dt, drift)) return rc;
    if (n == 0) return fail(LQCD_ERR_ZERO_INTEGRATION, "no integration steps");
    DeviceGuard dg(L->ctx->device);
    if (int rc = leap_steps(L, dt, n, drift, 1.0)) return rc;
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return LQCD_OK;
}

int lqcd_integrate_leap_sync(lqcd_lattice* L, double dt, int drift)
{
    if (int rc = check_integration(L, dt, drift)) return rc;
    DeviceGuard dg(L->ctx->device);
    if (int rc = leap_steps(L, dt, 1, drift, 0.5)) return rc;
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return LQCD_OK;
}

int lqcd_simulate_symplectic_n(lqcd_lattice* L, double dt, size_t n, int drift, int scheme)
{
    CHECK_ARG(L, "lattice is null");
    if (n == 0) return fail(LQCD_ERR_ZERO_INTEGRATION, "no integration steps");
    CHECK_ARG(std::isfinite(dt), "dt must be finite");
    CHECK_ARG(drift == LQCD_DRIFT_ADDITIVE || drift == LQCD_DRIFT_EXP, "unknown drift mode");
    CHECK_ARG(scheme == LQCD_SCHEME_REFERENCE || scheme == LQCD_SCHEME_LEAPFROG, "unknown scheme");
    DeviceGuard dg(L->ctx->device);
    int last = L->cur;
    if (int rc = integrate_steps(L, dt, n, drift, scheme, &last)) return rc;
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    L->cur = last;
    return LQCD_OK;
}

int lqcd_hmc_trajectory(lqcd_lattice* L, double beta, double dt, size_t n, int drift, int scheme,
                        uint64_t uniform, lqcd_hmc_stats* stats)
{
    CHECK_ARG(L, "lattice is null");
    if (n == 0) return fail(LQCD_ERR_ZERO_INTEGRATION, "no integration steps");
    CHECK_ARG(std::isfinite(dt) && std::isfinite(beta), "dt and beta must be finite");
    CHECK_ARG(drift == LQCD_DRIFT_ADDITIVE || drift == LQCD_DRIFT_EXP, "unknown drift mode");
    CHECK_ARG(scheme == LQCD_SCHEME_REFERENCE || scheme == LQCD_SCHEME_LEAPFROG, "unknown scheme");
    DeviceGuard dg(L->ctx->device);
    // The per-site MD kernels sum H_old's and H_new's terms on the way (the
    // first launch over the starting links and momenta, the final kick over
    // the final ones; integrate_steps): every plaquette enters
    // sum Re Tr(U Sigma) through its 4 links, so
    // H_U = beta (N_plaquettes - sum Re Tr(U Sigma) / 12), H_E = beta sum e^2 / 2.
    // Other work decompositions use the plaquette kernel before and after.
    // Either way the sums stay on the device until the MD is done: one host
    // round trip per trajectory.
    int last = L->cur;
    bool fused = false;
    if (int rc = integrate_steps(L, dt, n, drift, scheme, &last, &fused)) return rc;
    double h_old, h_new;
    if (fused) {
        if (int rc = read_results(L, 4)) return rc;
        const double nplaq = (double)L->g.nsite * (double)(L->g.dim * (L->g.dim - 1) / 2);
        h_old = beta * (nplaq - L->host_result[0] / 12.0) + beta * (0.5 * L->host_result[1]);
        h_new = beta * (nplaq - L->host_result[2] / 12.0) + beta * (0.5 * L->host_result[3]);
    } else {
        if (int rc = read_results(L, 8)) return rc;
        h_old = beta * L->host_result[2] + beta * L->host_result[3];
        h_new = beta * L->host_result[6] + beta * L->host_result[7];
    }
    // probability_of_replacement, hybrid_monte_carlo.rs:584-589 (clamped twice)
    double p = std::fmax(std::fmin(std::exp(h_old - h_new), 1.0), 0.0);
    p = std::fmax(std::fmin(p, 1.0), 0.0);
    bool accept;
    if (p >= 1.0) {
        accept = true;
    } else {
        const uint64_t p_int = (uint64_t)(p * 18446744073709551616.0);
        accept = uniform < p_int;
    }
    if (accept) L->cur = last;
    if (stats) {
        stats->h_old = h_old;
        stats->h_new = h_new;
        stats->delta_h = h_new - h_old;
        stats->probability = p;
        stats->accepted = accept ? 1 : 0;
    }
    return LQCD_OK;
}

int lqcd_wilson_forward(lqcd_lattice* L, double beta, double* action)
{
    return lqcd_hamiltonian_links(L, beta, action);
}

int lqcd_wilson_backward(lqcd_lattice* L, double beta, double* grad, size_t n_links)
{
    CHECK_ARG(std::isfinite(beta), "beta must be finite");
    CHECK_ARG(L, "lattice is null");
    return force_to_host(L, grad, n_links, -(beta * L->size) / std::sqrt(6.0));
}

int lqcd_wilson_backward_device(lqcd_lattice* L, double beta, double* d_grad, size_t n_links)
{
    CHECK_ARG(std::isfinite(beta), "beta must be finite");
    CHECK_ARG(L && d_grad, "null argument");
    if (int rc = check_links(L, n_links)) return rc;
    DeviceGuard dg(L->ctx->device);
    HIP_TRY(launch_force_store(L, L->U[L->cur], d_grad, -(beta * L->size) / std::sqrt(6.0)));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    return LQCD_OK;
}

static int metropolis_prepare(lqcd_lattice* L, double beta, size_t n_update)
{
    CHECK_ARG(n_update > 0, "n_update must be positive");
    CHECK_ARG(std::isfinite(beta), "beta must be finite");
    for (int i = 0; i < L->g.dim; ++i)
        if (L->g.extent[i] % 2)
            return fail(LQCD_ERR_UNSUPPORTED, "checkerboard sweep needs even extents");
    if (!L->mh_p) HIP_TRY(hipMalloc((void**)&L->mh_p, sizeof(double) * L->nlink));
    if (!L->mh_acc) HIP_TRY(hipMalloc((void**)&L->mh_acc, sizeof(uint32_t) * L->nlink));
    return LQCD_OK;
}

static int metropolis_finish(lqcd_lattice* L, lqcd_mh_stats* stats)
{
    HIP_TRY(launch_mh_stats(L, L->result));
    if (int rc = read_results(L, 2)) return rc;
    if (stats) {
        stats->prob_replace_mean = L->host_result[0] / (double)L->nlink;
        stats->number_replace_last = (size_t)L->host_result[1];
    }
    return LQCD_OK;
}

int lqcd_metropolis_sweep_random(lqcd_lattice* L, double beta, size_t n_update, double spread,
                                 uint64_t seed, uint64_t counter, lqcd_mh_stats* stats)
{
    CHECK_ARG(L, "lattice is null");
    CHECK_ARG(std::isfinite(spread) && spread > 0.0 && spread < 1.0, "spread must be in (0, 1)");
    CHECK_ARG(n_update <= kMaxMhUpdates, "n_update exceeds the device proposal stream");
    DeviceGuard dg(L->ctx->device);
    if (int rc = metropolis_prepare(L, beta, n_update)) return rc;
    // the device proposals are orthonormalised (in SU(3)): links in SU(3)
    // stay there, and the staples read them as two rows
    R12Scope scope{L};
    if (L->md_r12_mode) {
        bool ok;
        if (int rc = links_su3(L, (int64_t)n_update, &ok)) return rc;
        L->md_r12 = ok;
    }
    L->md_r12_last = L->md_r12;
    for (int mu = 0; mu < L->g.dim; ++mu)
        for (int par = 0; par < 2; ++par)
            HIP_TRY(launch_metropolis_class_rng(L, L->U[L->cur], mu, par, beta, n_update, spread, seed,
                                                counter));
    return metropolis_finish(L, stats);
}

int lqcd_metropolis_sweep(lqcd_lattice* L, double beta, size_t n_update, const double* proposals,
                          const uint64_t* uniforms, lqcd_mh_stats* stats)
{
    CHECK_ARG(L && proposals && uniforms, "null argument");
    DeviceGuard dg(L->ctx->device);
    if (int rc = metropolis_prepare(L, beta, n_update)) return rc;
    DevBuf dp, du;
    const size_t pbytes = sizeof(double) * 18 * L->nlink * n_update;
    HIP_TRY(hipMalloc(&dp.p, pbytes));
    HIP_TRY(hipMalloc(&du.p, sizeof(uint64_t) * L->nlink));
    HIP_TRY(hipMemcpyAsync(dp.p, proposals, pbytes, hipMemcpyHostToDevice, L->ctx->stream));
    HIP_TRY(hipMemcpyAsync(du.p, uniforms, sizeof(uint64_t) * L->nlink, hipMemcpyHostToDevice,
                           L->ctx->stream));
    for (int mu = 0; mu < L->g.dim; ++mu)
        for (int par = 0; par < 2; ++par)
            HIP_TRY(launch_metropolis_class(L, L->U[L->cur], mu, par, beta, n_update,
                                            (const double*)dp.p, (const uint64_t*)du.p));
    // the caller's proposals need not be in SU(3)
    L->su3_known = -1;
    return metropolis_finish(L, stats);
}

int lqcd_metropolis_link_random(lqcd_lattice* L, double beta, double spread, uint64_t seed, uint64_t counter,
                                double* probability, int* accepted, size_t* link)
{
    CHECK_ARG(L, "lattice is null");
    CHECK_ARG(std::isfinite(beta), "beta must be finite");
    // MetropolisHastingsDeltaDiagnostic::new returns None outside (0, 1)
    CHECK_ARG(spread > 0.0 && spread < 1.0, "spread must be in (0, 1)");
    DeviceGuard dg(L->ctx->device);
    HIP_TRY(launch_metropolis_link(L, L->U[L->cur], beta, spread, seed, counter, L->result));
    if (int rc = read_results(L, 3)) return rc;
    if (probability) *probability = L->host_result[0];
    if (accepted) *accepted = L->host_result[1] != 0.0 ? 1 : 0;
    if (link) *link = (size_t)L->host_result[2];
    // the proposal is orthonormalised: one SU(3)-preserving update
    if (L->host_result[1] != 0.0 && L->su3_known >= 0) ++L->su3_known;
    return LQCD_OK;
}

int lqcd_heat_bath_sweep(lqcd_lattice* L, double beta, uint64_t seed, uint64_t counter, size_t* exhausted)
{
    CHECK_ARG(L, "lattice is null");
    // HeatBathDistribution::new(beta * k) needs a positive finite parameter
    CHECK_ARG(std::isfinite(beta) && beta > 0.0, "beta must be positive and finite");
    DeviceGuard dg(L->ctx->device);
    if (int rc = metropolis_prepare(L, beta, 1)) return rc;
    for (int mu = 0; mu < L->g.dim; ++mu)
        for (int par = 0; par < 2; ++par)
            HIP_TRY(launch_heat_bath_class(L, L->U[L->cur], mu, par, beta, seed, counter));
    // the subgroup factors follow the reference's sigma_3 = identity: off SU(3)
    L->su3_known = -2;
    HIP_TRY(launch_mh_stats(L, L->result));
    if (int rc = read_results(L, 2)) return rc;
    if (exhausted) *exhausted = (size_t)L->host_result[1];
    return LQCD_OK;
}

static int overrelaxation(lqcd_lattice* L, bool reverse)
{
    CHECK_ARG(L, "lattice is null");
    for (int i = 0; i < L->g.dim; ++i)
        if (L->g.extent[i] % 2) return fail(LQCD_ERR_UNSUPPORTED, "checkerboard sweep needs even extents");
    DeviceGuard dg(L->ctx->device);
    for (int mu = 0; mu < L->g.dim; ++mu)
        for (int par = 0; par < 2; ++par) HIP_TRY(launch_overrelax_class(L, L->U[L->cur], mu, par, reverse));
    HIP_TRY(hipStreamSynchronize(L->ctx->stream));
    // unitary, but det U is not kept at 1
    L->su3_known = -2;
    return LQCD_OK;
}

int lqcd_overrelaxation_sweep(lqcd_lattice* L) { return overrelaxation(L, false); }

int lqcd_overrelaxation_reverse_sweep(lqcd_lattice* L) { return overrelaxation(L, true); }

int lqcd_normalize_link_matrices(lqcd_lattice* L)
{
    CHECK_ARG(L, "lattice is null");
    DeviceGuard dg(L->ctx->device);
    // enqueued only: later calls on the context are stream-ordered after it
    HIP_TRY(launch_reunitarize(L, L->U[L->cur]));
    L->su3_known = 0;
    return LQCD_OK;
}

int lqcd_stream_copy(lqcd_ctx* ctx, size_t bytes, int iters, double* gb_per_s)
{
    CHECK_ARG(ctx && gb_per_s, "null argument");
    CHECK_ARG(bytes >= 16 && bytes % 16 == 0, "bytes must be a positive multiple of 16");
    CHECK_ARG(iters > 0, "iters must be positive");
    DeviceGuard dg(ctx->device);
    DevBuf a, b;
    HIP_TRY(hipMalloc(&a.p, bytes));
    HIP_TRY(hipMalloc(&b.p, bytes));
    HIP_TRY(launch_stream_copy(ctx->stream, a.p, b.p, bytes));  // first touch
    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    hipError_t e = hipEventRecord(t0, ctx->stream);
    for (int i = 0; i < iters && e == hipSuccess; ++i) e = launch_stream_copy(ctx->stream, a.p, b.p, bytes);
    if (e == hipSuccess) e = hipEventRecord(t1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(t1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (e != hipSuccess) return fail(LQCD_ERR_DEVICE, std::string("stream copy: ") + hipGetErrorString(e));
    *gb_per_s = 2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9;
    return LQCD_OK;
}

int lqcd_stream_read(lqcd_ctx* ctx, size_t bytes, int iters, double* gb_per_s)
{
    CHECK_ARG(ctx && gb_per_s, "null argument");
    CHECK_ARG(bytes >= 16 && bytes % 16 == 0, "bytes must be a positive multiple of 16");
    CHECK_ARG(iters > 0, "iters must be positive");
    DeviceGuard dg(ctx->device);
    DevBuf a, sink;
    HIP_TRY(hipMalloc(&a.p, bytes));
    HIP_TRY(hipMalloc(&sink.p, sizeof(uint32_t) * (size_t)kStreamReadBlocks * kBlock));
    HIP_TRY(hipMemsetAsync(a.p, 0, bytes, ctx->stream));  // first touch
    HIP_TRY(launch_stream_read(ctx->stream, a.p, bytes, (uint32_t*)sink.p));
    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    hipError_t e = hipEventRecord(t0, ctx->stream);
    for (int i = 0; i < iters && e == hipSuccess; ++i) e = launch_stream_read(ctx->stream, a.p, bytes, (uint32_t*)sink.p);
    if (e == hipSuccess) e = hipEventRecord(t1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(t1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (e != hipSuccess) return fail(LQCD_ERR_DEVICE, std::string("stream read: ") + hipGetErrorString(e));
    *gb_per_s = (double)bytes * iters / (ms * 1e-3) / 1e9;
    return LQCD_OK;
}

int lqcd_md_autotune(lqcd_lattice* L, int drift, int* variant)
{
    CHECK_ARG(L, "lattice is null");
    CHECK_ARG(drift == LQCD_DRIFT_ADDITIVE || drift == LQCD_DRIFT_EXP, "unknown drift mode");
    DeviceGuard dg(L->ctx->device);
    return autotune_md(L, drift, variant);
}

int lqcd_md_variant_set(lqcd_lattice* L, int variant)
{
    CHECK_ARG(L, "lattice is null");
    CHECK_ARG(variant >= -1 && variant < VAR_COUNT, "unknown MD variant");
    L->md_variant = variant;
    return LQCD_OK;
}

int lqcd_md_variant_get(const lqcd_lattice* L, int* variant)
{
    CHECK_ARG(L && variant, "null argument");
    *variant = L->md_variant;
    return LQCD_OK;
}

int lqcd_md_reconstruct_set(lqcd_lattice* L, int mode)
{
    CHECK_ARG(L, "lattice is null");
    CHECK_ARG(mode == 0 || mode == 1, "mode must be 0 or 1");
    L->md_r12_mode = mode;
    return LQCD_OK;
}

int lqcd_md_graph_set(lqcd_lattice* L, int enable)
{
    CHECK_ARG(L, "lattice is null");
    CHECK_ARG(enable == 0 || enable == 1, "enable must be 0 or 1");
    L->md_graphs = enable == 1;
    return LQCD_OK;
}

int lqcd_md_reconstruct_get(const lqcd_lattice* L, int* mode, int* last_used)
{
    CHECK_ARG(L && mode && last_used, "null argument");
    *mode = L->md_r12_mode;
    *last_used = L->md_r12_last ? 1 : 0;
    return LQCD_OK;
}

int lqcd_su3_distance(lqcd_lattice* L, double* out)
{
    CHECK_ARG(L && out, "null argument");
    DeviceGuard dg(L->ctx->device);
    HIP_TRY(launch_su3_distance(L, L->U[L->cur], L->result));
    if (int rc = read_results(L, 1)) return rc;
    *out = L->host_result[0];
    return LQCD_OK;
}

int lqcd_profile_enable(lqcd_ctx* ctx, unsigned mask)
{
    CHECK_ARG(ctx, "ctx is null");
    DeviceGuard dg(ctx->device);
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (auto& e : ctx->events) ctx->event_pool.push_back(e);
    ctx->events.clear();
    ctx->profile_mask = mask;
    ctx->profiling = mask != 0;
    return LQCD_OK;
}

int lqcd_profile_read(lqcd_ctx* ctx, int kind, size_t* count, double* total_ms)
{
    CHECK_ARG(ctx && count && total_ms, "null argument");
    DeviceGuard dg(ctx->device);
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    size_t c = 0;
    double t = 0.0;
    for (auto& e : ctx->events) {
        if (e.kind != kind) continue;
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e.start, e.stop));
        t += ms;
        c += (size_t)e.launches;
    }
    *count = c;
    *total_ms = t;
    return LQCD_OK;
}

} // extern "C"
