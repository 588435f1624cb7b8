// capi_batch.cpp -- extern "C" functions of the batched chains (lqcd_batch_*).
//
// A trajectory of all chains is a fixed sequence of launches on the context
// stream with no host synchronisation inside (the accept decision and the
// trajectory counter live on the device), so it is captured once as a HIP
// graph and replayed; the graph holds kernel nodes only.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/lqcd.h"
#include "batch.hpp"
#include "capi_util.hpp"

using namespace lqcd;

namespace {

void free_batch(lqcd_batch* b)
{
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    for (int k = 0; k < 3; ++k)
        if (b->U[k]) (void)hipFree(b->U[k]);
    void* dev[] = {b->E, b->d_chains, b->d_tile_chain, b->d_partials, b->d_h_old, b->d_h_new,
                   b->d_accept, b->d_stats, b->d_scratch, b->d_counter};
    for (void* p : dev)
        if (p) (void)hipFree(p);
    if (b->h_stats) (void)hipHostFree(b->h_stats);
    if (b->h_scratch) (void)hipHostFree(b->h_scratch);
    delete b;
}

// a single-lattice view of chain c (its geometry and planes) for the
// layout-conversion kernels of the lattice path
lqcd_lattice chain_view(const lqcd_batch* b, size_t c)
{
    lqcd_lattice v;
    v.ctx = b->ctx;
    v.g = b->chains[c].g;
    v.size = b->chains[c].size;
    v.precision = b->precision;
    v.elem = b->elem;
    v.nlink = (size_t)v.g.nsite * (size_t)b->dim;
    return v;
}

void* chain_links(const lqcd_batch* b, int buf, size_t c)
{
    return (char*)b->U[buf] + (size_t)b->chains[c].u_off * 2 * b->elem;
}

int check_chain(const lqcd_batch* b, size_t c, size_t n_links)
{
    CHECK_ARG(c < b->chains.size(), "chain index out of range");
    const size_t nl = (size_t)b->chains[c].g.nsite * (size_t)b->dim;
    if (n_links != nl)
        return fail(LQCD_ERR_INCOMPATIBLE_SIZE, "size of lattice and data are incompatible: expected " +
                                                    std::to_string(nl) + " links, got " +
                                                    std::to_string(n_links));
    return LQCD_OK;
}

// the MD steps of one trajectory from U[0]; the final links end in U[*last]
hipError_t batch_integrate(const lqcd_batch* b, int config, double dt, size_t n, int drift, int scheme,
                           int* last)
{
    int src = 0, dst = 1;
    hipError_t e = hipSuccess;
    if (scheme == LQCD_SCHEME_REFERENCE) {
        for (size_t k = 0; k < n && e == hipSuccess; ++k) {
            e = launch_batch_md(b, config, b->U[src], b->U[dst], dt / 2.0, dt, drift);
            if (e == hipSuccess) e = launch_batch_md(b, config, b->U[dst], nullptr, dt / 2.0, dt, BATCH_KICK);
            src = dst;
            dst = dst == 1 ? 2 : 1;
        }
    } else {
        for (size_t k = 0; k < n && e == hipSuccess; ++k) {
            e = launch_batch_md(b, config, b->U[src], b->U[dst], k == 0 ? dt / 2.0 : dt, dt, drift);
            src = dst;
            dst = dst == 1 ? 2 : 1;
        }
        if (e == hipSuccess) e = launch_batch_md(b, config, b->U[src], nullptr, dt / 2.0, dt, BATCH_KICK);
    }
    *last = src;
    return e;
}

// every launch of one trajectory, in stream order
hipError_t batch_trajectory(const lqcd_batch* b, double dt, size_t n, int drift, int scheme, int reunit)
{
    int last = 0;
    hipError_t e = launch_batch_momenta(b);
    if (e == hipSuccess) e = launch_batch_sums(b, b->U[0], b->E, b->d_h_old);
    if (e == hipSuccess) e = batch_integrate(b, b->config, dt, n, drift, scheme, &last);
    // two-row MD kernels store rows 0, 1 of the links they drift
    if (e == hipSuccess && b->md_r12 && n > 0) e = launch_batch_complete_row2(b, b->U[last]);
    if (e == hipSuccess) e = launch_batch_sums(b, b->U[last], b->E, b->d_h_new);
    if (e == hipSuccess) e = launch_batch_accept(b);
    if (e == hipSuccess) e = launch_batch_accept_copy(b, b->U[last]);
    if (e == hipSuccess && reunit) e = launch_batch_reunitarize(b);
    if (e == hipSuccess) e = launch_batch_counter_advance(b);
    return e;
}

int capture_trajectory(lqcd_batch* b, double dt, size_t n, int drift, int scheme, int reunit)
{
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    HIP_TRY(prepare_batch_md(b, b->config, drift));
    HIP_TRY(prepare_batch_md(b, b->config, BATCH_KICK));
    hipStream_t st = b->ctx->stream;
    // timing events stay out of the graph: the graph holds kernel nodes only
    const bool profiling = b->ctx->profiling;
    b->ctx->profiling = false;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const hipError_t le = batch_trajectory(b, dt, n, drift, scheme, reunit);
    hipGraph_t graph = nullptr;
    const hipError_t ce = hipStreamEndCapture(st, &graph);
    b->ctx->profiling = profiling;
    if (le != hipSuccess || ce != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        return fail(LQCD_ERR_DEVICE, std::string("trajectory capture: ") +
                                         hipGetErrorString(le != hipSuccess ? le : ce));
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) return fail(LQCD_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(ie));
    b->graph = exec;
    b->g_dt = dt;
    b->g_steps = n;
    b->g_drift = drift;
    b->g_r12 = b->md_r12 ? 1 : 0;
    b->g_scheme = scheme;
    b->g_reunit = reunit;
    b->g_config = b->config;
    return LQCD_OK;
}

// Times every launch shape on one fused kick+drift step over all chains
// (kick 0; U[0] is only read, U[1] and E are scratch between trajectories)
// and keeps the fastest.
int autotune_batch(lqcd_batch* b, int drift, int* chosen)
{
    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    int best = 4;
    float best_ms = 1e30f;
    hipError_t e = hipSuccess;
    for (int c = 0; c < kBatchConfigs && e == hipSuccess; ++c) {
        e = launch_batch_md(b, c, b->U[0], b->U[1], 0.0, 1e-3, drift);
        if (e == hipSuccess) e = hipEventRecord(t0, b->ctx->stream);
        for (int r = 0; r < 3 && e == hipSuccess; ++r) e = launch_batch_md(b, c, b->U[0], b->U[1], 0.0, 1e-3, drift);
        if (e == hipSuccess) e = hipEventRecord(t1, b->ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(t1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
        if (e == hipSuccess && ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (e != hipSuccess) return fail(LQCD_ERR_DEVICE, std::string("batch autotune: ") + hipGetErrorString(e));
    b->config = best;
    if (chosen) *chosen = best;
    return LQCD_OK;
}

int ensure_config(lqcd_batch* b, int drift)
{
    if (b->config >= 0) return LQCD_OK;
    const char* env = std::getenv("LQCD_BATCH_CONFIG");
    if (env && *env) {
        const int c = std::atoi(env);
        if (c >= 0 && c < kBatchConfigs) {
            b->config = c;
            return LQCD_OK;
        }
    }
    return autotune_batch(b, drift, nullptr);
}

int reserve_host_stats(lqcd_batch* b, size_t n)
{
    if (n <= b->h_stats_cap) return LQCD_OK;
    if (b->h_stats) (void)hipHostFree(b->h_stats);
    b->h_stats = nullptr;
    b->h_stats_cap = 0;
    HIP_TRY(hipHostMalloc((void**)&b->h_stats, n * sizeof(lqcd_chain_stats), 0));
    b->h_stats_cap = n;
    return LQCD_OK;
}

} // namespace

extern "C" {

int lqcd_batch_create(lqcd_ctx* ctx, int dim, int precision, const lqcd_chain_spec* specs, size_t nchain,
                      lqcd_batch** out)
{
    CHECK_ARG(ctx && out, "null argument");
    if (dim == 0) return fail(LQCD_ERR_ZERO_DIMENSION, "the lattice has zero dimension");
    if (dim < 2 || dim > kMaxDim) return fail(LQCD_ERR_UNSUPPORTED, "dimension must be 2, 3 or 4");
    CHECK_ARG(precision == LQCD_F64 || precision == LQCD_F32, "unknown precision");
    CHECK_ARG(specs && nchain > 0, "at least one chain is needed");
    CHECK_ARG(nchain <= (1u << 20), "too many chains");
    for (size_t c = 0; c < nchain; ++c) {
        const lqcd_chain_spec& sp = specs[c];
        if (!(sp.size > 0.0) || !std::isfinite(sp.size))
            return fail(LQCD_ERR_NON_POSITIVE_SIZE, "the lattice spacing must be positive and finite");
        CHECK_ARG(std::isfinite(sp.beta) && sp.beta != 0.0, "beta must be finite and non-zero");
        for (int i = 0; i < dim; ++i) {
            if (sp.extent[i] < 2) return fail(LQCD_ERR_DIM_TOO_SMALL, "every extent must be at least 2");
            CHECK_ARG(sp.extent[i] <= (size_t)(1 << 30), "extent too large");
        }
    }
    DeviceGuard dg(ctx->device);
    lqcd_batch* b = new lqcd_batch;
    b->ctx = ctx;
    b->dim = dim;
    b->precision = precision;
    b->elem = precision == LQCD_F32 ? 4 : 8;
    int64_t u_off = 0, e_off = 0, tile0 = 0, nsite_all = 0;
    for (size_t c = 0; c < nchain; ++c) {
        ChainDesc d{};
        d.g.dim = dim;
        int64_t n = 1;
        for (int i = 0; i < kMaxDim; ++i) {
            d.g.extent[i] = 1;
            d.g.stride[i] = 0;
        }
        for (int i = 0; i < dim; ++i) {
            d.g.extent[i] = (int)specs[c].extent[i];
            d.g.stride[i] = n;
            n *= (int64_t)specs[c].extent[i];
            if (n > ((int64_t)1 << 36)) {
                delete b;
                return fail(LQCD_ERR_INVALID_ARGUMENT, "lattice too large");
            }
        }
        d.g.nsite = n;
        d.g.npad = (n + 255) / 256 * 256;
        if ((uint64_t)d.g.npad * 9 * 2 * b->elem > kMaxPlaneBytes) {
            delete b;
            return fail(LQCD_ERR_UNSUPPORTED,
                        "chain too large: the link planes of one direction exceed 4 GiB");
        }
        d.u_off = u_off;
        d.e_off = e_off;
        d.tile0 = tile0;
        d.ntile = (n + kTile - 1) / kTile;
        d.beta = specs[c].beta;
        d.size = specs[c].size;
        d.sigma = 0.5 / specs[c].beta;
        d.seed = specs[c].seed;
        u_off += (int64_t)dim * 9 * d.g.npad;
        e_off += (int64_t)dim * 8 * d.g.npad;
        tile0 += d.ntile;
        nsite_all += n;
        b->chains.push_back(d);
    }
    b->ntile = tile0;
    b->nsite = nsite_all;
    b->u_elems = (size_t)u_off;
    b->e_elems = (size_t)e_off;
    std::vector<int32_t> tile_chain((size_t)b->ntile);
    for (size_t c = 0; c < nchain; ++c)
        for (int64_t t = 0; t < b->chains[c].ntile; ++t) tile_chain[(size_t)(b->chains[c].tile0 + t)] = (int32_t)c;

    hipError_t e = hipSuccess;
    const size_t ubytes = b->u_elems * 2 * b->elem;
    for (int k = 0; k < 3 && e == hipSuccess; ++k) e = hipMalloc(&b->U[k], ubytes);
    if (e == hipSuccess) e = hipMalloc(&b->E, b->e_elems * b->elem);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_chains, nchain * sizeof(ChainDesc));
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_tile_chain, tile_chain.size() * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_partials, (size_t)b->ntile * 4 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_h_old, nchain * 4 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_h_new, nchain * 4 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_scratch, nchain * 4 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_accept, nchain * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_stats, nchain * sizeof(lqcd_chain_stats));
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_counter, sizeof(uint64_t));
    if (e == hipSuccess) e = hipHostMalloc((void**)&b->h_scratch, nchain * 4 * sizeof(double), 0);
    hipStream_t st = ctx->stream;
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_chains, b->chains.data(), nchain * sizeof(ChainDesc), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_tile_chain, tile_chain.data(), tile_chain.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice, st);
    // padding entries of the planes stay defined
    for (int k = 0; k < 3 && e == hipSuccess; ++k) e = hipMemsetAsync(b->U[k], 0, ubytes, st);
    if (e == hipSuccess) e = hipMemsetAsync(b->E, 0, b->e_elems * b->elem, st);
    if (e == hipSuccess) e = hipMemsetAsync(b->d_counter, 0, sizeof(uint64_t), st);
    if (e == hipSuccess) e = launch_batch_cold(b);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        free_batch(b);
        return fail(e == hipErrorOutOfMemory ? LQCD_ERR_OUT_OF_MEMORY : LQCD_ERR_DEVICE,
                    std::string("batch allocation: ") + hipGetErrorString(e));
    }
    *out = b;
    return LQCD_OK;
}

int lqcd_batch_destroy(lqcd_batch* b)
{
    if (!b) return LQCD_OK;
    DeviceGuard dg(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    free_batch(b);
    return LQCD_OK;
}

int lqcd_batch_info(const lqcd_batch* b, size_t* nchain, size_t* nsite_total)
{
    CHECK_ARG(b, "batch is null");
    if (nchain) *nchain = b->chains.size();
    if (nsite_total) *nsite_total = (size_t)b->nsite;
    return LQCD_OK;
}

int lqcd_batch_chain_info(const lqcd_batch* b, size_t c, size_t* nsite, size_t* nlink)
{
    CHECK_ARG(b, "batch is null");
    CHECK_ARG(c < b->chains.size(), "chain index out of range");
    if (nsite) *nsite = (size_t)b->chains[c].g.nsite;
    if (nlink) *nlink = (size_t)b->chains[c].g.nsite * (size_t)b->dim;
    return LQCD_OK;
}

int lqcd_batch_cold_start(lqcd_batch* b)
{
    CHECK_ARG(b, "batch is null");
    DeviceGuard dg(b->ctx->device);
    HIP_TRY(launch_batch_cold(b));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    return LQCD_OK;
}

int lqcd_batch_hot_start(lqcd_batch* b)
{
    CHECK_ARG(b, "batch is null");
    DeviceGuard dg(b->ctx->device);
    HIP_TRY(launch_batch_hot_start(b));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    return LQCD_OK;
}

int lqcd_batch_links_upload(lqcd_batch* b, size_t c, const double* links, size_t n_links)
{
    CHECK_ARG(b && links, "null argument");
    if (int rc = check_chain(b, c, n_links)) return rc;
    DeviceGuard dg(b->ctx->device);
    lqcd_lattice v = chain_view(b, c);
    DevBuf st;
    HIP_TRY(hipMalloc(&st.p, sizeof(double) * 18 * n_links));
    HIP_TRY(hipMemcpyAsync(st.p, links, sizeof(double) * 18 * n_links, hipMemcpyHostToDevice, b->ctx->stream));
    HIP_TRY(launch_links_from_aos(&v, (const double*)st.p, chain_links(b, 0, c)));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    return LQCD_OK;
}

int lqcd_batch_links_download(lqcd_batch* b, size_t c, double* links, size_t n_links)
{
    CHECK_ARG(b && links, "null argument");
    if (int rc = check_chain(b, c, n_links)) return rc;
    DeviceGuard dg(b->ctx->device);
    lqcd_lattice v = chain_view(b, c);
    DevBuf st;
    HIP_TRY(hipMalloc(&st.p, sizeof(double) * 18 * n_links));
    HIP_TRY(launch_links_to_aos(&v, chain_links(b, 0, c), (double*)st.p));
    HIP_TRY(hipMemcpyAsync(links, st.p, sizeof(double) * 18 * n_links, hipMemcpyDeviceToHost, b->ctx->stream));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    return LQCD_OK;
}

int lqcd_batch_plaquette(lqcd_batch* b, double* plaquette)
{
    CHECK_ARG(b && plaquette, "null argument");
    DeviceGuard dg(b->ctx->device);
    const size_t n = b->chains.size();
    HIP_TRY(launch_batch_sums(b, b->U[0], nullptr, b->d_scratch));
    HIP_TRY(hipMemcpyAsync(b->h_scratch, b->d_scratch, n * 4 * sizeof(double), hipMemcpyDeviceToHost,
                           b->ctx->stream));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    for (size_t c = 0; c < n; ++c) {
        const double nplaq = (double)b->chains[c].g.nsite * (double)(b->dim * (b->dim - 1) / 2);
        plaquette[c] = b->h_scratch[c * 4] / nplaq / 3.0;
    }
    return LQCD_OK;
}

int lqcd_batch_counter_get(const lqcd_batch* b, uint64_t* counter)
{
    CHECK_ARG(b && counter, "null argument");
    DeviceGuard dg(b->ctx->device);
    HIP_TRY(hipMemcpyAsync(b->h_scratch, b->d_counter, sizeof(uint64_t), hipMemcpyDeviceToHost, b->ctx->stream));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    std::memcpy(counter, b->h_scratch, sizeof(uint64_t));
    return LQCD_OK;
}

int lqcd_batch_counter_set(lqcd_batch* b, uint64_t counter)
{
    CHECK_ARG(b, "batch is null");
    DeviceGuard dg(b->ctx->device);
    std::memcpy(b->h_scratch, &counter, sizeof(uint64_t));
    HIP_TRY(hipMemcpyAsync(b->d_counter, b->h_scratch, sizeof(uint64_t), hipMemcpyHostToDevice, b->ctx->stream));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    return LQCD_OK;
}

int lqcd_batch_hmc(lqcd_batch* b, double dt, size_t n_steps, int drift, int scheme, size_t n_traj,
                   int reunitarize, lqcd_chain_stats* stats)
{
    CHECK_ARG(b, "batch is null");
    if (n_steps == 0) return fail(LQCD_ERR_ZERO_INTEGRATION, "no integration steps");
    CHECK_ARG(std::isfinite(dt), "dt must be finite");
    CHECK_ARG(drift == LQCD_DRIFT_ADDITIVE || drift == LQCD_DRIFT_EXP, "unknown drift mode");
    CHECK_ARG(scheme == LQCD_SCHEME_REFERENCE || scheme == LQCD_SCHEME_LEAPFROG, "unknown scheme");
    if (n_traj == 0) return LQCD_OK;
    DeviceGuard dg(b->ctx->device);
    // two-row link reads: exp drift, reunitarised trajectories (every
    // trajectory then starts from links in SU(3)) and links in SU(3) now
    b->md_r12 = false;
    if (drift == LQCD_DRIFT_EXP && reunitarize && b->md_r12_mode) {
        HIP_TRY(launch_batch_su3_distance(b, b->U[0], b->d_scratch));
        HIP_TRY(hipMemcpyAsync(b->h_scratch, b->d_scratch, sizeof(double), hipMemcpyDeviceToHost, b->ctx->stream));
        HIP_TRY(hipStreamSynchronize(b->ctx->stream));
        b->md_r12 = b->h_scratch[0] <= r12_distance(b->precision);
    }
    if (int rc = ensure_config(b, drift)) return rc;
    const size_t nc = b->chains.size();
    if (stats)
        if (int rc = reserve_host_stats(b, n_traj * nc)) return rc;
    const int reunit = reunitarize ? 1 : 0;
    if (b->use_graph) {
        const bool stale = !b->graph || b->g_dt != dt || b->g_steps != n_steps || b->g_drift != drift ||
                           b->g_scheme != scheme || b->g_reunit != reunit || b->g_config != b->config ||
                           b->g_r12 != (b->md_r12 ? 1 : 0);
        if (stale)
            if (int rc = capture_trajectory(b, dt, n_steps, drift, scheme, reunit)) return rc;
    }
    for (size_t t = 0; t < n_traj; ++t) {
        if (b->use_graph)
            HIP_TRY(hipGraphLaunch(b->graph, b->ctx->stream));
        else
            HIP_TRY(batch_trajectory(b, dt, n_steps, drift, scheme, reunit));
        if (stats)
            HIP_TRY(hipMemcpyAsync(b->h_stats + t * nc, b->d_stats, nc * sizeof(lqcd_chain_stats),
                                   hipMemcpyDeviceToHost, b->ctx->stream));
    }
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    if (stats) std::memcpy(stats, b->h_stats, n_traj * nc * sizeof(lqcd_chain_stats));
    return LQCD_OK;
}

int lqcd_batch_autotune(lqcd_batch* b, int drift, int* config)
{
    CHECK_ARG(b, "batch is null");
    CHECK_ARG(drift == LQCD_DRIFT_ADDITIVE || drift == LQCD_DRIFT_EXP, "unknown drift mode");
    DeviceGuard dg(b->ctx->device);
    return autotune_batch(b, drift, config);
}

int lqcd_batch_config_set(lqcd_batch* b, int config)
{
    CHECK_ARG(b, "batch is null");
    CHECK_ARG(config >= 0 && config < kBatchConfigs, "config must be in 0..11");
    b->config = config;
    return LQCD_OK;
}

int lqcd_batch_config_get(const lqcd_batch* b, int* config)
{
    CHECK_ARG(b && config, "null argument");
    *config = b->config;
    return LQCD_OK;
}

int lqcd_batch_reconstruct_set(lqcd_batch* b, int mode)
{
    CHECK_ARG(b, "batch is null");
    CHECK_ARG(mode == 0 || mode == 1, "mode must be 0 or 1");
    b->md_r12_mode = mode;
    return LQCD_OK;
}

int lqcd_batch_reconstruct_get(const lqcd_batch* b, int* mode, int* last_used)
{
    CHECK_ARG(b && mode && last_used, "null argument");
    *mode = b->md_r12_mode;
    *last_used = b->md_r12 ? 1 : 0;
    return LQCD_OK;
}

int lqcd_batch_graph_enable(lqcd_batch* b, int enable)
{
    CHECK_ARG(b, "batch is null");
    b->use_graph = enable != 0;
    return LQCD_OK;
}

} // extern "C"
