// batch.hpp -- many independent chains (lattices of their own extents, beta
// and seed) resident in one set of buffers and advanced together.
//
// Every chain keeps the single-lattice structure-of-arrays layout with its
// own plane length npad; the chains' planes follow each other in one link
// buffer (offset u_off, in 2-vectors) and one E buffer (offset e_off, in
// reals).  Work is cut into tiles of kTile = 64 consecutive sites of one
// chain (one wavefront); tile_chain[t] names the chain of global tile t, and
// the tiles of a chain are [tile0, tile0 + ntile).
#pragma once
#include <vector>

#include "../../include/lqcd.h"
#include "internal.hpp"

namespace lqcd {

constexpr int kTile = 64;

struct ChainDesc {
    Geom g;
    int64_t u_off;   // first 2-vector of the chain in a link buffer
    int64_t e_off;   // first real of the chain in the E buffer
    int64_t tile0;   // first global tile of the chain
    int64_t ntile;   // tiles of the chain, ceil(nsite / kTile)
    double beta;
    double size;     // lattice spacing a
    double sigma;    // momentum width 0.5 / beta
    uint64_t seed;
};

// Launch shape of the batched MD kernel.  config = block + 3 * site + 6 * persistent:
//   block      0, 1, 2 -> 64, 128, 256 threads per workgroup (1, 2, 4 tiles at a time)
//   site       1: one thread per site for all D links; 0: one thread per link
//   persistent 1: as many workgroups as fit on the device at once, each walking
//              a contiguous range of tiles; 0: one workgroup per 1-4 tiles
constexpr int kBatchConfigs = 12;
struct BatchShape {
    int block;
    bool site;
    bool persistent;
};
inline BatchShape batch_shape(int config)
{
    return BatchShape{64 << (config % 3), ((config / 3) & 1) != 0, (config / 6) != 0};
}

// which MD operation one launch performs
enum BatchMdOp { BATCH_KICK = -1 };

} // namespace lqcd

struct lqcd_batch {
    lqcd_ctx* ctx = nullptr;
    int dim = 4;
    int precision = 0;
    size_t elem = 8;
    std::vector<lqcd::ChainDesc> chains;   // host copy of d_chains
    lqcd::ChainDesc* d_chains = nullptr;
    int32_t* d_tile_chain = nullptr;
    int64_t ntile = 0;
    int64_t nsite = 0;                     // all chains
    size_t u_elems = 0;                    // 2-vectors per link buffer
    size_t e_elems = 0;                    // reals in the E buffer
    void* U[3] = {nullptr, nullptr, nullptr};  // U[0] holds the chains' current links
    void* E = nullptr;
    double* d_partials = nullptr;          // [ntile][4]
    double* d_h_old = nullptr;             // [nchain][4] sums before the trajectory
    double* d_h_new = nullptr;             // [nchain][4] sums after it
    uint32_t* d_accept = nullptr;          // [nchain]
    lqcd_chain_stats* d_stats = nullptr;   // [nchain], last trajectory
    lqcd_chain_stats* h_stats = nullptr;   // pinned, [h_stats_cap]
    size_t h_stats_cap = 0;
    double* d_scratch = nullptr;           // [nchain][4]
    double* h_scratch = nullptr;           // pinned [nchain][4]
    uint64_t* d_counter = nullptr;         // trajectory counter of the random streams
    int config = -1;                       // tuned launch shape (-1: not tuned yet)
    bool use_graph = true;
    hipGraphExec_t graph = nullptr;        // captured trajectory
    double g_dt = 0.0;                     // what the captured trajectory was built for
    size_t g_steps = 0;
    int g_drift = -1, g_scheme = -1, g_reunit = -1, g_config = -1, g_r12 = -1;
    int md_r12_mode = 1;                   // as lqcd_lattice::md_r12_mode
    bool md_r12 = false;                   // the current lqcd_batch_hmc reads links as two rows
};

namespace lqcd {

// ---- batched kernels (k_batch.hip); all on b->ctx->stream ----
hipError_t launch_batch_md(const lqcd_batch* b, int config, const void* U_in, void* U_out, double kick,
                           double dt, int op);
// settles the grid of launch_batch_md(config, op) without launching
hipError_t prepare_batch_md(const lqcd_batch* b, int config, int op);
hipError_t launch_batch_momenta(const lqcd_batch* b);
// largest distance from SU(3) over all chains' links in U, into d_out1
hipError_t launch_batch_su3_distance(const lqcd_batch* b, const void* U, double* d_out1);
hipError_t launch_batch_hot_start(const lqcd_batch* b);
hipError_t launch_batch_cold(const lqcd_batch* b);
// per chain: [Re sum Tr P, Im sum Tr P, sum (1 - Re Tr P / 3), sum e^2 / 2] (E may be null)
hipError_t launch_batch_sums(const lqcd_batch* b, const void* U, const void* E, double* d_out);
hipError_t launch_batch_accept(const lqcd_batch* b);
hipError_t launch_batch_accept_copy(const lqcd_batch* b, const void* U_final);
hipError_t launch_batch_reunitarize(const lqcd_batch* b);
// row 2 of every chain's links in U rebuilt from rows 0, 1 (SU(3) links)
hipError_t launch_batch_complete_row2(const lqcd_batch* b, void* U);
hipError_t launch_batch_counter_advance(const lqcd_batch* b);

} // namespace lqcd
