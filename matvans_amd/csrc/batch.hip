// Persistent batched fused decode-matvec: every matrix of a batch (uneven shapes, mixed
// encodings, 1..8 vectors each) in one launch that pulls row tiles from a device work
// queue.  The reference launches one kernel per matrix (matvansWarpPerRow.cu:156-164,
// buffered/ans-warp-rows.cu:143-164); this has no counterpart there.
//
// Work items: a "unit" is the group of rows one 4-wave block decodes in the one-matrix
// kernels (blk_v1.h, blk_b.h); a tile is `units_per_tile` consecutive units of one
// matrix.  The tile table is built on the host in the configured order (matrix-major,
// round-robin over matrices, or longest decode chain first) and copied to the device at
// configuration time.  Schedule 0 is a device work queue in 8 shards (block b draws from
// shard b % 8, one counter per shard on its own cache line): a block's first tile is
// fixed, later ones are drawn with one atomic add each, and the block that draws a
// shard's last ticket resets that shard's counter, so the next launch (or graph replay)
// starts from zero with no host-side or memset work on the launch path.  One shared
// counter, with a per-block exit count for the reset, serialised every block on two
// addresses and cost more than the decode.  Schedule 1 is static: block b of shard g
// takes the shard's tiles at positions b / 8, b / 8 + nb, ... and uses no atomics.
//
// The plan owns its device descriptors, tile table, counters and the decode tables it
// builds for .bin v0 items; matvans_batch_launch allocates nothing and is one kernel
// launch, so a caller can capture it in a HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "blk_b.h"
#include "blk_bwd.h"
#include "blk_v1.h"
#include "launch.h"
#include "matvans_internal.h"
#include "v0_stream.h"

using mv::fail;

namespace mv {
namespace {

// decode path of an item: v0, quad with 32 / 64 / 128 streams, family-B rows / stripes,
// quad with 32 streams and the products on the matrix cores (six or more vectors), and the
// quad streams back to back (coder 2) with 32 or 64 streams, or 32 on the matrix cores
enum : uint32_t {
    kPathV0 = 0, kPathQ32 = 1, kPathQ64 = 2, kPathBRows = 3, kPathBStripes = 4, kPathQ128 = 5,
    kPathQ32M = 6, kPathL32 = 7, kPathL64 = 8, kPathL32M = 9
};
// from this many vectors on, quad items with 32 streams per row take the matrix-core path
constexpr uint32_t kBatchMfmaMinVec = 6;

struct Item {
    uint32_t path;
    uint32_t units;
    blk::V1Job v1;
    blk::BJob b;
};

struct Tile {
    uint32_t item, first, count, pad;
};


constexpr uint32_t bit(uint32_t path) { return 1u << path; }

// LDS of a kernel instance: the union of the block bodies of the paths it contains only,
// so an instance without the matrix-core path does not pay for its stage
struct NoSmem {};
template <bool On, class T>
using Opt = typename std::conditional<On, T, NoSmem>::type;
template <int NV, uint32_t MASK>
union BatchSmem {
    Opt<(MASK & bit(kPathV0)) != 0, blk::V1Smem<NV, 32, 8, 0>> v0;
    Opt<(MASK & bit(kPathQ32)) != 0, blk::QuadSmem<NV, 32>> q32;
    Opt<(MASK & bit(kPathQ64)) != 0, blk::QuadSmem<NV, 64>> q64;
    Opt<(MASK & bit(kPathQ128)) != 0, blk::QuadSmem<NV, 128>> q128;
    Opt<(MASK & (bit(kPathBRows) | bit(kPathBStripes))) != 0, blk::BSmem> b;
    Opt<(MASK & bit(kPathQ32M)) != 0, blk::QuadMfmaSmem<4>> q32m;
    Opt<(MASK & bit(kPathL32)) != 0, blk::QuadSmem<NV, 32>> l32;
    Opt<(MASK & bit(kPathL64)) != 0, blk::QuadSmem<NV, 64>> l64;
    Opt<(MASK & bit(kPathL32M)) != 0, blk::QuadMfmaSmem<4>> l32m;
};
// kernel instances: the decode paths each one contains (a plan picks the first one that
// covers its items; fewer paths leave fewer registers live: the all-paths instance keeps
// coder 1, v0 and family B in 0x7F, since the coder-2 paths' registers cost a batch of
// family-B stripes, which needs every resident wave, 670 -> 1109 us)
constexpr uint32_t kQuad = bit(kPathQ32) | bit(kPathQ64) | bit(kPathQ128);
constexpr uint32_t kLanes = bit(kPathL32) | bit(kPathL64);
constexpr uint32_t kMasks[] = {bit(kPathV0), bit(kPathQ32), bit(kPathQ64), bit(kPathQ128),
                               bit(kPathQ32) | bit(kPathQ64), kQuad,
                               bit(kPathBRows) | bit(kPathBStripes), bit(kPathQ32M),
                               kQuad | bit(kPathQ32M), bit(kPathL32), bit(kPathL64), kLanes,
                               bit(kPathL32M), kLanes | bit(kPathL32M), 0x7Fu, 0x3FFu};
constexpr uint32_t kNumMasks = sizeof(kMasks) / sizeof(kMasks[0]);

// work-queue shards: block b draws from shard b % kShards, whose tiles are t = g, g + 8,
// g + 16, ...; each shard's counter sits on its own 128-byte line
constexpr uint32_t kShards = 8;
constexpr uint32_t kCounterStride = 32;

template <int NV, uint32_t MASK>
__global__ __launch_bounds__(256) void batch_fwd(const Item *__restrict__ items,
                                                 const Tile *__restrict__ tiles, uint32_t n_tiles,
                                                 uint32_t *counters, uint32_t queue) {
    __shared__ BatchSmem<NV, MASK> sm;
    __shared__ uint32_t s_pos;
    const uint32_t tid = threadIdx.x;
    // shard g holds nt tiles and nb blocks; a block's first tile is static (position
    // blockIdx / kShards in its shard), later ones come from the shard's counter as
    // position nb + ticket.  Every block that enters the loop draws tickets until one
    // is past the end, so the shard hands out exactly max(nt - nb, 0) + min(nb, nt)
    // tickets, and the block that draws the last one resets the counter for the next
    // launch: no block touches it after that.
    const uint32_t g = blockIdx.x % kShards;
    const uint32_t nb = (gridDim.x - g + kShards - 1) / kShards;
    const uint32_t nt = n_tiles > g ? (n_tiles - g + kShards - 1) / kShards : 0u;
    const uint32_t last_ticket = (nt > nb ? nt - nb : 0u) + (nb < nt ? nb : nt) - 1u;
    uint32_t *counter = counters + g * kCounterStride;
    uint32_t pos = blockIdx.x / kShards;
    while (pos < nt) {
        const Tile tl = tiles[g + kShards * pos];
        const Item &it = items[tl.item];
        const uint32_t path = __builtin_amdgcn_readfirstlane(it.path);
        for (uint32_t u = 0; u < tl.count; u++) {
            const uint32_t bid = tl.first + u;
            if constexpr ((MASK & bit(kPathV0)) != 0)
                if (path == kPathV0) blk::v1_block<NV, 32, 8, 0>(it.v1, bid, sm.v0);
            if constexpr ((MASK & bit(kPathQ32)) != 0)
                if (path == kPathQ32) blk::quad_block<NV, 32>(it.v1, bid, sm.q32);
            if constexpr ((MASK & bit(kPathQ64)) != 0)
                if (path == kPathQ64) blk::quad_block<NV, 64>(it.v1, bid, sm.q64);
            if constexpr ((MASK & bit(kPathQ128)) != 0)
                if (path == kPathQ128) blk::quad_block<NV, 128>(it.v1, bid, sm.q128);
            if constexpr ((MASK & bit(kPathQ32M)) != 0)
                if (path == kPathQ32M) blk::quad_block_mfma<NV, 4>(it.v1, bid, sm.q32m);
            if constexpr ((MASK & bit(kPathL32)) != 0)
                if (path == kPathL32) blk::quad_block<NV, 32, 0, 4, 1>(it.v1, bid, sm.l32);
            if constexpr ((MASK & bit(kPathL64)) != 0)
                if (path == kPathL64) blk::quad_block<NV, 64, 0, 4, 1>(it.v1, bid, sm.l64);
            if constexpr ((MASK & bit(kPathL32M)) != 0)
                if (path == kPathL32M) blk::quad_block_mfma<NV, 4, 1>(it.v1, bid, sm.l32m);
            if constexpr ((MASK & bit(kPathBRows)) != 0)
                if (path == kPathBRows) blk::b_block<NV, 0>(it.b, bid, sm.b);
            if constexpr ((MASK & bit(kPathBStripes)) != 0)
                if (path == kPathBStripes) blk::b_block<NV, 1>(it.b, bid, sm.b);
            __syncthreads();   // the next unit rebuilds the tables and rings
        }
        if (queue) {
            // every thread read s_pos before the barrier that ended the last unit
            if (tid == 0) {
                const uint32_t k = atomicAdd(counter, 1u);
                if (k == last_ticket) atomicExch(counter, 0u);
                s_pos = nb + k;
            }
            __syncthreads();
            pos = s_pos;
        } else {
            pos += nb;
        }
    }
}

// A / B decode tables of a .bin v0 item from its cdf and ppf (the fused entries of the
// one-matrix kernels, v0_stream.h).
__global__ void v0_tables(uint32_t *out, const uint8_t *cdf, const uint8_t *ppf, int min_value,
                          uint32_t G) {
    uint32_t q = threadIdx.x, f, sym, bias;
    v0_table_entry(q, min_value, G, cdf, ppf, f, sym, bias);
    out[q] = f | (sym << 24);
    out[256 + q] = bias;
}

// ---- backward: dx = alpha * W^T (s_w * g) for every matrix of a batch (blk_bwd.h)

// v0, quad with 32 / 64 streams, and the quad streams back to back (coder 2), 32 / 64
enum : uint32_t { kBwdV0 = 0, kBwdQ32 = 1, kBwdQ64 = 2, kBwdL32 = 3, kBwdL64 = 4 };
constexpr uint32_t kBwdMasks[] = {bit(kBwdV0), bit(kBwdQ32), bit(kBwdQ64),
                                  bit(kBwdQ32) | bit(kBwdQ64), bit(kBwdL32), bit(kBwdL64),
                                  bit(kBwdL32) | bit(kBwdL64), 0x1Fu};
constexpr uint32_t kNumBwdMasks = sizeof(kBwdMasks) / sizeof(kBwdMasks[0]);

struct BwdItem {
    uint32_t path;
    uint32_t units;
    uint32_t sgs;    // stream groups (units per row block)
    uint32_t nrb;    // row blocks of 64 rows
    blk::BwdJob job;
};

// reduce tile: 64 columns of one gradient vector of one item
struct RTile {
    uint32_t item, v, cb, pad;
};

// Waves (streams) per backward unit: 4.  Twice the units of the one-matrix default (8),
// half the LDS per block: finer work for the queue and more resident blocks (decoder
// block, q32: 91.1 -> 84.5 us at one vector, 141.3 -> 136.3 at eight).  Every wave
// computes the same whatever the unit width, so the results stay those of the
// one-matrix backward.
constexpr int kBatchBwdWaves = 4;

template <int NV, uint32_t MASK>
union BwdSmem {
    Opt<(MASK & bit(kBwdV0)) != 0, blk::V0BwdSmem<NV, kBatchBwdWaves>> v0;
    Opt<(MASK & bit(kBwdQ32)) != 0,
        blk::QuadBwdSmem<NV, 32, 0, blk::bwd_set_words<32, 0>(), kBatchBwdWaves>> q32;
    Opt<(MASK & bit(kBwdQ64)) != 0,
        blk::QuadBwdSmem<NV, 64, 0, blk::bwd_set_words<64, 0>(), kBatchBwdWaves>> q64;
    Opt<(MASK & bit(kBwdL32)) != 0,
        blk::QuadBwdSmem<NV, 32, 0, blk::bwd_set_words<32, 0>(), kBatchBwdWaves>> l32;
    Opt<(MASK & bit(kBwdL64)) != 0,
        blk::QuadBwdSmem<NV, 64, 0, blk::bwd_set_words<64, 0>(), kBatchBwdWaves>> l64;
};

// Same work queue as batch_fwd (sharded counters, static first tile, last-ticket reset);
// a unit is one (row block, stream group) of the one-matrix backward.
template <int NV, uint32_t MASK>
__global__ __launch_bounds__(64 * kBatchBwdWaves) void batch_bwd(const BwdItem *__restrict__ items,
                                                 const Tile *__restrict__ tiles, uint32_t n_tiles,
                                                 uint32_t *counters, uint32_t queue) {
    __shared__ BwdSmem<NV, MASK> sm;
    __shared__ uint32_t s_pos;
    const uint32_t tid = threadIdx.x;
    const uint32_t g = blockIdx.x % kShards;
    const uint32_t nb = (gridDim.x - g + kShards - 1) / kShards;
    const uint32_t nt = n_tiles > g ? (n_tiles - g + kShards - 1) / kShards : 0u;
    const uint32_t last_ticket = (nt > nb ? nt - nb : 0u) + (nb < nt ? nb : nt) - 1u;
    uint32_t *counter = counters + g * kCounterStride;
    uint32_t pos = blockIdx.x / kShards;
    while (pos < nt) {
        const Tile tl = tiles[g + kShards * pos];
        const BwdItem &it = items[tl.item];
        const uint32_t path = __builtin_amdgcn_readfirstlane(it.path);
        const uint32_t nrb = __builtin_amdgcn_readfirstlane(it.nrb);
        // a tile's units are consecutive row blocks of one stream group (units are
        // numbered stream-group-major): the first builds the table and synchronises, the
        // others reuse it with no barrier
        for (uint32_t u = 0; u < tl.count; u++) {
            const uint32_t unit = tl.first + u;
            const uint32_t sg = unit / nrb, rb = unit % nrb;
            const bool fresh = u == 0;
            if constexpr ((MASK & bit(kBwdV0)) != 0) {
                if (path == kBwdV0)
                    blk::v0_bwd_unit<NV, kBatchBwdWaves>(it.job, rb, sg, nrb, sm.v0, fresh);
            }
            if constexpr ((MASK & bit(kBwdQ32)) != 0) {
                if (path == kBwdQ32)
                    blk::quad_bwd_unit<NV, 32, 0, blk::bwd_set_words<32, 0>(), kBatchBwdWaves>(
                        it.job, rb, sg, nrb, sm.q32, fresh);
            }
            if constexpr ((MASK & bit(kBwdQ64)) != 0) {
                if (path == kBwdQ64)
                    blk::quad_bwd_unit<NV, 64, 0, blk::bwd_set_words<64, 0>(), kBatchBwdWaves>(
                        it.job, rb, sg, nrb, sm.q64, fresh);
            }
            if constexpr ((MASK & bit(kBwdL32)) != 0) {
                if (path == kBwdL32)
                    blk::quad_bwd_unit<NV, 32, 0, blk::bwd_set_words<32, 0>(), kBatchBwdWaves, 1>(
                        it.job, rb, sg, nrb, sm.l32, fresh);
            }
            if constexpr ((MASK & bit(kBwdL64)) != 0) {
                if (path == kBwdL64)
                    blk::quad_bwd_unit<NV, 64, 0, blk::bwd_set_words<64, 0>(), kBatchBwdWaves, 1>(
                        it.job, rb, sg, nrb, sm.l64, fresh);
            }
        }
        __syncthreads();   // the next tile rebuilds the table
        if (queue) {
            if (tid == 0) {
                const uint32_t k = atomicAdd(counter, 1u);
                if (k == last_ticket) atomicExch(counter, 0u);
                s_pos = nb + k;
            }
            __syncthreads();
            pos = s_pos;
        } else {
            pos += nb;
        }
    }
}

// Second launch of a batched backward: every (item, vector, 64-column block) reduce tile.
__global__ __launch_bounds__(1024) void batch_bwd_reduce(const BwdItem *__restrict__ items,
                                                         const RTile *__restrict__ rtiles) {
    __shared__ float part[blk::kRedPartsMax][64];
    const RTile rt = rtiles[blockIdx.x];
    const BwdItem &it = items[rt.item];
    blk::reduce_tile(it.job, it.nrb, rt.v, rt.cb, part);
}

}  // namespace
}  // namespace mv

struct matvans_batch {
    bool backward = false;
    std::vector<mv::Item> items;          // forward items (device pointers inside)
    std::vector<mv::BwdItem> bitems;      // backward items
    std::vector<uint32_t> units;          // units per item
    std::vector<uint32_t> chain;          // decode-chain length per unit, for ordering
    mv::Item *d_items = nullptr;
    mv::BwdItem *d_bitems = nullptr;
    mv::Tile *d_tiles = nullptr;
    mv::RTile *d_rtiles = nullptr;        // backward reduce tiles
    uint32_t *d_counters = nullptr;
    uint32_t *d_tables = nullptr;         // v0 decode tables, 512 dwords per v0 item
    float *d_partial = nullptr;           // backward 64-row partial sums, owned by the plan
    uint32_t n_tiles = 0, max_tiles = 0, n_rtiles = 0, nv = 1, n_cu = 0, mask = 0, kernel = 5;
    uint32_t red_parts = 1;               // backward reduce: threads per column (the items' most)
    uint32_t max_bpc = 32;                // resident blocks per CU of the plan's kernel
    matvans_batch_config cfg{1, 0, 8, 0};
};

namespace {

hipError_t build_tiles(matvans_batch *p) {
    struct T { uint32_t item, first, count, chain; };
    std::vector<T> t;
    const uint32_t k = std::max(1u, p->cfg.units_per_tile);
    for (uint32_t i = 0; i < p->units.size(); i++)
        for (uint32_t u = 0; u < p->units[i]; u += k)
            t.push_back({i, u, std::min(k, p->units[i] - u), p->chain[i]});
    if (p->cfg.order == 1) {            // round-robin over matrices
        std::stable_sort(t.begin(), t.end(), [&](const T &a, const T &b) {
            uint32_t ra = a.first / k, rb = b.first / k;
            return ra != rb ? ra < rb : a.item < b.item;
        });
    } else if (p->cfg.order == 2) {     // longest decode chain first
        std::stable_sort(t.begin(), t.end(),
                         [](const T &a, const T &b) { return a.chain * a.count > b.chain * b.count; });
    }
    std::vector<mv::Tile> h(t.size());
    for (size_t i = 0; i < t.size(); i++) h[i] = {t[i].item, t[i].first, t[i].count, 0};
    p->n_tiles = uint32_t(h.size());
    return hipMemcpy(p->d_tiles, h.data(), h.size() * sizeof(mv::Tile), hipMemcpyHostToDevice);
}

int check_item(const matvans_batch_item &it, uint32_t k) {
    const bool b = it.format == MATVANS_BATCH_B_ROWS || it.format == MATVANS_BATCH_B_STRIPES;
    if (!it.vecs || (!b && !it.dst && !it.out8))
        return fail(MATVANS_E_INVALID, "item %u: null vecs, or neither dst nor out8", k);
    if (it.n_vec < 1 || it.n_vec > 8) return fail(MATVANS_E_INVALID, "item %u: n_vec %u", k, it.n_vec);
    if (!it.rows || !it.cols) return fail(MATVANS_E_SHAPE, "item %u: empty matrix", k);
    uint32_t stride = it.n_vec == 1 && it.vec_stride < it.cols ? it.cols : it.vec_stride;
    if (stride < it.cols) return fail(MATVANS_E_INVALID, "item %u: vec_stride < cols", k);
    uint32_t align = b ? 4 : 16;
    if (reinterpret_cast<uintptr_t>(it.vecs) % align || (it.n_vec > 1 && stride % align))
        return fail(MATVANS_E_ALIGN, "item %u: vectors must be %u-byte aligned", k, align);
    switch (it.format) {
        case MATVANS_BATCH_V0:
            if (it.cols % 512) return fail(MATVANS_E_SHAPE, "item %u: cols %% 512", k);
            if (!it.cursors || !it.cdf || !it.ppf || !it.payload)
                return fail(MATVANS_E_INVALID, "item %u: null v0 arrays", k);
            break;
        case MATVANS_BATCH_QUAD:
        case MATVANS_BATCH_QUAD_LANES:
            if (it.format == MATVANS_BATCH_QUAD_LANES && it.streams != 32 && it.streams != 64)
                return fail(MATVANS_E_UNSUPPORTED,
                            "item %u: coder-2 quad items need 32 or 64 streams", k);
            if (it.streams != 32 && it.streams != 64 && it.streams != 128)
                return fail(MATVANS_E_UNSUPPORTED,
                            "item %u: quad items need 32, 64 or 128 streams", k);
            if (it.cols % (16 * it.streams)) return fail(MATVANS_E_SHAPE, "item %u: cols", k);
            if (!it.cursors || !it.lut || !it.payload)
                return fail(MATVANS_E_INVALID, "item %u: null quad arrays", k);
            if (it.format == MATVANS_BATCH_QUAD_LANES &&
                !mv::lanes_tables_fit(it.rows, it.streams, it.payload_size))
                return fail(MATVANS_E_SHAPE, "item %u: payload shorter than the row tables", k);
            if (!it.out_start != !it.out_rec)
                return fail(MATVANS_E_INVALID, "item %u: outlier offsets and records go together",
                            k);
            break;
        case MATVANS_BATCH_B_ROWS:
        case MATVANS_BATCH_B_STRIPES:
            if (it.format == MATVANS_BATCH_B_ROWS ? it.cols % 128 : (it.rows % 32 || it.cols % 4))
                return fail(MATVANS_E_SHAPE, "item %u: shape does not fit the layout", k);
            if (!it.compressed || !it.warp_offsets || !it.b_ppf || !it.b_cum || !it.b_prob)
                return fail(MATVANS_E_INVALID, "item %u: null family-B arrays", k);
            if (!it.acc32 && !it.out8) return fail(MATVANS_E_INVALID, "item %u: no output", k);
            break;
        default:
            return fail(MATVANS_E_INVALID, "item %u: format %u", k, it.format);
    }
    if ((it.payload && reinterpret_cast<uintptr_t>(it.payload) % 16) ||
        (it.compressed && reinterpret_cast<uintptr_t>(it.compressed) % 16))
        return fail(MATVANS_E_ALIGN, "item %u: compressed data not 16-byte aligned", k);
    return MATVANS_OK;
}

int hip_fail(hipError_t e, const char *what) {
    return e == hipSuccess ? MATVANS_OK : fail(MATVANS_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

}  // namespace

namespace {

using FwdKernel = void (*)(const mv::Item *, const mv::Tile *, uint32_t, uint32_t *, uint32_t);
using BwdKernel = void (*)(const mv::BwdItem *, const mv::Tile *, uint32_t, uint32_t *, uint32_t);

int nv_row(uint32_t nv) { return nv == 1 ? 0 : nv == 2 ? 1 : nv == 4 ? 2 : 3; }

FwdKernel fwd_kernel(const matvans_batch *p) {
#define MV_ROW(NV) {mv::batch_fwd<NV, mv::kMasks[0]>, mv::batch_fwd<NV, mv::kMasks[1]>, \
                    mv::batch_fwd<NV, mv::kMasks[2]>, mv::batch_fwd<NV, mv::kMasks[3]>, \
                    mv::batch_fwd<NV, mv::kMasks[4]>, mv::batch_fwd<NV, mv::kMasks[5]>, \
                    mv::batch_fwd<NV, mv::kMasks[6]>, mv::batch_fwd<NV, mv::kMasks[7]>, \
                    mv::batch_fwd<NV, mv::kMasks[8]>, mv::batch_fwd<NV, mv::kMasks[9]>, \
                    mv::batch_fwd<NV, mv::kMasks[10]>, mv::batch_fwd<NV, mv::kMasks[11]>, \
                    mv::batch_fwd<NV, mv::kMasks[12]>, mv::batch_fwd<NV, mv::kMasks[13]>, \
                    mv::batch_fwd<NV, mv::kMasks[14]>, mv::batch_fwd<NV, mv::kMasks[15]>}
    static_assert(mv::kNumMasks == 16, "kernel table");
    static const FwdKernel table[4][16] = {MV_ROW(1), MV_ROW(2), MV_ROW(4), MV_ROW(8)};
#undef MV_ROW
    return table[nv_row(p->nv)][p->kernel];
}

BwdKernel bwd_kernel(const matvans_batch *p) {
#define MV_ROW(NV) {mv::batch_bwd<NV, mv::kBwdMasks[0]>, mv::batch_bwd<NV, mv::kBwdMasks[1]>, \
                    mv::batch_bwd<NV, mv::kBwdMasks[2]>, mv::batch_bwd<NV, mv::kBwdMasks[3]>, \
                    mv::batch_bwd<NV, mv::kBwdMasks[4]>, mv::batch_bwd<NV, mv::kBwdMasks[5]>, \
                    mv::batch_bwd<NV, mv::kBwdMasks[6]>, mv::batch_bwd<NV, mv::kBwdMasks[7]>}
    static_assert(mv::kNumBwdMasks == 8, "kernel table");
    static const BwdKernel table[4][8] = {MV_ROW(1), MV_ROW(2), MV_ROW(4), MV_ROW(8)};
#undef MV_ROW
    return table[nv_row(p->nv)][p->kernel];
}

// Blocks of the plan's kernel that fit on one CU at once: a persistent grid larger than
// that would leave blocks (and their fixed first tiles) waiting for others to finish.
hipError_t set_max_bpc(matvans_batch *p) {
    int nb = 0;
    hipError_t e = p->backward
        ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
              &nb, reinterpret_cast<const void *>(bwd_kernel(p)), 64 * mv::kBatchBwdWaves, 0)
        : hipOccupancyMaxActiveBlocksPerMultiprocessor(
              &nb, reinterpret_cast<const void *>(fwd_kernel(p)), 256, 0);
    p->max_bpc = e == hipSuccess && nb > 0 ? uint32_t(nb) : 1u;
    return e;
}

hipError_t device_cus(uint32_t &n_cu) {
    int dev = 0;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
    n_cu = e == hipSuccess ? uint32_t(prop.multiProcessorCount) : 0u;
    return e;
}

}  // namespace

extern "C" int matvans_batch_create(const matvans_batch_item *items, uint32_t n,
                                    matvans_batch **out) {
    if (!items || !n || !out) return fail(MATVANS_E_INVALID, "null argument or empty batch");
    *out = nullptr;
    uint32_t n_v0 = 0, nv = 1;
    for (uint32_t k = 0; k < n; k++) {
        if (int st = check_item(items[k], k)) return st;
        n_v0 += items[k].format == MATVANS_BATCH_V0;
        nv = std::max(nv, items[k].n_vec);
    }
    auto *p = new matvans_batch();
    p->nv = nv <= 1 ? 1 : nv <= 2 ? 2 : nv <= 4 ? 4 : 8;
    hipError_t e = device_cus(p->n_cu);
    if (e != hipSuccess) { delete p; return hip_fail(e, "device query"); }
    if (n_v0) e = hipMalloc(&p->d_tables, size_t(n_v0) * 512 * sizeof(uint32_t));
    uint32_t v0_i = 0, max_tiles = 0;
    for (uint32_t k = 0; e == hipSuccess && k < n; k++) {
        const auto &it = items[k];
        mv::Item m{};
        uint32_t stride = it.n_vec == 1 && it.vec_stride < it.cols ? it.cols : it.vec_stride;
        const bool quad = it.format == MATVANS_BATCH_QUAD || it.format == MATVANS_BATCH_QUAD_LANES;
        if (it.format == MATVANS_BATCH_V0 || quad) {
            const uint32_t *lut = it.lut;
            if (it.format == MATVANS_BATCH_V0) {
                uint32_t *tab = p->d_tables + size_t(v0_i++) * 512;
                hipLaunchKernelGGL(mv::v0_tables, dim3(1), dim3(256), 0, 0, tab, it.cdf, it.ppf,
                                   int(it.min_value), it.G);
                e = hipGetLastError();
                lut = tab;
                m.path = mv::kPathV0;
                m.units = (it.rows + 7) / 8;
                p->chain.push_back(it.cols / 32);
            } else if (it.format == MATVANS_BATCH_QUAD_LANES) {
                m.path = it.streams == 32 ? (it.n_vec >= mv::kBatchMfmaMinVec ? mv::kPathL32M
                                                                               : mv::kPathL32)
                                          : mv::kPathL64;
                m.units = (it.rows + 256u / it.streams - 1) / (256u / it.streams);
                p->chain.push_back(it.cols / it.streams / 4);
            } else {
                m.path = it.streams == 32 ? (it.n_vec >= mv::kBatchMfmaMinVec ? mv::kPathQ32M
                                                                               : mv::kPathQ32)
                         : it.streams == 64 ? mv::kPathQ64 : mv::kPathQ128;
                const uint32_t rpb = 256u / it.streams;   // rows per 4-wave block: 8, 4, 2
                m.units = (it.rows + rpb - 1) / rpb;
                p->chain.push_back(it.cols / it.streams / 4);
            }
            m.v1 = mv::blk::V1Job{it.dst, it.vecs, stride, it.n_vec, it.rows, it.cols, it.cursors,
                                  it.min_value, it.G, lut, nullptr, it.payload, it.payload_size,
                                  it.out8, it.scale, quad ? it.out_start : nullptr,
                                  quad ? it.out_rec : nullptr};
        } else {
            const bool rows = it.format == MATVANS_BATCH_B_ROWS;
            m.path = rows ? mv::kPathBRows : mv::kPathBStripes;
            uint32_t warps = rows ? it.rows : it.rows / 32;
            m.units = (warps + 7) / 8;
            p->chain.push_back(rows ? it.cols / 128 : it.cols / 4);
            m.b = mv::blk::BJob{it.acc32, it.out8, it.scale, it.compressed,
                                uint32_t(std::min<uint64_t>(it.compressed_words, 0xFFFFFFF0ull)),
                                it.warp_offsets, it.end_offset, it.b_ppf, it.b_cum, it.b_prob,
                                it.vecs, stride, it.n_vec, it.cols, it.rows};
        }
        max_tiles += m.units;
        p->mask |= 1u << m.path;
        p->units.push_back(m.units);
        p->items.push_back(m);
    }
    if (e == hipSuccess) e = hipMalloc(&p->d_items, sizeof(mv::Item) * n);
    if (e == hipSuccess)
        e = hipMemcpy(p->d_items, p->items.data(), sizeof(mv::Item) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&p->d_tiles, sizeof(mv::Tile) * max_tiles);
    const size_t counter_bytes = mv::kShards * mv::kCounterStride * sizeof(uint32_t);
    if (e == hipSuccess) e = hipMalloc(&p->d_counters, counter_bytes);
    if (e == hipSuccess) e = hipMemset(p->d_counters, 0, counter_bytes);
    p->max_tiles = max_tiles;
    for (uint32_t i = 0; i < sizeof(mv::kMasks) / sizeof(mv::kMasks[0]); i++)
        if ((mv::kMasks[i] & p->mask) == p->mask) { p->kernel = i; break; }
    if (e == hipSuccess) e = set_max_bpc(p);
    if (e == hipSuccess) e = build_tiles(p);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        matvans_batch_destroy(p);
        return hip_fail(e, "batch plan");
    }
    *out = p;
    mv::clear_error();
    return MATVANS_OK;
}

namespace {

int check_bwd_item(const matvans_batch_bwd_item &it, uint32_t k) {
    if (!it.g || !it.s_w || !it.alpha || !it.dx || !it.cursors || !it.payload)
        return fail(MATVANS_E_INVALID, "item %u: null device pointer", k);
    if (it.n_vec < 1 || it.n_vec > 8) return fail(MATVANS_E_INVALID, "item %u: n_vec %u", k, it.n_vec);
    if (!it.rows || !it.cols) return fail(MATVANS_E_SHAPE, "item %u: empty matrix", k);
    if (!mv::bwd_partials_fit(it.rows, it.cols))
        return fail(MATVANS_E_SHAPE, "item %u: %u x %u: partial sums beyond 4 GiB", k, it.rows,
                    it.cols);
    if (it.n_vec > 1 && it.g_stride < it.rows)
        return fail(MATVANS_E_INVALID, "item %u: g_stride < rows", k);
    if (reinterpret_cast<uintptr_t>(it.payload) % 16)
        return fail(MATVANS_E_ALIGN, "item %u: payload not 16-byte aligned", k);
    if (it.format == MATVANS_BATCH_V0) {
        if (it.cols % 512) return fail(MATVANS_E_SHAPE, "item %u: cols %% 512", k);
        if (!it.cdf || !it.ppf) return fail(MATVANS_E_INVALID, "item %u: null v0 tables", k);
    } else if (it.format == MATVANS_BATCH_QUAD || it.format == MATVANS_BATCH_QUAD_LANES) {
        if (it.streams != 32 && it.streams != 64)
            return fail(MATVANS_E_UNSUPPORTED, "item %u: quad items need 32 or 64 streams", k);
        if (it.cols % (16 * it.streams)) return fail(MATVANS_E_SHAPE, "item %u: cols", k);
        if (!it.lut) return fail(MATVANS_E_INVALID, "item %u: null decode table", k);
        if (it.format == MATVANS_BATCH_QUAD_LANES &&
            !mv::lanes_tables_fit(it.rows, it.streams, it.payload_size))
            return fail(MATVANS_E_SHAPE, "item %u: payload shorter than the row tables", k);
    } else {
        return fail(MATVANS_E_UNSUPPORTED, "item %u: the backward takes .bin v0 and quad items", k);
    }
    return MATVANS_OK;
}

}  // namespace

extern "C" int matvans_batch_create_backward(const matvans_batch_bwd_item *items, uint32_t n,
                                             matvans_batch **out) {
    if (!items || !n || !out) return fail(MATVANS_E_INVALID, "null argument or empty batch");
    *out = nullptr;
    uint32_t nv = 1;
    size_t partial_floats = 0;
    for (uint32_t k = 0; k < n; k++) {
        if (int st = check_bwd_item(items[k], k)) return st;
        nv = std::max(nv, items[k].n_vec);
        const size_t nrb = (items[k].rows + 63) / 64;
        partial_floats += (nrb * items[k].cols * items[k].n_vec + 63) / 64 * 64;
    }
    auto *p = new matvans_batch();
    p->backward = true;
    p->nv = nv <= 1 ? 1 : nv <= 2 ? 2 : nv <= 4 ? 4 : 8;
    hipError_t e = device_cus(p->n_cu);
    if (e == hipSuccess) e = hipMalloc(&p->d_partial, partial_floats * sizeof(float));
    std::vector<mv::RTile> rt;
    size_t off = 0;
    uint32_t max_tiles = 0;
    for (uint32_t k = 0; e == hipSuccess && k < n; k++) {
        const auto &it = items[k];
        mv::BwdItem b{};
        const bool v0 = it.format == MATVANS_BATCH_V0;
        const bool lanes = it.format == MATVANS_BATCH_QUAD_LANES;
        b.path = v0 ? mv::kBwdV0
                    : it.streams == 32 ? (lanes ? mv::kBwdL32 : mv::kBwdQ32)
                                       : (lanes ? mv::kBwdL64 : mv::kBwdQ64);
        b.sgs = (v0 ? 32u : it.streams) / uint32_t(mv::kBatchBwdWaves);
        b.nrb = (it.rows + 63) / 64;
        b.units = b.nrb * b.sgs;
        b.job = mv::blk::BwdJob{p->d_partial + off, it.dx, it.g, it.s_w, it.alpha,
                                it.n_vec > 1 ? it.g_stride : it.rows, it.n_vec, it.rows, it.cols,
                                it.cursors, it.payload, it.payload_size, it.min_value, it.G,
                                it.cdf, it.ppf, it.lut, nullptr,
                                v0 ? nullptr : it.out_col_start, v0 ? nullptr : it.out_col_rec};
        off += (size_t(b.nrb) * it.cols * it.n_vec + 63) / 64 * 64;
        for (uint32_t v = 0; v < it.n_vec; v++)
            for (uint32_t cb = 0; cb < (it.cols + 63) / 64; cb++) rt.push_back({k, v, cb, 0});
        p->red_parts = std::max(p->red_parts, mv::blk::red_parts(b.nrb, it.n_vec));
        p->chain.push_back(v0 ? it.cols / 32 : it.cols / it.streams / 4);
        p->units.push_back(b.units);
        max_tiles += b.units;
        p->mask |= 1u << b.path;
        p->bitems.push_back(b);
    }
    if (e == hipSuccess) e = hipMalloc(&p->d_bitems, sizeof(mv::BwdItem) * n);
    if (e == hipSuccess)
        e = hipMemcpy(p->d_bitems, p->bitems.data(), sizeof(mv::BwdItem) * n, hipMemcpyHostToDevice);
    p->n_rtiles = uint32_t(rt.size());
    if (e == hipSuccess) e = hipMalloc(&p->d_rtiles, sizeof(mv::RTile) * rt.size());
    if (e == hipSuccess)
        e = hipMemcpy(p->d_rtiles, rt.data(), sizeof(mv::RTile) * rt.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&p->d_tiles, sizeof(mv::Tile) * max_tiles);
    const size_t counter_bytes = mv::kShards * mv::kCounterStride * sizeof(uint32_t);
    if (e == hipSuccess) e = hipMalloc(&p->d_counters, counter_bytes);
    if (e == hipSuccess) e = hipMemset(p->d_counters, 0, counter_bytes);
    p->max_tiles = max_tiles;
    p->kernel = mv::kNumBwdMasks - 1;
    for (uint32_t i = 0; i < mv::kNumBwdMasks; i++)
        if ((mv::kBwdMasks[i] & p->mask) == p->mask) { p->kernel = i; break; }
    p->cfg.blocks_per_cu = 4;
    if (e == hipSuccess) e = set_max_bpc(p);
    if (e == hipSuccess) e = build_tiles(p);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        matvans_batch_destroy(p);
        return hip_fail(e, "backward batch plan");
    }
    *out = p;
    mv::clear_error();
    return MATVANS_OK;
}

extern "C" void matvans_batch_destroy(matvans_batch *p) {
    if (!p) return;
    // release what the plan owns; a failure here (a device already in error) has no
    // caller to report to
    for (void *d : {static_cast<void *>(p->d_items), static_cast<void *>(p->d_bitems),
                    static_cast<void *>(p->d_tiles), static_cast<void *>(p->d_rtiles),
                    static_cast<void *>(p->d_counters), static_cast<void *>(p->d_tables),
                    static_cast<void *>(p->d_partial)})
        (void)hipFree(d);
    delete p;
}

extern "C" int matvans_batch_configure(matvans_batch *p, const matvans_batch_config *cfg) {
    if (!p || !cfg) return fail(MATVANS_E_INVALID, "null argument");
    if (cfg->units_per_tile < 1 || cfg->units_per_tile > 64 || cfg->order > 2 ||
        cfg->blocks_per_cu < 1 || cfg->blocks_per_cu > 32 || cfg->schedule > 1)
        return fail(MATVANS_E_INVALID, "config out of range");
    p->cfg = *cfg;
    return hip_fail(build_tiles(p), "tile table");
}

extern "C" int matvans_batch_get_config(const matvans_batch *p, matvans_batch_config *cfg) {
    if (!p || !cfg) return fail(MATVANS_E_INVALID, "null argument");
    *cfg = p->cfg;
    return MATVANS_OK;
}

extern "C" int matvans_batch_launch(const matvans_batch *p, void *stream) {
    if (!p) return fail(MATVANS_E_INVALID, "null plan");
    // every shard that holds a tile needs a block: at least min(n_tiles, kShards) blocks;
    // no more blocks per CU than fit at once
    const uint32_t bpc = std::min(p->cfg.blocks_per_cu, p->max_bpc);
    const uint32_t grid = std::max(1u, std::min(p->n_tiles, std::max(mv::kShards, p->n_cu * bpc)));
    const uint32_t queue = p->cfg.schedule == 0 ? 1u : 0u;
    auto s = reinterpret_cast<hipStream_t>(stream);
    if (p->backward) {
        hipLaunchKernelGGL(bwd_kernel(p), dim3(grid), dim3(64 * mv::kBatchBwdWaves), 0, s,
                           p->d_bitems, p->d_tiles,
                           p->n_tiles, p->d_counters, queue);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "backward batch launch");
        hipLaunchKernelGGL(mv::batch_bwd_reduce, dim3(p->n_rtiles), dim3(64 * p->red_parts),
                           0, s, p->d_bitems, p->d_rtiles);
        return hip_fail(hipGetLastError(), "backward reduce launch");
    }
    hipLaunchKernelGGL(fwd_kernel(p), dim3(grid), dim3(256), 0, s, p->d_items, p->d_tiles,
                       p->n_tiles, p->d_counters, queue);
    return hip_fail(hipGetLastError(), "batch launch");
}

// Times every configuration of a small grid (schedule x units per tile x tile order x
// blocks per CU) over `reps` launches each on `stream`, twice, and keeps the
// configuration with the fastest of its two timings.  Synchronises; not for use inside
// graph capture.  No reference counterpart (the closest is the hand-set
// warpsPerBlock of matvansWarpPerRow.cu:156).
extern "C" int matvans_batch_autotune(matvans_batch *p, void *stream, uint32_t reps,
                                      matvans_batch_config *best_out, float *best_ms) {
    if (!p) return fail(MATVANS_E_INVALID, "null plan");
    reps = std::max(1u, reps);
    auto s = reinterpret_cast<hipStream_t>(stream);
    hipEvent_t a, b;
    hipError_t e = hipEventCreate(&a);
    if (e != hipSuccess) return hip_fail(e, "event");
    if ((e = hipEventCreate(&b)) != hipSuccess) {
        (void)hipEventDestroy(a);
        return hip_fail(e, "event");
    }
    matvans_batch_config best = p->cfg;
    float best_t = 1e30f;
    std::vector<matvans_batch_config> grid;
    for (uint32_t q : {0u, 1u})
        for (uint32_t u : {1u, 2u, 4u})
            for (uint32_t o : {0u, 1u, 2u})
                for (uint32_t c : {4u, 8u}) grid.push_back({u, o, c, q});
    int st = MATVANS_OK;
    // two passes over the grid, the second in reverse order, keeping each configuration's
    // faster pass: a transient slow measurement (clock ramp, a neighbour's burst) then
    // does not decide the choice, and drift along the grid cancels
    std::vector<float> t(grid.size(), 1e30f);
    for (int pass = 0; pass < 2; pass++) {
        for (size_t n = 0; n < grid.size(); n++) {
            const size_t k = pass == 0 ? n : grid.size() - 1 - n;
            const matvans_batch_config &cfg = grid[k];
            if ((st = matvans_batch_configure(p, &cfg))) goto done;
            if ((st = matvans_batch_launch(p, stream))) goto done;   // warm-up
            if ((e = hipEventRecord(a, s)) != hipSuccess) {
                st = hip_fail(e, "event record");
                goto done;
            }
            for (uint32_t r = 0; r < reps && !st; r++) st = matvans_batch_launch(p, stream);
            if (st) goto done;
            if ((e = hipEventRecord(b, s)) != hipSuccess || (e = hipEventSynchronize(b)) != hipSuccess) {
                st = hip_fail(e, "event sync");
                goto done;
            }
            float ms = 0;
            if ((e = hipEventElapsedTime(&ms, a, b)) != hipSuccess) {
                st = hip_fail(e, "event time");
                goto done;
            }
            t[k] = std::min(t[k], ms / float(reps));
        }
    }
    for (size_t k = 0; k < grid.size(); k++)
        if (t[k] < best_t) {
            best_t = t[k];
            best = grid[k];
        }
done:
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (st) return st;
    if ((st = matvans_batch_configure(p, &best))) return st;
    if (best_out) *best_out = best;
    if (best_ms) *best_ms = best_t;
    mv::clear_error();
    return MATVANS_OK;
}
