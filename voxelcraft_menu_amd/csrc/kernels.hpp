// kernels.hpp -- kernel declarations shared by the host translation units.
#pragma once

#include "march_core.hpp"

namespace vxd {

// blockIdx -> work block so that consecutive work blocks land on one XCD (one L2):
// the dispatcher hands block b to XCD b % 8; grids are padded to a multiple of 8.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
    uint32_t per = nb >> 3;
    return (b & 7u) * per + (b >> 3);
}

__device__ __forceinline__ uint64_t ballot(bool v) { return __ballot(v); }

// Counters are produced by earlier kernels of the same frame; read them with an
// agent-scope atomic load so the value comes from L2, never from a scalar-cache line.
__device__ __forceinline__ uint32_t ld_counter(const uint32_t *c, uint32_t i) {
    return __hip_atomic_load(c + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Control words (the reference's counters[0..5] plus ours) used by the pairs path.
constexpr uint32_t CNT_PAIRS = 0, CNT_DEFERRED = 1, CNT_RECOVER = 2, CNT_SELECT = 3, CNT_SELECT_RAW = 4,
                   CNT_RECOVER_RAW = 5, CNT_SHADOW = 6, CNT_TRUNC = 7;
// Statistics accumulated by every tile's wave go to one of 64 slots, each on its own
// 128-byte line, so tens of thousands of end-of-wave atomics do not serialise on one
// address; readback sums the slots.
constexpr uint32_t STAT_SELECT = 0, STAT_DEFERRED = 1, STAT_RECOVER = 2, STAT_SHADOW = 3;
constexpr uint32_t CNT_STAT_BASE = 32, CNT_STAT_SLOTS = 64, CNT_STAT_STRIDE = 32;
constexpr uint32_t CNT_LEN = CNT_STAT_BASE + CNT_STAT_SLOTS * CNT_STAT_STRIDE;
// Tiles are ordered by (class, cost bucket): a batch holds up to four view classes
// (Hi-Z on/off x camera under water or not), each marched by its own specialised kernel
// over a contiguous range of the queue; inside a class, 32 buckets of log2 cycles, most
// expensive first.
constexpr uint32_t COST_BUCKETS = 32, N_CLASSES = 4, ORDER_BINS = COST_BUCKETS * N_CLASSES;
__host__ __device__ constexpr uint32_t view_class(bool hiz, bool underwater) {
    return (underwater ? 2u : 0u) + (hiz ? 0u : 1u);
}
// Control block after the statistics: the bin histogram and per-bin cursors of the
// launch ordering, and one work-queue head per class for the persistent batch march
// (each head on its own 128-byte line). Cleared with the counters at every frame.
constexpr uint32_t CNT_CTL = CNT_LEN;
constexpr uint32_t CTL_HIST = 0, CTL_CURSOR = ORDER_BINS, CTL_HEAD = 2 * ORDER_BINS, CTL_HEAD_STRIDE = 32,
                   CTL_LEN = 2 * ORDER_BINS + N_CLASSES * CTL_HEAD_STRIDE;
// Secondary-leg queues (two: water legs, glass legs) and their overflow flag, each word
// on its own line, after the control block.
// SEC_SUB: the water legs spawned where a glass transmit leg ends on water.
constexpr uint32_t CNT_SEC = CNT_CTL + CTL_LEN, SEC_LINE = 32, SEC_WATER = 0, SEC_GLASS = 1, SEC_OVERFLOW = 2,
                   SEC_SUN = 3, SEC_SUB = 4, SEC_LEN = 5 * SEC_LINE;
// A pixel's leg records: [0] refract / transmit, [1] reflect, [2] [3] the refract and
// reflect legs from where its glass transmit leg ends on water.
constexpr uint32_t SEC_RECORDS = 4;
// Leg-queue entry: view << 32 | pixel << 2 | record.
__host__ __device__ constexpr unsigned long long sec_entry(uint32_t view, uint32_t pix, uint32_t rec) {
    return ((unsigned long long)view << 32) | ((unsigned long long)pix << 2) | rec;
}
constexpr uint32_t CNT_TOTAL = CNT_LEN + CTL_LEN + SEC_LEN;

__device__ __forceinline__ void stat_add(uint32_t *counters, uint32_t slot, uint32_t k, uint32_t v) {
    if (v) atomicAdd(&counters[CNT_STAT_BASE + (slot & (CNT_STAT_SLOTS - 1)) * CNT_STAT_STRIDE + k], v);
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n);
__global__ void k_fill_u64(uint64_t *p, uint64_t v, size_t n);
__global__ void k_tile_select(Pools p, FrameK f, const uint32_t *__restrict__ hiz, uint32_t *__restrict__ pairs,
                              uint32_t *__restrict__ counters);
__global__ void k_finalize(FrameK f, uint32_t *counters);
__global__ void k_march_pairs(Pools p, FrameK f, Spec s, const uint32_t *__restrict__ pairs,
                              const uint32_t *__restrict__ counters, unsigned long long *__restrict__ vis,
                              uint32_t *__restrict__ heat);
__global__ void k_build_hiz(FrameK f, const unsigned long long *__restrict__ vis, uint32_t *__restrict__ hiz);
__global__ void k_recover_select(Pools p, FrameK f, const uint32_t *__restrict__ hiz, uint32_t *__restrict__ pairs,
                                 uint32_t *__restrict__ counters);
__global__ void k_finalize_recover(FrameK f, uint32_t *counters);
// Tile-march structure bits (vx_tuning.variant).
//   VAR_SPLIT: classify the whole chunk list into LDS bitmasks before marching; otherwise
//              each 64-chunk block is classified and marched in turn.
//   VAR_FUSED_SHADOW: trace each pixel's sun ray in the tile's wave right after its key.
//   VAR_BINS (batched forms): classify only the chunks on the candidate list of the
//              tile's screen bin (k_bin_chunks) instead of the view's whole list.
constexpr int VAR_SPLIT = 1, VAR_FUSED_SHADOW = 4, VAR_BINS = 8;

// Arguments of the single-view tile march (one frame of one view).
struct TileArgs {
    Pools p;
    FrameK f;
    Spec s;
    uint32_t *hiz;
    unsigned long long *vis;
    uint16_t *shadow;
    uint32_t *cnt;
    const uint32_t *order;
    uint32_t *cost;
};

template <int kOcc, int kVar, bool kUseHiz, bool kUnderwater>
__global__ void k_march_tiles(TileArgs a);

// Launch ordering (class, then most expensive cost bucket first). Entries are tile
// indices, or with a view table (n_views > 0) packed view << ENTRY_VIEW_SHIFT |
// tile-in-view. The element count is *n_dev when n_dev is non-null, else n. cb, when
// non-null, holds the class boundaries cb[0..4] of the tile index space (views are laid
// out class by class); without it every tile is class 0.
constexpr uint32_t ENTRY_VIEW_SHIFT = 20, ENTRY_TILE_MASK = (1u << ENTRY_VIEW_SHIFT) - 1u;
constexpr uint32_t BATCH_MAX_VIEWS = 1024;
constexpr uint32_t ORDER_BLOCKS = 256;
__global__ void k_cost_hist(const uint32_t *__restrict__ cost, const uint32_t *__restrict__ n_dev, uint32_t n,
                            const uint32_t *__restrict__ cb, uint32_t *__restrict__ ctl);
__global__ void k_cost_scatter(const uint32_t *__restrict__ cost, const uint32_t *__restrict__ n_dev, uint32_t n,
                               const uint32_t *__restrict__ cb, uint32_t *__restrict__ ctl,
                               const uint32_t *__restrict__ tile_off, const uint32_t *__restrict__ n_views_dev,
                               uint32_t *__restrict__ order);
__global__ void k_iota_views(uint32_t *__restrict__ order, const uint32_t *__restrict__ n_dev, uint32_t n,
                             const uint32_t *__restrict__ tile_off, const uint32_t *__restrict__ n_views_dev);

// One view of a batch: its frame and spec, and where its inputs and outputs start in the
// batch's concatenated buffers.
struct ViewK {
    FrameK f;
    Spec s;
    uint32_t chunk_off;  // first GpuChunk record of the view's list
    uint32_t n_chunks;   // records in the view's list (visible first)
    uint32_t grid_off;   // first word of the view's 32x8x32 chunk grid
    uint32_t tile_off;   // first Hi-Z / cost word of the view
    uint32_t px_off_lo, px_off_hi;  // first pixel of the view in vis / shadow
    uint32_t bin_off;     // first screen bin of the view (bins of BIN_TILES x BIN_TILES tiles)
    uint32_t bin_stride;  // list entries per bin: the records reserved for the view's list
    uint32_t binl_off;    // first list entry of the view's bins
    // The tree pools with the view's chunk list and grid (chunks = batch list + chunk_off,
    // grid = batch grids + grid_off, n_chunks): kernels read it in place, field by field.
    // The host rewrites it when a pool moves or grows.
    Pools p;
};

// Per-view chunk lists on the device: partition_for_grid + build_chunk_list
// (src/render/mod.rs:110-138, :1447-1564) over a table of LOD-0 GpuChunk records (fade 1)
// in render-set order, for a camera given by its frustum planes, position and grid origin.
struct ListCam {
    float planes[24];  // Frustum::from_camera: (nx, ny, nz, d) x 6
    float pos[3];
    int32_t gmin[3];
    uint32_t chunk_off;  // first record of the view's list in the output
    uint32_t view;       // the view's slot: its grid, counts and descriptor
};
static_assert(sizeof(ListCam) == 128, "ListCam layout");
constexpr uint32_t LIST_BLOCK = 1024, LIST_KEYS = 8192;

__global__ void k_view_lists(const vx_gpu_chunk *__restrict__ table, uint32_t n_table, const ListCam *__restrict__ cams,
                             vx_gpu_chunk *__restrict__ out_chunks, uint32_t *__restrict__ grids,
                             uint32_t *__restrict__ counts, ViewK *__restrict__ patch);

// Header words of a batch in device memory (read by graph-captured kernels, so a replay
// follows the current view set): tile and view counts, class boundaries cb[0..4].
// BH_BINS: screen bins of all views.
constexpr uint32_t BH_TILES = 0, BH_VIEWS = 1, BH_BINS = 2, BH_CB = 4, BH_LEN = 12;

struct BatchArgs {
    Pools p;                  // tree pools; chunks/grid are the batch's concatenated arrays
    const ViewK *__restrict__ views;
    const uint32_t *__restrict__ hdr;
    uint32_t *__restrict__ hiz;
    unsigned long long *__restrict__ vis;
    uint16_t *__restrict__ shadow;
    uint32_t *__restrict__ counters;
    const uint32_t *__restrict__ order;
    uint32_t *__restrict__ cost;
    // Screen-bin candidate lists (k_bin_chunks): per bin a count and, at binl_off +
    // bin * bin_stride of its view, the chunk indices in list order. bin_cnt null: every
    // tile classifies its view's whole list.
    const uint32_t *__restrict__ bin_cnt = nullptr;
    const uint16_t *__restrict__ bin_list = nullptr;
};

// Screen bins: a bin is BIN_TILES x BIN_TILES tiles (64 x 64 pixels), in rows of
// ceil(tiles_x / BIN_TILES) bins.
constexpr uint32_t BIN_TILES = 8;
__host__ __device__ constexpr uint32_t bins_x(uint32_t tiles_x) { return (tiles_x + BIN_TILES - 1) / BIN_TILES; }
__host__ __device__ constexpr uint32_t bins_y(uint32_t tiles_y) { return (tiles_y + BIN_TILES - 1) / BIN_TILES; }

// k_bin_chunks: one wave per screen bin of the batch (a grid-stride loop over BH_BINS),
// writes the bin's candidate list: every chunk of the view's visible list except those
// that lie wholly in front of the camera and wholly outside one of the bin's four side
// planes, each moved outwards by a margin. Such a chunk lies at least the margin outside
// the matching side plane of every tile in the bin: no pixel ray of those tiles meets it,
// and tile_select's f32 plane test rejects it unless that test's own rounding is larger
// than the margin (tile edges spanning far less than a milliradian).
__global__ void k_bin_chunks(BatchArgs a, uint16_t *__restrict__ bin_list, uint32_t *__restrict__ bin_cnt);

// Batched tile march over the queue range of one view class (kHiz, kUnderwater; the
// range is the header's cb[c]..cb[c+1]): FLAT = one workgroup per entry; PERSISTENT =
// resident waves taking `chunk` entries at a time from the class's queue head.
template <int kOcc, int kVar, bool kHiz, bool kUnderwater>
__global__ void k_batch_flat(BatchArgs a);
template <int kOcc, int kVar, bool kHiz, bool kUnderwater>
__global__ void k_batch_persistent(BatchArgs a, uint32_t chunk);

// Launches the batched march over every non-empty class range cb[c]..cb[c+1] (host
// copies of the header's class boundaries).
bool batch_form_ok(uint32_t mode, uint32_t occ, uint32_t var);
void launch_batch_classes(uint32_t mode, uint32_t occ, uint32_t var, uint32_t chunk, uint32_t wps, int n_cu,
                          const uint32_t cb[N_CLASSES + 1], const BatchArgs &a, hipStream_t st);
__global__ void k_shadow(Pools p, FrameK f, Spec s, const unsigned long long *__restrict__ vis,
                         uint16_t *__restrict__ out, uint32_t *__restrict__ counters);
__global__ void k_decode(Pools p, FrameK f, const unsigned long long *__restrict__ vis,
                         const uint16_t *__restrict__ shadow, vx_hit *__restrict__ out);

// Secondary bounce of a batch (one leg record per pixel and leg, 16 bytes; a record is
// zero when the pixel has no such leg):
//   k_sec_clear: before a frame resets the counters, zeroes the records of the previous
//     frame's queue entries, so that only pixels with legs are ever written;
//   k_sec_gather: one wave per tile (queue entry of `order`), lane = pixel: appends the
//     pixel's legs to the water / glass queues (entries view << 32 | pixel << 1 | leg);
//   k_sec_trace<kWater>: one leg per lane from its queue, grouped by view, trace_world.
struct SecArgs {
    uint4 *__restrict__ out;                  // SEC_RECORDS records per pixel of the batch
    unsigned long long *__restrict__ queue;   // queue q (SEC_WATER, SEC_GLASS, then the sub-leg queue) at q * cap
    const uint32_t *__restrict__ tile_off;    // n_views + 1 tile offsets (batch slot order)
    uint32_t cap;                             // entries per queue
};
__global__ void k_sec_clear(BatchArgs a, SecArgs s);
__global__ void k_sec_gather(BatchArgs a, SecArgs s);
template <bool kWater>
__global__ void k_sec_trace(BatchArgs a, SecArgs s);
__global__ void k_sec_trace_sub(BatchArgs a, SecArgs s);
// Queue slot of the sub-leg queue in SecArgs::queue (after the water and glass queues).
constexpr uint32_t SEC_Q_SUB = 2;

}  // namespace vxd
