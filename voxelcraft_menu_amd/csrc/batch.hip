// batch.hip -- extern "C" entry points of view batches (vx_batch_*, include/vxmarch.h).
//
// A batch renders many views of one world in one submission. It shares the context's
// tree pools and stream; every view brings its own frame, chunk list and chunk grid
// (the reference's per-view build_chunk_list output) and keeps its own Hi-Z, visibility
// and shadow buffers, laid out back to back:
//   chunks  [sum n_total]      GpuChunk records, view after view
//   grids   [n_views * 8192]   u32 chunk grids
//   vis     [sum W*H]          u64 keys        shadow [sum W*H] u16 (R16F bits)
//   hiz/cost/order [sum tiles] u32
// Views are laid out class by class (Hi-Z on/off x camera under water or not; the slot
// of a view is its place in that layout), so each class is one contiguous tile range
// marched by a kernel specialised for it.
// Three launch forms give the same bits (each tile runs the same body):
//   PLAIN       per view: order kernels + the single-view tile march (one dispatch per
//               view, the pattern of the reference's per-view Renderer::render);
//   FLAT        one launch over every view's tiles, one workgroup per tile, in cost order;
//   PERSISTENT  one launch of resident waves pulling tiles from a global queue.
// With graph = 1 the submission is captured once into a hipGraph and replayed; the
// kernels read the tile and view counts from a device header, so a replay follows the
// current view set while no buffer is reallocated.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "ctx.hpp"

using namespace vxd;
using vxh::DevBuf;
using vxh::ensure;

#define HIPX VX_HIPX
#define LAUNCHED VX_LAUNCHED

struct vx_batch {
    vx_ctx *ctx = nullptr;
    uint32_t max_views = 0;
    uint32_t n_views = 0, n_tiles = 0;
    uint64_t n_px = 0;
    int n_cu = 0;
    std::vector<ViewK> vk;         // host copies of the view descriptors, by slot
    std::vector<uint32_t> toff;    // n_views + 1 tile offsets, by slot
    std::vector<uint64_t> poff;    // n_views + 1 pixel offsets, by slot
    std::vector<uint32_t> slot_of; // caller's view index -> slot
    uint32_t cb[N_CLASSES + 1] = {0, 0, 0, 0, 0};  // class boundaries of the tile range
    DevBuf views, tile_off, hdr, chunks, grids, vis, shadow, hiz, cost, order, counters;
    // Secondary bounce: SEC_RECORDS 16-byte leg records per pixel, three leg queues of
    // sec_cap entries (water legs, glass legs, water legs behind glass).
    bool sec_on = false;
    DevBuf sec_out, sec_queue;
    DevBuf table, cams, list_counts;  // device chunk lists: chunk table, per-view cameras, counts
    // Screen-bin candidate lists (k_bin_chunks) of the installed view set: bins_fit when the
    // set's lists fit bin_list_max_bytes(); built by the first render of a form that uses them.
    DevBuf bin_cnt, bin_list;
    uint32_t n_bins = 0;
    uint64_t bin_entries = 0;
    bool bins_fit = false, bins_built = false;
    uint32_t n_table = 0;
    uint32_t sec_cap = 0;
    uint64_t vk_pool_gen = 0;  // pool generation the device descriptors' pool fields were written for
    vx_batch_tuning tn{};
    // Graph of the current submission (captured lazily, dropped when it would go stale).
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t graph_tiles = 0;
    uint32_t graph_cb[N_CLASSES + 1] = {0, 0, 0, 0, 0};
    uint64_t graph_pool_gen = 0;
    bool graph_valid = false;
    bool graph_bins = false;  // the captured launches classify by the bin lists
    static constexpr int RING = 256;
    hipEvent_t ring[RING][4] = {};  // [0] call, [1] submission start, [2] march kernel start, [3] end
    uint32_t ring_next = 0, ring_count = 0;
    bool rendered = false;
};

namespace {

// Largest screen-bin list allocation; a view set that needs more keeps whole lists.
// VX_BIN_LIST_MAX_BYTES in the environment lowers it (read when a view set is installed).
constexpr uint64_t BIN_LIST_MAX_BYTES = 1ull << 30;
uint64_t bin_list_max_bytes() {
    if (const char *e = std::getenv("VX_BIN_LIST_MAX_BYTES")) {
        const unsigned long long v = std::strtoull(e, nullptr, 10);
        if (v < BIN_LIST_MAX_BYTES) return v;
    }
    return BIN_LIST_MAX_BYTES;
}

// Per-view control blocks of the PLAIN form follow the shared counters.
size_t counter_words(uint32_t n_views) { return (size_t)CNT_TOTAL + (size_t)n_views * CTL_LEN; }

bool batch_tuning_ok(const vx_batch_tuning &t) {
    if (t.mode > VX_BATCH_PLAIN || t.ordered > 1 || t.graph > 1) return false;
    if (t.mode == VX_BATCH_PLAIN) return vxh::tuning_ok(t.occupancy, t.variant);
    if (!batch_form_ok(t.mode, t.occupancy, t.variant)) return false;
    if (t.mode == VX_BATCH_PERSISTENT) {
        if (t.queue_chunk < 1 || t.queue_chunk > 64) return false;
        if (t.waves_per_simd < 1 || t.waves_per_simd > t.occupancy) return false;
    }
    return true;
}

void drop_graph(vx_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_valid = false;
}

// Pool fields of a view descriptor: the context's pools with the view's chunk list and grid.
void view_pools(const vx_batch *b, ViewK &k) {
    k.p = vxh::make_pools(b->ctx);
    k.p.chunks = (const vx_gpu_chunk *)b->chunks.p + k.chunk_off;
    k.p.grid = (const uint32_t *)b->grids.p + k.grid_off;
    k.p.n_chunks = k.n_chunks;
}

// Rewrites the descriptors on the device when a pool moved or grew since they were written.
int refresh_view_pools(vx_batch *b, hipStream_t st) {
    if (b->vk_pool_gen == b->ctx->pool_gen || b->n_views == 0) return VX_OK;
    for (ViewK &k : b->vk) view_pools(b, k);
    HIPX(hipMemcpyAsync(b->views.p, b->vk.data(), (size_t)b->n_views * sizeof(ViewK), hipMemcpyHostToDevice, st));
    HIPX(hipStreamSynchronize(st));
    b->vk_pool_gen = b->ctx->pool_gen;
    return VX_OK;
}

// The tuning asks for bin lists (a batched form, not PLAIN, with VAR_BINS).
bool wants_bins(const vx_batch *b) { return (b->tn.variant & VAR_BINS) && b->tn.mode != VX_BATCH_PLAIN; }
bool uses_bins(const vx_batch *b) { return wants_bins(b) && b->bins_fit && b->bins_built; }

BatchArgs batch_args(const vx_batch *b) {
    BatchArgs a;
    a.p = vxh::make_pools(b->ctx);
    a.p.chunks = (const vx_gpu_chunk *)b->chunks.p;
    a.p.grid = (const uint32_t *)b->grids.p;
    a.p.n_chunks = 0;  // per view, from its descriptor
    a.views = (const ViewK *)b->views.p;
    a.hdr = (const uint32_t *)b->hdr.p;
    a.hiz = (uint32_t *)b->hiz.p;
    a.vis = (unsigned long long *)b->vis.p;
    a.shadow = (uint16_t *)b->shadow.p;
    a.counters = (uint32_t *)b->counters.p;
    a.order = (const uint32_t *)b->order.p;
    a.cost = (uint32_t *)b->cost.p;
    if (uses_bins(b)) {
        a.bin_cnt = (const uint32_t *)b->bin_cnt.p;
        a.bin_list = (const uint16_t *)b->bin_list.p;
    }
    return a;
}

// The secondary bounce over the frame's keys: gather the legs into the water and glass
// queues, trace each, then the water legs the glass trace queued where its transmit legs
// end on water (leg counts are read on the device, so a replay follows the frame).
int enqueue_secondary(vx_batch *b, hipStream_t st) {
    BatchArgs a = batch_args(b);
    SecArgs sa;
    sa.out = (uint4 *)b->sec_out.p;
    sa.queue = (unsigned long long *)b->sec_queue.p;
    sa.tile_off = (const uint32_t *)b->tile_off.p;
    sa.cap = b->sec_cap;
    k_sec_gather<<<std::max(b->n_tiles, 1u), 64, 0, st>>>(a, sa);
    const unsigned nb = (unsigned)b->n_cu * 16u;
    k_sec_trace<true><<<nb, 64, 0, st>>>(a, sa);
    k_sec_trace<false><<<nb, 64, 0, st>>>(a, sa);
    k_sec_trace_sub<<<nb, 64, 0, st>>>(a, sa);
    LAUNCHED();
    return VX_OK;
}

// Sizes the secondary buffers for the current views (the water-look ripples are not
// carried into the legs: such views are refused).
int secondary_buffers(vx_batch *b) {
    for (const ViewK &k : b->vk)
        if (k.s.bits & SPEC_WATER_LOOK) return VX_E_UNSUPPORTED;
    if (b->n_px * 2 > 0xFFFFFFFFull) return VX_E_OVERFLOW;
    const uint32_t cap = (uint32_t)std::max<uint64_t>(b->n_px * 2, 64);
    const void *o0 = b->sec_out.p, *q0 = b->sec_queue.p;
    int rc;
    if ((rc = ensure(b->sec_out, (size_t)std::max<uint64_t>(b->n_px, 1) * SEC_RECORDS * 16)) ||
        (rc = ensure(b->sec_queue, (size_t)cap * 3 * 8)))
        return rc;
    if (o0 != b->sec_out.p || q0 != b->sec_queue.p || cap != b->sec_cap) drop_graph(b);
    b->sec_cap = cap;
    // zero records and leg counters: the next frame's clear pass has nothing to undo
    hipStream_t st = b->ctx->stream;
    if ((rc = vxh::fill_u32(st, b->sec_out.p, 0, std::max<uint64_t>(b->n_px, 1) * SEC_RECORDS * 4)) ||
        (rc = vxh::fill_u32(st, (uint32_t *)b->counters.p + CNT_SEC, 0, SEC_LEN)))
        return rc;
    HIPX(hipStreamSynchronize(st));
    return VX_OK;
}

// The whole submission of one batch frame, on the context stream (also what a graph
// captures). PLAIN reads every view's geometry from the host copies, so its graph is
// tied to the view set; the other forms read counts from the device header.
int enqueue(vx_batch *b, hipStream_t st, hipEvent_t ev_march) {
    int rc;
    uint32_t *cnt = (uint32_t *)b->counters.p;
    if (b->sec_on) {  // the previous frame's leg records, from its queues, before the counters reset
        SecArgs sa;
        sa.out = (uint4 *)b->sec_out.p;
        sa.queue = (unsigned long long *)b->sec_queue.p;
        sa.tile_off = (const uint32_t *)b->tile_off.p;
        sa.cap = b->sec_cap;
        k_sec_clear<<<(unsigned)b->n_cu * 4u, 256, 0, st>>>(batch_args(b), sa);
        LAUNCHED();
    }
    if ((rc = vxh::fill_u32(st, cnt, 0, counter_words(b->n_views)))) return rc;
    if (b->tn.mode == VX_BATCH_PLAIN) {
        vxh::TileArgs ta;
        ta.p = vxh::make_pools(b->ctx);
        ta.cnt = cnt;
        for (uint32_t v = 0; v < b->n_views; v++) {
            const ViewK &k = b->vk[v];
            const uint32_t nt = k.f.tiles[0] * k.f.tiles[1];
            const size_t px0 = b->poff[v];
            uint32_t *order = (uint32_t *)b->order.p + k.tile_off;
            uint32_t *cost = (uint32_t *)b->cost.p + k.tile_off;
            if (b->tn.ordered)
                vxh::launch_order(st, cost, nt, cnt + CNT_TOTAL + (size_t)v * CTL_LEN, order);
            else
                k_iota_views<<<(nt + 255) / 256, 256, 0, st>>>(order, nullptr, nt, nullptr, nullptr);
            ta.p.chunks = (const vx_gpu_chunk *)b->chunks.p + k.chunk_off;
            ta.p.grid = (const uint32_t *)b->grids.p + k.grid_off;
            ta.p.n_chunks = k.n_chunks;
            ta.f = k.f;
            ta.s = k.s;
            ta.hiz = (uint32_t *)b->hiz.p + k.tile_off;
            ta.vis = (unsigned long long *)b->vis.p + px0;
            ta.shadow = (uint16_t *)b->shadow.p + px0;
            ta.order = order;
            ta.cost = cost;
            vxh::launch_tiles(b->tn.occupancy, b->tn.variant, (k.f.flags & VX_FLAG_HIZ) != 0u,
                              (k.f.flags & VX_FLAG_UNDERWATER) != 0u, nt, st, ta);
            if (!(b->tn.variant & VAR_FUSED_SHADOW))
                k_shadow<<<((nt + 3) / 4 + 7u) & ~7u, 256, 0, st>>>(ta.p, k.f, k.s, ta.vis, ta.shadow, cnt);
        }
        LAUNCHED();
        if (b->sec_on) return enqueue_secondary(b, st);
        return VX_OK;
    }
    const uint32_t *hdr = (const uint32_t *)b->hdr.p;
    if (b->tn.ordered) {
        k_cost_hist<<<ORDER_BLOCKS, 256, 0, st>>>((const uint32_t *)b->cost.p, hdr + BH_TILES, 0, hdr + BH_CB,
                                                  cnt + CNT_CTL);
        k_cost_scatter<<<ORDER_BLOCKS, 256, 0, st>>>((const uint32_t *)b->cost.p, hdr + BH_TILES, 0, hdr + BH_CB,
                                                     cnt + CNT_CTL, (const uint32_t *)b->tile_off.p, hdr + BH_VIEWS,
                                                     (uint32_t *)b->order.p);
    } else {
        k_iota_views<<<ORDER_BLOCKS, 256, 0, st>>>((uint32_t *)b->order.p, hdr + BH_TILES, 0,
                                                   (const uint32_t *)b->tile_off.p, hdr + BH_VIEWS);
    }
    LAUNCHED();
    if (ev_march) HIPX(hipEventRecord(ev_march, st));
    launch_batch_classes(b->tn.mode, b->tn.occupancy, b->tn.variant, b->tn.queue_chunk, b->tn.waves_per_simd, b->n_cu,
                         b->cb, batch_args(b), st);
    LAUNCHED();
    if (b->sec_on) return enqueue_secondary(b, st);
    return VX_OK;
}

// Builds the installed view set's screen-bin lists, the first time a form that classifies
// by bins renders the set. If their buffers cannot be had, the set keeps whole lists (the
// same bits). A captured graph whose bin pointers no longer hold is dropped.
int build_bins(vx_batch *b, hipStream_t st) {
    if (!wants_bins(b) || !b->bins_fit || b->bins_built) return VX_OK;
    const void *c0 = b->bin_cnt.p, *l0 = b->bin_list.p;
    if (ensure(b->bin_cnt, (size_t)b->n_bins * 4) ||
        ensure(b->bin_list, (size_t)std::max<uint64_t>(b->bin_entries, 1) * 2)) {
        (void)hipGetLastError();
        b->bins_fit = false;
        return VX_OK;
    }
    if (c0 != b->bin_cnt.p || l0 != b->bin_list.p) drop_graph(b);
    k_bin_chunks<<<(unsigned)b->n_cu * 4u, 256, 0, st>>>(batch_args(b), (uint16_t *)b->bin_list.p,
                                                         (uint32_t *)b->bin_cnt.p);
    LAUNCHED();
    b->bins_built = true;
    return VX_OK;
}

int capture(vx_batch *b) {
    drop_graph(b);
    hipStream_t st = b->ctx->stream;
    HIPX(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = enqueue(b, st, nullptr);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    HIPX(e);
    b->graph = g;
    hipError_t ei = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        drop_graph(b);
        return vxh_hip_fail(ei);
    }
    b->graph_tiles = b->n_tiles;
    std::memcpy(b->graph_cb, b->cb, sizeof b->cb);
    b->graph_pool_gen = b->ctx->pool_gen;
    b->graph_bins = uses_bins(b);
    b->graph_valid = true;
    return VX_OK;
}

int view_buffers(vx_batch *b, uint32_t n_views, uint32_t n_tiles, uint64_t n_px, size_t n_chunk_recs, bool &moved) {
    const void *before[11] = {b->views.p, b->tile_off.p, b->hdr.p,  b->chunks.p, b->grids.p,   b->vis.p,
                              b->shadow.p, b->hiz.p,     b->cost.p, b->order.p,  b->counters.p};
    int rc;
    if ((rc = ensure(b->views, (size_t)std::max(n_views, 1u) * sizeof(ViewK))) ||
        (rc = ensure(b->tile_off, ((size_t)n_views + 1) * 4)) || (rc = ensure(b->hdr, BH_LEN * 4)) ||
        (rc = ensure(b->chunks, std::max<size_t>(n_chunk_recs, 1) * sizeof(vx_gpu_chunk))) ||
        (rc = ensure(b->grids, (size_t)std::max(n_views, 1u) * VX_GRID_LEN * 4)) ||
        (rc = ensure(b->vis, std::max<uint64_t>(n_px, 1) * 8)) ||
        (rc = ensure(b->shadow, std::max<uint64_t>(n_px, 2) * 2)) ||
        (rc = ensure(b->hiz, (size_t)std::max(n_tiles, 1u) * 4)) ||
        (rc = ensure(b->cost, (size_t)std::max(n_tiles, 1u) * 4)) ||
        (rc = ensure(b->order, (size_t)std::max(n_tiles, 1u) * 4)) ||
        (rc = ensure(b->counters, counter_words(n_views) * 4)))
        return rc;
    const void *after[11] = {b->views.p, b->tile_off.p, b->hdr.p,  b->chunks.p, b->grids.p,   b->vis.p,
                             b->shadow.p, b->hiz.p,     b->cost.p, b->order.p,  b->counters.p};
    moved = std::memcmp(before, after, sizeof before) != 0;
    return VX_OK;
}


}  // namespace

namespace vxd {

// One block per view. Each table record is classified (visible: inside the six planes;
// off-screen caster: outside but on the 32x8x32 grid; else dropped) and keyed by
// (class, f32 bits of its AABB distance, table index): an ascending sort of the keys is
// the two stable distance sorts, visible first. The block then copies the records in that
// order, fills the grid (one cell per LOD-0 chunk) and finds the camera chunk. counts
// gets (records, visible, camera chunk) per view. The f32 arithmetic is the host's, in
// its order, so the lists are the host builder's bytes.
__global__ __launch_bounds__(LIST_BLOCK) void k_view_lists(const vx_gpu_chunk *__restrict__ table, uint32_t n_table,
                                                           const ListCam *__restrict__ cams,
                                                           vx_gpu_chunk *__restrict__ out_chunks,
                                                           uint32_t *__restrict__ grids, uint32_t *__restrict__ counts,
                                                           ViewK *__restrict__ patch) {
    __shared__ unsigned long long keys[LIST_KEYS];
    __shared__ uint32_t s_vis_end, s_off_end, s_cam;
    const uint32_t tid = threadIdx.x;
    const ListCam &cam = cams[blockIdx.x];
    const uint32_t v = cam.view;
    uint32_t n2 = 1;
    while (n2 < n_table) n2 <<= 1;
    if (tid == 0) {
        s_vis_end = 0;
        s_off_end = 0;
        s_cam = VX_NO_CHUNK;
    }
    for (uint32_t i = tid; i < n2; i += LIST_BLOCK) {
        unsigned long long key = ~0ull;
        if (i < n_table) {
            const vx_gpu_chunk &c = table[i];
            const float mn[3] = {c.aabb_min[0], c.aabb_min[1], c.aabb_min[2]};
            const float mx[3] = {c.aabb_max[0], c.aabb_max[1], c.aabb_max[2]};
            bool inside = true;
            for (int p = 0; p < 6; p++) {
                const float *pl = cam.planes + 4 * p;
                const float q0 = pl[0] > 0.0f ? mx[0] : mn[0], q1 = pl[1] > 0.0f ? mx[1] : mn[1],
                            q2 = pl[2] > 0.0f ? mx[2] : mn[2];
                const float dd = pl[0] * q0 + pl[1] * q1 + pl[2] * q2;
                if (dd + pl[3] < 0.0f) inside = false;
            }
            float g[3];
            for (int k = 0; k < 3; k++) g[k] = fmaxf(fmaxf(mn[k] - cam.pos[k], cam.pos[k] - mx[k]), 0.0f);
            const float dist = sqrtf(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
            const unsigned long long dbits = (unsigned long long)__float_as_uint(dist) << 32;
            if (inside) {
                key = dbits | i;
            } else {
                // grid_span (src/render/mod.rs:93-108) of the record's key
                const int32_t span = 1 << c.lod, size = 64 << c.lod;
                bool on_grid = true;
                const int32_t dims[3] = {(int32_t)VX_GRID_X, (int32_t)VX_GRID_Y, (int32_t)VX_GRID_Z};
                for (int k = 0; k < 3; k++) {
                    const int32_t base = ((int32_t)c.origin[k] / size) * span - cam.gmin[k];
                    on_grid = on_grid && max(base, 0) < min(base + span, dims[k]);
                }
                if (on_grid) key = (1ull << 63) | dbits | i;
            }
        }
        keys[i] = key;
    }
    __syncthreads();
    // bitonic sort, ascending
    for (uint32_t k = 2; k <= n2; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = tid; i < n2; i += LIST_BLOCK) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const unsigned long long a = keys[i], b = keys[l];
                    if ((a > b) == ((i & k) == 0)) {
                        keys[i] = b;
                        keys[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    // class boundaries: visible keys [0, vis_end), casters [vis_end, off_end)
    for (uint32_t i = tid; i < n2; i += LIST_BLOCK) {
        const unsigned long long k0 = keys[i];
        const uint32_t c0 = k0 == ~0ull ? 2u : (uint32_t)(k0 >> 63);
        const unsigned long long k1 = i + 1 < n2 ? keys[i + 1] : ~0ull;
        const uint32_t c1 = k1 == ~0ull ? 2u : (uint32_t)(k1 >> 63);
        if (c0 == 0u && c1 != 0u) s_vis_end = i + 1;
        if (c0 != 2u && c1 == 2u) s_off_end = i + 1;
    }
    uint32_t *grid = grids + (size_t)v * VX_GRID_LEN;
    for (uint32_t j = tid; j < VX_GRID_LEN; j += LIST_BLOCK) grid[j] = VX_NO_CHUNK;
    __syncthreads();
    const uint32_t vis_all = s_vis_end, off_all = s_off_end > vis_all ? s_off_end - vis_all : 0u;
    const uint32_t nv = min(vis_all, (uint32_t)VX_MAX_CHUNKS);
    const uint32_t no = min(off_all, (uint32_t)VX_MAX_CHUNKS - nv);
    vx_gpu_chunk *out = out_chunks + cam.chunk_off;
    for (uint32_t r = tid; r < nv + no; r += LIST_BLOCK) {
        const uint32_t at = r < nv ? r : vis_all + (r - nv);
        const uint32_t idx = (uint32_t)keys[at];
        const vx_gpu_chunk &c = table[idx];
        const uint4 *src = (const uint4 *)&c;
        uint4 *dst = (uint4 *)&out[r];
#pragma unroll
        for (int w = 0; w < 6; w++) dst[w] = src[w];
        const int32_t span = 1 << c.lod, size = 64 << c.lod;
        int32_t cell[3];
        bool on_grid = true;
        const int32_t dims[3] = {(int32_t)VX_GRID_X, (int32_t)VX_GRID_Y, (int32_t)VX_GRID_Z};
        for (int k = 0; k < 3; k++) {
            cell[k] = ((int32_t)c.origin[k] / size) * span - cam.gmin[k];
            on_grid = on_grid && cell[k] >= 0 && cell[k] < dims[k];
        }
        if (on_grid) grid[(uint32_t)cell[0] + (uint32_t)cell[1] * VX_GRID_X + (uint32_t)cell[2] * VX_GRID_X * VX_GRID_Y] = r;
        if (r < nv && cam.pos[0] >= c.aabb_min[0] && cam.pos[1] >= c.aabb_min[1] && cam.pos[2] >= c.aabb_min[2] &&
            cam.pos[0] <= c.aabb_max[0] && cam.pos[1] <= c.aabb_max[1] && cam.pos[2] <= c.aabb_max[2])
            atomicMin(&s_cam, r);
    }
    __syncthreads();
    if (tid == 0) {
        if (counts) {
            counts[3 * v] = nv + no;
            counts[3 * v + 1] = nv;
            counts[3 * v + 2] = s_cam;
        }
        if (patch) {  // the view's descriptor, read by the march launched after this kernel
            patch[v].n_chunks = nv + no;
            patch[v].p.n_chunks = nv + no;
            patch[v].f.chunk_count = nv;
            patch[v].f.camera_chunk = s_cam;
        }
    }
}

}  // namespace vxd

namespace {

// Slots, descriptors and offsets of a view set (the first half of vx_batch_set_views).
struct ViewPlan {
    std::vector<ViewK> vk;
    std::vector<uint32_t> toff, slot_of, view_at;
    std::vector<uint64_t> poff;
    uint32_t cb[N_CLASSES + 1] = {0, 0, 0, 0, 0};
    size_t recs = 0;
    uint64_t bins = 0, bin_entries = 0;  // screen bins of the set and their list entries
};

template <class FrameOf, class SpecOf, class SpecHiOf, class TotalOf>
int plan_views(const vx_batch *b, uint32_t n, FrameOf frame_of, SpecOf spec_of, SpecHiOf spec_hi_of, TotalOf total_of,
               ViewPlan &pl) {
    pl.vk.assign(n, ViewK{});
    pl.toff.assign(n + 1, 0);
    pl.poff.assign(n + 1, 0);
    pl.slot_of.assign(n, 0);
    pl.view_at.assign(n, 0);
    {
        // Slots: the views of class 0, then class 1, ... (caller order inside a class).
        uint32_t next = 0;
        for (uint32_t c = 0; c < N_CLASSES; c++)
            for (uint32_t v = 0; v < n; v++) {
                const uint32_t fl = frame_of(v)->flags;
                if (view_class((fl & VX_FLAG_HIZ) != 0u, (fl & VX_FLAG_UNDERWATER) != 0u) != c) continue;
                pl.slot_of[v] = next;
                pl.view_at[next++] = v;
            }
    }
    for (uint32_t v = 0; v < n; v++) {
        const uint32_t src = pl.view_at[v];
        const vx_gpu_frame *fr = frame_of(src);
        const uint32_t total = total_of(src);
        int rc = vxh::check_frame(fr, total);
        if (rc) return rc;
        const uint32_t cls = view_class((fr->flags & VX_FLAG_HIZ) != 0u, (fr->flags & VX_FLAG_UNDERWATER) != 0u);
        ViewK &k = pl.vk[v];
        std::memset(&k, 0, sizeof k);
        k.f = vxh::make_frame(fr);
        if ((rc = vxh::make_spec(spec_of(src), spec_hi_of(src), fr->flags, k.s))) return rc;
        k.chunk_off = (uint32_t)pl.recs;
        k.n_chunks = total;
        k.grid_off = v * VX_GRID_LEN;
        k.tile_off = pl.toff[v];
        k.px_off_lo = (uint32_t)pl.poff[v];
        k.px_off_hi = (uint32_t)(pl.poff[v] >> 32);
        const uint64_t nbv = (uint64_t)bins_x(k.f.tiles[0]) * bins_y(k.f.tiles[1]);
        k.bin_off = (uint32_t)std::min<uint64_t>(pl.bins, 0xFFFFFFFFull);
        k.bin_stride = total;
        k.binl_off = (uint32_t)std::min<uint64_t>(pl.bin_entries, 0xFFFFFFFFull);
        pl.bins += nbv;
        pl.bin_entries += nbv * total;
        const uint64_t nt = (uint64_t)k.f.tiles[0] * k.f.tiles[1];
        if (pl.toff[v] + nt > 0xFFFFFFFFull) return VX_E_OVERFLOW;
        pl.toff[v + 1] = pl.toff[v] + (uint32_t)nt;
        pl.poff[v + 1] = pl.poff[v] + (uint64_t)k.f.res[0] * k.f.res[1];
        pl.recs += total;
        for (uint32_t c = cls + 1; c <= N_CLASSES; c++) pl.cb[c] = pl.toff[v + 1];
    }
    if (pl.recs > 0xFFFFFFFFull) return VX_E_OVERFLOW;
    return VX_OK;
}

// The second half: descriptors, offsets and header to the device, targets cleared, the
// captured graph kept only when it still fits the view set.
int install_views(vx_batch *b, uint32_t n, ViewPlan &pl, bool moved, bool keep_targets = false) {
    vx_ctx *c = b->ctx;
    hipStream_t st = c->stream;
    int rc;
    for (ViewK &k : pl.vk) view_pools(b, k);  // after view_buffers: the chunk and grid buffers are placed
    const std::vector<uint32_t> &toff = pl.toff;
    const std::vector<uint64_t> &poff = pl.poff;
    const uint32_t *cb = pl.cb;
    // Screen-bin lists go with the view set: they depend only on the views' frames and
    // chunk lists. A set whose lists would not fit keeps whole lists (bins off). They are
    // built by the first render of a form that classifies by bins (build_bins).
    const bool fit = pl.bins > 0 && pl.bins <= 0xFFFFFFFFull && pl.bin_entries <= 0xFFFFFFFFull &&
                     pl.bin_entries * 2 <= bin_list_max_bytes();
    uint32_t hdr[BH_LEN] = {};
    hdr[BH_TILES] = toff[n];
    hdr[BH_VIEWS] = n;
    hdr[BH_BINS] = fit ? (uint32_t)pl.bins : 0u;
    for (uint32_t k = 0; k <= N_CLASSES; k++) hdr[BH_CB + k] = cb[k];
    HIPX(hipMemcpyAsync(b->views.p, pl.vk.data(), (size_t)n * sizeof(ViewK), hipMemcpyHostToDevice, st));
    HIPX(hipMemcpyAsync(b->tile_off.p, toff.data(), ((size_t)n + 1) * 4, hipMemcpyHostToDevice, st));
    HIPX(hipMemcpyAsync(b->hdr.p, hdr, sizeof hdr, hipMemcpyHostToDevice, st));
    b->bins_fit = fit;
    b->bins_built = false;
    b->n_bins = fit ? (uint32_t)pl.bins : 0u;
    b->bin_entries = fit ? pl.bin_entries : 0u;
    // New views start as a new Renderer would: empty targets, Hi-Z 0; costs 0. Moved
    // cameras keep them (the same views, one frame later).
    if (!keep_targets &&
        ((rc = vxh::fill_u64(st, b->vis.p, 0, poff[n])) || (rc = vxh::fill_u32(st, b->shadow.p, 0, (poff[n] + 1) / 2)) ||
         (rc = vxh::fill_u32(st, b->hiz.p, 0, toff[n])) || (rc = vxh::fill_u32(st, b->cost.p, 0, toff[n]))))
        return rc;
    HIPX(hipStreamSynchronize(st));
    // A captured graph stays valid for another view set only if it reads every count from
    // the header and launches the same class kernels: not PLAIN (per-view launches), not
    // FLAT with different class ranges (grid sizes), not PERSISTENT with a different set
    // of non-empty classes.
    bool same_classes = true, same_ranges = true;
    for (uint32_t c = 0; c < N_CLASSES; c++) {
        same_classes &= ((cb[c + 1] - cb[c]) != 0u) == ((b->graph_cb[c + 1] - b->graph_cb[c]) != 0u);
        same_ranges &= (cb[c + 1] - cb[c]) == (b->graph_cb[c + 1] - b->graph_cb[c]);
    }
    if (moved || b->tn.mode == VX_BATCH_PLAIN || (b->tn.mode == VX_BATCH_FLAT && !same_ranges) || !same_classes)
        drop_graph(b);
    b->vk.swap(pl.vk);
    b->vk_pool_gen = c->pool_gen;
    b->toff.swap(pl.toff);
    b->poff.swap(pl.poff);
    b->slot_of.swap(pl.slot_of);
    std::memcpy(b->cb, pl.cb, sizeof b->cb);
    b->n_views = n;
    b->n_tiles = b->toff[n];
    b->n_px = b->poff[n];
    b->rendered = false;
    if (b->sec_on && (rc = secondary_buffers(b))) {
        b->sec_on = false;
        drop_graph(b);
        return rc;
    }
    return VX_OK;
}

}  // namespace

extern "C" {

int vx_batch_create(vx_ctx *c, uint32_t max_views, vx_batch **out) {
    if (!c || !out || max_views == 0 || max_views > BATCH_MAX_VIEWS) return VX_E_ARG;
    if (c->path != VX_PATH_TILES) return VX_E_UNSUPPORTED;
    *out = nullptr;
    HIPX(hipSetDevice(c->device));
    vx_batch *b = new (std::nothrow) vx_batch();
    if (!b) return VX_E_OOM;
    b->ctx = c;
    b->max_views = max_views;
    int rc = VX_OK;
    if (hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess ||
        b->n_cu <= 0)
        rc = VX_E_HIP;
    for (auto &set : b->ring)
        for (auto &e : set)
            if (rc == VX_OK && hipEventCreate(&e) != hipSuccess) rc = VX_E_HIP;
    if (rc) {
        for (auto &set : b->ring)
            for (auto &e : set)
                if (e) (void)hipEventDestroy(e);
        delete b;
        return rc;
    }
    b->tn.mode = VX_BATCH_PERSISTENT;
    b->tn.occupancy = 4;
    b->tn.variant = VAR_FUSED_SHADOW;
    b->tn.ordered = 1;
    b->tn.queue_chunk = 1;
    b->tn.waves_per_simd = 4;
    b->tn.graph = 1;
    c->n_batches++;
    *out = b;
    return VX_OK;
}

int vx_batch_destroy(vx_batch *b) {
    if (!b) return VX_E_ARG;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    drop_graph(b);
    DevBuf *bufs[] = {&b->views,   &b->tile_off, &b->hdr,       &b->chunks, &b->grids, &b->vis,
                      &b->shadow,  &b->hiz,      &b->cost,      &b->order,  &b->counters, &b->sec_out,
                      &b->sec_queue, &b->table,  &b->cams,      &b->list_counts, &b->bin_cnt, &b->bin_list};
    for (DevBuf *d : bufs) vxh::release(*d);
    for (auto &set : b->ring)
        for (auto &e : set)
            if (e) (void)hipEventDestroy(e);
    b->ctx->n_batches--;
    delete b;
    return VX_OK;
}

int vx_batch_set_views(vx_batch *b, const vx_view *views, uint32_t n) {
    if (!b || !views || n == 0 || n > b->max_views) return VX_E_ARG;
    for (uint32_t v = 0; v < n; v++)
        if (!views[v].frame || !views[v].grid || (!views[v].chunks && views[v].n_total) ||
            views[v].n_total > VX_MAX_CHUNKS)
            return VX_E_ARG;
    ViewPlan pl;
    int rc = plan_views(
        b, n, [&](uint32_t v) { return views[v].frame; }, [&](uint32_t v) { return views[v].spec; },
        [&](uint32_t v) { return views[v].spec_hi; }, [&](uint32_t v) { return views[v].n_total; }, pl);
    if (rc) return rc;
    vx_ctx *c = b->ctx;
    HIPX(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    HIPX(hipStreamSynchronize(st));
    bool moved = false;
    if ((rc = view_buffers(b, n, pl.toff[n], pl.poff[n], pl.recs, moved))) {
        drop_graph(b);
        b->n_views = 0;
        return rc;
    }
    std::vector<vx_gpu_chunk> cat(std::max<size_t>(pl.recs, 1));
    std::vector<uint32_t> grids((size_t)n * VX_GRID_LEN);
    for (uint32_t v = 0; v < n; v++) {
        const vx_view &w = views[pl.view_at[v]];
        if (w.n_total) std::memcpy(&cat[pl.vk[v].chunk_off], w.chunks, (size_t)w.n_total * sizeof(vx_gpu_chunk));
        std::memcpy(&grids[(size_t)v * VX_GRID_LEN], w.grid, VX_GRID_LEN * 4);
    }
    if (pl.recs)
        HIPX(hipMemcpyAsync(b->chunks.p, cat.data(), pl.recs * sizeof(vx_gpu_chunk), hipMemcpyHostToDevice, st));
    HIPX(hipMemcpyAsync(b->grids.p, grids.data(), grids.size() * 4, hipMemcpyHostToDevice, st));
    return install_views(b, n, pl, moved);
}

int vx_batch_set_chunk_table(vx_batch *b, const vx_gpu_chunk *chunks, uint32_t n) {
    if (!b || (!chunks && n) || n > VX_CHUNK_TABLE_MAX) return VX_E_ARG;
    for (uint32_t i = 0; i < n; i++)  // LOD-0 records of a render set: one grid cell each, share 1
        if (chunks[i].lod != 0u || chunks[i].fade != 1.0f || chunks[i].voxel_size != 1.0f) return VX_E_ARG;
    vx_ctx *c = b->ctx;
    HIPX(hipSetDevice(c->device));
    HIPX(hipStreamSynchronize(c->stream));
    int rc;
    if ((rc = ensure(b->table, (size_t)std::max(n, 1u) * sizeof(vx_gpu_chunk)))) return rc;
    if (n) HIPX(hipMemcpyAsync(b->table.p, chunks, (size_t)n * sizeof(vx_gpu_chunk), hipMemcpyHostToDevice, c->stream));
    HIPX(hipStreamSynchronize(c->stream));
    b->n_table = n;
    return VX_OK;
}

static int set_cameras(vx_batch *b, const vx_camera_view *views, uint32_t n, bool keep_targets) {
    if (!b || !views || n == 0 || n > b->max_views) return VX_E_ARG;
    for (uint32_t v = 0; v < n; v++)
        if (!views[v].frame) return VX_E_ARG;
    const uint32_t nt = b->n_table;
    // The lists are built on the device: every view gets room for the whole table, and its
    // frame's chunk_count / camera_chunk are the device's.
    std::vector<vx_gpu_frame> frames(n);
    for (uint32_t v = 0; v < n; v++) {
        frames[v] = *views[v].frame;
        frames[v].chunk_count = 0;
        frames[v].camera_chunk = VX_NO_CHUNK;
    }
    ViewPlan pl;
    int rc = plan_views(
        b, n, [&](uint32_t v) { return &frames[v]; }, [&](uint32_t v) { return views[v].spec; },
        [&](uint32_t v) { return views[v].spec_hi; }, [&](uint32_t) { return nt; }, pl);
    if (rc) return rc;
    // Moved cameras: the views must sit where they sat (same count, sizes and classes),
    // so every pixel, tile and Hi-Z word keeps its place.
    if (keep_targets && (n != b->n_views || pl.toff != b->toff || pl.poff != b->poff || pl.slot_of != b->slot_of ||
                         std::memcmp(pl.cb, b->cb, sizeof b->cb) != 0))
        return VX_E_ARG;
    vx_ctx *c = b->ctx;
    HIPX(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    HIPX(hipStreamSynchronize(st));
    bool moved = false;
    if ((rc = view_buffers(b, n, pl.toff[n], pl.poff[n], pl.recs, moved)) ||
        (rc = ensure(b->cams, (size_t)n * sizeof(ListCam))) || (rc = ensure(b->list_counts, (size_t)n * 3 * 4))) {
        drop_graph(b);
        b->n_views = 0;
        return rc;
    }
    std::vector<ListCam> cams(n);
    for (uint32_t slot = 0; slot < n; slot++) {
        const vx_camera_view &w = views[pl.view_at[slot]];
        ListCam &lc = cams[slot];
        std::memset(&lc, 0, sizeof lc);
        std::memcpy(lc.planes, w.planes, sizeof lc.planes);
        for (int i = 0; i < 3; i++) {
            lc.pos[i] = w.frame->cam_pos[i];
            lc.gmin[i] = w.frame->grid_min[i];
        }
        lc.chunk_off = pl.vk[slot].chunk_off;
        lc.view = slot;
    }
    HIPX(hipMemcpyAsync(b->cams.p, cams.data(), (size_t)n * sizeof(ListCam), hipMemcpyHostToDevice, st));
    k_view_lists<<<n, LIST_BLOCK, 0, st>>>((const vx_gpu_chunk *)b->table.p, nt, (const ListCam *)b->cams.p,
                                          (vx_gpu_chunk *)b->chunks.p, (uint32_t *)b->grids.p,
                                          (uint32_t *)b->list_counts.p, nullptr);
    LAUNCHED();
    std::vector<uint32_t> counts((size_t)n * 3);
    HIPX(hipMemcpyAsync(counts.data(), b->list_counts.p, counts.size() * 4, hipMemcpyDeviceToHost, st));
    HIPX(hipStreamSynchronize(st));
    for (uint32_t slot = 0; slot < n; slot++) {
        ViewK &k = pl.vk[slot];
        k.n_chunks = counts[3 * slot];
        k.f.chunk_count = counts[3 * slot + 1];
        k.f.camera_chunk = counts[3 * slot + 2];
    }
    return install_views(b, n, pl, moved, keep_targets);
}

int vx_batch_set_cameras(vx_batch *b, const vx_camera_view *views, uint32_t n) {
    return set_cameras(b, views, n, false);
}

int vx_batch_move_cameras(vx_batch *b, const vx_camera_view *views, uint32_t n) {
    return set_cameras(b, views, n, true);
}

int vx_batch_view_counts(vx_batch *b, uint32_t view, uint32_t *n_total, uint32_t *n_visible, uint32_t *camera_chunk) {
    if (!b || view >= b->n_views) return VX_E_ARG;
    const ViewK &k = b->vk[b->slot_of[view]];
    if (n_total) *n_total = k.n_chunks;
    if (n_visible) *n_visible = k.f.chunk_count;
    if (camera_chunk) *camera_chunk = k.f.camera_chunk;
    return VX_OK;
}

int vx_batch_read_chunks(vx_batch *b, uint32_t view, vx_gpu_chunk *chunks, size_t cap, uint32_t *grid) {
    if (!b || view >= b->n_views) return VX_E_ARG;
    const ViewK &k = b->vk[b->slot_of[view]];
    if ((!chunks && k.n_chunks) || cap < k.n_chunks) return VX_E_ARG;
    vx_ctx *c = b->ctx;
    HIPX(hipSetDevice(c->device));
    if (k.n_chunks)
        HIPX(hipMemcpyAsync(chunks, (const vx_gpu_chunk *)b->chunks.p + k.chunk_off,
                            (size_t)k.n_chunks * sizeof(vx_gpu_chunk), hipMemcpyDeviceToHost, c->stream));
    if (grid)
        HIPX(hipMemcpyAsync(grid, (const uint32_t *)b->grids.p + k.grid_off, VX_GRID_LEN * 4, hipMemcpyDeviceToHost,
                            c->stream));
    HIPX(hipStreamSynchronize(c->stream));
    return VX_OK;
}

int vx_batch_set_tuning(vx_batch *b, const vx_batch_tuning *t) {
    if (!b || !t || !batch_tuning_ok(*t)) return VX_E_ARG;
    if (std::memcmp(&b->tn, t, sizeof *t) != 0) drop_graph(b);
    b->tn = *t;
    return VX_OK;
}

int vx_batch_get_tuning(vx_batch *b, vx_batch_tuning *t) {
    if (!b || !t) return VX_E_ARG;
    *t = b->tn;
    return VX_OK;
}

int vx_batch_render(vx_batch *b) {
    if (!b) return VX_E_ARG;
    if (b->n_views == 0) return VX_E_ARG;
    vx_ctx *c = b->ctx;
    HIPX(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    hipEvent_t *ev = b->ring[b->ring_next];
    b->ring_next = (b->ring_next + 1) % vx_batch::RING;
    if (b->ring_count < (uint32_t)vx_batch::RING) b->ring_count++;
    HIPX(hipEventRecord(ev[0], st));
    int rc = refresh_view_pools(b, st);
    if (rc) return rc;
    if ((rc = build_bins(b, st))) return rc;
    if (b->graph_valid && b->graph_bins != uses_bins(b)) drop_graph(b);
    if (b->tn.graph) {
        if ((!b->graph_valid || b->graph_pool_gen != c->pool_gen) && (rc = capture(b))) return rc;
        HIPX(hipEventRecord(ev[1], st));
        HIPX(hipEventRecord(ev[2], st));
        HIPX(hipGraphLaunch(b->exec, st));
    } else {
        HIPX(hipEventRecord(ev[1], st));
        const bool one_march = b->tn.mode != VX_BATCH_PLAIN;
        if (!one_march) HIPX(hipEventRecord(ev[2], st));
        if ((rc = enqueue(b, st, one_march ? ev[2] : nullptr))) return rc;
    }
    HIPX(hipEventRecord(ev[3], st));
    b->rendered = true;
    return VX_OK;
}

int vx_batch_sync(vx_batch *b) {
    if (!b) return VX_E_ARG;
    return vx_sync(b->ctx);
}

int vx_batch_reset_hiz(vx_batch *b) {
    if (!b) return VX_E_ARG;
    HIPX(hipSetDevice(b->ctx->device));
    int rc = vxh::fill_u32(b->ctx->stream, b->hiz.p, 0, b->n_tiles);
    if (rc) return rc;
    HIPX(hipStreamSynchronize(b->ctx->stream));
    return VX_OK;
}

static int batch_read(vx_batch *b, const DevBuf &buf, size_t off_bytes, void *out, size_t bytes) {
    HIPX(hipSetDevice(b->ctx->device));
    HIPX(hipMemcpyAsync(out, (const char *)buf.p + off_bytes, bytes, hipMemcpyDeviceToHost, b->ctx->stream));
    HIPX(hipStreamSynchronize(b->ctx->stream));
    return VX_OK;
}

int vx_batch_read_vis(vx_batch *b, uint32_t view, uint64_t *out, size_t n) {
    if (!b || !out || view >= b->n_views) return VX_E_ARG;
    const uint32_t v = b->slot_of[view];
    if (n != b->poff[v + 1] - b->poff[v]) return VX_E_ARG;
    return batch_read(b, b->vis, b->poff[v] * 8, out, n * 8);
}

int vx_batch_read_shadow(vx_batch *b, uint32_t view, uint16_t *out, size_t n) {
    if (!b || !out || view >= b->n_views) return VX_E_ARG;
    const uint32_t v = b->slot_of[view];
    if (n != b->poff[v + 1] - b->poff[v]) return VX_E_ARG;
    return batch_read(b, b->shadow, b->poff[v] * 2, out, n * 2);
}

int vx_batch_read_hiz(vx_batch *b, uint32_t view, uint32_t *out, size_t n) {
    if (!b || !out || view >= b->n_views) return VX_E_ARG;
    const uint32_t v = b->slot_of[view];
    if (n != (size_t)(b->toff[v + 1] - b->toff[v])) return VX_E_ARG;
    return batch_read(b, b->hiz, (size_t)b->toff[v] * 4, out, n * 4);
}

int vx_batch_demand_get(vx_batch *b, vx_pair_demand *out) {
    if (!b || !out) return VX_E_ARG;
    std::vector<uint32_t> cnt(CNT_LEN);
    int rc = batch_read(b, b->counters, 0, cnt.data(), CNT_LEN * 4);
    if (rc) return rc;
    uint64_t s[4] = {0, 0, 0, 0};
    for (uint32_t sl = 0; sl < CNT_STAT_SLOTS; sl++)
        for (uint32_t k = 0; k < 4; k++) s[k] += cnt[CNT_STAT_BASE + sl * CNT_STAT_STRIDE + k];
    std::memset(out, 0, sizeof *out);
    out->select = (uint32_t)s[STAT_SELECT];
    out->deferred = (uint32_t)s[STAT_DEFERRED];
    out->recover = (uint32_t)s[STAT_RECOVER];
    out->shadow_rays = (uint32_t)s[STAT_SHADOW];
    return VX_OK;
}

int vx_batch_timings_reset(vx_batch *b) {
    if (!b) return VX_E_ARG;
    HIPX(hipSetDevice(b->ctx->device));
    HIPX(hipStreamSynchronize(b->ctx->stream));
    b->ring_count = 0;
    return VX_OK;
}

int vx_batch_timings_sum(vx_batch *b, float *total_ms, float *march_ms, uint32_t *frames) {
    if (!b || !total_ms) return VX_E_ARG;
    HIPX(hipSetDevice(b->ctx->device));
    HIPX(hipStreamSynchronize(b->ctx->stream));
    float tot = 0, mar = 0;
    for (uint32_t k = 0; k < b->ring_count; k++) {
        uint32_t idx = (b->ring_next + vx_batch::RING - 1 - k) % vx_batch::RING;
        float ms = 0, mm = 0;
        HIPX(hipEventElapsedTime(&ms, b->ring[idx][1], b->ring[idx][3]));
        HIPX(hipEventElapsedTime(&mm, b->ring[idx][2], b->ring[idx][3]));
        tot += ms;
        mar += mm;
    }
    *total_ms = tot;
    if (march_ms) *march_ms = mar;
    if (frames) *frames = b->ring_count;
    return VX_OK;
}

// Times every batched launch form on the current views (reps frames each, after one
// warm frame that also refreshes the tile costs) and keeps the fastest. Outputs do not
// depend on the choice; the Hi-Z words reach their steady state after the first frame.
int vx_batch_autotune(vx_batch *b, uint32_t reps, vx_batch_tuning *best, float *best_ms) {
    if (!b || b->n_views == 0 || reps == 0 || reps > 64) return VX_E_ARG;
    std::vector<vx_batch_tuning> cand;
    auto add = [&](uint32_t mode, uint32_t occ, uint32_t var, uint32_t chunk, uint32_t wps) {
        vx_batch_tuning t{};
        t.mode = mode;
        t.occupancy = occ;
        t.variant = var;
        t.ordered = 1;
        t.queue_chunk = chunk;
        t.waves_per_simd = wps;
        t.graph = 1;
        cand.push_back(t);
    };
    for (uint32_t occ : {4u, 6u})
        for (uint32_t var : {4u, 5u, 12u, 13u}) {
            add(VX_BATCH_FLAT, occ, var, 1, occ);
            add(VX_BATCH_PERSISTENT, occ, var, 4, occ);
        }
    add(VX_BATCH_PERSISTENT, 6, 4, 2, 6);
    add(VX_BATCH_PERSISTENT, 6, 4, 8, 6);
    const vx_batch_tuning saved = b->tn;
    float best_t = 1e30f;
    vx_batch_tuning best_c = saved;
    int rc = VX_OK;
    for (const vx_batch_tuning &t : cand) {
        if ((rc = vx_batch_set_tuning(b, &t))) break;
        if ((rc = vx_batch_render(b))) break;
        if ((rc = vx_batch_timings_reset(b))) break;
        for (uint32_t r = 0; r < reps && rc == VX_OK; r++) rc = vx_batch_render(b);
        if (rc) break;
        float ms = 0;
        uint32_t frames = 0;
        if ((rc = vx_batch_timings_sum(b, &ms, nullptr, &frames))) break;
        ms /= (float)std::max(frames, 1u);
        if (ms < best_t) {
            best_t = ms;
            best_c = t;
        }
    }
    if (rc) {
        vx_batch_set_tuning(b, &saved);
        return rc;
    }
    if ((rc = vx_batch_set_tuning(b, &best_c))) return rc;
    if (best) *best = best_c;
    if (best_ms) *best_ms = best_t;
    return VX_OK;
}

int vx_batch_set_secondary(vx_batch *b, uint32_t on) {
    if (!b || on > 1) return VX_E_ARG;
    HIPX(hipSetDevice(b->ctx->device));
    HIPX(hipStreamSynchronize(b->ctx->stream));
    if (on && b->n_views) {
        int rc = secondary_buffers(b);
        if (rc) return rc;
    }
    if ((bool)on != b->sec_on) drop_graph(b);
    b->sec_on = on != 0;
    return VX_OK;
}

int vx_batch_read_secondary(vx_batch *b, uint32_t view, vx_sec_hit *out, size_t n) {
    if (!b || !out || view >= b->n_views || !b->sec_on || !b->rendered) return VX_E_ARG;
    const uint32_t v = b->slot_of[view];
    if (n != SEC_RECORDS * (b->poff[v + 1] - b->poff[v])) return VX_E_ARG;
    return batch_read(b, b->sec_out, b->poff[v] * SEC_RECORDS * 16, out, n * 16);
}

int vx_batch_secondary_count(vx_batch *b, uint64_t *water_legs, uint64_t *glass_legs, uint64_t *sun_rays,
                             uint32_t *overflow) {
    if (!b) return VX_E_ARG;
    uint32_t w[CNT_TOTAL - CNT_SEC];
    int rc = batch_read(b, b->counters, CNT_SEC * 4, w, sizeof w);
    if (rc) return rc;
    if (water_legs) *water_legs = (uint64_t)w[SEC_WATER * SEC_LINE] + w[SEC_SUB * SEC_LINE];
    if (glass_legs) *glass_legs = w[SEC_GLASS * SEC_LINE];
    if (sun_rays) *sun_rays = w[SEC_SUN * SEC_LINE];
    if (overflow) *overflow = w[SEC_OVERFLOW * SEC_LINE];
    return VX_OK;
}

int vx_batch_device_outputs(vx_batch *b, uint32_t view, uint64_t **vis, uint16_t **shadow) {
    if (!b || view >= b->n_views) return VX_E_ARG;
    const uint32_t v = b->slot_of[view];
    if (vis) *vis = (uint64_t *)b->vis.p + b->poff[v];
    if (shadow) *shadow = (uint16_t *)b->shadow.p + b->poff[v];
    return VX_OK;
}

}  // extern "C"
