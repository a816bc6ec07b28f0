// ctx.hpp -- host-side state of a march context, shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>

#include "kernels.hpp"
#include "vxmarch.h"

#define VX_HIPX(expr)                                    \
    do {                                                 \
        hipError_t e_ = (expr);                          \
        if (e_ != hipSuccess) return vxh_hip_fail(e_);   \
    } while (0)
#define VX_LAUNCHED() VX_HIPX(hipPeekAtLastError())

int vxh_hip_fail(hipError_t e);

namespace vxh {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

// Grow-only device allocation (contents are not preserved).
int ensure(DevBuf &b, size_t bytes);
void release(DevBuf &b);

int fill_u32(hipStream_t st, void *p, uint32_t v, size_t n);
int fill_u64(hipStream_t st, void *p, uint64_t v, size_t n);

vxd::FrameK make_frame(const vx_gpu_frame *g);
// Kernel spec bits for a (flags & SPEC_MASK, spec_hi) pipeline key; VX_E_UNSUPPORTED for
// the features that need the texture atlas or the refraction path.
int make_spec(uint32_t spec, uint32_t spec_hi, uint32_t frame_flags, vxd::Spec &s);
// Checks a frame against the fixed limits (tile counts, grid size, chunk count).
int check_frame(const vx_gpu_frame *g, uint32_t n_chunks);

// Launch of the tile-per-wave march for one view (one workgroup per tile, in `order`).
using vxd::TileArgs;
bool tuning_ok(uint32_t occ, uint32_t var);
void launch_tiles(uint32_t occ, uint32_t var, bool hz, bool uw, unsigned nb, hipStream_t st, const TileArgs &a);
// Cost-bucket launch order of n tiles into `order` (entries are tile indices); `ctl`
// is the CTL_LEN-word control block, cleared beforehand.
void launch_order(hipStream_t st, const uint32_t *cost, uint32_t n, uint32_t *ctl, uint32_t *order);
// Looks up the device address of vx_render_noalloc's working array on `device` (which
// must be current) and keeps it for later calls.
int noalloc_setup(int device);

}  // namespace vxh

struct vx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t w = 0, h = 0, tx = 0, ty = 0, max_pairs = 0, max_pairs_req = 0, path = VX_PATH_TILES;
    vxh::DevBuf pool[3];
    size_t pool_len[3] = {0, 0, 0};
    bool pool_synced[3] = {false, false, false};
    uint64_t pool_uploaded[3] = {0, 0, 0};  // GpuMirror::uploaded_bytes per pool
    vxh::DevBuf chunks, grid, vis, shadow, hiz, pairs, counters;
    vxh::DevBuf cost, order;  // per-tile cycles of the last frame, and the launch order
    vxh::DevBuf heat;         // FLAG_HEATMAP: marched iterations per pixel
    bool heat_valid = false;  // the last frame was a heatmap frame
    bool have_cost = false;
    uint32_t n_chunks = 0;
    // Ring of per-frame event sets: [0] start, [1..5] pass boundaries, [6] end.
    static constexpr int RING = 256;
    hipEvent_t ring[RING][7] = {};
    uint32_t ring_next = 0;   // set the next frame records into
    uint32_t ring_count = 0;  // frames recorded since the last reset (saturates at RING)
    hipEvent_t *ev = nullptr; // the set of the frame being recorded / last recorded
    uint32_t occ = 4;                          // tile-march occupancy variant (waves per SIMD)
    uint32_t variant = vxd::VAR_FUSED_SHADOW;  // tile-march structure variant (VAR_* bits)
    uint32_t ordered = 1;  // launch tiles most-expensive-first (1) or in raster order (0)
    bool have_frame = false;
    vxd::FrameK frame{};
    vxd::Spec spec{};
    int n_batches = 0;     // batches bound to this context (it must outlive them)
    // Bumped whenever a pool's device address or length changes: captured graphs hold
    // both in their kernel arguments and are re-captured when it moves.
    uint64_t pool_gen = 0;
    // vx_pool_set_flat: one buffer holding every chunk's nodes, leaves and attribute words
    // (vxw_flat_layout); while flat_len > 0 the kernels read all three pools from it.
    vxh::DevBuf flat;
    size_t flat_len = 0;
    // Device chunk build (vx_build_local): dense input staging, per-leaf and per-L1
    // scratch, record output staging; grown on demand, kept for the next call.
    vxh::DevBuf build_dense, build_leaf, build_words, build_l1, build_out;
    // The last vx_render's frame record and pipeline key (the resolve reads them).
    vx_gpu_frame gframe{};
    uint32_t spec_word = 0, spec_hi_word = 0;
    // Resolve: the caller's atlas and face table, and the colour target.
    vxh::DevBuf atlas, faces, color, image;
    // TAA: two f16 RGBA histories (the last output is hist[taa_last]) and the f32 copy of
    // one that the present reads; taa_valid while the last pass over the colour was vx_taa.
    vxh::DevBuf hist[2], taa_f32;
    vxh::DevBuf sec_r, sec_m;  // the resolve's water_sec legs (2x2 blocks, f16 RGBA)
    vxh::DevBuf shaft;         // the distant-shadow field: scan slices 0 and 1, heights in slice 2
    vxh::DevBuf probe;         // the ambient probe field: 128 x 768 x 128 RGBA16F
    vxh::DevBuf soft_h, soft_mask;  // soft shadows: the row pass (RG16F) and the mask (R8)
    bool probe_pinned = false; // filled with a value: writes are ignored
    size_t hist_len = 0;
    uint32_t taa_last = 0;
    bool taa_valid = false;
    uint32_t atlas_size = 0, atlas_layers = 0, atlas_mips = 0, n_faces = 0;
    bool color_valid = false, image_valid = false;
};

namespace vxh {
vxd::Pools make_pools(const vx_ctx *c);
// build_local of n dense chunks (device memory) into records (device memory) on the
// context's stream, without waiting for it.
int build_local_async(vx_ctx *c, const uint16_t *d_dense, uint32_t n, uint8_t *d_out, size_t slot_bytes);
}
