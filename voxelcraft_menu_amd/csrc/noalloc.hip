// noalloc.hip -- vx_render_noalloc: the tile march of vx_render on caller-held device
// buffers, for callers that cannot allocate device memory at call time (code under HIP
// graph capture, tight serving loops).
//
// The march's working memory -- the counter block (statistics and the launch-order
// control words), the per-tile launch order and the per-tile cycle counts the kernel
// writes -- lives in `scratch`, a file-scope device array of fixed size. Each GPU holds
// its own copy until the process ends. Host code never passes the array's name to a
// kernel: the name is a host-side stand-in, so every use goes through the device address
// hipGetSymbolAddress returns, looked up once per device (at vx_ctx_create for the
// context's device, or at the first call on a device) and kept per device.
//
// Every call on a device shares that device's array and leaves its contents behind; the
// counter block is cleared by a fill kernel at the start of each call. Two calls that
// overlap on one device (different streams, or threads) race on the array.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <mutex>

#include "ctx.hpp"

using namespace vxd;

namespace {

constexpr uint32_t kMaxTiles = VX_NOALLOC_MAX_TILES;
constexpr size_t kOrderOff = CNT_TOTAL;                 // u32 words
constexpr size_t kCostOff = kOrderOff + kMaxTiles;
constexpr size_t kScratchWords = kCostOff + kMaxTiles;  // ~1.01 MiB

static __device__ __align__(256) uint32_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::atomic<uint32_t *> g_scratch[kMaxDevices];
std::mutex g_lock;

// Device address of this device's `scratch` (the current device must be `dev`).
int scratch_for(int dev, uint32_t **out) {
    if (dev < 0 || dev >= kMaxDevices) return VX_E_ARG;
    uint32_t *p = g_scratch[dev].load(std::memory_order_acquire);
    if (!p) {
        std::lock_guard<std::mutex> g(g_lock);
        p = g_scratch[dev].load(std::memory_order_relaxed);
        if (!p) {
            void *ptr = nullptr;
            hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !ptr) {
                std::fprintf(stderr, "vxmarch: no device address for the no-alloc scratch on device %d (%s)\n", dev,
                             hipGetErrorString(e));
                return vxh_hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol);
            }
            p = (uint32_t *)ptr;
            g_scratch[dev].store(p, std::memory_order_release);
        }
    }
    *out = p;
    return VX_OK;
}

}  // namespace

namespace vxh {
// Called by vx_ctx_create with the context's device current.
int noalloc_setup(int device) {
    uint32_t *p = nullptr;
    return scratch_for(device, &p);
}
}  // namespace vxh

extern "C" int vx_render_noalloc(const vx_gpu_frame *g, uint32_t spec, uint32_t spec_hi, const vx_device_view *in,
                                 uint32_t *hiz, uint64_t *vis, uint16_t *shadow, void *stream) {
    if (!g || !in || !hiz || !vis || !shadow) return VX_E_ARG;
    if (!in->inners || !in->leaves || !in->bricks || !in->chunks || !in->grid) return VX_E_ARG;
    if ((in->inners_bytes | in->leaves_bytes | in->bricks_bytes) > 0xFFFFFFF0ull) return VX_E_OVERFLOW;
    if (int rc = vxh::check_frame(g, in->n_chunks)) return rc;
    if (g->flags & VX_FLAG_HEATMAP) return VX_E_UNSUPPORTED;  // the heatmap runs the pair list
    const uint32_t ntiles = g->tiles[0] * g->tiles[1];
    if (ntiles > kMaxTiles) {
        std::fprintf(stderr, "vxmarch: vx_render_noalloc: %u tiles exceed its limit of %u (VX_NOALLOC_MAX_TILES)\n",
                     ntiles, kMaxTiles);
        return VX_E_OVERFLOW;
    }
    Spec s;
    if (int rc = vxh::make_spec(spec, spec_hi, g->flags, s)) return rc;
    int dev = 0;
    VX_HIPX(hipGetDevice(&dev));
    uint32_t *sc = nullptr;
    if (int rc = scratch_for(dev, &sc)) return rc;

    Pools p;
    p.inners = (const uint4 *)in->inners;
    p.leaves = (const uint2 *)in->leaves;
    p.bricks = (const uint32_t *)in->bricks;
    p.chunks = in->chunks;
    p.grid = in->grid;
    p.n_inners = (uint32_t)(in->inners_bytes / 16);
    p.n_leaves = (uint32_t)(in->leaves_bytes / 8);
    p.n_bricks = (uint32_t)(in->bricks_bytes / 4);
    p.n_chunks = in->n_chunks;
    const FrameK f = vxh::make_frame(g);
    hipStream_t st = (hipStream_t)stream;
    uint32_t *counters = sc;
    uint32_t *order = sc + kOrderOff;
    uint32_t *cost = sc + kCostOff;
    if (int rc = vxh::fill_u32(st, counters, 0, CNT_TOTAL)) return rc;
    k_iota_views<<<(ntiles + 255) / 256, 256, 0, st>>>(order, nullptr, ntiles, nullptr, nullptr);
    VX_LAUNCHED();
    TileArgs a{p, f, s, hiz, (unsigned long long *)vis, shadow, counters, order, cost};
    vxh::launch_tiles(4, VAR_FUSED_SHADOW, (f.flags & VX_FLAG_HIZ) != 0u, (f.flags & VX_FLAG_UNDERWATER) != 0u,
                      ntiles, st, a);
    VX_LAUNCHED();
    return VX_OK;
}
