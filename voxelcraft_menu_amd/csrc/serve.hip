// serve.hip -- continuous-batching view server (vx_serve_*, include/vxmarch.h).
//
// No counterpart in the reference, whose headless loop renders one view per
// Renderer::render call and blocks on it (src/headless.rs:419-497, :680-837).
//
// Requests (one view each: frame, chunk list, grid, pipeline key) are copied at submit
// and queued FIFO. A worker thread owns `lanes` launch lanes, each with its own HIP
// stream and input buffers. Whenever a lane is free it admits the queued requests in
// order while (a) the lane has view slots left and (b) the request's device footprint
// fits the HBM budget and the output arena, then launches them as one batch (the tile
// march of vx_batch, every tile of every admitted view in one launch). Requests that
// arrive while other lanes are busy are admitted into the next free lane, so their
// work overlaps on the device with the batches still running. A request's outputs
// (key and shadow word per pixel) are written straight into its own region of the
// output arena and stay there until the caller releases the request, which also
// returns its footprint to the budget.
//
// The tree pools are the context's. A barrier (vx_serve_barrier) waits for every
// request submitted before it and holds later ones back, so the caller can apply an
// edit batch to the pools (vx_pool_sync of the dirty ranges) between frames; launches
// resume with vx_serve_resume.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ctx.hpp"

using namespace vxd;
using vxh::DevBuf;
using vxh::ensure;

#define HIPX VX_HIPX

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr uint64_t PX_UNIT = 256;  // output-arena granularity, pixels

// First-fit allocator over arena units with coalescing.
struct Arena {
    std::map<uint64_t, uint64_t> free_;  // start -> length
    void init(uint64_t units) {
        free_.clear();
        if (units) free_[0] = units;
    }
    bool alloc(uint64_t n, uint64_t &off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            off = it->first;
            const uint64_t rest = it->second - n, start = it->first + n;
            free_.erase(it);
            if (rest) free_[start] = rest;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t n) {
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

enum : int { RQ_PENDING = 0, RQ_RUNNING = 1, RQ_DONE = 2 };

struct Req {
    uint64_t id = 0, seq = 0;
    ViewK vk{};
    std::vector<vx_gpu_chunk> chunks;
    std::vector<uint32_t> grid;
    bool cam = false;  // chunk list and grid built on the device from the chunk table
    ListCam lc{};
    uint32_t list_recs = 0;  // room of its list: the table size at submit
    uint32_t w = 0, h = 0, tiles = 0;
    uint64_t units = 0, px_off = 0, bytes = 0;
    double t_arrival = 0, t_submit = 0, t_launch = 0, t_done = 0;
    int state = RQ_PENDING, status = VX_OK;
    uint32_t lane = 0, batch_views = 0;
    uint64_t batch = 0;
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct Lane {
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    DevBuf in, hiz, cost, order, counters;
    void *stage = nullptr;  // pinned staging of `in`
    size_t stage_cap = 0;
    uint32_t *stat = nullptr;  // host-mapped: shadow rays of the lane's last batch
    uint32_t *stat_dev = nullptr;
    std::vector<Req *> reqs;
    uint64_t batch = 0;
    bool busy = false;
};

}  // namespace

namespace vxd {
// Sums one statistic over the spread slots into a host-mapped word.
__global__ void k_stat_sum(const uint32_t *__restrict__ counters, uint32_t k, uint32_t *__restrict__ out) {
    uint32_t v = 0;
    for (uint32_t s = threadIdx.x; s < CNT_STAT_SLOTS; s += blockDim.x)
        v += counters[CNT_STAT_BASE + s * CNT_STAT_STRIDE + k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o);
    if (threadIdx.x == 0) *out = v;
}
}  // namespace vxd

struct vx_serve {
    vx_ctx *ctx = nullptr;
    vx_serve_desc d{};
    DevBuf vis, shadow;  // output arena
    Arena arena;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req *> pending;
    std::unordered_map<uint64_t, Req *> live;
    std::deque<Req *> finished;
    std::vector<Lane> lanes;
    uint64_t used = 0, seq = 0, fence = ~0ull, batches = 0;
    bool stop = false;
    std::thread worker;
    hipStream_t io = nullptr;
    vx_serve_stats stats{};
    DevBuf table;  // LOD-0 chunk records camera requests build their lists from
    uint32_t n_table = 0;
    struct Replay *replay = nullptr;  // a running vx_serve_replay_start, if any
};

// A native request replay: the driver thread submits camera requests at their scheduled
// arrival times, the collector thread takes completions and releases their outputs.
struct Replay {
    std::vector<vx_replay_req> reqs;    // copies, their frame pointers into `frames`
    std::vector<vx_gpu_frame> frames;
    std::vector<vx_serve_result> res;
    uint64_t first_id = 0;
    double t0 = 0;
    bool keep = false;  // VX_REPLAY_KEEP: outputs stay until the caller releases them
    std::atomic<uint32_t> submitted{0}, completed{0}, failed{0};
    std::atomic<bool> driver_done{false}, stop{false};
    std::thread driver, collector;
};

namespace {

int n_cus(int dev) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return n;
}

uint32_t req_class(const Req *r) {
    return view_class((r->vk.f.flags & VX_FLAG_HIZ) != 0u, (r->vk.f.flags & VX_FLAG_UNDERWATER) != 0u);
}

// Enqueues one batch (the lane's admitted requests) on the lane stream.
int launch(vx_serve *s, Lane &L) {
    // Class-major layout: each view class is one tile range marched by its own kernel.
    std::stable_sort(L.reqs.begin(), L.reqs.end(), [](const Req *x, const Req *y) { return req_class(x) < req_class(y); });
    const uint32_t n = (uint32_t)L.reqs.size();
    size_t recs = 0, host_recs = 0;
    uint32_t tiles = 0, n_cam = 0;
    for (Req *r : L.reqs) {
        host_recs += r->chunks.size();
        recs += r->cam ? r->list_recs : r->chunks.size();
        tiles += r->tiles;
        n_cam += r->cam ? 1u : 0u;
    }
    // [descriptors | tile offsets | header | list cameras | chunk lists: the uploaded ones,
    // then room for the device-built ones | grids]
    const size_t o_views = 0, o_toff = align256(n * sizeof(ViewK)), o_hdr = align256(o_toff + (n + 1) * 4),
                 o_cams = align256(o_hdr + BH_LEN * 4), o_chunks = align256(o_cams + (size_t)n_cam * sizeof(ListCam)),
                 o_grids = align256(o_chunks + recs * sizeof(vx_gpu_chunk)),
                 total = o_grids + (size_t)n * VX_GRID_LEN * 4;
    int rc;
    if ((rc = ensure(L.in, total)) || (rc = ensure(L.hiz, (size_t)tiles * 4)) ||
        (rc = ensure(L.cost, (size_t)tiles * 4)) || (rc = ensure(L.order, (size_t)tiles * 4)))
        return rc;
    if (L.stage_cap < total) {
        if (L.stage) (void)hipHostFree(L.stage);
        L.stage = nullptr;
        L.stage_cap = 0;
        HIPX(hipHostMalloc(&L.stage, total + total / 2, hipHostMallocDefault));
        L.stage_cap = total + total / 2;
    }
    char *stg = (char *)L.stage;
    ViewK *vk = (ViewK *)(stg + o_views);
    uint32_t *toff = (uint32_t *)(stg + o_toff);
    uint32_t *hdr = (uint32_t *)(stg + o_hdr);
    vx_gpu_chunk *ch = (vx_gpu_chunk *)(stg + o_chunks);
    uint32_t *grids = (uint32_t *)(stg + o_grids);
    ListCam *cams = (ListCam *)(stg + o_cams);
    uint32_t t = 0, k_cam = 0;
    size_t c = 0, c_cam = host_recs;
    uint32_t cb[N_CLASSES + 1] = {0, 0, 0, 0, 0};
    for (uint32_t v = 0; v < n; v++) {
        const Req *r = L.reqs[v];
        const size_t off = r->cam ? c_cam : c;
        vk[v] = r->vk;
        vk[v].chunk_off = (uint32_t)off;
        vk[v].grid_off = v * VX_GRID_LEN;
        vk[v].p = vxh::make_pools(s->ctx);
        vk[v].p.chunks = (const vx_gpu_chunk *)((const char *)L.in.p + o_chunks) + off;
        vk[v].p.grid = (const uint32_t *)((const char *)L.in.p + o_grids) + (size_t)v * VX_GRID_LEN;
        vk[v].p.n_chunks = r->cam ? 0u : (uint32_t)r->chunks.size();  // device-built: set by the list kernel
        vk[v].tile_off = t;
        vk[v].px_off_lo = (uint32_t)r->px_off;
        vk[v].px_off_hi = (uint32_t)(r->px_off >> 32);
        toff[v] = t;
        if (r->cam) {
            cams[k_cam] = r->lc;
            cams[k_cam].chunk_off = (uint32_t)off;
            cams[k_cam].view = v;
            k_cam++;
            c_cam += r->list_recs;
        } else {
            if (!r->chunks.empty()) std::memcpy(ch + c, r->chunks.data(), r->chunks.size() * sizeof(vx_gpu_chunk));
            std::memcpy(grids + (size_t)v * VX_GRID_LEN, r->grid.data(), VX_GRID_LEN * 4);
            c += r->chunks.size();
        }
        t += r->tiles;
        for (uint32_t k = req_class(r) + 1; k <= N_CLASSES; k++) cb[k] = t;
    }
    toff[n] = t;
    for (uint32_t k = 0; k < BH_LEN; k++) hdr[k] = 0;
    hdr[BH_TILES] = tiles;
    hdr[BH_VIEWS] = n;
    for (uint32_t k = 0; k <= N_CLASSES; k++) hdr[BH_CB + k] = cb[k];
    hipStream_t st = L.st;
    char *din = (char *)L.in.p;
    // everything up to the uploaded lists, then the grids (device-built lists and grids are
    // written by the list kernel)
    const size_t head = o_chunks + host_recs * sizeof(vx_gpu_chunk);
    HIPX(hipMemcpyAsync(din, stg, head, hipMemcpyHostToDevice, st));
    HIPX(hipMemcpyAsync(din + o_grids, stg + o_grids, total - o_grids, hipMemcpyHostToDevice, st));
    if (n_cam) {
        k_view_lists<<<n_cam, LIST_BLOCK, 0, st>>>((const vx_gpu_chunk *)s->table.p, s->n_table,
                                                   (const ListCam *)(din + o_cams), (vx_gpu_chunk *)(din + o_chunks),
                                                   (uint32_t *)(din + o_grids), nullptr, (ViewK *)(din + o_views));
        VX_LAUNCHED();
    }
    if ((rc = vxh::fill_u32(st, L.counters.p, 0, CNT_TOTAL)) || (rc = vxh::fill_u32(st, L.hiz.p, 0, tiles)))
        return rc;
    const uint32_t *hdr_d = (const uint32_t *)(din + o_hdr);
    const uint32_t *toff_d = (const uint32_t *)(din + o_toff);
    k_iota_views<<<ORDER_BLOCKS, 256, 0, st>>>((uint32_t *)L.order.p, hdr_d + BH_TILES, 0, toff_d, hdr_d + BH_VIEWS);
    BatchArgs a;
    a.p = vxh::make_pools(s->ctx);
    a.p.chunks = (const vx_gpu_chunk *)(din + o_chunks);
    a.p.grid = (const uint32_t *)(din + o_grids);
    a.p.n_chunks = 0;
    a.views = (const ViewK *)(din + o_views);
    a.hdr = hdr_d;
    a.hiz = (uint32_t *)L.hiz.p;
    a.vis = (unsigned long long *)s->vis.p;
    a.shadow = (uint16_t *)s->shadow.p;
    a.counters = (uint32_t *)L.counters.p;
    a.order = (const uint32_t *)L.order.p;
    a.cost = (uint32_t *)L.cost.p;
    const vx_batch_tuning &tn = s->d.tuning;
    launch_batch_classes(tn.mode, tn.occupancy, tn.variant, tn.queue_chunk, tn.waves_per_simd, n_cus(s->ctx->device), cb,
                         a, st);
    k_stat_sum<<<1, 64, 0, st>>>((const uint32_t *)L.counters.p, STAT_SHADOW, L.stat_dev);
    VX_LAUNCHED();
    HIPX(hipEventRecord(L.done, st));
    return VX_OK;
}

bool lanes_idle(const vx_serve *s) {
    for (const Lane &L : s->lanes)
        if (L.busy) return false;
    return true;
}

void finish(vx_serve *s, Lane &L, int status, double t) {
    const uint32_t sr = status == VX_OK ? *(volatile uint32_t *)L.stat : 0u;
    for (Req *r : L.reqs) {
        r->t_done = t;
        r->state = RQ_DONE;
        r->status = status;
        s->finished.push_back(r);
    }
    s->stats.completed += L.reqs.size();
    s->stats.shadow_rays += sr;
    L.reqs.clear();
    L.busy = false;
}

void worker_main(vx_serve *s) {
    (void)hipSetDevice(s->ctx->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (!s->stop) {
        bool progressed = false, busy = false;
        for (Lane &L : s->lanes) {
            if (!L.busy) continue;
            hipError_t q = hipEventQuery(L.done);
            if (q == hipErrorNotReady) {
                busy = true;
                continue;
            }
            finish(s, L, q == hipSuccess ? VX_OK : VX_E_HIP, now_s());
            progressed = true;
        }
        for (uint32_t li = 0; li < s->lanes.size(); li++) {
            Lane &L = s->lanes[li];
            if (L.busy || s->pending.empty()) continue;
            while (!s->pending.empty() && L.reqs.size() < s->d.max_views) {
                Req *r = s->pending.front();
                if (r->seq >= s->fence) break;              // held back by a barrier
                if (s->used + r->bytes > s->d.hbm_budget) break;  // FIFO: wait for releases
                uint64_t off = 0;
                if (!s->arena.alloc(r->units, off)) break;
                r->px_off = off * PX_UNIT;
                s->used += r->bytes;
                r->state = RQ_RUNNING;
                r->lane = li;
                L.reqs.push_back(r);
                s->pending.pop_front();
            }
            if (L.reqs.empty()) break;
            L.busy = true;
            L.batch = ++s->batches;
            const double t = now_s();
            for (Req *r : L.reqs) {
                r->t_launch = t;
                r->batch_views = (uint32_t)L.reqs.size();
                r->batch = L.batch;
            }
            s->stats.launches++;
            s->stats.max_batch_views = std::max<uint64_t>(s->stats.max_batch_views, L.reqs.size());
            lk.unlock();
            const int rc = launch(s, L);
            lk.lock();
            if (rc) finish(s, L, rc, now_s());
            else busy = true;
            progressed = true;
        }
        if (progressed) s->cv_done.notify_all();
        if (!busy) {
            if (!progressed) s->cv_work.wait_for(lk, std::chrono::milliseconds(2));
        } else if (!progressed) {
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
        }
    }
}

void replay_wait(Replay *R) {
    if (R->driver.joinable()) R->driver.join();
    if (R->collector.joinable()) R->collector.join();
}

// Sleeps to within 0.2 ms of `target`, then spins.
void wait_until(double target) {
    for (;;) {
        const double left = target - now_s();
        if (left <= 0) return;
        if (left > 3e-4) std::this_thread::sleep_for(std::chrono::duration<double>(left - 2e-4));
        else std::this_thread::yield();
    }
}

void replay_driver(vx_serve *s, Replay *R) {
    const uint32_t n = (uint32_t)R->reqs.size();
    for (uint32_t i = 0; i < n && !R->stop; i++) {
        const double t_arr = R->t0 + R->reqs[i].t;
        wait_until(t_arr);
        int rc;
        // VX_E_BUSY: the chunk table changed between reading its size and queueing
        while ((rc = vx_serve_submit_camera(s, R->first_id + i, &R->reqs[i].view, t_arr)) == VX_E_BUSY && !R->stop)
            std::this_thread::yield();
        if (rc) {
            vx_serve_result &o = R->res[i];
            o.id = R->first_id + i;
            o.status = rc;
            o.t_arrival = t_arr;
            R->failed++;
        } else {
            R->submitted++;
        }
    }
    R->driver_done = true;
}

void replay_collector(vx_serve *s, Replay *R) {
    vx_serve_result buf[256];
    const uint64_t n = R->reqs.size();
    for (;;) {
        uint32_t k = 0;
        if (vx_serve_poll(s, buf, 256, &k, 2) != VX_OK) break;
        for (uint32_t j = 0; j < k; j++) {
            const uint64_t i = buf[j].id - R->first_id;
            if (i < n) R->res[i] = buf[j];
            if (!R->keep) (void)vx_serve_release(s, buf[j].id);
            R->completed++;
        }
        if (R->driver_done && R->completed.load() >= R->submitted.load()) break;
        if (R->stop && k == 0) {
            std::lock_guard<std::mutex> lk(s->mu);
            if (lanes_idle(s) && s->pending.empty()) break;
        }
    }
}

// Device footprint of a camera request whose list has room for a table of n records.
uint64_t cam_bytes(const Req *r, uint32_t n) {
    return r->units * PX_UNIT * 10 + (uint64_t)n * sizeof(vx_gpu_chunk) + VX_GRID_LEN * 4 + (uint64_t)r->tiles * 12 +
           sizeof(ViewK) + sizeof(ListCam);
}

void free_req(vx_serve *s, Req *r) {
    if (r->state == RQ_DONE && r->units) {
        s->arena.release(r->px_off / PX_UNIT, r->units);
        s->used -= r->bytes;
    }
    s->live.erase(r->id);
    delete r;
}

}  // namespace

extern "C" {

double vx_serve_clock(void) { return now_s(); }

int vx_serve_create(vx_ctx *c, const vx_serve_desc *d, vx_serve **out) {
    if (!c || !d || !out) return VX_E_ARG;
    *out = nullptr;
    if (c->path != VX_PATH_TILES) return VX_E_UNSUPPORTED;
    if (d->lanes < 1 || d->lanes > 8 || d->max_views < 1 || d->max_views > BATCH_MAX_VIEWS) return VX_E_ARG;
    if (d->hbm_budget < ((uint64_t)1 << 20)) return VX_E_ARG;
    const vx_batch_tuning &t = d->tuning;
    // Screen bins (VAR_BINS) need per-set candidate lists the server does not build.
    if ((t.variant & VAR_BINS) || !batch_form_ok(t.mode, t.occupancy, t.variant)) return VX_E_ARG;
    if (t.mode == VX_BATCH_PERSISTENT &&
        (t.queue_chunk < 1 || t.queue_chunk > 64 || t.waves_per_simd < 1 || t.waves_per_simd > t.occupancy))
        return VX_E_ARG;
    HIPX(hipSetDevice(c->device));
    vx_serve *s = new (std::nothrow) vx_serve();
    if (!s) return VX_E_OOM;
    s->ctx = c;
    s->d = *d;
    int rc = VX_OK;
    // The output arena takes the whole budget at 10 bytes per pixel (8-byte key, 2-byte
    // shadow word); inputs are charged against the same budget.
    const uint64_t units = d->hbm_budget / (10 * PX_UNIT);
    s->lanes.resize(d->lanes);
    do {
        if ((rc = ensure(s->vis, units * PX_UNIT * 8)) || (rc = ensure(s->shadow, units * PX_UNIT * 2))) break;
        s->arena.init(units);
        if (hipStreamCreateWithFlags(&s->io, hipStreamNonBlocking) != hipSuccess) { rc = VX_E_HIP; break; }
        for (Lane &L : s->lanes) {
            if (hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&L.done, hipEventDisableTiming) != hipSuccess) {
                rc = VX_E_HIP;
                break;
            }
            if ((rc = ensure(L.counters, CNT_TOTAL * 4))) break;
            if (hipHostMalloc((void **)&L.stat, 64, hipHostMallocMapped) != hipSuccess ||
                hipHostGetDevicePointer((void **)&L.stat_dev, L.stat, 0) != hipSuccess) {
                rc = VX_E_HIP;
                break;
            }
            *L.stat = 0;
        }
    } while (0);
    if (rc) {
        s->stop = true;
        vx_serve_destroy(s);
        return rc;
    }
    c->n_batches++;
    s->worker = std::thread(worker_main, s);
    *out = s;
    return VX_OK;
}

int vx_serve_destroy(vx_serve *s) {
    if (!s) return VX_E_ARG;
    if (s->replay) {
        s->replay->stop = true;
        replay_wait(s->replay);
        delete s->replay;
        s->replay = nullptr;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    const bool started = s->worker.joinable();
    if (started) s->worker.join();
    (void)hipSetDevice(s->ctx->device);
    for (Lane &L : s->lanes) {
        if (L.st) (void)hipStreamSynchronize(L.st);
        DevBuf *bufs[] = {&L.in, &L.hiz, &L.cost, &L.order, &L.counters};
        for (DevBuf *b : bufs) vxh::release(*b);
        if (L.stage) (void)hipHostFree(L.stage);
        if (L.stat) (void)hipHostFree(L.stat);
        if (L.done) (void)hipEventDestroy(L.done);
        if (L.st) (void)hipStreamDestroy(L.st);
        for (Req *r : L.reqs) delete r;
    }
    for (Req *r : s->pending) delete r;
    for (auto &kv : s->live)
        if (kv.second->state == RQ_DONE) delete kv.second;
    if (s->io) (void)hipStreamDestroy(s->io);
    vxh::release(s->vis);
    vxh::release(s->shadow);
    vxh::release(s->table);
    if (started) s->ctx->n_batches--;
    delete s;
    return VX_OK;
}

int vx_serve_submit(vx_serve *s, uint64_t id, const vx_view *v, double t_arrival) {
    if (!s || !v || !v->frame || !v->grid || (!v->chunks && v->n_total) || v->n_total > VX_MAX_CHUNKS)
        return VX_E_ARG;
    int rc = vxh::check_frame(v->frame, v->n_total);
    if (rc) return rc;
    Req *r = new (std::nothrow) Req();
    if (!r) return VX_E_OOM;
    r->id = id;
    r->vk.f = vxh::make_frame(v->frame);
    if ((rc = vxh::make_spec(v->spec, v->spec_hi, v->frame->flags, r->vk.s))) {
        delete r;
        return rc;
    }
    r->vk.n_chunks = v->n_total;
    r->w = v->frame->res[0];
    r->h = v->frame->res[1];
    r->tiles = v->frame->tiles[0] * v->frame->tiles[1];
    r->chunks.assign(v->chunks, v->chunks + v->n_total);
    r->grid.assign(v->grid, v->grid + VX_GRID_LEN);
    const uint64_t px = (uint64_t)r->w * r->h;
    r->units = (px + PX_UNIT - 1) / PX_UNIT;
    r->bytes = r->units * PX_UNIT * 10 + (uint64_t)v->n_total * sizeof(vx_gpu_chunk) + VX_GRID_LEN * 4 +
               (uint64_t)r->tiles * 12 + sizeof(ViewK);
    r->t_submit = now_s();
    r->t_arrival = t_arrival > 0 ? t_arrival : r->t_submit;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (r->bytes > s->d.hbm_budget) {
            delete r;
            return VX_E_OVERFLOW;
        }
        if (s->live.count(id)) {
            delete r;
            return VX_E_ARG;
        }
        r->seq = s->seq++;
        s->live[id] = r;
        s->pending.push_back(r);
        s->stats.submitted++;
    }
    s->cv_work.notify_all();
    return VX_OK;
}

int vx_serve_set_chunk_table(vx_serve *s, const vx_gpu_chunk *chunks, uint32_t n) {
    if (!s || (!chunks && n) || n > VX_CHUNK_TABLE_MAX) return VX_E_ARG;
    for (uint32_t i = 0; i < n; i++)
        if (chunks[i].lod != 0u || chunks[i].fade != 1.0f || chunks[i].voxel_size != 1.0f) return VX_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    // Only while no launch reads the table: between vx_serve_barrier and vx_serve_resume,
    // or before any camera request. Camera requests held back by the barrier have not
    // built their lists yet: they take the new table, with list room for its size.
    for (const Lane &L : s->lanes)
        if (L.busy) return VX_E_BUSY;
    for (const Req *r : s->pending)
        if (r->cam && r->seq < s->fence) return VX_E_BUSY;
    for (const Req *r : s->pending)
        if (r->cam && cam_bytes(r, n) > s->d.hbm_budget) return VX_E_OVERFLOW;
    HIPX(hipSetDevice(s->ctx->device));
    int rc;
    if ((rc = ensure(s->table, (size_t)std::max(n, 1u) * sizeof(vx_gpu_chunk)))) return rc;
    if (n) HIPX(hipMemcpyAsync(s->table.p, chunks, (size_t)n * sizeof(vx_gpu_chunk), hipMemcpyHostToDevice, s->io));
    HIPX(hipStreamSynchronize(s->io));
    s->n_table = n;
    for (Req *r : s->pending) {  // not admitted yet: their bytes are not in s->used
        if (!r->cam) continue;
        r->list_recs = n;
        r->bytes = cam_bytes(r, n);
    }
    return VX_OK;
}

int vx_serve_submit_camera(vx_serve *s, uint64_t id, const vx_camera_view *v, double t_arrival) {
    if (!s || !v || !v->frame) return VX_E_ARG;
    vx_gpu_frame fr = *v->frame;
    fr.chunk_count = 0;
    fr.camera_chunk = VX_NO_CHUNK;
    uint32_t nt;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        nt = s->n_table;
    }
    int rc = vxh::check_frame(&fr, nt);
    if (rc) return rc;
    Req *r = new (std::nothrow) Req();
    if (!r) return VX_E_OOM;
    r->id = id;
    r->vk.f = vxh::make_frame(&fr);
    if ((rc = vxh::make_spec(v->spec, v->spec_hi, fr.flags, r->vk.s))) {
        delete r;
        return rc;
    }
    r->cam = true;
    r->list_recs = nt;
    std::memcpy(r->lc.planes, v->planes, sizeof r->lc.planes);
    for (int i = 0; i < 3; i++) {
        r->lc.pos[i] = fr.cam_pos[i];
        r->lc.gmin[i] = fr.grid_min[i];
    }
    r->w = fr.res[0];
    r->h = fr.res[1];
    r->tiles = fr.tiles[0] * fr.tiles[1];
    const uint64_t px = (uint64_t)r->w * r->h;
    r->units = (px + PX_UNIT - 1) / PX_UNIT;
    r->bytes = cam_bytes(r, nt);
    r->t_submit = now_s();
    r->t_arrival = t_arrival > 0 ? t_arrival : r->t_submit;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (r->bytes > s->d.hbm_budget || s->n_table != nt) {
            const int code = r->bytes > s->d.hbm_budget ? VX_E_OVERFLOW : VX_E_BUSY;
            delete r;
            return code;
        }
        if (s->live.count(id)) {
            delete r;
            return VX_E_ARG;
        }
        r->seq = s->seq++;
        s->live[id] = r;
        s->pending.push_back(r);
        s->stats.submitted++;
    }
    s->cv_work.notify_all();
    return VX_OK;
}

int vx_serve_poll(vx_serve *s, vx_serve_result *out, uint32_t cap, uint32_t *n, int wait_ms) {
    if (!s || !n || (!out && cap)) return VX_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->finished.empty() && wait_ms > 0)
        s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s->finished.empty(); });
    uint32_t k = 0;
    while (k < cap && !s->finished.empty()) {
        Req *r = s->finished.front();
        s->finished.pop_front();
        vx_serve_result &o = out[k++];
        o.id = r->id;
        o.status = r->status;
        o.lane = r->lane;
        o.batch_views = r->batch_views;
        o.batch = r->batch;
        o.width = r->w;
        o.height = r->h;
        o.t_arrival = r->t_arrival;
        o.t_submit = r->t_submit;
        o.t_launch = r->t_launch;
        o.t_done = r->t_done;
    }
    *n = k;
    return VX_OK;
}

static int serve_read(vx_serve *s, uint64_t id, const DevBuf &buf, size_t elem, void *out, size_t n) {
    size_t off = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->live.find(id);
        if (it == s->live.end() || it->second->state != RQ_DONE) return VX_E_ARG;
        if (it->second->status != VX_OK) return it->second->status;
        if (n != (size_t)it->second->w * it->second->h) return VX_E_ARG;
        off = it->second->px_off;
    }
    HIPX(hipSetDevice(s->ctx->device));
    HIPX(hipMemcpyAsync(out, (const char *)buf.p + off * elem, n * elem, hipMemcpyDeviceToHost, s->io));
    HIPX(hipStreamSynchronize(s->io));
    return VX_OK;
}

int vx_serve_read_vis(vx_serve *s, uint64_t id, uint64_t *out, size_t n) {
    if (!s || !out) return VX_E_ARG;
    return serve_read(s, id, s->vis, 8, out, n);
}

int vx_serve_read_shadow(vx_serve *s, uint64_t id, uint16_t *out, size_t n) {
    if (!s || !out) return VX_E_ARG;
    return serve_read(s, id, s->shadow, 2, out, n);
}

int vx_serve_release(vx_serve *s, uint64_t id) {
    if (!s) return VX_E_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->live.find(id);
        if (it == s->live.end()) return VX_E_ARG;
        if (it->second->state != RQ_DONE) return VX_E_BUSY;
        free_req(s, it->second);
    }
    s->cv_work.notify_all();
    return VX_OK;
}

int vx_serve_barrier(vx_serve *s) {
    if (!s) return VX_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    s->fence = s->seq;
    s->cv_work.notify_all();
    s->cv_done.wait(lk, [&] {
        if (!lanes_idle(s)) return false;
        for (const Req *r : s->pending)
            if (r->seq < s->fence) return false;
        return true;
    });
    return VX_OK;
}

int vx_serve_resume(vx_serve *s) {
    if (!s) return VX_E_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->fence = ~0ull;
    }
    s->cv_work.notify_all();
    return VX_OK;
}

int vx_serve_replay_start(vx_serve *s, const vx_replay_req *reqs, uint32_t n, uint64_t first_id, double t0,
                          uint32_t flags) {
    if (!s || (!reqs && n) || (flags & ~VX_REPLAY_KEEP)) return VX_E_ARG;
    if (s->replay) return VX_E_BUSY;
    for (uint32_t i = 0; i < n; i++)
        if (!reqs[i].view.frame || !(reqs[i].t >= 0.0) || (i && reqs[i].t < reqs[i - 1].t)) return VX_E_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (uint32_t i = 0; i < n; i++)
            if (s->live.count(first_id + i)) return VX_E_ARG;
    }
    Replay *R = new (std::nothrow) Replay();
    if (!R) return VX_E_OOM;
    try {
        R->reqs.assign(reqs, reqs + n);
        R->frames.resize(n);
        R->res.assign(n, vx_serve_result{});
    } catch (...) {
        delete R;
        return VX_E_OOM;
    }
    for (uint32_t i = 0; i < n; i++) {
        R->frames[i] = *reqs[i].view.frame;
        R->reqs[i].view.frame = &R->frames[i];
    }
    R->first_id = first_id;
    R->keep = (flags & VX_REPLAY_KEEP) != 0u;
    R->t0 = t0 > 0 ? t0 : now_s();
    s->replay = R;
    R->collector = std::thread(replay_collector, s, R);
    R->driver = std::thread(replay_driver, s, R);
    return VX_OK;
}

int vx_serve_replay_progress(vx_serve *s, uint32_t *submitted, uint32_t *completed) {
    if (!s || !s->replay) return VX_E_ARG;
    if (submitted) *submitted = s->replay->submitted + s->replay->failed;
    if (completed) *completed = s->replay->completed;
    return VX_OK;
}

int vx_serve_replay_join(vx_serve *s, vx_serve_result *out, uint32_t n) {
    if (!s || !s->replay || (!out && n)) return VX_E_ARG;
    Replay *R = s->replay;
    if (n != R->reqs.size()) return VX_E_ARG;
    replay_wait(R);
    if (n) std::memcpy(out, R->res.data(), (size_t)n * sizeof(vx_serve_result));
    delete R;
    s->replay = nullptr;
    return VX_OK;
}

int vx_serve_stats_get(vx_serve *s, vx_serve_stats *out) {
    if (!s || !out) return VX_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->pending = s->pending.size();
    out->used_bytes = s->used;
    out->budget_bytes = s->d.hbm_budget;
    return VX_OK;
}

}  // extern "C"
