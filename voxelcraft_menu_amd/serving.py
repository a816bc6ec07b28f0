"""Open-loop request stream for the view server: Poisson view requests and edit batches.

The reference's request loop (src/headless.rs:419-497) takes one view at a time: build
the chunk list, render, wait. Here requests arrive on a seeded Poisson clock whether or
not earlier ones have finished; each is stamped with its scheduled arrival time, so
latency = completion - arrival includes every host and queueing delay.

  view request  by default the camera alone (Server.submit_camera, under a shared lock):
                the server builds its chunk list and grid on the device from the chunk
                table of the current snapshot; with device_lists off, a host chunk list
                from the snapshot (ViewSnapshot.build_view) and Server.submit. The native
                server batches it with whatever else is queued and launches it on a free
                lane.
  edit batch    1000 random voxel set/clear ops, on their own thread: applied to the
                host world while views keep being set up from the snapshot and rendered
                from the device pools, both still the world before the edit; a new
                snapshot is taken; then, under the exclusive lock, barrier -> dirty-range
                pool upload -> snapshot swap (and its chunk table to the server) -> resume. Views submitted before the swap
                see the old world, views after it the new one.

A poller thread collects completions and releases them (returning their HBM budget).

With device lists (the default) the arrival path is native: the server's replay driver
thread submits every camera request at its scheduled time and its collector thread
releases completions (vx_serve_replay_start), so Python only prepares the frames up
front and applies the edit batches, which a barrier serialises with the launches.
Requests that arrive while an edit holds the barrier are queued and render the edited
world. sweep() runs the stream at increasing rates and reports the highest rate whose
p99 latency stays within a bound.
"""
from __future__ import annotations

import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _abi as A
from . import workloads as W
from .renderer import Renderer, Server
from .world import ViewSnapshot, World

EDIT_IDS = [0, 1, 2, 5, 8, 11]  # air, stone, dirt, planks, glass, glowstone


class _RWLock:
    def __init__(self):
        self._c = threading.Condition()
        self._readers = 0
        self._writer = False

    def acquire_read(self):
        with self._c:
            while self._writer:
                self._c.wait()
            self._readers += 1

    def release_read(self):
        with self._c:
            self._readers -= 1
            if self._readers == 0:
                self._c.notify_all()

    def acquire_write(self):
        with self._c:
            while self._writer:
                self._c.wait()
            self._writer = True
            while self._readers:
                self._c.wait()

    def release_write(self):
        with self._c:
            self._writer = False
            self._c.notify_all()


def random_edits(rng, wl: W.Workload, n: int = 1000) -> np.ndarray:
    e = np.zeros(n, A.EDIT_DTYPE)
    g = wl.gen
    e["x"] = rng.integers(g.chunk_min[0] * 64, g.chunk_max[0] * 64, n)
    e["y"] = rng.integers(g.chunk_min[1] * 64, g.chunk_max[1] * 64, n)
    e["z"] = rng.integers(g.chunk_min[2] * 64, g.chunk_max[2] * 64, n)
    e["id"] = rng.choice(EDIT_IDS, n)
    return e


def _busy_fraction(intervals, t_start: float, t_end: float) -> float:
    """Fraction of [t_start, t_end] covered by the union of the (launch, done) intervals."""
    covered, cur_s, cur_e = 0.0, None, None
    for a, b in sorted(intervals):
        a, b = max(a, t_start), min(b, t_end)
        if b <= a:
            continue
        if cur_e is None or a > cur_e:
            if cur_e is not None:
                covered += cur_e - cur_s
            cur_s, cur_e = a, b
        else:
            cur_e = max(cur_e, b)
    if cur_e is not None:
        covered += cur_e - cur_s
    return covered / max(t_end - t_start, 1e-9)


def _summary(qps, duration_s, lat_ms, queue_ms, batches, edit_ms, st, t0, t_end, total_px, lanes, hbm_budget):
    l, q = np.asarray(lat_ms), np.asarray(queue_ms)
    rays = total_px + st["shadow_rays"]
    span = t_end - t0
    return {
        "qps_offered": qps, "duration_s": duration_s, "requests": int(l.size),
        "qps_completed": round(l.size / span, 1),
        "p50_ms": round(float(np.percentile(l, 50)), 3), "p99_ms": round(float(np.percentile(l, 99)), 3),
        "mean_ms": round(float(l.mean()), 3), "max_ms": round(float(l.max()), 3),
        "queue_p50_ms": round(float(np.percentile(q, 50)), 3), "queue_p99_ms": round(float(np.percentile(q, 99)), 3),
        "edit_batches": len(edit_ms), "edit_ops_per_batch": 1000,
        "edit_apply_ms_mean": round(float(np.mean(edit_ms)), 3) if edit_ms else None,
        "batches": len(batches), "mean_batch_views": round(float(np.mean([v[0] for v in batches.values()])), 2),
        "max_batch_views": int(st["max_batch_views"]), "lanes": lanes, "hbm_budget_bytes": hbm_budget,
        "lane_busy": round(_busy_fraction([(v[1], v[2]) for v in batches.values()], t0, t_end), 3),
        "Grays_per_s": round(rays / span / 1e9, 4), "primary_rays": int(total_px), "shadow_rays": int(st["shadow_rays"]),
    }


def _stream(wl: W.Workload, qps: float, duration_s: float, edit_rate: float, seed: int):
    rng = np.random.default_rng(seed)
    n_views = max(1, int(qps * duration_s * 1.5) + 16)
    t_views = np.cumsum(rng.exponential(1.0 / qps, n_views))
    t_views = t_views[t_views < duration_s]
    t_edits = np.cumsum(rng.exponential(1.0 / edit_rate, max(2, int(edit_rate * duration_s * 2) + 4))) \
        if edit_rate > 0 else np.zeros(0)
    t_edits = t_edits[t_edits < duration_s]
    cams = W.random_cameras(wl, len(t_views), seed=seed)
    edits = [random_edits(rng, wl) for _ in range(len(t_edits))]
    return t_views, cams, t_edits, edits


def run_native(world: World, renderer: Renderer, wl: W.Workload, qps: float, duration_s: float, edit_rate: float,
               seed: int = 4, lanes: int = 3, max_views: int = 32, hbm_budget: int = 8 << 30,
               mode: int = A.BATCH_FLAT, occupancy: int = 6, variant: int = 5, log=None) -> dict:
    """The stream with camera requests submitted by the server's native replay driver;
    edit batches are applied from this thread at their scheduled times."""
    t_views, cams, t_edits, edits = _stream(wl, qps, duration_s, edit_rate, seed)
    srv = Server(renderer, lanes=lanes, max_views=max_views, hbm_budget=hbm_budget, mode=mode, occupancy=occupancy,
                 variant=variant)
    snaps = [ViewSnapshot(world), ViewSnapshot()]  # [current, spare]
    srv.set_chunk_table(snaps[0].chunk_table())
    specs = [W.spec_words(c.flags, snaps[0]) for c in cams]
    t0 = srv.replay_start(t_views, cams, specs, first_id=0)
    edit_ms = []
    for t, e in zip(t_edits, edits):
        target = t0 + float(t)
        while (left := target - srv.clock()) > 0:
            time.sleep(min(left, 0.002))
        te = srv.clock()
        world.apply_edits(e)  # views meanwhile render from the device pools, the world before the edit
        snaps[1].take(world)
        srv.barrier()
        renderer.sync_world(world)
        snaps[0], snaps[1] = snaps[1], snaps[0]
        srv.set_chunk_table(snaps[0].chunk_table())
        srv.resume()
        edit_ms.append((srv.clock() - te) * 1e3)
    res = srv.replay_join()
    st = srv.stats()
    srv.close()
    for sn in snaps:
        sn.close()
    bad = [r for r in res if r["status"] != 0]
    if bad:
        raise RuntimeError(f"{len(bad)} requests failed, first: {bad[0]}")
    lat = [(r["t_done"] - r["t_arrival"]) * 1e3 for r in res]
    queue = [(r["t_launch"] - r["t_arrival"]) * 1e3 for r in res]
    batches = {}
    for r in res:
        batches[r["batch"]] = (r["batch_views"], r["t_launch"], r["t_done"])
    t_end = max(r["t_done"] for r in res)
    total_px = sum(c.width * c.height for c in cams)
    out = _summary(qps, duration_s, lat, queue, batches, edit_ms, st, t0, t_end, total_px, lanes, hbm_budget)
    out["arrival_path"] = "native replay driver"
    out["tuning"] = {"mode": mode, "occupancy": occupancy, "variant": variant, "max_views": max_views}
    out["submit_late_p99_ms"] = round(float(np.percentile([(r["t_submit"] - r["t_arrival"]) * 1e3 for r in res], 99)), 3)
    if log:
        log(f"serving @ {qps} qps: {out}")
    return out


def sweep(world: World, renderer: Renderer, wl: W.Workload, rates, duration_s: float, edit_rate: float,
          p99_bound_ms: float = 5.0, seed: int = 4, log=None, **kw) -> dict:
    """run_native at each rate in increasing order, stopping after the first rate whose
    p99 exceeds the bound; reports every point and the highest rate within it."""
    points = []
    best = None
    for q in sorted(rates):
        r = run_native(world, renderer, wl, q, duration_s, edit_rate, seed=seed, log=log, **kw)
        points.append({k: r[k] for k in ("qps_offered", "qps_completed", "requests", "p50_ms", "p99_ms",
                                         "mean_batch_views", "lane_busy", "Grays_per_s", "submit_late_p99_ms")})
        if r["p99_ms"] <= p99_bound_ms:
            best = r
        else:
            break
    return {"p99_bound_ms": p99_bound_ms, "points": points,
            "max_qps_within_bound": best["qps_offered"] if best else None,
            "at_max": {k: best[k] for k in ("p50_ms", "p99_ms", "mean_batch_views", "lane_busy", "Grays_per_s")}
            if best else None}


def run(world: World, renderer: Renderer, wl: W.Workload, qps: float, duration_s: float, edit_rate: float,
        seed: int = 4, workers: int = 6, lanes: int = 3, max_views: int = 32, hbm_budget: int = 8 << 30,
        mode: int = A.BATCH_FLAT, device_lists: bool = True, log=None) -> dict:
    """Runs the stream for duration_s seconds of arrivals and returns latency and
    throughput figures of the completed view requests. Camera requests (device lists)
    go through the native replay driver; host-list requests through a Python arrival
    loop and thread pool."""
    if device_lists:
        return run_native(world, renderer, wl, qps, duration_s, edit_rate, seed=seed, lanes=lanes,
                          max_views=max_views, hbm_budget=hbm_budget, mode=mode, log=log)
    rng = np.random.default_rng(seed)
    n_views = max(1, int(qps * duration_s * 1.5) + 16)
    t_views = np.cumsum(rng.exponential(1.0 / qps, n_views))
    t_views = t_views[t_views < duration_s]
    t_edits = np.cumsum(rng.exponential(1.0 / edit_rate, max(2, int(edit_rate * duration_s * 2) + 4))) \
        if edit_rate > 0 else np.zeros(0)
    t_edits = t_edits[t_edits < duration_s]
    cams = W.random_cameras(wl, len(t_views), seed=seed)
    edits = [random_edits(rng, wl) for _ in range(len(t_edits))]
    events = sorted([(float(t), 0, i) for i, t in enumerate(t_views)] + [(float(t), 1, i) for i, t in enumerate(t_edits)])

    srv = Server(renderer, lanes=lanes, max_views=max_views, hbm_budget=hbm_budget, mode=mode)
    lock = _RWLock()
    snaps = [ViewSnapshot(world), ViewSnapshot()]  # [current, spare]
    if device_lists:
        srv.set_chunk_table(snaps[0].chunk_table())
    lat = {}
    px = {}
    batches = {}
    errors = []
    done_evt = threading.Event()
    submitted = [0]
    sub_lock = threading.Lock()

    def poller():
        while True:
            for r in srv.poll(wait_ms=5):
                if r["status"] != 0:
                    errors.append(r)
                lat[r["id"]] = (r["t_done"] - r["t_arrival"], r["t_launch"] - r["t_arrival"], r["t_done"])
                batches[r["batch"]] = r["batch_views"]
                srv.release(r["id"])
            if done_evt.is_set() and len(lat) >= submitted[0]:
                return

    def prep_submit(i: int, t_arr: float):
        lock.acquire_read()
        try:
            cam = cams[i]
            snap = snaps[0]
            spec, hi = W.spec_words(cam.flags, snap)
            if device_lists:
                srv.submit_camera(i, cam, spec, hi, t_arr)
            else:
                srv.submit(i, snap.build_view(cam), spec, hi, t_arr)
            px[i] = cam.width * cam.height
            with sub_lock:
                submitted[0] += 1
        finally:
            lock.release_read()

    edit_ms = []

    def apply_edit(e: np.ndarray):
        te = srv.clock()
        world.apply_edits(e)  # views meanwhile come from snaps[0] and the device pools
        snaps[1].take(world)
        lock.acquire_write()
        try:
            srv.barrier()
            renderer.sync_world(world)
            snaps[0], snaps[1] = snaps[1], snaps[0]
            if device_lists:
                srv.set_chunk_table(snaps[0].chunk_table())
            srv.resume()
        finally:
            lock.release_write()
        edit_ms.append((srv.clock() - te) * 1e3)

    pt = threading.Thread(target=poller, daemon=True)
    pt.start()
    futs = []
    edit_futs = []
    # edits in order on one thread of their own, so the arrival loop never waits on one
    with ThreadPoolExecutor(max_workers=workers) as pool, ThreadPoolExecutor(max_workers=1) as editor:
        t0 = srv.clock() + 0.05
        for t, kind, i in events:
            target = t0 + t
            while True:
                now = srv.clock()
                if now >= target:
                    break
                time.sleep(min(target - now, 0.0005))
            if kind == 0:
                futs.append(pool.submit(prep_submit, i, target))
            else:
                edit_futs.append(editor.submit(apply_edit, edits[i]))
        for f in futs + edit_futs:
            f.result()
    done_evt.set()
    pt.join()
    t_end = max(v[2] for v in lat.values())
    st = srv.stats()
    srv.close()
    for sn in snaps:
        sn.close()
    if errors:
        raise RuntimeError(f"{len(errors)} requests failed, first: {errors[0]}")
    l = np.array([v[0] for v in lat.values()]) * 1e3
    q = np.array([v[1] for v in lat.values()]) * 1e3
    span = t_end - t0
    total_px = sum(px.values())
    rays = total_px + st["shadow_rays"]
    out = {
        "qps_offered": qps, "duration_s": duration_s, "requests": len(lat),
        "p50_ms": round(float(np.percentile(l, 50)), 3), "p99_ms": round(float(np.percentile(l, 99)), 3),
        "mean_ms": round(float(l.mean()), 3), "max_ms": round(float(l.max()), 3),
        "queue_p50_ms": round(float(np.percentile(q, 50)), 3), "queue_p99_ms": round(float(np.percentile(q, 99)), 3),
        "edit_batches": len(edit_ms), "edit_ops_per_batch": 1000,
        "edit_apply_ms_mean": round(float(np.mean(edit_ms)), 3) if edit_ms else None,
        "batches": len(batches), "mean_batch_views": round(float(np.mean(list(batches.values()))), 2),
        "max_batch_views": int(st["max_batch_views"]), "lanes": lanes, "hbm_budget_bytes": hbm_budget,
        "Grays_per_s": round(rays / span / 1e9, 4), "primary_rays": int(total_px), "shadow_rays": int(st["shadow_rays"]),
    }
    if log:
        log(f"serving @ {qps} qps: {out}")
    return out
