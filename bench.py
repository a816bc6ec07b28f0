#!/usr/bin/env python3
"""Benchmark: Grays/s (primary + sun-shadow rays) of the voxel ray march on one MI355X.

Default workload (--workload 2): the 1024x1024x256 world (seed 7) seen by 64 cameras of
uneven resolution (uniform over 256x144, 640x360, 1280x720, 1920x1080; seeded poses
20-60 blocks above the terrain), all rendered per step: primary DDA, one sun shadow ray
per lit pixel, two-phase Hi-Z on 8x8 tiles. A step is one frame of every view. The
batched form (the launch form picked by the autotuner, replayed from a hipGraph) is
`value`; the same views rendered with one dispatch per view (inputs equally resident)
are reported beside it as `plain`, with the ratio. Every view keeps its Hi-Z words
between steps, so steps after the first run in the Hi-Z steady state.

The same line carries `incoherent` (config 3: 32 views at 1280x720 of the 2048x2048x512
world, waves on, with one secondary bounce per water / glass hit) and `serving`: p50/p99 request latency and Grays/s of an open-loop
Poisson stream of view requests (1000/s by default, the four resolutions) mixed with
edit batches of 1000 voxel ops (2/s) on the device-resident 2048x2048x512 world,
through the continuous-batching server.

--workload 1 measures one 1920x1080 camera per step instead; --workload 0 the small
640x360 case. Inputs (pools, chunk lists, grids, frames) are in HBM before the timed
region starts.

Output: one JSON line with the metric, a roofline object for the dominant kernel and a
cpu_baseline object (the C oracle timed on the host cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# MI355X published FP32 vector peak, 157.3 TFLOP/s: 78.65 T lane FMAs/s, i.e. two wave64
# VALU instructions per 4 cycles per SIMD (1024 SIMDs at 2.4 GHz, dual issue).
VALU_PEAK_TLANEOPS = 157.3 / 2.0
METRIC = "Grays/s primary+shadow, batched vs per-view plain; p50/p99 view latency at QPS"


def _fail(msg: str, code: int = 2):
    print(msg, file=sys.stderr, flush=True)
    sys.exit(code)


def _log(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def profile_figures(kernel: str, workload: str) -> dict:
    """The `name: value` lines of the committed rocprofv3 --pmc summary of a kernel
    (per-launch counters: traffic_bytes_per_launch, valu_insts_per_launch, ...), or {}."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_pmc.md")
    out = {}
    if not os.path.exists(path):
        return out
    with open(path) as fh:
        for line in fh:
            if ":" in line and not line.startswith((" ", "|", "#")):
                k, v = line.split(":", 1)
                if k.replace("_", "").isalnum():
                    out[k.strip()] = v.strip()
    return out


def traffic_from_profiles(kernel: str, workload: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if present."""
    try:
        return float(profile_figures(kernel, workload)["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def valu_probe():
    """Measured VALU issue ceiling: wave64 VALU instructions per second of a grid that fills
    every SIMD with independent 32-bit VALU chains (libvxcalib.so), f32 FMA and integer."""
    import ctypes as C

    from voxelcraft_menu_amd import _abi as A
    L = A.calib()
    blocks, iters = 2048, 40000  # 8192 waves of 64 lanes: eight per SIMD on 256 CUs
    res = {}
    for kind, name in ((0, "fma_f32"), (1, "int32")):
        ms, per_wave = C.c_double(), C.c_double()
        if L.vxc_valu_probe(kind, blocks, iters, 3, C.byref(ms), C.byref(per_wave)) != 0:
            return None
        res[name] = blocks * 4 * per_wave.value / (ms.value * 1e-3) / 1e12  # T wave64 instructions / s
    return {k: round(v * 64.0, 2) for k, v in res.items()}  # T lane-ops / s


def _threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(n, 64))


def cpu_baseline(world, views, specs, budget_s: float):
    """The oracle (plain-C restatement of the WGSL passes) on this host's cores.

    First runs one frame of every view (Hi-Z from zero) and, per view, the steady-state
    frame split into the primary pass and the shadow pass to get the exact algorithmic
    byte counts; then times whole steady-state frames view after view until budget_s
    has elapsed. Returns (cpu_baseline object, summed primary stats, summed shadow stats)."""
    from oracle import oracle as O

    threads = _threads()
    scenes, hizs = [], []
    prim = {"node_bytes": 0}
    sh = {"node_bytes": 0, "rays_shadow": 0}
    for v, (spec, spec_hi) in zip(views, specs):
        sc = O.Scene(world.pool(0), world.pool(1), world.pool(2), v.chunks, v.grid)
        sp = O.spec_from(spec, spec_hi)
        hz = np.zeros(v.tiles[0] * v.tiles[1], np.uint32)
        _, _, hz, _ = O.render(sc, v.frame_bytes, sp, hz, shadow=False, threads=threads)
        vis, _, _, p = O.render(sc, v.frame_bytes, sp, hz, shadow=False, threads=threads)
        _, s = O.primary_shadow(sc, v.frame_bytes, sp, vis, threads=threads)
        prim["node_bytes"] += p["node_bytes"]
        sh["node_bytes"] += s["node_bytes"]
        sh["rays_shadow"] += s["rays_shadow"]
        scenes.append((sc, sp))
        hizs.append(hz)
    _log("cpu baseline: byte counts done")
    nviews, rays, t0 = 0, 0, time.perf_counter()
    while True:
        i = nviews % len(views)
        v = views[i]
        sc, sp = scenes[i]
        _, _, hizs[i], st = O.render(sc, v.frame_bytes, sp, hizs[i], shadow=True, threads=threads)
        nviews += 1
        rays += int(v.frame.res[0]) * int(v.frame.res[1]) + st["rays_shadow"]
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    sizes = sorted({(int(v.frame.res[0]), int(v.frame.res[1])) for v in views})
    return {
        "value": rays / dt / 1e9, "unit": "Grays/s", "cores": threads, "kind": "port",
        "sample": f"{nviews} view-frames of the same {len(views)} views (sizes {sizes}; primary + shadow + "
                  f"Hi-Z, steady state) in {dt:.1f} s",
    }, prim, sh


def _time_batch(b, steps):
    b.sync()
    b.timings_reset()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.render()
    b.sync()
    dt = time.perf_counter() - t0
    tot, mar, n = b.timings_sum()
    return dt, tot / max(n, 1), mar / max(n, 1)


def run_batch(args, out):
    from voxelcraft_menu_amd import _abi as A
    from voxelcraft_menu_amd import workloads as W
    from voxelcraft_menu_amd.renderer import Batch, Renderer, device_name
    from voxelcraft_menu_amd.world import build_view

    wl = W.WORKLOADS[2]
    world = W.make_world(wl)
    cams = W.random_cameras(wl, wl.views, seed=2)
    views = [build_view(world, c) for c in cams]
    specs = [W.spec_words(c.flags, world) for c in cams]
    px = sum(int(v.frame.res[0]) * int(v.frame.res[1]) for v in views)
    tiles = sum(v.tiles[0] * v.tiles[1] for v in views)
    _log(f"{len(views)} views, {px} pixels")
    r = Renderer(256, 144)
    r.sync_world(world)
    b = Batch(r, len(views))
    if args.pool_layout == "flat":
        # Every measurement of this run on the per-chunk block layout (for counter runs).
        table, flat, fviews = flat_views(world, views)
        r.set_flat_pools(flat)
        views = fviews
    b.set_views(views, specs)
    best, best_ms = b.autotune(3)
    _log(f"autotuned: {best} ({best_ms:.3f} ms)")
    for _ in range(max(args.warmup, 1)):
        b.render()
    b.sync()
    d = b.demand()
    rays_per_step = px + d["shadow_rays"]

    dt, _, _ = _time_batch(b, args.steps)
    value = rays_per_step * args.steps / dt / 1e9
    _log(f"batched: {dt / args.steps * 1e3:.3f} ms/step")

    # The march kernel alone (HIP events around it on the context stream): the same
    # launch form without graph replay, a few frames.
    tuned = dict(best)
    kt = dict(best)
    kt["graph"] = 0
    b.set_tuning(**kt)
    _, _, march_ms = _time_batch(b, min(args.steps, 10))

    # Per-view plain: one dispatch per view (cost-order kernels + the single-view tile
    # march), the same resident inputs, no host synchronisation between views; the
    # fastest single-view kernel form of a short trial.
    plain_best = None
    for occ, var in ((4, 4), (4, 5), (6, 5)):
        b.set_tuning(mode=A.BATCH_PLAIN, occupancy=occ, variant=var, ordered=1, graph=0)
        b.render()
        t_try, _, _ = _time_batch(b, 3)
        if plain_best is None or t_try < plain_best[0]:
            plain_best = (t_try, occ, var)
    b.set_tuning(mode=A.BATCH_PLAIN, occupancy=plain_best[1], variant=plain_best[2], ordered=1, graph=0)
    b.render()
    pdt, _, _ = _time_batch(b, args.steps)
    plain = rays_per_step * args.steps / pdt / 1e9
    _log(f"plain: {pdt / args.steps * 1e3:.3f} ms/step")
    b.set_tuning(**tuned)
    layout_ab = None if args.no_layout_ab else run_layout_ab(args, r, b, world, views, specs, tuned)
    b.close()

    kname = {A.BATCH_FLAT: "k_batch_flat", A.BATCH_PERSISTENT: "k_batch_persistent"}.get(best["mode"], "k_march_tiles")
    out.update({
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Grays/s",
        "ms_per_step": round(dt / args.steps * 1e3, 4),
        "plain": {"value": round(plain, 4), "ms_per_step": round(pdt / args.steps * 1e3, 4),
                  "form": "one dispatch per view (cost-order kernels + tile march), inputs resident",
                  "tuning": {"occupancy": plain_best[1], "variant": plain_best[2]}},
        "batched_over_plain": round(value / plain, 3),
        "config": {
            "workload": wl.name, "views": len(views), "resolutions": "uniform over " + str(W.RESOLUTIONS),
            "camera_seed": 2, "world_chunks": [wl.gen.chunk_max[i] - wl.gen.chunk_min[i] for i in range(3)],
            "seed": wl.gen.seed, "hiz": True, "primary_rays": px, "shadow_rays": d["shadow_rays"],
            "tiles": tiles, "tuning": tuned, "device": device_name(),
        },
    })

    if layout_ab:
        out["pool_layout"] = layout_ab
    if not args.no_moving and args.pool_layout == "reference":  # its device lists carry reference pointers
        out["moving_camera"] = run_moving(args, world, cams, specs, tuned, px)

    # Single 1920x1080 view (the plain-at-scale case), for reference.
    wl1 = W.WORKLOADS[1]
    cam1 = W.default_camera(wl1)
    v1 = build_view(world, cam1)
    s1 = W.spec_words(cam1.flags, world)
    r1 = Renderer(wl1.width, wl1.height)
    r1.sync_world(world)
    r1.set_view(v1)
    for _ in range(3):
        r1.render(v1.frame, *s1)
    r1.sync()
    sr = r1.pair_demand()["shadow_rays"]
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r1.render(v1.frame, *s1)
    r1.sync()
    dt1 = time.perf_counter() - t0
    out["single_view_1080p"] = {"value": round((wl1.width * wl1.height + sr) * args.steps / dt1 / 1e9, 4),
                                "ms_per_step": round(dt1 / args.steps * 1e3, 4)}
    out["resolve_1080p"] = run_resolve(r1, world, wl1, args)
    r1.close()
    r.close()

    if not args.no_serving:
        out["serving"] = run_serving(args)
    if not args.no_incoherent:
        out["incoherent"] = run_incoherent(args)
    if not args.no_build:
        out["chunk_build"] = run_chunk_build(args)
    if not args.no_view_setup:
        out["view_setup"] = run_view_setup(args)
    if not args.no_device_edits:
        out["device_edits"] = run_device_edits(args)

    if args.no_cpu_baseline:
        out["cpu_baseline"] = None
        out["roofline"] = None
        return
    cb, prim, sh = cpu_baseline(world, views, specs, args.cpu_budget)
    out["cpu_baseline"] = cb
    # Algorithmic bytes of one batch launch: the oracle's exact per-(ray, chunk) node
    # accounting of the primary and shadow rays of every view, plus per pixel the 8-byte
    # key and 2-byte R16F shadow word, plus the 4-byte Hi-Z word per tile.
    alg = prim["node_bytes"] + sh["node_bytes"] + 10 * px + 4 * tiles
    out["roofline"] = roofline_valu(kname, "workload2", march_ms, alg, px)


def flat_views(world, views):
    """The world's pools laid out one block per chunk (world.flat_layout) and the views
    with their chunk records moved into it."""
    from voxelcraft_menu_amd.world import ViewSnapshot, flat_layout, relocate_view

    snap = ViewSnapshot(world)
    table = snap.chunk_table()
    snap.close()
    flat, moved = flat_layout(world.pool(0), world.pool(1), world.pool(2), table)
    return table, flat, [relocate_view(v, table, moved) for v in views]


def run_layout_ab(args, r, b, world, views, specs, tuned) -> dict:
    """The tuned batch on the reference pools and on the per-chunk block layout, the same
    views and launch form, alternating in one process; keys and shadow words of a sample
    of views are compared between the two."""
    table, flat, fviews = flat_views(world, views)
    sample = sorted({0, len(views) // 3, 2 * len(views) // 3, len(views) - 1})
    ms = {"reference": [], "flat": []}
    ref = {}
    same = True
    for rep in range(3):
        for name, vs_, words in (("reference", views, None), ("flat", fviews, flat)):
            r.set_flat_pools(words)
            b.set_views(vs_, specs)
            b.set_tuning(**tuned)
            for _ in range(3):
                b.render()
            dt, _, _ = _time_batch(b, args.steps)
            ms[name].append(dt / args.steps * 1e3)
            if rep == 0:
                for i in sample:
                    o = (b.read_vis(i), b.read_shadow(i))
                    if name == "reference":
                        ref[i] = o
                    else:
                        same &= bool(np.array_equal(o[0], ref[i][0]) and np.array_equal(o[1], ref[i][1]))
    r.set_flat_pools(None)
    b.set_views(views, specs)
    b.set_tuning(**tuned)
    pool_bytes = sum(int(world.pool(k).nbytes) for k in range(3))
    res = {"reference_ms": round(float(np.median(ms["reference"])), 4), "flat_ms": round(float(np.median(ms["flat"])), 4),
           "reference_pool_bytes": pool_bytes, "flat_bytes": int(flat.nbytes), "chunks": int(table.shape[0]),
           "identical_on_sampled_views": same, "sampled_views": sample,
           "form": "per-chunk block (L1 nodes, leaf runs, attribute table, palettes), 128-byte aligned"}
    _log(f"pool layout: reference {res['reference_ms']} ms, flat {res['flat_ms']} ms per step, identical {same}")
    if not same:
        _fail("pool layout A/B: the block layout changed keys or shadow words")
    return res


def roofline_valu(kname: str, workload: str, kernel_ms: float, alg: float, px: int) -> dict:
    """The march kernel is VALU-issue bound (integer mask logic and f32 stepping, no dense
    contraction; the ~20 MB tree is served from L2 and the Infinity Cache): its roofline is
    VALU lane-operations per second against the FP32 vector peak. The VALU instruction
    count per launch is the committed rocprofv3 SQ_INSTS_VALU figure for this workload,
    valid only for the kernel build it was measured on (the digest of the tile kernels'
    sources and flags is recorded with it); the kernel time is measured here with HIP events."""
    from voxelcraft_menu_amd import build as B

    prof = profile_figures(kname, workload)
    kernel_s = kernel_ms * 1e-3
    valu = float(prof["valu_insts_per_launch"]) if "valu_insts_per_launch" in prof else None
    same_build = prof.get("library_digest") == B.source_digest("tile_kernels")
    achieved = valu * 64.0 / kernel_s / 1e12 if valu and same_build else None
    traffic = float(prof["traffic_bytes_per_launch"]) if "traffic_bytes_per_launch" in prof else None
    probe = valu_probe()
    return {
        "bound": "valu", "achieved": round(achieved, 2) if achieved else None, "peak": VALU_PEAK_TLANEOPS,
        "unit": "T VALU lane-ops/s (wave64 VALU instructions x 64)",
        "frac": round(achieved / VALU_PEAK_TLANEOPS, 4) if achieved else None,
        "traffic": traffic, "kernel": kname, "kernel_ms": round(kernel_ms, 4),
        "valu_insts_per_launch": valu, "profile_matches_build": same_build,
        "peak_source": "MI355X FP32 vector 157.3 TFLOP/s = 78.65 T lane FMAs/s (2 wave64 VALU issues / 4 cycles / SIMD)",
        "measured_valu_ceiling": probe,
        "lane_utilization": float(prof["lane_utilization"]) if "lane_utilization" in prof else None,
        "hbm_GBps": round(traffic / kernel_s / 1e9, 1) if traffic else None,
        "hbm_frac": round(traffic / kernel_s / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
        "cache_served_node_bytes_per_launch": int(alg),
        "cache_served_node_bytes_note": "the oracle's count for the reference's unpruned pair decomposition; "
                                        "served from L2 / Infinity Cache, not an HBM figure",
        "cache_served_bytes_per_pixel": round(alg / px, 2),
    }


def run_moving(args, world, cams, specs, tuned, px):
    """The same 64 views with every camera turning by 0.01 rad of yaw per step, as the
    reference's bench_frames turns its camera each frame (src/headless.rs:718-720). A step
    rebuilds every view's chunk list and grid on the device for its moved camera
    (vx_batch_move_cameras: Hi-Z words and tile costs carried from the previous frame,
    screen bins rebuilt) and renders the batch in the tuned form."""
    import dataclasses

    from voxelcraft_menu_amd.renderer import Batch, Renderer
    from voxelcraft_menu_amd.world import ViewSnapshot

    r = Renderer(256, 144)
    r.sync_world(world)
    b = Batch(r, len(cams))
    b.set_chunk_table(ViewSnapshot(world).chunk_table())
    b.set_tuning(**tuned)
    b.set_cameras(cams, specs)

    def turned(k):
        return [dataclasses.replace(c, yaw=c.yaw + 0.01 * k) for c in cams]

    k = 0
    for _ in range(max(args.warmup, 2)):
        k += 1
        b.set_cameras(turned(k), specs, moved=True)
        b.render()
    b.sync()
    rays = rec = dfr = 0
    b.timings_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        k += 1
        b.set_cameras(turned(k), specs, moved=True)
        b.render()
        d = b.demand()
        rays += px + d["shadow_rays"]
        rec += d["recover"]
        dfr += d["deferred"]
    dt = time.perf_counter() - t0
    tot, _, n = b.timings_sum()
    b.close()
    r.close()
    res = {"value": round(rays / dt / 1e9, 4), "unit": "Grays/s", "ms_per_step": round(dt / args.steps * 1e3, 4),
           "render_ms_per_step": round(tot / max(n, 1), 4), "yaw_per_step_rad": 0.01,
           "deferred_pairs_per_step": dfr // args.steps, "recovered_pairs_per_step": rec // args.steps,
           "step": "device chunk lists for the moved cameras + batched render (Hi-Z carried)"}
    _log(f"moving camera: {res['ms_per_step']} ms/step ({res['render_ms_per_step']} ms render)")
    return res


def run_resolve(r1, world, wl1, args):
    """The resolve (vx_resolve: lanes 0, 1 and 2) of the 1080p view, then TAA: the workload's
    flags (all of them): the probe cube, bounce and sun over the probe field baked for the
    chunks within the upload window, the biome tint (world seed, strength 1), distant
    terrain shadows from the shaft window around the camera (its eight scan passes timed
    apart), sky tint, sky specular and fog on, a synthetic 27-layer 16x16 atlas with its
    mip chain. The march runs once; the resolve alone is timed over the steps."""
    from voxelcraft_menu_amd import workloads as W
    from voxelcraft_menu_amd.world import ShaftField, build_view, probe_bake, probe_uploadable, set_tint

    cam = W.default_camera(wl1, flags=W.DEFAULT_FLAGS)
    v = build_view(world, cam)
    set_tint(v.frame, wl1.gen.seed)
    baked = 0
    for key in world.keys():
        origin = tuple(int(k) * 64 for k in key)
        if probe_uploadable(origin, v.frame.cam_pos):
            occl, bounce = probe_bake(wl1.gen, origin)
            r1.write_probe(origin, occl, bounce)
            baked += 1
    v.frame.ambient, v.frame.mip_bias, v.frame.fog_density, v.frame.fog_falloff = 0.05, 0.0, 0.002, 0.01
    v.frame.fog_scatter, v.frame.fog_g, v.frame.fog_height = 0.3, 0.6, 80.0
    sf = ShaftField()
    sf.update(wl1.gen, v.frame.cam_pos[0], v.frame.cam_pos[2])
    sf.set_frame(v.frame)
    r1.upload_shaft(sf.heights())
    sf.close()
    spec, hi = W.spec_words(cam.flags, world)
    hi = (hi | 1) & ~(64 | 8192 | 16384 | 65536)
    rng = np.random.default_rng(0)
    texels = sum(27 * (16 >> l) ** 2 for l in range(4))
    faces = (np.arange(32 * 8, dtype=np.uint32) % 27) | ((np.arange(32 * 8, dtype=np.uint32) % 3) << 16)
    r1.set_resolve_assets(rng.integers(0, 256, texels * 4, dtype=np.uint8), 16, 27, 4, faces)
    r1.set_view(v)
    r1.shaft_scan(v.frame)
    r1.render(v.frame, spec, hi)
    r1.resolve()
    r1.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r1.shaft_scan(v.frame)
    r1.sync()
    dt_scan = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r1.resolve()
    r1.sync()
    dt = time.perf_counter() - t0
    px = wl1.width * wl1.height
    # TAA after it (feedback 0.1; the previous frame's camera 0.3 blocks to the left).
    from voxelcraft_menu_amd.world import set_prev_view_proj, view_proj
    import copy
    prev = copy.copy(v.frame)
    for i in range(3):
        prev.cam_pos[i] = v.frame.cam_pos[i] - 0.3 * v.frame.cam_right[i]
    set_prev_view_proj(v.frame, view_proj(prev))
    v.frame.taa_feedback = 0.1
    r1.render(v.frame, spec, hi)
    r1.resolve()
    r1.taa()
    r1.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r1.taa()
    r1.sync()
    dt_taa = time.perf_counter() - t0
    res = {"ms_per_frame": round(dt / args.steps * 1e3, 4), "Gpix_per_s": round(px * args.steps / dt / 1e9, 3),
           "taa_ms_per_frame": round(dt_taa / args.steps * 1e3, 4),
           "shaft_scan_ms": round(dt_scan / args.steps * 1e3, 4),
           "config": f"1920x1080, the workload's flags: probe cube / bounce / sun over {baked} baked chunks, "
                     "biome tint (world seed, strength 1), distant shadows from the 512x512 shaft window (texel 4) "
                     "around the camera; sky tint, sky specular and fog on; synthetic 16x16x27 atlas, 4 mips; "
                     "TAA with feedback 0.1"}
    _log(f"resolve 1080p: {res['ms_per_frame']} ms, taa {res['taa_ms_per_frame']} ms, "
         f"shaft scan {res['shaft_scan_ms']} ms")
    return res


def run_serving(args):
    """The serving workload: the 2048x2048x512 world resident on the device, an open-loop
    Poisson stream of view requests (resolutions uniform over the four sizes, seeded
    poses) mixed with Poisson edit batches of 1000 random voxel set/clear ops, through
    the continuous-batching server. Latency = completion - scheduled arrival. Camera
    requests are submitted by the server's native replay driver. Then a rate sweep
    (--serve-sweep): the same stream at increasing rates, up to the first whose p99
    exceeds 5 ms, reporting the highest rate within it with its mean batch size and the
    fraction of time a lane had a batch on the device."""
    from voxelcraft_menu_amd import serving as S
    from voxelcraft_menu_amd import workloads as W
    from voxelcraft_menu_amd.renderer import Renderer

    wl = W.WORKLOADS[4]
    world = W.make_world(wl)
    r = Renderer(64, 64)
    r.sync_world(world)
    dev = not args.serve_host_lists
    S.run(world, r, wl, qps=200, duration_s=1.0, edit_rate=0, seed=99, device_lists=dev)  # warm-up stream
    res = S.run(world, r, wl, qps=args.serve_qps, duration_s=args.serve_seconds, edit_rate=args.serve_edit_rate,
                seed=4, device_lists=dev)
    res["workload"] = wl.name
    res["chunk_lists"] = "device" if dev else "host"
    _log(f"serving: p50 {res['p50_ms']} ms p99 {res['p99_ms']} ms at {args.serve_qps} qps")
    rates = [float(x) for x in args.serve_sweep.split(",") if x.strip()] if dev else []
    if rates:
        sw = S.sweep(world, r, wl, rates, duration_s=args.serve_sweep_seconds, edit_rate=args.serve_edit_rate,
                     p99_bound_ms=5.0, seed=5, log=lambda m: _log(m[:160]))
        res["sweep"] = sw
        _log(f"serving sweep: max {sw['max_qps_within_bound']} qps with p99 <= 5 ms")
    r.close()
    return res


def run_view_setup(args):
    """Setting up the 64 views of the batch workload on the 8192-chunk world: chunk lists
    and grids built on the host (build_view per view, then vx_batch_set_views) against
    built on the device from the chunk table (vx_batch_set_cameras). The device lists are
    checked against the host's for every view."""
    from voxelcraft_menu_amd import workloads as W
    from voxelcraft_menu_amd.renderer import Batch, Renderer
    from voxelcraft_menu_amd.world import ViewSnapshot, build_view

    wl = W.WORKLOADS[4]
    world = W.make_world(wl)
    cams = W.random_cameras(wl, 64, seed=2)
    specs = [W.spec_words(c.flags, world) for c in cams]
    snap = ViewSnapshot(world)
    table = snap.chunk_table()
    r = Renderer(64, 64)
    r.sync_world(world)
    b = Batch(r, 64)
    b.set_chunk_table(table)

    def host_setup():
        b.set_views([snap.build_view(c) for c in cams], specs)

    def device_setup():
        b.set_cameras(cams, specs)

    res = {}
    for name, fn in (("host", host_setup), ("device", device_setup)):
        fn()
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        res[name] = (time.perf_counter() - t0) / reps
    b.set_cameras(cams, specs)
    for i, c in enumerate(cams):
        v = build_view(world, c)
        ch, grid = b.read_chunks(i)
        if b.view_counts(i) != (v.n_total, v.n_visible, v.camera_chunk) or not np.array_equal(
                ch, v.chunks[:v.n_total]) or not np.array_equal(grid, v.grid):
            _fail(f"device chunk list of view {i} differs from the host's")
    b.close()
    r.close()
    out = {"workload": "64-views-of-2048x2048x512", "views": 64, "table_chunks": int(len(table)),
           "host_ms": round(res["host"] * 1e3, 3), "device_ms": round(res["device"] * 1e3, 3),
           "host_over_device": round(res["host"] / res["device"], 1), "checked_against_host": 64}
    _log(f"view setup: 64 views in {out['host_ms']} ms with host lists, {out['device_ms']} ms built on the device")
    return out


def run_chunk_build(args):
    """build_local of the 1024 chunks of the 1024x1024x256 world: on the device (dense
    chunks resident in HBM, vx_build_local_device, records written to HBM) and on one host
    core (the host build_local into the same flat records, a sample of chunks). The
    device records are checked against the host's on the sample."""
    import ctypes as C

    from voxelcraft_menu_amd import _abi as A
    from voxelcraft_menu_amd import workloads as W
    from voxelcraft_menu_amd.renderer import Renderer
    from voxelcraft_menu_amd.world import build_local_flat

    wl = W.WORKLOADS[1]
    world = W.make_world(wl)
    keys = [tuple(int(v) for v in k) for k in world.keys()]
    dense = np.stack([world.extract_dense(k) for k in keys])
    n = dense.shape[0]
    r = Renderer(64, 64)
    L = A.march()
    d_dense, d_out = C.c_void_p(), C.c_void_p()
    A.check(L.vx_dev_alloc(r.handle, dense.nbytes, C.byref(d_dense)), "vx_dev_alloc")
    A.check(L.vx_dev_alloc(r.handle, n * A.LOCAL_TREE_MAX, C.byref(d_out)), "vx_dev_alloc")
    A.check(L.vx_dev_upload(r.handle, d_dense, A.ptr(dense), dense.nbytes), "vx_dev_upload")

    def device_build():
        A.check(L.vx_build_local_device(r.handle, d_dense, n, d_out, A.LOCAL_TREE_MAX), "vx_build_local_device")

    for _ in range(2):
        device_build()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        device_build()
    dt = (time.perf_counter() - t0) / reps
    sample = list(range(0, n, max(1, n // 64)))
    host = []
    t0 = time.perf_counter()
    for i in sample:
        host.append(build_local_flat(dense[i]))
    host_dt = (time.perf_counter() - t0) / len(sample)
    out_h = np.zeros(n * A.LOCAL_TREE_MAX, dtype=np.uint8)
    A.check(L.vx_dev_download(r.handle, A.ptr(out_h), d_out, out_h.nbytes), "vx_dev_download")
    A.check(L.vx_dev_free(r.handle, d_dense), "vx_dev_free")
    A.check(L.vx_dev_free(r.handle, d_out), "vx_dev_free")
    for i, want in zip(sample, host):
        got = out_h[i * A.LOCAL_TREE_MAX:i * A.LOCAL_TREE_MAX + len(want)].tobytes()
        if got != want:
            _fail(f"device chunk build differs from the host build for chunk {keys[i]}")
    rec_bytes = int(sum(int(out_h[i * A.LOCAL_TREE_MAX + 60:i * A.LOCAL_TREE_MAX + 64].view(np.uint32)[0])
                        for i in range(n)))
    r.close()
    res = {
        "workload": "chunks-of-1024x1024x256", "chunks": n, "device_ms": round(dt * 1e3, 3),
        "device_chunks_per_s": round(n / dt, 1), "host_ms_per_chunk_1core": round(host_dt * 1e3, 3),
        "host_chunks_per_s_1core": round(1.0 / host_dt, 1), "device_over_host_1core": round(host_dt * n / dt, 1),
        "dense_GBps": round(n * 64 ** 3 * 2 / dt / 1e9, 1), "record_bytes": rec_bytes,
        "checked_against_host": len(sample),
    }
    _log(f"chunk build: {n} chunks in {res['device_ms']} ms on the device; host {res['host_ms_per_chunk_1core']} ms/chunk")
    return res


def run_device_edits(args):
    """Edit batches of 1000 random voxel set/clear ops on the 2048x2048x512 world (8192
    chunks): applied on the GPU to the device-resident world (vx_dworld_edit: decode the
    edited chunks, replay each voxel's edits, rebuild the chunks, place their blocks)
    against applied to the host world (World::apply_edits on one core, without the upload
    of the edited pools that follows it). Timed per batch, after two warm-up batches; the
    device world is then checked against the host world: chunk records (box, dry mask,
    flags) of every chunk and the dense ids of a sample of edited chunks."""
    from voxelcraft_menu_amd import serving as S
    from voxelcraft_menu_amd import workloads as W
    from voxelcraft_menu_amd.renderer import DeviceWorld, Renderer
    from voxelcraft_menu_amd.world import ViewSnapshot

    wl = W.WORKLOADS[4]
    world = W.make_world(wl)
    r = Renderer(64, 64)
    dw = DeviceWorld(r, world, capacity_factor=8.0)  # compacts itself when the batches fill it
    rng = np.random.default_rng(12)
    batches = [S.random_edits(rng, wl, 1000) for _ in range(12)]
    dev_ms, dev_gpu_ms, host_ms, rebuilt = [], [], [], []
    for k, e in enumerate(batches):
        t0 = time.perf_counter()
        dw.edit(e)
        t1 = time.perf_counter()
        world.apply_edits(e)
        t2 = time.perf_counter()
        if k >= 2:
            st = dw.stats()
            dev_ms.append((t1 - t0) * 1e3)
            dev_gpu_ms.append(st["last_ms"])
            rebuilt.append(st["chunks_rebuilt"])
            host_ms.append((t2 - t1) * 1e3)
    snap = ViewSnapshot(world)
    host = snap.chunk_table().view(np.uint32).reshape(-1, 24)
    snap.close()
    dev = dw.chunks().view(np.uint32).reshape(-1, 24)
    keep = np.uint32(0xFFFFFFFE)
    if host.shape != dev.shape or not (np.array_equal(host[:, 0:11], dev[:, 0:11]) and np.array_equal(
            host[:, 20:22], dev[:, 20:22]) and np.array_equal(host[:, 17] & keep, dev[:, 17] & keep)):
        _fail("device world records differ from the host world's after the edit batches")
    e = batches[-1]
    keys = sorted({(int(x) // 64, int(y) // 64, int(z) // 64) for x, y, z in zip(e["x"], e["y"], e["z"])})
    sample = keys[::max(1, len(keys) // 32)]
    for key in sample:
        if world.contains(key) and not np.array_equal(dw.extract(key), world.extract_dense(key)):
            _fail(f"device world chunk {key} differs from the host world's after the edit batches")
    st = dw.stats()
    dw.close()
    r.close()
    med = lambda v: float(np.median(v))
    res = {"workload": "1000-op-edit-batches-on-2048x2048x512", "batches": len(dev_ms), "ops_per_batch": 1000,
           "chunks_rebuilt_per_batch": int(med(rebuilt)), "device_ms": round(med(dev_ms), 3),
           "device_gpu_ms": round(med(dev_gpu_ms), 3), "host_ms_1core": round(med(host_ms), 3),
           "host_over_device": round(med(host_ms) / med(dev_ms), 2), "used_bytes": int(st["used_bytes"]),
           "checked_against_host": {"records": int(dev.shape[0]), "chunks_dense": len(sample)}}
    _log(f"device edits: {res['device_ms']} ms per 1000-op batch ({res['device_gpu_ms']} ms GPU), "
         f"host {res['host_ms_1core']} ms")
    return res


def run_incoherent(args):
    """The incoherent batch: the 2048x2048x512 world (caves, sea level 96, 1 % glass), 32
    seeded views at 1280x720, primary + sun shadow + one secondary bounce per water / glass
    hit (water refract and reflect legs, glass transmit leg and the water legs behind it,
    each traced with trace_world through compact leg queues), waves on with the reference's
    defaults (amp 0.12, scale 32, speed 0.32, reflect slope 0; aniso, shoal and fill on) at
    animation time 2.5. A step is one frame of all 32 views; rays = primary + shadow +
    secondary legs + sun rays at their hits (the shoal probes under waved reflect legs
    are not counted)."""
    from voxelcraft_menu_amd import workloads as W
    from voxelcraft_menu_amd.renderer import Batch, Renderer
    from voxelcraft_menu_amd.world import build_view

    wl = W.WORKLOADS[3]
    world = W.make_world(wl)
    cams = W.random_cameras(wl, wl.views, seed=3, resolutions=[(wl.width, wl.height)])
    for c in cams:
        c.anim_time = 2.5
    views = [build_view(world, c) for c in cams]
    specs = [W.spec_words(c.flags, world) for c in cams]
    px = sum(int(v.frame.res[0]) * int(v.frame.res[1]) for v in views)
    r = Renderer(64, 64)
    r.sync_world(world)
    b = Batch(r, len(views))
    b.set_views(views, specs)
    best, _ = b.autotune(2)
    b.set_secondary(False)
    for _ in range(2):
        b.render()
    prim_dt, _, _ = _time_batch(b, args.steps)
    b.set_secondary(True)
    for _ in range(2):
        b.render()
    b.sync()
    d = b.demand()
    sec = b.secondary_count()
    if sec["overflow"]:
        _fail("secondary leg queue overflowed")
    dt, _, _ = _time_batch(b, args.steps)
    b.close()
    r.close()
    rays = px + d["shadow_rays"] + sec["water_legs"] + sec["glass_legs"] + sec["sun_rays"]
    res = {
        "workload": wl.name, "views": len(views), "resolution": [wl.width, wl.height], "camera_seed": 3,
        "value": round(rays * args.steps / dt / 1e9, 4), "unit": "Grays/s", "ms_per_step": round(dt / args.steps * 1e3, 4),
        "ms_per_step_without_secondary": round(prim_dt / args.steps * 1e3, 4),
        "primary_rays": px, "shadow_rays": d["shadow_rays"], "secondary_rays": sec["water_legs"] + sec["glass_legs"],
        "water_legs": sec["water_legs"], "glass_legs": sec["glass_legs"], "secondary_sun_rays": sec["sun_rays"],
        "tuning": best,
        "water": "waves on (reference defaults, time 2.5)",
    }
    _log(f"incoherent: {res['ms_per_step']} ms/step ({res['ms_per_step_without_secondary']} without the bounce)")
    return res


def run_single(args, out):
    from voxelcraft_menu_amd import _abi as A
    from voxelcraft_menu_amd import workloads as W
    from voxelcraft_menu_amd.renderer import Renderer, device_name
    from voxelcraft_menu_amd.world import build_view

    wl = W.WORKLOADS[args.workload]
    world = W.make_world(wl)
    cam = W.default_camera(wl)
    view = build_view(world, cam)
    spec, spec_hi = W.spec_words(cam.flags, world)
    path = A.PATH_TILES if args.path == "tiles" else A.PATH_PAIRS
    r = Renderer(wl.width, wl.height, path=path)
    r.sync_world(world)
    r.set_view(view)
    for _ in range(max(args.warmup, 2)):
        r.render(view.frame, spec, spec_hi)
    r.sync()
    demand = r.pair_demand()
    if demand["truncated"]:
        _fail("pair list overflowed its capacity; results would not be deterministic")
    shadow_rays = demand["shadow_rays"]
    rays_per_step = wl.width * wl.height + shadow_rays

    r.timings_reset()
    r.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.render(view.frame, spec, spec_hi)
    r.sync()
    dt = time.perf_counter() - t0
    sums, nframes = r.timings_sum()
    value = rays_per_step * args.steps / dt / 1e9
    march_ms = (sums["march"] + sums["march2"]) / max(nframes, 1)
    shadow_ms = sums["shadow"] / max(nframes, 1)
    fused = path == A.PATH_TILES and (r.tuning()["variant"] & 4) != 0
    r.close()
    out.update({
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Grays/s",
        "ms_per_step": round(dt / args.steps * 1e3, 4),
        "config": {
            "workload": wl.name, "resolution": [wl.width, wl.height],
            "world_chunks": [wl.gen.chunk_max[i] - wl.gen.chunk_min[i] for i in range(3)],
            "seed": wl.gen.seed, "hiz": True, "path": args.path, "primary_rays": wl.width * wl.height,
            "shadow_rays": shadow_rays, "chunks_visible": view.n_visible, "chunks_total": view.n_total,
            "device": device_name(),
        },
        "passes_ms": {k: round(v / max(nframes, 1), 4) for k, v in sums.items()},
    })
    if args.no_cpu_baseline:
        out["cpu_baseline"] = None
        out["roofline"] = None
        return
    cb, prim, sh = cpu_baseline(world, [view], [(spec, spec_hi)], args.cpu_budget)
    out["cpu_baseline"] = cb
    tiles = view.tiles[0] * view.tiles[1]
    px = wl.width * wl.height
    alg = prim["node_bytes"] + 8 * px + 4 * tiles
    kernel_ms = march_ms
    if fused:
        alg += sh["node_bytes"] + 2 * px
        kernel_ms = march_ms + shadow_ms
    achieved = alg / (kernel_ms * 1e-3) / 1e9
    out["roofline"] = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic_from_profiles("k_march_tiles", f"workload{args.workload}") if path == A.PATH_TILES
        else None,
        "kernel": "k_march_tiles" if path == A.PATH_TILES else "k_march_pairs",
        "kernel_ms": round(kernel_ms, 4), "fused_shadow": fused, "alg_bytes_per_launch": int(alg),
        "alg_bytes_per_pixel": round(alg / px, 2),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", type=int, default=2, choices=[0, 1, 2])
    ap.add_argument("--path", choices=["tiles", "pairs"], default="tiles", help="single-view workloads only")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of oracle work for cpu_baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving stream (workload 2 only)")
    ap.add_argument("--no-incoherent", action="store_true", help="skip the incoherent batch (workload 2 only)")
    ap.add_argument("--no-build", action="store_true", help="skip the device chunk build (workload 2 only)")
    ap.add_argument("--no-view-setup", action="store_true", help="skip the view-setup comparison (workload 2 only)")
    ap.add_argument("--no-device-edits", action="store_true", help="skip the device edit batches (workload 2 only)")
    ap.add_argument("--no-moving", action="store_true", help="skip the turning-camera steps (workload 2 only)")
    ap.add_argument("--no-layout-ab", action="store_true", help="skip the pool-layout A/B (workload 2 only)")
    ap.add_argument("--pool-layout", choices=["reference", "flat"], default="reference",
                    help="pool layout of the batched measurements (workload 2)")
    ap.add_argument("--serve-qps", type=float, default=1000.0, help="offered view requests per second")
    ap.add_argument("--serve-seconds", type=float, default=4.0, help="seconds of request arrivals")
    ap.add_argument("--serve-edit-rate", type=float, default=2.0, help="edit batches (1000 ops) per second")
    ap.add_argument("--serve-sweep", default="1000,2000,2500,3000,3500,4000,5000,6000,8000",
                    help="comma-separated request rates of the sweep (empty: no sweep)")
    ap.add_argument("--serve-sweep-seconds", type=float, default=3.0, help="seconds of arrivals per sweep rate")
    ap.add_argument("--serve-host-lists", action="store_true",
                    help="serving requests carry host-built chunk lists instead of cameras")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one 256 MiB device-to-device copy (scale reference for PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        _fail("bench.py measures one GPU: this build has no multi-GPU path (views would shard across GPUs by "
              "view, with the pools replicated). Run with --gpus 1.")
    if args.steps < 1 or args.steps > 256:
        _fail("--steps must be in [1, 256]")

    if args.calibrate_copy:
        import torch
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    out = {"metric": METRIC, "value": None, "unit": "Grays/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32",
           "data": "synthetic seeded world (project terrain generator), seeded cameras"}
    if args.workload == 2:
        run_batch(args, out)
    else:
        run_single(args, out)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
