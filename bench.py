#!/usr/bin/env python3
"""Benchmark: Grays/s (primary + sun-shadow rays) of the voxel ray march on one MI355X.

Default workload: the 1024x1024x256 world (seed 7) rendered by one 1920x1080 camera at
the default pose, primary DDA + one sun shadow ray per lit pixel, two-phase Hi-Z on 8x8
tiles. A step is one frame: march + Hi-Z + shadow over the resident pools; the camera is
fixed, so frames after the first run in the Hi-Z steady state the reference's
--bench-frames reports. Inputs (pools, chunk list, grid, frame) are in HBM before the
timed region starts.

Output: one JSON line with the metric, a roofline object for the dominant kernel
(the tile march) and a cpu_baseline object (the C oracle timed on the host cores).
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


def _fail(msg: str, code: int = 2):
    print(msg, file=sys.stderr, flush=True)
    sys.exit(code)


def traffic_from_profiles(kernel: str, workload: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if present."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_pmc.md")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith("traffic_bytes_per_launch:"):
                try:
                    return float(line.split(":", 1)[1].strip())
                except ValueError:
                    return None
    return None


def cpu_baseline(world, view, spec, spec_hi, budget_s: float):
    """The oracle (plain-C restatement of the WGSL passes) on this host's cores.

    Runs whole frames (the same view, Hi-Z carried) until budget_s has elapsed and
    returns the rate plus the oracle's algorithmic byte counts of one steady-state
    frame, split into the primary pass and the shadow pass."""
    from oracle import oracle as O

    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(threads, 64))
    sc = O.Scene(world.pool(0), world.pool(1), world.pool(2), view.chunks, view.grid)
    sp = O.spec_from(spec, spec_hi)
    hz = np.zeros(view.tiles[0] * view.tiles[1], np.uint32)
    _, _, hz, _ = O.render(sc, view.frame_bytes, sp, hz, shadow=False, threads=threads)
    vis, _, _, prim = O.render(sc, view.frame_bytes, sp, hz, shadow=False, threads=threads)
    _, sh_stats = O.primary_shadow(sc, view.frame_bytes, sp, vis, threads=threads)
    frames, rays, t0 = 0, 0, time.perf_counter()
    while True:
        _, _, hz, st = O.render(sc, view.frame_bytes, sp, hz, shadow=True, threads=threads)
        frames += 1
        rays += view.width * view.height + st["rays_shadow"]
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    return {
        "value": rays / dt / 1e9, "unit": "Grays/s", "cores": threads, "kind": "port",
        "sample": f"{frames} full {view.width}x{view.height} frames of the same view "
                  f"(primary + shadow + Hi-Z) in {dt:.1f} s",
    }, prim, sh_stats


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", type=int, default=1, choices=[0, 1])
    ap.add_argument("--path", choices=["tiles", "pairs"], default="tiles")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of oracle work for cpu_baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one 256 MiB device-to-device copy (scale reference for PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        _fail("bench.py measures one GPU: this build has no multi-GPU path (views would shard by GPU; "
              "see the design notes). Run with --gpus 1.")
    if args.steps < 1 or args.steps > 256:
        _fail("--steps must be in [1, 256]")

    from voxelcraft_menu_amd import _abi as A
    from voxelcraft_menu_amd import workloads as W
    from voxelcraft_menu_amd.renderer import Renderer, device_name
    from voxelcraft_menu_amd.world import build_view

    if args.calibrate_copy:
        import torch
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

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
    t1 = time.perf_counter()
    sums, nframes = r.timings_sum()
    dt = t1 - t0
    value = rays_per_step * args.steps / dt / 1e9
    march_ms = (sums["march"] + sums["march2"]) / max(nframes, 1)
    shadow_ms = sums["shadow"] / max(nframes, 1)

    out = {
        "metric": "Grays/s primary+shadow (plain march, one 1920x1080 view per step)"
        if args.workload == 1 else "Grays/s primary+shadow (plain march, 640x360)",
        "value": round(value, 4),
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic seeded world (project terrain generator), fixed camera",
        "config": {
            "workload": wl.name,
            "resolution": [wl.width, wl.height],
            "world_chunks": [wl.gen.chunk_max[i] - wl.gen.chunk_min[i] for i in range(3)],
            "seed": wl.gen.seed,
            "hiz": True,
            "path": args.path,
            "primary_rays": wl.width * wl.height,
            "shadow_rays": shadow_rays,
            "chunks_visible": view.n_visible,
            "chunks_total": view.n_total,
            "device": device_name(),
        },
        "passes_ms": {k: round(v / max(nframes, 1), 4) for k, v in sums.items()},
    }

    prim = sh = None
    if not args.no_cpu_baseline:
        cb, prim, sh = cpu_baseline(world, view, spec, spec_hi, args.cpu_budget)
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None

    # Roofline of the dominant kernel: algorithmic bytes per launch (the oracle's exact
    # per-(ray, chunk) node accounting plus the 8-byte key per pixel and 4-byte Hi-Z per
    # tile; with the fused shadow also the shadow rays' node traffic and the 2-byte R16F
    # word per pixel) over the kernel's average duration measured above with HIP events.
    fused = path == A.PATH_TILES and (r.tuning()["variant"] & 4) != 0
    if prim is not None:
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
            "kernel_ms": round(kernel_ms, 4),
            "fused_shadow": fused,
            "alg_bytes_per_launch": int(alg),
            "alg_bytes_per_pixel": round(alg / px, 2),
        }
        if not fused:
            sh_alg = sh["node_bytes"] + 8 * px + 2 * px
            out["roofline_shadow"] = {
                "kernel": "k_shadow", "kernel_ms": round(shadow_ms, 4),
                "achieved": round(sh_alg / (shadow_ms * 1e-3) / 1e9, 1), "unit": "GB/s",
            }
    else:
        out["roofline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
