"""Build recipe for the two in-tree shared objects.

  libvxhost.so   host world store + view setup (g++, CPU only)
  libvxmarch.so  HIP kernels + C ABI for gfx950 (hipcc)

Floating point: the kernels are compiled without FMA contraction, with correctly
rounded f32 division and square root and with denormals kept, so every f32 operation
rounds once in source order. Packed-FP32 instruction selection is disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INC = os.path.join(ROOT, "include")
CSRC = os.path.join(HERE, "csrc")

HOST_SOURCES = ["host/world.cpp", "host/worldgen.cpp", "host/view.cpp", "host/capi.cpp", "host/journal.cpp",
                "host/flat.cpp", "host/lod.cpp", "host/flat_layout.cpp", "host/shaft.cpp",
                "host/probe.cpp"]
HIP_SOURCES = ["kernels.hip", "tiles.hip", "secondary.hip", "vxmarch.hip", "batch.hip", "serve.hip", "build.hip", "noalloc.hip",
               "resolve.hip", "dworld.hip"]
HIP_EXTRA = []  # further .hip translation units are appended by later modules

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]


def _include_closure(src: str) -> list:
    """The file and every quoted header it includes, transitively (looked up beside the
    including file, in the library's sources and in include/)."""
    import re

    seen, todo = [], [src]
    while todo:
        f = todo.pop()
        if f in seen:
            continue
        seen.append(f)
        with open(f) as fh:
            for name in re.findall(r'^\s*#\s*include\s+"([^"]+)"', fh.read(), re.M):
                for d in (os.path.dirname(f), CSRC, INC):
                    cand = os.path.join(d, name)
                    if os.path.exists(cand):
                        todo.append(os.path.normpath(cand))
                        break
    return seen


def source_digest(kind: str) -> str:
    """Digest of what a library is built from: its sources, every header they can include
    and the compile flags. The build writes it beside the library (<lib>.srchash); the
    loader refuses a library whose digest differs from the current sources."""
    import glob
    import hashlib

    if kind == "host":
        files = [os.path.join(CSRC, s) for s in HOST_SOURCES] + glob.glob(os.path.join(CSRC, "host", "*.hpp"))
        flags = "g++ -std=c++17 -O2"
    elif kind == "tile_kernels":
        # What the tile-march kernels (single view and batched) are compiled from: their
        # translation unit, the headers it includes (transitively) and the flags. Counter
        # profiles of those kernels carry this digest, so edits elsewhere in the library
        # leave them valid.
        files = _include_closure(os.path.join(CSRC, "tiles.hip"))
        h = hashlib.sha256(" ".join(HIPCC_FLAGS).encode())
        for f in sorted(set(files)):
            h.update(os.path.relpath(f, ROOT).encode())
            with open(f, "rb") as fh:
                h.update(fh.read())
        return h.hexdigest()
    elif kind == "calib":
        return hashlib.sha256(" ".join(HIPCC_FLAGS).encode() +
                              open(os.path.join(CSRC, "calib.hip"), "rb").read()).hexdigest()
    else:
        files = [os.path.join(CSRC, s) for s in HIP_SOURCES + HIP_EXTRA] + glob.glob(os.path.join(CSRC, "*.hpp"))
        flags = " ".join(HIPCC_FLAGS)
    files += glob.glob(os.path.join(INC, "*.h"))
    h = hashlib.sha256(flags.encode())
    for f in sorted(set(files)):
        h.update(os.path.relpath(f, ROOT).encode())
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def _stamp(lib: str, kind: str):
    with open(lib + ".srchash", "w") as fh:
        fh.write(source_digest(kind) + "\n")


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... -o {cmd[-1]}")
    return r


def build_host(verbose: bool = False) -> str:
    out = os.path.join(HERE, "libvxhost.so")
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-Wextra", f"-I{INC}"]
    cmd += [os.path.join(CSRC, s) for s in HOST_SOURCES] + ["-lpthread", "-o", out]
    _run(cmd, verbose)
    _stamp(out, "host")
    return out


def build_march(verbose: bool = False) -> str:
    """Each translation unit compiles to its own object, several at once (they share no
    device symbols: kernels are launched through their host stubs), then one link."""
    from concurrent.futures import ThreadPoolExecutor

    out = os.path.join(HERE, "libvxmarch.so")
    tmp = out + ".tmp"
    objdir = os.path.join(ROOT, "_build", "march")
    os.makedirs(objdir, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES + HIP_EXTRA]
    objs = [os.path.join(objdir, os.path.basename(s) + ".o") for s in srcs]
    cflags = [f for f in HIPCC_FLAGS if f != "-shared"]
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", "0") or 0) or (os.cpu_count() or 1), 16))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda so: _run([_hipcc()] + cflags + [f"-I{INC}", f"-I{CSRC}", "-c", so[0], "-o", so[1]], verbose),
                    zip(srcs, objs)))
    _run([_hipcc()] + HIPCC_FLAGS + objs + ["-pthread", "-o", tmp], verbose)
    os.replace(tmp, out)
    _stamp(out, "march")
    return out


def build_calib(verbose: bool = False) -> str:
    """libvxcalib.so: the VALU issue-rate probe the benchmark's roofline is checked with."""
    out = os.path.join(HERE, "libvxcalib.so")
    cmd = [_hipcc()] + HIPCC_FLAGS + [os.path.join(CSRC, "calib.hip"), "-o", out]
    _run(cmd, verbose)
    _stamp(out, "calib")
    return out


def build(verbose: bool = False):
    return build_host(verbose), build_march(verbose), build_calib(verbose)


if __name__ == "__main__":
    print(build(verbose=True))
