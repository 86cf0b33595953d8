"""Build libquickfock.so (gfx950) in-tree.

    python -m quick_amd.build [-j N]

Steps: regenerate the per-class recursions (quick_amd/codegen/gen_eri.py),
compile every translation unit with hipcc for gfx950 in parallel, link one
shared library at quick_amd/libquickfock.so.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys
import time

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libquickfock.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # no packed FP32 instructions (v_pk_*_f32) anywhere
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # hardware FP64 atomic add on coarse-grained device memory
    "-munsafe-fp-atomics",
    "-Wno-unused-variable", "-Wno-unused-function",
]


def sources():
    srcs = sorted(glob.glob(os.path.join(CSRC, "gen", "eri_*.hip")))
    srcs += sorted(glob.glob(os.path.join(CSRC, "gen", "egrad_*.hip")))
    srcs += [os.path.join(CSRC, "engine.hip"), os.path.join(CSRC, "oei.hip"),
             os.path.join(CSRC, "quick_compat.cpp")]
    return srcs


def _deps(src):
    deps = [src, os.path.join(CSRC, "qf_types.h"), os.path.join(CSRC, "qf_device.h"), os.path.join(CSRC, "qf_coop.h"),
            os.path.join(os.path.dirname(PKG), "include", "quickfock.h"), os.path.join(CSRC, "gen", "launchers.h")]
    base = os.path.basename(src)
    if base.startswith("eri_"):
        deps.append(os.path.join(CSRC, "gen", base[4:-4] + ".h"))
        deps.append(os.path.join(CSRC, "gen", "coop_" + base[4:-4] + ".h"))
    if base.startswith("egrad_"):
        deps += [os.path.join(CSRC, "qf_grad.h"), os.path.join(CSRC, "gen", "grad_maps.h"),
                 os.path.join(CSRC, "gen", "grad_" + base[6:-4] + ".h")]
    if base == "engine.hip":
        deps.append(os.path.join(CSRC, "gen", "grad_launchers.h"))
    return deps


def compile_one(src, force=True):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    if not force and os.path.exists(obj):
        t = os.path.getmtime(obj)
        if all(os.path.getmtime(d) <= t for d in _deps(src) if os.path.exists(d)):
            return obj, 0.0
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-O2", "-fPIC", "-std=c++17", "-c", src, "-o", obj]
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr[-6000:]))
    return obj, time.time() - t0


def build(jobs=None, verbose=True, force=True):
    from quick_amd.codegen import gen_eri
    import contextlib
    import io
    with contextlib.redirect_stdout(io.StringIO()):
        gen_eri.main()
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), os.cpu_count() or 4, 16)
    objs = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        for obj, dt in ex.map(lambda x: compile_one(x, force), srcs):
            objs.append(obj)
            if verbose:
                print("  %-28s %6.1fs" % (os.path.basename(obj), dt), flush=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--incremental", action="store_true", help="skip objects newer than their sources")
    a = ap.parse_args()
    build(a.j, force=not a.incremental)


if __name__ == "__main__":
    sys.exit(main())
