"""Builds libquantkos.so (gfx950 HIP kernels + C ABI + C++ host) and the CLI in-tree.

    python -m quantkos_amd.build          # compile everything
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libquantkos.so")
CLI = os.path.join(PKG, "quantkos")

ARCH = os.environ.get("QK_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 (v_pk_*_f32) is disabled for all device code.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++20", "-fPIC", "-Wall",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include")]
HOST_FLAGS = ["-O2", "-std=c++20", "-fPIC", "-Wall", "-I" + os.path.join(ROOT, "include"),
              "-I" + HOST]

HIP_SOURCES = ["qk_kernels.hip", "qk_capi.hip", "qk_server.hip"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "not a recognized feature" not in l
                    and not l.startswith("1 warning generated"))
    if p.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (p.returncode, " ".join(cmd), out))
    return out


def _host_sources():
    if not os.path.isdir(HOST):
        return []
    return sorted(f for f in os.listdir(HOST) if f.endswith(".cpp") and f != "main.cpp")


def build(verbose: bool = False) -> str:
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found: the HIP extension cannot be built")
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    for src in _host_sources():
        obj = os.path.join(BUILD, "host_" + src.replace(".cpp", ".o"))
        jobs.append(["g++"] + HOST_FLAGS + ["-c", os.path.join(HOST, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    objs = [j[-1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    main = os.path.join(HOST, "main.cpp")
    if os.path.exists(main):
        _run(["g++"] + HOST_FLAGS + [main, "-o", CLI + ".tmp", "-L" + PKG, "-lquantkos",
                                     "-Wl,-rpath,$ORIGIN"])
        os.replace(CLI + ".tmp", CLI)
    return LIB


if __name__ == "__main__":
    t = build(verbose="-v" in sys.argv)
    print("built", t)
