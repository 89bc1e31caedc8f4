"""Build libgigpu.so (HIP, gfx950) in-tree.

    python -m gigpu.build [--incremental]

Compiles every source of ../csrc with hipcc for gfx950 into
``gigpu/libgigpu.so``; objects go to ``../_build``.  Translation units are
compiled in parallel.  Without --incremental every object is rebuilt.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(HERE, "libgigpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-Wno-pass-failed",
          "-Wno-unused-result", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I" + CSRC, "-I" + os.path.join(os.path.dirname(PKG), "include")]
LIBS = ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]

SOURCES = ["eri.hip", "chol.hip", "mo.hip", "grad.hip", "split.hip", "onee.hip", "aomap.hip", "thc.hip", "gi_tables.cpp",
           "gi_context.cpp", "sort.cpp", "capi.cpp"]
HEADERS = ["gi_common.h", "gi_context.h", "md_kernel.h", "md4_kernel.h", "md_generic.h", "md_classes.inc", "grad_kernels.h", "onee_kernels.h", "thc_kernels.h"]


def _compile(src: str, incremental: bool) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    deps = [path] + [os.path.join(CSRC, h) for h in HEADERS] + [os.path.join(os.path.dirname(PKG), "include", "gigpu.h")]
    if incremental and os.path.exists(obj) and os.path.getmtime(obj) >= max(os.path.getmtime(d) for d in deps):
        return obj
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC] + COMMON + lang + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout[-4000:], r.stderr[-8000:]))
    return obj


def build(incremental: bool = False, jobs: int = 8) -> str:
    sys.path.insert(0, PKG)
    from gigpu import eri_gen  # noqa: E402  (emits csrc/md_classes.inc)
    eri_gen.emit()
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, incremental), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + LIBS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(incremental="--incremental" in sys.argv))
