"""Build the native parts of harvest_cuda in place.

  * libharvest_hip.so - hand-written gfx950 kernels + C ABI (hipcc), and the
                        host staging copy (g++)
  * _core*.so         - C++ host engine + pybind11 module (g++), linked to
                        libharvest_hip.so with an $ORIGIN rpath

Every compile disables floating-point contraction (WORLD's reference build
does the same, CMakeLists.txt:37,46) and hipcc never emits packed FP32 ops.
Run directly (python harvest-cuda_amd/build_native.py) or via
__graft_entry__.build().
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent
CSRC = ROOT / "csrc"
PKG = ROOT / "harvest_cuda"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(map(str, cmd))}")
    return r


def build_backend() -> Path:
    out = PKG / "libharvest_hip.so"
    objs = []
    srcs = ["kernels.hip", "refine_table.hip", "capi.hip"]
    tmp = ROOT / "_build"
    tmp.mkdir(exist_ok=True)

    def compile_one(src):
        obj = tmp / (src + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        return obj

    def compile_host(src):  # host-only code: the host compiler, no device pass
        obj = tmp / (src + ".o")
        _run(["g++", "-O3", "-std=c++17", "-fPIC", "-pthread", "-ffp-contract=off", "-c", str(CSRC / src),
              "-o", str(obj)])
        return obj

    with ThreadPoolExecutor(len(srcs) + 1) as pool:
        host = pool.submit(compile_host, "host_copy.cpp")
        objs = list(pool.map(compile_one, srcs)) + [host.result()]
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(out), *map(str, objs)])
    return out


def build_module() -> Path:
    import pybind11

    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = PKG / f"_core{suffix}"
    cmd = [
        "g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-ffp-contract=off",
        "-fvisibility=hidden",
        f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}",
        str(CSRC / "engine.cpp"), str(CSRC / "adjoint.cpp"), str(CSRC / "bindings.cpp"),
        f"-L{PKG}", "-lharvest_hip", "-Wl,-rpath,$ORIGIN", "-o", str(out),
    ]
    _run(cmd)
    return out


def build() -> None:
    PKG.mkdir(exist_ok=True)
    build_backend()
    build_module()


if __name__ == "__main__":
    build()
    print("built", *sorted(p.name for p in PKG.glob("*.so")))
