"""In-tree build of the native parts (hipcc for gfx950, g++ for host-only code).

Outputs (git-ignored, shipped with the tree):
  lightsim2grid_amd/libpfgpu.so            C-ABI library: HIP kernels + host analysis
  lightsim2grid_amd/lightsim2grid_amd_cpp*.so  pybind11 module mirroring lightsim2grid_cpp
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpfgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PFGPU_ARCH", "gfx950")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             # packed FP32 ops stay off (hipcc's host pass warns that it does not know the feature)
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]


def _run(cmd, **kw):
    print("[build]", " ".join(cmd), flush=True)
    subprocess.check_call(cmd, **kw)


# translation units of libpfgpu.so: the HIP kernels and their C ABI, the host analysis, the
# host mirror of the solver surface (over the C ABI) and the serving loop with its C ABI
LIB_SOURCES = ["pfgpu.hip", "wide.hip", "batchprep.hip", "analysis.cpp", "gridmodel.cpp", "serve.cpp", "serve_capi.cpp"]


def build_lib() -> str:
    """hipcc per translation unit (in parallel), then one link."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        _run([HIPCC, *HIP_FLAGS, f"-I{ROOT}/include", "-c", "-o", obj, os.path.join(CSRC, src)])
        return obj

    with ThreadPoolExecutor(max_workers=len(LIB_SOURCES)) as ex:
        objs = list(ex.map(compile_one, LIB_SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *objs, "-Wl,-soname,libpfgpu.so", "-pthread"])
    os.replace(tmp, LIB)
    return LIB


def module_path() -> str:
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, "lightsim2grid_amd_cpp" + suffix)


def build_module() -> str:
    import pybind11
    src = os.path.join(CSRC, "module.cpp")
    if not os.path.exists(src):
        return ""
    out = module_path()
    inc = [f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}", f"-I{ROOT}/include", f"-I{CSRC}"]
    srcs = [src]
    tmp = out + ".tmp"
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden", *inc, *srcs, "-o", tmp,
          f"-L{HERE}", "-lpfgpu", "-Wl,-rpath,$ORIGIN", "-pthread"])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_lib()
    build_module()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
