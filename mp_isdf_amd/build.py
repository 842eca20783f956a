"""Build the gfx950 HIP library and the ``mp_isdf`` pybind11 module in-tree.

Outputs (next to this file, git-ignored, shipped with the source tree):

* ``libmp_isdf_amd.so``  -- HIP kernels + C ABI (include/mp_isdf_amd.h)
* ``mp_isdf<EXT_SUFFIX>`` -- pybind11 module, host C++ only, linked against
  the C ABI with an ``$ORIGIN`` rpath.

Every HIP translation unit is compiled by ``hipcc --offload-arch=gfx950`` with
packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libmp_isdf_amd.so"
EXT = PKG / ("mp_isdf" + sysconfig.get_config_var("EXT_SUFFIX"))

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-I", str(ROOT / "include"),
]
SOURCES = [
    "common.cpp", "context.cpp", "axis.cpp", "select.cpp", "online.cpp", "capi.cpp", "grid.cpp", "lebedev.cpp", "overlap.cpp", "rbrp.cpp",
    "ao_eval.hip", "gemm.hip", "gemm_tn.hip", "qrcp.hip", "sketch.hip", "util.hip", "online_kernels.hip", "sketch_mfma.hip", "fold_bf16x3.hip", "fold_i8.hip", "fit.hip", "coulomb.hip", "cvt.hip", "audit.hip", "eri.hip",
]


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")
    return proc


def _compile(src: str) -> Path:
    out = OBJ / (src.replace(".", "_") + ".o")
    cmd = [HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(out)]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]  # host files include hip_runtime for the device API
    _run(cmd)
    return out


def build(verbose: bool = False) -> None:
    OBJ.mkdir(exist_ok=True)
    jobs = min(len(SOURCES), max(1, min(16, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as pool:
        objects = list(pool.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objects), "-lpthread"])
    os.replace(tmp, LIB)
    import pybind11
    ext_tmp = EXT.with_name(EXT.name + ".tmp")
    _run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
          "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"],
          str(CSRC / "pybind_module.cpp"), "-o", str(ext_tmp),
          "-L", str(PKG), "-lmp_isdf_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(ext_tmp, EXT)
    if verbose:
        print(f"built {LIB.name} and {EXT.name}")


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
