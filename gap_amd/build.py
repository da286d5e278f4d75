"""Builds the native libraries of gap_amd in-tree (gfx950 only).

  gap_amd/lib/libgap_hip.so          HIP kernels + engine + C ABI (include/gap_hip.h)
  gap_amd/lib/libgap_hip_plugins.so  C++20 plugin classes over the C ABI, exporting
                                     create_gpu_* (include/gap_hip_plugins.h), also
                                     linked as libgap_{admittance,lu,powerflow}_gpu.so
  gap_amd/gap_solver*.so             pybind11 module with the reference's Python surface
  gap_amd/lib/gap_hip_plugin_check   loads the plugins as the reference factory does and
                                     solves through the interfaces (integration check)

Run ``python -m gap_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import sysconfig
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
SRC = PKG / "csrc"
LIB = PKG / "lib"
OBJ = ROOT / "_build"
INC = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# FP32 packed math is disabled for all device code: packed fp32 adds/muls/fmas
# have produced wrong results on these boxes.
HIP_FLAGS = ["-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature",
             "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-result"]
CXX_FLAGS = ["-O3", "-std=c++20", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def _run(cmd):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True)


def _newer(target: Path, *deps: Path) -> bool:
    if not target.exists():
        return True
    t = target.stat().st_mtime
    return any(d.stat().st_mtime > t for d in deps)


def build(force: bool = True) -> None:
    LIB.mkdir(exist_ok=True)
    OBJ.mkdir(exist_ok=True)
    headers = list(SRC.glob("*.hpp")) + list(INC.glob("*.h")) + list(INC.glob("gap/**/*.h"))
    sym_o = OBJ / "symbolic.o"
    if force or _newer(sym_o, SRC / "symbolic.cpp", *headers):
        _run(["g++", *CXX_FLAGS, "-c", SRC / "symbolic.cpp", "-o", sym_o])
    eng_o = OBJ / "engine.o"
    if force or _newer(eng_o, SRC / "engine.hip", *headers):
        _run([HIPCC, *HIP_FLAGS, "-I", INC, "-c", SRC / "engine.hip", "-o", eng_o])
    core = LIB / "libgap_hip.so"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", core, eng_o, sym_o,
          "-Wl,-soname,libgap_hip.so"])
    plugins = LIB / "libgap_hip_plugins.so"
    _run(["g++", *CXX_FLAGS, "-shared", "-I", INC, SRC / "backends.cpp", SRC / "pgm_io.cpp", "-o", plugins,
          "-L", LIB, "-lgap_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-soname,libgap_hip_plugins.so"])
    # command-line front end (PGM JSON in and out)
    _run(["g++", *CXX_FLAGS, "-I", INC, SRC / "gap_hip_main.cpp", "-o", LIB / "gap_hip",
          "-L", LIB, "-lgap_hip_plugins", "-lgap_hip", "-Wl,-rpath,$ORIGIN"])
    # plugin-load check: the reference factory's dlopen / dlsym path, linked
    # against nothing of this project (finds the plugins by bare name)
    _run(["g++", *CXX_FLAGS, "-I", INC, SRC / "plugin_check.cpp", "-o", LIB / "gap_hip_plugin_check", "-ldl"])
    for alias in ("libgap_admittance_gpu.so", "libgap_lu_gpu.so", "libgap_powerflow_gpu.so"):
        dst = LIB / alias
        if dst.exists() or dst.is_symlink():
            dst.unlink()
        shutil.copyfile(plugins, dst)
    import pybind11
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    mod = PKG / f"gap_solver{ext}"
    py_inc = sysconfig.get_paths()["include"]
    _run(["g++", *CXX_FLAGS, "-shared", "-I", INC, "-I", pybind11.get_include(), "-I", py_inc,
          SRC / "pymodule.cpp", "-o", mod, "-L", LIB, "-lgap_hip_plugins", "-lgap_hip",
          "-Wl,-rpath,$ORIGIN/lib"])


if __name__ == "__main__":
    build(force="--incremental" not in sys.argv)
