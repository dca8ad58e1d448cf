"""Build the MI355X backend in-tree.

Produces
  hpdex_amd/lib/libhpdex_mi355x.so        C-ABI + gfx950 kernels (hipcc)
  hpdex_amd/backend/bin/kernel*.so        Python extension over the C-ABI (g++, pybind11)

Run as ``python -m hpdex_amd.build`` or through ``__graft_entry__.build()``.
Every call recompiles.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "hpdex_amd")
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
BINDIR = os.path.join(PKG, "backend", "bin")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(LIBDIR, "libhpdex_mi355x.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _run(cmds):
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
             for c in cmds]
    failed = []
    for c, p in procs:
        out, _ = p.communicate()
        lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln]
        if p.returncode != 0:
            failed.append((c, "\n".join(lines)))
    if failed:
        msg = "\n\n".join(" ".join(c) + "\n" + o for c, o in failed)
        raise RuntimeError("build failed:\n" + msg)


def ext_suffix() -> str:
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


def build(verbose: bool = False) -> None:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(BINDIR, exist_ok=True)
    k_obj = os.path.join(BUILD, "hpdex_kernels.o")
    a_obj = os.path.join(BUILD, "hpdex_api.o")
    s_obj = os.path.join(BUILD, "hpdex_session.o")
    d_obj = os.path.join(BUILD, "hpdex_de.o")
    i_obj = os.path.join(BUILD, "hpdex_ingest.o")
    _run([
        [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "hpdex_kernels.hip"), "-o", k_obj],
        [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "hpdex_de.hip"), "-o", d_obj],
        [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "hpdex_ingest.hip"), "-o", i_obj],
        [HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, "hpdex_api.cpp"), "-o", a_obj],
        [HIPCC, *HIP_FLAGS, "-pthread", "-x", "hip", "-c", os.path.join(CSRC, "hpdex_session.cpp"),
         "-o", s_obj],
    ])
    tmp = LIB + ".tmp"
    _run([[HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", k_obj, d_obj, i_obj, a_obj,
           s_obj, "-o", tmp, "-Wl,-soname,libhpdex_mi355x.so"]])
    os.replace(tmp, LIB)

    import pybind11  # noqa: WPS433 (build-time only)
    py_inc = sysconfig.get_paths()["include"]
    ext = os.path.join(BINDIR, "kernel" + ext_suffix())
    tmp = ext + ".tmp"
    _run([["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
           "-I", pybind11.get_include(), "-I", py_inc, "-I", os.path.join(ROOT, "include"),
           os.path.join(CSRC, "kernel_module.cpp"), "-o", tmp,
           "-L", LIBDIR, "-lhpdex_mi355x", "-Wl,-rpath,$ORIGIN/../../lib"]])
    os.replace(tmp, ext)
    if verbose:
        print("built", LIB, "and", ext)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
