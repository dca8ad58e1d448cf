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


# Register budget of every kernel, checked on the compiler's resource report
# at each build: spilled SGPRs go through VGPR lanes and spilled VGPRs to
# scratch memory.  A kernel far beyond these (hundreds of SGPR spills) is the
# profile of one that once stalled the device; none may come near it again.
SGPR_SPILL_BUDGET = 48
VGPR_SPILL_BUDGET = 8
RESOURCE_REMARKS = "-Rpass-analysis=kernel-resource-usage"


def _run(cmds):
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
             for c in cmds]
    failed, outs = [], []
    for c, p in procs:
        out, _ = p.communicate()
        lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln]
        outs.append("\n".join(lines))
        if p.returncode != 0:
            failed.append((c, "\n".join(lines)))
    if failed:
        msg = "\n\n".join(" ".join(c) + "\n" + o for c, o in failed)
        raise RuntimeError("build failed:\n" + msg)
    return outs


def kernel_resources(report: str) -> dict:
    """{kernel: {"VGPRs": n, "SGPRs Spill": n, "VGPRs Spill": n, ...}} from a
    compile's resource-usage remarks."""
    import re
    out, cur = {}, None
    for ln in report.splitlines():
        m = re.search(r"Function Name: (\S+)", ln)
        if m:
            cur = out.setdefault(m.group(1), {})
            continue
        m = re.search(r"remark:\s+([^:]+): (\d+)", ln)
        if m and cur is not None:
            cur[m.group(1).strip()] = int(m.group(2))
    return out


def check_resources(report: str) -> dict:
    """Raise if a kernel spills more registers than the budgets allow."""
    res = kernel_resources(report)
    if not res:
        raise RuntimeError("no kernel resource report in the compiler output")
    over = [f"{k}: {v.get('SGPRs Spill', 0)} SGPR / {v.get('VGPRs Spill', 0)} VGPR spills"
            for k, v in res.items()
            if v.get("SGPRs Spill", 0) > SGPR_SPILL_BUDGET or
            v.get("VGPRs Spill", 0) > VGPR_SPILL_BUDGET]
    if over:
        raise RuntimeError(f"kernels over the register budget ({SGPR_SPILL_BUDGET} SGPR, "
                           f"{VGPR_SPILL_BUDGET} VGPR spills): " + "; ".join(over))
    return res


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
    g_obj = os.path.join(BUILD, "hpdex_stage.o")
    outs = _run([
        [HIPCC, *HIP_FLAGS, RESOURCE_REMARKS, "-c", os.path.join(CSRC, "hpdex_kernels.hip"), "-o",
         k_obj],
        [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "hpdex_de.hip"), "-o", d_obj],
        [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "hpdex_ingest.hip"), "-o", i_obj],
        [HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, "hpdex_api.cpp"), "-o", a_obj],
        [HIPCC, *HIP_FLAGS, "-pthread", "-x", "hip", "-c", os.path.join(CSRC, "hpdex_session.cpp"),
         "-o", s_obj],
        [HIPCC, *HIP_FLAGS, "-pthread", "-x", "hip", "-c", os.path.join(CSRC, "hpdex_stage.cpp"),
         "-o", g_obj],
    ])
    res = check_resources(outs[0])
    if verbose:
        worst = max(res.items(), key=lambda kv: kv[1].get("SGPRs Spill", 0))
        print(f"{len(res)} kernels within the register budget (most SGPR spills: "
              f"{worst[1].get('SGPRs Spill', 0)})")
    tmp = LIB + ".tmp"
    _run([[HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", k_obj, d_obj, i_obj, a_obj,
           s_obj, g_obj, "-o", tmp, "-Wl,-soname,libhpdex_mi355x.so", "-Wl,--no-undefined"]])
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
