"""Build the host simulator of the gfx950 kernels (test/debug tool).

The device part of hpdex_kernels.hip (everything above its launcher section)
is copied, with the two device-only constructs replaced -- the dynamic LDS
declaration becomes the simulator's buffer and the empty register-pinning
asm statements are dropped -- and compiled with g++ and AddressSanitizer
together with sim_main.cpp into _build/cpu_sim/libhpdex_cpusim.so.
"""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "hpdex_amd", "csrc")
OUT = os.path.join(ROOT, "_build", "cpu_sim")
LIB = os.path.join(OUT, "libhpdex_cpusim.so")


def _device_text(path, cut_marker=None):
    s = open(path).read()
    if cut_marker:
        s = s.split(cut_marker)[0]
    s = s.replace("extern __shared__ __align__(16) unsigned char smem[];",
                  "unsigned char* smem = sim::g_lds;")
    s = re.sub(r'asm volatile\(""\s*:\s*"\+v"\((\w+)\)\);', r"(void)\1;", s)
    # address spaces, clang vector types and non-temporal loads have no g++
    # spelling: plain pointers, structs with the same members, plain loads
    s = s.replace("#define HPDEX_GLOBAL __attribute__((address_space(1)))", "#define HPDEX_GLOBAL")
    s = s.replace("#define HPDEX_LDS __attribute__((address_space(3)))", "#define HPDEX_LDS")
    s = re.sub(r"typedef (\w+) (\w+)x4 __attribute__\(\(ext_vector_type\(4\)\)\);",
               r"struct \2x4 { \1 x, y, z, w; };", s)
    s = re.sub(r"typedef (\w+) (\w+)x2 __attribute__\(\(ext_vector_type\(2\)\)\);",
               r"struct \2x2 { \1 x, y; };", s)
    s = s.replace("__builtin_nontemporal_load(", "sim_plain_load(")
    # workgroup-static LDS objects live in a simulator buffer that is filled
    # with a garbage pattern before every workgroup, as device LDS is
    s = re.sub(r"__shared__ (\w+) (\w+);",
               r"\1& \2 = *reinterpret_cast<\1*>(sim::static_lds(sizeof(\1)));", s)
    return s


def build(sanitize=True, compiler="g++", opt="-O1", lib=None):
    """Build under an exclusive lock (parallel test workers may ask at once);
    the library is written beside its target and renamed over it."""
    import fcntl
    os.makedirs(OUT, exist_ok=True)
    with open(os.path.join(OUT, ".build.lock"), "w") as lk:
        fcntl.flock(lk, fcntl.LOCK_EX)
        return _build(sanitize, compiler, opt, lib)


def _build(sanitize, compiler, opt, lib):
    target = lib or LIB
    tag = f".{os.getpid()}"
    dev = _device_text(os.path.join(CSRC, "hpdex_device.h"))
    ker = _device_text(os.path.join(CSRC, "hpdex_kernels.hip"),
                       "// ---------------------------------------------------------------------------\n// launchers")
    ker = ker.replace('#include "hpdex_device.h"', dev) + "\n}  // namespace hpdex\n"
    inc = os.path.join(OUT, "hpdex_kernels_device.inc")
    with open(inc + tag, "w") as f:
        f.write(ker)
    os.replace(inc + tag, inc)
    flags = [opt, "-g", "-std=c++20", "-fPIC", "-shared", "-pthread", "-ffp-contract=off",
             "-Wno-unknown-pragmas", "-Wno-attributes"]
    if sanitize:
        flags += ["-fsanitize=address,undefined", "-fno-omit-frame-pointer"]
    cmd = [compiler, *flags, "-I", HERE, "-I", OUT, "-I", CSRC,
           os.path.join(HERE, "sim_main.cpp"), "-o", target + tag]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-6000:])
    os.replace(target + tag, target)
    return target


if __name__ == "__main__":
    print(build(sanitize="--no-asan" not in sys.argv))
