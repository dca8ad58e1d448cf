"""The host side of large uploads (hpdex_amd/csrc/hpdex_stage.cpp) on the CPU:
the copy pool that fills the pinned staging buffers, under ThreadSanitizer
(concurrent callers, jobs of random widths, every part run exactly once), and
the compact upload's fill under Address and UndefinedBehavior sanitizers
(round trips through the device's decoding rule, every disqualifying entry
refused)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(os.path.dirname(HERE), "hpdex_amd", "csrc", "hpdex_stage.cpp")

HEADERS = """#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <functional>
#include <mutex>
#include <limits>
#include <random>
#include <thread>
#include <vector>
"""


def test_copy_pool_under_thread_sanitizer(tmp_path):
    src = open(SRC).read()
    pool = src[src.index("class CopyPool {"):src.index("struct Ring {")]
    main = open(os.path.join(HERE, "copy_pool", "pool_main.cpp")).read()
    cpp = tmp_path / "pool.cpp"
    cpp.write_text(HEADERS + pool + main)
    exe = tmp_path / "pool"
    r = subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=thread", "-pthread",
                        str(cpp), "-o", str(exe)], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr[-3000:])
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr[-3000:]
    assert "bad 0" in r.stdout


def test_compact_fill_round_trip_and_refusals(tmp_path):
    src = open(SRC).read()
    fill = src[src.index("// Compact words (row & 0xFFFF)"):
               src.index("}  // namespace\n\ntemplate <typename IDX>\nhipError_t upload_compact(")]
    main = open(os.path.join(HERE, "copy_pool", "compact_main.cpp")).read()
    cpp = tmp_path / "compact.cpp"
    cpp.write_text(HEADERS + "#include <cstdint>\n" + fill + main)
    exe = tmp_path / "compact"
    r = subprocess.run(["g++", "-O1", "-g", "-std=c++20", "-Wno-unknown-pragmas",
                        "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                        str(cpp), "-o", str(exe)], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr[-3000:])
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr[-3000:]
    assert "bad 0" in r.stdout
