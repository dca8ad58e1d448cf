"""The host copy pool that fills the pinned staging buffers of large uploads
(hpdex_amd/csrc/hpdex_stage.cpp), run on the CPU under ThreadSanitizer:
concurrent callers, jobs of random widths, every part run exactly once."""
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
