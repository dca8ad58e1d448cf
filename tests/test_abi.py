"""The C-ABI library and the Python wrapper on the CPU: the library loads and
exports every function include/hpdex_mi355x.h declares; argument validation
matches the reference wrapper (src/hpdex/backend/mannwhitneyu_.py:44-67)."""
import ctypes
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

from hpdex_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HAS_GPU = os.path.exists("/dev/kfd")


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names.update(re.findall(r"\b(hpdex_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_header():
    L = _capi.lib()
    names = declared_functions()
    assert names, "no declarations found"
    for n in sorted(names):
        assert hasattr(L, n), f"{n} not exported"
    assert set(_capi.EXPORTED) <= names


def test_version_string():
    assert b"gfx950" in _capi.lib().hpdex_version()


def test_wrapper_validation():
    from hpdex_amd.backend import group_mean, mannwhitneyu
    M = sp.csc_matrix(np.ones((6, 2), dtype=np.float32))
    with pytest.raises(ValueError, match="sparse_matrix must be CSR or CSC"):
        mannwhitneyu(np.ones((6, 2)), np.zeros(6), 1)
    with pytest.raises(ValueError, match=r"group_id must be 1D with length 6 \(rows of CSC / cols of CSR\)"):
        mannwhitneyu(M, np.zeros(5), 1)
    with pytest.raises(ValueError, match="group_id must be 1D with length 2"):
        group_mean(M.tocsr(), np.zeros(6), 2)
    with pytest.raises(KeyError):
        mannwhitneyu(M, np.zeros(6), 1, zero_handling="median")
    with pytest.raises(KeyError):
        mannwhitneyu(M, np.zeros(6), 1, alternative="two_sided")
    with pytest.raises(RuntimeError, match=r"n_targets must be >= 1"):
        mannwhitneyu(M, np.zeros(6), 0)
    with pytest.raises(RuntimeError, match="Unsupported dtype for data"):
        mannwhitneyu(sp.csc_matrix(np.ones((6, 2), dtype=np.bool_)), np.zeros(6), 1)


@pytest.mark.skipif(HAS_GPU, reason="checks the no-device error")
def test_no_cpu_fallback():
    from hpdex_amd.backend import mannwhitneyu
    M = sp.csc_matrix(np.arange(12, dtype=np.float32).reshape(6, 2))
    with pytest.raises(RuntimeError, match="no HIP device available"):
        mannwhitneyu(M, np.array([0, 0, 0, 1, 1, 1]), 1)


def test_kernel_module_defaults():
    from hpdex_amd.backend.bin import kernel
    doc = kernel.mannwhitneyu.__doc__
    for frag in ("ref_sorted: bool = False", "fast_norm: bool = True", "= 0, alternative",
                 "SupportsIndex = 2, method", "layout: str = 'csc'", "show_progress: bool = False"):
        assert frag in doc, frag
    assert "include_zeros: bool = True" in kernel.group_mean.__doc__


def test_register_budget_check_reads_the_compiler_report():
    """The build's register check parses the compiler's resource remarks and
    refuses a kernel over the spill budget (a report, not a device run)."""
    from hpdex_amd import build
    report = "\n".join([
        "x.hip:1:1: remark: Function Name: _Z1kv [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:1:1: remark:     VGPRs: 100 [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:1:1: remark:     SGPRs Spill: 2 [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:1:1: remark:     VGPRs Spill: 0 [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:2:1: remark: Function Name: _Z1jv [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:2:1: remark:     SGPRs Spill: 280 [-Rpass-analysis=kernel-resource-usage]",
    ])
    res = build.kernel_resources(report)
    assert res["_Z1kv"] == {"VGPRs": 100, "SGPRs Spill": 2, "VGPRs Spill": 0}
    with pytest.raises(RuntimeError, match="_Z1jv: 280 SGPR"):
        build.check_resources(report)
    build.check_resources("\n".join(report.splitlines()[:4]))
