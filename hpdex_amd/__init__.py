"""MI355X-native Mann-Whitney U differential-expression backend.

``hpdex_amd.backend.mannwhitneyu`` / ``group_mean`` keep the signatures of
hpdex's ``hpdex.backend`` functions; underneath they call the C-ABI in
include/hpdex_mi355x.h (gfx950 HIP kernels).  Import of the backend fails if
the native library is not built -- there is no CPU fallback.

The package names of the reference (src/hpdex/__init__.py:1-15) resolve on
first use: ``parallel_differential_expression`` / ``pde`` (hpdex.de),
``pden`` (the numba pipeline, hpdex.backend.de_) and
``parallel_differential_expression_stream`` (hpdex.stream).
"""
__version__ = "0.1.0"

_LAZY = {
    "parallel_differential_expression": ("hpdex_amd.de", "parallel_differential_expression"),
    "pde": ("hpdex_amd.de", "parallel_differential_expression"),
    "pden": ("hpdex_amd.backend.de_", "parallel_differential_expression"),
    "parallel_differential_expression_stream": ("hpdex_amd.stream",
                                                "parallel_differential_expression"),
}

__all__ = ["parallel_differential_expression", "parallel_differential_expression_stream", "pden",
           "pde"]


def __getattr__(name):
    if name in _LAZY:
        import importlib
        mod, attr = _LAZY[name]
        return getattr(importlib.import_module(mod), attr)
    raise AttributeError(f"module 'hpdex_amd' has no attribute {name!r}")
