"""MI355X-native Mann-Whitney U differential-expression backend.

``hpdex_amd.backend.mannwhitneyu`` / ``group_mean`` keep the signatures of
hpdex's ``hpdex.backend`` functions; underneath they call the C-ABI in
include/hpdex_mi355x.h (gfx950 HIP kernels).  Import of the backend fails if
the native library is not built -- there is no CPU fallback.
"""
__version__ = "0.1.0"
