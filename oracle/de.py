"""CPU restatement of the DE post-processing the reference runs after the test
(src/hpdex/de.py:126-184).  Test infrastructure only: used by tests/ as the
checker of the device path, never by the product.

fold_change: de.py:126-142 -- group means with zeros reshaped (C, G).T, fold
change target / reference with the clip rules, log2 with numpy.

fdr_bh: the Benjamini-Hochberg step of de.py:172-184, which calls
scipy.stats.false_discovery_control(P, method="bh") (scipy 1.15.3, present
here: scipy/stats/_morestats.py, false_discovery_control): sort ascending,
scale the k-th (1-based) of m by fl(m / k), running minimum from the end,
restore order, clip to [0, 1]; any value outside [0, 1] or NaN makes scipy
raise and de.py fall back to NaN everywhere.  Pinned against scipy itself in
tests/test_oracle_de.py.
"""
from __future__ import annotations

import numpy as np


def fold_change(means_cg: np.ndarray, n_cols: int, n_targets: int, clip_value: float = 20.0):
    """means_cg: group means as group_mean returns them (c*G + g).  Returns
    (fc, log2fc), each n_targets x n_cols target-major, flattened (de.py:
    fold_changes.reshape(-1))."""
    means = np.asarray(means_cg, dtype=np.float64).reshape(n_cols, n_targets + 1).T
    ref, tar = means[0], means[1:]
    with np.errstate(divide="ignore", invalid="ignore"):
        fc = tar / ref
        fc = np.where(ref < 1e-10, clip_value, fc)
        fc = np.where(tar < 1e-10, 1.0 / clip_value, fc)
        lfc = np.log2(fc)
    return fc.reshape(-1), lfc.reshape(-1)


def fdr_bh(p: np.ndarray) -> np.ndarray:
    p = np.asarray(p, dtype=np.float64).reshape(-1)
    m = p.size
    if m == 0:
        return p.copy()
    if not np.all(p == np.clip(p, 0.0, 1.0)):
        return np.full(m, np.nan)
    if m == 1:
        return p.copy()
    order = np.argsort(p, kind="stable")
    adj = p[order] * (m / np.arange(1, m + 1))
    adj = np.minimum.accumulate(adj[::-1])[::-1]
    out = np.empty(m)
    out[order] = adj
    return np.clip(out, 0.0, 1.0)
