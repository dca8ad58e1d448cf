"""parallel_differential_expression on the MI355X backend.

Same signature, steps, messages and output DataFrame as the reference's
driver (src/hpdex/de.py:35-196).  Its numeric part -- group means with zeros,
fold change and log2 fold change, the Mann-Whitney U test with zero_handling
"min", and the Benjamini-Hochberg FDR -- runs in one device call
(hpdex_de_csc) instead of group_mean + mannwhitneyu + numpy +
scipy.stats.false_discovery_control on the host.

`adata` is anything with AnnData's `obs` (pandas DataFrame), `X` (numpy array
or scipy CSR/CSC matrix, cells x genes), `var_names` and `n_vars`.  X goes
to the device as it is and becomes a CSC there (hpdex_matrix_create_from).

The row labels keep the reference's layout: target/feature labels and fold
changes are target-major, while p_value, u_statistic and fdr are gene-major as
the backend returns them (de.py:160-196), so rows mix pairs exactly as the
reference's do whenever there is more than one target.

Steps 0-2 (metric and key checks, reference handling, the min_samples filter,
the group-id map) and the label/DataFrame assembly reproduce the reference's
text (de.py:53-103, 160-196): Copyright (c) 2025 Wang Zhongqi, MIT License --
the full notice is in the repository's NOTICE file.
"""
from __future__ import annotations

import logging

import numpy as np
import pandas as pd
import scipy.sparse

from . import _capi

logger = logging.getLogger(__name__)

supported_metrics = ["wilcoxon"]

COLUMNS = ["target", "feature", "p_value", "u_statistic", "fold_change", "log2_fold_change", "fdr"]


def parallel_differential_expression(
    adata,
    groupby_key: str,
    reference: str | None,   # None means pd.NA
    groups: list[str] | None = None,
    metric: str = "wilcoxon",
    tie_correction: bool = True,
    use_continuity: bool = True,
    min_samples: int = 2,
    threads: int = -1,
    clip_value: float = 20.0,
    show_progress: bool = True,
) -> pd.DataFrame:
    """Differential expression of every group against `reference` with the
    Mann-Whitney U test (de.py:35-196)."""
    # Step 0 (de.py:53-57)
    if metric not in supported_metrics:
        raise ValueError(f"Unsupported metric: {metric}; supported: {supported_metrics}")
    if groupby_key not in adata.obs.columns:
        raise ValueError(f"Groupby key `{groupby_key}` not found in adata.obs")
    obs = adata.obs.copy()

    # Step 1: reference (de.py:62-76)
    if reference is None:
        logger.info("No reference provided, using NA -> `non-targeting` as baseline")
        obs[groupby_key] = obs[groupby_key].cat.add_categories("non-targeting")
        obs[groupby_key] = obs[groupby_key].fillna("non-targeting")
        reference = "non-targeting"
    uniq = obs[groupby_key].unique().tolist()
    if groups is None:
        groups = [g for g in uniq if g != reference]
    else:
        groups = [g for g in groups if g in uniq and g != reference]
    if reference not in uniq:
        raise ValueError(f"Reference `{reference}` not found in `{groupby_key}`")
    if not groups:
        raise ValueError("No valid target groups found")

    # at least min_samples cells per group (de.py:79-91)
    valid_groups = []
    for g in groups:
        count = np.sum(obs[groupby_key] == g)
        if count >= min_samples:
            valid_groups.append(g)
        else:
            logger.warning(f"Group `{g}` has only {count} samples, skipping")
    groups = valid_groups
    if not groups:
        logger.warning(f"No groups have at least {min_samples} samples")
        return pd.DataFrame([], columns=COLUMNS)

    # Step 2: groups -> ids (de.py:96-103)
    group_map = {reference: 0}
    for i, g in enumerate(groups):
        group_map[g] = i + 1
    group_id = np.full(len(obs), -1, dtype=np.int32)
    for g, idx in group_map.items():
        group_id[np.asarray(obs[groupby_key] == g)] = idx
    n_targets = len(groups)

    # Step 3: CSC (de.py:106-112), built on the device from X as it is: a
    # dense array, CSR or CSC, with its own index width (no host tocsc, no
    # int64 copies)
    if isinstance(adata.X, np.ndarray):
        matrix = _capi.DeviceMatrix.from_dense(adata.X)
    elif scipy.sparse.issparse(adata.X) and adata.X.format in ("csr", "csc"):
        matrix = _capi.DeviceMatrix.from_scipy(adata.X)
    else:
        raise ValueError("sparse_matrix must be CSR or CSC")

    # Steps 4, 5, 7 on the device (de.py:126-184)
    try:
        U1, P, fc, lfc, fdr = _capi.de_matrix(matrix, group_id, n_targets,
                                              tie_correction=tie_correction,
                                              use_continuity=use_continuity,
                                              clip_value=clip_value)
    finally:
        matrix.close()

    # Step 6/8: labels and output (de.py:160-196)
    n_genes = adata.n_vars
    features = np.asarray(adata.var_names, dtype=object)
    targets_flat = np.repeat(np.asarray(groups, dtype=object), n_genes)
    features_flat = np.tile(features, n_targets)
    return pd.DataFrame({
        "target": targets_flat,
        "feature": features_flat,
        "p_value": P,
        "u_statistic": U1,
        "fold_change": fc,
        "log2_fold_change": lfc,
        "fdr": fdr,
    })
