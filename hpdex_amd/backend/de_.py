"""hpdex.pden -- the reference's numba pipeline -- on the MI355X backend.

Same signature, steps, messages and output DataFrame as
src/hpdex/backend/de_.py:508-658 (exported as hpdex.pden and
hpdex.backend.parallel_differential_expression_numba).  Where the reference
densifies X, sorts the reference on the host and runs numba kernels over
(group, gene chunk) tasks in worker processes, here X goes to the device as
it is (a dense array, CSR or CSC) and one call, hpdex_pden_matrix, returns
every group's target U, p-value, fold change, log2 fold change and FDR.

What the device reproduces of the numba path:
  * ranks of the densified matrix: zeros (stored or not) at their true
    position, n = the group's cell count, no minimum sample size check;
  * u_statistic is the TARGET group's U (pykernel_.py:74-76, 394-396), not
    the reference group's U that hpdex.backend.mannwhitneyu returns;
  * the per-task choice of the exact or asymptotic p (pykernel_.py:267-271,
    489-493): a task is one group and one chunk of genes, chunk_size as
    _auto_schedule_chunk_size picks it (de_.py:291-333) -- all genes with
    one worker;
  * the histogram kernel (prefer_hist_if_int, an integer X, num_workers > 1)
    and its fallback to the float merge with the asymptotic p forced when a
    task's value range passes max_bins (pykernel_.py:466-480);
  * fold changes from numpy's means (float32 arithmetic for a float32 X,
    float64 otherwise) with the clip rules of de_.py:189-230, and the BH FDR
    over all p-values (de_.py:250-288).

Known differences, all at inputs where the reference does not return:
non-finite values (its merge loops forever on NaN), a task whose worker
raises (the reference logs it and leaves NaN; here the call raises), exact
tables of more than 2^27 cells (the reference's int64 DP runs out of memory
or overflows), and duplicate names in `groups` (the reference fails while
assigning fold changes; here each name is tested once).

Steps of the public function (group selection, its messages, the task
size rule, the DataFrame assembly) follow the reference's text
(de_.py:291-333, 508-658): Copyright (c) 2025 Wang Zhongqi, MIT License --
the full notice is in the repository's NOTICE file.
"""
from __future__ import annotations

import logging
from typing import List, Optional

import numpy as np
import pandas as pd
import scipy.sparse

from .. import _capi

logger = logging.getLogger(__name__)

SUPPORTED_METRICS = ["wilcoxon"]


def _auto_schedule_chunk_size(n_genes: int, n_groups: int, num_workers: int,
                              data_size_mb: float) -> int:
    """Genes per task (de_.py:291-333)."""
    if num_workers == 1:
        chunk_size = n_genes
        logger.info(f"Auto-scheduling: Single worker mode, chunk_size={chunk_size} (all genes)")
    else:
        target_chunks_per_worker = 1
        if data_size_mb > 1024:
            target_chunks_per_worker = 2
        target_total_chunks = num_workers * target_chunks_per_worker
        chunk_size = max(4096, n_genes // target_total_chunks)
        logger.info(f"Auto-scheduling: {num_workers} workers, target {target_chunks_per_worker} "
                    f"chunks/worker, chunk_size={chunk_size} genes/chunk")
    return chunk_size


def _estimate_data_size(adata) -> float:
    """The dense size in MB that scheduling uses, for sparse and dense X alike
    (de_.py:336-367)."""
    n_cells, n_genes = adata.shape
    return n_cells * n_genes * 8 / 1024 / 1024


def _check_dtype(dtype: np.dtype) -> None:
    # pykernel_.py:202-217: float16 and non-numeric dtypes are refused
    if dtype == np.float16:
        raise TypeError("float16 is not supported (Numba & numerical stability).")
    if not (np.issubdtype(dtype, np.integer) or np.issubdtype(dtype, np.floating)):
        raise TypeError(f"Unsupported dtype {dtype}.")


def _device_matrix(X):
    if isinstance(X, np.ndarray):
        return _capi.DeviceMatrix.from_dense(X)
    if scipy.sparse.issparse(X):
        if X.format not in ("csr", "csc"):
            X = X.tocsr()
        return _capi.DeviceMatrix.from_scipy(X)
    return _capi.DeviceMatrix.from_dense(np.asarray(X))


def parallel_differential_expression(
    adata,
    groupby_key: str,
    reference: str,
    groups: Optional[List[str]] = None,
    metric: str = "wilcoxon",
    tie_correction: bool = True,
    continuity_correction: bool = True,
    use_asymptotic: Optional[bool] = None,
    min_samples: int = 2,
    max_bins: int = 100_000,
    prefer_hist_if_int: bool = False,
    num_workers: int = 1,
    clip_value: float = 20.0,
) -> pd.DataFrame:
    """Differential expression of every group against `reference` with the
    Mann-Whitney U test, as hpdex.pden computes it (de_.py:508-658).

    Returns a DataFrame with columns target, feature, p_value, u_statistic
    (the target group's U), fold_change, log2_fold_change and fdr, one block
    of n_vars rows per group."""
    # group selection (de_.py:565-597)
    obs_vals = adata.obs[groupby_key].values
    uniq = np.unique(obs_vals)
    if reference not in uniq:
        raise ValueError(f"reference `{reference}` not found in `{groupby_key}`")
    if groups is None:
        unique_set = set(uniq)
        unique_set.discard(reference)
        groups = list(unique_set)
    else:
        groups = [g for g in groups if g != reference and g in uniq]
    row_idx_ref = np.where(obs_vals == reference)[0].astype(np.int64)
    group_rows = []
    group_names = []
    for g in groups:
        if g in group_names:
            continue
        idx = np.where(obs_vals == g)[0].astype(np.int64)
        if len(idx) >= min_samples:
            group_rows.append(idx)
            group_names.append(g)
    if not group_rows:
        raise ValueError("No target groups meet minimum sample size requirement")
    n_genes = adata.n_vars
    if metric not in SUPPORTED_METRICS:
        raise ValueError(f"Unsupported statistical test method: {metric}. "
                         f"Available: {SUPPORTED_METRICS}")
    dtype = np.dtype(adata.X.dtype)
    _check_dtype(dtype)
    use_hist = bool(prefer_hist_if_int and np.issubdtype(dtype, np.integer) and num_workers > 1)

    # task size (de_.py:608-614); one worker runs each group over all genes
    data_size_mb = _estimate_data_size(adata)
    chunk_size = _auto_schedule_chunk_size(n_genes, len(group_rows), num_workers, data_size_mb)
    chunk_genes = chunk_size if num_workers > 1 else n_genes

    group_id = np.full(adata.n_obs, -1, dtype=np.int32)
    group_id[row_idx_ref] = 0
    for i, idx in enumerate(group_rows):
        group_id[idx] = i + 1
    n_targets = len(group_rows)
    logger.info(f"Data: {adata.n_obs} cells x {n_genes} genes, {n_targets} treatment groups on "
                "the device")

    matrix = _device_matrix(adata.X)
    try:
        U, P, fc, lfc, fdr = _capi.pden_matrix(
            matrix, group_id, n_targets, tie_correction=tie_correction,
            continuity_correction=continuity_correction, use_asymptotic=use_asymptotic,
            use_hist=use_hist, max_bins=max_bins, chunk_genes=max(int(chunk_genes), 1),
            clip_value=float(clip_value))
    finally:
        matrix.close()

    # one block of genes per group, in group order (de_.py:414-428, 487-502)
    gene_names = np.asarray(adata.var.index.values)
    return pd.DataFrame({
        "target": np.repeat(np.asarray(group_names, dtype=object), n_genes),
        "feature": np.tile(gene_names, n_targets),
        "p_value": P,
        "u_statistic": U,
        "fold_change": fc,
        "log2_fold_change": lfc,
        "fdr": fdr,
    })
