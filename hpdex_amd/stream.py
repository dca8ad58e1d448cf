"""Memory-bounded streaming differential expression on the MI355X backend.

The reference's streaming driver (src/hpdex/stream.py:365-687) handles datasets
larger than memory by reading a backed .h5ad group by group: the reference
group once, then batches of target groups packed greedily under
`memory_limit_gb` (stream.py:161-262), each batch through the numba pipeline,
with one BH FDR over all p-values at the end (stream.py:682-684).

Here the stream runs over gene columns instead: columns are independent
(src/cpp/mannwhitneyu.cpp:903), so a CSC matrix whose arrays may live on disk
(numpy memmaps, or any arrays that slice) is cut into runs of consecutive
genes, each small enough that its host copy and device footprint fit
memory_limit_gb / num_workers (the same greedy packing, over contiguous gene
ranges so the gene-major outputs stay contiguous; a single gene that does not
fit raises MemoryError, as the reference does for a group).  Each run is one
hpdex_de_csc call; num_workers calls are in flight at once, so one run's read
and upload overlap another's kernels.  The FDR is recomputed once over all
p-values.  Every value equals the one-shot hpdex_de_csc result on the whole
matrix bit for bit.
"""
from __future__ import annotations

import logging
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pandas as pd

from . import _capi

logger = logging.getLogger(__name__)

COLUMNS = ["target", "feature", "p_value", "u_statistic", "fold_change", "log2_fold_change", "fdr"]

# Host bytes of a run in flight: its values read from the source and their
# int64 row indices and column pointers (copies made for the call), and the
# five float64 outputs of every (gene, target) pair the call returns.
HOST_BYTES_PER_VALUE_ROW = 8
HOST_BYTES_PER_PAIR = 40


def run_bytes(n_values, n_genes, n_targets, n_rows=0, dtype=np.float32):
    """Upper bound of the memory a run of n_genes genes holding n_values stored
    values takes while in flight: its host copies and outputs, plus the device
    memory of its hpdex_de_csc call as the library bounds it
    (hpdex_de_csc_device_bytes: the upload, the workspace of every gene and
    of the n_rows labels, group means, fold changes and FDR)."""
    dev = _capi.de_csc_device_bytes(dtype, n_rows, n_genes, n_values, n_targets)
    host = (n_values * (np.dtype(dtype).itemsize + HOST_BYTES_PER_VALUE_ROW) +
            (n_genes + 1) * 8 + n_genes * n_targets * HOST_BYTES_PER_PAIR)
    return dev + host


def plan_gene_runs(indptr, n_targets, budget_bytes, n_rows=0, dtype=np.float32):
    """Consecutive gene ranges [c0, c1), each within budget_bytes (greedy, in
    gene order); MemoryError naming the first gene that does not fit alone."""
    indptr = np.asarray(indptr, dtype=np.int64)
    C = len(indptr) - 1
    runs, c0, vals = [], 0, 0
    cost = lambda nv, ng: run_bytes(nv, ng, n_targets, n_rows, dtype)
    for c in range(C):
        n = int(indptr[c + 1] - indptr[c])
        if cost(n, 1) > budget_bytes:
            raise MemoryError(f"Gene {c} requires {cost(n, 1) / 2**30:.2f}GB but "
                              f"the budget per run is {budget_bytes / 2**30:.2f}GB; consider "
                              "increasing memory_limit_gb or reducing num_workers")
        if c > c0 and cost(vals + n, c + 1 - c0) > budget_bytes:
            runs.append((c0, c))
            c0, vals = c, 0
        vals += n
    if C > c0:
        runs.append((c0, C))
    return runs


def stream_de(data, indices, indptr, n_rows, group_id, n_targets, memory_limit_gb=16.0,
              num_workers=2, tie_correction=True, use_continuity=True, clip_value=20.0):
    """hpdex_de_csc over gene runs of a CSC matrix (cells x genes) whose
    arrays may be memmaps: (U1, P, fold_change, log2_fold_change, fdr) in
    hpdex_de_csc's layouts (fold changes target-major, the rest gene-major)."""
    if num_workers < 1:
        raise ValueError("num_workers must be >= 1")
    T = int(n_targets)
    indptr = np.asarray(indptr)
    C = len(indptr) - 1
    budget = int(memory_limit_gb * 2**30) // num_workers
    runs = plan_gene_runs(indptr, T, budget, n_rows, np.asarray(data[:0]).dtype)
    logger.info(f"streaming {C} genes in {len(runs)} runs, {num_workers} in flight")
    group_id = np.ascontiguousarray(group_id, dtype=np.int32)
    U1 = np.empty(C * T)
    P = np.empty(C * T)
    fc = np.empty((T, C))
    lfc = np.empty((T, C))

    def one(run):
        c0, c1 = run
        p0, p1 = int(indptr[c0]), int(indptr[c1])
        d = np.ascontiguousarray(data[p0:p1])          # reads the run from disk for memmaps
        r = np.ascontiguousarray(indices[p0:p1], dtype=np.int64)
        ptr = np.asarray(indptr[c0:c1 + 1], dtype=np.int64) - p0
        u, p, f, l, _ = _capi.de_csc(d, r, ptr, group_id, T, n_rows=n_rows,
                                     tie_correction=tie_correction,
                                     use_continuity=use_continuity, clip_value=clip_value)
        n = c1 - c0
        U1[c0 * T:c1 * T] = u
        P[c0 * T:c1 * T] = p
        fc[:, c0:c1] = f.reshape(T, n)
        lfc[:, c0:c1] = l.reshape(T, n)

    if num_workers == 1:
        for run in runs:
            one(run)
    else:
        with ThreadPoolExecutor(num_workers) as ex:
            for f in [ex.submit(one, run) for run in runs]:
                f.result()
    fdr = _capi.fdr_bh(P)
    return U1, P, fc.reshape(-1), lfc.reshape(-1), fdr


def parallel_differential_expression(
    adata,
    groupby_key: str,
    reference: str,
    groups: list[str] | None = None,
    memory_limit_gb: float = 16.0,
    metric: str = "wilcoxon",
    tie_correction: bool = True,
    continuity_correction: bool = True,
    min_samples: int = 2,
    num_workers: int = 2,
    clip_value: float = 20.0,
) -> pd.DataFrame:
    """Streaming counterpart of stream.py:365-687 on an AnnData-shaped object
    whose X is a CSC matrix (cells x genes; its arrays may be memmaps).  Rows
    are target-major, as the reference's streaming driver concatenates them
    per group (stream.py:674-684); u_statistic is U1, as this backend's
    mannwhitneyu returns it."""
    if metric not in ("wilcoxon", "wilcoxon-hist"):
        raise ValueError(f"Unsupported statistical test method: {metric}. "
                         "Available: ['wilcoxon', 'wilcoxon-hist']")
    if groupby_key not in adata.obs.columns:
        raise ValueError(f"Groupby key `{groupby_key}` not found in adata.obs")
    labels = adata.obs[groupby_key]
    uniq = labels.unique().tolist()
    if reference not in uniq:
        raise ValueError(f"Reference group '{reference}' not found in {groupby_key}")
    counts = labels.value_counts()
    if groups is None:
        targets = [g for g in uniq if g != reference and counts[g] >= min_samples]
    else:
        targets = [g for g in groups if g != reference and g in uniq and counts[g] >= min_samples]
    if not targets:
        raise ValueError("No valid target groups found")
    group_id = np.full(len(labels), -1, dtype=np.int32)
    group_id[np.asarray(labels == reference)] = 0
    for i, g in enumerate(targets):
        group_id[np.asarray(labels == g)] = i + 1
    X = adata.X
    if getattr(X, "format", None) != "csc":
        raise ValueError("streaming needs X as a CSC matrix (cells x genes)")
    T, C = len(targets), X.shape[1]
    U1, P, fc, lfc, fdr = stream_de(X.data, X.indices, X.indptr, X.shape[0], group_id, T,
                                    memory_limit_gb=memory_limit_gb, num_workers=num_workers,
                                    tie_correction=tie_correction,
                                    use_continuity=continuity_correction, clip_value=clip_value)
    # gene-major (c*T + t) -> target-major (t*C + c)
    tm = lambda a: a.reshape(C, T).T.reshape(-1)
    return pd.DataFrame({
        "target": np.repeat(np.asarray(targets, dtype=object), C),
        "feature": np.tile(np.asarray(adata.var_names, dtype=object), T),
        "p_value": tm(P),
        "u_statistic": tm(U1),
        "fold_change": fc,
        "log2_fold_change": lfc,
        "fdr": tm(fdr),
    }, columns=COLUMNS)
