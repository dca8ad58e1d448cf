"""CPU restatement of the reference's numba pipeline, hpdex.pden
(src/hpdex/backend/de_.py:508-658 with the kernels of
src/hpdex/backend/pykernel_.py) -- TEST INFRASTRUCTURE ONLY: tests/ use it as
the checker of hpdex_amd.backend.de_, the product never imports it.

  merge_stats       _merge_many_sorted_numba (pykernel_.py:8-81) and
                    _hist_merge_and_stats_kernel (:302-406): per row the
                    TARGET rank sum minus n_tar (n_tar + 1) / 2, the tie sum
                    sum c (c^2 - 1) over runs of c > 1 equal values, and
                    whether there is a run.  Restated over the sorted distinct
                    values (numpy unique) rather than a two-pointer merge;
                    every term is a multiple of 1/2 below 2^53, so the float64
                    sums are exact in any order.
  p_asymptotic      _p_asymptotic_batch_numba (:84-147), operation by
                    operation (n1 = n_tar, n2 = n_ref), math.erfc.
  exact_tail_table  _exact_tail_table (:150-199): the integer DP of the
                    Mann-Whitney null's counts, sf[k] = P(U >= k).
  rank_sum_float    rank_sum_chunk_kernel_float (:220-290).
  rank_sum_hist     rank_sum_chunk_kernel_hist (:410-509), its max_bins
                    fallback to the float merge with the asymptotic p forced.
  pden              parallel_differential_expression (de_.py:508-658): the
                    dense matrix, the sorted reference, one task per (group,
                    gene chunk) (de_.py:113-124, 291-333), numpy means and the
                    fold-change rules (de_.py:189-247, 640-652), BH FDR
                    (de_.py:250-288, scipy.stats.false_discovery_control).

Pinned in tests/test_oracle_pden.py against scipy.stats.mannwhitneyu
(asymptotic and exact), a brute-force enumeration of the null, closed-form
means and scipy's BH.
"""
from __future__ import annotations

import math

import numpy as np
import pandas as pd
import scipy.sparse
from scipy.stats import false_discovery_control


def merge_stats(ref2: np.ndarray, tar2: np.ndarray):
    """(U_tar, tie_sum, has_ties) of each row of ref2 [B, n_ref] and tar2
    [B, n_tar] (pykernel_.py:24-81)."""
    B, n_ref = ref2.shape
    n_tar = tar2.shape[1]
    U = np.zeros(B)
    tie = np.zeros(B)
    has = np.zeros(B, dtype=np.int64)
    if n_ref == 0 or n_tar == 0:
        return U, tie, has
    for b in range(B):
        v = np.concatenate([np.asarray(ref2[b], dtype=np.float64),
                            np.asarray(tar2[b], dtype=np.float64)])
        _, inv, cnt = np.unique(v, return_inverse=True, return_counts=True)
        ct = np.bincount(inv[n_ref:], minlength=len(cnt))
        start = np.cumsum(cnt) - cnt + 1           # the run's first rank
        rank_sum_t = float(np.sum(ct * (start + 0.5 * (cnt - 1))))
        U[b] = rank_sum_t - 0.5 * n_tar * (n_tar + 1.0)
        runs = cnt[cnt > 1].astype(np.int64)
        tie[b] = float(np.sum(runs * (runs * runs - 1)))
        has[b] = 1 if runs.size else 0
    return U, tie, has


def p_asymptotic(U1, tie_sum, n1: int, n2: int, tie_correction: int, continuity_correction: int):
    """pykernel_.py:84-147 with n1 = n_tar, n2 = n_ref."""
    U1 = np.asarray(U1, dtype=np.float64)
    p = np.empty(U1.shape[0])
    N = n1 + n2
    if N <= 1 or n1 == 0 or n2 == 0:
        p.fill(1.0)
        return p
    mu = 0.5 * n1 * n2
    base = (n1 * n2) * (N + 1.0) / 12.0
    use_tie = (tie_correction == 1) and (N > 1)
    k_tie = (n1 * n2) / (12.0 * N * (N - 1.0)) if use_tie else 0.0
    use_cc = continuity_correction == 1
    inv_sqrt2 = 0.7071067811865475
    for i in range(U1.shape[0]):
        sigma2 = base - k_tie * tie_sum[i] if use_tie else base
        if sigma2 <= 0.0 or not np.isfinite(sigma2):
            p[i] = 1.0
            continue
        num = U1[i] - mu
        if use_cc and num != 0.0:
            num = num - 0.5 if num > 0.0 else num + 0.5
        zabs = abs(num) / math.sqrt(sigma2)
        pj = math.erfc(zabs * inv_sqrt2)
        if not np.isfinite(pj) or pj > 1.0:
            pj = 1.0
        elif pj < 0.0:
            pj = 0.0
        p[i] = pj
    return p


def exact_tail_table(n1: int, n2: int) -> np.ndarray:
    """sf[k] = P(U >= k), k = 0 .. n1 n2 (pykernel_.py:150-199): f[j][k]
    counts the arrangements of i first-sample and j second-sample values
    with U = k, built row by row over i."""
    m, n = n1, n2
    Ucap = m * n
    f_prev = np.zeros((n + 1, Ucap + 1), dtype=np.int64)
    f_prev[:, 0] = 1
    for i in range(1, m + 1):
        f_curr = np.zeros_like(f_prev)
        f_curr[0, 0] = 1
        cap = i * n
        for j in range(1, n + 1):
            f_curr[j, :j] = f_curr[j - 1, :j]
            f_curr[j, j:cap + 1] = f_curr[j - 1, j:cap + 1] + f_prev[j, 0:cap + 1 - j]
        f_prev = f_curr
    counts = f_prev[n]
    total = int(np.sum(counts))
    sf = np.empty(Ucap + 1)
    acc = 0
    for k in range(Ucap, -1, -1):
        acc += int(counts[k])
        sf[k] = acc / float(total)
    return sf


def _p_of(U1, tie_sum, has_ties, n_ref, n_tar, tie_correction, continuity_correction,
          use_asymptotic):
    if use_asymptotic is None:
        use_asym = bool(np.any(has_ties != 0)) or (min(n_ref, n_tar) > 8)
    else:
        use_asym = bool(use_asymptotic)
    if use_asym:
        return p_asymptotic(U1, tie_sum, n_tar, n_ref, 1 if tie_correction else 0,
                            1 if continuity_correction else 0)
    total_U = int(n_ref * n_tar)
    U2 = total_U - U1
    Umax_idx = np.rint(np.maximum(U1, U2)).astype(np.int64)
    sf = exact_tail_table(n_tar, n_ref)
    p = 2.0 * sf[Umax_idx]
    return np.clip(p, 0.0, 1.0)


def rank_sum_float(ref_sorted, tar_sorted, tie_correction=True, continuity_correction=True,
                   use_asymptotic=None):
    """(p, U_tar) per row (pykernel_.py:220-290)."""
    U1, tie, has = merge_stats(ref_sorted, tar_sorted)
    return _p_of(U1, tie, has, ref_sorted.shape[1], tar_sorted.shape[1], tie_correction,
                 continuity_correction, use_asymptotic), U1


def rank_sum_hist(ref_sorted, tar, tie_correction=True, continuity_correction=True,
                  use_asymptotic=None, max_bins=200_000):
    """pykernel_.py:410-509: the same statistics from bin counts, or the float
    merge with the asymptotic p forced when vmax - vmin + 1 passes max_bins."""
    n_ref, n_tar = ref_sorted.shape[1], tar.shape[1]
    vmin_ref = int(np.min(ref_sorted)) if n_ref > 0 else 0
    vmax_ref = int(np.max(ref_sorted)) if n_ref > 0 else 0
    vmin_tar = int(np.min(tar)) if n_tar > 0 else 0
    vmax_tar = int(np.max(tar)) if n_tar > 0 else 0
    vmin = min(vmin_ref, vmin_tar)
    vmax = max(vmax_ref, vmax_tar)
    Kp1 = int(vmax - vmin + 1) if (n_ref > 0 and n_tar > 0) else 1
    if Kp1 > max_bins:
        return rank_sum_float(ref_sorted.astype(np.float64), np.sort(tar.astype(np.float64), axis=1),
                              tie_correction, continuity_correction, use_asymptotic=True)
    U1, tie, has = merge_stats(ref_sorted, tar)
    return _p_of(U1, tie, has, n_ref, n_tar, tie_correction, continuity_correction,
                 use_asymptotic), U1


def chunk_size(n_genes, num_workers, n_cells):
    """de_.py:291-333 with the dense size estimate of de_.py:336-367."""
    if num_workers == 1:
        return n_genes
    data_size_mb = n_cells * n_genes * 8 / 1024 / 1024
    per = 2 if data_size_mb > 1024 else 1
    return max(4096, n_genes // (num_workers * per))


def fold_changes(target_means, ref_means, clip_value=20.0):
    """de_.py:189-230."""
    target_means = np.asarray(target_means)
    ref_means = np.asarray(ref_means)
    ref_zero_mask = ref_means < 1e-10
    target_zero_mask = target_means < 1e-10
    with np.errstate(divide="ignore", invalid="ignore"):
        fc = target_means / ref_means
        fc = np.where(target_zero_mask, 1.0 / clip_value, fc)
        fc = np.where(ref_zero_mask, clip_value, fc)
    return fc


def pden(obs_vals, X, gene_names, reference, groups=None, tie_correction=True,
         continuity_correction=True, use_asymptotic=None, min_samples=2, max_bins=100_000,
         prefer_hist_if_int=False, num_workers=1, clip_value=20.0) -> pd.DataFrame:
    """The whole pipeline on cell labels obs_vals and X (cells x genes, dense
    or scipy sparse), with the task loop run in-process."""
    obs_vals = np.asarray(obs_vals)
    uniq = np.unique(obs_vals)
    if groups is None:
        s = set(uniq)
        s.discard(reference)
        groups = list(s)
    else:
        groups = [g for g in groups if g != reference and g in uniq]
    ref_rows = np.where(obs_vals == reference)[0]
    names, rows = [], []
    for g in groups:
        if g in names:
            continue
        idx = np.where(obs_vals == g)[0]
        if len(idx) >= min_samples:
            names.append(g)
            rows.append(idx)
    X_dense = X.toarray() if scipy.sparse.issparse(X) else np.array(X, copy=True)
    n_cells, n_genes = X_dense.shape
    use_hist = prefer_hist_if_int and np.issubdtype(X_dense.dtype, np.integer) and num_workers > 1
    ref_sorted = np.sort(X_dense[ref_rows, :], axis=0)
    cs = chunk_size(n_genes, num_workers, n_cells) if num_workers > 1 else n_genes
    P = np.full((len(names), n_genes), np.nan)
    U = np.full((len(names), n_genes), np.nan)
    for gi, idx in enumerate(rows):
        tar_dense = X_dense[idx]
        for g0 in range(0, n_genes, max(cs, 1)):
            g1 = min(n_genes, g0 + cs)
            ref_chunk = ref_sorted[:, g0:g1].T
            if use_hist:
                p, u = rank_sum_hist(ref_chunk, tar_dense[:, g0:g1].T.astype(np.int64),
                                     tie_correction, continuity_correction, use_asymptotic,
                                     max_bins)
            else:
                p, u = rank_sum_float(ref_chunk, np.sort(tar_dense[:, g0:g1].T, axis=1),
                                      tie_correction, continuity_correction, use_asymptotic)
            P[gi, g0:g1] = p
            U[gi, g0:g1] = u
    ref_means = np.mean(X_dense[ref_rows], axis=0)
    fc = np.empty((len(names), n_genes))
    lfc = np.empty((len(names), n_genes))
    for gi, idx in enumerate(rows):
        f = fold_changes(np.mean(X_dense[idx], axis=0), ref_means, clip_value)
        fc[gi] = f
        with np.errstate(divide="ignore", invalid="ignore"):
            lfc[gi] = np.log2(f)
    p_flat = P.reshape(-1)
    fdr = np.full_like(p_flat, np.nan)
    ok = np.isfinite(p_flat)
    if np.any(ok):
        fdr[ok] = false_discovery_control(p_flat[ok], method="bh")
    return pd.DataFrame({
        "target": np.repeat(np.asarray(names, dtype=object), n_genes),
        "feature": np.tile(np.asarray(gene_names), len(names)),
        "p_value": p_flat,
        "u_statistic": U.reshape(-1),
        "fold_change": fc.reshape(-1),
        "log2_fold_change": lfc.reshape(-1),
        "fdr": fdr,
    })
